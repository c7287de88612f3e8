"""GPU denoiser (csrc/denoise.hip through the C ABI) against its CPU oracle
(oracle/denoiser_ref.py, float64 with BF16 rounding where the GPU rounds):
loss and every layer's gradient, 2D and 3D; Adam training lowers the loss;
the in-step correction keeps the deposit's charge continuity."""
import os
import sys

import numpy as np
import pytest

import orc
from gpu_util import make_context, random_state

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import denoiser_ref as dr  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tmpg():
    import tmpg as t
    t.load_library()
    return t


def smooth_pair(rng, dim, shape):
    """a smooth target and a noisy version of it"""
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    jt = np.stack([np.sin(2 * np.pi * (x / nx + 2 * y / ny) + k) * (1 + 0.3 * np.cos(2 * np.pi * z / max(nz, 1)))
                   for k in range(3)]).astype(np.float32)
    jlo = (jt + 0.7 * rng.standard_normal(jt.shape)).astype(np.float32)
    return jlo, jt


@pytest.mark.parametrize("dim,shape,batch,lam", [(2, (1, 24, 32), 2, 0.0), (3, (8, 10, 12), 1, 0.0),
                                                 (2, (1, 24, 32), 2, 3.0), (3, (8, 10, 12), 1, 3.0)])
def test_loss_and_gradients_match_oracle(tmpg, dim, shape, batch, lam):
    rng = np.random.default_rng(dim)
    m = tmpg.Model(dim, seed=5)
    p = m.get_params().astype(np.float64)
    # make the last layer non-trivial so every gradient path is exercised
    p = p + 0.05 * rng.standard_normal(p.size)
    m.set_params(p.astype(np.float32))
    p32 = m.get_params()
    pairs = [smooth_pair(rng, dim, shape) for _ in range(batch)]
    jlo = np.stack([a for a, _ in pairs])
    jt = np.stack([b for _, b in pairs])
    s = 0.9
    m.set_scale(s)
    m.set_amplitude_weight(lam)
    loss, g = m.train_batch(jlo, jt)
    net = dr.Net(dim, params=p32)
    lo, go = 0.0, np.zeros_like(g, dtype=np.float64)
    for b in range(batch):
        l_, g_, _ = net.loss_and_grad(jlo[b], jt[b], s, bf=True, lam=lam)
        lo += l_ / batch
        go += g_ / batch
    assert abs(loss - lo) < 1e-4 * lo, (loss, lo)
    o = 0
    for li, (c, t, i) in enumerate(net.shapes):
        for n in (c * t * i, c):
            a, b = g[o:o + n].astype(np.float64), go[o:o + n]
            o += n
            if li == len(net.shapes) - 1 and n == c:
                # the last layer's bias is a constant A, whose curl is zero:
                # its gradient vanishes in exact arithmetic and is left with
                # BF16 rounding residue only
                wn = np.linalg.norm(go[o - n - c * t * i:o - n])
                assert np.linalg.norm(a) < 1e-3 * wn and np.linalg.norm(b) < 1e-3 * wn, (a, b, wn)
                continue
            err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            cos = a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30)
            assert err < 2e-2 and cos > 0.9995, (li, n, err, cos)


def test_training_lowers_loss(tmpg):
    rng = np.random.default_rng(3)
    m = tmpg.Model(2, seed=1)
    for _ in range(4):
        a, b = smooth_pair(rng, 2, (1, 32, 32))
        m.add_sample(a, b)
    a, b = smooth_pair(rng, 2, (1, 32, 32))
    l0 = m.eval(a, b)
    losses = m.train(300, lr=2e-3, seed=2)
    l1 = m.eval(a, b)
    assert np.all(np.isfinite(losses))
    assert losses[-20:].mean() < 0.7 * losses[:20].mean()
    assert l1 < 0.7 * l0, (l0, l1)


@pytest.mark.parametrize("dim", [2, 3])
def test_denoised_current_keeps_continuity(tmpg, dim):
    nx, ny, nz = (16, 16, 12) if dim == 3 else (48, 32, 1)
    order = 2
    p = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.5, ppc0=4.0, me=1.0, mi=1.0)
    f, pi, pe = random_state(dim, order, nx, ny, nz, 4000, seed=9)
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    ctx = make_context(tmpg, dim, order, nx, ny, nz, p)
    ctx.set_fields(orc.fill_from_interior(cfg, f, np.float32)[:6])
    ctx.set_particles(0, pi)
    ctx.set_particles(1, pe)
    m = tmpg.Model(dim, seed=3)
    q = m.get_params()
    m.set_params(q + np.float32(0.05) * np.random.default_rng(0).standard_normal(q.size).astype(np.float32))
    ctx.track_continuity(True)
    ctx.reset_currents()
    ctx.push_deposit()
    j0 = ctx.get_currents().copy()
    m.set_scale(float(np.sqrt(np.mean(j0 ** 2))))
    ctx.denoise(m)
    j1 = ctx.get_currents()
    d = ctx.diag()
    qscale = d[7] / np.sqrt(nx * ny * nz)
    assert np.abs(j1 - j0).max() > 1e-3 * np.abs(j0).max()     # the correction is not zero
    assert d[6] < 5e-6 * qscale, (d[6], qscale)
