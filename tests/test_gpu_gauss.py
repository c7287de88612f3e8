"""Gauss's law on the GPU (tmpg_gauss_residual).

* The library's residual equals the same formula evaluated in numpy from the
  library's own state (tmpg_get_fields, tmpg_charge_density): max
  |div_b E - F^n rho| with F^n the step's binomial passes.
* From a neutral start (E = B = 0, both species of the pair plasma at the
  same places, independent thermal momenta) the main loop keeps the residual
  at FP32 round-off for 200 steps, 2D and 3D, shape orders 1-3: the E update
  adds the filtered Esirkepov current, whose divergence is the filtered
  change of rho, and a curl of no divergence.  The bound is stated relative
  to the rms of the filtered charge density the species separation builds
  up; FP32 fields carry about 1e-7 of |E| per addition.
The oracle's main loop meets the same identity in FP64 to 1e-11
(tests/test_oracle_invariants.py::test_step_keeps_gauss_law).
"""
import numpy as np
import pytest

from gpu_util import make_context

pytestmark = pytest.mark.gpu

PHYS = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.3, ppc0=8.0, me=1.0, mi=1.0)


def binomial(a, npasses, dim):
    for _ in range(npasses):
        for ax in ([2, 1, 0] if dim == 3 else [2, 1]):
            a = 0.25 * (np.roll(a, 1, axis=ax) + np.roll(a, -1, axis=ax)) + 0.5 * a
    return a


def div_b(e3, dim):
    d = (e3[0] - np.roll(e3[0], 1, axis=2)) + (e3[1] - np.roll(e3[1], 1, axis=1))
    if dim == 3:
        d = d + (e3[2] - np.roll(e3[2], 1, axis=0))
    return d


def neutral_pairs(ctx, dim, nx, ny, nz, ppc, seed):
    """both species at the same positions, independent momenta"""
    rng = np.random.default_rng(seed)
    h, hz = ctx.ng // 2, ctx.ngz // 2
    n = ppc * nx * ny * (nz if dim == 3 else 1)
    pos = np.vstack([h + 1 + rng.random(n) * nx, h + 1 + rng.random(n) * ny,
                     (hz + 1 + rng.random(n) * nz) if dim == 3 else np.full(n, hz + 1.5)])
    for s in range(2):
        mom = rng.normal(0.0, 0.2, (3, n))
        mom[0] += 0.3 if s == 0 else -0.3
        ctx.set_particles(s, np.vstack([pos, mom, np.ones((1, n))]).astype(np.float32))


CASES = [(2, o) for o in (1, 2, 3)] + [(3, o) for o in (1, 2, 3)]


@pytest.mark.parametrize("dim,order", CASES)
def test_gauss_law_holds_from_a_neutral_start(tmpg, dim, order):
    nx, ny, nz = (24, 20, 16) if dim == 3 else (64, 48, 1)
    ntimes = 4
    ctx = make_context(tmpg, dim, order, nx, ny, nz, PHYS, ntimes=ntimes, track_ids=False)
    neutral_pairs(ctx, dim, nx, ny, nz, 4, 11 + 10 * dim + order)
    g0 = ctx.gauss_residual()
    q0 = abs(ctx.cfg.q[0])
    # rho sums both species with FP64 atomics in any order: zero up to that round-off
    assert g0[0] < 1e-12 * q0 and g0[1] < 1e-12 * q0
    worst = 0.0
    for _ in range(4):
        ctx.step(50)
        g = ctx.gauss_residual()
        worst = max(worst, g[0] / g[1])
    print(f"dim {dim} order {order}: worst residual / rms F^n rho {worst:.3e}, rms {g[1]:.3e}")
    # the residual is FP32 round-off of E, far below the charge separation
    assert g[1] > 1e-3 * q0, g
    assert worst < 2e-4, worst
    # the same number from the library's state, in numpy
    h, hz = ctx.ng // 2, ctx.ngz // 2
    sl = (slice(hz, hz + nz) if dim == 3 else slice(0, 1), slice(h, h + ny), slice(h, h + nx))
    f = ctx.get_fields()
    e3 = np.stack([f[q][sl] for q in range(3)]).astype(np.float64)
    rho = binomial(ctx.charge_density(), ntimes, dim)
    res = div_b(e3, dim) - rho
    g = ctx.gauss_residual()
    np.testing.assert_allclose(g[0], np.abs(res).max(), rtol=1e-6, atol=1e-12 * g[1])
    np.testing.assert_allclose(g[1], np.sqrt(np.mean(rho ** 2)), rtol=1e-9)
    np.testing.assert_allclose(g[2], np.abs(rho).max(), rtol=1e-9)
    np.testing.assert_allclose(g[3], np.abs(div_b(e3, dim)).max(), rtol=1e-9)
