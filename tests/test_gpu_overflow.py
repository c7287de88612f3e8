"""Tile segments that overflow: force-free particles (q/m = 0, q = 0) stream
toward the plane x = nx/2 from both sides, so the tiles next to it hold
about twice their particles after a few steps, more than their segment
capacity (1.25 x the fullest tile at binning).  The excess goes to the
overflow list, which the flat kernel serves the next step, and a filling
list re-bins the species.  No particle may be lost or duplicated, and every
particle must follow its free-streaming path."""
import numpy as np
import pytest

from gpu_util import periodic_diff

pytestmark = pytest.mark.gpu

C = 0.45


@pytest.mark.parametrize("dim,order,shape,ppc", [(3, 2, (32, 16, 16), 8), (3, 3, (32, 16, 16), 8),
                                                 (2, 1, (64, 32, 1), 24)])
def test_converging_streams_keep_every_particle(dim, order, shape, ppc):
    import orc
    import tmpg
    nx, ny, nz = shape
    ctx = tmpg.Context(dim, order, nx, ny, nz, ntimes=0, c=C, qm=(0.0, 0.0), q=(0.0, 0.0), mass=(1.0, 1.0),
                       track_ids=True)
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    _, _, _, ng, ngz = orc.extent(cfg)
    lo = (ng // 2 + 1, ng // 2 + 1, ngz // 2 + 1)
    rng = np.random.default_rng(17 + dim + order)
    ncell = nx * ny * (nz if dim == 3 else 1)
    state = []
    for s in range(2):
        n = ppc * ncell
        p = np.zeros((7, n), np.float32)
        p[0] = lo[0] + rng.random(n) * nx
        p[1] = lo[1] + rng.random(n) * ny
        p[2] = (lo[2] + rng.random(n) * nz) if dim == 3 else lo[2] + 0.5
        u = np.where(p[0] < lo[0] + 0.5 * nx, 0.9, -0.9).astype(np.float32)
        p[3] = u
        p[6] = 1.0
        ctx.set_particles(s, p)
        state.append(p)
    v = 0.9 / np.sqrt(1.0 + 0.81) * C          # cells per step
    steps = 24
    for k in range(steps):
        ctx.step(1)
        d = ctx.diag()
        assert d[4] == state[0].shape[1] and d[5] == state[1].shape[1], (k, d[4], d[5])
    for s in range(2):
        p, ids = ctx.get_particles(s, with_ids=True)
        assert np.array_equal(np.sort(ids), np.arange(state[s].shape[1]))
        p0 = state[s][:, ids]
        x_expect = p0[0] + np.sign(p0[3]) * v * steps
        assert np.abs(periodic_diff(p[0], x_expect, lo[0], nx)).max() < 2e-3
        assert np.abs(periodic_diff(p[1], p0[1], lo[1], ny)).max() < 1e-5
        assert np.array_equal(p[3], p0[3])
    # the strip next to the plane really was over capacity: its density
    # doubled (particles from both sides within v*steps of the plane)
    x = ctx.get_particles(0)[0] - lo[0]
    near = np.abs(x - 0.5 * nx) < 0.8 * v * steps
    ctx.close()
    assert near.sum() > 1.8 * ppc * (ncell / nx) * (2 * 0.8 * v * steps)
