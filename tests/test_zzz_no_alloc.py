"""tmpg_gauss_residual_noalloc: the Gauss's-law diagnostic on the library's
static per-GPU scratch, with no device allocation, against
tmpg_gauss_residual on the same state.  Two calls in a row, the second on a
smaller grid (the scratch keeps the first call's contents, so anything not
cleared would show), then a grid beyond the scratch's size, which is refused.
The residuals are FP64 maxima and sums of FP64 atomic accumulations, which may
add in another order: the bar is 1e-9 relative."""
import numpy as np
import pytest

from gpu_util import make_context

pytestmark = pytest.mark.gpu

PHYS = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.3, ppc0=8.0, me=1.0, mi=1.0)


def _state(tmpg, dim, order, shape, steps, seed):
    nx, ny, nz = shape
    ctx = make_context(tmpg, dim, order, nx, ny, nz, PHYS, ntimes=4, track_ids=False)
    rng = np.random.default_rng(seed)
    h, hz = ctx.ng // 2, ctx.ngz // 2
    n = 4 * nx * ny * (nz if dim == 3 else 1)
    pos = np.vstack([h + 1 + rng.random(n) * nx, h + 1 + rng.random(n) * ny,
                     (hz + 1 + rng.random(n) * nz) if dim == 3 else np.full(n, hz + 1.5)])
    for s in range(2):
        mom = rng.normal(0.0, 0.2, (3, n))
        mom[0] += 0.3 if s == 0 else -0.3
        ctx.set_particles(s, np.vstack([pos, mom, np.ones((1, n))]).astype(np.float32))
    ctx.step(steps)
    return ctx


def _compare(ctx):
    a = ctx.gauss_residual()
    b = ctx.gauss_residual_noalloc()
    print("allocating", a, "allocation-free", b)
    assert a[1] > 0.0 and a[3] > 0.0
    np.testing.assert_allclose(b, a, rtol=1e-9, atol=1e-12 * a[1])


def test_noalloc_gauss_matches_the_allocating_entry_point(tmpg):
    big = _state(tmpg, 3, 3, (32, 24, 16), 30, 1)
    small = _state(tmpg, 2, 1, (40, 24, 1), 30, 2)
    _compare(big)
    _compare(small)      # the second call, on fewer cells than the first
    _compare(big)        # and back
    big.close()
    small.close()


def test_noalloc_gauss_refuses_grids_beyond_its_scratch(tmpg):
    ctx = make_context(tmpg, 2, 1, 2049, 2048, 1, PHYS, ntimes=4, track_ids=False)
    with pytest.raises(tmpg.TmpgError, match="exceed its scratch"):
        ctx.gauss_residual_noalloc()
    # the allocating entry point still serves this grid
    assert ctx.gauss_residual()[0] == 0.0
    ctx.close()
