"""Long main loops keep a fixed set of HIP resources.

tmpg_step records eight timing events per step into a fixed ring and folds
the oldest step into running sums when its slot is reused, so a run of many
steps with no tmpg_last_timings in between neither grows nor fails; the sums
then cover every step.  Reference loop: code/tristanmainloop.F90:90-346
(a production run makes 10^4-10^5 laps).
"""
import numpy as np
import pytest

from gpu_util import make_context

pytestmark = pytest.mark.gpu

PHYS = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.3, ppc0=4.0, me=1.0, mi=1.0)


def test_twenty_thousand_steps_without_reading_timings(tmpg):
    nx, ny = 16, 16
    ctx = make_context(tmpg, 2, 1, nx, ny, 1, PHYS, track_ids=False)
    ctx.load_uniform(0, 2, 0.3, 1e-3, 1)
    ctx.load_uniform(1, 2, -0.3, 1e-3, 2)
    n0 = ctx.count(0) + ctx.count(1)
    e0 = ctx.diag()
    nsteps = 20000
    for _ in range(nsteps // 1000):
        ctx.step(1000)
    ctx.sync()
    ms = ctx.last_timings()
    assert ms[3] == nsteps
    assert np.all(np.isfinite(ms)) and ms[4] > 0.0 and ms[5] > 0.0
    # the phases add up to the total, and the push kernel is part of its phase
    assert abs(ms[0] + ms[1] + ms[2] - ms[4]) < 1e-6 * ms[4] + 1e-3
    assert ms[5] <= ms[1] * (1.0 + 1e-6)
    # a second read starts from zero
    ctx.step(3)
    ms2 = ctx.last_timings()
    assert ms2[3] == 3
    d = ctx.diag()
    assert ctx.count(0) + ctx.count(1) == n0
    assert np.all(np.isfinite(d[:4]))
    # energy stays bounded over the run (no runaway from a broken state)
    tot0, tot = e0[0] + e0[1] + e0[2] + e0[3], d[0] + d[1] + d[2] + d[3]
    assert abs(tot - tot0) < 0.05 * tot0
