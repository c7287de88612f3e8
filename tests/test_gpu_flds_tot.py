"""flds.tot.NNN from the GPU state (tmpg_write_flds_tot) and from the Fortran
host (<gpu> flds_tot = 1, output_tot at the top of every interval-th lap,
code/tristanmainloop.F90:102-109, code/output.F90:2546-2556).

* the context's file equals the host writer's on the arrays the getters
  return (the same particle order, so bit-equal);
* the Fortran host's files at laps 0, 5, 10 of a 2D Weibel run: the 17
  datasets (the 2D two-stream run: the in-plane list of 15), the fields and
  currents zero at lap 0, and the sampled densities summing to the
  particle weights (every particle adds ch to the 25 nodes of its box, the
  fold brings the ghost parts home, and every interior node divides by 25).
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import orc
from gpu_util import make_context, random_state
from test_flds_tot_hdf5 import H5LIB, H5Reader

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd", "bin", "tristan-mp-gpu")
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import flds_tot_ref as ref  # noqa: E402

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(H5LIB), reason="no HDF5 C library")]


@pytest.mark.parametrize("dim,order,shape", [(2, 2, (24, 20, 1)), (3, 3, (12, 10, 8))])
def test_context_file_equals_host_writer(tmpg, tmp_path, dim, order, shape):
    nx, ny, nz = shape
    p = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.5, ppc0=4.0, me=1.0, mi=16.0)
    f, pi, pe = random_state(dim, order, nx, ny, nz, 900, seed=dim + order)
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    ctx = make_context(tmpg, dim, order, nx, ny, nz, p)
    ctx.set_fields(orc.fill_from_interior(cfg, f, np.float32)[:6])
    ctx.set_particles(0, pi)
    ctx.set_particles(1, pe)
    ctx.step(2)
    a = tmp_path / "ctx.h5"
    b = tmp_path / "host.h5"
    ctx.write_flds_tot(str(a), istep=2)
    tmpg.flds_tot_write(str(b), dim, order, nx, ny, nz, 2, ctx.get_fields(), ctx.get_currents(),
                        ctx.get_particles(0), ctx.get_particles(1))
    ctx.close()
    r = H5Reader()
    ga, gb = r.read(a), r.read(b)
    assert list(ga) == ref.DATASETS_3D
    for k in ga:
        np.testing.assert_array_equal(ga[k], gb[k], err_msg=k)
    assert np.abs(ga["jx"]).max() > 0 and np.abs(ga["v3y"]).max() > 0


INPUT = """<time>
last = 10
c = .45
<grid>
mx0 = {nx}
my0 = {ny}
mz0 = 1
<algorithm>
highorder = 0
ntimes = 4
<output>
interval = 5
istep = 1
<particles>
ppc0 = 8
delgam = 1e-4
me = 1.
mi = 16.
gamma0 = {gamma0}
c_omp = 8
<gpu>
dim = 2
order = 2
problem = {problem}
flds_tot = 1
"""


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
@pytest.mark.parametrize("problem,nsets", [(0, 17), (1, 15)])
def test_fortran_host_writes_flds_tot(tmp_path, problem, nsets):
    nx, ny = 32, 24
    (tmp_path / "input").write_text(INPUT.format(nx=nx, ny=ny, problem=problem, gamma0=0.2 if problem else 0.5))
    r = subprocess.run([EXE, "-i", "input"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    words = r.stdout.replace("  ", " ").split()
    ni = int(words[words.index("ions") + 1])
    ne = int(words[words.index("electrons") + 1])
    files = sorted(os.listdir(tmp_path / "output"))
    assert files == ["flds.tot.000", "flds.tot.001", "flds.tot.002"]
    rd = H5Reader()
    mx, my, ng = nx + 7, ny + 7, 7
    h = ng // 2
    inner = (0, slice(h, h + ny), slice(h, h + nx))
    for n, name in enumerate(files):
        d = rd.read(tmp_path / "output" / name)
        assert len(d) == nsets and all(a.shape == (1, my, mx) for a in d.values())
        for k in ("ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"):
            assert np.all(np.isfinite(d[k]))
            if n == 0:
                assert np.all(d[k] == 0.0), k
        np.testing.assert_allclose(d["dens"][inner].sum(dtype=np.float64), ni + ne, rtol=1e-5)
        np.testing.assert_allclose(d["densi"][inner].sum(dtype=np.float64), ni, rtol=1e-5)
        assert np.abs(d["v3x"]).max() < 1.0
    assert np.abs(d["jx"]).max() > 0
