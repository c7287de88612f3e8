"""The Fortran host (bin/tristan-mp-gpu) drives the library through m_tmpg:
reads a reference-format input file, loads the two-region Weibel setup
(user/user_weibel.F90:252-292) and runs the main loop; the fused lap
(tmpg_step) and the reference's one-routine-at-a-time sequence agree."""
import os
import subprocess

import numpy as np
import pytest

from gpu_util import periodic_diff

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tristan-mp-pu-master_densdecomp_amd", "bin", "tristan-mp-gpu")

pytestmark = pytest.mark.gpu

INPUT = """<node_configuration>
sizex = 1
sizey = 1
<time>
last = {last}
c = .45
<grid>
mx0 = 64
my0 = 48
mz0 = 1
<algorithm>
highorder = {highorder}
Corr = 1.025
ntimes = 4
<output>
interval = 10
<particles>
ppc0 = 16
delgam = 2.5e-3
me = 1.
mi = 1.
gamma0 = 0.5
c_omp = 5
<problem>
temperature_ratio = 1
<gpu>
order = 2
fused = {fused}
seed = 3
pusher = {pusher}
"""


def run(tmp_path, fused, last=30, highorder=0, pusher=0):
    tag = f"{fused}{highorder}{pusher}"
    inp = tmp_path / f"input{tag}"
    inp.write_text(INPUT.format(last=last, fused=fused, highorder=highorder, pusher=pusher))
    out = tmp_path / f"energy{tag}.txt"
    r = subprocess.run([EXE, "-i", str(inp), "-o", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return np.loadtxt(out), r.stdout


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
def test_fortran_host_runs_and_sequences_agree(tmp_path):
    a, out = run(tmp_path, 1)
    assert "ions 24576" in out.replace("  ", " ") or "ions" in out
    b, _ = run(tmp_path, 0)
    assert a.shape == b.shape and a.shape[0] == 3
    assert np.all(np.isfinite(a))
    # kinetic energy dominates and is nearly constant early on
    assert abs(a[-1, 3] - a[0, 3]) < 1e-2 * a[0, 3]
    # fused and split laps: same physics up to FP32 summation order
    np.testing.assert_allclose(a[:, 3:], b[:, 3:], rtol=1e-5)
    np.testing.assert_allclose(a[:, 1:3], b[:, 1:3], rtol=2e-2, atol=1e-12 * a[0, 3])


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
def test_fortran_host_passes_solver_and_pusher_choices(tmp_path):
    """highorder = 1 (4th-order field solver) and <gpu> pusher = 1 (Vay)
    reach the library, and the fused and split laps still agree.  The 4th-
    order solver visibly changes the field energies; in this weak-field start
    Boris and Vay differ far below them (their difference is pinned by the
    parity tests against the reference's -Dvay build)."""
    base, out = run(tmp_path, 1, last=60)
    assert "field solver 2nd order" in out and "pusher Boris" in out
    for ho, pu in ((1, 0), (0, 1)):
        a, out = run(tmp_path, 1, last=60, highorder=ho, pusher=pu)
        b, _ = run(tmp_path, 0, last=60, highorder=ho, pusher=pu)
        assert ("4th order" in out) == (ho == 1) and ("pusher Vay" in out) == (pu == 1)
        np.testing.assert_allclose(a[:, 3:], b[:, 3:], rtol=1e-5)
        assert np.all(np.isfinite(a))
        if ho:
            # the 4th-order solver changes the fields well beyond round-off
            assert not np.allclose(a[:, 1:3], base[:, 1:3], rtol=1e-6, atol=0)


RESTART = """<restart>
irestart = {irestart}
intrestart = {intrestart}
laprestart = {laprestart}
lnamedrestartint = {named}
"""


def run_in(d, last, irestart=0, intrestart=2000, laprestart=0, named=1000000):
    """the host in directory d (its restart/ folder lives there), ids tracked"""
    d.mkdir(parents=True, exist_ok=True)
    inp = d / "input"
    inp.write_text(INPUT.format(last=last, fused=1, highorder=0, pusher=0) + "track_ids = 1\n"
                   + RESTART.format(irestart=irestart, intrestart=intrestart, laprestart=laprestart, named=named))
    out = d / "energy.txt"
    r = subprocess.run([EXE, "-i", str(inp), "-o", str(out)], capture_output=True, text=True, timeout=120, cwd=d)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    return np.loadtxt(out, ndmin=2), r.stdout


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
def test_fortran_host_checkpoint_and_resume(tmp_path):
    """write_restart / restart() (code/output.F90:2085-2262,
    code/restart.F90:209-254) in the reference's file format: a 40-lap run
    that checkpoints every 20 laps, and a run resumed from its lap-20 files,
    reach the same state at lap 40, up to the order of the deposit's FP32
    atomics, which differs between any two runs (the bars are those of the
    fused/split comparison above)."""
    import shutil

    import restart_io as rio
    a_dir, b_dir = tmp_path / "a", tmp_path / "b"
    a, out_a = run_in(a_dir, 40, intrestart=20, named=20)
    assert "wrote restart at lap 20" in out_a and "wrote restart at lap 40" in out_a
    names = {lap: (a_dir / f"restart/restflds.default.lap{lap:06d}.000.d",
                   a_dir / f"restart/restprtl.default.lap{lap:06d}.000.d") for lap in (20, 40)}
    # the lap-20 files as the reference's reader sees them
    f20 = rio.read_fields(names[20][0])
    assert (f20["mx"], f20["my"], f20["mz"], f20["lap"]) == (64 + 7, 48 + 7, 1, 20)
    p20 = rio.read_particles(names[20][1])
    n_ion, n_lec = p20["ions"], p20["lecs"]
    assert n_ion == n_lec == 64 * 48 * 16 // 2 and p20["totalpartnum"] == n_ion + n_lec
    for s in p20["species"]:
        assert np.unique(s["ind"]).size == s["ind"].size
        assert np.all(s["ch"] > 0) and np.all(np.isfinite(s["u"]))
    # the unnamed file holds the last checkpoint
    assert rio.read_fields(a_dir / "restart/restflds.default.000.d")["lap"] == 40

    (b_dir / "restart").mkdir(parents=True)
    for p in names[20]:
        shutil.copy(p, b_dir / "restart" / p.name)
    b, out_b = run_in(b_dir, 40, irestart=1, intrestart=20, laprestart=20, named=20)
    assert "read restart: lap 20" in out_b
    assert list(b[:, 0]) == [30, 40] and list(a[:, 0]) == [10, 20, 30, 40]
    np.testing.assert_allclose(b[:, 3:], a[2:, 3:], rtol=1e-5)
    np.testing.assert_allclose(b[:, 1:3], a[2:, 1:3], rtol=2e-2, atol=1e-12 * a[0, 3])

    fa, fb = rio.read_fields(names[40][0]), rio.read_fields(b_dir / "restart/restflds.default.lap000040.000.d")
    for k in rio.FIELD_NAMES:
        scale = np.abs(fa[k]).max()
        assert np.abs(fa[k] - fb[k]).max() <= 2e-2 * scale + 1e-30, k
    pa, pb = rio.read_particles(names[40][1]), rio.read_particles(b_dir / "restart/restprtl.default.lap000040.000.d")
    for sa, sb in zip(pa["species"], pb["species"]):
        ia, ib = np.argsort(sa["ind"]), np.argsort(sb["ind"])
        assert np.array_equal(sa["ind"][ia], sb["ind"][ib])
        for k in ("x", "y", "z"):
            assert np.abs(sa[k][ia] - sb[k][ib]).max() < 1e-3, k
        for k in ("u", "v", "w"):
            assert np.abs(sa[k][ia] - sb[k][ib]).max() < 1e-3 * np.abs(sa[k]).max(), k


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
def test_fortran_host_write_restart_file_trigger(tmp_path):
    """a file named write_restart asks for a checkpoint at the next lap and
    is deleted (code/output.F90:2099-2110, 2262-2265)"""
    import restart_io as rio
    d = tmp_path / "c"
    d.mkdir()
    (d / "write_restart").write_text("")
    run_in(d, 3)
    assert not (d / "write_restart").exists()
    assert rio.read_fields(d / "restart/restflds.default.000.d")["lap"] == 1
    # intrestart = 0 turns the periodic checkpoints off
    e = tmp_path / "e"
    run_in(e, 3, intrestart=0, named=0)
    assert not (e / "restart").exists()


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
@pytest.mark.parametrize("dim,order", [(2, 2), (3, 3)])
def test_fortran_host_resumes_reference_state(tmp_path, dim, order):
    """Fixture exchange through the restart format: the reference's state
    before a step (tests/golden/ref_<dim>d_ord<order>.npz, from the compiled
    reference) is written as restflds/restprtl files, the Fortran host
    resumes from them, runs one lap and checkpoints; that checkpoint matches
    the reference FP64 build's state after the step."""
    import ast

    import restart_io as rio
    g = np.load(os.path.join(ROOT, "tests", "golden", f"ref_{dim}d_ord{order}.npz"))
    prm = ast.literal_eval(str(g["params"]))
    nx, ny, nz = prm["nx"], prm["ny"], prm["nz"]
    ng = 5 if order == 1 else 7
    h = ng // 2
    pad = ((h, ng - h) if dim == 3 else (0, 0), (h, ng - h), (h, ng - h))
    fields = {k: np.pad(g["f_in"][q], pad, mode="wrap") for q, k in enumerate(rio.FIELD_NAMES)}
    d = tmp_path / "r"
    (d / "restart").mkdir(parents=True)
    rio.write_fields(d / "restart/restflds.default.000.d", fields, lap=0)
    sp = []
    for key in ("pi_in", "pe_in"):
        p = g[key]
        sp.append(dict(zip(rio.PART_REALS, p)) | dict(ind=np.arange(p.shape[1], dtype=np.int32)))
    rio.write_particles(d / "restart/restprtl.default.000.d", sp[0], sp[1])
    inp = d / "input"
    inp.write_text(f"""<time>
last = 1
c = {prm['c']}
<grid>
mx0 = {nx}
my0 = {ny}
mz0 = {nz}
<algorithm>
highorder = 0
Corr = {prm['corr']}
ntimes = {prm['ntimes']}
<output>
interval = 1
<particles>
ppc0 = {prm['ppc0']}
me = {prm['me']}
mi = {prm['mi']}
gamma0 = {prm['gamma0']}
c_omp = {prm['c_omp']}
<gpu>
dim = {dim}
order = {order}
track_ids = 1
<restart>
irestart = 1
intrestart = 1
""")
    r = subprocess.run([EXE, "-i", str(inp)], capture_output=True, text=True, timeout=120, cwd=d)
    print(r.stdout[-2000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "read restart: lap 0" in r.stdout and "wrote restart at lap 1" in r.stdout
    f = rio.read_fields(d / "restart/restflds.default.000.d")
    assert f["lap"] == 1
    inner = (slice(h, h + nz) if dim == 3 else slice(0, 1), slice(h, h + ny), slice(h, h + nx))
    ref = g["step_f64"]
    for q, k in enumerate(rio.FIELD_NAMES):
        scale = np.abs(ref[q]).max()
        err = np.abs(f[k][inner] - ref[q]).max()
        assert err <= 2e-6 * scale, (k, err, scale)
    pr = rio.read_particles(d / "restart/restprtl.default.000.d")
    for s, key in enumerate(("step_pi_f64", "step_pe_f64")):
        want = g[key]
        got = pr["species"][s]
        order_ = np.argsort(got["ind"])
        assert np.array_equal(got["ind"][order_], np.arange(want.shape[1]))
        # positions modulo the periodic box; in 2D the library keeps no z
        # (the reference's 2D z drifts with w and is never read)
        for q, (k, n) in enumerate(zip(("x", "y", "z")[:dim], (nx, ny, nz))):
            assert np.abs(periodic_diff(got[k][order_], want[q], h + 1, n)).max() < 2e-5, k
        for q, k in enumerate(("u", "v", "w"), start=3):
            assert np.abs(got[k][order_] - want[q]).max() <= 2e-6 * np.abs(want[q]).max(), k
