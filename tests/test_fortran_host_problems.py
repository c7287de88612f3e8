"""The Fortran host's other user problems (<gpu> problem), with the plasma
drawn by the reference's own random stream (<gpu> loader = 1):

* plasma oscillation (user/user_plasmaosc.F90): electrons kicked by
  beta sin(2 pi x / (mx0 - 5)) over an immobile background oscillate at the
  plasma frequency of the leapfrog scheme, sin(w dt / 2) = w_pe dt / 2 with
  w_pe^2 = (c/c_omp)^2 gamma0 / (me (1 + me/mi)) (the reference's charge
  normalisation, code/particles.F90:206-235).  The field energy oscillates at
  2w; its frequency, fitted over 20 periods, matches to 1%.  The filter's
  transfer at the kick's wavelength (4 passes) moves w by 6e-4 and a
  thermal (Bohm-Gross) shift by 1e-5, both inside the bar.
* two-stream (user/user_twostream.F90): two electron beams, ions removed;
  the run starts at the loader's particle count, the beams give up kinetic
  energy to the unstable field growing from noise, and the total energy
  stays within 1%.
The Weibel problem and the fused/split sequences are in
tests/test_fortran_host.py."""
import os
import subprocess

import numpy as np
import pytest

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
mx0 = {nx}
my0 = {ny}
mz0 = {nz}
<algorithm>
highorder = 0
Corr = 1.025
ntimes = 4
<output>
interval = 1
<particles>
ppc0 = {ppc0}
delgam = {delgam}
me = 1.
mi = {mi}
gamma0 = {gamma0}
c_omp = {c_omp}
<problem>
temperature_ratio = 1
<gpu>
order = {order}
problem = {problem}
loader = 1
"""


def run(tmp_path, **kw):
    inp = tmp_path / "input"
    inp.write_text(INPUT.format(**kw))
    out = tmp_path / "energy.txt"
    r = subprocess.run([EXE, "-i", str(inp), "-o", str(out)], capture_output=True, text=True, timeout=300)
    print(r.stdout[-1500:], r.stderr[-1500:])
    assert r.returncode == 0
    return np.loadtxt(out, ndmin=2), r.stdout


def fitted_frequency(t, y):
    """angular frequency of the dominant oscillation of y(t): least squares
    of a + b cos(w t) + c sin(w t) over a fine scan of w around the
    periodogram's peak"""
    y = y - y.mean()
    ws = np.linspace(0.02, 1.0, 4000)
    power = [abs(np.sum(y * np.exp(-1j * w * t))) for w in ws]
    w0 = ws[int(np.argmax(power))]
    best, wbest = None, w0
    for w in np.linspace(0.97 * w0, 1.03 * w0, 3001):
        A = np.stack([np.ones_like(t), np.cos(w * t), np.sin(w * t)], axis=1)
        res = np.linalg.lstsq(A, y, rcond=None)[1]
        r = res[0] if len(res) else np.inf
        if best is None or r < best:
            best, wbest = r, w
    return wbest


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
@pytest.mark.parametrize("order", [1, 3])
def test_plasma_oscillation_frequency(tmp_path, order):
    c, c_omp, gamma0, mi = 0.45, 5.0, 0.02, 1836.0
    h, _ = run(tmp_path, last=700, nx=256, ny=4, nz=1, ppc0=16, delgam=1e-7, mi=mi, gamma0=gamma0, c_omp=c_omp,
               order=order, problem=2)
    t, ee = h[:, 0], h[:, 1]
    assert np.all(np.isfinite(h))
    g0 = 1.0 / np.sqrt(1.0 - gamma0 ** 2)
    wpe = (c / c_omp) * np.sqrt(g0 / (1.0 + 1.0 / mi))
    w_leapfrog = 2.0 * np.arcsin(0.5 * wpe)
    w = 0.5 * fitted_frequency(t, ee)        # the energy oscillates at 2w
    print(f"order {order}: measured w {w:.6f}, leapfrog {w_leapfrog:.6f}, continuum {wpe:.6f}")
    assert abs(w / w_leapfrog - 1.0) < 0.01
    # the ions were removed: the kinetic energy is the electrons' alone
    assert np.all(h[:, 3] == 0.0) and np.all(h[:, 4] > 0.0)


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built")
def test_two_stream_problem_runs_electron_beams(tmp_path):
    h, out = run(tmp_path, last=200, nx=16, ny=16, nz=16, ppc0=8, delgam=1e-4, mi=16.0, gamma0=0.2, c_omp=8.0,
                 order=2, problem=1)
    words = out.replace("  ", " ").split()
    ions = int(words[words.index("ions") + 1])
    lecs = int(words[words.index("electrons") + 1])
    assert ions == 0 and lecs == 16 ** 3 * 4     # ppc0/2 electrons per cell, two beams
    assert np.all(np.isfinite(h))
    assert np.all(h[:, 3] == 0.0)
    # the beams give up kinetic energy to the growing electrostatic field;
    # the total stays within 1%
    tot = h[:, 1] + h[:, 2] + h[:, 4]
    print(f"two-stream: E energy {h[0, 1]:.3e} -> {h[-1, 1]:.3e}, electron KE {h[0, 4]:.4e} -> {h[-1, 4]:.4e}, "
          f"total drift {tot[-1] / tot[0] - 1:.2e}")
    assert h[-1, 1] > 10.0 * h[0, 1] and h[-1, 1] > 0.02 * h[-1, 4]
    assert abs(tot[-1] / tot[0] - 1.0) < 0.01
