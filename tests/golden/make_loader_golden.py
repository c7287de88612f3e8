"""The reference's own initial plasma (tests/golden/loader_*.npz).

The compiled reference (oracle/_ref/tristan_ref_<dim>d_dd3_f32_user_<setup>,
made by `make -C oracle ref-user`) runs its user module's
init_particle_distribution_user (ref_driver.F90 mode 6): the random seed
123457 (code/communications.F90:224-231), inject_plasma_region,
init_maxw_table, maxwell_dist and the Park-Miller random
(code/particles.F90:2126-2938, code/aux.F90:76-99), then the setup's own
changes (the two-stream and plasma-oscillation setups drop the ions; the
plasma oscillation kicks the electrons) and reorder_particles.  The fixture
keeps every particle (x, y, z, u, v, w, ch per species) and the input
parameters.

    python tests/golden/make_loader_golden.py
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
import orc  # noqa: E402
from make_golden import REF_DIR, read_state, write_state  # noqa: E402

# setup -> input parameters (the reference's input keys)
CASES = {
    "weibel2d": dict(setup="weibel", dim=2, nx=40, ny=32, nz=1, ppc0=8, gamma0=0.5, delgam=2.5e-3, me=1.0, mi=1.0,
                     c_omp=5.0),
    "weibel3d": dict(setup="weibel", dim=3, nx=16, ny=12, nz=10, ppc0=4, gamma0=1.414, delgam=0.01, me=1.0, mi=1.0,
                     c_omp=6.0),
    "twostream3d": dict(setup="twostream", dim=3, nx=16, ny=12, nz=10, ppc0=8, gamma0=0.2, delgam=1e-4, me=1.0,
                        mi=16.0, c_omp=8.0),
    "plasmaosc2d": dict(setup="plasmaosc", dim=2, nx=64, ny=8, nz=1, ppc0=16, gamma0=0.02, delgam=1e-7, me=1.0,
                        mi=1836.0, c_omp=10.0),
}
C = 0.45


def input_text(case, maxptl):
    return f"""<node_configuration>
sizex = 1
sizey = 1
<time>
last = 1
c = {C}
<grid>
mx0 = {case['nx']}
my0 = {case['ny']}
mz0 = {case['nz']}
<algorithm>
highorder = 0
Corr = 1.025
ntimes = 4
<boundaries>
periodicx = 1
periodicy = 1
periodicz = 1
<fields>
btheta = 0
bphi = 90
<domain>
enlarge = 0
<particles>
sigma = 0.
maxptl0 = {maxptl}
ppc0 = {case['ppc0']}
delgam = {case['delgam']}
me = {case['me']}
mi = {case['mi']}
gamma0 = {case['gamma0']}
c_omp = {case['c_omp']}
<problem>
distr_dim = 3
temperature_ratio = 1
external_fields = 0
sigma_ext = 0.
user_part_bcs = 0
"""


def run(case):
    dim = case["dim"]
    cfg = orc.make_cfg(dim, 3, case["nx"], case["ny"], case["nz"])
    mx, my, mz, _, _ = orc.extent(cfg)
    exe = os.path.join(REF_DIR, f"tristan_ref_{dim}d_dd3_f32_user_{case['setup']}")
    ncell = case["nx"] * case["ny"] * case["nz"]
    maxptl = 4 * ncell * case["ppc0"] + 10000
    empty = np.zeros((7, 0), np.float32)
    with tempfile.TemporaryDirectory() as td:
        fin = os.path.join(td, "input")
        open(fin, "w").write(input_text(case, maxptl))
        st = os.path.join(td, "state.bin")
        write_state(st, np.zeros((6, mz, my, mx), np.float32), empty, empty, np.float32)
        out = os.path.join(td, "out.bin")
        import subprocess
        subprocess.run([exe, fin, st, out, "0", "6"], check=True, cwd=td, stdout=subprocess.DEVNULL)
        _, ((a_i, _), (a_e, _)) = read_state(out, cfg, np.float32)
    return a_i, a_e


def main(names=None):
    for name, case in CASES.items():
        if names and name not in names:
            continue
        pi, pe = run(case)
        out = os.path.join(HERE, f"loader_{name}.npz")
        np.savez_compressed(out, ions=pi, electrons=pe, params=np.array(repr({**case, "c": C})))
        print(out, pi.shape, pe.shape, os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1:])
