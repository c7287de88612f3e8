"""Energy histories of the compiled reference over many main-loop steps.

A counter-streaming plasma (pair Weibel, or electron beams over cold ions for
the two-stream case) is run by the reference's own
routines (oracle/_ref/tristan_ref_*, FP32 and FP64 builds), writing every
EVERY steps 0.5*sum(E^2), 0.5*sum(B^2) and the kinetic energy of each
species.  The initial state is regenerated from its seed by `initial_state`
(numpy's PCG64 stream), so the fixture stores only the parameters and the
histories; the GPU test rebuilds the same state and compares its own history
with both.
"""
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, HERE)
import orc  # noqa: E402
from make_golden import REF_DIR, write_state  # noqa: E402

CASES = {
    # 2D: 64 x 48 cells, 8 particles per cell per species, 2nd order
    "weibel2d_ord2": dict(dim=2, order=2, nx=64, ny=48, nz=1, ppc=8, steps=1000, seed=71),
    # 3D: 16^3 cells, 4 particles per cell per species, 3rd order
    "weibel3d_ord3": dict(dim=3, order=3, nx=16, ny=16, nz=16, ppc=4, steps=500, seed=72),
    # 2D: 256^2 cells, 8 particles per cell per species, 1st order, c/wp = 10
    # (the size of the reference's CPU-runnable 2D configuration)
    "weibel2d_ord1_256": dict(dim=2, order=1, nx=256, ny=256, nz=1, ppc=8, steps=1000, seed=73, c_omp=10.0),
    # 3D: 16^3 cells, 12 particles per cell per species, 3rd order: 6144
    # particles per 8x8x4 tile, beyond the push's record space (the
    # multi-pass push)
    "weibel3d_ord3_dense": dict(dim=3, order=3, nx=16, ny=16, nz=16, ppc=12, steps=300, seed=74),
    # 2D order 2 with the 4th-order field solver (highorder = 1)
    "weibel2d_ord2_highorder": dict(dim=2, order=2, nx=64, ny=48, nz=1, ppc=8, steps=600, seed=75, highorder=1),
    # 2D order 1 with the Vay pusher (the reference built with -Dvay)
    "weibel2d_ord1_vay": dict(dim=2, order=1, nx=64, ny=48, nz=1, ppc=8, steps=600, seed=76, pusher="vay"),
    # the same two in 3D, order 3, 16^3 cells, 4 ppc per species
    "weibel3d_ord3_highorder": dict(dim=3, order=3, nx=16, ny=16, nz=16, ppc=4, steps=300, seed=77, highorder=1),
    "weibel3d_ord3_vay": dict(dim=3, order=3, nx=16, ny=16, nz=16, ppc=4, steps=300, seed=78, pusher="vay"),
    # the 3D pair Weibel precision study (512x256x256, 4 ppc per species,
    # gamma*beta = 1, theta = 0.01, c/wp = 6 cells, 3rd order) on 16^3 cells
    # for its full 2000 steps
    "weibel3d_ord3_c2": dict(dim=3, order=3, nx=16, ny=16, nz=16, ppc=4, steps=2000, seed=79,
                             gamma0=2.0 ** 0.5, gb=1.0, theta=0.01, c_omp=6.0),
    # the 3D electron-ion two-stream (192^3, 2nd order, m_i/m_e = 16,
    # c/wpe = 8, electron beams +-0.2c, theta_e = 1e-4, cold ions at rest,
    # 16 ppc per species, 500 steps) on 16^3 cells
    "twostream3d_ord2": dict(dim=3, order=2, nx=16, ny=16, nz=16, ppc=16, steps=500, seed=80, kind="twostream",
                             gamma0=0.2, gb=0.2 / (1.0 - 0.04) ** 0.5, theta=1e-4, c_omp=8.0, mi=16.0),
    # the original scheme (-Dzzag: trilinear mover, zigzag deposit, the
    # first-order ghost layout): 2D 64 x 48 and 3D 16^3
    "weibel2d_zzag": dict(dim=2, order=1, nx=64, ny=48, nz=1, ppc=8, steps=1000, seed=81, scheme="zzag"),
    "weibel3d_zzag": dict(dim=3, order=1, nx=16, ny=16, nz=16, ppc=4, steps=500, seed=82, scheme="zzag"),
}
PHYS = dict(c=0.45, corr=1.025, ntimes=4, c_omp=5.0, gamma0=0.5, me=1.0, mi=1.0, gb=0.5, theta=2.5e-3)
EVERY = 10


def input_text(case, maxptl):
    p = PHYS
    return f"""<node_configuration>
sizex = 1
sizey = 1
<time>
last = {case['steps']}
c = {p['c']}
<grid>
mx0 = {case['nx']}
my0 = {case['ny']}
mz0 = {case['nz']}
<algorithm>
highorder = {case.get('highorder', 0)}
Corr = {p['corr']}
ntimes = {p['ntimes']}
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
ppc0 = {2 * case['ppc']}
delgam = {case.get('theta', p['theta'])}
me = {case.get('me', p['me'])}
mi = {case.get('mi', p['mi'])}
gamma0 = {case.get('gamma0', p['gamma0'])}
c_omp = {case.get('c_omp', p['c_omp'])}
<problem>
distr_dim = 3
temperature_ratio = 1
external_fields = 0
sigma_ext = 0.
user_part_bcs = 0
"""


def physics(case):
    """normalisation of the reference (code/particles.F90:206-235), ppc0 = 2 ppc"""
    p = dict(PHYS)
    p["ppc0"] = 2.0 * case["ppc"]
    for k in ("c_omp", "gamma0", "gb", "theta", "me", "mi"):
        p[k] = case.get(k, PHYS[k])
    return p


def initial_state(case):
    """uniform positions; half of each species drifts at -gb, half at +gb
    along x, with a Gaussian thermal spread sqrt(theta) per component"""
    rng = np.random.default_rng(case["seed"])
    cfg = orc.make_cfg(case["dim"], case["order"], case["nx"], case["ny"], case["nz"])
    mx, my, mz, ng, ngz = orc.extent(cfg)
    h, hz = ng // 2, ngz // 2
    ncell = case["nx"] * case["ny"] * case["nz"]
    n = ncell * case["ppc"]
    p = physics(case)
    out = []
    for s in range(2):
        pos = np.empty((3, n))
        pos[0] = h + 1 + rng.random(n) * case["nx"]
        pos[1] = h + 1 + rng.random(n) * case["ny"]
        pos[2] = (hz + 1 + rng.random(n) * case["nz"]) if case["dim"] == 3 else hz + 1.5
        mom = rng.standard_normal((3, n)) * np.sqrt(p["theta"])
        if case.get("kind") == "twostream" and s == 0:
            mom[:] = 0.0                    # cold ions at rest
        else:
            mom[0, : n // 2] -= p["gb"]
            mom[0, n // 2:] += p["gb"]
        out.append(np.concatenate([pos, mom, np.ones((1, n))]).astype(np.float32))
    return cfg, out[0], out[1]


def run_history(case, prec):
    import subprocess
    cfg, pi, pe = initial_state(case)
    dt = np.float32 if prec == "f32" else np.float64
    suffix = "_vay" if case.get("pusher") == "vay" else ""
    exe = os.path.join(REF_DIR, f"tristan_ref_{case['dim']}d_dd{case['order']}_{prec}{suffix}")
    if case.get("scheme") == "zzag":
        exe = os.path.join(REF_DIR, f"tristan_ref_{case['dim']}d_{prec}_zzag")
    maxptl = 2 * (pi.shape[1] + pe.shape[1]) + 1000
    mx, my, mz, _, _ = orc.extent(cfg)
    with tempfile.TemporaryDirectory() as td:
        open(os.path.join(td, "input"), "w").write(input_text(case, maxptl))
        st = os.path.join(td, "state.bin")
        write_state(st, np.zeros((6, mz, my, mx), dt), pi, pe, dt)
        hist = os.path.join(td, "hist.txt")
        subprocess.run([exe, "input", st, os.path.join(td, "out.bin"), str(case["steps"]), "0", hist, str(EVERY)],
                       check=True, cwd=td, stdout=subprocess.DEVNULL)
        return np.loadtxt(hist)


def main(names=None):
    for name, case in CASES.items():
        if names and name not in names:
            continue
        fix = dict(params=np.array(repr({**case, **physics(case), "every": EVERY})))
        for prec in ("f32", "f64"):
            fix[f"hist_{prec}"] = run_history(case, prec)
        out = os.path.join(HERE, f"energy_{name}.npz")
        np.savez_compressed(out, **fix)
        print(out, os.path.getsize(out))


if __name__ == "__main__":
    main(sys.argv[1:])
