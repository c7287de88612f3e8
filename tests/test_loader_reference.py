"""The library's host loader (tmpg_inject_plasma_region) against the
reference's own initial plasma (tests/golden/loader_*.npz, written by
tests/golden/make_loader_golden.py from the compiled reference's
init_particle_distribution_user of the Weibel, two-stream and
plasma-oscillation user modules).

The same calls as the user modules, seeded like the reference (123457 +
rank 0, code/communications.F90:224-231):
* user/user_weibel.F90:252-292: two regions over the whole interior,
  drifting -gamma0 and +gamma0, both species;
* user/user_twostream.F90:236-305: the same from 3 to mx0-2 with
  pcosthmult = 0, then the ions dropped;
* user/user_plasmaosc.F90:236-305: one region at rest with pcosthmult = 0,
  the electrons kicked by beta sin(2 pi x / (mx0 - 5)), the ions dropped.
The reference sorts its particles by cell afterwards (reorder_particles), so
both sets are compared in one sort order.  Positions must be bit-equal;
momenta come through cos, sin, exp and sqrt of two math libraries and agree
to 2e-6 of their scale.  Runs on the CPU (the loader is host code)."""
import ast
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(os.path.dirname(HERE), "tristan-mp-pu-master_densdecomp_amd", "lib", "libtmpg.so")

NAMES = ["weibel2d", "weibel3d", "twostream3d", "plasmaosc2d"]


def f32(v):
    return float(np.float32(v))


def generate(tmpg, p):
    """the user module's calls for one fixture (the reference's grid sizes
    mx0 etc. include the ghost cells, code/fields.F90:186)"""
    dim, order = p["dim"], 3
    ng = 7
    ngz = 5 if dim == 2 else ng
    mx0, my0 = p["nx"] + ng, p["ny"] + ng
    mz0 = p["nz"] + ngz if dim == 3 else 1
    g0 = np.float32(p["gamma0"])
    if g0 < 1:
        g0 = np.float32(np.sqrt(np.float32(1.0) / (np.float32(1.0) - g0 * g0)))
    delgam = f32(p["delgam"])
    delgam_e = f32(np.float32(np.float32(delgam * np.float32(p["mi"])) / np.float32(p["me"])) * np.float32(1.0))
    seed = 123457.0
    ions, lecs = [], []

    def region(x1, x2, y1, y2, z1, z2, ppc, gd, pcm):
        nonlocal seed
        r = tmpg.PlasmaRegion(dim=dim, order=order, nx=p["nx"], ny=p["ny"], nz=p["nz"], c=p["c"],
                              x1=x1, x2=x2, y1=y1, y2=y2, z1=z1, z2=z2, ppc=ppc, gamma_drift=gd,
                              delgam_i=delgam, delgam_e=delgam_e, weight=1.0, direction=1, pcosthmult=pcm,
                              sigma=0.0, upsamp_e=1)
        a, b, seed = tmpg.inject_plasma_region(r, seed)
        ions.append(a)
        lecs.append(b)

    ppc0 = p["ppc0"]
    if p["setup"] == "weibel":
        h = ng // 2
        for gd in (-g0, g0):
            region(h + 1, mx0 - h, h + 1, my0 - h, 3.0, mz0 - 2.0, ppc0 / 2.0, gd, 1)
    elif p["setup"] == "twostream":
        for gd in (-g0, g0):
            region(3.0, mx0 - 2.0, 3.0, my0 - 2.0, 3.0, mz0 - 2.0, ppc0 / 2.0, gd, 0)
        ions = []
    else:
        region(3.0, mx0 - 2.0, 3.0, my0 - 2.0, 3.0, mz0 - 2.0, float(ppc0), 0.0, 0)
        e = lecs[0]
        beta = np.float32(np.sqrt(np.float32(1.0) - np.float32(1.0) / (g0 * g0)))
        pi = np.float64(np.float32(3.1415927))
        kick = np.float64(beta) * np.sin(2.0 * pi / np.float64(np.float32(mx0 - 5.0)) * e[0].astype(np.float64))
        e[3] = (e[3].astype(np.float64) + kick).astype(np.float32)
        ions = []
    cat = lambda L: np.concatenate(L, axis=1) if L else np.zeros((7, 0), np.float32)
    return cat(ions), cat(lecs)


def by_position(a):
    return a[:, np.lexsort((a[2], a[1], a[0]))]


@pytest.mark.skipif(not os.path.exists(LIB), reason="libtmpg.so not built")
@pytest.mark.parametrize("name", NAMES)
def test_loader_reproduces_reference_plasma(name):
    import tmpg
    fx = np.load(os.path.join(HERE, "golden", f"loader_{name}.npz"))
    p = ast.literal_eval(str(fx["params"]))
    gi, ge = generate(tmpg, p)
    for got, ref, what in ((gi, fx["ions"], "ions"), (ge, fx["electrons"], "electrons")):
        assert got.shape == ref.shape, (what, got.shape, ref.shape)
        if ref.shape[1] == 0:
            continue
        g, r = by_position(got), by_position(ref)
        np.testing.assert_array_equal(g[:3], r[:3], err_msg=f"{name} {what} positions")
        np.testing.assert_array_equal(g[6], r[6])
        scale = np.abs(r[3:6]).max()
        err = np.abs(g[3:6] - r[3:6]).max() / scale
        print(f"{name} {what}: {ref.shape[1]} particles, momenta max error {err:.2e} of {scale:.3e}")
        assert err < 2e-6, (name, what, err)
