"""The CPU oracle against the golden fixtures made by the compiled reference.

The fixtures (tests/golden/ref_*.npz, written by tests/golden/make_golden.py)
hold one main-loop step and one deposit+fold of the reference's own Fortran
routines, in FP32 and FP64, for 2D/3D and shape orders 1-3.

Orders 2 and 3 must agree bit for bit: the oracle restates the reference's
arithmetic in the same evaluation order.  Order 1 differs only by the
round-off residual that the reference's uninitialised prefix accumulators
carry from one particle to the next (code/particles.F90:691-732; the
reference build compiles that file with -fno-automatic so they start at 0):
FP64 agrees to 1e-12, FP32 to 2e-5 norm-wise.
"""
import ast
import os

import numpy as np
import pytest

import orc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rel(a, b):
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def load(dim, order):
    fx = np.load(os.path.join(GOLD, f"ref_{dim}d_ord{order}.npz"))
    p = ast.literal_eval(str(fx["params"]))
    cfg = orc.make_cfg(dim, order, p["nx"], p["ny"], p["nz"], ntimes=p["ntimes"], c=p["c"], corr=p["corr"])
    return fx, p, cfg


CASES = [(d, o) for d in (2, 3) for o in (1, 2, 3)]


@pytest.mark.parametrize("dim,order", CASES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_step_matches_reference(dim, order, prec):
    fx, p, cfg = load(dim, order)
    dt = np.float32 if prec == "f32" else np.float64
    nm = orc.normalize(dt, p["c"], p["c_omp"], p["gamma0"], p["ppc0"], p["me"], p["mi"])
    f6 = orc.fill_from_interior(cfg, fx["f_in"].astype(dt), dt)
    j3 = orc.zeros_grid(cfg, 3, dt)
    pi = fx["pi_in"].astype(dt).copy()
    pe = fx["pe_in"].astype(dt).copy()
    orc.step(cfg, f6, j3, pi, pe, nm, 1)
    sl = orc.interior_slices(cfg)
    got = np.concatenate([f6, j3])[(slice(None),) + sl]
    ref = fx[f"step_{prec}"]
    if order in (2, 3):
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(pi, fx[f"step_pi_{prec}"])
        np.testing.assert_array_equal(pe, fx[f"step_pe_{prec}"])
    else:
        tol = 1e-12 if prec == "f64" else 2e-5
        assert rel(got[:6], ref[:6]) < tol
        assert rel(got[6:], ref[6:]) < tol
        np.testing.assert_array_equal(pi, fx[f"step_pi_{prec}"])
        np.testing.assert_array_equal(pe, fx[f"step_pe_{prec}"])


@pytest.mark.parametrize("dim,order", CASES)
@pytest.mark.parametrize("prec", ["f32", "f64"])
def test_deposit_matches_reference(dim, order, prec):
    fx, p, cfg = load(dim, order)
    dt = np.float32 if prec == "f32" else np.float64
    nm = orc.normalize(dt, p["c"], p["c_omp"], p["gamma0"], p["ppc0"], p["me"], p["mi"])
    j3 = orc.zeros_grid(cfg, 3, dt)
    pi = fx["pi_in"].astype(dt).copy()
    pe = fx["pe_in"].astype(dt).copy()
    orc.deposit(cfg, j3, pi, nm["qi"])
    orc.deposit(cfg, j3, pe, nm["qe"])
    orc.fold(cfg, j3)
    got = j3[(slice(None),) + orc.interior_slices(cfg)]
    ref = fx[f"dep_{prec}"]
    if order in (2, 3):
        np.testing.assert_array_equal(got, ref)
    else:
        assert rel(got, ref) < (1e-12 if prec == "f64" else 2e-5)


def test_fixture_mismatch_is_detected():
    """A planted mistake in the restatement (a dropped 1/3 term) must fail."""
    fx, p, cfg = load(3, 3)
    dt = np.float64
    nm = orc.normalize(dt, p["c"], p["c_omp"], p["gamma0"], p["ppc0"], p["me"], p["mi"])
    j3 = orc.zeros_grid(cfg, 3, dt)
    pi = fx["pi_in"].astype(dt).copy()
    orc.deposit(cfg, j3, pi, nm["qi"])
    # perturb the way a transposed operand would: swap x/y currents
    bad = j3[[1, 0, 2]]
    pe = fx["pe_in"].astype(dt).copy()
    orc.deposit(cfg, bad, pe, nm["qe"])
    orc.fold(cfg, bad)
    assert rel(bad[(slice(None),) + orc.interior_slices(cfg)], fx["dep_f64"]) > 1e-3
