"""The 4th-order field solver's CPU restatement (oracle/yee42_ref.py) against
the compiled reference's FP64 build, and the identity that pins its
structure independently: the B step leaves the matching 4th-order
divergence of B unchanged."""
import ast
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import yee42_ref as y  # noqa: E402


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 3), (3, 1), (3, 3)])
def test_restatement_matches_reference_fp64(dim, order):
    fx = np.load(os.path.join(HERE, "golden", f"ref_{dim}d_ord{order}_ho.npz"))
    p = ast.literal_eval(str(fx["params"]))
    assert p["highorder"] == 1
    got = y.yee_only(fx["f_in"].astype(np.float64), p["c"], p["corr"], dim)
    ref = fx["yee_f64"]
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-13, err
    # and the 4th-order step differs from a 2nd-order one (a dropped c2 term
    # would not be caught by a tolerance this loose otherwise)
    assert np.linalg.norm(got - fx["f_in"]) > 1e-3 * np.linalg.norm(ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_b_step_keeps_div_b(dim):
    rng = np.random.default_rng(3 + dim)
    shape = (3, 6 if dim == 3 else 1, 10, 12)
    e = rng.standard_normal(shape)
    b = rng.standard_normal(shape)
    d0 = y.div4_b(b, dim)
    y.bhalf42(e, b, 0.45, 1.025, dim)
    d1 = y.div4_b(b, dim)
    assert np.abs(d1 - d0).max() < 1e-12


def test_vacuum_mode_dispersion():
    """one cycle (B half, B half, E) on a vacuum mode e^{ikx} (E_y, B_z) is
    the matrix [[1 - a^2, -i a], [-i a, 1]] in (E_y, B_z e^{-ik/2}) with
    a = c corr K, K = (9/8) 2 sin(k/2) - (1/24) 2 sin(3k/2): trace 2 - a^2,
    determinant 1, so the wave turns by w with sin(w/2) = a/2"""
    nx, c, corr = 32, 0.45, 1.025
    x = np.arange(nx)
    for m in (1, 3, 7, 12):
        k = 2 * np.pi * m / nx
        mode = np.exp(1j * k * x)
        cols = []
        for comp in (1, 2):   # unit E_y, unit B_z (scaled by e^{ik/2})
            e = np.zeros((3, 1, 1, nx), complex)
            b = np.zeros((3, 1, 1, nx), complex)
            if comp == 1:
                e[1, 0, 0] = mode
            else:
                b[2, 0, 0] = mode * np.exp(1j * k / 2)
            y.bhalf42(e, b, c, corr, 2)
            y.bhalf42(e, b, c, corr, 2)
            y.efull42(e, b, c, corr, 2)
            u = np.mean(e[1, 0, 0] * np.conj(mode))
            v = np.mean(b[2, 0, 0] * np.conj(mode)) * np.exp(-1j * k / 2)
            cols.append((u, v))
        mat = np.array(cols).T
        a = c * corr * (9.0 / 8.0 * 2 * np.sin(k / 2) - 1.0 / 24.0 * 2 * np.sin(3 * k / 2))
        assert np.allclose(mat, [[1 - a * a, -1j * a], [-1j * a, 1]], atol=1e-12), (m, mat)
