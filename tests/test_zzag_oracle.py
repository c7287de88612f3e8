"""The restatement of the reference's original scheme (-Dzzag: trilinear
mover, zigzag deposit; oracle/zzag_ref.py) against the compiled reference's
-Dzzag FP64 build (tests/golden/ref_{2,3}d_zzag.npz, written by
tests/golden/make_golden.py zzag), and the deposit's charge conservation."""
import ast
import os
import sys

import numpy as np
import pytest

import orc

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "oracle"))
import zzag_ref as zz  # noqa: E402


def _load(dim):
    g = np.load(os.path.join(HERE, "golden", f"ref_{dim}d_zzag.npz"))
    p = ast.literal_eval(str(g["params"]))
    cfg = orc.make_cfg(dim, 1, p["nx"], p["ny"], p["nz"], ntimes=p["ntimes"])
    nm = orc.normalize(np.float64, p["c"], p["c_omp"], p["gamma0"], p["ppc0"], p["me"], p["mi"])
    f6 = orc.fill_from_interior(cfg, g["f_in"].astype(np.float64), np.float64)
    return g, p, cfg, nm, f6


@pytest.mark.parametrize("dim", [2, 3])
def test_mover_matches_reference_fp64(dim):
    g, p, cfg, nm, f6 = _load(dim)
    for key, qm in (("pi", nm["qmi"]), ("pe", nm["qme"])):
        prt = g[f"{key}_in"].astype(np.float64)
        zz.mover(f6[:6], prt, qm, p["c"], dim)
        ref = g[f"move_{key}_f64"]
        np.testing.assert_allclose(prt[3:6], ref[3:6], rtol=0, atol=1e-12 * np.abs(ref[3:6]).max())
        np.testing.assert_allclose(prt[:dim], ref[:dim], rtol=0, atol=1e-12)


@pytest.mark.parametrize("dim", [2, 3])
def test_deposit_matches_reference_step_fp64(dim):
    """the step's filtered current from the particles the step ended with
    (their start points follow from their momenta)"""
    g, p, cfg, nm, f6 = _load(dim)
    j3 = orc.zeros_grid(cfg, 3, np.float64)
    for key, q in (("pi", nm["qi"]), ("pe", nm["qe"])):
        zz.zigzag(j3, g[f"step_{key}_f64"].astype(np.float64), q, p["c"], dim)
    orc.fold(cfg, j3)
    orc.filt(cfg, j3, p["ntimes"])
    got = j3[(slice(None),) + orc.interior_slices(cfg)]
    ref = g["step_f64"][6:9]
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def _cic_rho(shape, x, y, z, q, dim):
    """area-weighted charge at the integer nodes (the density whose change
    the zigzag current balances)"""
    rho = np.zeros(shape)
    i, j, k = np.trunc(x).astype(int), np.trunc(y).astype(int), np.trunc(z).astype(int)
    dx, dy, dz = x - i, y - j, z - k
    for a, wa in ((0, 1 - dx), (1, dx)):
        for b, wb in ((0, 1 - dy), (1, dy)):
            if dim == 3:
                for c, wc in ((0, 1 - dz), (1, dz)):
                    np.add.at(rho, (k - 1 + c, j - 1 + b, i - 1 + a), q * wa * wb * wc)
            else:
                np.add.at(rho, (0 * k, j - 1 + b, i - 1 + a), q * wa * wb)
    return rho


@pytest.mark.parametrize("dim,planar", [(2, True), (3, True), (3, False)])
def test_zigzag_conserves_charge(dim, planar):
    """backward-difference divergence of the deposited current = rho(end) -
    rho(start) node by node (the current carries the sign of -q dx), for
    moves within and across cells.  Exact in 2D and for 3D moves in a
    coordinate plane; a general 3D move leaves the residual of the midpoint
    weights, (1/12) dy dz per segment and component, which the reference's
    form of the scheme does not carry."""
    rng = np.random.default_rng(3 + dim)
    n, c = 400, 0.45
    shape = (12, 10, 9) if dim == 3 else (1, 10, 9)
    lo = 3.0
    x2 = lo + rng.random(n) * 4
    y2 = lo + rng.random(n) * 4
    z2 = lo + rng.random(n) * 4 if dim == 3 else np.full(n, 1.5)
    mom = rng.uniform(-2, 2, (3, n))
    if planar:
        mom[2] = 0.0
    prt = np.stack([x2, y2, z2, mom[0], mom[1], mom[2], np.ones(n)])
    j3 = np.zeros((3,) + shape)
    zz.zigzag(j3, prt, 1.0, c, dim)
    ig = 1.0 / np.sqrt(1.0 + (mom ** 2).sum(0))
    x1, y1, z1 = x2 - mom[0] * ig * c, y2 - mom[1] * ig * c, z2 - mom[2] * ig * c
    drho = _cic_rho(shape, x2, y2, z2, 1.0, dim) - _cic_rho(shape, x1, y1, z1, 1.0, dim)
    div = (j3[0] - np.roll(j3[0], 1, axis=2)) + (j3[1] - np.roll(j3[1], 1, axis=1))
    if dim == 3:
        div = div + (j3[2] - np.roll(j3[2], 1, axis=0))
    assert np.abs(drho).max() > 0.1            # the moves did change the density
    if planar:
        np.testing.assert_allclose(div, drho, rtol=0, atol=1e-12)
    else:
        # per segment at most |dy dz| / 12 <= 0.45^2 / 12 on each of its nodes
        res = np.abs(div - drho).max()
        assert 1e-6 < res < 0.05, res
