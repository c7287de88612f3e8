"""The original scheme on the GPU (scheme = 1, the reference's -Dzzag build):
trilinear mover of the staggered fields (code/particles_movedeposit.F90:
98-347) and the zigzag deposit (code/particles.F90:550-669).

* mover alone (tmpg_move) and one main-loop step against the compiled
  reference's -Dzzag FP64 build (tests/golden/ref_{2,3}d_zzag.npz), at the
  bars of the Esirkepov orders;
* fused push+deposit on larger random states against the FP64 restatement
  oracle/zzag_ref.py (itself pinned to that build to 1e-12), including the
  2D z offset, which the reference tracks modulo one cell and whose relay
  point splits the z flux between the two halves of a move;
* the split move/deposit path agrees with the fused one.
"""
import ast
import os
import sys

import numpy as np
import pytest

import orc
from gpu_util import make_context, norm64, periodic_diff, random_state, rel, sorted_by_id

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import zzag_ref as zz  # noqa: E402

pytestmark = pytest.mark.gpu

TOL_FIELD = 2e-6
TOL_MOM = 1e-6


def _interior(ctx, a):
    h, hz = ctx.ng // 2, ctx.ngz // 2
    if ctx.dim == 3:
        return a[..., hz:hz + ctx.nz, h:h + ctx.ny, h:h + ctx.nx]
    return a[..., :, h:h + ctx.ny, h:h + ctx.nx]


def _check(ctx, got, ids, ref, what, zoff=False):
    g = sorted_by_id(got, ids)
    e_u = rel(g[3:6], ref[3:6])
    assert e_u < TOL_MOM, f"{what} momenta {e_u}"
    h, hz = ctx.ng // 2, ctx.ngz // 2
    err = max(np.abs(periodic_diff(g[0], ref[0], h + 1, ctx.nx)).max(),
              np.abs(periodic_diff(g[1], ref[1], h + 1, ctx.ny)).max())
    if ctx.dim == 3:
        err = max(err, np.abs(periodic_diff(g[2], ref[2], hz + 1, ctx.nz)).max())
    elif zoff:
        # 2D: z modulo one cell
        err = max(err, np.abs(periodic_diff(g[2], ref[2], hz + 1, 1)).max())
    assert err < 8e-6, f"{what} positions {err}"


def _ctx_from_fixture(tmpg, dim):
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"ref_{dim}d_zzag.npz"))
    p = ast.literal_eval(str(g["params"]))
    ctx = make_context(tmpg, dim, 1, p["nx"], p["ny"], p["nz"], p, ntimes=p["ntimes"], scheme=1)
    cfg = orc.make_cfg(dim, 1, p["nx"], p["ny"], p["nz"])
    ctx.set_fields(orc.fill_from_interior(cfg, g["f_in"], np.float32)[:6])
    ctx.set_particles(0, g["pi_in"])
    ctx.set_particles(1, g["pe_in"])
    return ctx, g, p


@pytest.mark.parametrize("dim", [2, 3])
def test_zigzag_mover_vs_reference_fp64(tmpg, dim):
    ctx, g, p = _ctx_from_fixture(tmpg, dim)
    ctx.move()
    for s, key in ((0, "move_pi_f64"), (1, "move_pe_f64")):
        got, ids = ctx.get_particles(s, with_ids=True)
        _check(ctx, got, ids, g[key], f"zigzag mover species {s}", zoff=True)
    ctx.deposit()
    ctx.close()


def _oracle_step_zz(g, p, dim, exact_start):
    """one main-loop step of the -Dzzag build (code/tristanmainloop.F90:
    107-274) in FP64: the field routines of oracle/oracle.cpp, the mover and
    deposit of oracle/zzag_ref.py; the deposit's start point either
    recomputed from the end point like the reference or the exact start"""
    cfg = orc.make_cfg(dim, 1, p["nx"], p["ny"], p["nz"], ntimes=p["ntimes"])
    nm = norm64(p)
    f6 = np.ascontiguousarray(orc.fill_from_interior(cfg, g["f_in"].astype(np.float64), np.float64))

    def fill(first):
        blk = np.ascontiguousarray(f6[first:first + 3])
        orc.ghost_fill(cfg, blk)
        f6[first:first + 3] = blk

    fill(3); fill(0)
    orc.bhalf(cfg, f6)
    fill(3)
    pi, pe = g["pi_in"].astype(np.float64), g["pe_in"].astype(np.float64)
    starts = [pi[:3].copy(), pe[:3].copy()]
    zz.mover(f6, pi, nm["qmi"], p["c"], dim)
    zz.mover(f6, pe, nm["qme"], p["c"], dim)
    orc.bhalf(cfg, f6)
    fill(3)
    orc.efull(cfg, f6)
    fill(0)
    j3 = orc.zeros_grid(cfg, 3, np.float64)
    for prt, q, st in ((pi, nm["qi"], starts[0]), (pe, nm["qe"], starts[1])):
        zz.zigzag(j3, prt, q, p["c"], dim, start=st if exact_start else None)
    orc.fold(cfg, j3)
    orc.filt(cfg, j3, p["ntimes"])
    f6[:3] += j3
    sl = (slice(None),) + orc.interior_slices(cfg)
    return np.concatenate([f6[sl], j3[sl]])


@pytest.mark.parametrize("dim", [2, 3])
def test_zigzag_step_vs_reference_fp64(tmpg, dim):
    """One step.  A sixteenth of the fixture's particles start exactly on a
    cell face (make_golden.make_particles).  There the zigzag split jumps
    between its crossing and non-crossing forms, and the reference, which
    recomputes the start as x - u c / gamma, lands on either side of the face
    by round-off; the library keeps the exact start.  B and the particles are
    compared with the reference; J and E with the reference's step restated
    in FP64 with the exact start, after checking that this restatement with
    the recomputed start reproduces the reference."""
    ctx, g, p = _ctx_from_fixture(tmpg, dim)
    ctx.step(1)
    f = _interior(ctx, ctx.get_fields())
    j = _interior(ctx, ctx.get_currents())
    ref = g["step_f64"]
    recomputed = _oracle_step_zz(g, p, dim, exact_start=False)
    exact = _oracle_step_zz(g, p, dim, exact_start=True)
    assert rel(recomputed, ref) < 1e-12
    e_b = rel(f[3:6], ref[3:6])
    e_e, e_j = rel(f[:3], exact[:3]), rel(j, exact[6:9])
    print(f"zigzag dim{dim}: B {e_b:.2e} vs reference; E {e_e:.2e} J {e_j:.2e} vs the exact start "
          f"(the reference's J is {rel(ref[6:9], exact[6:9]):.2e} from it)")
    assert e_e < TOL_FIELD and e_b < TOL_FIELD and e_j < TOL_FIELD
    for s, key in ((0, "step_pi_f64"), (1, "step_pe_f64")):
        got, ids = ctx.get_particles(s, with_ids=True)
        _check(ctx, got, ids, g[key], f"zigzag step species {s}", zoff=True)
    ctx.close()


@pytest.mark.parametrize("dim,shape,npart", [(2, (48, 40, 1), 8000), (3, (16, 12, 10), 6000),
                                             (3, (8, 8, 4), 2600)])
def test_zigzag_push_deposit_vs_oracle_fp64(tmpg, dim, shape, npart):
    nx, ny, nz = shape
    p = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.5, ppc0=4.0, me=1.0, mi=16.0)
    f, pi, pe = random_state(dim, 1, nx, ny, nz, npart, seed=17 + dim + nx)
    cfg = orc.make_cfg(dim, 1, nx, ny, nz)
    mx, my, mz, ng, ngz = orc.extent(cfg)
    if dim == 2:
        # z anywhere in its cell, so some moves cross a z cell boundary
        rng = np.random.default_rng(5)
        for a in (pi, pe):
            a[2] = (ngz // 2 + 1 + rng.random(a.shape[1])).astype(np.float32)
    f6 = orc.fill_from_interior(cfg, f.astype(np.float64), np.float64)
    nm = norm64(p)
    ctx = make_context(tmpg, dim, 1, nx, ny, nz, p, scheme=1)
    ctx.set_fields(orc.fill_from_interior(cfg, f, np.float32)[:6])
    ctx.set_particles(0, pi)
    ctx.set_particles(1, pe)
    ctx.reset_currents()
    ctx.push_deposit()
    ri, re = pi.astype(np.float64), pe.astype(np.float64)
    zz.mover(f6[:6], ri, nm["qmi"], p["c"], dim)
    zz.mover(f6[:6], re, nm["qme"], p["c"], dim)
    j3 = orc.zeros_grid(cfg, 3, np.float64)
    zz.zigzag(j3, ri, nm["qi"], p["c"], dim)
    zz.zigzag(j3, re, nm["qe"], p["c"], dim)
    orc.fold(cfg, j3)
    sl = orc.interior_slices(cfg)
    jg = _interior(ctx, ctx.get_currents())
    e_j = rel(jg, j3[(slice(None),) + sl])
    print(f"zigzag push+deposit dim{dim} {shape}: J {e_j:.2e}")
    assert e_j < TOL_FIELD
    for s, ref in ((0, ri), (1, re)):
        got, ids = ctx.get_particles(s, with_ids=True)
        _check(ctx, got, ids, ref, f"species {s}", zoff=True)
    ctx.close()


@pytest.mark.parametrize("dim", [2, 3])
def test_zigzag_split_equals_fused(tmpg, dim):
    nx, ny, nz = (16, 12, 10) if dim == 3 else (32, 24, 1)
    p = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.5, ppc0=4.0, me=1.0, mi=4.0)
    f, pi, pe = random_state(dim, 1, nx, ny, nz, 3000, seed=3)
    cfg = orc.make_cfg(dim, 1, nx, ny, nz)
    f32 = orc.fill_from_interior(cfg, f, np.float32)[:6]
    outs = []
    for split in (False, True):
        ctx = make_context(tmpg, dim, 1, nx, ny, nz, p, scheme=1)
        ctx.set_fields(f32)
        ctx.set_particles(0, pi)
        ctx.set_particles(1, pe)
        ctx.reset_currents()
        if split:
            ctx.move()
            ctx.deposit()
        else:
            ctx.push_deposit()
        a0, i0 = ctx.get_particles(0, with_ids=True)
        a1, i1 = ctx.get_particles(1, with_ids=True)
        outs.append((sorted_by_id(a0, i0), sorted_by_id(a1, i1), ctx.get_currents()))
        ctx.close()
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
    assert rel(outs[1][2], outs[0][2]) < 1e-6


def test_zigzag_rejects_other_orders_and_vay(tmpg):
    p = dict(c=0.45, corr=1.025, c_omp=4.0, gamma0=0.5, ppc0=4.0, me=1.0, mi=4.0)
    for kw in (dict(order=2), dict(order=1, pusher=1)):
        with pytest.raises(tmpg.TmpgError):
            make_context(tmpg, 2, kw["order"], 16, 16, 1, p, scheme=1, pusher=kw.get("pusher", 0))
