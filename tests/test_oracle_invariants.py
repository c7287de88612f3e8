"""Mathematical invariants of the CPU oracle (oracle/oracle.cpp), in FP64.

These pin the parts of the oracle that the reference fixtures leave open,
against what the mathematics fixes rather than against the oracle's own
formulas:

* orc_rho (the charge density with the deposit's shape) against the closed
  form of the cardinal B-splines M2/M3/M4 at the nodes, for single particles
  at dense offsets (this includes the partition of unity, sum S = 1);
* orc_rho together with orc_deposit: the discrete continuity equation
  rho(new) - rho(old) = div_b(cur) of the Esirkepov deposit
  (code/particles.F90:678-1358), per particle batch, to round-off;
* orc_filter: n binomial passes act on a single Fourier mode as the transfer
  function prod_i cos^(2n)(k_i/2) (code/filter.F90:8-221: 3x3 weights
  {4,2,1}/16 in 2D, 3x3x3 weights {8,4,2,1}/64 in 3D);
* orc_mover in a pure, uniform magnetic field: |u| is conserved and u turns
  about B by 2 atan(|b|), b = qm B / (2 c gamma) (the Boris rotation,
  code/particles_movedeposit.F90:1224-1244);
* orc_step end to end: the main loop keeps Gauss's law div_b E = F^n rho
  (F^n the filter passes applied to the current) at round-off from a
  neutral start.

A plausible slip (a node off by one, a missing half-cell shift, a wrong sign
of the current, a transposed filter weight, a dropped Boris factor) breaks
one of these.
"""
import numpy as np
import pytest

import orc


def bspline(order, t):
    """centred cardinal B-spline of degree `order` (M2, M3, M4) at t."""
    a = np.abs(np.asarray(t, np.float64))
    if order == 1:
        return np.where(a < 1.0, 1.0 - a, 0.0)
    if order == 2:
        return np.where(a <= 0.5, 0.75 - a * a, np.where(a < 1.5, 0.5 * (1.5 - a) ** 2, 0.0))
    return np.where(a <= 1.0, 2.0 / 3.0 - a * a + 0.5 * a ** 3, np.where(a < 2.0, (2.0 - a) ** 3 / 6.0, 0.0))


def periodic_nodes_weight(order, x, base, n):
    """weight of every interior node g = 0..n-1 (absolute node g + base) for a
    particle at absolute x on a periodic axis of n nodes."""
    g = np.arange(n)
    node = g + base
    d = x - node
    d = d - n * np.round(d / n)   # nearest periodic image
    return bspline(order, d)


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_rho_is_the_closed_form_bspline(dim, order):
    nx, ny, nz = 9, 8, 7
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    mx, my, mz, ng, ngz = orc.extent(cfg)
    h, hz = ng // 2, ngz // 2
    rng = np.random.default_rng(100 + 10 * dim + order)
    # offsets on a dense grid plus the knots, where the pieces meet
    offs = np.concatenate([np.linspace(0.0, 1.0, 41, endpoint=False), [0.5, 0.25, 0.75], rng.random(20)])
    for o in offs:
        x = h + 1 + 3 + o
        y = h + 1 + 5 + rng.random()
        z = (hz + 1 + 2 + rng.random()) if dim == 3 else hz + 1.5
        prt = np.array([[x], [y], [z], [0.1], [0.2], [0.3], [1.0]], np.float64)
        q = -0.37
        rho = orc.rho(cfg, prt, q)
        wx = periodic_nodes_weight(order, x, h + 1, nx)
        wy = periodic_nodes_weight(order, y, h + 1, ny)
        wz = periodic_nodes_weight(order, z, hz + 1, nz) if dim == 3 else np.array([1.0])
        ref = q * wz[:, None, None] * wy[None, :, None] * wx[None, None, :]
        np.testing.assert_allclose(rho, ref, rtol=0, atol=1e-15)
        # partition of unity: the particle's whole charge is on the grid
        assert abs(rho.sum() - q) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("order", [1, 2, 3])
def test_deposit_satisfies_continuity_with_rho(dim, order):
    """rho(new) - rho(old) = div_b(cur): the oracle deposit recomputes the old
    position as x0 = x - u c / gamma (code/particles_movedeposit.F90:1384-1388)"""
    nx, ny, nz = 12, 10, 8
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    mx, my, mz, ng, ngz = orc.extent(cfg)
    h, hz = ng // 2, ngz // 2
    rng = np.random.default_rng(7 + dim + order)
    n = 3000
    x = h + 1 + rng.random(n) * nx
    y = h + 1 + rng.random(n) * ny
    z = (hz + 1 + rng.random(n) * nz) if dim == 3 else np.full(n, hz + 1.5)
    u = rng.uniform(-3.0, 3.0, (3, n))
    ch = rng.uniform(0.5, 1.5, n)
    prt = np.ascontiguousarray(np.vstack([x, y, z, u, ch[None]]).astype(np.float64))
    q = 0.0123
    c = cfg.c
    invg = 1.0 / np.sqrt(1.0 + u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
    old = prt.copy()
    old[0] = x - u[0] * invg * c
    old[1] = y - u[1] * invg * c
    if dim == 3:
        old[2] = z - u[2] * invg * c
    rho_old = orc.rho(cfg, old, q)
    rho_new = orc.rho(cfg, prt, q)
    j3 = orc.zeros_grid(cfg, 3, np.float64)
    orc.deposit(cfg, j3, prt.copy(), q)
    orc.fold(cfg, j3)
    sl = orc.interior_slices(cfg)
    jx, jy, jz = (j3[(k,) + sl] for k in range(3))
    div = (jx - np.roll(jx, 1, axis=2)) + (jy - np.roll(jy, 1, axis=1))
    if dim == 3:
        div = div + (jz - np.roll(jz, 1, axis=0))
    res = rho_new - rho_old - div
    scale = np.abs(rho_new).max()
    assert np.abs(res).max() < 1e-12 * scale, np.abs(res).max() / scale
    # the same particles with a sign slip in the current fail the check
    res_bad = rho_new - rho_old + div
    assert np.abs(res_bad).max() > 1e-3 * scale


@pytest.mark.parametrize("dim,shape", [(2, (16, 12, 1)), (3, (12, 10, 8))])
@pytest.mark.parametrize("npasses", [1, 2, 4])
def test_filter_transfer_function(dim, shape, npasses):
    nx, ny, nz = shape
    cfg = orc.make_cfg(dim, 2, nx, ny, nz)
    ii = np.arange(nx)[None, None, :]
    jj = np.arange(ny)[None, :, None]
    kk = np.arange(nz if dim == 3 else 1)[:, None, None]
    for mode in [(1, 0, 0), (0, 1, 0), (2, 3, 1), (3, 1, 2), (nx // 2, ny // 2, 0)]:
        kx, ky, kz = (2 * np.pi * mode[0] / nx, 2 * np.pi * mode[1] / ny,
                      2 * np.pi * mode[2] / nz if dim == 3 else 0.0)
        phase = kx * ii + ky * jj + kz * kk
        interior = np.stack([np.cos(phase + 0.3 * s) for s in range(3)])
        j3 = np.ascontiguousarray(orc.fill_from_interior(cfg, interior, np.float64))
        orc.filt(cfg, j3, npasses)
        got = j3[(slice(None),) + orc.interior_slices(cfg)]
        g = (np.cos(kx / 2) ** 2 * np.cos(ky / 2) ** 2 * (np.cos(kz / 2) ** 2 if dim == 3 else 1.0)) ** npasses
        np.testing.assert_allclose(got, g * interior, rtol=0, atol=1e-13)


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 3), (3, 1), (3, 3)])
def test_boris_rotation_in_pure_b(dim, order):
    nx, ny, nz = 10, 9, 8
    cfg = orc.make_cfg(dim, order, nx, ny, nz)
    mx, my, mz, ng, ngz = orc.extent(cfg)
    h, hz = ng // 2, ngz // 2
    B = np.array([0.031, -0.047, 0.022])
    f6 = orc.zeros_grid(cfg, 6, np.float64)
    for q in range(3):
        f6[3 + q] = B[q]       # uniform, ghosts included
    rng = np.random.default_rng(31 + dim + order)
    n = 500
    u0 = rng.uniform(-2.0, 2.0, (3, n))
    pos = np.vstack([h + 1 + rng.random(n) * nx, h + 1 + rng.random(n) * ny,
                     (hz + 1 + rng.random(n) * nz) if dim == 3 else np.full(n, hz + 1.5)])
    prt = np.ascontiguousarray(np.vstack([pos, u0, np.ones((1, n))]))
    qm = -0.83
    orc.mover(cfg, f6, prt, qm)
    u1 = prt[3:6]
    # |u| is conserved
    np.testing.assert_allclose(np.linalg.norm(u1, axis=0), np.linalg.norm(u0, axis=0), rtol=1e-14)
    # the parallel component is unchanged, the perpendicular one turns by 2 atan|b|
    bh = B / np.linalg.norm(B)
    par0, par1 = bh @ u0, bh @ u1
    np.testing.assert_allclose(par1, par0, rtol=0, atol=1e-14)
    p0 = u0 - np.outer(bh, par0)
    p1 = u1 - np.outer(bh, par1)
    cosang = np.sum(p0 * p1, axis=0) / (np.linalg.norm(p0, axis=0) * np.linalg.norm(p1, axis=0))
    sinang = (bh @ np.cross(p0.T, p1.T).T) / (np.linalg.norm(p0, axis=0) * np.linalg.norm(p1, axis=0))
    ang = np.arctan2(sinang, cosang)
    gamma = np.sqrt(1.0 + np.sum(u0 * u0, axis=0))
    bmag = 0.5 * abs(qm) * np.linalg.norm(B) / (cfg.c * gamma)
    # a particle of negative q/m turns positively about B (right-handed)
    np.testing.assert_allclose(ang, -np.sign(qm) * 2.0 * np.arctan(bmag), rtol=1e-11, atol=1e-14)


def binomial(a, npasses, dim):
    """npasses periodic (1,2,1)/4 passes along every axis of a (z, y, x) grid"""
    for _ in range(npasses):
        for ax in ([2, 1, 0] if dim == 3 else [2, 1]):
            a = 0.25 * (np.roll(a, 1, axis=ax) + np.roll(a, -1, axis=ax)) + 0.5 * a
    return a


def div_b(e3, dim):
    """backward-difference divergence at the nodes of the staggered E"""
    d = (e3[0] - np.roll(e3[0], 1, axis=2)) + (e3[1] - np.roll(e3[1], 1, axis=1))
    if dim == 3:
        d = d + (e3[2] - np.roll(e3[2], 1, axis=0))
    return d


@pytest.mark.parametrize("dim,order", [(2, 1), (2, 2), (2, 3), (3, 1), (3, 2), (3, 3)])
def test_step_keeps_gauss_law(dim, order):
    """the main loop (E += filtered Esirkepov current, E += c Corr curl_b B)
    keeps div_b E - F^n rho at its initial value: from a neutral start
    (E = 0, both species at the same places) it stays at round-off"""
    nx, ny, nz = 12, 10, 8
    ntimes = 3
    cfg = orc.make_cfg(dim, order, nx, ny, nz, ntimes=ntimes)
    mx, my, mz, ng, ngz = orc.extent(cfg)
    h, hz = ng // 2, ngz // 2
    rng = np.random.default_rng(5 + dim * 3 + order)
    n = 4 * nx * ny * (nz if dim == 3 else 1)
    pos = np.vstack([h + 1 + rng.random(n) * nx, h + 1 + rng.random(n) * ny,
                     (hz + 1 + rng.random(n) * nz) if dim == 3 else np.full(n, hz + 1.5)])
    pi = np.ascontiguousarray(np.vstack([pos, rng.normal(0, 0.3, (3, n)), np.ones((1, n))]))
    pe = np.ascontiguousarray(np.vstack([pos, rng.normal(0, 0.3, (3, n)), np.ones((1, n))]))
    nm = orc.normalize(np.float64, 0.45, 4.0, 0.3, 8.0, 1.0, 1.0)
    f6 = orc.zeros_grid(cfg, 6, np.float64)
    j3 = orc.zeros_grid(cfg, 3, np.float64)
    sl = orc.interior_slices(cfg)
    for _ in range(2):
        orc.step(cfg, f6, j3, pi, pe, nm, 10)
        rho = binomial(orc.rho(cfg, pi, nm["qi"]) + orc.rho(cfg, pe, nm["qe"]), ntimes, dim)
        e3 = np.stack([f6[(q,) + sl] for q in range(3)])
        res = div_b(e3, dim) - rho
        rms = np.sqrt(np.mean(rho ** 2))
        assert rms > 1e-6 * abs(nm["qi"])          # the species have separated
        assert np.abs(res).max() < 1e-11 * rms, np.abs(res).max() / rms
    # against the unfiltered rho the check fails: the step adds F^n cur to E
    raw = orc.rho(cfg, pi, nm["qi"]) + orc.rho(cfg, pe, nm["qe"])
    assert np.abs(div_b(e3, dim) - raw).max() > 1e-3 * rms
