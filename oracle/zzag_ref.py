"""CPU restatement of the reference's original scheme, the -Dzzag build (test
infrastructure).

mover (code/particles_movedeposit.F90:98-347): trilinear field interpolation
with the half-cell staggering folded in, then the Boris push and the
position advance; zigzag (code/particles.F90:550-669, called from
code/particles_movedeposit.F90:1392-1393): the charge-conserving first-order
deposit of Umeda et al. 2003, which splits a move at the relay point
(xr, yr, zr) and deposits each half as a straight segment.

Arrays are the padded module arrays (6 or 3, mz, my, mx) in float64,
indexed by the reference's absolute 1-based cell indices minus one; the
ghost cells must hold the periodic images (nghost = 5, code/fields.F90:
166-169).  Particles are (7, n): x, y, z, u, v, w, ch.

Pinned by the compiled reference's -Dzzag FP64 build (tests/golden/
ref_{2,3}d_zzag.npz: the mover alone and a full step), by charge
conservation of the deposit (the discrete divergence of the deposited
current equals the change of the CIC charge density), and by the mover
reproducing a uniform field exactly (tests/test_zzag_oracle.py).
"""
import numpy as np


def _at(f, i, j, k, di, dj, dk, dim):
    """f at absolute 1-based cell (i+di, j+dj, k+dk); in 2D the z offset is
    dropped (the reference's iz = 0)"""
    if dim == 2:
        dk = 0
    return f[k - 1 + dk, j - 1 + dj, i - 1 + di]


def gather(f6, x, y, z, dim):
    """(ex0, ey0, ez0, bx0, by0, bz0) before the qm scalings: the E values
    carry a factor 4 and the B values a factor 8 of the interpolated field"""
    ex, ey, ez, bx, by, bz = f6
    i = np.trunc(x).astype(np.int64)
    j = np.trunc(y).astype(np.int64)
    dx, dy = x - i, y - j
    if dim == 3:
        k = np.trunc(z).astype(np.int64)
        dz = z - k
    else:
        k = np.ones_like(i)
        dz = np.zeros_like(x)

    def a(F, di, dj, dk):
        return _at(F, i, j, k, di, dj, dk, dim)

    # a field staggered along one axis: the average of the two half-cell
    # neighbours plus the offset times their difference across two cells
    def sx(F, dj, dk):
        return a(F, 0, dj, dk) + a(F, -1, dj, dk) + dx * (a(F, 1, dj, dk) - a(F, -1, dj, dk))

    def sy(F, di, dk):
        return a(F, di, 0, dk) + a(F, di, -1, dk) + dy * (a(F, di, 1, dk) - a(F, di, -1, dk))

    def sz(F, di, dj):
        return a(F, di, dj, 0) + a(F, di, dj, -1) + dz * (a(F, di, dj, 1) - a(F, di, dj, -1))

    def lerp(p, q, t):
        return p + t * (q - p)

    e_x = lerp(lerp(sx(ex, 0, 0), sx(ex, 1, 0), dy), lerp(sx(ex, 0, 1), sx(ex, 1, 1), dy), dz)
    e_y = lerp(lerp(sy(ey, 0, 0), sy(ey, 0, 1), dz), lerp(sy(ey, 1, 0), sy(ey, 1, 1), dz), dx)
    e_z = lerp(lerp(sz(ez, 0, 0), sz(ez, 1, 0), dx), lerp(sz(ez, 0, 1), sz(ez, 1, 1), dx), dy)

    # B is staggered along two axes: z-staggered Bx sums over y as well
    def bxz(di, dj):
        return sz(bx, di, dj)

    f = bxz(0, 0) + bxz(0, -1) + dy * (bxz(0, 1) - bxz(0, -1))
    g = bxz(1, 0) + bxz(1, -1) + dy * (bxz(1, 1) - bxz(1, -1))
    b_x = lerp(f, g, dx)

    def byx(dj, dk):
        return sx(by, dj, dk)

    f = byx(0, 0) + byx(0, -1) + dz * (byx(0, 1) - byx(0, -1))
    g = byx(1, 0) + byx(1, -1) + dz * (byx(1, 1) - byx(1, -1))
    b_y = lerp(f, g, dy)

    def bzy(di, dk):
        return sy(bz, di, dk)

    f = bzy(0, 0) + bzy(-1, 0) + dx * (bzy(1, 0) - bzy(-1, 0))
    g = bzy(0, 1) + bzy(-1, 1) + dx * (bzy(1, 1) - bzy(-1, 1))
    b_z = lerp(f, g, dz)
    return e_x, e_y, e_z, b_x, b_y, b_z


def mover(f6, prt, qm, c, dim):
    """one push of prt (7, n) float64 in place: Boris, then x += v"""
    x, y, z, u, v, w = prt[:6]
    e_x, e_y, e_z, b_x, b_y, b_z = gather(f6, x, y, z, dim)
    se, sb = 0.25 * qm, 0.125 * qm / c
    ex0, ey0, ez0 = e_x * se, e_y * se, e_z * se
    bx0, by0, bz0 = b_x * sb, b_y * sb, b_z * sb
    u0, v0, w0 = c * u + ex0, c * v + ey0, c * w + ez0
    g = c / np.sqrt(c * c + u0 * u0 + v0 * v0 + w0 * w0)
    bx0, by0, bz0 = g * bx0, g * by0, g * bz0
    f = 2.0 / (1.0 + bx0 * bx0 + by0 * by0 + bz0 * bz0)
    u1 = (u0 + v0 * bz0 - w0 * by0) * f
    v1 = (v0 + w0 * bx0 - u0 * bz0) * f
    w1 = (w0 + u0 * by0 - v0 * bx0) * f
    u0 = u0 + v1 * bz0 - w1 * by0 + ex0
    v0 = v0 + w1 * bx0 - u1 * bz0 + ey0
    w0 = w0 + u1 * by0 - v1 * bx0 + ez0
    prt[3], prt[4], prt[5] = u0 / c, v0 / c, w0 / c
    g = c / np.sqrt(c * c + u0 * u0 + v0 * v0 + w0 * w0)
    prt[0] = x + prt[3] * g * c
    prt[1] = y + prt[4] * g * c
    prt[2] = z + prt[5] * g * c
    return prt


def zigzag(j3, prt, q, c, dim, start=None):
    """deposit of the moves that ended at prt (7, n) float64 into j3 (3, mz,
    my, mx): the start point is x - u c / gamma (code/particles_movedeposit.
    F90:1385-1388), or the given start (3, n) -- the same point in exact
    arithmetic, without the round-off of the recomputation -- and the charge
    q * ch"""
    x2, y2, z2, u, v, w, ch = prt
    if start is None:
        ig = 1.0 / np.sqrt(1.0 + u * u + v * v + w * w)
        x1, y1, z1 = x2 - u * ig * c, y2 - v * ig * c, z2 - w * ig * c
    else:
        x1, y1, z1 = start
    qq = q * ch
    i1, i2 = np.trunc(x1).astype(np.int64), np.trunc(x2).astype(np.int64)
    j1, j2 = np.trunc(y1).astype(np.int64), np.trunc(y2).astype(np.int64)
    k1, k2 = np.trunc(z1).astype(np.int64), np.trunc(z2).astype(np.int64)

    # the relay point: the midpoint, clamped to the cell boundary the move
    # crosses (or left inside the cell when it crosses none)
    def relay(a1, a2, n1, n2):
        return np.minimum(np.minimum(n1, n2) + 1.0, np.maximum(np.maximum(n1, n2), 0.5 * (a1 + a2)))

    xr, yr, zr = relay(x1, x2, i1, i2), relay(y1, y2, j1, j2), relay(z1, z2, k1, k2)
    # flux of each straight half, and the cell-relative centre of the half
    fx1, fy1, fz1 = -qq * (xr - x1), -qq * (yr - y1), -qq * (zr - z1)
    fx2, fy2, fz2 = -qq * (x2 - xr), -qq * (y2 - yr), -qq * (z2 - zr)
    wx1, wy1 = 0.5 * (x1 + xr) - i1, 0.5 * (y1 + yr) - j1
    wx2, wy2 = 0.5 * (x2 + xr) - i2, 0.5 * (y2 + yr) - j2
    if dim == 3:
        wz1, wz2 = 0.5 * (z1 + zr) - k1, 0.5 * (z2 + zr) - k2
    else:
        k1 = np.ones_like(k1)
        k2 = k1
        wz1 = wz2 = np.zeros_like(x1)

    def add(comp, ii, jj, kk, val):
        np.add.at(j3[comp], (kk - 1, jj - 1, ii - 1), val)

    for i, j, k, fx, fy, fz, wx, wy, wz in ((i1, j1, k1, fx1, fy1, fz1, wx1, wy1, wz1),
                                            (i2, j2, k2, fx2, fy2, fz2, wx2, wy2, wz2)):
        add(0, i, j, k, fx * (1 - wy) * (1 - wz))
        add(0, i, j + 1, k, fx * wy * (1 - wz))
        add(1, i, j, k, fy * (1 - wx) * (1 - wz))
        add(1, i + 1, j, k, fy * wx * (1 - wz))
        if dim == 3:
            add(0, i, j, k + 1, fx * (1 - wy) * wz)
            add(0, i, j + 1, k + 1, fx * wy * wz)
            add(1, i, j, k + 1, fy * (1 - wx) * wz)
            add(1, i + 1, j, k + 1, fy * wx * wz)
        add(2, i, j, k, fz * (1 - wx) * (1 - wy))
        add(2, i + 1, j, k, fz * wx * (1 - wy))
        add(2, i, j + 1, k, fz * (1 - wx) * wy)
        add(2, i + 1, j + 1, k, fz * wx * wy)
    return j3
