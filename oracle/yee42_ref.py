"""CPU restatement of the reference's 4th-order field solver (test
infrastructure).

advance_b_halfstep_42 (code/fields.F90:1039-1212) and advance_e_fullstep_42
(code/fields.F90:1223-1361) on a periodic box, in float64 numpy.  Arrays are
interior-only (3, nz, ny, nx); a shift by s cells along an axis is the
periodic image, which is what the reference's ghost fills (bc_b1/bc_e1,
nghost/2 layers per side) provide for every index these stencils read.

The difference operator of both steps is
    D f = (9/8) (f[+1] - f[0]) - (1/24) (f[+2] - f[-1])     (B step, forward)
    D f = (9/8) (f[-1] - f[0]) - (1/24) (f[-2] - f[+1])     (E step, as written)
with coef1 = 9/8 corr (c/2), coef2 = -1/24 corr (c/2) for B and the same
with c for E.  The 2D forms drop the z differences.  Pinned by the compiled
reference's FP64 build (tests/golden/ref_*_ho.npz, Yee-only run), by the
matching divergence of B being left unchanged by the B step, and by the
one-cycle matrix of a vacuum mode (tests/test_yee42_oracle.py).
"""
import numpy as np

X, Y, Z = 2, 1, 0   # axes of an (nz, ny, nx) array


def _s(f, axis, k):
    """f shifted so that element i holds f[i + k] along axis (periodic)"""
    return np.roll(f, -k, axis=axis)


def bhalf42(e, b, c, corr, dim):
    """B += (c/2) corr curl_4(E); e, b: (3, nz, ny, nx) float64, b updated"""
    c1 = 9.0 / 8.0 * corr * (0.5 * c)
    c2 = -1.0 / 24.0 * corr * (0.5 * c)
    ex, ey, ez = e
    if dim == 3:
        b[0] += c1 * (_s(ey, Z, 1) - ey - _s(ez, Y, 1) + ez) + c2 * (_s(ey, Z, 2) - _s(ey, Z, -1) - _s(ez, Y, 2)
                                                                    + _s(ez, Y, -1))
        b[1] += c1 * (_s(ez, X, 1) - ez - _s(ex, Z, 1) + ex) + c2 * (_s(ez, X, 2) - _s(ez, X, -1) - _s(ex, Z, 2)
                                                                    + _s(ex, Z, -1))
    else:
        b[0] += c1 * (-_s(ez, Y, 1) + ez) + c2 * (-_s(ez, Y, 2) + _s(ez, Y, -1))
        b[1] += c1 * (_s(ez, X, 1) - ez) + c2 * (_s(ez, X, 2) - _s(ez, X, -1))
    b[2] += c1 * (_s(ex, Y, 1) - ex - _s(ey, X, 1) + ey) + c2 * (_s(ex, Y, 2) - _s(ex, Y, -1) - _s(ey, X, 2)
                                                                + _s(ey, X, -1))
    return b


def efull42(e, b, c, corr, dim):
    """E += c corr curl_4(B) (no current); e updated"""
    c1 = 9.0 / 8.0 * corr * c
    c2 = -1.0 / 24.0 * corr * c
    bx, by, bz = b
    if dim == 3:
        e[0] += c1 * (_s(by, Z, -1) - by - _s(bz, Y, -1) + bz) + c2 * (_s(by, Z, -2) - _s(by, Z, 1)
                                                                      - _s(bz, Y, -2) + _s(bz, Y, 1))
        e[1] += c1 * (_s(bz, X, -1) - bz - _s(bx, Z, -1) + bx) + c2 * (_s(bz, X, -2) - _s(bz, X, 1)
                                                                      - _s(bx, Z, -2) + _s(bx, Z, 1))
    else:
        e[0] += c1 * (-_s(bz, Y, -1) + bz) + c2 * (-_s(bz, Y, -2) + _s(bz, Y, 1))
        e[1] += c1 * (_s(bz, X, -1) - bz) + c2 * (_s(bz, X, -2) - _s(bz, X, 1))
    e[2] += c1 * (_s(bx, Y, -1) - bx - _s(by, X, -1) + by) + c2 * (_s(bx, Y, -2) - _s(bx, Y, 1)
                                                                  - _s(by, X, -2) + _s(by, X, 1))
    return e


def div4_b(b, dim):
    """the 4th-order divergence matching the B step's differences (forward,
    as the curl of the B step): D f = (9/8)(f[+1] - f[0]) - (1/24)(f[+2] - f[-1])"""
    def d(f, axis):
        return 9.0 / 8.0 * (_s(f, axis, 1) - f) - 1.0 / 24.0 * (_s(f, axis, 2) - _s(f, axis, -1))
    out = d(b[0], X) + d(b[1], Y)
    if dim == 3:
        out = out + d(b[2], Z)
    return out


def yee_only(f6, c, corr, dim):
    """the reference driver's Yee-only run: B half, B half, E (no current)"""
    e = np.array(f6[:3], dtype=np.float64)
    b = np.array(f6[3:6], dtype=np.float64)
    bhalf42(e, b, c, corr, dim)
    bhalf42(e, b, c, corr, dim)
    efull42(e, b, c, corr, dim)
    return np.concatenate([e, b])
