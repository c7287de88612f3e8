"""CPU restatement of the reference's field output flds.tot.NNN for one
periodic domain (test infrastructure).

output_tot (code/output.F90:2494-3875) samples the padded arrays at the
nodes i = istep, 2 istep, ... <= mx (j and k likewise; k = 1 in 2D), the
loop istart..ifinish of :2645-2700 for a domain that is both first and last
along every axis.  Per dataset (:2604-2628, selected at :2860-2960):
  dens, densi, v3x .. v3zi  meanq_fld_cur (:5229-5482) with idx = idy = 2,
      idz = 2 (0 in 2D) (:189-194): every particle adds its term to every
      node of the box [i-idx, i+idx] x ... clamped to the padded array,
      exchange_current folds the ghosts onto their periodic images
      (code/fieldboundaries.F90:1792-1801, 1929-1940 and the z block), each
      node is divided by its clamped box volume, and the velocity datasets
      are curx / cury (0 where cury = 0);
  ex .. bz  interpolfld (:4995-5127) at real(i), real(j), real(k), computed
      here with its full trilinear expression in FP32;
  jx, jy, jz  the current arrays.
The restatement is written the reference's way (per-particle boxes, the
literal interpolation formula), not the library's (per-cell sums and
separable box sums, the neighbour averages alone).

Arrays are (mz, my, mx) float32 C order (x fastest, as the Fortran
arrays); particles (7, n): x, y, z, u, v, w, ch.  Returns a dict
name -> (mz//istep, my//istep, mx//istep) float32.

Pinned by closed forms in tests/test_flds_tot_hdf5.py: uniform fields give
their value at every node, one particle gives ch / box volume inside its box
and its own 3-velocity there, a field linear in x gives the node value of
the line.
"""
import numpy as np

DATASETS_3D = ["dens", "densi", "v3x", "v3xi", "v3y", "v3yi", "v3z", "v3zi",
               "ex", "ey", "ez", "bx", "by", "bz", "jx", "jy", "jz"]
DATASETS_2D_IN_PLANE = [d for d in DATASETS_3D if d not in ("v3z", "v3zi")]


def _ghosts(dim, order):
    ng = 5 if order == 1 else 7
    return ng, (5 if dim == 2 else ng)


def _fold(a, dim, ng, ngz):
    """exchange_current with sizex = sizey = sizez = 1 and periodic axes:
    cur(h+1:ng) += cur(n-h:n); cur(n-ng+1:n-h-1) += cur(1:h), per axis;
    the ghosts keep their values"""
    for axis, gh in ((2, ng), (1, ng), (0, ngz))[:dim]:
        h = gh // 2
        n = a.shape[axis]
        sl = [slice(None)] * 3

        def take(lo, hi):
            s = list(sl)
            s[axis] = slice(lo - 1, hi)
            return tuple(s)

        up = a[take(n - h, n)].copy()
        lo = a[take(1, h)].copy()
        a[take(h + 1, gh)] += up
        a[take(n - gh + 1, n - h - 1)] += lo
    return a


def _meanq(which, dim, shape, ions, lecs):
    mz, my, mx = shape
    idx, idy, idz = 2, 2, (2 if dim == 3 else 0)
    curx = np.zeros(shape, np.float64)
    cury = np.zeros(shape, np.float64)
    for s, p in enumerate((ions, lecs)):
        for q in range(p.shape[1]):
            x, y, z, u, v, w, ch = (np.float32(t) for t in p[:, q])
            gam = np.float32(1.0) / np.sqrt(np.float32(1.0) + u * u + v * v + w * w)
            ax = ay = 0.0
            if which == "tdens":
                ax = ch
            elif which == "idens":
                ax = ch if s == 0 else 0.0
            else:
                comp = {"x": u, "y": v, "z": w}[which[-1]]
                if which[0] == "t" or s == 0:
                    ax, ay = comp * gam * ch, ch
            i, j = int(x), int(y)
            k = int(z) if dim == 3 else 1
            lz1, lz2 = (max(k - idz, 1), min(k + idz, mz)) if dim == 3 else (1, 1)
            ly1, ly2 = max(j - idy, 1), min(j + idy, my)
            lx1, lx2 = max(i - idx, 1), min(i + idx, mx)
            curx[lz1 - 1:lz2, ly1 - 1:ly2, lx1 - 1:lx2] += ax
            cury[lz1 - 1:lz2, ly1 - 1:ly2, lx1 - 1:lx2] += ay
    return curx, cury


def _normalise(curx, cury, dim, ratio):
    mz, my, mx = curx.shape
    idx, idy, idz = 2, 2, (2 if dim == 3 else 0)
    k = np.arange(1, mz + 1)[:, None, None]
    j = np.arange(1, my + 1)[None, :, None]
    i = np.arange(1, mx + 1)[None, None, :]
    vol = ((np.minimum(i + idx, mx) - np.maximum(i - idx, 1) + 1) *
           (np.minimum(j + idy, my) - np.maximum(j - idy, 1) + 1))
    vol = vol * ((np.minimum(k + idz, mz) - np.maximum(k - idz, 1) + 1) if dim == 3 else 1)
    curx = curx / vol
    cury = cury / vol
    if ratio:
        out = np.zeros_like(curx)
        nz = cury != 0.0
        out[nz] = curx[nz] / cury[nz]
        return out
    return curx


def interpolfld(f6, dim, i, j, k):
    """interpolfld (code/output.F90:4995-5127) at the integer point (i, j, k)
    (1-based padded), FP32, the full trilinear expression with
    dx = dy = dz = 0; returns ex0, ey0, ez0, bx0, by0, bz0"""
    ex, ey, ez, bx, by, bz = (a.reshape(-1) for a in f6)
    mz, my, mx = f6[0].shape
    z0 = np.float32(0.0)
    dx = dy = dz = z0
    i = 2 if i == 1 else (mx - 1 if i == mx else i)
    j = 2 if j == 1 else (my - 1 if j == my else j)
    if dim == 3:
        k = 2 if k == 1 else (mz - 1 if k == mz else k)
    else:
        k = 1
    ix, iy, iz = 1, mx, (mx * my if dim == 3 else 0)
    l = (i - 1) + iy * (j - 1) + iz * (k - 1)
    half, quarter = np.float32(0.5), np.float32(0.25)

    def e_comp(a, s0, s1, s2, d0, d1, d2):
        # f(s0) along the staggered axis, then d1 along s1, g at + s2
        f = a[l] + a[l - s0] + d0 * (a[l + s0] - a[l - s0])
        f = f + d1 * (a[l + s1] + a[l - s0 + s1] + d0 * (a[l + s0 + s1] - a[l - s0 + s1]) - f)
        g = a[l + s2] + a[l - s0 + s2] + d0 * (a[l + s0 + s2] - a[l - s0 + s2])
        g = g + d1 * (a[l + s1 + s2] + a[l - s0 + s1 + s2] + d0 * (a[l + s0 + s1 + s2] - a[l - s0 + s1 + s2]) - g)
        return (f + d2 * (g - f)) * half

    def b_comp(a, s0, s1, s2, d0, d1, d2):
        # a staggered in s1 and s2 (the face normal is the third axis)
        f = a[l - s1] + a[l - s1 - s0] + d0 * (a[l - s1 + s0] - a[l - s1 - s0])
        f = (a[l] + a[l - s0] + d0 * (a[l + s0] - a[l - s0]) + f +
             d1 * (a[l + s1] + a[l + s1 - s0] + d0 * (a[l + s1 + s0] - a[l + s1 - s0]) - f))
        g = a[l + s2 - s1] + a[l + s2 - s1 - s0] + d0 * (a[l + s2 - s1 + s0] - a[l + s2 - s1 - s0])
        g = (a[l + s2] + a[l + s2 - s0] + d0 * (a[l + s2 + s0] - a[l + s2 - s0]) + g +
             d1 * (a[l + s2 + s1] + a[l + s2 + s1 - s0] + d0 * (a[l + s2 + s1 + s0] - a[l + s2 + s1 - s0]) - g))
        return (f + d2 * (g - f)) * quarter

    ex0 = e_comp(ex, ix, iy, iz, dx, dy, dz)
    ey0 = e_comp(ey, iy, iz, ix, dy, dz, dx)
    ez0 = e_comp(ez, iz, ix, iy, dz, dx, dy)
    bx0 = b_comp(bx, iz, iy, ix, dz, dy, dx)
    by0 = b_comp(by, ix, iz, iy, dx, dz, dy)
    bz0 = b_comp(bz, iy, ix, iz, dy, dx, dz)
    return ex0, ey0, ez0, bx0, by0, bz0


def flds_tot(dim, order, nx, ny, nz, istep, fields, currents, ions, lecs, in_plane=False):
    ng, ngz = _ghosts(dim, order)
    mx, my = nx + ng, ny + ng
    mz = nz + ngz if dim == 3 else 1
    shape = (mz, my, mx)
    f6 = [np.asarray(a, np.float32).reshape(shape) for a in fields]
    j3 = [np.asarray(a, np.float32).reshape(shape) for a in currents]
    ox, oy = mx // istep, my // istep
    oz = mz // istep if dim == 3 else 1
    ii = [(a + 1) * istep for a in range(ox)]
    jj = [(b + 1) * istep for b in range(oy)]
    kk = [(c + 1) * istep for c in range(oz)] if dim == 3 else [1]
    names = DATASETS_2D_IN_PLANE if (dim == 2 and in_plane) else DATASETS_3D
    tot = {"dens": "tdens", "densi": "idens", "v3x": "tbetx", "v3y": "tbety", "v3z": "tbetz",
           "v3xi": "ibetx", "v3yi": "ibety", "v3zi": "ibetz"}
    out = {}
    node = {}
    for c, k in enumerate(kk):
        for b, j in enumerate(jj):
            for a, i in enumerate(ii):
                node[(c, b, a)] = interpolfld(f6, dim, i, j, k)
    for name in names:
        res = np.zeros((oz, oy, ox), np.float32)
        if name in tot:
            curx, cury = _meanq(tot[name], dim, shape, ions, lecs)
            curx = _fold(curx, dim, ng, ngz)
            cury = _fold(cury, dim, ng, ngz)
            val = _normalise(curx, cury, dim, ratio=name not in ("dens", "densi"))
            for c, k in enumerate(kk):
                for b, j in enumerate(jj):
                    for a, i in enumerate(ii):
                        res[c, b, a] = val[k - 1, j - 1, i - 1]
        elif name[0] in "eb":
            q = ["ex", "ey", "ez", "bx", "by", "bz"].index(name)
            for key, v in node.items():
                res[key] = v[q]
        else:
            q = ["jx", "jy", "jz"].index(name)
            for c, k in enumerate(kk):
                for b, j in enumerate(jj):
                    for a, i in enumerate(ii):
                        res[c, b, a] = j3[q][k - 1, j - 1, i - 1]
        out[name] = res
    return out
