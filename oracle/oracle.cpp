/*
 * oracle.cpp -- CPU restatement of the reference's per-step PIC loop.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline).  The product
 * library never links or calls this code.
 *
 * Each routine follows the reference routine cited next to it, in the same
 * arithmetic (template T = float reproduces real(sprec)=4; T = double is the
 * accuracy yardstick) and, where the reference spells an expression out, in
 * the same evaluation order, so that the FP32 build can be compared term by
 * term with the compiled reference (oracle/_ref).  Quirks of the reference
 * are reproduced when the matching ORC_Q_* bit is set in cfg->compat.
 *
 * Pinning (see tests/): golden fixtures produced by the compiled reference
 * (tests/golden/), plus closed-form checks (B-spline partition of unity,
 * discrete charge continuity, Boris invariants, vacuum-wave dispersion,
 * filter transfer function).
 */
#include "oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#include <algorithm>

namespace {

struct Dims {
    int dim, ord, nx, ny, nz, mx, my, mz, ng, ngz, h, hz;
    int64_t sy, sz;  /* linear strides of the padded arrays */
    int64_t ncell;
};

/* code/fields.F90:166-184 (nghost by order, 2D nghostz=5), :221-229 (2D mz=1) */
Dims dims_of(const orc_cfg* c)
{
    Dims d;
    d.dim = c->dim;
    d.ord = c->order;
    d.ng = (c->order == 1) ? 5 : 7;
    d.ngz = (c->dim == 2) ? 5 : d.ng;
    d.nx = c->nx;
    d.ny = c->ny;
    d.nz = (c->dim == 2) ? 1 : c->nz;
    d.mx = d.nx + d.ng;
    d.my = d.ny + d.ng;
    d.mz = (c->dim == 2) ? 1 : d.nz + d.ngz;
    d.h = d.ng / 2;
    d.hz = d.ngz / 2;
    d.sy = d.mx;
    d.sz = (int64_t)d.mx * d.my;
    d.ncell = d.sz * d.mz;
    return d;
}

/* 1-based (i,j,k) -> linear offset */
inline int64_t L(const Dims& d, int i, int j, int k)
{
    return (int64_t)(i - 1) + d.sy * (j - 1) + d.sz * (k - 1);
}

template <typename T>
struct Consts {
    /* code/particles.F90:239-251 */
    T three = T(3), two = T(2), thhalf = T(3) / T(2), nineighth = T(9) / T(8),
      one = T(1), threeq = T(3) / T(4), twoth = T(2) / T(3), half = T(1) / T(2),
      third = T(1) / T(3), quart = T(1) / T(4), sixth = T(1) / T(6),
      negsixth = T(-1) / T(6), negone = T(-1);
};

/* Fortran aint() of a positive coordinate */
template <typename T>
inline int aint_(T x) { return (int)x; }

template <typename T>
inline T fsign(T a, T b) { return b >= T(0) ? std::fabs(a) : -std::fabs(a); }

/* ------------------------------------------------------------------ */
/* shape functions on the 6-slot frame (slot 3 = cell aint(x))          */
/* ord 1: code/particles_movedeposit.F90:447-468, code/particles.F90:738-773 */
/* ord 2: code/particles_movedeposit.F90:709-790, code/particles.F90:925-1020 */
/* ord 3: code/particles_movedeposit.F90:1035-1128, code/particles.F90:1172-1277 */
/* S[0..7] holds slots 0..7 (slots 1..6 used); lo/hi = slot range.     */
template <typename T>
void shape(int ord, T d, int shift, T* S, int& lo, int& hi)
{
    Consts<T> K;
    const int s = shift;
    if (ord == 1) {
        S[3 + s] = K.one - d;   /* 1.-d */
        S[4 + s] = d;
        lo = 3 + s; hi = 4 + s;
    } else if (ord == 2) {
        if (d <= K.half) {
            S[2 + s] = K.half * (d * d - d + K.quart);
            S[4 + s] = S[2 + s] + d;
            S[3 + s] = K.one - S[4 + s] - S[2 + s];
            lo = 2 + s; hi = 4 + s;
        } else {
            S[3 + s] = K.nineighth - K.thhalf * d + K.half * d * d;
            S[5 + s] = S[3 + s] - K.one + d;
            S[4 + s] = K.one - S[5 + s] - S[3 + s];
            lo = 3 + s; hi = 5 + s;
        }
    } else {
        if (d <= K.half) {
            S[2 + s] = K.negsixth * (d - K.one) * (d - K.one) * (d - K.one);
            S[3 + s] = K.twoth + K.half * (d - K.two) * d * d;
            S[5 + s] = K.sixth * d * d * d;
            S[4 + s] = K.one - S[5 + s] - S[3 + s] - S[2 + s];
        } else {
            S[5 + s] = K.sixth * d * d * d;
            S[4 + s] = K.twoth + K.half * (K.negone - d) * (K.one - d) * (K.one - d);
            S[2 + s] = K.sixth * (K.one - d) * (K.one - d) * (K.one - d);
            S[3 + s] = K.one - S[5 + s] - S[4 + s] - S[2 + s];
        }
        lo = 2 + s; hi = 5 + s;
    }
}

/* ------------------------------------------------------------------ */
/* periodic ghost fill: code/fieldboundaries.F90:181-263 (bc_b1), 306-392
 * (bc_e1) via copylayrx/y (:714-785) and copy_layrz1_opt (:1616-1680).
 * nghost/2 layers per side; the last layer m is not filled.           */
template <typename T>
void ghost_fill_one(const Dims& d, T* a)
{
    for (int it = 1; it <= d.h; ++it) {
        int lt = d.h + 1 - it, ls = d.mx - (d.h + it), nt = d.mx - (d.h + 1) + it, ns = d.h + it;
        for (int k = 1; k <= d.mz; ++k)
            for (int j = 1; j <= d.my; ++j) a[L(d, lt, j, k)] = a[L(d, ls, j, k)];
        for (int k = 1; k <= d.mz; ++k)
            for (int j = 1; j <= d.my; ++j) a[L(d, nt, j, k)] = a[L(d, ns, j, k)];
    }
    for (int it = 1; it <= d.h; ++it) {
        int lt = d.h + 1 - it, ls = d.my - (d.h + it), nt = d.my - (d.h + 1) + it, ns = d.h + it;
        for (int k = 1; k <= d.mz; ++k)
            for (int i = 1; i <= d.mx; ++i) a[L(d, i, lt, k)] = a[L(d, i, ls, k)];
        for (int k = 1; k <= d.mz; ++k)
            for (int i = 1; i <= d.mx; ++i) a[L(d, i, nt, k)] = a[L(d, i, ns, k)];
    }
    if (d.dim == 3) {
        for (int it = 1; it <= d.hz; ++it) {
            int lt = d.hz + 1 - it, ls = d.mz - (d.hz + it), nt = d.mz - (d.hz + 1) + it, ns = d.hz + it;
            for (int j = 1; j <= d.my; ++j)
                for (int i = 1; i <= d.mx; ++i) a[L(d, i, j, lt)] = a[L(d, i, j, ls)];
            for (int j = 1; j <= d.my; ++j)
                for (int i = 1; i <= d.mx; ++i) a[L(d, i, j, nt)] = a[L(d, i, j, ns)];
        }
    }
}

/* one-layer refresh used by each filter pass: copy_layr{x,y,z}1_opt with
 * (lt,ls,nt,ns) = (h, m-(h+1), m-h, h+1)   (code/filter.F90:71-99)      */
template <typename T>
void filter_refresh(const Dims& d, T* a)
{
    {
        int lt = d.h, ls = d.mx - (d.h + 1), nt = d.mx - d.h, ns = d.h + 1;
        for (int k = 1; k <= d.mz; ++k)
            for (int j = 1; j <= d.my; ++j) {
                a[L(d, lt, j, k)] = a[L(d, ls, j, k)];
            }
        for (int k = 1; k <= d.mz; ++k)
            for (int j = 1; j <= d.my; ++j) a[L(d, nt, j, k)] = a[L(d, ns, j, k)];
    }
    {
        int lt = d.h, ls = d.my - (d.h + 1), nt = d.my - d.h, ns = d.h + 1;
        for (int k = 1; k <= d.mz; ++k)
            for (int i = 1; i <= d.mx; ++i) a[L(d, i, lt, k)] = a[L(d, i, ls, k)];
        for (int k = 1; k <= d.mz; ++k)
            for (int i = 1; i <= d.mx; ++i) a[L(d, i, nt, k)] = a[L(d, i, ns, k)];
    }
    if (d.dim == 3) {
        int lt = d.hz, ls = d.mz - (d.hz + 1), nt = d.mz - d.hz, ns = d.hz + 1;
        for (int j = 1; j <= d.my; ++j)
            for (int i = 1; i <= d.mx; ++i) a[L(d, i, j, lt)] = a[L(d, i, j, ls)];
        for (int j = 1; j <= d.my; ++j)
            for (int i = 1; i <= d.mx; ++i) a[L(d, i, j, nt)] = a[L(d, i, j, ns)];
    }
}

/* ------------------------------------------------------------------ */
/* code/fields.F90:586-728 advance_b_halfstep (periodic ranges)        */
template <typename T>
void bhalf(const orc_cfg* cfg, T* f)
{
    Dims d = dims_of(cfg);
    T *ex = f, *ey = f + d.ncell, *ez = f + 2 * d.ncell, *bx = f + 3 * d.ncell,
      *by = f + 4 * d.ncell, *bz = f + 5 * d.ncell;
    const T c = (T)cfg->c, corr = (T)cfg->corr;
    const T cst = corr * (T(0.5) * c);
    int i1 = d.h + 1, i2 = d.mx - (d.h + 1), j1 = d.h + 1, j2 = d.my - (d.h + 1);
    if (d.dim == 3) {
        int k1 = d.hz + 1, k2 = d.mz - (d.hz + 1);
        for (int k = k1; k <= k2; ++k)
            for (int j = j1; j <= j2; ++j)
                for (int i = i1; i <= i2; ++i) {
                    int64_t o = L(d, i, j, k);
                    bx[o] = bx[o] + cst * (ey[o + d.sz] - ey[o] - ez[o + d.sy] + ez[o]);
                    by[o] = by[o] + cst * (ez[o + 1] - ez[o] - ex[o + d.sz] + ex[o]);
                    bz[o] = bz[o] + cst * (ex[o + d.sy] - ex[o] - ey[o + 1] + ey[o]);
                }
    } else {
        for (int j = j1; j <= j2; ++j)
            for (int i = i1; i <= i2; ++i) {
                int64_t o = L(d, i, j, 1);
                bx[o] = bx[o] + cst * (-ez[o + d.sy] + ez[o]);
                by[o] = by[o] + cst * (ez[o + 1] - ez[o]);
                bz[o] = bz[o] + cst * (ex[o + d.sy] - ex[o] - ey[o + 1] + ey[o]);
            }
    }
}

/* code/fields.F90:739-870 advance_e_fullstep; 3D k range forced to
 * [nghostz/2, mz] (:818-819)                                           */
template <typename T>
void efull(const orc_cfg* cfg, T* f)
{
    Dims d = dims_of(cfg);
    T *ex = f, *ey = f + d.ncell, *ez = f + 2 * d.ncell, *bx = f + 3 * d.ncell,
      *by = f + 4 * d.ncell, *bz = f + 5 * d.ncell;
    const T c = (T)cfg->c, corr = (T)cfg->corr;
    const T cst = corr * c;
    int i1 = d.h + 1, i2 = d.mx - (d.h + 1), j1 = d.h + 1, j2 = d.my - (d.h + 1);
    if (d.dim == 3) {
        int k1 = d.hz, k2 = d.mz;
        for (int k = k1; k <= k2; ++k)
            for (int j = j1; j <= j2; ++j)
                for (int i = i1; i <= i2; ++i) {
                    int64_t o = L(d, i, j, k);
                    ex[o] = ex[o] + cst * (by[o - d.sz] - by[o] - bz[o - d.sy] + bz[o]);
                    ey[o] = ey[o] + cst * (bz[o - 1] - bz[o] - bx[o - d.sz] + bx[o]);
                    ez[o] = ez[o] + cst * (bx[o - d.sy] - bx[o] - by[o - 1] + by[o]);
                }
    } else {
        for (int j = j1; j <= j2; ++j)
            for (int i = i1; i <= i2; ++i) {
                int64_t o = L(d, i, j, 1);
                ex[o] = ex[o] + cst * (-bz[o - d.sy] + bz[o]);
                ey[o] = ey[o] + cst * (bz[o - 1] - bz[o]);
                ez[o] = ez[o] + cst * (bx[o - d.sy] - bx[o] - by[o - 1] + by[o]);
            }
    }
}

/* code/fieldboundaries.F90:1768-2189 exchange_current, single periodic
 * domain: x, then y, then z (z through the self-send path)            */
template <typename T>
void fold(const orc_cfg* cfg, T* j3)
{
    Dims d = dims_of(cfg);
    for (int comp = 0; comp < 3; ++comp) {
        T* a = j3 + comp * d.ncell;
        /* x */
        for (int k = 1; k <= d.mz; ++k)
            for (int j = 1; j <= d.my; ++j) {
                T b1[8], b2[8];
                for (int t = 0; t <= d.h; ++t) b1[t] = a[L(d, d.mx - d.h + t, j, k)];
                for (int t = 0; t < d.h; ++t) b2[t] = a[L(d, 1 + t, j, k)];
                for (int t = 0; t <= d.h; ++t) a[L(d, d.h + 1 + t, j, k)] += b1[t];
                for (int t = 0; t < d.h; ++t) a[L(d, d.mx - (d.ng - 1) + t, j, k)] += b2[t];
            }
        /* y */
        for (int k = 1; k <= d.mz; ++k)
            for (int i = 1; i <= d.mx; ++i) {
                T b1[8], b2[8];
                for (int t = 0; t <= d.h; ++t) b1[t] = a[L(d, i, d.my - d.h + t, k)];
                for (int t = 0; t < d.h; ++t) b2[t] = a[L(d, i, 1 + t, k)];
                for (int t = 0; t <= d.h; ++t) a[L(d, i, d.h + 1 + t, k)] += b1[t];
                for (int t = 0; t < d.h; ++t) a[L(d, i, d.my - (d.ng - 1) + t, k)] += b2[t];
            }
        /* z */
        if (d.dim == 3) {
            for (int j = 1; j <= d.my; ++j)
                for (int i = 1; i <= d.mx; ++i) {
                    T b1[8], b2[8];
                    for (int t = 0; t <= d.hz; ++t) b1[t] = a[L(d, i, j, d.mz - d.hz + t)];
                    for (int t = 0; t <= d.hz; ++t) a[L(d, i, j, d.hz + 1 + t)] += b1[t];
                    for (int t = 0; t < d.hz; ++t) b2[t] = a[L(d, i, j, 1 + t)];
                    for (int t = 0; t < d.hz; ++t) a[L(d, i, j, d.mz - (d.ngz - 1) + t)] += b2[t];
                }
        }
    }
}

/* code/filter.F90:8-221 apply_filter1_opt                              */
template <typename T>
void filter(const orc_cfg* cfg, T* j3, int ntimes)
{
    Dims d = dims_of(cfg);
    std::vector<T> temp(d.ncell);
    T wtl, wt, wtr, wtu, wtb, wtlt, wtrt, wtlb, wtrb;
    T wtlz = 0, wtz = 0, wtrz = 0, wtuz = 0, wtbz = 0, wtltz = 0, wtrtz = 0, wtlbz = 0, wtrbz = 0;
    if (d.dim == 2) {
        T winv = T(1) / T(16);
        wtl = T(2) * winv; wt = T(4) * winv; wtr = T(2) * winv; wtu = T(2) * winv; wtb = T(2) * winv;
        wtlt = winv; wtrt = winv; wtlb = winv; wtrb = winv;
    } else {
        T winv = T(1) / T(64);
        wtl = T(4) * winv; wt = T(8) * winv; wtr = T(4) * winv; wtu = T(4) * winv; wtb = T(4) * winv;
        wtlt = T(2) * winv; wtrt = T(2) * winv; wtlb = T(2) * winv; wtrb = T(2) * winv;
        wtlz = T(2) * winv; wtz = T(4) * winv; wtrz = T(2) * winv; wtuz = T(2) * winv; wtbz = T(2) * winv;
        wtltz = winv; wtrtz = winv; wtlbz = winv; wtrbz = winv;
    }
    const int istr = d.h + 1, ifin = d.mx - (d.h + 1);
    const int k1 = (d.dim == 3) ? d.hz + 1 : 1, k2 = (d.dim == 3) ? d.mz - (d.hz + 1) : 1;
    for (int n = 1; n <= ntimes; ++n) {
        std::fill(temp.begin(), temp.end(), T(0));
        for (int comp = 0; comp < 3; ++comp) filter_refresh(d, j3 + comp * d.ncell);
        for (int comp = 0; comp < 3; ++comp) {
            T* cu = j3 + comp * d.ncell;
            /* quirk: the curz loops run to my-nghost/2+1 (code/filter.F90:186,211) */
            const int j1 = d.h + 1, j2 = (comp == 2) ? d.my - d.h + 1 : d.my - (d.h + 1);
            for (int k = k1; k <= k2; ++k)
                for (int j = j1; j <= j2; ++j)
                    for (int i = istr; i <= ifin; ++i) {
                        int64_t o = L(d, i, j, k);
                        T t = wtl * cu[o - 1] + wt * cu[o] + wtr * cu[o + 1] +
                              wtb * cu[o - d.sy] + wtu * cu[o + d.sy] +
                              wtlt * cu[o - 1 + d.sy] + wtrt * cu[o + 1 + d.sy] +
                              wtlb * cu[o - 1 - d.sy] + wtrb * cu[o + 1 - d.sy];
                        if (d.dim == 3) {
                            int64_t m = o - d.sz, p = o + d.sz;
                            t = t + wtlz * cu[m - 1] + wtz * cu[m] + wtrz * cu[m + 1] +
                                wtbz * cu[m - d.sy] + wtuz * cu[m + d.sy] +
                                wtltz * cu[m - 1 + d.sy] + wtrtz * cu[m + 1 + d.sy] +
                                wtlbz * cu[m - 1 - d.sy] + wtrbz * cu[m + 1 - d.sy] +
                                wtlz * cu[p - 1] + wtz * cu[p] + wtrz * cu[p + 1] +
                                wtbz * cu[p - d.sy] + wtuz * cu[p + d.sy] +
                                wtltz * cu[p - 1 + d.sy] + wtrtz * cu[p + 1 + d.sy] +
                                wtlbz * cu[p - 1 - d.sy] + wtrbz * cu[p + 1 - d.sy];
                        }
                        temp[o] = t;
                    }
            for (int k = k1; k <= k2; ++k)
                for (int j = j1; j <= j2; ++j)
                    for (int i = istr; i <= ifin; ++i) {
                        int64_t o = L(d, i, j, k);
                        cu[o] = temp[o];
                    }
        }
    }
}

/* ------------------------------------------------------------------ */
/* mover_{1,2,3}ord: code/particles_movedeposit.F90:356-610, 619-933,
 * 943-1271 (Boris branch, no external fields)                          */
template <typename T>
void mover(const orc_cfg* cfg, const T* f, T* prt, int64_t n, T qm)
{
    Dims d = dims_of(cfg);
    Consts<T> K;
    const T half = K.half;
    const T c = (T)cfg->c;
    const T cinv = T(1) / c;
    const T *ex = f, *ey = f + d.ncell, *ez = f + 2 * d.ncell, *bx = f + 3 * d.ncell,
            *by = f + 4 * d.ncell, *bz = f + 5 * d.ncell;
    T *X = prt, *Y = prt + n, *Z = prt + 2 * n, *U = prt + 3 * n, *V = prt + 4 * n, *W = prt + 5 * n;

    /* 3D primal temporaries with whole-array cshift (:982-991) */
    std::vector<T> P[6];
    if (d.dim == 3) {
        for (auto& v : P) v.assign(d.ncell, T(0));
        for (int k = 1; k <= d.mz; ++k) {
            int km = (k == 1) ? d.mz : k - 1;
            for (int j = 1; j <= d.my; ++j) {
                int jm = (j == 1) ? d.my : j - 1;
                for (int i = 1; i <= d.mx; ++i) {
                    int im = (i == 1) ? d.mx : i - 1;
                    int64_t o = L(d, i, j, k);
                    P[0][o] = T(0.5) * (ex[o] + ex[L(d, im, j, k)]);
                    P[1][o] = T(0.5) * (ey[o] + ey[L(d, i, jm, k)]);
                    P[2][o] = T(0.5) * (ez[o] + ez[L(d, i, j, km)]);
                    if (cfg->compat & ORC_Q_BXY_Z) {
                        P[3][o] = T(0.5) * (bx[o] + bx[L(d, i, jm, k)]);
                        P[4][o] = T(0.5) * (by[o] + by[L(d, im, j, k)]);
                    } else {
                        P[3][o] = T(0.5) * (T(0.5) * (bx[o] + bx[L(d, i, jm, k)]) +
                                            T(0.5) * (bx[L(d, i, j, km)] + bx[L(d, i, jm, km)]));
                        P[4][o] = T(0.5) * (T(0.5) * (by[o] + by[L(d, im, j, k)]) +
                                            T(0.5) * (by[L(d, i, j, km)] + by[L(d, im, j, km)]));
                    }
                    P[5][o] = T(0.5) * (T(0.5) * (bz[o] + bz[L(d, im, j, k)]) +
                                        T(0.5) * (bz[L(d, i, jm, k)] + bz[L(d, im, jm, k)]));
                }
            }
        }
    }

    for (int64_t p = 0; p < n; ++p) {
        T Sxp[8] = {0}, Syp[8] = {0}, Szp[8] = {0}, Sxd[8] = {0}, Syd[8] = {0}, Szd[8] = {0};
        int ip = aint_(X[p]);
        T dxp = X[p] - (T)ip;
        int jp = aint_(Y[p]);
        T dyp = Y[p] - (T)jp;
        int kp = aint_(Z[p]);
        int id = aint_(X[p] - half);
        T dxd = X[p] - half - (T)id;
        int jd = aint_(Y[p] - half);
        T dyd = Y[p] - half - (T)jd;
        int kd = aint_(Z[p] - half);
        T dzp = Z[p] - (T)kp;
        T dzd = (Z[p] - half) - (T)kd;
        int xlo, xhi, ylo, yhi, zlo = 3, zhi = 3, dlo, dhi;
        shape<T>(d.ord, dxp, 0, Sxp, xlo, xhi);
        shape<T>(d.ord, dyp, 0, Syp, ylo, yhi);
        if (d.dim == 3) shape<T>(d.ord, dzp, 0, Szp, zlo, zhi);
        shape<T>(d.ord, dxd, 0, Sxd, dlo, dhi);
        int x1min = xlo, x1max = xhi, dx1min = dlo, dx1max = dhi;
        shape<T>(d.ord, dyd, 0, Syd, dlo, dhi);
        int y1min = ylo, y1max = yhi, dy1min = dlo, dy1max = dhi;
        int z1min = zlo, z1max = zhi;
        if (d.dim == 3) {
            shape<T>(d.ord, dzd, 0, Szd, dlo, dhi);
            if (d.ord == 2 && (cfg->compat & ORC_Q_GATHER2)) { z1min = dlo; z1max = dhi; }
        }
        /* quirk 1: the loop bounds are those set last, by the dual blocks */
        int gx0, gx1, gy0, gy1;
        if (d.ord == 2 && (cfg->compat & ORC_Q_GATHER2)) {
            x1min = dx1min; x1max = dx1max; y1min = dy1min; y1max = dy1max;
            gx0 = dx1min; gx1 = dx1max; gy0 = dy1min; gy1 = dy1max;
        } else {
            /* corrected: every weight family over its own support
               (in the 2D branch the reference loops over one range for all
               components; corrected mode uses the union, which adds zeros) */
            gx0 = std::min(x1min, dx1min); gx1 = std::max(x1max, dx1max);
            gy0 = std::min(y1min, dy1min); gy1 = std::max(y1max, dy1max);
        }

        T ex0 = 0, ey0 = 0, ez0 = 0, bx0 = 0, by0 = 0, bz0 = 0;
        if (d.dim == 3) {
            for (int i3 = z1min; i3 <= z1max; ++i3)
                for (int i2 = y1min; i2 <= y1max; ++i2) {
                    T acc[6];
                    for (int q = 0; q < 6; ++q) {
                        T s = 0;
                        for (int i1 = x1min; i1 <= x1max; ++i1)
                            s = s + P[q][L(d, ip - 3 + i1, jp - 3 + i2, kp - 3 + i3)] * Sxp[i1];
                        acc[q] = s;
                    }
                    ex0 = ex0 + acc[0] * Syp[i2] * Szp[i3];
                    ey0 = ey0 + acc[1] * Syp[i2] * Szp[i3];
                    ez0 = ez0 + acc[2] * Syp[i2] * Szp[i3];
                    bx0 = bx0 + acc[3] * Syp[i2] * Szp[i3];
                    by0 = by0 + acc[4] * Syp[i2] * Szp[i3];
                    bz0 = bz0 + acc[5] * Syp[i2] * Szp[i3];
                }
        } else {
            for (int i2 = gy0; i2 <= gy1; ++i2)
                for (int i1 = gx0; i1 <= gx1; ++i1) {
                    int64_t lpp = L(d, ip - 3 + i1, jp - 3 + i2, 1);
                    int64_t lpd = L(d, ip - 3 + i1, jd - 3 + i2, 1);
                    int64_t ldp = L(d, id - 3 + i1, jp - 3 + i2, 1);
                    int64_t ldd = L(d, id - 3 + i1, jd - 3 + i2, 1);
                    ex0 = ex0 + ex[ldp] * Sxd[i1] * Syp[i2];
                    ey0 = ey0 + ey[lpd] * Sxp[i1] * Syd[i2];
                    ez0 = ez0 + ez[lpp] * Sxp[i1] * Syp[i2];
                    bx0 = bx0 + bx[lpd] * Sxp[i1] * Syd[i2];
                    by0 = by0 + by[ldp] * Sxd[i1] * Syp[i2];
                    bz0 = bz0 + bz[ldd] * Sxd[i1] * Syd[i2];
                }
        }
        ex0 = T(0.5) * ex0 * qm;
        ey0 = T(0.5) * ey0 * qm;
        ez0 = T(0.5) * ez0 * qm;
        bx0 = T(0.5) * bx0 * qm * cinv;
        by0 = T(0.5) * by0 * qm * cinv;
        bz0 = T(0.5) * bz0 * qm * cinv;

        /* Boris: :1224-1244 */
        T u0 = c * U[p] + ex0;
        T v0 = c * V[p] + ey0;
        T w0 = c * W[p] + ez0;
        T g = c / std::sqrt(c * c + u0 * u0 + v0 * v0 + w0 * w0);
        bx0 = g * bx0;
        by0 = g * by0;
        bz0 = g * bz0;
        T ff = T(2) / (T(1) + bx0 * bx0 + by0 * by0 + bz0 * bz0);
        T u1 = (u0 + v0 * bz0 - w0 * by0) * ff;
        T v1 = (v0 + w0 * bx0 - u0 * bz0) * ff;
        T w1 = (w0 + u0 * by0 - v0 * bx0) * ff;
        u0 = u0 + v1 * bz0 - w1 * by0 + ex0;
        v0 = v0 + w1 * bx0 - u1 * bz0 + ey0;
        w0 = w0 + u1 * by0 - v1 * bx0 + ez0;
        U[p] = u0 * cinv;
        V[p] = v0 * cinv;
        W[p] = w0 * cinv;
        /* position: :1264-1267 */
        g = c / std::sqrt(c * c + u0 * u0 + v0 * v0 + w0 * w0);
        X[p] = X[p] + U[p] * g * c;
        Y[p] = Y[p] + V[p] * g * c;
        Z[p] = Z[p] + W[p] * g * c;
    }
}

/* densdecomp_{1,2,3}ord: code/particles.F90:678-854, 864-1102, 1112-1358.
 * (x2,y2,z2) new, (x1,y1,z1) old.                                      */
template <typename T>
void densdecomp(const orc_cfg* cfg, const Dims& d, T* cx, T* cy, T* cz, T q,
                T x2, T y2, T z2, T x1, T y1, T z1)
{
    Consts<T> K;
    const T half = K.half, third = K.third;
    int i1 = aint_(x1), j1 = aint_(y1), k1 = aint_(z1);
    int shifti = aint_(x2) - aint_(x1);
    int shiftj = aint_(y2) - aint_(y1);
    int shiftk = aint_(z2) - aint_(z1);
    T dx1 = x1 - (T)aint_(x1), dy1 = y1 - (T)aint_(y1);
    T dx2 = x2 - (T)aint_(x2), dy2 = y2 - (T)aint_(y2);
    T dz1 = z1 - (T)aint_(z1), dz2 = z2 - (T)aint_(z2);
    T deltaz = z2 - z1;
    T Sx1[8] = {0}, Sx2[8] = {0}, Sy1[8] = {0}, Sy2[8] = {0}, Sz1[8] = {0}, Sz2[8] = {0};
    int a, b, c2, e;
    shape<T>(d.ord, dx1, 0, Sx1, a, b);
    shape<T>(d.ord, dx2, shifti, Sx2, c2, e);
    int xmin = std::min(a, c2), xmax = std::max(b, e);
    shape<T>(d.ord, dy1, 0, Sy1, a, b);
    shape<T>(d.ord, dy2, shiftj, Sy2, c2, e);
    int ymin = std::min(a, c2), ymax = std::max(b, e);
    const bool carry = (cfg->compat & ORC_Q_CARRY) != 0;
    if (d.dim == 3) {
        shape<T>(d.ord, dz1, 0, Sz1, a, b);
        shape<T>(d.ord, dz2, shiftk, Sz2, c2, e);
        int zmin = std::min(a, c2), zmax = std::max(b, e);
        T curx_add = 0, curx_prev = 0;
        T cury_adds[8] = {0}, cury_prevs[8] = {0};
        T curz_adds[8][8] = {{0}}, curz_prevs[8][8] = {{0}};
        for (int i2 = zmin; i2 <= zmax; ++i2) {
            if (!carry) std::fill(cury_prevs, cury_prevs + 8, T(0));
            for (int i1_ = ymin; i1_ <= ymax; ++i1_) {
                if (!carry) curx_prev = 0;
                for (int it = xmin; it <= xmax; ++it) {
                    int64_t l2 = L(d, i1 - 3 + it, j1 - 3 + i1_, k1 - 3 + i2);
                    curx_add = q * ((Sx2[it] - Sx1[it]) *
                                    (Sy1[i1_] * Sz1[i2] + half * (Sy2[i1_] - Sy1[i1_]) * Sz1[i2] +
                                     half * Sy1[i1_] * (Sz2[i2] - Sz1[i2]) +
                                     third * (Sy2[i1_] - Sy1[i1_]) * (Sz2[i2] - Sz1[i2]))) +
                               curx_prev;
                    cury_adds[it] = q * ((Sy2[i1_] - Sy1[i1_]) *
                                         (Sx1[it] * Sz1[i2] + half * (Sx2[it] - Sx1[it]) * Sz1[i2] +
                                          half * Sx1[it] * (Sz2[i2] - Sz1[i2]) +
                                          third * (Sx2[it] - Sx1[it]) * (Sz2[i2] - Sz1[i2]))) +
                                    cury_prevs[it];
                    curz_adds[it][i1_] = q * ((Sz2[i2] - Sz1[i2]) *
                                              (Sx1[it] * Sy1[i1_] + half * (Sx2[it] - Sx1[it]) * Sy1[i1_] +
                                               half * Sx1[it] * (Sy2[i1_] - Sy1[i1_]) +
                                               third * (Sx2[it] - Sx1[it]) * (Sy2[i1_] - Sy1[i1_]))) +
                                         curz_prevs[it][i1_];
                    cx[l2] = cx[l2] + curx_add;
                    cy[l2] = cy[l2] + cury_adds[it];
                    cz[l2] = cz[l2] + curz_adds[it][i1_];
                    curx_prev = curx_add;
                }
                curx_add = 0;
                std::copy(cury_adds, cury_adds + 8, cury_prevs);
                std::fill(cury_adds, cury_adds + 8, T(0));
            }
            std::memcpy(curz_prevs, curz_adds, sizeof(curz_adds));
            std::memset(curz_adds, 0, sizeof(curz_adds));
        }
    } else {
        T curx_add = 0, curx_prev = 0;
        T cury_adds[8] = {0}, cury_prevs[8] = {0};
        for (int i1_ = ymin; i1_ <= ymax; ++i1_) {
            if (!carry) curx_prev = 0;
            for (int it = xmin; it <= xmax; ++it) {
                int64_t l2 = L(d, i1 - 3 + it, j1 - 3 + i1_, 1);
                curx_add = q * ((Sx2[it] - Sx1[it]) * (Sy1[i1_] + half * (Sy2[i1_] - Sy1[i1_]))) + curx_prev;
                cury_adds[it] = q * ((Sy2[i1_] - Sy1[i1_]) * (Sx1[it] + half * (Sx2[it] - Sx1[it]))) + cury_prevs[it];
                T curz_add = T(-1) * q * deltaz *
                             (Sx1[it] * Sy1[i1_] + half * (Sx2[it] - Sx1[it]) * Sy1[i1_] +
                              half * Sx1[it] * (Sy2[i1_] - Sy1[i1_]) +
                              third * (Sx2[it] - Sx1[it]) * (Sy2[i1_] - Sy1[i1_]));
                cx[l2] = cx[l2] + curx_add;
                cy[l2] = cy[l2] + cury_adds[it];
                cz[l2] = cz[l2] + curz_add;
                curx_prev = curx_add;
            }
            curx_add = 0;
            std::copy(cury_adds, cury_adds + 8, cury_prevs);
            std::fill(cury_adds, cury_adds + 8, T(0));
        }
    }
}

/* deposit_particles: code/particles_movedeposit.F90:1281-1710 (one species):
 * old position from the new momentum, Esirkepov deposit, periodic wrap.
 * Particle order is kept (the reference's swap-compaction and z-crosser
 * re-append only permute particles).                                   */
template <typename T>
void deposit(const orc_cfg* cfg, T* j3, T* prt, int64_t n, T qsp)
{
    Dims d = dims_of(cfg);
    const T c = (T)cfg->c;
    T *cx = j3, *cy = j3 + d.ncell, *cz = j3 + 2 * d.ncell;
    T *X = prt, *Y = prt + n, *Z = prt + 2 * n, *U = prt + 3 * n, *V = prt + 4 * n, *W = prt + 5 * n,
      *CH = prt + 6 * n;
    for (int64_t p = 0; p < n; ++p) {
        T invgam = T(1) / std::sqrt(T(1) + U[p] * U[p] + V[p] * V[p] + W[p] * W[p]);
        T x0 = X[p] - U[p] * invgam * c;
        T y0 = Y[p] - V[p] * invgam * c;
        T z0 = Z[p] - W[p] * invgam * c;
        T q = CH[p] * qsp;
        densdecomp<T>(cfg, d, cx, cy, cz, q, X[p], Y[p], Z[p], x0, y0, z0);
    }
    /* wrap: :1359-1374, 1553-1633 */
    const T maxx = (T)d.mx - T(1) * (T)(d.ng / 2), minx = T(1) * (T)(d.ng / 2 + 1);
    const T maxy = (T)d.my - T(1) * (T)(d.ng / 2), miny = T(1) * (T)(d.ng / 2 + 1);
    T maxz, minz;
    if (d.dim == 3) { maxz = (T)d.mz - T(1) * (T)(d.ngz / 2); minz = T(1) * (T)(d.ngz / 2 + 1); }
    else { minz = T(1) * (T)(d.ngz / 2 + 1); maxz = T(1) * (T)(d.ngz / 2 + 1) + T(1); }
    const T midx = T(0.5) * (maxx - minx), midy = T(0.5) * (maxy - miny), midz = T(0.5) * (maxz - minz);
    for (int64_t p = 0; p < n; ++p) {
        T perx = 0, pery = 0, perz = 0;
        if (X[p] < minx || X[p] > maxx) perx = fsign(midx, X[p] - minx) + fsign(midx, X[p] - maxx);
        if (Y[p] < miny || Y[p] > maxy) pery = fsign(midy, Y[p] - miny) + fsign(midy, Y[p] - maxy);
        if (Z[p] < minz || Z[p] > maxz) perz = fsign(midz, Z[p] - minz) + fsign(midz, Z[p] - maxz);
        X[p] = X[p] - perx;
        Y[p] = Y[p] - pery;
        if (d.dim == 3 && perz < 0) perz = -((T)d.mz - T(1) * (T)d.ngz);
        Z[p] = Z[p] - perz;
    }
}

/* charge density at integer nodes with the deposit's own shape (S1 of the
 * densdecomp routines), accumulated in double on the interior grid with
 * periodic wrap.  Not a reference routine: the continuity check of the
 * Esirkepov deposit uses it (rho_new - rho_old = div_b cur).            */
template <typename T>
void rho(const orc_cfg* cfg, const T* prt, int64_t n, T qsp, double* out)
{
    Dims d = dims_of(cfg);
    const T *X = prt, *Y = prt + n, *Z = prt + 2 * n, *CH = prt + 6 * n;
    for (int64_t p = 0; p < n; ++p) {
        T Sx[8] = {0}, Sy[8] = {0}, Sz[8] = {0};
        int xl, xh, yl, yh, zl = 3, zh = 3;
        int i1 = aint_(X[p]), j1 = aint_(Y[p]), k1 = aint_(Z[p]);
        shape<T>(d.ord, X[p] - (T)i1, 0, Sx, xl, xh);
        shape<T>(d.ord, Y[p] - (T)j1, 0, Sy, yl, yh);
        if (d.dim == 3) shape<T>(d.ord, Z[p] - (T)k1, 0, Sz, zl, zh);
        else Sz[3] = T(1);
        double q = (double)CH[p] * (double)qsp;
        for (int c3 = zl; c3 <= zh; ++c3)
            for (int c2 = yl; c2 <= yh; ++c2)
                for (int c1 = xl; c1 <= xh; ++c1) {
                    /* interior index (0-based) with periodic wrap */
                    int gi = i1 - 3 + c1 - (d.h + 1), gj = j1 - 3 + c2 - (d.h + 1);
                    int gk = (d.dim == 3) ? k1 - 3 + c3 - (d.hz + 1) : 0;
                    gi = ((gi % d.nx) + d.nx) % d.nx;
                    gj = ((gj % d.ny) + d.ny) % d.ny;
                    gk = ((gk % d.nz) + d.nz) % d.nz;
                    out[gi + (int64_t)d.nx * (gj + (int64_t)d.ny * gk)] +=
                        q * (double)Sx[c1] * (double)Sy[c2] * (double)Sz[c3];
                }
    }
}

/* tristanmainloop.F90:107-274 for a periodic box without user hooks     */
template <typename T>
void step(const orc_cfg* cfg, T* f6, T* j3, T* pi, int64_t ni, T* pe, int64_t ne,
          T qmi, T qme, T qi, T qe, int nsteps)
{
    Dims d = dims_of(cfg);
    auto fill6 = [&](int first) {  /* bc_b1 (first=3) / bc_e1 (first=0) */
        for (int q = first; q < first + 3; ++q) ghost_fill_one(d, f6 + q * d.ncell);
    };
    for (int s = 0; s < nsteps; ++s) {
        fill6(3); fill6(0);
        bhalf<T>(cfg, f6);
        fill6(3);
        mover<T>(cfg, f6, pi, ni, qmi);
        mover<T>(cfg, f6, pe, ne, qme);
        bhalf<T>(cfg, f6);
        fill6(3); fill6(3);
        efull<T>(cfg, f6);
        fill6(0);
        std::fill(j3, j3 + 3 * d.ncell, T(0));
        fill6(0); fill6(3);
        deposit<T>(cfg, j3, pi, ni, qi);
        deposit<T>(cfg, j3, pe, ne, qe);
        fold<T>(cfg, j3);
        filter<T>(cfg, j3, cfg->ntimes);
        for (int q = 0; q < 3; ++q) {
            T* e = f6 + q * d.ncell;
            const T* jj = j3 + q * d.ncell;
            for (int64_t o = 0; o < d.ncell; ++o) e[o] = e[o] + jj[o];
        }
    }
}

template <typename T>
void normalize(T c, T c_omp, T gamma0, T ppc0, T me, T mi, T* out)
{
    /* code/particles.F90:206-235 */
    if (gamma0 < T(1)) gamma0 = std::sqrt(T(1) / (T(1) - gamma0 * gamma0));
    T omp = c / c_omp;
    T qe = -(omp * omp * gamma0) / ((ppc0 * T(.5)) * (T(1) + me / mi));
    T qi = -qe;
    me = me * std::fabs(qi);
    mi = mi * std::fabs(qi);
    out[0] = qe; out[1] = qi; out[2] = me; out[3] = mi; out[4] = qe / me; out[5] = qi / mi;
}

}  // namespace

extern "C" {

void orc_extent(const orc_cfg* cfg, int* mx, int* my, int* mz, int* nghost, int* nghostz)
{
    Dims d = dims_of(cfg);
    *mx = d.mx; *my = d.my; *mz = d.mz; *nghost = d.ng; *nghostz = d.ngz;
}

void orc_normalize_f32(float c, float c_omp, float gamma0, float ppc0, float me, float mi, float out[6])
{ normalize<float>(c, c_omp, gamma0, ppc0, me, mi, out); }
void orc_normalize_f64(double c, double c_omp, double gamma0, double ppc0, double me, double mi, double out[6])
{ normalize<double>(c, c_omp, gamma0, ppc0, me, mi, out); }

#define ORC_DEF(T, S)                                                                          \
void orc_ghost_fill_##S(const orc_cfg* cfg, T* a3) {                                          \
    Dims d = dims_of(cfg); for (int q = 0; q < 3; ++q) ghost_fill_one(d, a3 + q * d.ncell); } \
void orc_bhalf_##S(const orc_cfg* cfg, T* f6) { bhalf<T>(cfg, f6); }                         \
void orc_efull_##S(const orc_cfg* cfg, T* f6) { efull<T>(cfg, f6); }                         \
void orc_add_current_##S(const orc_cfg* cfg, T* f6, const T* j3) {                            \
    Dims d = dims_of(cfg);                                                                    \
    for (int64_t o = 0; o < 3 * d.ncell; ++o) f6[o] = f6[o] + j3[o]; }                        \
void orc_fold_##S(const orc_cfg* cfg, T* j3) { fold<T>(cfg, j3); }                           \
void orc_filter_##S(const orc_cfg* cfg, T* j3, int nt) { filter<T>(cfg, j3, nt); }           \
void orc_mover_##S(const orc_cfg* cfg, const T* f6, T* prt, int64_t n, T qm)                  \
{ mover<T>(cfg, f6, prt, n, qm); }                                                            \
void orc_deposit_##S(const orc_cfg* cfg, T* j3, T* prt, int64_t n, T q)                       \
{ deposit<T>(cfg, j3, prt, n, q); }                                                           \
void orc_step_##S(const orc_cfg* cfg, T* f6, T* j3, T* pi, int64_t ni, T* pe, int64_t ne,     \
                  T qmi, T qme, T qi, T qe, int nsteps)                                       \
{ step<T>(cfg, f6, j3, pi, ni, pe, ne, qmi, qme, qi, qe, nsteps); }                           \
void orc_rho_##S(const orc_cfg* cfg, const T* prt, int64_t n, T q, double* out)               \
{ rho<T>(cfg, prt, n, q, out); }

ORC_DEF(float, f32)
ORC_DEF(double, f64)

}  // extern "C"
