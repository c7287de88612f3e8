// fields.hip -- Yee stencils, current filter and primal averaging on the
// periodic interior grid (HBM-bound stencils).
//
// B half step: code/fields.F90:586-728 (advance_b_halfstep)
// E step     : code/fields.F90:739-870 (advance_e_fullstep), optionally fused
//              with add_current (code/fields.F90:1372-1395): E = (E + c*curl B) + J,
//              the reference's two roundings in its order
// highorder  : code/fields.F90:1039-1361 (advance_b_halfstep_42,
//              advance_e_fullstep_42): differences over 1 and 3 cells,
//              weights 9/8 and -1/24
// filter     : code/filter.F90:8-221, (1,2,1)^d / 4^d per pass, periodic;
//              up to four passes per sweep as the separable binomial
// The reference fills ghost layers (bc_b1/bc_e1) and folds ghost currents;
// here every neighbour index wraps periodically instead.
#include "tmpg_internal.h"

namespace tmpg {
namespace {

constexpr int BX_ = 64, BY_ = 4;   // block: 64 x 4 threads, one x-row segment per wave

__device__ __forceinline__ int wrapm(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <int DIM>
__global__ void __launch_bounds__(256) k_bhalf(float* __restrict__ ex, float* __restrict__ ey,
                                               float* __restrict__ ez, float* __restrict__ bx,
                                               float* __restrict__ by, float* __restrict__ bz,
                                               int nx, int ny, int nz, float cst)
{
    int i = blockIdx.x * BX_ + threadIdx.x;
    int j = blockIdx.y * BY_ + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    int64_t o = i + (int64_t)nx * (j + (int64_t)ny * k);
    int ip = (i + 1 == nx) ? 0 : i + 1;
    int jp = (j + 1 == ny) ? 0 : j + 1;
    int64_t oi = ip + (int64_t)nx * (j + (int64_t)ny * k);
    int64_t oj = i + (int64_t)nx * (jp + (int64_t)ny * k);
    float exo = ex[o], eyo = ey[o], ezo = ez[o];
    if (DIM == 3) {
        int kp = (k + 1 == nz) ? 0 : k + 1;
        int64_t ok = i + (int64_t)nx * (j + (int64_t)ny * kp);
        bx[o] = bx[o] + cst * (ey[ok] - eyo - ez[oj] + ezo);
        by[o] = by[o] + cst * (ez[oi] - ezo - ex[ok] + exo);
        bz[o] = bz[o] + cst * (ex[oj] - exo - ey[oi] + eyo);
    } else {
        bx[o] = bx[o] + cst * (-ez[oj] + ezo);
        by[o] = by[o] + cst * (ez[oi] - ezo);
        bz[o] = bz[o] + cst * (ex[oj] - exo - ey[oi] + eyo);
    }
}

template <int DIM, bool ADDJ>
__global__ void __launch_bounds__(256) k_efull(float* __restrict__ ex, float* __restrict__ ey,
                                               float* __restrict__ ez, const float* __restrict__ bx,
                                               const float* __restrict__ by, const float* __restrict__ bz,
                                               const float* __restrict__ jx, const float* __restrict__ jy,
                                               const float* __restrict__ jz, int nx, int ny, int nz,
                                               float cst)
{
    int i = blockIdx.x * BX_ + threadIdx.x;
    int j = blockIdx.y * BY_ + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    int64_t o = i + (int64_t)nx * (j + (int64_t)ny * k);
    int im = (i == 0) ? nx - 1 : i - 1;
    int jm = (j == 0) ? ny - 1 : j - 1;
    int64_t oi = im + (int64_t)nx * (j + (int64_t)ny * k);
    int64_t oj = i + (int64_t)nx * (jm + (int64_t)ny * k);
    float bxo = bx[o], byo = by[o], bzo = bz[o];
    float nex, ney, nez;
    if (DIM == 3) {
        int km = (k == 0) ? nz - 1 : k - 1;
        int64_t ok = i + (int64_t)nx * (j + (int64_t)ny * km);
        nex = ex[o] + cst * (by[ok] - byo - bz[oj] + bzo);
        ney = ey[o] + cst * (bz[oi] - bzo - bx[ok] + bxo);
        nez = ez[o] + cst * (bx[oj] - bxo - by[oi] + byo);
    } else {
        nex = ex[o] + cst * (-bz[oj] + bzo);
        ney = ey[o] + cst * (bz[oi] - bzo);
        nez = ez[o] + cst * (bx[oj] - bxo - by[oi] + byo);
    }
    if (ADDJ) {
        nex = nex + jx[o];
        ney = ney + jy[o];
        nez = nez + jz[o];
    }
    ex[o] = nex;
    ey[o] = ney;
    ez[o] = nez;
}

__global__ void k_add(float* __restrict__ e, const float* __restrict__ j, int64_t n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) e[t] = e[t] + j[t];
}

// One binomial pass of one component: out = (1,2,1)^d/4^d * in, periodic.
// Tile of 64 x 4 (x, y) cells per block, staged with a 1-cell halo in LDS for
// the three z planes k-1, k, k+1; each output reads its 27 (9 in 2D)
// neighbours from LDS.  The sum is the reference's (code/filter.F90:105-120).
// NP binomial passes in one sweep.  On the periodic grid the pass is
// (1,2,1)/4 along every axis, so NP passes are the separable 1D binomial
// C(2NP,t)/4^NP, t = 0..2NP, along x, y and (3D) z.  A workgroup owns an
// FT_X x FT_Y column of the grid over FT_Z planes: each plane (with an NP
// halo) is staged in LDS and filtered along x then y; the z filter runs
// over a ring of the last 2NP+1 xy-filtered planes held in registers.  One
// read (plus halo) and one write per cell instead of NP of each.
constexpr int FT_X = 64, FT_Y = 32, FT_Z = 32, FT_TY = 4;   // 64 x 4 threads, 8 rows each

__device__ __forceinline__ int wrapany(int i, int n) { i %= n; return i < 0 ? i + n : i; }

template <int NP> struct Binom;
template <> struct Binom<1> { static constexpr float w[3] = {0.25f, 0.5f, 0.25f}; };
template <> struct Binom<2> { static constexpr float w[5] = {1 / 16.f, 4 / 16.f, 6 / 16.f, 4 / 16.f, 1 / 16.f}; };
template <> struct Binom<3> {
    static constexpr float w[7] = {1 / 64.f, 6 / 64.f, 15 / 64.f, 20 / 64.f, 15 / 64.f, 6 / 64.f, 1 / 64.f};
};
template <> struct Binom<4> {
    static constexpr float w[9] = {1 / 256.f,  8 / 256.f, 28 / 256.f, 56 / 256.f, 70 / 256.f,
                                   56 / 256.f, 28 / 256.f, 8 / 256.f,  1 / 256.f};
};

// the three current components of one sweep in one launch: blockIdx.z
// runs over nzb z-blocks of component 0, then of 1, then of 2
struct FilterIO {
    const float* in[3];
    float* out[3];
};

template <int DIM, int NP>
__global__ void __launch_bounds__(FT_X * FT_TY) k_filter_fused(FilterIO io, int nx, int ny, int nz, int nzb)
{
    constexpr int T = 2 * NP + 1, RX = FT_X + 2 * NP, RY = FT_Y + 2 * NP, ROWS = FT_Y / FT_TY;
    __shared__ float raw[RY][RX];
    __shared__ float xs[RY][FT_X];
    const int comp = blockIdx.z / nzb, bz = blockIdx.z - comp * nzb;
    const float* __restrict__ in = io.in[comp];
    float* __restrict__ out = io.out[comp];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * FT_X + tx;
    const int i0 = blockIdx.x * FT_X, j0 = blockIdx.y * FT_Y;
    const int k0 = (DIM == 3) ? bz * FT_Z : 0;
    const int k1 = (DIM == 3) ? min(k0 + FT_Z, nz) : 1;
    const int i = i0 + tx;
    float ring[ROWS][T] = {};
    const int kfirst = (DIM == 3) ? k0 - NP : 0, klast = (DIM == 3) ? k1 - 1 + NP : 0;
    // each thread's share of a staged plane, loaded one plane ahead into
    // registers so that the loads of plane kk+1 are in flight while plane kk
    // is filtered
    constexpr int NTH = FT_X * FT_TY, NL = (RY * RX + NTH - 1) / NTH;
    float pre[NL];
    auto fetch = [&](int kk) {
        const int64_t plane = (int64_t)nx * ny * ((DIM == 3) ? wrapany(kk, nz) : 0);
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int e = tid + u * NTH;
            if (e < RY * RX) {
                const int r = e / RX, c = e - r * RX;
                pre[u] = in[plane + wrapany(i0 + c - NP, nx) + (int64_t)nx * wrapany(j0 + r - NP, ny)];
            }
        }
    };
    fetch(kfirst);
    for (int kk = kfirst; kk <= klast; ++kk) {
#pragma unroll
        for (int u = 0; u < NL; ++u) {
            const int e = tid + u * NTH;
            if (e < RY * RX) {
                const int r = e / RX, c = e - r * RX;
                raw[r][c] = pre[u];
            }
        }
        if (kk < klast) fetch(kk + 1);
        __syncthreads();
        for (int r = ty; r < RY; r += FT_TY) {
            float v = 0.0f;
#pragma unroll
            for (int t = 0; t < T; ++t) v += Binom<NP>::w[t] * raw[r][tx + t];
            xs[r][tx] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < ROWS; ++q) {
            const int r = ty + FT_TY * q;
            float v = 0.0f;
#pragma unroll
            for (int t = 0; t < T; ++t) v += Binom<NP>::w[t] * xs[r + t][tx];
            if (DIM == 3) {
#pragma unroll
                for (int t = 0; t + 1 < T; ++t) ring[q][t] = ring[q][t + 1];
                ring[q][T - 1] = v;
            } else {
                ring[q][0] = v;
            }
        }
        const int k = (DIM == 3) ? kk - NP : 0;
        if ((DIM == 2 || kk - kfirst >= 2 * NP) && i < nx) {
#pragma unroll
            for (int q = 0; q < ROWS; ++q) {
                const int j = j0 + ty + FT_TY * q;
                if (j >= ny) continue;
                float v;
                if (DIM == 3) {
                    v = 0.0f;
#pragma unroll
                    for (int t = 0; t < T; ++t) v += Binom<NP>::w[t] * ring[q][t];
                } else {
                    v = ring[q][0];
                }
                out[i + (int64_t)nx * (j + (int64_t)ny * k)] = v;
            }
        }
        __syncthreads();   // raw and xs are rewritten by the next plane
    }
}

// primal-averaged fields for the 3D gather
// (code/particles_movedeposit.F90:982-991): Ex, Ey, Ez averaged along their
// own axis, Bx over y (and z unless Q_BXY_Z), By over x (and z), Bz over x,y.
__global__ void __launch_bounds__(256) k_primal(const float* __restrict__ ex, const float* __restrict__ ey,
                                                const float* __restrict__ ez, const float* __restrict__ bx,
                                                const float* __restrict__ by, const float* __restrict__ bz,
                                                float* __restrict__ pex, float* __restrict__ pey,
                                                float* __restrict__ pez, float* __restrict__ pbx,
                                                float* __restrict__ pby, float* __restrict__ pbz,
                                                int nx, int ny, int nz, int quirk_z)
{
    int i = blockIdx.x * BX_ + threadIdx.x;
    int j = blockIdx.y * BY_ + threadIdx.y;
    int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    int im = (i == 0) ? nx - 1 : i - 1;
    int jm = (j == 0) ? ny - 1 : j - 1;
    int km = (k == 0) ? nz - 1 : k - 1;
    auto I = [&](int a, int b, int c) { return a + (int64_t)nx * (b + (int64_t)ny * c); };
    int64_t o = I(i, j, k);
    pex[o] = 0.5f * (ex[o] + ex[I(im, j, k)]);
    pey[o] = 0.5f * (ey[o] + ey[I(i, jm, k)]);
    pez[o] = 0.5f * (ez[o] + ez[I(i, j, km)]);
    if (quirk_z) {
        pbx[o] = 0.5f * (bx[o] + bx[I(i, jm, k)]);
        pby[o] = 0.5f * (by[o] + by[I(im, j, k)]);
    } else {
        pbx[o] = 0.5f * (0.5f * (bx[o] + bx[I(i, jm, k)]) + 0.5f * (bx[I(i, j, km)] + bx[I(i, jm, km)]));
        pby[o] = 0.5f * (0.5f * (by[o] + by[I(im, j, k)]) + 0.5f * (by[I(i, j, km)] + by[I(im, j, km)]));
    }
    pbz[o] = 0.5f * (0.5f * (bz[o] + bz[I(im, j, k)]) + 0.5f * (bz[I(i, jm, k)] + bz[I(im, jm, k)]));
}

// the three current components zeroed in one launch
__global__ void k_zero3(float* __restrict__ a, float* __restrict__ b, float* __restrict__ c, int64_t n)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        a[t] = 0.0f;
        b[t] = 0.0f;
        c[t] = 0.0f;
    }
}

__global__ void k_zero(float* __restrict__ a, int64_t n)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (; t < n; t += stride) a[t] = 0.0f;
}

// 4th-order curl: d f / d x at the half point = (9/8)(f[+1] - f[0]) - (1/24)(f[+2] - f[-1])
// (forward for B, backward for E), summed in the reference's order:
// b + c1*(pair of 1-cell differences) + c2*(pair of 3-cell differences)
template <int DIM>
__global__ void __launch_bounds__(256) k_bhalf42(const float* __restrict__ ex, const float* __restrict__ ey,
                                                 const float* __restrict__ ez, float* __restrict__ bx,
                                                 float* __restrict__ by, float* __restrict__ bz, int nx, int ny,
                                                 int nz, float c1, float c2)
{
    const int i = blockIdx.x * BX_ + threadIdx.x;
    const int j = blockIdx.y * BY_ + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    auto at = [&](int ii, int jj, int kk) {
        return wrapm(ii, nx) + (int64_t)nx * (wrapm(jj, ny) + (int64_t)ny * (DIM == 3 ? wrapm(kk, nz) : 0));
    };
    const int64_t o = at(i, j, k);
    const int64_t ip1 = at(i + 1, j, k), ip2 = at(i + 2, j, k), im1 = at(i - 1, j, k);
    const int64_t jp1 = at(i, j + 1, k), jp2 = at(i, j + 2, k), jm1 = at(i, j - 1, k);
    if (DIM == 3) {
        const int64_t kp1 = at(i, j, k + 1), kp2 = at(i, j, k + 2), km1 = at(i, j, k - 1);
        bx[o] = bx[o] + c1 * (ey[kp1] - ey[o] - ez[jp1] + ez[o]) + c2 * (ey[kp2] - ey[km1] - ez[jp2] + ez[jm1]);
        by[o] = by[o] + c1 * (ez[ip1] - ez[o] - ex[kp1] + ex[o]) + c2 * (ez[ip2] - ez[im1] - ex[kp2] + ex[km1]);
        bz[o] = bz[o] + c1 * (ex[jp1] - ex[o] - ey[ip1] + ey[o]) + c2 * (ex[jp2] - ex[jm1] - ey[ip2] + ey[im1]);
    } else {
        bx[o] = bx[o] + c1 * (-ez[jp1] + ez[o]) + c2 * (-ez[jp2] + ez[jm1]);
        by[o] = by[o] + c1 * (ez[ip1] - ez[o]) + c2 * (ez[ip2] - ez[im1]);
        bz[o] = bz[o] + c1 * (ex[jp1] - ex[o] - ey[ip1] + ey[o]) + c2 * (ex[jp2] - ex[jm1] - ey[ip2] + ey[im1]);
    }
}

template <int DIM, bool ADDJ>
__global__ void __launch_bounds__(256) k_efull42(float* __restrict__ ex, float* __restrict__ ey,
                                                 float* __restrict__ ez, const float* __restrict__ bx,
                                                 const float* __restrict__ by, const float* __restrict__ bz,
                                                 const float* __restrict__ jx, const float* __restrict__ jy,
                                                 const float* __restrict__ jz, int nx, int ny, int nz, float c1,
                                                 float c2)
{
    const int i = blockIdx.x * BX_ + threadIdx.x;
    const int j = blockIdx.y * BY_ + threadIdx.y;
    const int k = blockIdx.z;
    if (i >= nx || j >= ny) return;
    auto at = [&](int ii, int jj, int kk) {
        return wrapm(ii, nx) + (int64_t)nx * (wrapm(jj, ny) + (int64_t)ny * (DIM == 3 ? wrapm(kk, nz) : 0));
    };
    const int64_t o = at(i, j, k);
    const int64_t im1 = at(i - 1, j, k), im2 = at(i - 2, j, k), ip1 = at(i + 1, j, k);
    const int64_t jm1 = at(i, j - 1, k), jm2 = at(i, j - 2, k), jp1 = at(i, j + 1, k);
    float nex, ney, nez;
    if (DIM == 3) {
        const int64_t km1 = at(i, j, k - 1), km2 = at(i, j, k - 2), kp1 = at(i, j, k + 1);
        nex = ex[o] + c1 * (by[km1] - by[o] - bz[jm1] + bz[o]) + c2 * (by[km2] - by[kp1] - bz[jm2] + bz[jp1]);
        ney = ey[o] + c1 * (bz[im1] - bz[o] - bx[km1] + bx[o]) + c2 * (bz[im2] - bz[ip1] - bx[km2] + bx[kp1]);
        nez = ez[o] + c1 * (bx[jm1] - bx[o] - by[im1] + by[o]) + c2 * (bx[jm2] - bx[jp1] - by[im2] + by[ip1]);
    } else {
        nex = ex[o] + c1 * (-bz[jm1] + bz[o]) + c2 * (-bz[jm2] + bz[jp1]);
        ney = ey[o] + c1 * (bz[im1] - bz[o]) + c2 * (bz[im2] - bz[ip1]);
        nez = ez[o] + c1 * (bx[jm1] - bx[o] - by[im1] + by[o]) + c2 * (bx[jm2] - bx[jp1] - by[im2] + by[ip1]);
    }
    if (ADDJ) {
        nex = nex + jx[o];
        ney = ney + jy[o];
        nez = nez + jz[o];
    }
    ex[o] = nex;
    ey[o] = ney;
    ez[o] = nez;
}

dim3 grid_of(const Grid& g) { return dim3((g.nx + BX_ - 1) / BX_, (g.ny + BY_ - 1) / BY_, g.nz); }

}  // namespace

int fields_bhalf(tmpg_ctx* ctx)
{
    const Grid& g = ctx->g;
    float cst = ctx->cfg.corr * (0.5f * ctx->cfg.c);
    float** f = ctx->f;
    if (ctx->cfg.highorder) {
        // code/fields.F90:1050-1052: coef1 = 9/8*corr*(c/2), coef2 = -1/24*corr*(c/2)
        const float c1 = 9.0f / 8.0f * ctx->cfg.corr * (0.5f * ctx->cfg.c);
        const float c2 = -1.0f / 24.0f * ctx->cfg.corr * (0.5f * ctx->cfg.c);
        if (g.dim == 3)
            hipLaunchKernelGGL(k_bhalf42<3>, grid_of(g), dim3(BX_, BY_), 0, ctx->stream, f[EX], f[EY], f[EZ],
                               f[BX], f[BY], f[BZ], g.nx, g.ny, g.nz, c1, c2);
        else
            hipLaunchKernelGGL(k_bhalf42<2>, grid_of(g), dim3(BX_, BY_), 0, ctx->stream, f[EX], f[EY], f[EZ],
                               f[BX], f[BY], f[BZ], g.nx, g.ny, g.nz, c1, c2);
        TMPG_HIP_CHECK(ctx, hipGetLastError());
        return TMPG_OK;
    }
    if (g.dim == 3)
        hipLaunchKernelGGL(k_bhalf<3>, grid_of(g), dim3(BX_, BY_), 0, ctx->stream, f[EX], f[EY], f[EZ], f[BX],
                           f[BY], f[BZ], g.nx, g.ny, g.nz, cst);
    else
        hipLaunchKernelGGL(k_bhalf<2>, grid_of(g), dim3(BX_, BY_), 0, ctx->stream, f[EX], f[EY], f[EZ], f[BX],
                           f[BY], f[BZ], g.nx, g.ny, g.nz, cst);
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    return TMPG_OK;
}

int fields_efull(tmpg_ctx* ctx, bool add_j)
{
    const Grid& g = ctx->g;
    float cst = ctx->cfg.corr * ctx->cfg.c;
    float** f = ctx->f;
    if (ctx->cfg.highorder) {
        // code/fields.F90:1234-1236: coef1 = 9/8*corr*c, coef2 = -1/24*corr*c
        const float c1 = 9.0f / 8.0f * ctx->cfg.corr * ctx->cfg.c;
        const float c2 = -1.0f / 24.0f * ctx->cfg.corr * ctx->cfg.c;
#define TMPG_EF42(D, A)                                                                                   \
    hipLaunchKernelGGL((k_efull42<D, A>), grid_of(g), dim3(BX_, BY_), 0, ctx->stream, f[EX], f[EY], f[EZ], \
                       f[BX], f[BY], f[BZ], f[JX], f[JY], f[JZ], g.nx, g.ny, g.nz, c1, c2)
        if (g.dim == 3) { if (add_j) TMPG_EF42(3, true); else TMPG_EF42(3, false); }
        else { if (add_j) TMPG_EF42(2, true); else TMPG_EF42(2, false); }
#undef TMPG_EF42
        TMPG_HIP_CHECK(ctx, hipGetLastError());
        return TMPG_OK;
    }
#define TMPG_EF(D, A)                                                                                   \
    hipLaunchKernelGGL((k_efull<D, A>), grid_of(g), dim3(BX_, BY_), 0, ctx->stream, f[EX], f[EY], f[EZ], \
                       f[BX], f[BY], f[BZ], f[JX], f[JY], f[JZ], g.nx, g.ny, g.nz, cst)
    if (g.dim == 3) { if (add_j) TMPG_EF(3, true); else TMPG_EF(3, false); }
    else { if (add_j) TMPG_EF(2, true); else TMPG_EF(2, false); }
#undef TMPG_EF
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    return TMPG_OK;
}

int fields_add_current(tmpg_ctx* ctx)
{
    const Grid& g = ctx->g;
    int blocks = (int)std::min<int64_t>((g.n + 255) / 256, 8192);
    for (int q = 0; q < 3; ++q)
        hipLaunchKernelGGL(k_add, dim3(blocks), dim3(256), 0, ctx->stream, ctx->f[EX + q], ctx->f[JX + q], g.n);
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    return TMPG_OK;
}

int fields_filter(tmpg_ctx* ctx, int npasses)
{
    const Grid& g = ctx->g;
    // npasses = 4a + b: a sweeps of four passes, then one of b
    int left = npasses;
    while (left > 0) {
        const int np = left >= 4 ? 4 : left;
        left -= np;
        const int nzb = g.dim == 3 ? (g.nz + FT_Z - 1) / FT_Z : 1;
        const dim3 grid((g.nx + FT_X - 1) / FT_X, (g.ny + FT_Y - 1) / FT_Y, 3 * nzb);
        FilterIO io;
        for (int q = 0; q < 3; ++q) {
            io.in[q] = ctx->f[JX + q];
            io.out[q] = ctx->tmp[q];
        }
#define TMPG_FF(D, N) hipLaunchKernelGGL((k_filter_fused<D, N>), grid, dim3(FT_X, FT_TY), 0, ctx->stream, io, \
                                         g.nx, g.ny, g.nz, nzb)
        if (g.dim == 3) {
            switch (np) { case 1: TMPG_FF(3, 1); break; case 2: TMPG_FF(3, 2); break;
                          case 3: TMPG_FF(3, 3); break; default: TMPG_FF(3, 4); }
        } else {
            switch (np) { case 1: TMPG_FF(2, 1); break; case 2: TMPG_FF(2, 2); break;
                          case 3: TMPG_FF(2, 3); break; default: TMPG_FF(2, 4); }
        }
#undef TMPG_FF
        // ping-pong: the filtered arrays become the currents, the old ones the scratch
        for (int q = 0; q < 3; ++q) std::swap(ctx->f[JX + q], ctx->tmp[q]);
    }
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    return TMPG_OK;
}

int fields_primal(tmpg_ctx* ctx)
{
    const Grid& g = ctx->g;
    float** f = ctx->f;
    float** p = ctx->prim;
    hipLaunchKernelGGL(k_primal, grid_of(g), dim3(BX_, BY_), 0, ctx->stream, f[EX], f[EY], f[EZ], f[BX], f[BY],
                       f[BZ], p[0], p[1], p[2], p[3], p[4], p[5], g.nx, g.ny, g.nz,
                       (ctx->cfg.compat & TMPG_Q_BXY_Z) ? 1 : 0);
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    return TMPG_OK;
}

int fields_zero_j(tmpg_ctx* ctx)
{
    const Grid& g = ctx->g;
    int blocks = (int)std::min<int64_t>((g.n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_zero3, dim3(blocks), dim3(256), 0, ctx->stream, ctx->f[JX], ctx->f[JY], ctx->f[JZ], g.n);
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    return TMPG_OK;
}

}  // namespace tmpg
