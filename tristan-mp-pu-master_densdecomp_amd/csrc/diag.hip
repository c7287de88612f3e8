// diag.hip -- FP64 diagnostics on the device and the plasma loader.
//
// Energies and the charge-continuity residual are new (the reference's
// energy() at code/tristanmainloop.F90:384-494 is never called); the charge
// density uses the deposit's own B-spline at integer nodes, so the
// Esirkepov identity rho_new - rho_old = div_b(cur) holds to round-off.
// The loader follows the reference's inject_plasma_region/maxwell_dist
// (code/particles.F90:2549-2938): uniform positions per cell, a tabulated
// Maxwell-Juttner speed distribution (init_maxw_table) and Zenitani's
// flipping method for the drift.
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "pic_math.h"
#include "tmpg_internal.h"

namespace tmpg {
namespace {

__device__ __forceinline__ int wrapd(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// all live particles of a species: tile segments of a, then the overflow list
struct SlotView {
    PBuf a, oa;
    const int32_t* cnt;
    const int32_t* ocnt;
    int32_t tcap;
    int64_t slots, ocap;
    __host__ __device__ int64_t total() const { return slots + ocap; }
    // returns the buffer and index of slot s, or false when s is empty
    __device__ __forceinline__ bool at(int64_t s, const PBuf*& b, int64_t& idx) const
    {
        if (s < slots) {
            const int64_t t = s / tcap;
            const int l = (int)(s - t * tcap);
            if (l >= min(cnt[t], tcap)) return false;
            b = &a;
            idx = s;
            return true;
        }
        const int64_t o = s - slots;
        if (o >= min((int64_t)ocnt[0], ocap)) return false;
        b = &oa;
        idx = o;
        return true;
    }
};

SlotView slot_view(const Species& sp)
{
    SlotView v;
    v.a = sp.a;
    v.oa = sp.oa;
    v.cnt = sp.cnt_a;
    v.ocnt = sp.ocnt;
    v.tcap = sp.tcap;
    v.slots = sp.slots;
    v.ocap = sp.ocap;
    return v;
}

template <int BS>
__device__ __forceinline__ double block_sum(double v, double* sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < BS / 64; ++k) t += sh[k];
    return t;
}

__global__ void __launch_bounds__(256) k_sumsq(const float* __restrict__ a, const float* __restrict__ b,
                                               const float* __restrict__ c, int64_t n, double* out)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        double x = a[t], y = b[t], z = c[t];
        s += x * x + y * y + z * z;
    }
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) atomicAdd(out, 0.5 * s);
}

__global__ void __launch_bounds__(256) k_kinetic(SlotView sv, double mc2, double* out)
{
    __shared__ double sh[4];
    double s = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < sv.total();
         t += (int64_t)gridDim.x * blockDim.x) {
        const PBuf* b;
        int64_t i;
        if (!sv.at(t, b, i)) continue;
        double uu = (double)b->u[i] * b->u[i] + (double)b->v[i] * b->v[i] + (double)b->w[i] * b->w[i];
        double gm1 = uu / (sqrt(1.0 + uu) + 1.0);
        s += (b->ch ? (double)b->ch[i] : 1.0) * gm1;
    }
    s = block_sum<256>(s, sh);
    if (threadIdx.x == 0) atomicAdd(out, mc2 * s);
}

template <int DIM, int ORD>
__global__ void __launch_bounds__(256) k_rho(SlotView sv, PackInfo pk, int nx, int ny, int nz, double q,
                                             double* __restrict__ rho)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const PBuf* bp;
    int64_t p;
    if (s >= sv.total() || !sv.at(s, bp, p)) return;
    const PBuf& B = *bp;
    int i, j, k;
    unpack_cell(pk, B.cell[p], i, j, k);
    using P = Piece<ORD>;
    float wx[P::cnt], wy[P::cnt], wz[P::cnt];
    const float ox = B.ox[p], oy = B.oy[p];
    const int mx = piece_of<ORD>(ox), my = piece_of<ORD>(oy);
    P::w(ox - (float)mx, wx);
    P::w(oy - (float)my, wy);
    int mz = 0;
    if (DIM == 3) {
        const float oz = B.oz[p];
        mz = piece_of<ORD>(oz);
        P::w(oz - (float)mz, wz);
    }
    const double qq = B.ch ? q * (double)B.ch[p] : q;
    const int nzr = (DIM == 3) ? P::cnt : 1;
    for (int c = 0; c < nzr; ++c) {
        int kk = (DIM == 3) ? wrapd(k + mz + P::first + c, nz) : 0;
        double wzz = (DIM == 3) ? (double)wz[c] : 1.0;
        for (int b = 0; b < P::cnt; ++b) {
            int jj = wrapd(j + my + P::first + b, ny);
            for (int a = 0; a < P::cnt; ++a) {
                int ii = wrapd(i + mx + P::first + a, nx);
                atomicAdd(rho + ii + (int64_t)nx * (jj + (int64_t)ny * kk),
                          qq * (double)wx[a] * (double)wy[b] * wzz);
            }
        }
    }
}

// residual r = rho_new - rho_old - div_b(cur); max |r| and sum rho^2
__global__ void __launch_bounds__(256) k_continuity(const double* __restrict__ rn, const double* __restrict__ ro,
                                                    const float* __restrict__ jx, const float* __restrict__ jy,
                                                    const float* __restrict__ jz, int dim, int nx, int ny, int nz,
                                                    double* out)
{
    __shared__ double sh[4];
    const int64_t n = (int64_t)nx * ny * nz;
    double mx = 0.0, s2 = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        int i = (int)(t % nx);
        int64_t r = t / nx;
        int j = (int)(r % ny);
        int k = (int)(r / ny);
        int im = i == 0 ? nx - 1 : i - 1, jm = j == 0 ? ny - 1 : j - 1;
        double div = ((double)jx[t] - jx[im + (int64_t)nx * (j + (int64_t)ny * k)]) +
                     ((double)jy[t] - jy[i + (int64_t)nx * (jm + (int64_t)ny * k)]);
        if (dim == 3) {
            int km = k == 0 ? nz - 1 : k - 1;
            div += (double)jz[t] - jz[i + (int64_t)nx * (j + (int64_t)ny * km)];
        }
        double res = fabs(rn[t] - ro[t] - div);
        mx = res > mx ? res : mx;
        s2 += rn[t] * rn[t];
    }
    // max over the block
    for (int o = 32; o > 0; o >>= 1) {
        double v = __shfl_down(mx, o, 64);
        mx = v > mx ? v : mx;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sh[wv] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int k = 0; k < 4; ++k) m = sh[k] > m ? sh[k] : m;
        // non-negative doubles order like their bit patterns
        atomicMax((unsigned long long*)out, (unsigned long long)__double_as_longlong(m));
    }
    __syncthreads();
    s2 = block_sum<256>(s2, sh);
    if (threadIdx.x == 0) atomicAdd(out + 1, s2);
}

// one binomial (1,2,1)/4 pass along `axis` of a periodic FP64 grid
__global__ void __launch_bounds__(256) k_binomial_axis(const double* __restrict__ in, double* __restrict__ out,
                                                       int nx, int ny, int nz, int axis)
{
    const int64_t n = (int64_t)nx * ny * nz;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx);
        const int64_t r = t / nx;
        const int j = (int)(r % ny);
        const int k = (int)(r / ny);
        int64_t m, p;
        if (axis == 0) {
            m = (i == 0 ? nx - 1 : i - 1) + (int64_t)nx * (j + (int64_t)ny * k);
            p = (i == nx - 1 ? 0 : i + 1) + (int64_t)nx * (j + (int64_t)ny * k);
        } else if (axis == 1) {
            m = i + (int64_t)nx * ((j == 0 ? ny - 1 : j - 1) + (int64_t)ny * k);
            p = i + (int64_t)nx * ((j == ny - 1 ? 0 : j + 1) + (int64_t)ny * k);
        } else {
            m = i + (int64_t)nx * (j + (int64_t)ny * (k == 0 ? nz - 1 : k - 1));
            p = i + (int64_t)nx * (j + (int64_t)ny * (k == nz - 1 ? 0 : k + 1));
        }
        out[t] = 0.25 * (in[m] + in[p]) + 0.5 * in[t];
    }
}

__device__ __forceinline__ double block_max(double v, double* sh)
{
    for (int o = 32; o > 0; o >>= 1) {
        const double w = __shfl_down(v, o, 64);
        v = w > v ? w : v;
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double m = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) m = sh[k] > m ? sh[k] : m;
    __syncthreads();
    return m;
}

// Gauss's law on the nodes: r = div_b E - rho with the backward differences
// of the E update's divergence (E components staggered as the Yee grid:
// ex(i) sits at i + 1/2).  out[0] max|r|, out[1] sum rho^2, out[2] max|rho|,
// out[3] max|div_b E| (maxima of non-negative doubles via their bit
// patterns, which order alike)
__global__ void __launch_bounds__(256) k_gauss(const float* __restrict__ ex, const float* __restrict__ ey,
                                               const float* __restrict__ ez, const double* __restrict__ rho,
                                               int dim, int nx, int ny, int nz, double* out)
{
    __shared__ double sh[4];
    const int64_t n = (int64_t)nx * ny * nz;
    double mr = 0.0, s2 = 0.0, mq = 0.0, md = 0.0;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x) {
        const int i = (int)(t % nx);
        const int64_t r = t / nx;
        const int j = (int)(r % ny);
        const int k = (int)(r / ny);
        const int im = i == 0 ? nx - 1 : i - 1, jm = j == 0 ? ny - 1 : j - 1;
        double div = ((double)ex[t] - ex[im + (int64_t)nx * (j + (int64_t)ny * k)]) +
                     ((double)ey[t] - ey[i + (int64_t)nx * (jm + (int64_t)ny * k)]);
        if (dim == 3) {
            const int km = k == 0 ? nz - 1 : k - 1;
            div += (double)ez[t] - ez[i + (int64_t)nx * (j + (int64_t)ny * km)];
        }
        const double q = rho[t];
        mr = fmax(mr, fabs(div - q));
        md = fmax(md, fabs(div));
        mq = fmax(mq, fabs(q));
        s2 += q * q;
    }
    const double vals[3] = {mr, mq, md};
    const int slots[3] = {0, 2, 3};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const double m = block_max(vals[c], sh);
        if (threadIdx.x == 0) atomicMax((unsigned long long*)(out + slots[c]), (unsigned long long)__double_as_longlong(m));
    }
    s2 = block_sum<256>(s2, sh);
    if (threadIdx.x == 0) atomicAdd(out + 1, s2);
}

// ------------------------------------------------------------------ loader
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
__device__ __forceinline__ float urand(uint64_t key, uint32_t draw)
{
    uint64_t r = mix64(key ^ mix64(0x5851f42d4c957f2dull + draw));
    return ((float)(r >> 40) + 0.5f) * (1.0f / 16777216.0f);   // (0,1)
}

__global__ void __launch_bounds__(256) k_load(int32_t* cell, float* ox, float* oy, float* oz, float* u, float* v,
                                              float* w, int32_t* id, int64_t first, int64_t count, int ppc,
                                              int dim, int nx, int ny, PackInfo pk, float gdrift,
                                              const float* __restrict__ cdf_u, int ntab, uint64_t seed,
                                              int32_t id0)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int64_t c = t / ppc;
    const int i = (int)(c % nx);
    const int64_t r = c / nx;
    const int j = (int)(r % ny);
    const int k = (int)(r / ny);
    const uint64_t key = mix64(seed * 0x2545F4914F6CDD1Dull + (uint64_t)t);
    const int64_t p = first + t;
    cell[p] = pack_cell(pk, i, j, k);
    ox[p] = fminf(urand(key, 0), 0.99999994f);
    oy[p] = fminf(urand(key, 1), 0.99999994f);
    if (oz) oz[p] = fminf(urand(key, 2), 0.99999994f);   // 3D, or the 2D zigzag scheme's z offset
    // |u'| from the tabulated Maxwell-Juttner CDF (uniform grid in CDF)
    float x = urand(key, 3) * (float)(ntab - 1);
    int b = (int)x;
    if (b >= ntab - 1) b = ntab - 2;
    float fr = x - (float)b;
    float um = cdf_u[b] + fr * (cdf_u[b + 1] - cdf_u[b]);
    // isotropic direction
    float ct = 2.0f * urand(key, 4) - 1.0f;
    float st = sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    float ph = 6.2831853f * urand(key, 5);
    float ux = um * ct, uy = um * st * cosf(ph), uz = um * st * sinf(ph);
    // Zenitani's flipping, then boost along x with gamma*beta = gdrift
    if (gdrift != 0.0f) {
        float G = sqrtf(1.0f + gdrift * gdrift);
        float beta = gdrift / G;
        float gp = sqrtf(1.0f + ux * ux + uy * uy + uz * uz);
        if (-beta * ux / gp > urand(key, 6)) ux = -ux;
        ux = G * (ux + beta * gp);
    }
    u[p] = ux;
    v[p] = uy;
    w[p] = uz;
    if (id) id[p] = id0 + (int32_t)t;
}

}  // namespace

int diag_energies(tmpg_ctx* ctx, double out[4])
{
    double* d = ctx->red;
    TMPG_HIP_CHECK(ctx, hipMemsetAsync(d, 0, 4 * sizeof(double), ctx->stream));
    const int64_t n = ctx->g.n;
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(256), 0, ctx->stream, ctx->f[EX], ctx->f[EY], ctx->f[EZ], n, d);
    hipLaunchKernelGGL(k_sumsq, dim3(blocks), dim3(256), 0, ctx->stream, ctx->f[BX], ctx->f[BY], ctx->f[BZ], n,
                       d + 1);
    for (int s = 0; s < 2; ++s) {
        Species& sp = ctx->sp[s];
        if (sp.n == 0) continue;
        SlotView sv = slot_view(sp);
        int pb = (int)std::min<int64_t>((sv.total() + 255) / 256, 4096);
        double mc2 = (double)ctx->cfg.mass[s] * (double)ctx->cfg.c * (double)ctx->cfg.c;
        hipLaunchKernelGGL(k_kinetic, dim3(pb), dim3(256), 0, ctx->stream, sv, mc2, d + 2 + s);
    }
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(out, d, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TMPG_OK;
}

// zero n doubles with a kernel (no memset node, so a captured sequence
// replays it like any other kernel)
__global__ void __launch_bounds__(256) k_zero_f64(double* __restrict__ a, int64_t n)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
        a[t] = 0.0;
}

static int zero_f64(tmpg_ctx* ctx, double* a, int64_t n, bool by_kernel)
{
    if (!by_kernel) {
        TMPG_HIP_CHECK(ctx, hipMemsetAsync(a, 0, n * sizeof(double), ctx->stream));
        return TMPG_OK;
    }
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096));
    hipLaunchKernelGGL(k_zero_f64, dim3(blocks), dim3(256), 0, ctx->stream, a, n);
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    return TMPG_OK;
}

static int rho_deposit(tmpg_ctx* ctx, double* rho, bool zero_by_kernel)
{
    const Grid& g = ctx->g;
    int rc = zero_f64(ctx, rho, g.n, zero_by_kernel);
    if (rc) return rc;
    for (int s = 0; s < 2; ++s) {
        Species& sp = ctx->sp[s];
        if (sp.n == 0) continue;
        SlotView sv = slot_view(sp);
        dim3 grid((unsigned)((sv.total() + 255) / 256));
        const double q = ctx->cfg.q[s];
#define TMPG_RHO(D, O)                                                                                       \
    hipLaunchKernelGGL((k_rho<D, O>), grid, dim3(256), 0, ctx->stream, sv, ctx->pk, g.nx, g.ny, g.nz, q, rho)
        switch (g.dim * 10 + ctx->cfg.order) {
        case 21: TMPG_RHO(2, 1); break;
        case 22: TMPG_RHO(2, 2); break;
        case 23: TMPG_RHO(2, 3); break;
        case 31: TMPG_RHO(3, 1); break;
        case 32: TMPG_RHO(3, 2); break;
        case 33: TMPG_RHO(3, 3); break;
        }
#undef TMPG_RHO
        TMPG_HIP_CHECK(ctx, hipGetLastError());
    }
    return TMPG_OK;
}

int diag_rho(tmpg_ctx* ctx, double* rho) { return rho_deposit(ctx, rho, false); }

int diag_continuity(tmpg_ctx* ctx, double* maxres, double* rhonorm)
{
    const Grid& g = ctx->g;
    int rc = diag_rho(ctx, ctx->rho);
    if (rc) return rc;
    double* d = ctx->red + 4;
    TMPG_HIP_CHECK(ctx, hipMemsetAsync(d, 0, 2 * sizeof(double), ctx->stream));
    int blocks = (int)std::min<int64_t>((g.n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_continuity, dim3(blocks), dim3(256), 0, ctx->stream, ctx->rho, ctx->rho_prev, ctx->f[JX],
                       ctx->f[JY], ctx->f[JZ], g.dim, g.nx, g.ny, g.nz, d);
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    double h[2];
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(h, d, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *maxres = h[0];
    *rhonorm = std::sqrt(h[1]);
    return TMPG_OK;
}

int diag_gauss(tmpg_ctx* ctx, double* rho_a, double* rho_b, double* red, int npasses, bool zero_by_kernel,
               double out[4])
{
    const Grid& g = ctx->g;
    int rc = rho_deposit(ctx, rho_a, zero_by_kernel);
    if (rc) return rc;
    // the n binomial passes the step applies to the deposited current
    double* a = rho_a;
    double* b = rho_b;
    const int blocks = (int)std::min<int64_t>((g.n + 255) / 256, 4096);
    for (int p = 0; p < npasses; ++p)
        for (int ax = 0; ax < g.dim; ++ax) {
            hipLaunchKernelGGL(k_binomial_axis, dim3(blocks), dim3(256), 0, ctx->stream, a, b, g.nx, g.ny, g.nz, ax);
            std::swap(a, b);
        }
    double* d = red;
    rc = zero_f64(ctx, d, 4, zero_by_kernel);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gauss, dim3(blocks), dim3(256), 0, ctx->stream, ctx->f[EX], ctx->f[EY], ctx->f[EZ], a, g.dim,
                       g.nx, g.ny, g.nz, d);
    TMPG_HIP_CHECK(ctx, hipGetLastError());
    double h[4];
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(h, d, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    out[0] = h[0];
    out[1] = std::sqrt(h[1] / (double)g.n);
    out[2] = h[2];
    out[3] = h[3];
    return TMPG_OK;
}

// ---------------------------------------------- allocation-free variant
// Scratch of tmpg_gauss_residual_noalloc: two FP64 grids of at most
// kNoAllocCells cells (the charge density and its filtered copy), then the
// four reduction slots.  One copy per GPU, held until the process ends.  In
// host code `scratch` names a host-side stand-in: the device address comes
// from hipGetSymbolAddress, once per device, and only that address is used.
constexpr int64_t kNoAllocCells = (int64_t)1 << 22;
static __device__ __align__(256) double scratch[2 * kNoAllocCells + 32];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};   // device address of `scratch`, per device
}  // namespace

// the current device's address of the scratch array (looked up on the first
// call on that device, with that device current)
int noalloc_scratch(tmpg_ctx* ctx, double** out)
{
    int dev = 0;
    TMPG_HIP_CHECK(ctx, hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        ctx->err = "allocation-free scratch: device index out of range";
        return TMPG_EINVAL;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void* p = nullptr;
        TMPG_HIP_CHECK(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_dev[dev];
    return TMPG_OK;
}

int diag_gauss_noalloc(tmpg_ctx* ctx, int npasses, double out[4])
{
    if (ctx->g.n > kNoAllocCells) {
        ctx->err = "tmpg_gauss_residual_noalloc: " + std::to_string(ctx->g.n) + " cells exceed its scratch of " +
                   std::to_string(kNoAllocCells) + " cells";
        return TMPG_EINVAL;
    }
    double* base = nullptr;
    int rc = noalloc_scratch(ctx, &base);
    if (rc) return rc;
    return diag_gauss(ctx, base, base + kNoAllocCells, base + 2 * kNoAllocCells, npasses, true, out);
}

// Maxwell-Juttner speed table: inverse CDF of f(u) ~ u^2 exp(-(gamma-1)/theta)
// sampled on a uniform grid of the CDF (the reference's init_maxw_table,
// code/particles.F90:2593-2660, tabulates the same distribution)
static std::vector<float> mj_inverse_cdf(double theta, int ntab)
{
    std::vector<float> out(ntab, 0.0f);
    if (theta <= 0.0) return out;
    const int nq = 200000;
    // support: up to where the density has fallen by e^-40
    double umax = std::sqrt(std::pow(1.0 + 40.0 * theta, 2) - 1.0);
    std::vector<double> cdf(nq + 1, 0.0);
    double du = umax / nq;
    auto f = [&](double u) { double g = std::sqrt(1.0 + u * u); return u * u * std::exp(-(u * u / (g + 1.0)) / theta); };
    for (int k = 1; k <= nq; ++k) cdf[k] = cdf[k - 1] + 0.5 * du * (f((k - 1) * du) + f(k * du));
    double tot = cdf[nq];
    int k = 0;
    for (int t = 0; t < ntab; ++t) {
        double target = tot * (double)t / (ntab - 1);
        while (k < nq && cdf[k + 1] < target) ++k;
        double seg = cdf[k + 1] - cdf[k];
        double fr = seg > 0 ? (target - cdf[k]) / seg : 0.0;
        if (fr > 1.0) fr = 1.0;
        out[t] = (float)((k + fr) * du);
    }
    return out;
}

int particles_load_uniform(tmpg_ctx* ctx, int s, int ppc, float drift, float theta, uint64_t seed)
{
    Species& sp = ctx->sp[s];
    const Grid& g = ctx->g;
    const int64_t count = g.n * (int64_t)ppc;
    // existing particles first, the new ones appended, then binned together
    PBuf old;
    int64_t n0 = 0;
    int rc = particles_flatten(ctx, s, old, &n0);
    if (rc) { free_pbuf(old); return rc; }
    const bool ids = sp.has_id;
    if (ids && n0 + count > INT32_MAX) { free_pbuf(old); ctx->err = "too many particles for int32 ids"; return TMPG_EINVAL; }
    PBuf flat;
    rc = alloc_pbuf(ctx, flat, std::max<int64_t>(n0 + count, 1), false, ids);
    if (rc) { free_pbuf(old); return rc; }
    const size_t fb = (size_t)n0 * sizeof(float), ib = (size_t)n0 * sizeof(int32_t);
    hipError_t e = hipSuccess;
    if (n0 > 0) {
        // the existing particles move to the front of the new buffer
        auto cp = [&](void* d, const void* s, size_t bytes) {
            if (e == hipSuccess) e = hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, ctx->stream);
        };
        cp(flat.cell, old.cell, ib);
        cp(flat.ox, old.ox, fb);
        cp(flat.oy, old.oy, fb);
        if (flat.oz) cp(flat.oz, old.oz, fb);
        cp(flat.u, old.u, fb);
        cp(flat.v, old.v, fb);
        cp(flat.w, old.w, fb);
        if (ids) cp(flat.id, old.id, ib);
    }
    const int ntab = 4096;
    std::vector<float> tab = mj_inverse_cdf(theta, ntab);
    float* dtab = nullptr;
    if (e == hipSuccess) e = hipMalloc(&dtab, ntab * sizeof(float));
    if (e == hipSuccess) e = hipMemcpyAsync(dtab, tab.data(), ntab * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && count > 0) {
        dim3 grid((unsigned)((count + 255) / 256));
        hipLaunchKernelGGL(k_load, grid, dim3(256), 0, ctx->stream, flat.cell, flat.ox, flat.oy, flat.oz, flat.u,
                           flat.v, flat.w, flat.id, n0, count, ppc, g.dim, g.nx, g.ny, ctx->pk, drift, dtab, ntab,
                           seed * 2 + (uint64_t)s, (int32_t)n0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    else (void)hipStreamSynchronize(ctx->stream);
    if (dtab) (void)hipFree(dtab);
    free_pbuf(old);
    if (e != hipSuccess) { free_pbuf(flat); ctx->err = hipGetErrorString(e); return TMPG_EHIP; }
    rc = particles_bin(ctx, s, flat, n0 + count, false, ids);
    free_pbuf(flat);
    return rc;
}

}  // namespace tmpg
