// denoise.hip -- the learned current denoiser: a small periodic CNN whose
// output is a vector potential A on the B grid; the current correction is
// dJ = s * curl_b(A) with the backward-difference curl of the E update
// (code/fields.F90:836-838), so div_b(dJ) = 0 and the Esirkepov charge
// continuity of the deposit is kept to round-off.
//
// Network (DIM 2): 4 layers of 3x3 conv, width 32, GELU between layers;
//         (DIM 3): 3 layers of 3x3x3 conv, width 16.
// Input: the raw deposited current J / s (3 channels, padded to 4).
//
// Convolutions are implicit GEMMs on MFMA (v_mfma_f32_16x16x32_bf16):
// activations channels-last in BF16, FP32 accumulation.  The same kernel
// computes the backward pass w.r.t. the input (flipped, transposed weights);
// the weight gradient is its own MFMA kernel reducing over pixels.
// Training (Adam) runs on the device; master weights are FP32.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "tmpg_internal.h"

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace tmpg {
namespace dn {

__host__ __device__ inline uint16_t f2bf(float f)
{
    // round to nearest even
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t b)
{
    uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

struct Geo {
    int dim, nx, ny, nz;
    int64_t npix;       // pixels per image
    int batch;
};

// periodic neighbour of pixel p (within its image) for tap t (3x3 or 3x3x3,
// offsets -1..1, x fastest)
__device__ __forceinline__ int64_t neighbour(const Geo& g, int64_t p, int tap)
{
    const int64_t img = p / g.npix;
    int64_t r = p - img * g.npix;
    int x = (int)(r % g.nx);
    r /= g.nx;
    int y = (int)(r % g.ny);
    int z = (int)(r / g.ny);
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = (g.dim == 3) ? tap / 9 - 1 : 0;
    x += dx; y += dy; z += dz;
    x = x < 0 ? x + g.nx : (x >= g.nx ? x - g.nx : x);
    y = y < 0 ? y + g.ny : (y >= g.ny ? y - g.ny : y);
    if (g.dim == 3) z = z < 0 ? z + g.nz : (z >= g.nz ? z - g.nz : z);
    return img * g.npix + x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
}

// out[p][co] = bias[co] + sum_{tap,ci} W[co][tap*cin + ci] * in[nb(p,tap)][ci]
// in: BF16 [P][cin] (cin multiple of 4); W: BF16 [cout_pad][kpad] with
// kpad = taps*cin rounded up to 32 (zeros beyond); out: FP32 [P][cout]
// (cout <= cout_pad, multiple of 16 for cout_pad).
// gelu_out (optional): BF16 gelu(out) [P][cout]; raw_out (optional): FP32.
// GELU, tanh form: 0.5 z (1 + tanh u) with u = k0 (z + k1 z^3), evaluated
// as z sigmoid(2u) = z / (1 + exp(-2u)) with the hardware exponential (a few
// instructions instead of a library tanh; its relative error, ~1e-7, is far
// below the BF16 rounding of the activations).  exp(-2u) overflows to +inf
// for large negative z, which gives -0, the limit.
__device__ __forceinline__ float gelu_sig(float z)
{
    const float k0 = 0.7978845608f, k1 = 0.044715f;
    const float u = k0 * z * (1.0f + k1 * z * z);
    return __builtin_amdgcn_rcpf(1.0f + __expf(-2.0f * u));
}
__device__ __forceinline__ float gelu(float z)
{
    return z * gelu_sig(z);
}
__device__ __forceinline__ float gelu_grad(float z)
{
    const float k0 = 0.7978845608f, k1 = 0.044715f;
    const float sg = gelu_sig(z);
    const float th = 2.0f * sg - 1.0f;    // tanh u
    return sg + 0.5f * z * (1.0f - th * th) * k0 * (1.0f + 3.0f * k1 * z * z);
}

// A persistent workgroup of 8 waves loops over tiles of 512 pixels (3D
// 16x8x4, 2D 64x8; half as many rows when the layer has 32 output channels)
// and computes all cout_pad channels of each with the gfx950 MFMA
// v_mfma_f32_16x16x32_bf16: M = 16 pixels of a row segment, N = 16 output
// channels, K = 32 of (tap, input channel).  Its weights (KS K steps of 32)
// stay in registers for all its tiles.  The input of the next tile (with a
// one-pixel periodic halo) is loaded from memory into registers while the
// MFMAs of this tile run on the LDS copy of this one; the outputs go through
// a second LDS region, so the stores are whole 16-byte units of consecutive
// pixels.  A lane of K group kq supplies 8 consecutive K values: 8 channels
// of one tap (16 bytes) for 16 or 32 input channels, two taps of 4 channels
// for 4.  The LDS pixel stride is CIN + 8 halfwords for CIN >= 16 (48 or 80
// bytes), so each group of 8 lanes of a 16-byte read covers distinct banks.
// A and B take their K values in the same lane order, so the product does
// not depend on how the instruction orders K inside a lane group.
constexpr int kConvThreads = 512;
template <int DIM, int NT>
struct ConvTile {
    static constexpr int TX = DIM == 3 ? 16 : 64, TY = NT == 1 ? 8 : 4, TZ = DIM == 3 ? 4 : 1;
    static constexpr int HX = TX + 2, HY = TY + 2, HZ = DIM == 3 ? TZ + 2 : 1;
    static constexpr int MPR = TX / 16;                                      // 16-pixel segments per row
    static constexpr int NPIX = TX * TY * TZ;
    static constexpr int SPW = NPIX / 16 / (kConvThreads / 64);              // segments per wave
};

__host__ __device__ constexpr int conv_pixel_stride(int cin) { return cin + (cin >= 16 ? 8 : 0); }
// FP32 words per pixel of the output staging (16 * NT channels + 4: the 64
// lanes of an MFMA result write fall in distinct banks)
__host__ __device__ constexpr int conv_out_stride(int nt) { return 16 * nt + 4; }
// the input region of the LDS, in halfwords (a multiple of 8)
template <int DIM, int NT, int CIN>
__host__ __device__ constexpr int conv_in_halves()
{
    return (ConvTile<DIM, NT>::HX * ConvTile<DIM, NT>::HY * ConvTile<DIM, NT>::HZ * conv_pixel_stride(CIN) + 7) / 8 * 8;
}
template <int DIM, int NT, int CIN>
__host__ __device__ constexpr size_t conv_lds_bytes()
{
    return sizeof(uint16_t) * (size_t)conv_in_halves<DIM, NT, CIN>() +
           sizeof(float) * (size_t)ConvTile<DIM, NT>::NPIX * conv_out_stride(NT);
}

__device__ __forceinline__ int wrap_once(int x, int n)
{
    x += x < 0 ? n : 0;
    if (x >= n) x -= n;
    if (x >= n) x %= n;   // a tile wider than the grid
    return x;
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ bf16x8 bf16x8_of(uint4 v)
{
    return __builtin_bit_cast(bf16x8, v);
}

// BF16 bits of x, round to nearest even (the hardware conversion; f2bf's
// rounding for finite values)
__device__ __forceinline__ uint16_t bf16_bits(float x)
{
    return __builtin_bit_cast(uint16_t, (__bf16)x);
}

template <int NT, int DIM, int CIN>   // NT = cout_pad / 16
__global__ void __launch_bounds__(kConvThreads) k_conv(Geo g, const uint16_t* __restrict__ in,
                                                       const uint16_t* __restrict__ W,
                                                       const float* __restrict__ bias, int cout,
                                                       float* __restrict__ raw_out, uint16_t* __restrict__ gelu_out,
                                                       int gelu_stride)
{
    using T = ConvTile<DIM, NT>;
    constexpr int TAPS = DIM == 3 ? 27 : 9, K = TAPS * CIN, KS = (K + 31) / 32, KPAD = KS * 32;
    constexpr int PS = conv_pixel_stride(CIN), SPW = T::SPW, NPIX = T::NPIX, OS = conv_out_stride(NT);
    constexpr int NTH = kConvThreads;
    static_assert(CIN == 4 || CIN % 8 == 0, "input channels: 4 or a multiple of 8");
    extern __shared__ uint16_t sm[];
    uint16_t* si = sm;                                                         // input tile
    float* so = reinterpret_cast<float*>(sm + conv_in_halves<DIM, NT, CIN>());  // output tile
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    bf16x8 bw[KS][NT];
#pragma unroll
    for (int s = 0; s < KS; ++s)
#pragma unroll
        for (int n = 0; n < NT; ++n)
            bw[s][n] = bf16x8_of(*reinterpret_cast<const uint4*>(W + (n * 16 + i) * KPAD + s * 32 + 8 * kq));
    const int ntx = (g.nx + T::TX - 1) / T::TX, nty = (g.ny + T::TY - 1) / T::TY;
    const int ntz = (g.nz + T::TZ - 1) / T::TZ;
    const int64_t ntiles = (int64_t)ntx * nty * ntz * g.batch;
    struct Org { int x0, y0, z0; int64_t ibase; };
    auto origin = [&](int64_t t) {
        Org o;
        o.x0 = (int)(t % ntx) * T::TX;
        t /= ntx;
        o.y0 = (int)(t % nty) * T::TY;
        t /= nty;
        o.z0 = (int)(t % ntz) * T::TZ;
        o.ibase = (t / ntz) * g.npix;
        return o;
    };
    // the haloed input tile in units of 8 channels (16 bytes), or 4 (8 bytes)
    constexpr int CU = CIN >= 8 ? 8 : 4, NU = CIN / CU;
    constexpr int NUNITS = T::HX * T::HY * T::HZ * NU, PER = (NUNITS + NTH - 1) / NTH;
    using Unit = typename std::conditional<CU == 8, u32x4, u32x2>::type;
    Unit stage[PER];
    // this thread's units, the same in every tile: halo coordinates packed
    // as hx | hy << 8 | hz << 16 (-1: none) and the LDS destination
    int hpos[PER], hdst[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int u = threadIdx.x + NTH * j;
        hpos[j] = -1;
        hdst[j] = 0;
        if (u < NUNITS) {
            const int pix = u / NU, cu = u - pix * NU;
            const int hx = pix % T::HX, hy = (pix / T::HX) % T::HY, hz = pix / (T::HX * T::HY);
            hpos[j] = hx | (hy << 8) | (hz << 16);
            hdst[j] = pix * PS + CU * cu;
        }
    }
    auto load_tile = [&](const Org& o) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (hpos[j] >= 0) {
                const int hx = hpos[j] & 0xff, hy = (hpos[j] >> 8) & 0xff, hz = hpos[j] >> 16;
                const int x = wrap_once(o.x0 + hx - 1, g.nx), y = wrap_once(o.y0 + hy - 1, g.ny);
                const int z = DIM == 3 ? wrap_once(o.z0 + hz - 1, g.nz) : 0;
                const int64_t q = o.ibase + x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
                // the unit's channels: its LDS offset within the pixel
                const int cu = (hdst[j] % PS) / CU;
                stage[j] = *reinterpret_cast<const Unit*>(in + q * CIN + CU * cu);
            }
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int j = 0; j < PER; ++j)
            if (hpos[j] >= 0) *reinterpret_cast<Unit*>(si + hdst[j]) = stage[j];
    };
    int cen[SPW];   // LDS pixel of this lane's A row in each segment
#pragma unroll
    for (int mt = 0; mt < SPW; ++mt) {
        const int m = wave * SPW + mt, row = m / T::MPR, lx = (m % T::MPR) * 16 + i;
        const int ly = row % T::TY, lz = row / T::TY;
        cen[mt] = (((DIM == 3 ? lz + 1 : 0) * T::HY + ly + 1) * T::HX + lx + 1) * PS;
    }
    auto tap_off = [](int tap) {
        const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = (DIM == 3) ? tap / 9 - 1 : 0;
        return ((dz * T::HY + dy) * T::HX + dx) * PS;
    };
    int64_t t = blockIdx.x;
    if (t >= ntiles) return;
    Org cur = origin(t);
    load_tile(cur);
    store_tile();
    __syncthreads();
    for (;;) {
        const int64_t tn = t + gridDim.x;
        const bool more = tn < ntiles;
        Org nxt = cur;
        if (more) {
            nxt = origin(tn);
            load_tile(nxt);   // in flight during the MFMAs below
        }
        f32x4 acc[SPW][NT];
#pragma unroll
        for (int mt = 0; mt < SPW; ++mt)
#pragma unroll
            for (int n = 0; n < NT; ++n) acc[mt][n] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = s * 32 + 8 * kq;    // this lane's first K value
            // LDS offsets of this lane's K values (out-of-range K reads tap 0
            // and is masked to zero)
            int off0, off1 = 0;
            bool v0, v1 = false;
            if constexpr (CIN >= 8) {
                v0 = k < K;
                const int tap = v0 ? k / CIN : 0;
                off0 = tap_off(tap) + (v0 ? k - tap * CIN : 0);
            } else {
                const int t0 = k / 4, t1 = t0 + 1;
                v0 = t0 < TAPS;
                v1 = t1 < TAPS;
                off0 = tap_off(v0 ? t0 : 0);
                off1 = tap_off(v1 ? t1 : 0);
            }
#pragma unroll
            for (int mt = 0; mt < SPW; ++mt) {
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                if constexpr (CIN >= 8) {
                    if (v0) v = *reinterpret_cast<const uint4*>(si + cen[mt] + off0);
                } else {
                    if (v0) { const uint2 h = *reinterpret_cast<const uint2*>(si + cen[mt] + off0); v.x = h.x; v.y = h.y; }
                    if (v1) { const uint2 h = *reinterpret_cast<const uint2*>(si + cen[mt] + off1); v.z = h.x; v.w = h.y; }
                }
                const bf16x8 a = bf16x8_of(v);
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[mt][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bw[s][n], acc[mt][n], 0, 0, 0);
            }
        }
        // all waves are done with the input tile and with the previous
        // tile's outputs: D[i][j] (j = lane % 16 the channel, i = 4*(lane/16)
        // + r the pixel) goes to the output region, the next input to the
        // input region
        __syncthreads();
#pragma unroll
        for (int mt = 0; mt < SPW; ++mt) {
            const int m = wave * SPW + mt;
#pragma unroll
            for (int n = 0; n < NT; ++n)
#pragma unroll
                for (int r = 0; r < 4; ++r) so[(m * 16 + 4 * kq + r) * OS + n * 16 + i] = acc[mt][n][r];
        }
        if (more) store_tile();
        __syncthreads();
        // pixel e of the tile (x fastest) -> its global index, or -1 outside the grid
        auto gpix = [&](int e) -> int64_t {
            const int x = cur.x0 + e % T::TX, y = cur.y0 + (e / T::TX) % T::TY, z = cur.z0 + e / (T::TX * T::TY);
            if (x >= g.nx || y >= g.ny || z >= g.nz) return -1;
            return cur.ibase + x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
        };
        if (gelu_out) {
            if (cout % 8 == 0 && gelu_stride % 8 == 0) {
                // 8 channels (16 bytes) per thread and unit; consecutive
                // threads take consecutive pixels of one channel group, so
                // their LDS reads (stride OS words) fall in distinct banks
                const int nu = cout / 8;
                for (int u = threadIdx.x; u < NPIX * nu; u += NTH) {
                    const int cu = u / NPIX, e = u - cu * NPIX, c0 = 8 * cu;
                    const int64_t pp = gpix(e);
                    if (pp < 0) continue;
                    uint32_t w[4];
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const float z0v = so[e * OS + c0 + 2 * h] + (bias ? bias[c0 + 2 * h] : 0.0f);
                        const float z1v = so[e * OS + c0 + 2 * h + 1] + (bias ? bias[c0 + 2 * h + 1] : 0.0f);
                        w[h] = (uint32_t)bf16_bits(gelu(z0v)) | ((uint32_t)bf16_bits(gelu(z1v)) << 16);
                    }
                    *reinterpret_cast<uint4*>(gelu_out + pp * gelu_stride + c0) = make_uint4(w[0], w[1], w[2], w[3]);
                }
            } else {
                for (int u = threadIdx.x; u < NPIX * cout; u += NTH) {
                    const int e = u / cout, co = u - e * cout;
                    const int64_t pp = gpix(e);
                    if (pp >= 0) gelu_out[pp * gelu_stride + co] = bf16_bits(gelu(so[e * OS + co] + (bias ? bias[co] : 0.0f)));
                }
            }
        }
        if (raw_out) {
            // one element per thread: consecutive threads, consecutive addresses
            for (int u = threadIdx.x; u < NPIX * cout; u += NTH) {
                const int e = u / cout, co = u - e * cout;
                const int64_t pp = gpix(e);
                if (pp >= 0) raw_out[pp * cout + co] = so[e * OS + co] + (bias ? bias[co] : 0.0f);
            }
        }
        if (!more) break;
        t = tn;
        cur = nxt;
    }
}

// The 3D convolution with 16 input and 16 output channels (the hidden and
// last layers of the 3D network), register-blocked over output rows.  A
// workgroup of 4 waves computes a 16x8x4-pixel tile; wave w owns the output
// rows y = 2w, 2w+1 of all 4 z planes (8 row segments of 16 pixels).  An
// A operand (16 pixels of one input row, K = 32: the 16 channels at two x
// shifts, dx = -1 and 0, or dx = +1 and zeros) is read from LDS once and
// used by every output row of the wave within one row of it in y and z,
// with the weights of that (dy, dz): 6 LDS reads per output row instead of
// 14, for 18 MFMAs instead of 14 (the zero half of the dx = +1 operand).
// The workgroups are persistent: the 18 weight fragments (9 (dy, dz) x 2
// halves) are loaded into registers once, and the next tile's input is
// loaded into registers while this tile's MFMAs and stores run.  The
// output goes through the same LDS region after the MFMAs, for 16-byte
// stores of consecutive pixels.
struct RbTile {
    static constexpr int TX = 16, TY = 8, TZ = 4, HX = 18, HY = 10, HZ = 6, NPIX = 512;
    static constexpr int PS = 24;   // halfwords per staged pixel (16 channels + 8)
    static constexpr int OS = 20;   // FP32 words per output pixel
};
constexpr size_t conv_rb_lds_bytes()
{
    return std::max(sizeof(uint16_t) * (size_t)RbTile::HX * RbTile::HY * RbTile::HZ * RbTile::PS,
                    sizeof(float) * (size_t)RbTile::NPIX * RbTile::OS);
}

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
k_conv3_rb(Geo g, const uint16_t* __restrict__ in,
                                                  const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                                  int cout, float* __restrict__ raw_out,
                                                  uint16_t* __restrict__ gelu_out, int gelu_stride)
{
    using T = RbTile;
    constexpr int CIN = 16, KPAD = (27 * CIN + 31) / 32 * 32, PS = T::PS, OS = T::OS, NPIX = T::NPIX;
    extern __shared__ uint16_t sm[];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = lane & 15, kq = lane >> 4;
    // weights: fragment [dy+1][dz+1][h]; lane group kq < 2 holds dx = -1
    // (h = 0) or +1 (h = 1), kq >= 2 dx = 0 (h = 0) or zeros (h = 1);
    // channels 8 * (kq & 1) .. +7; tap = (dz+1)*9 + (dy+1)*3 + (dx+1)
    bf16x8 bw[3][3][2];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                uint4 v = make_uint4(0u, 0u, 0u, 0u);
                const int dxs = kq < 2 ? (h == 0 ? 0 : 2) : (h == 0 ? 1 : -1);   // dx + 1, -1: zeros
                if (dxs >= 0) {
                    const int tap = c * 9 + a * 3 + dxs;
                    v = *reinterpret_cast<const uint4*>(W + i * KPAD + tap * CIN + 8 * (kq & 1));
                }
                bw[a][c][h] = bf16x8_of(v);
            }
    const int ntx = (g.nx + T::TX - 1) / T::TX, nty = (g.ny + T::TY - 1) / T::TY;
    const int ntz = (g.nz + T::TZ - 1) / T::TZ;
    const int64_t ntiles = (int64_t)ntx * nty * ntz * g.batch;
    struct Org { int x0, y0, z0; int64_t ibase; };
    auto origin = [&](int64_t t) {
        Org o;
        o.x0 = (int)(t % ntx) * T::TX;
        t /= ntx;
        o.y0 = (int)(t % nty) * T::TY;
        t /= nty;
        o.z0 = (int)(t % ntz) * T::TZ;
        o.ibase = (t / ntz) * g.npix;
        return o;
    };
    // the haloed input tile in 16-byte units (8 channels); a thread's units
    // are the same in every tile: halo coordinates packed hx | hy<<8 | hz<<16
    constexpr int NUNITS = T::HX * T::HY * T::HZ * 2, PER = (NUNITS + 255) / 256;
    u32x4 stage[PER];
    int hpos[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
        const int u = threadIdx.x + 256 * j;
        hpos[j] = -1;
        if (u < NUNITS) {
            const int pix = u >> 1;
            hpos[j] = (pix % T::HX) | (((pix / T::HX) % T::HY) << 8) | ((pix / (T::HX * T::HY)) << 16);
        }
    }
    auto load_tile = [&](const Org& o) {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            if (hpos[j] >= 0) {
                const int u = threadIdx.x + 256 * j;
                const int hx = hpos[j] & 0xff, hy = (hpos[j] >> 8) & 0xff, hz = hpos[j] >> 16;
                const int x = wrap_once(o.x0 + hx - 1, g.nx), y = wrap_once(o.y0 + hy - 1, g.ny);
                const int z = wrap_once(o.z0 + hz - 1, g.nz);
                const int64_t q = o.ibase + x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
                stage[j] = *reinterpret_cast<const u32x4*>(in + q * CIN + 8 * (u & 1));
            }
        }
    };
    auto store_tile = [&]() {
#pragma unroll
        for (int j = 0; j < PER; ++j) {
            const int u = threadIdx.x + 256 * j;
            if (hpos[j] >= 0) *reinterpret_cast<u32x4*>(sm + (u >> 1) * PS + 8 * (u & 1)) = stage[j];
        }
    };
    // this lane's A column: pixel i of the row, shifted by the half's dx
    const int dxl0 = kq < 2 ? -1 : 0, dxl1 = kq < 2 ? 1 : 0;
    const bool live1 = kq < 2;
    const int ci = 8 * (kq & 1);
    float* so = reinterpret_cast<float*>(sm);
    // persistent: tiles blockIdx.x, + gridDim.x, ...; the next tile's input
    // is loaded into registers while this tile's MFMAs and stores run, and
    // the one LDS region holds the input, then the output, of each tile
    int64_t t = blockIdx.x;
    if (t >= ntiles) return;
    Org cur = origin(t);
    load_tile(cur);
    store_tile();
    __syncthreads();
    for (;;) {
        const int64_t tn = t + gridDim.x;
        const bool more = tn < ntiles;
        Org nxt = cur;
        if (more) {
            nxt = origin(tn);
            load_tile(nxt);
        }
        f32x4 acc[2][4];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 4; ++c) acc[a][c] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int zi = 0; zi < 6; ++zi)          // input plane z = zi - 1 (halo coordinate zi)
#pragma unroll
            for (int ry = 0; ry < 4; ++ry) {    // input row y = 2 wave + ry - 1 (halo coordinate 2 wave + ry)
                const int rowoff = ((zi * T::HY) + 2 * wave + ry) * T::HX;
                const uint4 v0 = *reinterpret_cast<const uint4*>(sm + (rowoff + i + 1 + dxl0) * PS + ci);
                uint4 v1 = make_uint4(0u, 0u, 0u, 0u);
                if (live1) v1 = *reinterpret_cast<const uint4*>(sm + (rowoff + i + 1 + dxl1) * PS + ci);
                const bf16x8 a0 = bf16x8_of(v0), a1 = bf16x8_of(v1);
#pragma unroll
                for (int yl = 0; yl < 2; ++yl)
#pragma unroll
                    for (int zo = 0; zo < 4; ++zo) {
                        const int dy = ry - 1 - yl, dz = zi - 1 - zo;
                        if (dy < -1 || dy > 1 || dz < -1 || dz > 1) continue;
                        acc[yl][zo] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(a0, bw[dy + 1][dz + 1][0], acc[yl][zo], 0, 0, 0);
                        acc[yl][zo] =
                            __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, bw[dy + 1][dz + 1][1], acc[yl][zo], 0, 0, 0);
                    }
            }
        // outputs through the LDS region (the input is no longer read):
        // D[i][j] with j = lane % 16 the channel, i = 4*(lane/16) + r the pixel
        __syncthreads();
#pragma unroll
        for (int yl = 0; yl < 2; ++yl)
#pragma unroll
            for (int zo = 0; zo < 4; ++zo) {
                const int row = zo * T::TY + 2 * wave + yl;
#pragma unroll
                for (int r = 0; r < 4; ++r) so[(row * 16 + 4 * kq + r) * OS + i] = acc[yl][zo][r];
            }
        __syncthreads();
        auto gpix = [&](int e) -> int64_t {
            const int x = cur.x0 + e % T::TX, y = cur.y0 + (e / T::TX) % T::TY, z = cur.z0 + e / (T::TX * T::TY);
            if (x >= g.nx || y >= g.ny || z >= g.nz) return -1;
            return cur.ibase + x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
        };
        if (gelu_out) {
            if (cout % 8 == 0 && gelu_stride % 8 == 0) {
                const int nu = cout / 8;
                for (int u = threadIdx.x; u < NPIX * nu; u += 256) {
                    const int cu = u / NPIX, e = u - cu * NPIX, c0 = 8 * cu;
                    const int64_t pp = gpix(e);
                    if (pp < 0) continue;
                    uint32_t w[4];
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const float z0v = so[e * OS + c0 + 2 * h] + (bias ? bias[c0 + 2 * h] : 0.0f);
                        const float z1v = so[e * OS + c0 + 2 * h + 1] + (bias ? bias[c0 + 2 * h + 1] : 0.0f);
                        w[h] = (uint32_t)bf16_bits(gelu(z0v)) | ((uint32_t)bf16_bits(gelu(z1v)) << 16);
                    }
                    *reinterpret_cast<uint4*>(gelu_out + pp * gelu_stride + c0) = make_uint4(w[0], w[1], w[2], w[3]);
                }
            } else {
                for (int u = threadIdx.x; u < NPIX * cout; u += 256) {
                    const int e = u / cout, co = u - e * cout;
                    const int64_t pp = gpix(e);
                    if (pp >= 0)
                        gelu_out[pp * gelu_stride + co] = bf16_bits(gelu(so[e * OS + co] + (bias ? bias[co] : 0.0f)));
                }
            }
        }
        if (raw_out) {
            for (int u = threadIdx.x; u < NPIX * cout; u += 256) {
                const int e = u / cout, co = u - e * cout;
                const int64_t pp = gpix(e);
                if (pp >= 0) raw_out[pp * cout + co] = so[e * OS + co] + (bias ? bias[co] : 0.0f);
            }
        }
        if (!more) break;
        // the stores above read the region; then it takes the next input
        __syncthreads();
        store_tile();
        __syncthreads();
        t = tn;
        cur = nxt;
    }
}

// dW[co][k] += sum_p dz[p][co] * patch(in)[p][k], db[co] += sum_p dz[p][co].
// MFMA with M = co, N = k, reduction over 16 pixels per instruction; a block
// sums a chunk of pixels for all output tiles and adds to dW atomically.
__global__ void __launch_bounds__(256) k_conv_wgrad(Geo g, const uint16_t* __restrict__ dz, int dz_stride,
                                                    int cout, int cout_pad, const uint16_t* __restrict__ in, int cin,
                                                    int taps, int kpad, int64_t chunk, float* __restrict__ dW,
                                                    float* __restrict__ db)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t P = g.npix * g.batch;
    const int64_t pbeg = (int64_t)blockIdx.x * chunk;
    const int64_t pend = pbeg + chunk < P ? pbeg + chunk : P;
    const int nco = cout_pad / 16, nk = kpad / 16;
    const int ntile = nco * nk;
    const int K = taps * cin;
    const int i = lane & 15, kq = lane >> 4;
    // each wave owns tiles wave, wave+4, ...
    for (int tile = wave; tile < ntile; tile += 4) {
        const int ct = tile / nk, kt = tile % nk;
        const int co = ct * 16 + i;           // A row (M)
        const int kk = kt * 16 + i;           // B column (N)
        int tap = 0, ci = 0;
        const bool kv = kk < K;
        if (kv) { tap = kk / cin; ci = kk - tap * cin; }
        f32x4 acc = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
        for (int64_t q0 = pbeg; q0 < pend; q0 += 16) {
            bf16x4 a, b;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t p = q0 + 4 * kq + r;
                short av = 0, bv = 0;
                if (p < pend) {
                    if (co < cout) av = (short)dz[p * dz_stride + co];
                    if (kv) bv = (short)in[neighbour(g, p, tap) * cin + ci];
                }
                a[r] = av;
                b[r] = bv;
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
        // D[m][n]: n = lane % 16 (k), m = 4*(lane/16) + r (co)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = ct * 16 + 4 * kq + r;
            const int n = kt * 16 + (lane & 15);
            if (m < cout && n < K) atomicAdd(dW + (int64_t)m * kpad + n, acc[r]);
        }
    }
    // bias gradient: FP32 sum of the BF16 dz over the chunk
    if (db) {
        for (int c = threadIdx.x; c < cout; c += blockDim.x) {
            float s = 0.0f;
            for (int64_t p = pbeg; p < pend; ++p) s += bf2f(dz[p * dz_stride + c]);
            atomicAdd(db + c, s);
        }
    }
}

// dW[co][k] += sum_p dz[p][co] in[nb(p, tap(k))][ci(k)], db[co] += sum_p
// dz[p][co], on v_mfma_f32_16x16x32_bf16: M = 16 output channels, N = 16
// weight columns (tap, input channel), K = 32 pixels of a tile.  Persistent
// workgroups of 8 waves walk the forward convolution's 512-pixel tiles; per
// tile they stage the haloed input (pixel-major, as the forward does) and
// dz transposed (channel-major, so a lane's 8 pixels are one 16-byte read)
// in LDS.  Wave w owns the (co tile, column tile) pairs w, w + 8, ... (at
// most WG_MT of them) and sums them in registers over all its tiles.  Each
// workgroup then writes its sums (and its bias sums) to its own slot of
// part, [grid][cout_pad * kpad + cout_pad]; k_wgrad_reduce adds the slots
// to dW and db.  (Atomics from every workgroup onto the same few KB of dW
// serialise on their cache lines and took most of the time.)
constexpr int WG_THREADS = 512, WG_MT = 5;
template <int DIM, int CIN>
__host__ __device__ constexpr size_t wgrad_lds_bytes(int cout_pad)
{
    return sizeof(uint16_t) * ((size_t)conv_in_halves<DIM, 1, CIN>() +
                               (size_t)cout_pad * (ConvTile<DIM, 1>::NPIX + 8));
}

template <int DIM, int CIN>
__global__ void __launch_bounds__(WG_THREADS) k_wgrad(Geo g, const uint16_t* __restrict__ dz, int dz_stride,
                                                      int cout, int cout_pad, const uint16_t* __restrict__ in,
                                                      int taps, int kpad, float* __restrict__ part)
{
    using T = ConvTile<DIM, 1>;
    constexpr int NPIX = T::NPIX, PS = conv_pixel_stride(CIN), DS = NPIX + 8;
    constexpr int HALO = T::HX * T::HY * T::HZ, UPP = CIN / 4;     // 8-byte units per staged pixel
    static_assert(NPIX % 32 == 0 && T::TX % 8 == 0, "32-pixel steps of whole 8-pixel row pieces");
    extern __shared__ uint16_t sm[];
    uint16_t* sin = sm;
    uint16_t* sdz = sm + conv_in_halves<DIM, 1, CIN>();
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i = lane & 15, kq = lane >> 4;
    const int K = taps * CIN, nkt = kpad / 16, ntile = (cout_pad / 16) * nkt;
    const int ntx = (g.nx + T::TX - 1) / T::TX, nty = (g.ny + T::TY - 1) / T::TY;
    const int ntz = (g.nz + T::TZ - 1) / T::TZ;
    const int64_t ntiles = (int64_t)ntx * nty * ntz * g.batch;
    // this lane's A row (output channel) and B column (tap, input channel)
    // of each owned pair: the column's offset in the staged input, in
    // halfwords, relative to the pixel's own position
    int aco[WG_MT], boff[WG_MT];
    bool bval[WG_MT];
    f32x4 acc[WG_MT];
#pragma unroll
    for (int m = 0; m < WG_MT; ++m) {
        const int j = wave + 8 * m;
        const int ct = j / nkt, kt = j - ct * nkt;
        const int k = kt * 16 + i;
        aco[m] = ct * 16 + i;
        bval[m] = j < ntile && k < K;
        const int tap = bval[m] ? k / CIN : 0, ci = k - tap * CIN;
        const int dx = tap % 3, dy = (tap / 3) % 3, dzo = DIM == 3 ? tap / 9 : 0;   // offsets + 1
        boff[m] = (dx + T::HX * (dy + T::HY * dzo)) * PS + ci;
        acc[m] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    }
    float bsum[2] = {0.0f, 0.0f};
    for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        int64_t r = t;
        const int x0 = (int)(r % ntx) * T::TX;
        r /= ntx;
        const int y0 = (int)(r % nty) * T::TY;
        r /= nty;
        const int z0 = (int)(r % ntz) * T::TZ;
        const int64_t ibase = (r / ntz) * g.npix;
        __syncthreads();    // the previous tile's reads are done
        for (int u = tid; u < HALO * UPP; u += WG_THREADS) {
            const int hp = u / UPP, uc = u - hp * UPP;
            const int hx = hp % T::HX, hy = (hp / T::HX) % T::HY, hz = hp / (T::HX * T::HY);
            const int x = wrap_once(x0 + hx - 1, g.nx), y = wrap_once(y0 + hy - 1, g.ny);
            const int z = DIM == 3 ? wrap_once(z0 + hz - 1, g.nz) : 0;
            const int64_t q = ibase + x + (int64_t)g.nx * (y + (int64_t)g.ny * z);
            *reinterpret_cast<u32x2*>(sin + hp * PS + 4 * uc) =
                *reinterpret_cast<const u32x2*>(in + q * CIN + 4 * uc);
        }
        for (int u = tid; u < NPIX * cout_pad; u += WG_THREADS) {
            const int p = u / cout_pad, co = u - p * cout_pad;
            const int x = x0 + p % T::TX, y = y0 + (p / T::TX) % T::TY, z = z0 + p / (T::TX * T::TY);
            uint16_t v = 0;
            if (co < cout && x < g.nx && y < g.ny && z < g.nz)
                v = dz[(ibase + x + (int64_t)g.nx * (y + (int64_t)g.ny * z)) * dz_stride + co];
            sdz[co * DS + p] = v;
        }
        __syncthreads();
#pragma unroll
        for (int jb = 0; jb < 2; ++jb) {
            const int u = tid + WG_THREADS * jb;
            if (u < cout_pad * 32) {
                const int co = u >> 5, p0 = (u & 31) * (NPIX / 32);
                float s = 0.0f;
                for (int p = p0; p < p0 + NPIX / 32; ++p) s += bf2f(sdz[co * DS + p]);
                bsum[jb] += s;
            }
        }
        for (int st = 0; st < NPIX / 32; ++st) {
            // this lane's 8 pixels: one piece of a tile row
            const int p0 = st * 32 + 8 * kq;
            const int px = p0 % T::TX, py = (p0 / T::TX) % T::TY, pz = p0 / (T::TX * T::TY);
            const int hbase = (px + T::HX * (py + T::HY * pz)) * PS;
#pragma unroll
            for (int m = 0; m < WG_MT; ++m) {
                if (wave + 8 * m >= ntile) break;
                const bf16x8 a = bf16x8_of(*reinterpret_cast<const uint4*>(sdz + aco[m] * DS + p0));
                uint4 bv = make_uint4(0u, 0u, 0u, 0u);
                if (bval[m]) {
                    const uint16_t* b = sin + hbase + boff[m];
                    bv.x = (unsigned)b[0] | ((unsigned)b[PS] << 16);
                    bv.y = (unsigned)b[2 * PS] | ((unsigned)b[3 * PS] << 16);
                    bv.z = (unsigned)b[4 * PS] | ((unsigned)b[5 * PS] << 16);
                    bv.w = (unsigned)b[6 * PS] | ((unsigned)b[7 * PS] << 16);
                }
                acc[m] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bf16x8_of(bv), acc[m], 0, 0, 0);
            }
        }
    }
    // D[m][n]: n = lane % 16 (column), m = 4 (lane / 16) + r (output
    // channel); the owned pairs cover the whole [cout_pad][kpad] slot
    float* pw = part + (int64_t)blockIdx.x * ((int64_t)cout_pad * kpad + cout_pad);
#pragma unroll
    for (int m = 0; m < WG_MT; ++m) {
        const int j = wave + 8 * m;
        if (j >= ntile) break;
        const int ct = j / nkt, kt = j - ct * nkt;
        const int n = kt * 16 + i;
#pragma unroll
        for (int rr = 0; rr < 4; ++rr) pw[(ct * 16 + 4 * kq + rr) * kpad + n] = acc[m][rr];
    }
    // the 32 partial bias sums of each channel, added in LDS
    __syncthreads();
    float* sb = reinterpret_cast<float*>(sm);
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const int u = tid + WG_THREADS * jb;
        if (u < cout_pad * 32) sb[u] = bsum[jb];
    }
    __syncthreads();
    if (tid < cout_pad) {
        float sum = 0.0f;
        for (int q = 0; q < 32; ++q) sum += sb[tid * 32 + q];
        pw[cout_pad * kpad + tid] = sum;
    }
}

// dW[co][k] += sum over the nblk slots of part (k < K, co < cout), and
// db[co] += the slots' bias sums, without atomics: a block of 256 threads
// takes 16 values, 16 threads per value each summing every 16th slot, and
// adds the 16 partial sums in LDS
__global__ void __launch_bounds__(256) k_wgrad_reduce(const float* __restrict__ part, int nblk, int cout,
                                                      int cout_pad, int kpad, int K, float* __restrict__ dW,
                                                      float* __restrict__ db)
{
    __shared__ float red[16][17];
    const int E = cout_pad * kpad, stride = E + cout_pad;
    const int le = threadIdx.x & 15, sl = threadIdx.x >> 4;
    const int e = blockIdx.x * 16 + le;
    float sum = 0.0f;
    if (e < stride)
        for (int b = sl; b < nblk; b += 16) sum += part[(int64_t)b * stride + e];
    red[sl][le] = sum;
    __syncthreads();
    if (threadIdx.x >= 16) return;
    const int e2 = blockIdx.x * 16 + threadIdx.x;
    float tot = 0.0f;
#pragma unroll
    for (int q = 0; q < 16; ++q) tot += red[q][threadIdx.x];
    if (e2 < E) {
        const int co = e2 / kpad, n = e2 - co * kpad;
        if (co < cout && n < K) dW[(int64_t)co * kpad + n] += tot;
    } else if (e2 < stride && db) {
        const int co = e2 - E;
        if (co < cout) db[co] += tot;
    }
}

// x (BF16, 4 channels) = J / s from the three FP32 current arrays of a batch
__global__ void k_pack_input(const float* __restrict__ j, int64_t n, int batch, float inv_s,
                             uint16_t* __restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * batch) return;
    const int64_t b = t / n, p = t - b * n;
    const float* jb = j + b * 3 * n;
    x[t * 4 + 0] = f2bf(jb[p] * inv_s);
    x[t * 4 + 1] = f2bf(jb[n + p] * inv_s);
    x[t * 4 + 2] = f2bf(jb[2 * n + p] * inv_s);
    x[t * 4 + 3] = 0;
}

// the same from the context's three current arrays (one image)
__global__ void k_pack_input3(const float* __restrict__ jx, const float* __restrict__ jy,
                              const float* __restrict__ jz, int64_t n, float inv_s, uint16_t* __restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    x[t * 4 + 0] = f2bf(jx[t] * inv_s);
    x[t * 4 + 1] = f2bf(jy[t] * inv_s);
    x[t * 4 + 2] = f2bf(jz[t] * inv_s);
    x[t * 4 + 3] = 0;
}

// backward-difference curl on the periodic grid: c = curl_b(A), A [P][3]
__global__ void k_curl(Geo g, const float* __restrict__ A, float* __restrict__ c)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.npix * g.batch) return;
    // taps: x-1 -> 3 (dz=0,dy=0,dx=-1 -> tap 1+3+9?)  computed explicitly
    const int64_t img = p / g.npix;
    int64_t r = p - img * g.npix;
    const int x = (int)(r % g.nx);
    r /= g.nx;
    const int y = (int)(r % g.ny);
    const int z = (int)(r / g.ny);
    auto at = [&](int xx, int yy, int zz) {
        xx = xx < 0 ? xx + g.nx : xx;
        yy = yy < 0 ? yy + g.ny : yy;
        zz = zz < 0 ? zz + g.nz : zz;
        return img * g.npix + xx + (int64_t)g.nx * (yy + (int64_t)g.ny * zz);
    };
    const int64_t px = at(x - 1, y, z), py = at(x, y - 1, z);
    float cx = A[p * 3 + 2] - A[py * 3 + 2];
    float cy = -(A[p * 3 + 2] - A[px * 3 + 2]);
    float cz = (A[p * 3 + 1] - A[px * 3 + 1]) - (A[p * 3 + 0] - A[py * 3 + 0]);
    if (g.dim == 3) {
        const int64_t pz = at(x, y, z - 1);
        cx -= A[p * 3 + 1] - A[pz * 3 + 1];
        cy += A[p * 3 + 0] - A[pz * 3 + 0];
    }
    c[p * 3 + 0] = cx;
    c[p * 3 + 1] = cy;
    c[p * 3 + 2] = cz;
}

// adjoint of k_curl: dA = curl_b^T(gc)
__global__ void k_curl_adj(Geo g, const float* __restrict__ gc, float* __restrict__ dA)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= g.npix * g.batch) return;
    const int64_t img = p / g.npix;
    int64_t r = p - img * g.npix;
    const int x = (int)(r % g.nx);
    r /= g.nx;
    const int y = (int)(r % g.ny);
    const int z = (int)(r / g.ny);
    auto at = [&](int xx, int yy, int zz) {
        xx = xx >= g.nx ? xx - g.nx : xx;
        yy = yy >= g.ny ? yy - g.ny : yy;
        zz = zz >= g.nz ? zz - g.nz : zz;
        return img * g.npix + xx + (int64_t)g.nx * (yy + (int64_t)g.ny * zz);
    };
    const int64_t qx = at(x + 1, y, z), qy = at(x, y + 1, z);
    // cx(p) = Az(p) - Az(p-y) [- Ay(p) + Ay(p-z)]; cy(p) = -Az(p) + Az(p-x) [+ Ax(p) - Ax(p-z)];
    // cz(p) = Ay(p) - Ay(p-x) - Ax(p) + Ax(p-y)
    float ax = -gc[p * 3 + 2] + gc[qy * 3 + 2];
    float ay = gc[p * 3 + 2] - gc[qx * 3 + 2];
    float az = gc[p * 3 + 0] - gc[qy * 3 + 0] - gc[p * 3 + 1] + gc[qx * 3 + 1];
    if (g.dim == 3) {
        const int64_t qz = at(x, y, z + 1);
        ay += -gc[p * 3 + 0] + gc[qz * 3 + 0];
        ax += gc[p * 3 + 1] - gc[qz * 3 + 1];
    }
    dA[p * 3 + 0] = ax;
    dA[p * 3 + 1] = ay;
    dA[p * 3 + 2] = az;
}

// per image b: dots[2b] = <d, t>, dots[2b+1] = <t, t> with d = pred - t
// (FP64 sums); a wave's pixels all belong to one image when n % 64 == 0,
// else each lane adds its own terms
__global__ void k_dots(int64_t n, int batch, const float* __restrict__ jlo, const float* __restrict__ jt,
                       float inv_s, const float* __restrict__ curl, double* dots)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double dt = 0.0, tt = 0.0;
    int64_t b = 0;
    if (t < n * batch) {
        b = t / n;
        const int64_t p = t - b * n;
        for (int c = 0; c < 3; ++c) {
            const float pred = jlo[(b * 3 + c) * n + p] * inv_s + curl[t * 3 + c];
            const float tv = jt[(b * 3 + c) * n + p] * inv_s;
            const float d = pred - tv;
            dt += (double)d * tv;
            tt += (double)tv * tv;
        }
    }
    if (n % 64 == 0) {
        for (int o = 32; o > 0; o >>= 1) {
            dt += __shfl_down(dt, o, 64);
            tt += __shfl_down(tt, o, 64);
        }
        if ((threadIdx.x & 63) == 0 && t < n * batch) {
            atomicAdd(dots + 2 * b, dt);
            atomicAdd(dots + 2 * b + 1, tt);
        }
    } else if (t < n * batch) {
        atomicAdd(dots + 2 * b, dt);
        atomicAdd(dots + 2 * b + 1, tt);
    }
}

// pred = x + curl(A), d = pred - t; loss += sum d^2 / N, g = 2 d / N; with
// the amplitude term (lam > 0) the image's loss gains lam <d,t>^2/(N <t,t>)
// and g gains 2 lam (<d,t>/<t,t>) t / N.
// x, t FP32 [batch][3][n] (the FP32 currents / s); curl [P][3]
__global__ void k_loss(int64_t n, int batch, const float* __restrict__ jlo, const float* __restrict__ jt,
                       float inv_s, const float* __restrict__ curl, float* __restrict__ grad, double* loss,
                       double invN, float lam, const double* __restrict__ dots)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (t < n * batch) {
        const int64_t b = t / n, p = t - b * n;
        float bias = 0.0f;
        if (lam > 0.0f && dots[2 * b + 1] > 0.0) bias = (float)((double)lam * dots[2 * b] / dots[2 * b + 1]);
        for (int c = 0; c < 3; ++c) {
            const float pred = jlo[(b * 3 + c) * n + p] * inv_s + curl[t * 3 + c];
            const float tv = jt[(b * 3 + c) * n + p] * inv_s;
            const float d = pred - tv;
            s += (double)d * d;
            if (grad) grad[t * 3 + c] = (float)(2.0 * invN) * (d + bias * tv);
        }
        if (p == 0 && lam > 0.0f && dots[2 * b + 1] > 0.0)
            s += (double)lam * dots[2 * b] * dots[2 * b] / dots[2 * b + 1];
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, s * invN);
}

// dz = dh * gelu'(z) in BF16 (dh FP32 [P][c], z FP32 [P][c])
__global__ void k_gelu_back(int64_t m, const float* __restrict__ dh, const float* __restrict__ z,
                            uint16_t* __restrict__ dz)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    dz[t] = f2bf(dh[t] * gelu_grad(z[t]));
}

__global__ void k_to_bf16(int64_t m, const float* __restrict__ a, int ca, uint16_t* __restrict__ b, int cb)
{
    // a [P][ca] -> b [P][cb] (cb >= ca, padded with zeros)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t P = m;
    if (t >= P * cb) return;
    const int64_t p = t / cb;
    const int c = (int)(t - p * cb);
    b[t] = c < ca ? f2bf(a[p * ca + c]) : (uint16_t)0;
}

// J += s * curl(A) on the library's three current arrays (batch 1)
__global__ void k_apply(int64_t n, const float* __restrict__ curl, float s, float* jx, float* jy, float* jz)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    jx[p] += s * curl[p * 3 + 0];
    jy[p] += s * curl[p * 3 + 1];
    jz[p] += s * curl[p * 3 + 2];
}

__global__ void k_adam(int64_t m, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                       const uint8_t* __restrict__ mask)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m) return;
    if (mask && !mask[t]) return;
    const float gg = g[t];
    const float a = b1 * m1[t] + (1.0f - b1) * gg;
    const float b = b2 * m2[t] + (1.0f - b2) * gg * gg;
    m1[t] = a;
    m2[t] = b;
    w[t] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// master FP32 weights [cout][kpad] -> BF16 forward weights [cout_pad][kpad]
// and BF16 transposed/flipped weights for the input gradient:
// WT[ci][tap'*cout_in + co] = W[co][tap*cin + ci] with tap' the mirrored tap
__global__ void k_prep_weights(const float* __restrict__ w, int cout, int cout_pad, int cin, int taps, int kpad,
                               uint16_t* __restrict__ wf, uint16_t* __restrict__ wt, int cinT, int cinT_pad,
                               int kpadT)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)cout_pad * kpad) {
        const int co = (int)(t / kpad), k = (int)(t % kpad);
        wf[t] = (co < cout && k < taps * cin) ? f2bf(w[(int64_t)co * kpad + k]) : (uint16_t)0;
    }
    // transposed: rows ci (cinT_pad of them), columns tap'*cinT + co  (cinT = cout rounded to 4)
    if (wt && t < (int64_t)cinT_pad * kpadT) {
        const int ci = (int)(t / kpadT), kk = (int)(t % kpadT);
        uint16_t v = 0;
        if (ci < cin && kk < taps * cinT) {
            const int tapm = kk / cinT, co = kk - tapm * cinT;
            const int tap = taps - 1 - tapm;   // mirrored offset
            if (co < cout) v = f2bf(w[(int64_t)co * kpad + tap * cin + ci]);
        }
        wt[t] = v;
    }
}

}  // namespace dn
}  // namespace tmpg

// ====================================================================== host
using namespace tmpg;
using namespace tmpg::dn;

struct DLayer {
    int cin, cout;          // real channel counts (cin includes the input pad to 4)
    int cout_pad;           // multiple of 16 (MFMA N tiles)
    int taps, kpad;         // K = taps*cin rounded to 32
    int cinT, kpadT, cinT_pad;   // transposed conv for the input gradient
    float* w = nullptr;     // FP32 master [cout][kpad]
    float* b = nullptr;     // [cout]
    float *gw = nullptr, *gb = nullptr, *m1w = nullptr, *m2w = nullptr, *m1b = nullptr, *m2b = nullptr;
    uint16_t* wf = nullptr; // BF16 [cout_pad][kpad]
    uint16_t* wt = nullptr; // BF16 [cinT_pad][kpadT]
};

struct tmpg_model {
    int dim, width, nlayers;
    int ncu = 256;          // compute units of the model's device
    std::vector<DLayer> L;
    float scale = 0.0f;
    int base_passes = 0;    // binomial passes applied to J before the correction
    int adam_t = 0;
    hipStream_t stream = nullptr;
    std::string err;
    // work buffers, sized for the largest (pixels) seen
    int64_t cap_pix = 0;
    uint16_t* x0 = nullptr;              // input [P][4]
    std::vector<float*> z;               // pre-activations [P][width]
    std::vector<uint16_t*> h;            // activations [P][width]
    float *A = nullptr, *curl = nullptr, *gcurl = nullptr, *dA = nullptr, *dh = nullptr;
    uint16_t* dz = nullptr;
    float *jlo = nullptr, *jt = nullptr;
    double* dloss = nullptr;
    double* dots = nullptr;              // per image <d,t>, <t,t> (amplitude term)
    float* wpart = nullptr;              // per-workgroup weight-gradient partial sums
    int64_t wpart_cap = 0;               // (floats)
    float lam = 0.0f;                    // weight of the amplitude term
    // ordering between a context's stream and the model's stream in
    // denoise_apply (created once with the model)
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    // device-resident training set
    std::vector<float*> set_lo, set_t;
    int set_nx = 0, set_ny = 0, set_nz = 0;
};

namespace {

int64_t cdiv(int64_t a, int64_t b) { return (a + b - 1) / b; }
constexpr int kMaxBatch = 64;        // images per training batch with the amplitude term

#define DN_CHECK(call)                                                                 \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            m->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return TMPG_EHIP;                                                          \
        }                                                                              \
    } while (0)

template <class T>
int dn_alloc(tmpg_model* m, T** p, int64_t n)
{
    hipError_t e = hipMalloc((void**)p, (size_t)std::max<int64_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) { m->err = std::string("hipMalloc: ") + hipGetErrorString(e); *p = nullptr; return TMPG_ENOMEM; }
    return TMPG_OK;
}

void free_work(tmpg_model* m)
{
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    f(m->x0); f(m->A); f(m->curl); f(m->gcurl); f(m->dA); f(m->dh); f(m->dz); f(m->jlo); f(m->jt);
    for (auto p : m->z) f(p);
    for (auto p : m->h) f(p);
    m->z.clear(); m->h.clear();
    m->x0 = nullptr; m->A = m->curl = m->gcurl = m->dA = m->dh = nullptr; m->dz = nullptr;
    m->jlo = m->jt = nullptr;
    m->cap_pix = 0;
}

int ensure_work(tmpg_model* m, int64_t P)
{
    if (P <= m->cap_pix) return TMPG_OK;
    free_work(m);
    int rc = 0;
    const int w = m->width;
    rc |= dn_alloc(m, &m->x0, P * 4);
    for (int l = 0; l < m->nlayers - 1; ++l) {
        float* zz; uint16_t* hh;
        rc |= dn_alloc(m, &zz, P * w);
        rc |= dn_alloc(m, &hh, P * w);
        m->z.push_back(zz); m->h.push_back(hh);
    }
    rc |= dn_alloc(m, &m->A, P * 3);
    rc |= dn_alloc(m, &m->curl, P * 3);
    rc |= dn_alloc(m, &m->gcurl, P * 3);
    rc |= dn_alloc(m, &m->dA, P * 3);
    rc |= dn_alloc(m, &m->dh, P * w);
    rc |= dn_alloc(m, &m->dz, P * w);
    rc |= dn_alloc(m, &m->jlo, P * 3);
    rc |= dn_alloc(m, &m->jt, P * 3);
    if (rc) return TMPG_ENOMEM;
    m->cap_pix = P;
    return TMPG_OK;
}

int prep_weights(tmpg_model* m)
{
    for (auto& l : m->L) {
        const int64_t n1 = (int64_t)l.cout_pad * l.kpad, n2 = (int64_t)l.cinT_pad * l.kpadT;
        hipLaunchKernelGGL(k_prep_weights, dim3((unsigned)cdiv(std::max(n1, n2), 256)), dim3(256), 0, m->stream,
                           l.w, l.cout, l.cout_pad, l.cin, l.taps, l.kpad, l.wf, l.wt, l.cinT, l.cinT_pad, l.kpadT);
    }
    DN_CHECK(hipGetLastError());
    return TMPG_OK;
}

int launch_conv(tmpg_model* m, const Geo& g, const uint16_t* in, int cin, int taps, const uint16_t* W, int kpad,
                int cout_pad, const float* bias, int cout, float* raw, uint16_t* gout, int gstride)
{
    auto go = [&](auto dim, auto nt, auto ci) {
        constexpr int D = decltype(dim)::value, NT = decltype(nt)::value, CIN = decltype(ci)::value;
        if constexpr (D == 3 && NT == 1 && CIN == 16) {
            // the register-blocked kernel: persistent workgroups, as many as
            // fit on the device (three per CU by LDS)
            const int64_t tiles = (int64_t)cdiv(g.nx, RbTile::TX) * cdiv(g.ny, RbTile::TY) *
                                  cdiv(g.nz, RbTile::TZ) * g.batch;
            constexpr size_t lds = conv_rb_lds_bytes();
            (void)hipFuncSetAttribute((const void*)k_conv3_rb, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
            int per_cu = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv3_rb, 256, lds) != hipSuccess ||
                per_cu < 1)
                per_cu = 1;
            const int64_t grid = std::min<int64_t>(tiles, (int64_t)m->ncu * per_cu);
            hipLaunchKernelGGL(k_conv3_rb, dim3((unsigned)grid), dim3(256), lds, m->stream, g, in, W, bias, cout,
                               raw, gout, gstride);
            return;
        }
        using T = ConvTile<D, NT>;
        const int64_t tiles = (int64_t)cdiv(g.nx, T::TX) * cdiv(g.ny, T::TY) * cdiv(g.nz, T::TZ) * g.batch;
        constexpr size_t lds = conv_lds_bytes<D, NT, CIN>();
        static_assert(lds <= 160 * 1024, "conv LDS");
        (void)hipFuncSetAttribute((const void*)k_conv<NT, D, CIN>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        // persistent workgroups: as many as are resident on the device at once
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_conv<NT, D, CIN>, kConvThreads, lds) !=
                hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const int64_t grid = std::min<int64_t>(tiles, (int64_t)m->ncu * per_cu);
        hipLaunchKernelGGL((k_conv<NT, D, CIN>), dim3((unsigned)grid), dim3(kConvThreads), lds, m->stream, g, in, W,
                           bias, cout, raw, gout, gstride);
    };
    using I1 = std::integral_constant<int, 1>;
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using C4 = std::integral_constant<int, 4>;
    using C16 = std::integral_constant<int, 16>;
    using C32 = std::integral_constant<int, 32>;
    // the layer shapes of the two networks: inputs of 4 channels (J) or the
    // width (2D 32, 3D 16), and 16 or 32 output channels after padding
    auto by_cin = [&](auto dim, auto nt) {
        if (cin == 4) go(dim, nt, C4{});
        else if constexpr (decltype(dim)::value == 3) go(dim, nt, C16{});
        else go(dim, nt, C32{});
    };
    auto by_nt = [&](auto dim) {
        if (cout_pad == 16) by_cin(dim, I1{}); else by_cin(dim, I2{});
    };
    const int width = g.dim == 3 ? 16 : 32;
    const int taps_dim = g.dim == 3 ? 27 : 9;
    if ((cin != 4 && cin != width) || taps != taps_dim || kpad != (taps * cin + 31) / 32 * 32 ||
        (cout_pad != 16 && cout_pad != 32)) {
        // unreachable for the two networks built here
        m->err = "no conv kernel for cin " + std::to_string(cin) + ", cout_pad " + std::to_string(cout_pad);
        return TMPG_EINVAL;
    }
    if (g.dim == 3) by_nt(I3{}); else by_nt(I2{});
    return TMPG_OK;
}

// forward from m->x0: fills h, A and curl, and the pre-activations z that
// the backward pass needs when keep_z
int forward(tmpg_model* m, const Geo& g, bool keep_z = true)
{
    const uint16_t* in = m->x0;
    int cin = 4;
    for (int l = 0; l < m->nlayers; ++l) {
        DLayer& L = m->L[l];
        const bool last = l == m->nlayers - 1;
        const int rc = launch_conv(m, g, in, cin, L.taps, L.wf, L.kpad, L.cout_pad, L.b, L.cout,
                                   last ? m->A : (keep_z ? m->z[l] : nullptr), last ? nullptr : m->h[l], L.cout);
        if (rc) return rc;
        if (!last) { in = m->h[l]; cin = L.cout; }
    }
    const int64_t P = g.npix * g.batch;
    hipLaunchKernelGGL(k_curl, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, m->stream, g, m->A, m->curl);
    return TMPG_OK;
}

// backward from m->gcurl (dL/dcurl): accumulates gw, gb
// the tiled weight gradient k_wgrad for (dim, cin) in {2, 3} x {4, 16, 32}
// with at most 8 * WG_MT output pairs; false if the shape is not one of them
template <int D, int CIN>
int launch_wgrad_t(tmpg_model* m, const Geo& g, int dz_c, int cout, int cout_pad, const uint16_t* in, int taps,
                    int kpad, float* gw, float* gb)
{
    using T = ConvTile<D, 1>;
    const size_t lds = wgrad_lds_bytes<D, CIN>(cout_pad);
    (void)hipFuncSetAttribute((const void*)k_wgrad<D, CIN>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_wgrad<D, CIN>, WG_THREADS, lds) != hipSuccess ||
        per_cu < 1) {
        (void)hipGetLastError();
        per_cu = 1;
    }
    const int64_t tiles = (int64_t)cdiv(g.nx, T::TX) * cdiv(g.ny, T::TY) * cdiv(g.nz, T::TZ) * g.batch;
    const int64_t grid = std::min<int64_t>(tiles, (int64_t)m->ncu * per_cu);
    const int64_t stride = (int64_t)cout_pad * kpad + cout_pad;
    if (grid * stride > m->wpart_cap) {
        if (m->wpart) (void)hipFree(m->wpart);
        m->wpart = nullptr;
        m->wpart_cap = 0;
        if (dn_alloc(m, &m->wpart, grid * stride)) return TMPG_ENOMEM;
        m->wpart_cap = grid * stride;
    }
    hipLaunchKernelGGL((k_wgrad<D, CIN>), dim3((unsigned)grid), dim3(WG_THREADS), lds, m->stream, g, m->dz, dz_c,
                       cout, cout_pad, in, taps, kpad, m->wpart);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3((unsigned)cdiv(stride, 16)), dim3(256), 0, m->stream, m->wpart,
                       (int)grid, cout, cout_pad, kpad, taps * CIN, gw, gb);
    return TMPG_OK;
}

// returns 1 when the shape has no tiled kernel (the caller then runs the
// per-chunk one), else TMPG_OK or an error code
int launch_wgrad(tmpg_model* m, const Geo& g, int dz_c, int cout, int cout_pad, const uint16_t* in, int cin,
                 int taps, int kpad, float* gw, float* gb)
{
    if (cout_pad % 16 || cout_pad > 32 || (cout_pad / 16) * (kpad / 16) > 8 * WG_MT || kpad % 16) return 1;
    if (g.dim == 3 && cin == 4) return launch_wgrad_t<3, 4>(m, g, dz_c, cout, cout_pad, in, taps, kpad, gw, gb);
    if (g.dim == 3 && cin == 16) return launch_wgrad_t<3, 16>(m, g, dz_c, cout, cout_pad, in, taps, kpad, gw, gb);
    if (g.dim == 2 && cin == 4) return launch_wgrad_t<2, 4>(m, g, dz_c, cout, cout_pad, in, taps, kpad, gw, gb);
    if (g.dim == 2 && cin == 32) return launch_wgrad_t<2, 32>(m, g, dz_c, cout, cout_pad, in, taps, kpad, gw, gb);
    return 1;
}

int backward(tmpg_model* m, const Geo& g)
{
    const int64_t P = g.npix * g.batch;
    const int64_t chunk = 1024;
    hipLaunchKernelGGL(k_curl_adj, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, m->stream, g, m->gcurl, m->dA);
    // dZ of the last layer: BF16 [P][4]
    hipLaunchKernelGGL(k_to_bf16, dim3((unsigned)cdiv(P * 4, 256)), dim3(256), 0, m->stream, P, m->dA, 3, m->dz, 4);
    int dz_c = 4;
    for (int l = m->nlayers - 1; l >= 0; --l) {
        DLayer& L = m->L[l];
        const uint16_t* in = (l == 0) ? m->x0 : m->h[l - 1];
        const int cin = (l == 0) ? 4 : m->L[l - 1].cout;
        // weight gradient (dz stored with stride dz_c; cout real channels):
        // the tiled MFMA kernel for the model's layer shapes, the per-chunk
        // one otherwise
        const int wrc = launch_wgrad(m, g, dz_c, L.cout, L.cout_pad, in, cin, L.taps, L.kpad, L.gw, L.gb);
        if (wrc < 0) return wrc;
        if (wrc == 1)
            hipLaunchKernelGGL(k_conv_wgrad, dim3((unsigned)cdiv(P, chunk)), dim3(256), 0, m->stream, g, m->dz,
                               dz_c, L.cout, L.cout_pad, in, cin, L.taps, L.kpad, chunk, L.gw, L.gb);
        if (l == 0) break;
        // input gradient: conv of dz with the transposed, flipped weights
        const int rc = launch_conv(m, g, m->dz, dz_c, L.taps, L.wt, L.kpadT, L.cinT_pad, nullptr, L.cin, m->dh,
                                   nullptr, 0);
        if (rc) return rc;
        const int c = m->L[l - 1].cout;
        hipLaunchKernelGGL(k_gelu_back, dim3((unsigned)cdiv(P * c, 256)), dim3(256), 0, m->stream, P * c, m->dh,
                           m->z[l - 1], m->dz);
        dz_c = c;
    }
    return TMPG_OK;
}

int set_geometry(tmpg_model* m, Geo& g, int nx, int ny, int nz, int batch)
{
    g.dim = m->dim;
    g.nx = nx; g.ny = ny; g.nz = (m->dim == 3) ? nz : 1;
    g.npix = (int64_t)g.nx * g.ny * g.nz;
    g.batch = batch;
    if (nx < 3 || ny < 3 || (m->dim == 3 && nz < 3) || batch < 1) { m->err = "bad denoiser geometry"; return TMPG_EINVAL; }
    return ensure_work(m, g.npix * batch);
}

// forward + loss (+ backward) on device arrays jlo, jt [batch][3][n]
int run_batch(tmpg_model* m, const Geo& g, const float* jlo, const float* jt, bool grad, double* loss)
{
    const int64_t n = g.npix, P = n * g.batch;
    if (m->scale <= 0.0f) { m->err = "denoiser scale not set"; return TMPG_ESTATE; }
    const float inv_s = 1.0f / m->scale;
    hipLaunchKernelGGL(k_pack_input, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, m->stream, jlo, n, g.batch, inv_s,
                       m->x0);
    int rc = forward(m, g);
    if (rc) return rc;
    DN_CHECK(hipMemsetAsync(m->dloss, 0, sizeof(double), m->stream));
    if (m->lam > 0.0f) {
        if (g.batch > kMaxBatch) { m->err = "batch too large for the amplitude term"; return TMPG_EINVAL; }
        DN_CHECK(hipMemsetAsync(m->dots, 0, 2 * (size_t)g.batch * sizeof(double), m->stream));
        hipLaunchKernelGGL(k_dots, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, m->stream, n, g.batch, jlo, jt,
                           inv_s, m->curl, m->dots);
    }
    hipLaunchKernelGGL(k_loss, dim3((unsigned)cdiv(P, 256)), dim3(256), 0, m->stream, n, g.batch, jlo, jt, inv_s,
                       m->curl, grad ? m->gcurl : nullptr, m->dloss, 1.0 / (double)(3 * P), m->lam,
                       (const double*)m->dots);
    if (grad) {
        for (auto& L : m->L) {
            DN_CHECK(hipMemsetAsync(L.gw, 0, (size_t)L.cout * L.kpad * sizeof(float), m->stream));
            DN_CHECK(hipMemsetAsync(L.gb, 0, (size_t)L.cout * sizeof(float), m->stream));
        }
        rc = backward(m, g);
        if (rc) return rc;
    }
    DN_CHECK(hipGetLastError());
    if (loss) {
        DN_CHECK(hipMemcpyAsync(loss, m->dloss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
        DN_CHECK(hipStreamSynchronize(m->stream));
    }
    return TMPG_OK;
}

int adam_step(tmpg_model* m, float lr)
{
    m->adam_t++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
    for (auto& L : m->L) {
        const int64_t nw = (int64_t)L.cout * L.kpad;
        hipLaunchKernelGGL(k_adam, dim3((unsigned)cdiv(nw, 256)), dim3(256), 0, m->stream, nw, L.w, L.gw, L.m1w, L.m2w,
                           lr, b1, b2, eps, c1, c2, (const uint8_t*)nullptr);
        hipLaunchKernelGGL(k_adam, dim3((unsigned)cdiv(L.cout, 256)), dim3(256), 0, m->stream, (int64_t)L.cout, L.b,
                           L.gb, L.m1b, L.m2b, lr, b1, b2, eps, c1, c2, (const uint8_t*)nullptr);
    }
    DN_CHECK(hipGetLastError());
    return prep_weights(m);
}

// flat parameter order: per layer W[cout][taps*cin_real] (cin_real: 3 for
// the first layer, else the previous width) then b[cout]
int64_t nparams(const tmpg_model* m)
{
    int64_t n = 0;
    for (size_t l = 0; l < m->L.size(); ++l) {
        const DLayer& L = m->L[l];
        const int cr = (l == 0) ? 3 : L.cin;
        n += (int64_t)L.cout * L.taps * cr + L.cout;
    }
    return n;
}

}  // namespace

extern "C" {

int tmpg_model_create(int dim, uint64_t seed, tmpg_model** out)
{
    if (!out || (dim != 2 && dim != 3)) return TMPG_EINVAL;
    tmpg_model* m = new (std::nothrow) tmpg_model();
    if (!m) return TMPG_ENOMEM;
    m->dim = dim;
    m->width = (dim == 2) ? 32 : 16;
    m->nlayers = (dim == 2) ? 4 : 3;
    const int taps = (dim == 2) ? 9 : 27;
    if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess) { delete m; return TMPG_EHIP; }
    {
        int dev = 0, ncu = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && ncu > 0)
            m->ncu = ncu;
    }
    uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
    auto rnd = [&]() {  // uniform (-1,1), LCG + xorshift
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        uint64_t x = st ^ (st >> 33);
        return (double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0;
    };
    int rc = 0;
    for (int l = 0; l < m->nlayers; ++l) {
        DLayer L;
        L.cin = (l == 0) ? 4 : m->width;
        L.cout = (l == m->nlayers - 1) ? 3 : m->width;
        L.cout_pad = (L.cout + 15) / 16 * 16;
        L.taps = taps;
        L.kpad = (taps * L.cin + 31) / 32 * 32;
        L.cinT = (L.cout + 3) / 4 * 4;
        L.cinT_pad = (L.cin + 15) / 16 * 16;
        L.kpadT = (taps * L.cinT + 31) / 32 * 32;
        const int64_t nw = (int64_t)L.cout * L.kpad;
        rc |= dn_alloc(m, &L.w, nw); rc |= dn_alloc(m, &L.b, L.cout);
        rc |= dn_alloc(m, &L.gw, nw); rc |= dn_alloc(m, &L.gb, L.cout);
        rc |= dn_alloc(m, &L.m1w, nw); rc |= dn_alloc(m, &L.m2w, nw);
        rc |= dn_alloc(m, &L.m1b, L.cout); rc |= dn_alloc(m, &L.m2b, L.cout);
        rc |= dn_alloc(m, &L.wf, (int64_t)L.cout_pad * L.kpad);
        rc |= dn_alloc(m, &L.wt, (int64_t)L.cinT_pad * L.kpadT);
        if (rc) break;
        // He-uniform init over the real fan-in; the last layer starts small
        // so the untrained correction is close to zero
        const int cr = (l == 0) ? 3 : L.cin;
        const double a = std::sqrt(6.0 / (taps * cr)) * ((l == m->nlayers - 1) ? 0.01 : 1.0);
        std::vector<float> hw((size_t)nw, 0.0f);
        for (int co = 0; co < L.cout; ++co)
            for (int t = 0; t < taps; ++t)
                for (int ci = 0; ci < cr; ++ci) hw[(size_t)co * L.kpad + t * L.cin + ci] = (float)(a * rnd());
        m->L.push_back(L);
        // weights, biases and Adam moments: a failed copy is an error, not
        // a silently uninitialised model
        hipError_t e = hipMemcpy(L.w, hw.data(), nw * sizeof(float), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemset(L.b, 0, L.cout * sizeof(float));
        if (e == hipSuccess) e = hipMemset(L.m1w, 0, nw * sizeof(float));
        if (e == hipSuccess) e = hipMemset(L.m2w, 0, nw * sizeof(float));
        if (e == hipSuccess) e = hipMemset(L.m1b, 0, L.cout * sizeof(float));
        if (e == hipSuccess) e = hipMemset(L.m2b, 0, L.cout * sizeof(float));
        if (e != hipSuccess) {
            tmpg_model_destroy(m);
            return TMPG_EHIP;
        }
    }
    rc |= dn_alloc(m, &m->dloss, 1);
    rc |= dn_alloc(m, &m->dots, 2 * kMaxBatch);
    if (rc || prep_weights(m)) { tmpg_model_destroy(m); return TMPG_ENOMEM; }
    if (hipEventCreateWithFlags(&m->ev_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ev_out, hipEventDisableTiming) != hipSuccess ||
        hipStreamSynchronize(m->stream) != hipSuccess) {
        tmpg_model_destroy(m);
        return TMPG_EHIP;
    }
    *out = m;
    return TMPG_OK;
}

int tmpg_model_destroy(tmpg_model* m)
{
    if (!m) return TMPG_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    free_work(m);
    auto f = [](void* p) { if (p) (void)hipFree(p); };
    for (auto& L : m->L) { f(L.w); f(L.b); f(L.gw); f(L.gb); f(L.m1w); f(L.m2w); f(L.m1b); f(L.m2b); f(L.wf); f(L.wt); }
    for (auto p : m->set_lo) f(p);
    for (auto p : m->set_t) f(p);
    f(m->dloss);
    f(m->wpart);
    f(m->dots);
    if (m->ev_in) (void)hipEventDestroy(m->ev_in);
    if (m->ev_out) (void)hipEventDestroy(m->ev_out);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return TMPG_OK;
}

const char* tmpg_model_last_error(const tmpg_model* m) { return m ? m->err.c_str() : ""; }

int tmpg_model_num_params(const tmpg_model* m, int64_t* n)
{
    if (!m || !n) return TMPG_EINVAL;
    *n = nparams(m);
    return TMPG_OK;
}

static int params_io(tmpg_model* m, float* host, bool get, bool grads)
{
    (void)hipStreamSynchronize(m->stream);
    int64_t off = 0;
    for (size_t l = 0; l < m->L.size(); ++l) {
        DLayer& L = m->L[l];
        const int cr = (l == 0) ? 3 : L.cin;
        const int64_t nw = (int64_t)L.cout * L.kpad;
        std::vector<float> w((size_t)nw), b((size_t)L.cout);
        float* dw = grads ? L.gw : L.w;
        float* dbv = grads ? L.gb : L.b;
        if (get || true) {
            DN_CHECK(hipMemcpy(w.data(), dw, nw * sizeof(float), hipMemcpyDeviceToHost));
            DN_CHECK(hipMemcpy(b.data(), dbv, L.cout * sizeof(float), hipMemcpyDeviceToHost));
        }
        for (int co = 0; co < L.cout; ++co)
            for (int t = 0; t < L.taps; ++t)
                for (int ci = 0; ci < cr; ++ci) {
                    float& v = w[(size_t)co * L.kpad + t * L.cin + ci];
                    if (get) host[off] = v; else v = host[off];
                    ++off;
                }
        for (int co = 0; co < L.cout; ++co) {
            if (get) host[off] = b[co]; else b[co] = host[off];
            ++off;
        }
        if (!get) {
            DN_CHECK(hipMemcpy(dw, w.data(), nw * sizeof(float), hipMemcpyHostToDevice));
            DN_CHECK(hipMemcpy(dbv, b.data(), L.cout * sizeof(float), hipMemcpyHostToDevice));
        }
    }
    if (!get && !grads) return prep_weights(m);
    return TMPG_OK;
}

int tmpg_model_get_params(tmpg_model* m, float* out) { return (m && out) ? params_io(m, out, true, false) : TMPG_EINVAL; }
int tmpg_model_set_params(tmpg_model* m, const float* in)
{
    return (m && in) ? params_io(m, const_cast<float*>(in), false, false) : TMPG_EINVAL;
}

int tmpg_model_set_scale(tmpg_model* m, float s)
{
    if (!m || !(s > 0.0f)) return TMPG_EINVAL;
    m->scale = s;
    return TMPG_OK;
}

int tmpg_model_set_amplitude_weight(tmpg_model* m, float lam)
{
    if (!m || !(lam >= 0.0f)) return TMPG_EINVAL;
    m->lam = lam;
    return TMPG_OK;
}

int tmpg_model_set_base_passes(tmpg_model* m, int npasses)
{
    if (!m || npasses < 0) return TMPG_EINVAL;
    m->base_passes = npasses;
    return TMPG_OK;
}

int tmpg_model_get_base_passes(const tmpg_model* m, int* npasses)
{
    if (!m || !npasses) return TMPG_EINVAL;
    *npasses = m->base_passes;
    return TMPG_OK;
}

int tmpg_model_get_scale(const tmpg_model* m, float* s)
{
    if (!m || !s) return TMPG_EINVAL;
    *s = m->scale;
    return TMPG_OK;
}

// one batch from host arrays [batch][3][n]: loss, optional flat gradient,
// optional Adam update with learning rate lr (update != 0)
int tmpg_model_train_batch(tmpg_model* m, int nx, int ny, int nz, int batch, const float* jlo, const float* jt,
                           float lr, int update, double* loss, float* grad_out)
{
    if (!m || !jlo || !jt) return TMPG_EINVAL;
    Geo g;
    int rc = set_geometry(m, g, nx, ny, nz, batch);
    if (rc) return rc;
    const size_t bytes = (size_t)g.npix * 3 * batch * sizeof(float);
    DN_CHECK(hipMemcpyAsync(m->jlo, jlo, bytes, hipMemcpyHostToDevice, m->stream));
    DN_CHECK(hipMemcpyAsync(m->jt, jt, bytes, hipMemcpyHostToDevice, m->stream));
    rc = run_batch(m, g, m->jlo, m->jt, true, loss);
    if (rc) return rc;
    if (grad_out) {
        rc = params_io(m, grad_out, true, true);
        if (rc) return rc;
    }
    if (update) rc = adam_step(m, lr);
    return rc;
}

// loss only, and optionally the corrected current J + s*curl(A) [batch][3][n]
int tmpg_model_eval(tmpg_model* m, int nx, int ny, int nz, int batch, const float* jlo, const float* jt, double* loss,
                    float* jout)
{
    if (!m || !jlo || !jt) return TMPG_EINVAL;
    Geo g;
    int rc = set_geometry(m, g, nx, ny, nz, batch);
    if (rc) return rc;
    const size_t bytes = (size_t)g.npix * 3 * batch * sizeof(float);
    DN_CHECK(hipMemcpyAsync(m->jlo, jlo, bytes, hipMemcpyHostToDevice, m->stream));
    DN_CHECK(hipMemcpyAsync(m->jt, jt, bytes, hipMemcpyHostToDevice, m->stream));
    rc = run_batch(m, g, m->jlo, m->jt, false, loss);
    if (rc) return rc;
    if (jout) {
        std::vector<float> c((size_t)g.npix * 3 * batch);
        DN_CHECK(hipMemcpy(c.data(), m->curl, c.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int b = 0; b < batch; ++b)
            for (int q = 0; q < 3; ++q)
                for (int64_t p = 0; p < g.npix; ++p) {
                    const size_t o = ((size_t)b * 3 + q) * g.npix + p;
                    jout[o] = jlo[o] + m->scale * c[((size_t)b * g.npix + p) * 3 + q];
                }
    }
    return TMPG_OK;
}

// add one training pair (host arrays [3][n]) to the device-resident set; the
// first pair sets the input scale to the rms of its low-ppc current
int tmpg_model_add_sample(tmpg_model* m, int nx, int ny, int nz, const float* jlo, const float* jt)
{
    if (!m || !jlo || !jt) return TMPG_EINVAL;
    const int nzz = (m->dim == 3) ? nz : 1;
    if (!m->set_lo.empty() && (nx != m->set_nx || ny != m->set_ny || nzz != m->set_nz)) {
        m->err = "training samples must share one grid";
        return TMPG_EINVAL;
    }
    m->set_nx = nx; m->set_ny = ny; m->set_nz = nzz;
    const int64_t n = (int64_t)nx * ny * nzz * 3;
    if (m->scale <= 0.0f) {
        double s2 = 0.0;
        for (int64_t i = 0; i < n; ++i) s2 += (double)jlo[i] * jlo[i];
        m->scale = (float)std::sqrt(s2 / (double)n);
        if (!(m->scale > 0.0f)) m->scale = 1.0f;
    }
    float *a, *b;
    if (dn_alloc(m, &a, n) || dn_alloc(m, &b, n)) return TMPG_ENOMEM;
    DN_CHECK(hipMemcpy(a, jlo, n * sizeof(float), hipMemcpyHostToDevice));
    DN_CHECK(hipMemcpy(b, jt, n * sizeof(float), hipMemcpyHostToDevice));
    m->set_lo.push_back(a);
    m->set_t.push_back(b);
    return TMPG_OK;
}

// iters Adam steps over the device-resident set, one sample per step in a
// fixed shuffled order; losses[i] (optional) = loss of step i
int tmpg_model_train(tmpg_model* m, int iters, float lr, uint64_t seed, double* losses)
{
    if (!m || iters < 0) return TMPG_EINVAL;
    if (m->set_lo.empty()) { m->err = "no training samples"; return TMPG_ESTATE; }
    Geo g;
    int rc = set_geometry(m, g, m->set_nx, m->set_ny, m->set_nz, 1);
    if (rc) return rc;
    uint64_t st = seed * 0x9E3779B97F4A7C15ull + 1;
    const size_t ns = m->set_lo.size();
    for (int it = 0; it < iters; ++it) {
        st ^= st << 13; st ^= st >> 7; st ^= st << 17;
        const size_t k = (size_t)(st % ns);
        double l = 0.0;
        rc = run_batch(m, g, m->set_lo[k], m->set_t[k], true, losses ? &l : nullptr);
        if (rc) return rc;
        if (losses) losses[it] = l;
        // cosine decay of the learning rate
        const float f = 0.5f * (1.0f + std::cos(3.14159265f * (float)it / (float)std::max(iters, 1)));
        rc = adam_step(m, lr * (0.05f + 0.95f * f));
        if (rc) return rc;
    }
    DN_CHECK(hipStreamSynchronize(m->stream));
    return TMPG_OK;
}

}  // extern "C"

// ---------------------------------------------- use inside the step loop
namespace tmpg {
int denoise_apply(tmpg_ctx* ctx, tmpg_model* m)
{
    if (m->base_passes > 0) {
        const int rc = fields_filter(ctx, m->base_passes);
        if (rc) return rc;
    }
    Geo g;
    int rc = set_geometry(m, g, ctx->g.nx, ctx->g.ny, ctx->g.nz, 1);
    if (rc) { ctx->err = m->err; return rc; }
    if (m->scale <= 0.0f) { ctx->err = "denoiser scale not set"; return TMPG_ESTATE; }
    // the model's stream waits for the context's stream and vice versa (the
    // model's two events; a failure still joins the streams, so no work of
    // this call is left running behind the caller)
    TMPG_HIP_CHECK(ctx, hipEventRecord(m->ev_in, ctx->stream));
    TMPG_HIP_CHECK(ctx, hipStreamWaitEvent(m->stream, m->ev_in, 0));
    const int64_t n = g.npix;
    hipLaunchKernelGGL(k_pack_input3, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, m->stream, ctx->f[JX],
                       ctx->f[JY], ctx->f[JZ], n, 1.0f / m->scale, m->x0);
    rc = forward(m, g, false);
    if (rc) {
        ctx->err = m->err;
    } else {
        hipLaunchKernelGGL(k_apply, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, m->stream, n, m->curl, m->scale,
                           ctx->f[JX], ctx->f[JY], ctx->f[JZ]);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            ctx->err = std::string("denoiser launch: ") + hipGetErrorString(e);
            rc = TMPG_EHIP;
        }
    }
    const hipError_t e1 = hipEventRecord(m->ev_out, m->stream);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamWaitEvent(ctx->stream, m->ev_out, 0) : e1;
    if (e2 != hipSuccess) {
        (void)hipStreamSynchronize(m->stream);
        if (!rc) {
            ctx->err = std::string("denoiser stream join: ") + hipGetErrorString(e2);
            rc = TMPG_EHIP;
        }
    }
    return rc;
}
}  // namespace tmpg

extern "C" {
int tmpg_denoise(tmpg_ctx* ctx, tmpg_model* m)
{
    if (!ctx || !m) return TMPG_EINVAL;
    if (m->dim != ctx->g.dim) { ctx->err = "denoiser dimensionality differs from the grid's"; return TMPG_EINVAL; }
    return tmpg::denoise_apply(ctx, m);
}

int tmpg_set_denoiser(tmpg_ctx* ctx, tmpg_model* m)
{
    if (!ctx) return TMPG_EINVAL;
    if (m && m->dim != ctx->g.dim) { ctx->err = "denoiser dimensionality differs from the grid's"; return TMPG_EINVAL; }
    ctx->denoiser = m;
    return TMPG_OK;
}
}  // extern "C"
