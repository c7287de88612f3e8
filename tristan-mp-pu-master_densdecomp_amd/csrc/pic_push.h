// pic_push.h -- gather and Esirkepov deposit of one particle, written against
// a field source and a current sink so the same arithmetic serves the tiled
// kernel (LDS-staged fields, LDS currents) and the flat kernel (HBM fields,
// global atomics).
//
// Reference: mover_{1,2,3}ord gathers (code/particles_movedeposit.F90:
// 1137-1170 and the order-1/2 equivalents) and densdecomp_{1,2,3}ord
// (code/particles.F90:678-1358).
#pragma once
#include "pic_math.h"

namespace tmpg {

__device__ __forceinline__ int wrap1(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// gather nodes with nonzero weights: order 1 uses nodes 0..1, others -1..2
template <int ORD> struct GatherRange { static constexpr int N0 = (ORD == 1) ? 1 : 0, N1 = (ORD == 1) ? 3 : 4; };

// Field sources.  (i,j,k) are the particle's cell in the source's own
// coordinates; node offsets are -2..+3.
struct GlobalSrc {
    const float* f[6];
    int nx, ny, nz;
    __device__ __forceinline__ int64_t row(int j, int k) const
    {
        return (int64_t)nx * (wrap1(j, ny) + (int64_t)ny * wrap1(k, nz));
    }
    __device__ __forceinline__ int col(int i) const { return wrap1(i, nx); }
    __device__ __forceinline__ float get(int q, int64_t r, int c) const { return f[q][r + c]; }
    __device__ __forceinline__ void get6(int64_t r, int c, float* v) const
    {
#pragma unroll
        for (int q = 0; q < 6; ++q) v[q] = f[q][r + c];
    }
};

// 3D primal fields staged in LDS node-major: the six components of a node are
// adjacent (24 bytes), so one node is three 8-byte LDS reads instead of six
// 4-byte reads from six planes.  All six use the same primal weights.
struct LdsSrc6 {
    const float* s;     // [node][6], node = i + fx*(j + fy*k)
    int fx, fy;
    __device__ __forceinline__ int row(int j, int k) const { return (j + fy * k) * fx; }
    __device__ __forceinline__ int col(int i) const { return i; }
    __device__ __forceinline__ void get6(int r, int c, float* v) const
    {
        const float2* p = reinterpret_cast<const float2*>(s) + 3 * (r + c);
        const float2 a = p[0], b = p[1], d = p[2];
        v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y; v[4] = d.x; v[5] = d.y;
    }
};

struct LdsSrc {
    const float* s;     // 6 component planes of fx*fy*fz floats
    int fx, fy, plane;
    __device__ __forceinline__ int row(int j, int k) const { return (j + fy * k) * fx; }
    __device__ __forceinline__ int col(int i) const { return i; }
    __device__ __forceinline__ float get(int q, int r, int c) const { return s[q * plane + r + c]; }
    __device__ __forceinline__ void get6(int r, int c, float* v) const
    {
#pragma unroll
        for (int q = 0; q < 6; ++q) v[q] = s[q * plane + r + c];
    }
};

struct GlobalSink {
    float* j[3];
    int nx, ny, nz;
    __device__ __forceinline__ int64_t row(int jj, int kk) const
    {
        return (int64_t)nx * (wrap1(jj, ny) + (int64_t)ny * wrap1(kk, nz));
    }
    __device__ __forceinline__ int col(int i) const { return wrap1(i, nx); }
    __device__ __forceinline__ void add(int q, int64_t r, int c, float v) const { atomicAdd(j[q] + r + c, v); }
};

// Tile currents in LDS are accumulated in FP64: on gfx950 ds_add_f64 runs
// about 3x faster than ds_add_f32 (measured on the 3D order-3 push), and the
// tile sum is rounded to FP32 once, at the flush.
struct LdsSink {
    double* s;          // 3 component planes of jx*jy*jz doubles
    int jx, jy, plane;
    __device__ __forceinline__ int row(int jj, int kk) const { return (jj + jy * kk) * jx; }
    __device__ __forceinline__ int col(int i) const { return i; }
    __device__ __forceinline__ void add(int q, int r, int c, float v) const
    {
        atomicAdd(s + q * plane + r + c, (double)v);
    }
};

// E, B at the particle; 3D from primal-averaged fields with primal weights
// (code/particles_movedeposit.F90:1137-1153), 2D from the staggered fields
// with matching primal/dual weights (:1155-1170).
template <int DIM, int ORD, class Src>
__device__ __forceinline__ void gather(const Src& src, bool quirk2, int i, int j, int k, float ox, float oy,
                                       float oz, float* E)
{
    float Spx[4], Sdx[4], Spy[4], Sdy[4];
    int dix, diy;
    gather_weights<ORD>(ox, quirk2, Spx, Sdx, dix);
    gather_weights<ORD>(oy, quirk2, Spy, Sdy, diy);
    constexpr int N0 = GatherRange<ORD>::N0, N1 = GatherRange<ORD>::N1;
#pragma unroll
    for (int q = 0; q < 6; ++q) E[q] = 0.0f;
    if constexpr (DIM == 3) {
        float Spz[4], Sdz[4];
        int diz;
        gather_weights<ORD>(oz, quirk2, Spz, Sdz, diz);
        int cx[4];
#pragma unroll
        for (int t = N0; t < N1; ++t) cx[t] = src.col(i - 1 + t);
        // each z plane summed with the x-y weights, then scaled by its z
        // weight (the order of every sum is the one of gather3_lds)
        float wxy[4][4];
#pragma unroll
        for (int b = N0; b < N1; ++b)
#pragma unroll
            for (int t = N0; t < N1; ++t) wxy[b][t] = Spx[t] * Spy[b];
#pragma unroll
        for (int c = N0; c < N1; ++c) {
            float acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
            for (int b = N0; b < N1; ++b) {
                const auto r = src.row(j - 1 + b, k - 1 + c);
#pragma unroll
                for (int t = N0; t < N1; ++t) {
                    float v[6];
                    src.get6(r, cx[t], v);
#pragma unroll
                    for (int q = 0; q < 6; ++q) acc[q] = fmaf(v[q], wxy[b][t], acc[q]);
                }
            }
#pragma unroll
            for (int q = 0; q < 6; ++q) E[q] = fmaf(acc[q], Spz[c], E[q]);
        }
    } else {
        int cp[4], cd[4];
#pragma unroll
        for (int t = N0; t < N1; ++t) {
            cp[t] = src.col(i - 1 + t);
            cd[t] = src.col(i + dix - 1 + t);
        }
#pragma unroll
        for (int b = N0; b < N1; ++b) {
            const auto rp = src.row(j - 1 + b, 0), rd = src.row(j + diy - 1 + b, 0);
#pragma unroll
            for (int t = N0; t < N1; ++t) {
                E[0] += src.get(0, rp, cd[t]) * Sdx[t] * Spy[b];
                E[1] += src.get(1, rd, cp[t]) * Spx[t] * Sdy[b];
                E[2] += src.get(2, rp, cp[t]) * Spx[t] * Spy[b];
                E[3] += src.get(3, rd, cp[t]) * Spx[t] * Sdy[b];
                E[4] += src.get(4, rp, cd[t]) * Sdx[t] * Spy[b];
                E[5] += src.get(5, rd, cd[t]) * Sdx[t] * Sdy[b];
            }
        }
    }
}

// The 3D gather from node-major primal fields in LDS whose strides FX, FY
// (nodes per row, rows per plane) are compile-time constants.  A node's six
// values sit in two arrays of 16-byte entries: (Ex, Ey, Ez, Bx) in s4, and
// in sp (By, Bz) of the node followed by (By, Bz) of its +x neighbour.  A row
// of the block is NB reads of s4 and NB/2 of sp, all 16-byte reads with a
// 16-bit immediate offset, which covers the particle's whole 4x4x4 node
// block (at most 6.4 KB from its first node for the 8x8x4 tile): one base
// address per array and no per-node address arithmetic.
// (i,j,k) is the particle's cell in staged-node coordinates, at least 1 on
// every axis.  Same arithmetic as gather<3, ORD>.  RB node rows go into each
// batch of reads (two batches in flight): 2 hides LDS latency within the
// wave, 1 keeps the registers for more waves.
template <int ORD, int FX, int FY, int RB = 2>
__device__ __forceinline__ void gather3_lds(const float4* __restrict__ s4, const float4* __restrict__ sp, bool quirk2,
                                            int i, int j, int k, float ox, float oy, float oz, float* E)
{
    float Spx[4], Sdx[4], Spy[4], Sdy[4], Spz[4], Sdz[4];
    int dix, diy, diz;
    gather_weights<ORD>(ox, quirk2, Spx, Sdx, dix);
    gather_weights<ORD>(oy, quirk2, Spy, Sdy, diy);
    gather_weights<ORD>(oz, quirk2, Spz, Sdz, diz);
    constexpr int N0 = GatherRange<ORD>::N0, N1 = GatherRange<ORD>::N1, NB = N1 - N0;
    constexpr int NR = NB * NB;          // node rows (y, z) of the block
    constexpr int NBATCH = NR / RB;
    constexpr int BW = RB * NB;          // nodes per batch
#pragma unroll
    for (int q = 0; q < 6; ++q) E[q] = 0.0f;
    const int n0 = (i - 1) + FX * ((j - 1) + FY * (k - 1));
    static_assert(NB % 2 == 0, "node pairs");
    const float4* p4 = s4 + n0;
    const float4* pp = sp + n0;
    // the reads of batch n + 1 are issued before batch n is summed, so a
    // wave waits for LDS once per batch instead of once per node
    float4 b4[2][BW];
    float4 b2[2][BW / 2];
    auto load = [&](int n, float4* d4, float4* d2) {
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) {
            const int r = n * RB + rr;
            const int b = N0 + r % NB, c = N0 + r / NB;
            const int row = FX * (b + FY * c) + N0;
#pragma unroll
            for (int t = 0; t < NB; ++t) d4[rr * NB + t] = p4[row + t];
#pragma unroll
            for (int t = 0; t < NB; t += 2) d2[(rr * NB + t) / 2] = pp[row + t];
        }
    };
    load(0, b4[0], b2[0]);
    // a z plane of the block is summed with the x-y weights, then scaled by
    // its z weight: NB*NB products plus 6 (NB^3 + NB) multiply-adds
    float wxy[NB][NB];
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int t = 0; t < NB; ++t) wxy[b][t] = Spx[N0 + t] * Spy[N0 + b];
    float acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int n = 0; n < NBATCH; ++n) {
        if (n + 1 < NBATCH) load(n + 1, b4[(n + 1) & 1], b2[(n + 1) & 1]);
        __builtin_amdgcn_sched_barrier(0);
        const float4* d4 = b4[n & 1];
        const float4* d2 = b2[n & 1];
#pragma unroll
        for (int rr = 0; rr < RB; ++rr) {
            const int r = n * RB + rr;
            const int b = r % NB, c = r / NB;
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const float w = wxy[b][t];
                const float4 a = d4[rr * NB + t];
                const float4 pr = d2[(rr * NB + t) / 2];
                const float2 m = (t & 1) ? make_float2(pr.z, pr.w) : make_float2(pr.x, pr.y);
                acc[0] = fmaf(a.x, w, acc[0]);
                acc[1] = fmaf(a.y, w, acc[1]);
                acc[2] = fmaf(a.z, w, acc[2]);
                acc[3] = fmaf(a.w, w, acc[3]);
                acc[4] = fmaf(m.x, w, acc[4]);
                acc[5] = fmaf(m.y, w, acc[5]);
            }
            if (b == NB - 1) {
                const float wz = Spz[N0 + c];
#pragma unroll
                for (int q = 0; q < 6; ++q) {
                    E[q] = fmaf(acc[q], wz, E[q]);
                    acc[q] = 0.0f;
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Esirkepov deposit of one particle from old cell (i,j,k) / offsets o*, with
// displacement d*, charge q.  The stored current follows the reference's
// sign (E += cur): x current at node a = q * sum_{a'<=a} dSx(a') *
// (Ay Az + dSy dSz/12), A = S1 + dS/2, which is the reference's
// S1 S1 + dS S1/2 + S1 dS/2 + dS dS/3.  The last node of each prefix sum is
// the telescoped zero and is not written.
template <int DIM, int ORD, class Sink, bool SKIP_ZERO = true>
__device__ __forceinline__ void deposit(const Sink& snk, int i, int j, int k, float ox, float oy, float oz,
                                        float dx, float dy, float dz, float q)
{
    constexpr int W = Win<ORD>::W;
    float S1x[W], dSx[W], S1y[W], dSy[W];
    const int bx = axis_deposit<ORD>(ox, dx, S1x, dSx);
    const int by = axis_deposit<ORD>(oy, dy, S1y, dSy);
    float Ax[W], Ay[W], Cx[W], Cy[W];
    float run_x = 0.0f, run_y = 0.0f;
#pragma unroll
    for (int t = 0; t < W; ++t) {
        Ax[t] = S1x[t] + 0.5f * dSx[t];
        Ay[t] = S1y[t] + 0.5f * dSy[t];
        run_x += dSx[t];
        run_y += dSy[t];
        Cx[t] = q * run_x;
        Cy[t] = q * run_y;
    }
    const float twelfth = 1.0f / 12.0f;
    int cc[W];
#pragma unroll
    for (int t = 0; t < W; ++t) cc[t] = snk.col(i + bx + t);
    if (DIM == 3) {
        float S1z[W], dSz[W], Az[W], Cz[W];
        const int bz = axis_deposit<ORD>(oz, dz, S1z, dSz);
        float run_z = 0.0f;
#pragma unroll
        for (int t = 0; t < W; ++t) {
            Az[t] = S1z[t] + 0.5f * dSz[t];
            run_z += dSz[t];
            Cz[t] = q * run_z;
        }
#pragma unroll
        for (int c = 0; c < W; ++c) {
#pragma unroll
            for (int b = 0; b < W; ++b) {
                const auto r = snk.row(j + by + b, k + bz + c);
                const float wyz = Ay[b] * Az[c] + twelfth * dSy[b] * dSz[c];
#pragma unroll
                for (int t = 0; t < W; ++t) {
                    if (t < W - 1) {
                        const float v = Cx[t] * wyz;
                        if (!SKIP_ZERO || v != 0.0f) snk.add(0, r, cc[t], v);
                    }
                    if (b < W - 1) {
                        const float v = Cy[b] * (Ax[t] * Az[c] + twelfth * dSx[t] * dSz[c]);
                        if (!SKIP_ZERO || v != 0.0f) snk.add(1, r, cc[t], v);
                    }
                    if (c < W - 1) {
                        const float v = Cz[c] * (Ax[t] * Ay[b] + twelfth * dSx[t] * dSy[b]);
                        if (!SKIP_ZERO || v != 0.0f) snk.add(2, r, cc[t], v);
                    }
                }
            }
        }
    } else {
        // 2D: Jz directly from the z displacement (code/particles.F90:822-848)
        const float qdz = -q * dz;
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const auto r = snk.row(j + by + b, 0);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                if (t < W - 1) {
                    const float v = Cx[t] * Ay[b];
                    if (!SKIP_ZERO || v != 0.0f) snk.add(0, r, cc[t], v);
                }
                if (b < W - 1) {
                    const float v = Cy[b] * Ax[t];
                    if (!SKIP_ZERO || v != 0.0f) snk.add(1, r, cc[t], v);
                }
                const float v = qdz * (Ax[t] * Ay[b] + twelfth * dSx[t] * dSy[b]);
                if (!SKIP_ZERO || v != 0.0f) snk.add(2, r, cc[t], v);
            }
        }
    }
}

// ------------------------------------------------- the original scheme
// The -Dzzag build of the reference: its mover interpolates the staggered
// fields trilinearly with the half-cell staggering folded in
// (code/particles_movedeposit.F90:98-347) and its deposit is the zigzag
// scheme of Umeda et al. 2003 (code/particles.F90:550-669).

// E, B at the particle for the zigzag mover: the staggered component along an
// axis is the two half-cell neighbours' sum plus the offset times their
// difference across two cells, then bilinear in the other axes (E) or
// summed once more over the second staggered axis (B).  In 2D every z offset
// is the particle's own plane.  The sums carry a factor 2 (E) and 4 (B),
// removed here, so E[] is on the scale of gather()'s output.
template <int DIM, class Src>
__device__ __forceinline__ void gather_zz(const Src& src, int i, int j, int k, float dx, float dy, float dz,
                                          float* E)
{
    auto at = [&](int q, int a, int b, int c) {
        return src.get(q, src.row(j + b, DIM == 3 ? k + c : 0), src.col(i + a));
    };
    auto sx = [&](int q, int b, int c) {
        return at(q, 0, b, c) + at(q, -1, b, c) + dx * (at(q, 1, b, c) - at(q, -1, b, c));
    };
    auto sy = [&](int q, int a, int c) {
        return at(q, a, 0, c) + at(q, a, -1, c) + dy * (at(q, a, 1, c) - at(q, a, -1, c));
    };
    auto sz = [&](int q, int a, int b) {
        return at(q, a, b, 0) + at(q, a, b, -1) + dz * (at(q, a, b, 1) - at(q, a, b, -1));
    };
    auto lerp = [](float p, float r, float t) { return p + t * (r - p); };
    const float ex = lerp(lerp(sx(0, 0, 0), sx(0, 1, 0), dy), lerp(sx(0, 0, 1), sx(0, 1, 1), dy), dz);
    const float ey = lerp(lerp(sy(1, 0, 0), sy(1, 0, 1), dz), lerp(sy(1, 1, 0), sy(1, 1, 1), dz), dx);
    const float ez = lerp(lerp(sz(2, 0, 0), sz(2, 1, 0), dx), lerp(sz(2, 0, 1), sz(2, 1, 1), dx), dy);
    // Bx: staggered in y and z; By: in z and x; Bz: in x and y
    float f = sz(3, 0, 0) + sz(3, 0, -1) + dy * (sz(3, 0, 1) - sz(3, 0, -1));
    float g = sz(3, 1, 0) + sz(3, 1, -1) + dy * (sz(3, 1, 1) - sz(3, 1, -1));
    const float bx = lerp(f, g, dx);
    f = sx(4, 0, 0) + sx(4, 0, -1) + dz * (sx(4, 0, 1) - sx(4, 0, -1));
    g = sx(4, 1, 0) + sx(4, 1, -1) + dz * (sx(4, 1, 1) - sx(4, 1, -1));
    const float by = lerp(f, g, dy);
    f = sy(5, 0, 0) + sy(5, -1, 0) + dx * (sy(5, 1, 0) - sy(5, -1, 0));
    g = sy(5, 0, 1) + sy(5, -1, 1) + dx * (sy(5, 1, 1) - sy(5, -1, 1));
    const float bz = lerp(f, g, dz);
    E[0] = 0.5f * ex;
    E[1] = 0.5f * ey;
    E[2] = 0.5f * ez;
    E[3] = 0.25f * bx;
    E[4] = 0.25f * by;
    E[5] = 0.25f * bz;
}

// relay point of one axis (positions relative to the old cell: a1 in [0,1),
// a2 = a1 + displacement, in cell n2 = floor(a2)): the midpoint, clamped to
// the cell boundary the move crosses
__device__ __forceinline__ float zz_relay(float a1, float a2, int n2)
{
    const float lo = (float)((n2 < 0 ? n2 : 0) + 1), hi = (float)(n2 > 0 ? n2 : 0);
    return fminf(lo, fmaxf(hi, 0.5f * (a1 + a2)));
}

// zigzag deposit of one particle from old cell (i,j,k) / offsets o*, with
// displacement d*, charge q: each half of the move (start -> relay, relay ->
// end) deposits its flux with area weights at its centre.  Same stored sign
// as deposit() (E += cur).  In 2D, z only splits the z flux at the relay
// point (the reference tracks z modulo one cell there).
template <int DIM, class Sink>
__device__ __forceinline__ void deposit_zz(const Sink& snk, int i, int j, int k, float ox, float oy, float oz,
                                           float dx, float dy, float dz, float q)
{
    const float x2 = ox + dx, y2 = oy + dy, z2 = oz + dz;
    const int i2 = (int)floorf(x2), j2 = (int)floorf(y2), k2 = (int)floorf(z2);
    const float xr = zz_relay(ox, x2, i2), yr = zz_relay(oy, y2, j2), zr = zz_relay(oz, z2, k2);
    const float fx1 = -q * (xr - ox), fy1 = -q * (yr - oy), fz1 = -q * (zr - oz);
    const float fx2 = -q * (x2 - xr), fy2 = -q * (y2 - yr), fz2 = -q * (z2 - zr);
    const float wx1 = 0.5f * (ox + xr), wy1 = 0.5f * (oy + yr);
    const float wx2 = 0.5f * (x2 + xr) - (float)i2, wy2 = 0.5f * (y2 + yr) - (float)j2;
    const float wz1 = DIM == 3 ? 0.5f * (oz + zr) : 0.0f;
    const float wz2 = DIM == 3 ? 0.5f * (z2 + zr) - (float)k2 : 0.0f;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        const int ci = i + (h ? i2 : 0), cj = j + (h ? j2 : 0), ck = DIM == 3 ? k + (h ? k2 : 0) : 0;
        const float fx = h ? fx2 : fx1, fy = h ? fy2 : fy1, fz = h ? fz2 : fz1;
        const float wx = h ? wx2 : wx1, wy = h ? wy2 : wy1, wz = h ? wz2 : wz1;
        const auto r00 = snk.row(cj, ck), r10 = snk.row(cj + 1, ck);
        const int c0 = snk.col(ci), c1 = snk.col(ci + 1);
        snk.add(0, r00, c0, fx * (1.0f - wy) * (1.0f - wz));
        snk.add(0, r10, c0, fx * wy * (1.0f - wz));
        snk.add(1, r00, c0, fy * (1.0f - wx) * (1.0f - wz));
        snk.add(1, r00, c1, fy * wx * (1.0f - wz));
        if (DIM == 3) {
            const auto r01 = snk.row(cj, ck + 1), r11 = snk.row(cj + 1, ck + 1);
            snk.add(0, r01, c0, fx * (1.0f - wy) * wz);
            snk.add(0, r11, c0, fx * wy * wz);
            snk.add(1, r01, c0, fy * (1.0f - wx) * wz);
            snk.add(1, r01, c1, fy * wx * wz);
        }
        snk.add(2, r00, c0, fz * (1.0f - wx) * (1.0f - wy));
        snk.add(2, r00, c1, fz * wx * (1.0f - wy));
        snk.add(2, r10, c0, fz * (1.0f - wx) * wy);
        snk.add(2, r10, c1, fz * wx * wy);
    }
}

// One z node plane (window slot zc, 0 <= zc < W) of the 3D Esirkepov deposit
// of one particle: exactly the values deposit<3, ORD, Sink, false> adds in
// that plane.  Splitting a particle's deposit over its W planes lets W
// threads share it.
template <int ORD, class Sink>
__device__ __forceinline__ void deposit3_plane(const Sink& snk, int i, int j, int k, float ox, float oy, float oz,
                                               float dx, float dy, float dz, float q, int zc)
{
    constexpr int W = Win<ORD>::W;
    float S1x[W], dSx[W], S1y[W], dSy[W], S1z[W], dSz[W];
    const int bx = axis_deposit<ORD>(ox, dx, S1x, dSx);
    const int by = axis_deposit<ORD>(oy, dy, S1y, dSy);
    const int bz = axis_deposit<ORD>(oz, dz, S1z, dSz);
    float Ax[W], Ay[W], Cx[W], Cy[W];
    float run_x = 0.0f, run_y = 0.0f, run_z = 0.0f;
    float az = 0.0f, dsz = 0.0f, cz = 0.0f;
#pragma unroll
    for (int t = 0; t < W; ++t) {
        Ax[t] = S1x[t] + 0.5f * dSx[t];
        Ay[t] = S1y[t] + 0.5f * dSy[t];
        run_x += dSx[t];
        run_y += dSy[t];
        run_z += dSz[t];
        Cx[t] = q * run_x;
        Cy[t] = q * run_y;
        if (t == zc) {
            az = S1z[t] + 0.5f * dSz[t];
            dsz = dSz[t];
            cz = q * run_z;
        }
    }
    // a plane at the end of the window that the particle's z shape does not
    // reach adds nothing (a move that does not cross in z leaves one such
    // plane of the five)
    if (az == 0.0f && dsz == 0.0f && (cz == 0.0f || zc == W - 1)) return;
    const float twelfth = 1.0f / 12.0f;
    const float dszt = twelfth * dsz;
    // the plane's weights, each a fused multiply-add: wyz(b) = Ay az + dSy
    // dsz / 12, wxz(t) = Ax az + dSx dsz / 12; Jz as cz Ax Ay + (cz / 12) dSx dSy
    float wxz[W], ga[W], gd[W];
    const float czt = twelfth * cz;
#pragma unroll
    for (int t = 0; t < W; ++t) {
        wxz[t] = fmaf(Ax[t], az, dSx[t] * dszt);
        ga[t] = cz * Ax[t];
        gd[t] = czt * dSx[t];
    }
    int cc[W];
#pragma unroll
    for (int t = 0; t < W; ++t) cc[t] = snk.col(i + bx + t);
#pragma unroll
    for (int b = 0; b < W; ++b) {
        const auto r = snk.row(j + by + b, k + bz + zc);
        const float wyz = fmaf(Ay[b], az, dSy[b] * dszt);
#pragma unroll
        for (int t = 0; t < W; ++t) {
            // (the window's end nodes on an axis the move does not cross
            // carry exact zeros, which need no atomic)
            const float vx = Cx[t] * wyz, vy = Cy[b] * wxz[t];
            if (t < W - 1 && vx != 0.0f) snk.add(0, r, cc[t], vx);
            if (b < W - 1 && vy != 0.0f) snk.add(1, r, cc[t], vy);
        }
    }
    // the last plane's z prefix is the telescoped zero
    if (zc < W - 1) {
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const auto r = snk.row(j + by + b, k + bz + zc);
#pragma unroll
            for (int t = 0; t < W; ++t) {
                const float vz = fmaf(gd[t], dSy[b], ga[t] * Ay[b]);
                if (vz != 0.0f) snk.add(2, r, cc[t], vz);
            }
        }
    }
}

}  // namespace tmpg
