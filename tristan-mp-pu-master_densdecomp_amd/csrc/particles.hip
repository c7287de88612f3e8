// particles.hip -- the particle push: gather + Boris + Esirkepov deposit with
// the per-step re-sort by cell tile fused in.
//
// Reference: move_particles / mover_{1,2,3}ord
// (code/particles_movedeposit.F90:63-88, 356-1271) and deposit_particles /
// densdecomp_{1,2,3}ord (code/particles_movedeposit.F90:1281-2051,
// code/particles.F90:678-1358); the cell sort of reorder_particles
// (code/particles.F90:394-497) is replaced by binning into tiles every step.
//
// k_push_tiled: one workgroup per cell tile.  The tile's E,B with the halo of
//   the shape order are staged in LDS, the tile's particles (both species)
//   stream through with coalesced SoA loads, currents accumulate in LDS with
//   LDS atomics and are added to HBM once per tile.  Each pushed particle is
//   written straight into the segment of the tile it lands in (wave-level
//   slot reservation for the ones that stay), so the next step finds the
//   particles binned again without a separate sort pass.
// k_push_flat: one thread per particle with HBM fields and global atomics;
//   serves the overflow list and the split move/deposit calls.
//
// The deposit uses the stored old position and the push's displacement; the
// reference recomputes the old position from the new momentum
// (x0 = x - u c/gamma), which agrees in exact arithmetic.
#include <algorithm>
#include <vector>

#include "pic_push.h"
#include "tmpg_internal.h"

namespace tmpg {

// Per-phase cycle stamps of the cell-sum push (a profiling build only, built
// with -DTMPG_PHASE_STAMPS): thread 0 of each workgroup adds the shader
// cycles since its previous stamp to the phase's sum at every phase's
// closing barrier.  Phases: 0 stage, 1 push/emit/records (phase A), 2 y/z
// crossers and scans, 3 record index scatter, 4 per-cell sums (phase C),
// 5 record reset, 6 flush to the global currents.
#ifdef TMPG_PHASE_STAMPS
__device__ unsigned long long g_phase_cycles[8];
__device__ unsigned long long g_phase_count[8];
#define PHASE_STAMP_START() unsigned long long phase_t0 = clock64()
#define PHASE_STAMP(n)                                               \
    do {                                                             \
        if (threadIdx.x == 0) {                                      \
            const unsigned long long phase_now = clock64();          \
            atomicAdd(&g_phase_cycles[(n)], phase_now - phase_t0);   \
            atomicAdd(&g_phase_count[(n)], 1ull);                    \
            phase_t0 = phase_now;                                    \
        }                                                            \
    } while (0)
#else
#define PHASE_STAMP_START() do {} while (0)
#define PHASE_STAMP(n) do {} while (0)
#endif

namespace {

struct SpecArgs {
    float qm, q;
    int32_t tcap;
    int64_t ocap;
    PBuf a, b, ob;
    const int32_t* cnt_a;
    int32_t* cnt_b;
    int32_t* ocnt;
};

struct TiledArgs {
    int nx, ny, nz;
    PackInfo pk;
    Tiling tl;
    int fo;                 // field region origin: tile origin - fo
    int FX, FY, FZ;         // staged field region
    int JX, JY, JZ;         // LDS current region (origin: tile origin - 2)
    float c, cinv;
    int flags;              // PUSH_GATHER2 | PUSH_VAY
    int remap;              // XCD-aware tile order
    const float* f6[6];
    float* j3[3];
    SpecArgs sp[2];
};

struct Part {
    int i, j, k;
    float ox, oy, oz, u, v, w, ch;
    int32_t id;
};

template <int DIM, bool HAS_CH, bool HAS_ID>
__device__ __forceinline__ void load_part(const PBuf& b, int64_t s, const PackInfo& pk, Part& p)
{
    unpack_cell(pk, b.cell[s], p.i, p.j, p.k);
    p.ox = b.ox[s];
    p.oy = b.oy[s];
    // 2D keeps a z offset only in the zigzag scheme (b.oz allocated)
    p.oz = (DIM == 3 || b.oz != nullptr) ? b.oz[s] : 0.0f;
    p.u = b.u[s];
    p.v = b.v[s];
    p.w = b.w[s];
    p.ch = HAS_CH ? b.ch[s] : 1.0f;
    p.id = HAS_ID ? b.id[s] : 0;
}

template <int DIM, bool HAS_CH, bool HAS_ID>
__device__ __forceinline__ void store_part(const PBuf& b, int64_t s, const PackInfo& pk, const Part& p)
{
    b.cell[s] = pack_cell(pk, p.i, p.j, p.k);
    b.ox[s] = p.ox;
    b.oy[s] = p.oy;
    if (DIM == 3 || b.oz != nullptr) b.oz[s] = p.oz;
    b.u[s] = p.u;
    b.v[s] = p.v;
    b.w[s] = p.w;
    if (HAS_CH) b.ch[s] = p.ch;
    if (HAS_ID) b.id[s] = p.id;
}

// write a pushed particle to slot `slot` of tile `t` in b, or to the
// overflow list when the segment is full; returns the index used
// (t*tcap+slot, or slots + list index, or -1 when lost)
template <int DIM, bool HAS_CH, bool HAS_ID>
__device__ __forceinline__ int64_t emit(const SpecArgs& s, const PackInfo& pk, int64_t slots, int64_t t,
                                        int32_t slot, const Part& p)
{
    if (slot < s.tcap) {
        const int64_t d = t * s.tcap + slot;
        store_part<DIM, HAS_CH, HAS_ID>(s.b, d, pk, p);
        return d;
    }
    const int oi = atomicAdd(s.ocnt + 1, 1);
    if (oi < s.ocap) {
        store_part<DIM, HAS_CH, HAS_ID>(s.ob, oi, pk, p);
        return slots + oi;
    }
    atomicAdd(s.ocnt + 2, 1);
    return -1;
}

// push flags: the 2nd-order gather quirk; the Vay pusher instead of Boris
constexpr int PUSH_GATHER2 = 1, PUSH_VAY = 2;

// field gather and momentum push of p; returns the displacement.
// VAY_OK = false compiles the Boris push only (the Boris instantiations of
// the 3D cell-sum kernel, whose register budget is tight)
template <bool VAY_OK = true>
__device__ __forceinline__ void push_fields(const float* E, int flags, float c, float cinv, float qm, Part& p,
                                            float& dx, float& dy, float& dz)
{
    const float se = 0.5f * qm, sb = 0.5f * qm * cinv;
    if (VAY_OK && (flags & PUSH_VAY))
        vay(c, cinv, se * E[0], se * E[1], se * E[2], sb * E[3], sb * E[4], sb * E[5], p.u, p.v, p.w, dx, dy, dz);
    else
        boris(c, cinv, se * E[0], se * E[1], se * E[2], sb * E[3], sb * E[4], sb * E[5], p.u, p.v, p.w, dx, dy,
              dz);
}

// ORD 0 is the original scheme (trilinear gather of the staggered fields)
template <int DIM, int ORD, bool VAY_OK = true, class Src>
__device__ __forceinline__ void push_one(const Src& src, int flags, float c, float cinv, float qm, int si, int sj,
                                         int sk, Part& p, float& dx, float& dy, float& dz)
{
    float E[6];
    if constexpr (ORD == 0) gather_zz<DIM>(src, si, sj, sk, p.ox, p.oy, p.oz, E);
    else gather<DIM, ORD>(src, (flags & PUSH_GATHER2) != 0, si, sj, sk, p.ox, p.oy, p.oz, E);
    const float se = 0.5f * qm, sb = 0.5f * qm * cinv;
    if (VAY_OK && (flags & PUSH_VAY))
        vay(c, cinv, se * E[0], se * E[1], se * E[2], sb * E[3], sb * E[4], sb * E[5], p.u, p.v, p.w, dx, dy, dz);
    else
        boris(c, cinv, se * E[0], se * E[1], se * E[2], sb * E[3], sb * E[4], sb * E[5], p.u, p.v, p.w, dx, dy,
              dz);
}

__device__ __forceinline__ void advance(Part& p, int dim, int nx, int ny, int nz, float dx, float dy, float dz,
                                        bool z_offset_2d = false)
{
    p.i = wrap1(p.i + advance_offset(p.ox, dx), nx);
    p.j = wrap1(p.j + advance_offset(p.oy, dy), ny);
    if (dim == 3) p.k = wrap1(p.k + advance_offset(p.oz, dz), nz);
    // the 2D zigzag scheme tracks z modulo one cell, like the reference's
    // wrap of z into its single plane (code/particles_movedeposit.F90:1553-1633)
    else if (z_offset_2d) (void)advance_offset(p.oz, dz);
}

// Esirkepov deposit of the shape ORD, or the zigzag deposit for ORD 0
template <int DIM, int ORD, class Sink, bool SKIP_ZERO = true>
__device__ __forceinline__ void deposit_any(const Sink& snk, int i, int j, int k, float ox, float oy, float oz,
                                            float dx, float dy, float dz, float q)
{
    if constexpr (ORD == 0) deposit_zz<DIM>(snk, i, j, k, ox, oy, oz, dx, dy, dz, q);
    else deposit<DIM, ORD, Sink, SKIP_ZERO>(snk, i, j, k, ox, oy, oz, dx, dy, dz, q);
}

// --------------------------------------------------------------- tiled
template <int DIM, int ORD, bool HAS_CH, bool HAS_ID>
__global__ void __launch_bounds__(512) k_push_tiled(TiledArgs a)
{
    extern __shared__ float sm[];
    const int nb = gridDim.x;
    int64_t t = blockIdx.x;
    if (a.remap) t = (int64_t)(blockIdx.x % 8) * (nb / 8) + blockIdx.x / 8;
    const Tiling& tl = a.tl;
    const int ti = (int)(t % tl.ntx);
    const int tj = (int)((t / tl.ntx) % tl.nty);
    const int tk = (int)(t / ((int64_t)tl.ntx * tl.nty));
    const int i0 = ti * tl.tx, j0 = tj * tl.ty, k0 = tk * tl.tz;
    const int fplane = a.FX * a.FY * a.FZ;
    const int jplane = a.JX * a.JY * a.JZ;
    float* sf = sm;
    double* sj = reinterpret_cast<double*>(sm + ((6 * fplane + 1) & ~1));
    const int tid = threadIdx.x, NT = blockDim.x;
    // stage E,B (primal in 3D; the staggered fields of the zigzag scheme,
    // ORD 0) with the halo, wrapping periodically
    constexpr bool NODE_MAJOR = DIM == 3 && ORD != 0;
    for (int e = tid; e < fplane; e += NT) {
        const int li = e % a.FX;
        const int r = e / a.FX;
        const int lj = r % a.FY;
        const int lk = r / a.FY;
        const int gi = wrap1(i0 - a.fo + li, a.nx);
        const int gj = wrap1(j0 - a.fo + lj, a.ny);
        const int gk = (DIM == 3) ? wrap1(k0 - a.fo + lk, a.nz) : 0;
        const int64_t g = gi + (int64_t)a.nx * (gj + (int64_t)a.ny * gk);
        if (NODE_MAJOR) {
            // node-major: six components per node
            float2* d = reinterpret_cast<float2*>(sf) + 3 * e;
            d[0] = make_float2(a.f6[0][g], a.f6[1][g]);
            d[1] = make_float2(a.f6[2][g], a.f6[3][g]);
            d[2] = make_float2(a.f6[4][g], a.f6[5][g]);
        } else {
#pragma unroll
            for (int q = 0; q < 6; ++q) sf[q * fplane + e] = a.f6[q][g];
        }
    }
    for (int e = tid; e < 3 * jplane; e += NT) sj[e] = 0.0;
    __syncthreads();

    using Src = typename std::conditional<NODE_MAJOR, LdsSrc6, LdsSrc>::type;
    Src src;
    if constexpr (NODE_MAJOR) src = LdsSrc6{sf, a.FX, a.FY};
    else src = LdsSrc{sf, a.FX, a.FY, fplane};
    const LdsSink snk{sj, a.JX, a.JY, jplane};
    const int lane = __lane_id();
    const uint64_t ltmask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll 1
    for (int s = 0; s < 2; ++s) {
        const SpecArgs& sp = a.sp[s];
        const int n = min(sp.cnt_a[t], sp.tcap);
        const int64_t base = t * sp.tcap;
        const int64_t slots = (int64_t)tl.ntiles * sp.tcap;
        for (int p0 = 0; p0 < n; p0 += NT) {
            const int pi = p0 + tid;
            const bool active = pi < n;
            Part p;
            int64_t td = -1;
            if (active) {
                load_part<DIM, HAS_CH, HAS_ID>(sp.a, base + pi, a.pk, p);
                const int li = p.i - i0, lj = p.j - j0, lk = p.k - k0;
                const int oi = p.i, oj = p.j, ok = p.k;
                const float ox = p.ox, oy = p.oy, oz = p.oz;
                float dx, dy, dz;
                push_one<DIM, ORD>(src, a.flags, a.c, a.cinv, sp.qm, li + a.fo, lj + a.fo,
                                   (DIM == 3) ? lk + a.fo : 0, p, dx, dy, dz);
                advance(p, DIM, a.nx, a.ny, a.nz, dx, dy, dz, ORD == 0);
                deposit_any<DIM, ORD>(snk, li + 2, lj + 2, (DIM == 3) ? lk + 2 : 0, ox, oy, oz, dx, dy, dz,
                                      sp.q * p.ch);
                (void)oi; (void)oj; (void)ok;
                td = tile_of(tl, p.i, p.j, p.k);
            }
            // slots in b: one reservation per wave for the particles that stay
            const bool stay = active && td == t;
            const uint64_t m = __ballot(stay);
            int32_t slot = -1;
            if (m) {
                const int leader = __ffsll((unsigned long long)m) - 1;
                int32_t b0 = 0;
                if (lane == leader) b0 = atomicAdd(sp.cnt_b + t, __popcll(m));
                b0 = __shfl(b0, leader, 64);
                if (stay) slot = b0 + __popcll(m & ltmask);
            }
            if (active && !stay) slot = atomicAdd(sp.cnt_b + td, 1);
            if (active) emit<DIM, HAS_CH, HAS_ID>(sp, a.pk, slots, td, slot, p);
        }
    }
    __syncthreads();
    // one pass of global atomics for the tile's currents
    for (int e = tid; e < jplane; e += NT) {
        const int li = e % a.JX;
        const int r = e / a.JX;
        const int lj = r % a.JY;
        const int lk = r / a.JY;
        const int gi = wrap1(i0 - 2 + li, a.nx);
        const int gj = wrap1(j0 - 2 + lj, a.ny);
        const int gk = (DIM == 3) ? wrap1(k0 - 2 + lk, a.nz) : 0;
        const int64_t g = gi + (int64_t)a.nx * (gj + (int64_t)a.ny * gk);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const float v = (float)sj[q * jplane + e];
            if (v != 0.0f) atomicAdd(a.j3[q] + g, v);
        }
    }
}

// --------------------------------------------------- tiled, per-cell sums
// 3D push whose deposit is reduced per cell before touching LDS atomics.
// Phase A pushes the tile's particles like k_push_tiled, but keeps each
// particle's deposit record (old offsets, displacement, charge, old cell) in
// LDS instead of depositing it, unless its y/z shape window leaves the
// 4-node window [-1,2] (a y or z cell crossing), in which case it deposits at
// once.  Phase B counting-sorts the records by cell.  Phase C gives every
// cell to four threads (one per z slot), which sum the Esirkepov stencils of
// all of the cell's particles in registers -- x window [-2,3] (6 nodes, x
// crossings included), y/z windows [-1,2] -- and add the sums to the tile's
// LDS currents once.  That replaces ~300 LDS atomics per particle by ~250 per
// cell.
constexpr int CELLS_T = 1024;       // threads per workgroup
constexpr int CELLS_R = 2816;       // records per pass (both species)

struct CellsLds {
    int fplane, jplane, nc;
    size_t f_off, j_off, rec_off, rq_off, rc_off, idx_off, perm_off, hist_off, cnt_off, start_off, fill_off, bytes;
};

__host__ __device__ inline CellsLds cells_lds(int FX, int FY, int FZ, int JX, int JY, int JZ, int nc)
{
    CellsLds L;
    L.fplane = FX * FY * FZ;
    L.jplane = JX * JY * JZ;
    L.nc = nc;
    size_t o = 0;
    L.j_off = o;   o += sizeof(double) * 3 * (size_t)L.jplane;
    o = (o + 15) & ~(size_t)15;    // the staged fields are read 16 bytes at a time
    L.f_off = o;   o += sizeof(float) * 8 * (size_t)L.fplane;
    L.rec_off = o; o += sizeof(float) * 6 * (size_t)CELLS_R;
    L.rq_off = o;  o += sizeof(float) * (size_t)CELLS_R;
    L.cnt_off = o; o += sizeof(int) * (size_t)nc;
    L.start_off = o; o += sizeof(int) * (size_t)nc;
    L.fill_off = o; o += sizeof(int) * (size_t)nc;
    L.rc_off = o;  o += sizeof(uint16_t) * (size_t)CELLS_R;
    L.idx_off = o; o += sizeof(uint16_t) * (size_t)CELLS_R;
    L.perm_off = o; o += sizeof(uint16_t) * (size_t)((nc + 1) & ~1);
    L.hist_off = o; o += sizeof(int) * 64;
    o = (o + 15) & ~(size_t)15;
    o += 16;       // record counter
    L.bytes = o;
    return L;
}

// place the W window values starting at node `base` into an NW-node window
// starting at node `first` (values outside are dropped: callers guarantee
// they are zero)
template <int W, int NW>
__device__ __forceinline__ void place(const float* v, int base, int first, float* out)
{
#pragma unroll
    for (int s = 0; s < NW; ++s) {
        float acc = 0.0f;
#pragma unroll
        for (int t = 0; t < W; ++t) acc = (s == base - first + t) ? v[t] : acc;
        out[s] = acc;
    }
}

// Shape profile of one axis of a cell-summed record on a fixed node window:
// S = old weights, dS = new - old.  profile6: nodes [-2,3] (x, cell
// crossings included); profile4: nodes [-1,2] (y, z: the record's move keeps
// the support inside the window).  Order 3 writes the pieces straight into
// their slots: the old piece is the particle's own cell (m0 = 0), the new
// one is m1 = floor(o + d) in {-1, 0, 1}, and a crossing splits dS at the
// knot K = max(m0, m1) exactly as axis_deposit does.  Other orders go
// through axis_deposit and place the window.
template <int ORD>
__device__ __forceinline__ void profile6(float o, float d, float* S6, float* dS6)
{
    if constexpr (ORD == 3) {
        using P = Piece<3>;
        float w[4], t0[4], t1[4];
        P::w(o, w);
        const int m1 = (int)floorf(o + d);
        const float K = (m1 > 0) ? 1.0f : 0.0f;
        const float b0 = (m1 == 0) ? o + d : K;
        const float d0 = (m1 == 0) ? d : K - o;
        P::dw(o, b0, d0, t0);
        // second segment on piece m1 (length zero, so all zero, without a crossing)
        const float d2 = d - d0;
        const float a1 = K - (float)m1;
        P::dw(a1, a1 + d2, d2, t1);
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            S6[u] = (u >= 1 && u <= 4) ? w[u - 1] : 0.0f;
            const float v = (u >= 1 && u <= 4) ? t0[u - 1] : 0.0f;
            const float up = (u >= 2) ? t1[u - 2] : 0.0f;    // m1 = +1: nodes 0..3
            const float dn = (u <= 3) ? t1[u] : 0.0f;        // m1 = -1: nodes -2..1
            dS6[u] = v + ((m1 > 0) ? up : dn);
        }
    } else {
        constexpr int W = Win<ORD>::W;
        float S1[W], dS[W];
        const int b = axis_deposit<ORD>(o, d, S1, dS);
        place<W, 6>(S1, b, -2, S6);
        place<W, 6>(dS, b, -2, dS6);
    }
}

template <int ORD>
__device__ __forceinline__ void profile4(float o, float d, float* S4, float* dS4)
{
    if constexpr (ORD == 3) {
        Piece<3>::w(o, S4);
        Piece<3>::dw(o, o + d, d, dS4);
    } else {
        constexpr int W = Win<ORD>::W;
        float S1[W], dS[W];
        const int b = axis_deposit<ORD>(o, d, S1, dS);
        place<W, 4>(S1, b, -1, S4);
        place<W, 4>(dS, b, -1, dS4);
    }
}

// y/z shape window of a move inside nodes [-1,2]?
template <int ORD>
__device__ __forceinline__ bool fits4(float o, float d)
{
    const int m0 = piece_of<ORD>(o), m1 = piece_of<ORD>(o + d);
    const int lo = (m0 < m1 ? m0 : m1) + Piece<ORD>::first;
    const int hi = (m0 > m1 ? m0 : m1) + Piece<ORD>::first + Piece<ORD>::cnt - 1;
    return lo >= -1 && hi <= 2;
}

// lane S of every group of four lanes, to all four (DPP quad permutation)
template <int S>
__device__ __forceinline__ float quad_bcast(float v)
{
    constexpr int ctrl = S | (S << 2) | (S << 4) | (S << 6);
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, true));
}

// Per-cell register sums of one cell's records (idx[r0 .. r0+n) index the
// records) for slot Q (0..3) of the cell: Jx and Jy at z node slot Q, and Jz
// at y node slot Q of all three z prefix slots (so the Jz work is spread over
// the four lanes), added to the tile's LDS currents once.  The four slots of
// a cell are four adjacent lanes (a quad) that walk the cell's records
// together: lane Q computes the shape profile of axis Q (x, y, z; lane 3
// repeats x) and the quad shares the three profiles by DPP broadcasts, so
// each record's shapes are computed once.  A lane holds 56 sums, which keeps
// the kernel in 128 VGPRs (four waves per SIMD).  The y and z profiles come from the 6-node window, whose nodes
// -1..2 are the 4-node window of a record that keeps its y/z piece.
template <int ORD>
__device__ __forceinline__ void cell_sums(int Q, int c, int r0, int n, int tx, int ty,
                                          const uint16_t* __restrict__ idx, const float* __restrict__ rec,
                                          const float* __restrict__ rq, const LdsSink& snk)
{
    const float twelfth = 1.0f / 12.0f;
    const int r1 = r0 + n;
    if (r1 == r0) return;
    const int li = c % tx, lj = (c / tx) % ty, lk = c / (tx * ty);
    const int axis = Q == 3 ? 0 : Q;
    float ax[5][4], ay[6][3], az[3][6];
#pragma unroll
    for (int u = 0; u < 5; ++u)
#pragma unroll
        for (int b = 0; b < 4; ++b) ax[u][b] = 0.0f;
#pragma unroll
    for (int u = 0; u < 6; ++u)
#pragma unroll
        for (int b = 0; b < 3; ++b) ay[u][b] = 0.0f;
#pragma unroll
    for (int s = 0; s < 3; ++s)
#pragma unroll
        for (int u = 0; u < 6; ++u) az[s][u] = 0.0f;
    for (int rr = r0; rr < r1; ++rr) {
        const int r = idx[rr];
        const float o = rec[axis * CELLS_R + r], d = rec[(3 + axis) * CELLS_R + r];
        const float q = rq[r];
        // each lane forms A = S + dS/2 of its own axis, and the quad shares A
        // and dS
        float A6[6], dS6[6];
        profile6<ORD>(o, d, A6, dS6);
#pragma unroll
        for (int u = 0; u < 6; ++u) A6[u] = A6[u] + 0.5f * dS6[u];
        float dSx[6], Ax[6], dSy[4], Ay[4];
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            Ax[u] = quad_bcast<0>(A6[u]);
            dSx[u] = quad_bcast<0>(dS6[u]);
        }
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            Ay[b] = quad_bcast<1>(A6[1 + b]);
            dSy[b] = quad_bcast<1>(dS6[1 + b]);
        }
        float azc, dzc, cz0, cz1, cz2;
        {
            float Az[4], dSz[4];
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                Az[b] = quad_bcast<2>(A6[1 + b]);
                dSz[b] = quad_bcast<2>(dS6[1 + b]);
            }
            const float s0 = dSz[0], s1 = s0 + dSz[1], s2 = s1 + dSz[2];
            dzc = Q == 0 ? dSz[0] : Q == 1 ? dSz[1] : Q == 2 ? dSz[2] : dSz[3];
            azc = Q == 0 ? Az[0] : Q == 1 ? Az[1] : Q == 2 ? Az[2] : Az[3];
            cz0 = q * s0;
            cz1 = q * s1;
            cz2 = q * s2;
        }
        float Cx[5], Cy[3];
        {
            float run = 0.0f;
#pragma unroll
            for (int u = 0; u < 5; ++u) { run += dSx[u]; Cx[u] = q * run; }
            run = 0.0f;
#pragma unroll
            for (int u = 0; u < 3; ++u) { run += dSy[u]; Cy[u] = q * run; }
        }
        // every update is a chain of fused multiply-adds into the sum (the
        // compiler does not reassociate "s += a*b + c*d" into two of them)
        const float dzt = twelfth * dzc;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const float wyz = fmaf(Ay[b], azc, dSy[b] * dzt);
#pragma unroll
            for (int u = 0; u < 5; ++u) ax[u][b] = fmaf(Cx[u], wyz, ax[u][b]);
        }
#pragma unroll
        for (int u = 0; u < 6; ++u) {
            const float wxz = fmaf(Ax[u], azc, dSx[u] * dzt);
#pragma unroll
            for (int b = 0; b < 3; ++b) ay[u][b] = fmaf(Cy[b], wxz, ay[u][b]);
        }
        {
            // Jz at y slot Q: cz(s) (Ax Ay(Q) + dSx dSy(Q) / 12) for the three
            // z prefix slots s
            const float ayq = Q == 0 ? Ay[0] : Q == 1 ? Ay[1] : Q == 2 ? Ay[2] : Ay[3];
            const float dyq = twelfth * (Q == 0 ? dSy[0] : Q == 1 ? dSy[1] : Q == 2 ? dSy[2] : dSy[3]);
#pragma unroll
            for (int u = 0; u < 6; ++u) {
                const float w = fmaf(Ax[u], ayq, dSx[u] * dyq);
                az[0][u] = fmaf(cz0, w, az[0][u]);
                az[1][u] = fmaf(cz1, w, az[1][u]);
                az[2][u] = fmaf(cz2, w, az[2][u]);
            }
        }
    }
    // add the cell's sums to the tile currents (the window's outer nodes
    // collect only from records that cross there, so many sums are exact
    // zeros, which need no atomic): node (li-2+u, lj-1+b,
    // lk-1+Q) sits at LDS (li+u, lj+1+b, lk+1+Q)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int row = snk.row(lj + 1 + b, lk + 1 + Q);
#pragma unroll
        for (int u = 0; u < 5; ++u)
            if (ax[u][b] != 0.0f) snk.add(0, row, li + u, ax[u][b]);
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const int row = snk.row(lj + 1 + b, lk + 1 + Q);
#pragma unroll
        for (int u = 0; u < 6; ++u)
            if (ay[u][b] != 0.0f) snk.add(1, row, li + u, ay[u][b]);
    }
#pragma unroll
    for (int s = 0; s < 3; ++s) {
        const int row = snk.row(lj + 1 + Q, lk + 1 + s);
#pragma unroll
        for (int u = 0; u < 6; ++u)
            if (az[s][u] != 0.0f) snk.add(2, row, li + u, az[s][u]);
    }
}

// The tile extents TX x TY x TZ are compile-time constants (the default
// 8x8x4 tile): the staged-field and current strides are then constants too,
// so the 64 nodes of a gather and the sums of a cell address LDS with one
// base register plus immediate offsets, and cell <-> tile coordinates are
// shifts instead of integer divisions.
template <int ORD, bool HAS_CH, bool HAS_ID, bool MULTI, bool VAY, int TX, int TY, int TZ>
__global__ void __launch_bounds__(CELLS_T) k_push_cells(TiledArgs a, CellsLds L)
{
    constexpr int FX = TX + 3, FY = TY + 3, FZ = TZ + 3;   // staged fields: nodes -1 .. T+1
    constexpr int JX = TX + 5, JY = TY + 5, JZ = TZ + 5;   // tile currents: nodes -2 .. T+2
    constexpr int fplane = FX * FY * FZ, jplane = JX * JY * JZ, nc = TX * TY * TZ;
    extern __shared__ __align__(16) unsigned char smb[];
    double* sj = reinterpret_cast<double*>(smb + L.j_off);
    float4* sf4 = reinterpret_cast<float4*>(smb + L.f_off);
    float4* sfp = reinterpret_cast<float4*>(smb + L.f_off + sizeof(float4) * (size_t)fplane);
    float* rec = reinterpret_cast<float*>(smb + L.rec_off);    // [6][R]: ox oy oz dx dy dz
    float* rq = reinterpret_cast<float*>(smb + L.rq_off);
    int* cnt = reinterpret_cast<int*>(smb + L.cnt_off);
    int* start = reinterpret_cast<int*>(smb + L.start_off);
    int* fill = reinterpret_cast<int*>(smb + L.fill_off);
    uint16_t* rc = reinterpret_cast<uint16_t*>(smb + L.rc_off);
    uint16_t* idx = reinterpret_cast<uint16_t*>(smb + L.idx_off);
    uint16_t* perm = reinterpret_cast<uint16_t*>(smb + L.perm_off);   // cells by descending record count
    int* hist = reinterpret_cast<int*>(smb + L.hist_off);
    // record counts packed in one word: low 16 bits cell-summed records
    // (stored from 0 up), high 16 bits y/z crossers (stored from the top down)
    unsigned* nrec = reinterpret_cast<unsigned*>(smb + L.bytes - 16);

    const Tiling& tl = a.tl;
    int tid = threadIdx.x;
    constexpr int NT = CELLS_T;
    static_assert(fplane <= NT, "one staged node per thread");
    PHASE_STAMP_START();
    // persistent workgroups: workgroup b takes the tiles tbase, tbase +
    // tstride, ... below tend.  With a.remap (the grid and the tile count
    // multiples of 8) the workgroups of one XCD (b mod 8) share one
    // contiguous eighth of the tiles, so neighbouring tiles' halos meet in
    // that XCD's L2.
    // (tile indices in 32 bits: fewer than 2^31 tiles of 256 cells)
    const int G = gridDim.x;
    const int ntiles = (int)tl.ntiles;
    int tbase = blockIdx.x, tstride = G, tend = ntiles;
    if (a.remap) {
        const int per = ntiles / 8;
        tbase = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
        tstride = G / 8;
        tend = (int)(blockIdx.x % 8 + 1) * per;
    }
    // tile tt's first cell
    auto tile_origin = [&](int tt, int& x0, int& y0, int& z0) {
        const unsigned u = (unsigned)tt, ntx = (unsigned)tl.ntx, nty = (unsigned)tl.nty;
        const unsigned q = u / ntx;
        x0 = (int)(u - q * ntx) * TX;
        y0 = (int)(q % nty) * TY;
        z0 = (int)(q / nty) * TZ;
    };
    // the six field values of staged node `tid` of tile tt: the next tile's
    // are read before this tile's flush and written to the (by then unused)
    // field arrays after it, so their latency hides behind the flush's
    // atomics
    float nf[6];
    auto fetch = [&](int tt) {
        int x0, y0, z0;
        tile_origin(tt, x0, y0, z0);
        const int e = min(tid, fplane - 1);     // threads past the staged nodes read a valid one
        const int li = e % FX;
        const int r = e / FX;
        const int lj = r % FY;
        const int lk = r / FY;
        const int gi = wrap1(x0 - 1 + li, a.nx);
        const int gj = wrap1(y0 - 1 + lj, a.ny);
        const int gk = wrap1(z0 - 1 + lk, a.nz);
        const int64_t g = gi + (int64_t)a.nx * (gj + (int64_t)a.ny * gk);
#pragma unroll
        for (int q = 0; q < 6; ++q) nf[q] = a.f6[q][g];
    };
    // two node-major arrays of 16-byte entries: (Ex, Ey, Ez, Bx) of the
    // node, and (By, Bz) of the node and of its +x neighbour
    auto put = [&]() {
        sf4[tid] = make_float4(nf[0], nf[1], nf[2], nf[3]);
        const float2 byz = make_float2(nf[4], nf[5]);
        float2* half = reinterpret_cast<float2*>(sfp);
        half[2 * tid] = byz;
        if (tid > 0) half[2 * tid - 1] = byz;
    };
    int t = tbase;
    if (t < tend && tid < fplane) {
        fetch(t);
        put();
    }
    // the currents and record counters start cleared; afterwards the flush
    // clears each current entry it reads and the tile's final drain clears
    // the counters, so a tile needs no clearing pass of its own
    for (int e = tid; e < 3 * jplane; e += NT) sj[e] = 0.0;
    for (int e = tid; e < nc; e += NT) { cnt[e] = 0; fill[e] = 0; }
    if (tid < 64) hist[tid] = 0;
    if (tid == 0) *nrec = 0u;
    __syncthreads();
    int i0 = 0, j0 = 0, k0 = 0;

    const LdsSink snk{sj, JX, JY, jplane};
    uint64_t ltmask = 0;
    int lane = __lane_id();
    // tile of a (wrapped, so non-negative) cell: the tile extents are powers
    // of two in the default tile, and a grid of fewer than 2^31 cells has
    // fewer than 2^23 such tiles, so 32-bit arithmetic holds the index
    auto tile_at = [&](int i, int j, int k) -> int {
        return (int)((unsigned)i / TX + (unsigned)tl.ntx * ((unsigned)j / TY + (unsigned)tl.nty * ((unsigned)k / TZ)));
    };

    // ---- the records so far: crossers deposited directly, the rest
    // counting-sorted by cell and summed per cell (phases B and C); the
    // record space is then empty again
    // (last: the tile's final drain: the fields are no longer read, so the
    // next tile's are read at its start and stored after the crossers)
    bool more = false;
    auto drain = [&](bool last) {
        if (last) {
            // read unconditionally (the last tile's own fields again when
            // there is no next tile), so that the values are not held in
            // scratch through the crossers' branches
            more = t + tstride < tend;
            fetch(more ? t + tstride : t);
        }
        // ---- direct deposit of the y/z crossers, one z node plane of one
        // crosser per thread (so every wave shares the work)
        const int nr = (int)(*nrec & 0xffffu);
        const int nd = (int)(*nrec >> 16);
        constexpr int WZ = Win<ORD>::W;
        for (int w = tid; w < nd * WZ; w += NT) {
            const int k = w / WZ, zc = w - k * WZ;
            const int r = CELLS_R - 1 - k;
            const int c = rc[r];
            const int li = c % TX, lj = (c / TX) % TY, lk = c / (TX * TY);
            deposit3_plane<ORD>(snk, li + 2, lj + 2, lk + 2, rec[r], rec[CELLS_R + r], rec[2 * CELLS_R + r],
                                rec[3 * CELLS_R + r], rec[4 * CELLS_R + r], rec[5 * CELLS_R + r], rq[r], zc);
        }

        // ---- phase B: counting sort of the cell-summed records by cell (the
        // crossers above run on the first waves; the scans and the cell
        // order on the last two)
        const int wv = tid >> 6, nwv = NT >> 6, lane6 = tid & 63;
        if (wv == nwv - 1) {
            // exclusive scan of cnt[0..nc)
            int carry = 0;
            for (int c0 = 0; c0 < nc; c0 += 64) {
                const int c = c0 + lane6;
                const int v = c < nc ? cnt[c] : 0;
                int x = v;
    #pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    // lane lane6 - o's x (the lane index comes from the
                    // per-tile thread index, so no address is held across
                    // the tile loop)
                    const int y = __builtin_amdgcn_ds_bpermute((lane6 - o) << 2, x);
                    if (lane6 >= o) x += y;
                }
                if (c < nc) start[c] = carry + x - v;
                carry += __builtin_amdgcn_readlane(x, 63);
            }
        } else if (wv == nwv - 2) {
            // histogram of the record counts (bins 0..63), then the cells in
            // descending record count: bin b starts after all larger bins;
            // the cells a wave then sums together have similar loop lengths.
            // One wave does both (its LDS operations complete in order), in
            // the shadow of the crossers.
            for (int c = lane6; c < nc; c += 64) atomicAdd(hist + min(cnt[c], 63), 1);
            __builtin_amdgcn_wave_barrier();
            const int h = hist[lane6];
            int x = h;
    #pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __builtin_amdgcn_ds_bpermute((lane6 + o) << 2, x);
                if (lane6 + o < 64) x += y;
            }
            hist[lane6] = x - h;
            __builtin_amdgcn_wave_barrier();
            for (int c = lane6; c < nc; c += 64) perm[atomicAdd(hist + min(cnt[c], 63), 1)] = (uint16_t)c;
        }
        __syncthreads();
        PHASE_STAMP(2);
        if (last && more && tid < fplane) put();
        for (int r = tid; r < nr; r += NT) {
            const int c = rc[r];
            const int pos = atomicAdd(fill + c, 1);
            idx[start[c] + pos] = (uint16_t)r;
        }
        __syncthreads();
        PHASE_STAMP(3);

        // ---- phase C: per-cell register sums, four adjacent lanes per cell
        // (one per z slot)
        static_assert(NT % 4 == 0, "quads of lanes");
        for (int wk = tid; wk < 4 * nc; wk += NT) {
            const int c = perm[wk >> 2];
            cell_sums<ORD>(wk & 3, c, start[c], cnt[c], TX, TY, idx, rec, rq, snk);
        }
        __syncthreads();
        PHASE_STAMP(4);
        for (int e = tid; e < nc; e += NT) { cnt[e] = 0; fill[e] = 0; }
        if (tid < 64) hist[tid] = 0;
        if (tid == 0) *nrec = 0u;
        // after the tile's final drain the barrier closing the tile orders
        // these stores before the next tile's phase A
        if (!last) __syncthreads();
        PHASE_STAMP(5);
    };

    // ---- phase A: push, emit, record or deposit at once.  A tile with more
    // particles than records drains the records whenever a round of NT
    // might not fit (the test is block-uniform: read between two barriers)
    // one round of phase A: NT particles of species `sp` starting at p0;
    // `pn`/`an` carry the prefetched particle of the next round
    auto round = [&](const SpecArgs& sp, int n, int64_t base, int64_t slots, int p0, Part& pn, bool& an) {
        const bool active = an;
        Part p = pn;
        an = p0 + NT + tid < n;
        if (an) load_part<3, HAS_CH, HAS_ID>(sp.a, base + p0 + NT + tid, a.pk, pn);
        int td = -1;
        float ox = 0.0f, oy = 0.0f, oz = 0.0f, dx = 0.0f, dy = 0.0f, dz = 0.0f, q = 0.0f;
        int li = 0, lj = 0, lk = 0;
        bool fits = false;
        if (active) {
            li = p.i - i0; lj = p.j - j0; lk = p.k - k0;
            ox = p.ox; oy = p.oy; oz = p.oz;
            {
                float E[6];
                gather3_lds<ORD, FX, FY, 1>(sf4, sfp, (a.flags & PUSH_GATHER2) != 0, li + 1, lj + 1, lk + 1, p.ox, p.oy,
                                         p.oz, E);
                push_fields<VAY>(E, a.flags, a.c, a.cinv, sp.qm, p, dx, dy, dz);
            }
            advance(p, 3, a.nx, a.ny, a.nz, dx, dy, dz);
            q = sp.q * p.ch;
            fits = fits4<ORD>(oy, dy) && fits4<ORD>(oz, dz);
            td = tile_at(p.i, p.j, p.k);
        }
        // slots in b: leavers reserve one each in their new tile, the
        // wave's stayers one block in this tile.  Both global atomics are
        // issued here and their results used only after the deposit record
        // is written, so the record work covers their latency.
        const bool stay = active && td == t;
        const uint64_t m = __ballot(stay);
        int32_t lslot = 0, b0 = 0;
        if (active && !stay) lslot = atomicAdd(sp.cnt_b + td, 1);
        const int leader = m ? __ffsll((unsigned long long)m) - 1 : 0;
        if (m && lane == leader) b0 = atomicAdd(sp.cnt_b + t, __popcll(m));
        // record slots, one LDS atomic per wave: cell-summed records grow
        // from 0 up, y/z crossers from the top down.  A wave whose block
        // would make the two meet takes its block back and deposits at
        // once; any block reserved while such a failed one is pending
        // fails too, so the reserved blocks stay contiguous.
        int r = -1;
        {
            const uint64_t mf = __ballot(active && fits), mc = __ballot(active && !fits);
            const unsigned inc = (unsigned)__popcll(mf) + ((unsigned)__popcll(mc) << 16);
            unsigned old = 0;
            if (inc && lane == 0) old = atomicAdd(nrec, inc);
            old = __shfl(old, 0, 64);
            const int lo = (int)(old & 0xffffu) + __popcll(mf), hi = (int)(old >> 16) + __popcll(mc);
            const bool ok = lo + hi <= CELLS_R;
            if (inc && !ok && lane == 0) atomicSub(nrec, inc);
            if (ok && active)
                r = fits ? (int)(old & 0xffffu) + __popcll(mf & ltmask)
                         : CELLS_R - 1 - ((int)(old >> 16) + __popcll(mc & ltmask));
        }
        if (active) {
            const int c = li + TX * (lj + TY * lk);
            if (r >= 0) {
                rec[0 * CELLS_R + r] = ox;
                rec[1 * CELLS_R + r] = oy;
                rec[2 * CELLS_R + r] = oz;
                rec[3 * CELLS_R + r] = dx;
                rec[4 * CELLS_R + r] = dy;
                rec[5 * CELLS_R + r] = dz;
                rq[r] = q;
                rc[r] = (uint16_t)c;
                if (fits) atomicAdd(cnt + c, 1);
            } else {
                // record space exhausted: deposit this particle at once
                deposit<3, ORD>(snk, li + 2, lj + 2, lk + 2, ox, oy, oz, dx, dy, dz, q);
            }
        }
        b0 = __shfl(b0, leader, 64);
        const int32_t slot = stay ? b0 + __popcll(m & ltmask) : lslot;
        if (active) emit<3, HAS_CH, HAS_ID>(sp, a.pk, slots, td, slot, p);
    };

#pragma unroll 1
    for (; t < tend; t += tstride) {
    // the thread and lane indices are opaque to the compiler from here on,
    // so what the tile's phases derive from them is computed per tile
    // instead of being hoisted out of the tile loop and held in registers
    // through every phase
    asm volatile("" : "+v"(tid), "+v"(lane));
    ltmask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    tile_origin(t, i0, j0, k0);
    PHASE_STAMP(0);

    if constexpr (!MULTI) {
#pragma unroll 1
    for (int s = 0; s < 2; ++s) {
        const SpecArgs& sp = a.sp[s];
        const int n = min(sp.cnt_a[t], sp.tcap);
        const int64_t base = (int64_t)t * sp.tcap;
        const int64_t slots = (int64_t)tl.ntiles * sp.tcap;
        // the next round's particles are loaded while this round is pushed
        Part pn;
        bool an = tid < n;
        if (an) load_part<3, HAS_CH, HAS_ID>(sp.a, base + tid, a.pk, pn);
        for (int p0 = 0; p0 < n; p0 += NT) round(sp, n, base, slots, p0, pn, an);
    }
    __syncthreads();
    PHASE_STAMP(1);
    drain(true);
    } else {
    // dense tiles: phase A stops when a round might not fit, the records
    // are drained, and phase A resumes at the block-uniform point
    // (species s, round p0); the test reads the record count between two
    // barriers
    const bool multipass = min(a.sp[0].cnt_a[t], a.sp[0].tcap) + min(a.sp[1].cnt_a[t], a.sp[1].tcap) > CELLS_R;
    int s = 0, p0 = 0;
#pragma unroll 1
    for (;;) {
#pragma unroll 1
    for (; s < 2; ++s, p0 = 0) {
        const SpecArgs& sp = a.sp[s];
        const int n = min(sp.cnt_a[t], sp.tcap);
        const int64_t base = (int64_t)t * sp.tcap;
        const int64_t slots = (int64_t)tl.ntiles * sp.tcap;
        Part pn;
        bool an = p0 + tid < n;
        if (an) load_part<3, HAS_CH, HAS_ID>(sp.a, base + p0 + tid, a.pk, pn);
        bool full = false;
        for (; p0 < n; p0 += NT) {
            if (multipass) {
                __syncthreads();
                const unsigned used = *nrec;
                __syncthreads();
                full = (int)(used & 0xffffu) + (int)(used >> 16) + NT > CELLS_R;
                if (full) break;
            }
            round(sp, n, base, slots, p0, pn, an);
        }
        if (full) break;
    }
    __syncthreads();
    PHASE_STAMP(1);
    drain(s >= 2);
    if (s >= 2) break;
    }
    }
    // (the flush's addresses are computed from here on, not carried from
    // the earlier phases)
    asm volatile("" : "+v"(tid));
    // a thread's (at most two) nodes: all their values are read from LDS
    // and cleared for the next tile first, then the atomics are issued, so
    // the LDS reads overlap instead of each waiting before its atomic
    constexpr int FI = (jplane + NT - 1) / NT;
    float fv[FI][3];
    int64_t fg[FI];
#pragma unroll
    for (int it = 0; it < FI; ++it) {
        const int e = tid + it * NT;
        fg[it] = -1;
#pragma unroll
        for (int q = 0; q < 3; ++q) fv[it][q] = 0.0f;
        if (e < jplane) {
            const int li = e % JX;
            const int r = e / JX;
            const int lj = r % JY;
            const int lk = r / JY;
            const int gi = wrap1(i0 - 2 + li, a.nx);
            const int gj = wrap1(j0 - 2 + lj, a.ny);
            const int gk = wrap1(k0 - 2 + lk, a.nz);
            fg[it] = gi + (int64_t)a.nx * (gj + (int64_t)a.ny * gk);
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                fv[it][q] = (float)sj[q * jplane + e];
                sj[q * jplane + e] = 0.0;
            }
        }
    }
#pragma unroll
    for (int it = 0; it < FI; ++it)
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (fv[it][q] != 0.0f) atomicAdd(a.j3[q] + fg[it], fv[it][q]);
    // closes the tile: the fields, the cleared currents and counters are
    // in place for the next tile's phase A
    __syncthreads();
    PHASE_STAMP(6);
    }
}

// ---------------------------------------------------------------- flat
struct FlatArgs {
    int nx, ny, nz;
    PackInfo pk;
    Tiling tl;
    float c, cinv;
    int flags;              // PUSH_GATHER2 | PUSH_VAY
    GlobalSrc src;
    GlobalSink snk;
    SpecArgs sp;
    // input: a tile-segmented buffer (cnt_in != null) or a list (count at *lcount)
    PBuf in;
    const int32_t* cnt_in;
    int32_t tcap_in;
    const int32_t* lcount;
    int64_t lcap;
    int64_t scratch_off;     // index offset of the input in the scratch arrays
    int64_t nthreads;
    // scratch of the split move/deposit
    int32_t* cell_old;
    float *ox_old, *oy_old, *oz_old, *dxs, *dys, *dzs;
};

// MODE 0: push + deposit + re-bin, 1: push + re-bin keeping the old position
// and displacement, 2: deposit of the last MODE-1 push
template <int DIM, int ORD, int MODE, bool HAS_CH, bool HAS_ID>
__global__ void __launch_bounds__(256) k_push_flat(FlatArgs a)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= a.nthreads) return;
    if (a.cnt_in) {
        const int64_t t = s / a.tcap_in;
        const int l = (int)(s - t * a.tcap_in);
        if (l >= min(a.cnt_in[t], a.tcap_in)) return;
    } else {
        const int64_t cnt = min((int64_t)*a.lcount, a.lcap);
        if (s >= cnt) return;
    }
    Part p;
    load_part<DIM, HAS_CH, HAS_ID>(a.in, s, a.pk, p);
    const int64_t sc = a.scratch_off + s;
    if (MODE == 2) {
        int i, j, k;
        unpack_cell(a.pk, a.cell_old[sc], i, j, k);
        deposit_any<DIM, ORD>(a.snk, i, j, k, a.ox_old[sc], a.oy_old[sc], a.oz_old ? a.oz_old[sc] : 0.0f, a.dxs[sc],
                          a.dys[sc], a.dzs[sc], a.sp.q * p.ch);
        return;
    }
    const int oi = p.i, oj = p.j, ok = p.k;
    const float ox = p.ox, oy = p.oy, oz = p.oz;
    float dx, dy, dz;
    push_one<DIM, ORD>(a.src, a.flags, a.c, a.cinv, a.sp.qm, p.i, p.j, p.k, p, dx, dy, dz);
    advance(p, DIM, a.nx, a.ny, a.nz, dx, dy, dz, ORD == 0);
    if (MODE == 0) deposit_any<DIM, ORD>(a.snk, oi, oj, ok, ox, oy, oz, dx, dy, dz, a.sp.q * p.ch);
    const int64_t td = tile_of(a.tl, p.i, p.j, p.k);
    const int32_t slot = atomicAdd(a.sp.cnt_b + td, 1);
    const int64_t slots = (int64_t)a.tl.ntiles * a.sp.tcap;
    const int64_t d = emit<DIM, HAS_CH, HAS_ID>(a.sp, a.pk, slots, td, slot, p);
    if (MODE == 1 && d >= 0) {
        a.cell_old[d] = pack_cell(a.pk, oi, oj, ok);
        a.ox_old[d] = ox;
        a.oy_old[d] = oy;
        if (a.oz_old) a.oz_old[d] = oz;
        a.dxs[d] = dx;
        a.dys[d] = dy;
        a.dzs[d] = dz;
    }
}

// ------------------------------------------------------------- binning
__global__ void k_bin_count(const int32_t* __restrict__ cell, int64_t n, PackInfo pk, Tiling tl,
                            int32_t* __restrict__ cnt)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    int i, j, k;
    unpack_cell(pk, cell[s], i, j, k);
    atomicAdd(cnt + tile_of(tl, i, j, k), 1);
}

template <int DIM, bool HAS_CH, bool HAS_ID>
__global__ void k_bin_scatter(PBuf in, int64_t n, PackInfo pk, Tiling tl, PBuf out, int32_t tcap,
                              int32_t* __restrict__ cnt)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    Part p;
    load_part<DIM, HAS_CH, HAS_ID>(in, s, pk, p);
    const int64_t t = tile_of(tl, p.i, p.j, p.k);
    const int32_t slot = atomicAdd(cnt + t, 1);
    store_part<DIM, HAS_CH, HAS_ID>(out, t * tcap + slot, pk, p);
}

// in-tile cell sort: the particles of tile t's segment of `in` are written to
// the same segment of `out` in cell order (counting sort in LDS).  A push
// whose consecutive lanes hold particles of the same or neighbouring cells
// reads the staged fields with far fewer LDS bank conflicts; the push keeps
// most of that order from step to step, the sort restores it.  Every thread
// loads its R particles at once; each array then goes through an LDS
// staging buffer (two, used in turn: one barrier per array), so the HBM
// reads and writes are both coalesced.  The rank of a particle within its
// cell is the return of its counting atomic, so particles of one cell keep
// no particular order among themselves.
constexpr int SORT_THREADS = 512;

// R: particles per thread, for segments of up to SORT_THREADS * R particles
template <int DIM, bool HAS_CH, bool HAS_ID, int SORT_R>
__global__ void __launch_bounds__(SORT_THREADS) k_tile_sort(PBuf in, const int32_t* __restrict__ cnt,
                                                           int32_t tcap, PackInfo pk, Tiling tl, PBuf out)
{
    extern __shared__ int sh[];
    const int nc = tl.tx * tl.ty * tl.tz;
    int* ccount = sh;          // [nc] particles per cell
    int* cstart = sh + nc;     // [nc] exclusive prefix sum
    int* stage0 = sh + 2 * nc; // [tcap] staging of one array
    int* stage1 = stage0 + tcap;
    const int64_t t = blockIdx.x;
    const int ti = (int)(t % tl.ntx), tj = (int)((t / tl.ntx) % tl.nty), tk = (int)(t / ((int64_t)tl.ntx * tl.nty));
    const int i0 = ti * tl.tx, j0 = tj * tl.ty, k0 = tk * tl.tz;
    const int n = min(cnt[t], tcap);
    const int64_t base = t * tcap;
    const int tid = threadIdx.x;
    constexpr int NA = 6 + (DIM == 3) + HAS_CH + HAS_ID;
    int32_t v[NA][SORT_R];
    int dst[SORT_R];
#pragma unroll
    for (int r = 0; r < SORT_R; ++r) {
        const int e = tid + r * SORT_THREADS;
        if (e < n) {
            const int64_t s = base + e;
            int a = 0;
            v[a++][r] = in.cell[s];
            v[a++][r] = __float_as_int(in.ox[s]);
            v[a++][r] = __float_as_int(in.oy[s]);
            if (DIM == 3) v[a++][r] = __float_as_int(in.oz[s]);
            v[a++][r] = __float_as_int(in.u[s]);
            v[a++][r] = __float_as_int(in.v[s]);
            v[a++][r] = __float_as_int(in.w[s]);
            if (HAS_CH) v[a++][r] = __float_as_int(in.ch[s]);
            if (HAS_ID) v[a++][r] = in.id[s];
        }
    }
    for (int c = tid; c < nc; c += SORT_THREADS) ccount[c] = 0;
    __syncthreads();
    int lc[SORT_R];
#pragma unroll
    for (int r = 0; r < SORT_R; ++r) {
        lc[r] = 0;
        dst[r] = 0;
        if (tid + r * SORT_THREADS < n) {
            int i, j, k;
            unpack_cell(pk, v[0][r], i, j, k);
            lc[r] = (i - i0) + tl.tx * ((j - j0) + tl.ty * (k - k0));
            dst[r] = atomicAdd(ccount + lc[r], 1);
        }
    }
    __syncthreads();
    if (tid < 64) {
        int carry = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
            const int c = c0 + tid;
            const int x0 = c < nc ? ccount[c] : 0;
            int x = x0;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const int y = __shfl_up(x, o, 64);
                if (tid >= o) x += y;
            }
            if (c < nc) cstart[c] = carry + x - x0;
            carry += __shfl(x, 63, 64);
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SORT_R; ++r)
        if (tid + r * SORT_THREADS < n) dst[r] += cstart[lc[r]];
    int32_t* outs[NA];
    {
        int a = 0;
        outs[a++] = out.cell + base;
        outs[a++] = reinterpret_cast<int32_t*>(out.ox + base);
        outs[a++] = reinterpret_cast<int32_t*>(out.oy + base);
        if (DIM == 3) outs[a++] = reinterpret_cast<int32_t*>(out.oz + base);
        outs[a++] = reinterpret_cast<int32_t*>(out.u + base);
        outs[a++] = reinterpret_cast<int32_t*>(out.v + base);
        outs[a++] = reinterpret_cast<int32_t*>(out.w + base);
        if (HAS_CH) outs[a++] = reinterpret_cast<int32_t*>(out.ch + base);
        if (HAS_ID) outs[a++] = out.id + base;
    }
    // array a is scattered into stage[a % 2] and written out after the next
    // barrier, while array a + 1 is scattered into the other buffer
#pragma unroll
    for (int a = 0; a <= NA; ++a) {
        if (a < NA) {
            int* st = (a & 1) ? stage1 : stage0;
#pragma unroll
            for (int r = 0; r < SORT_R; ++r)
                if (tid + r * SORT_THREADS < n) st[dst[r]] = v[a][r];
        }
        if (a > 0) {
            const int* st = ((a - 1) & 1) ? stage1 : stage0;
            for (int e = tid; e < n; e += SORT_THREADS) outs[a - 1][e] = st[e];
        }
        __syncthreads();
    }
}

// before a push: the output tile counts of both species and their new
// overflow-list counts start at zero (one launch instead of four memsets)
__global__ void k_reset_counts(int32_t* __restrict__ cb0, int32_t* __restrict__ cb1, int64_t ntiles,
                               int32_t* __restrict__ oc0, int32_t* __restrict__ oc1)
{
    const int64_t t0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t t = t0; t < ntiles; t += (int64_t)gridDim.x * blockDim.x) {
        if (cb0) cb0[t] = 0;
        if (cb1) cb1[t] = 0;
    }
    if (t0 == 0) {
        if (oc0) oc0[1] = 0;
        if (oc1) oc1[1] = 0;
    }
}

// after a push: the overflow list just filled (count in slot 1) becomes the
// one the next push serves (slot 0)
__global__ void k_promote_lists(int32_t* __restrict__ oc0, int32_t* __restrict__ oc1)
{
    if (threadIdx.x == 0) {
        if (oc0) oc0[0] = oc0[1];
        if (oc1) oc1[0] = oc1[1];
    }
}

// compaction: tile segments of `in` -> consecutive positions of `out`
// starting at off[t]
template <int DIM, bool HAS_CH, bool HAS_ID>
__global__ void k_flatten(PBuf in, const int32_t* __restrict__ cnt, int32_t tcap, int64_t nslots,
                          const int64_t* __restrict__ off, PackInfo pk, PBuf out)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nslots) return;
    const int64_t t = s / tcap;
    const int l = (int)(s - t * tcap);
    if (l >= min(cnt[t], tcap)) return;
    Part p;
    load_part<DIM, HAS_CH, HAS_ID>(in, s, pk, p);
    store_part<DIM, HAS_CH, HAS_ID>(out, off[t] + l, pk, p);
}

template <int DIM, bool HAS_CH, bool HAS_ID>
__global__ void k_copy_list(PBuf in, int64_t n, int64_t dst0, PackInfo pk, PBuf out)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    Part p;
    load_part<DIM, HAS_CH, HAS_ID>(in, s, pk, p);
    store_part<DIM, HAS_CH, HAS_ID>(out, dst0 + s, pk, p);
}

// ------------------------------------------------------- dispatch helpers
template <class F>
void dispatch(int dim, int order, bool ch, bool id, F&& f)
{
    // instantiate every (dim, order, ch, id) once; f is a generic lambda
    auto go = [&](auto D, auto O) {
        if (ch) { if (id) f(D, O, std::true_type{}, std::true_type{}); else f(D, O, std::true_type{}, std::false_type{}); }
        else { if (id) f(D, O, std::false_type{}, std::true_type{}); else f(D, O, std::false_type{}, std::false_type{}); }
    };
    using I2 = std::integral_constant<int, 2>;
    using I3 = std::integral_constant<int, 3>;
    using O0 = std::integral_constant<int, 0>;   // the zigzag scheme
    using O1 = std::integral_constant<int, 1>;
    using O2 = std::integral_constant<int, 2>;
    using O3 = std::integral_constant<int, 3>;
    if (dim == 2) {
        if (order == 0) go(I2{}, O0{});
        else if (order == 1) go(I2{}, O1{}); else if (order == 2) go(I2{}, O2{}); else go(I2{}, O3{});
    } else {
        if (order == 0) go(I3{}, O0{});
        else if (order == 1) go(I3{}, O1{}); else if (order == 2) go(I3{}, O2{}); else go(I3{}, O3{});
    }
}

// template order of the push kernels: the shape order, or 0 for the zigzag scheme
int push_order(const tmpg_ctx* ctx) { return ctx->zz ? 0 : ctx->cfg.order; }

int push_flags(const tmpg_ctx* ctx)
{
    return ((ctx->cfg.compat & TMPG_Q_GATHER2) ? PUSH_GATHER2 : 0) | (ctx->cfg.pusher == 1 ? PUSH_VAY : 0);
}

SpecArgs spec_args(tmpg_ctx* ctx, int s)
{
    Species& sp = ctx->sp[s];
    SpecArgs r;
    r.qm = ctx->cfg.qm[s];
    r.q = ctx->cfg.q[s];
    r.tcap = sp.tcap;
    r.ocap = sp.ocap;
    r.a = sp.a;
    r.b = sp.b;
    r.ob = sp.ob;
    r.cnt_a = sp.cnt_a;
    r.cnt_b = sp.cnt_b;
    r.ocnt = sp.ocnt;
    return r;
}

FlatArgs flat_args(tmpg_ctx* ctx, int s)
{
    const Grid& g = ctx->g;
    FlatArgs a;
    a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
    a.pk = ctx->pk;
    a.tl = ctx->tl;
    a.c = ctx->cfg.c;
    a.cinv = 1.0f / ctx->cfg.c;
    a.flags = push_flags(ctx);
    for (int q = 0; q < 6; ++q) a.src.f[q] = (g.dim == 3 && !ctx->zz) ? ctx->prim[q] : ctx->f[EX + q];
    a.src.nx = g.nx; a.src.ny = g.ny; a.src.nz = g.nz;
    for (int q = 0; q < 3; ++q) a.snk.j[q] = ctx->f[JX + q];
    a.snk.nx = g.nx; a.snk.ny = g.ny; a.snk.nz = g.nz;
    a.sp = spec_args(ctx, s);
    Species& sp = ctx->sp[s];
    a.cell_old = sp.cell_old; a.ox_old = sp.ox_old; a.oy_old = sp.oy_old; a.oz_old = sp.oz_old;
    a.dxs = sp.dx; a.dys = sp.dy; a.dzs = sp.dz;
    a.cnt_in = nullptr; a.lcount = nullptr; a.tcap_in = 0; a.lcap = 0; a.scratch_off = 0; a.nthreads = 0;
    return a;
}

template <int DIM, int ORD, bool C, bool I>
void launch_flat_mode(const FlatArgs& a, int mode, hipStream_t st)
{
    if (a.nthreads <= 0) return;
    dim3 grid((unsigned)((a.nthreads + 255) / 256));
    if (mode == 0) hipLaunchKernelGGL((k_push_flat<DIM, ORD, 0, C, I>), grid, dim3(256), 0, st, a);
    else if (mode == 1) hipLaunchKernelGGL((k_push_flat<DIM, ORD, 1, C, I>), grid, dim3(256), 0, st, a);
    else hipLaunchKernelGGL((k_push_flat<DIM, ORD, 2, C, I>), grid, dim3(256), 0, st, a);
}

void swap_buffers(Species& sp)
{
    std::swap(sp.a, sp.b);
    std::swap(sp.cnt_a, sp.cnt_b);
    std::swap(sp.oa, sp.ob);
}

}  // namespace

// ---------------------------------------------------------------- host
int alloc_pbuf(tmpg_ctx* ctx, PBuf& b, int64_t n, bool has_ch, bool has_id)
{
    b = PBuf();
    if (n <= 0) return TMPG_OK;
    const size_t fb = (size_t)n * sizeof(float), ib = (size_t)n * sizeof(int32_t);
    auto al = [&](void** p, size_t bytes) -> int {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) {
            ctx->err = "hipMalloc(" + std::to_string(bytes) + " B): " + hipGetErrorString(e);
            *p = nullptr;
            return TMPG_ENOMEM;
        }
        return TMPG_OK;
    };
    int rc = 0;
    rc |= al((void**)&b.cell, ib);
    rc |= al((void**)&b.ox, fb);
    rc |= al((void**)&b.oy, fb);
    if (ctx->g.dim == 3 || ctx->zz) rc |= al((void**)&b.oz, fb);   // 2D zigzag: z modulo one cell
    rc |= al((void**)&b.u, fb);
    rc |= al((void**)&b.v, fb);
    rc |= al((void**)&b.w, fb);
    if (has_ch) rc |= al((void**)&b.ch, fb);
    if (has_id) rc |= al((void**)&b.id, ib);
    if (rc) { free_pbuf(b); return TMPG_ENOMEM; }
    return TMPG_OK;
}

void free_pbuf(PBuf& b)
{
    void* ps[] = {b.cell, b.ox, b.oy, b.oz, b.u, b.v, b.w, b.ch, b.id};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    b = PBuf();
}

void free_species(Species& s)
{
    free_pbuf(s.a);
    free_pbuf(s.b);
    free_pbuf(s.oa);
    free_pbuf(s.ob);
    void* ps[] = {s.cnt_a, s.cnt_b, s.ocnt, s.cell_old, s.ox_old, s.oy_old, s.oz_old, s.dx, s.dy, s.dz};
    for (void* p : ps)
        if (p) (void)hipFree(p);
    s = Species();
}

int particles_bin(tmpg_ctx* ctx, int s, const PBuf& flat, int64_t n, bool has_ch, bool has_id)
{
    Species& sp = ctx->sp[s];
    const Tiling& tl = ctx->tl;
    free_species(sp);
    sp.has_ch = has_ch;
    sp.has_id = has_id;
    TMPG_HIP_CHECK(ctx, hipMalloc(&sp.cnt_a, tl.ntiles * sizeof(int32_t)));
    TMPG_HIP_CHECK(ctx, hipMalloc(&sp.cnt_b, tl.ntiles * sizeof(int32_t)));
    TMPG_HIP_CHECK(ctx, hipMalloc(&sp.ocnt, 4 * sizeof(int32_t)));
    TMPG_HIP_CHECK(ctx, hipMemsetAsync(sp.cnt_a, 0, tl.ntiles * sizeof(int32_t), ctx->stream));
    TMPG_HIP_CHECK(ctx, hipMemsetAsync(sp.ocnt, 0, 4 * sizeof(int32_t), ctx->stream));
    if (n > 0)
        hipLaunchKernelGGL(k_bin_count, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, flat.cell, n,
                           ctx->pk, tl, sp.cnt_a);
    std::vector<int32_t> cnt((size_t)tl.ntiles);
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(cnt.data(), sp.cnt_a, cnt.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                       ctx->stream));
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t mx = cnt.empty() ? 0 : *std::max_element(cnt.begin(), cnt.end());
    const int64_t cells_per_tile = (int64_t)tl.tx * tl.ty * tl.tz;
    const int64_t mean = (n + tl.ntiles - 1) / std::max<int64_t>(tl.ntiles, 1);
    // slack for density changes: 25% above the fullest tile, at least 32
    // slots or one particle per cell more
    int64_t tcap = std::max<int64_t>({mx + mx / 4 + 32, mean + mean / 2, mx + cells_per_tile});
    if (n == 0) tcap = 64;
    tcap = (tcap + 63) / 64 * 64;
    if (tcap > (int64_t)1 << 30) { ctx->err = "tile capacity too large"; return TMPG_EINVAL; }
    sp.tcap = (int32_t)tcap;
    sp.slots = tl.ntiles * tcap;
    sp.ocap = std::max<int64_t>(65536, n / 16);
    int rc = 0;
    rc |= alloc_pbuf(ctx, sp.a, sp.slots, has_ch, has_id);
    rc |= alloc_pbuf(ctx, sp.b, sp.slots, has_ch, has_id);
    rc |= alloc_pbuf(ctx, sp.oa, sp.ocap, has_ch, has_id);
    rc |= alloc_pbuf(ctx, sp.ob, sp.ocap, has_ch, has_id);
    if (rc) return TMPG_ENOMEM;
    TMPG_HIP_CHECK(ctx, hipMemsetAsync(sp.cnt_a, 0, tl.ntiles * sizeof(int32_t), ctx->stream));
    if (n > 0) {
        dim3 grid((unsigned)((n + 255) / 256));
        const int dim = ctx->g.dim;
        dispatch(dim, 1, has_ch, has_id, [&](auto D, auto, auto C, auto I) {
            hipLaunchKernelGGL((k_bin_scatter<decltype(D)::value, decltype(C)::value, decltype(I)::value>), grid,
                               dim3(256), 0, ctx->stream, flat, n, ctx->pk, tl, sp.a, sp.tcap, sp.cnt_a);
        });
        TMPG_HIP_CHECK(ctx, hipGetLastError());
    }
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    sp.n = n;
    return TMPG_OK;
}

int particles_flatten(tmpg_ctx* ctx, int s, PBuf& flat, int64_t* n_out)
{
    Species& sp = ctx->sp[s];
    const Tiling& tl = ctx->tl;
    *n_out = 0;
    if (!sp.cnt_a) return TMPG_OK;
    std::vector<int32_t> cnt((size_t)tl.ntiles);
    int32_t oc[4];
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(cnt.data(), sp.cnt_a, cnt.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                       ctx->stream));
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(oc, sp.ocnt, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (oc[2] > 0) {
        ctx->err = "particles lost: " + std::to_string(oc[2]) + " overflowed the overflow list";
        return TMPG_EOVERFLOW;
    }
    std::vector<int64_t> off((size_t)tl.ntiles);
    int64_t tot = 0;
    for (size_t t = 0; t < cnt.size(); ++t) {
        off[t] = tot;
        tot += std::min(cnt[t], sp.tcap);
    }
    const int64_t nov = std::min<int64_t>(oc[0], sp.ocap);
    const int64_t n = tot + nov;
    int rc = alloc_pbuf(ctx, flat, std::max<int64_t>(n, 1), sp.has_ch, sp.has_id);
    if (rc) return rc;
    int64_t* doff = nullptr;
    TMPG_HIP_CHECK(ctx, hipMalloc(&doff, off.size() * sizeof(int64_t)));
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(doff, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice,
                                       ctx->stream));
    dispatch(ctx->g.dim, 1, sp.has_ch, sp.has_id, [&](auto D, auto, auto C, auto I) {
        constexpr int Dv = decltype(D)::value;
        constexpr bool Cv = decltype(C)::value, Iv = decltype(I)::value;
        if (sp.slots > 0)
            hipLaunchKernelGGL((k_flatten<Dv, Cv, Iv>), dim3((unsigned)((sp.slots + 255) / 256)), dim3(256), 0,
                               ctx->stream, sp.a, sp.cnt_a, sp.tcap, sp.slots, doff, ctx->pk, flat);
        if (nov > 0)
            hipLaunchKernelGGL((k_copy_list<Dv, Cv, Iv>), dim3((unsigned)((nov + 255) / 256)), dim3(256), 0,
                               ctx->stream, sp.oa, nov, tot, ctx->pk, flat);
    });
    hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(doff);
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return TMPG_EHIP; }
    *n_out = n;
    return TMPG_OK;
}

int particles_rebalance(tmpg_ctx* ctx, int s)
{
    Species& sp = ctx->sp[s];
    if (!sp.cnt_a) return TMPG_OK;
    PBuf flat;
    int64_t n = 0;
    int rc = particles_flatten(ctx, s, flat, &n);
    if (rc) { free_pbuf(flat); return rc; }
    rc = particles_bin(ctx, s, flat, n, sp.has_ch, sp.has_id);
    free_pbuf(flat);
    return rc;
}

int particles_sort_tiles(tmpg_ctx* ctx)
{
    const Grid& g = ctx->g;
    const Tiling& tl = ctx->tl;
    // the sort carries the 3D offsets; the 2D zigzag scheme's z offset is
    // not among its arrays, so its particles keep their order
    if (g.dim == 2 && ctx->zz) return TMPG_OK;
    for (int s = 0; s < 2; ++s) {
        Species& sp = ctx->sp[s];
        if (!sp.cnt_a || sp.n == 0) continue;
        const size_t lds = sizeof(int) * (2 * (size_t)tl.tx * tl.ty * tl.tz + 2 * (size_t)sp.tcap);
        // the sort only speeds up the push: segments longer than one
        // workgroup holds stay as they are
        const int r = sp.tcap <= 4 * SORT_THREADS ? 4 : sp.tcap <= 8 * SORT_THREADS ? 8 : 16;
        if (sp.tcap > 16 * SORT_THREADS || lds > 80 * 1024) continue;
        dispatch(g.dim, 1, sp.has_ch, sp.has_id, [&](auto D, auto O, auto C, auto I) {
            (void)O;
            constexpr int DV = decltype(D)::value;
            constexpr bool CV = decltype(C)::value, IV = decltype(I)::value;
            const dim3 grid((unsigned)tl.ntiles), block(SORT_THREADS);
            if (r == 4)
                hipLaunchKernelGGL((k_tile_sort<DV, CV, IV, 4>), grid, block, lds, ctx->stream, sp.a, sp.cnt_a,
                                   sp.tcap, ctx->pk, tl, sp.b);
            else if (r == 8)
                hipLaunchKernelGGL((k_tile_sort<DV, CV, IV, 8>), grid, block, lds, ctx->stream, sp.a, sp.cnt_a,
                                   sp.tcap, ctx->pk, tl, sp.b);
            else
                hipLaunchKernelGGL((k_tile_sort<DV, CV, IV, 16>), grid, block, lds, ctx->stream, sp.a, sp.cnt_a,
                                   sp.tcap, ctx->pk, tl, sp.b);
        });
        TMPG_HIP_CHECK(ctx, hipGetLastError());
        // the sorted copy becomes the live buffer; counts are unchanged
        std::swap(sp.a, sp.b);
    }
    return TMPG_OK;
}

int particles_check(tmpg_ctx* ctx)
{
    // called at sync points: lost particles are an error, a filling overflow
    // list triggers a re-bin with larger tiles segments
    for (int s = 0; s < 2; ++s) {
        Species& sp = ctx->sp[s];
        if (!sp.ocnt) continue;
        int32_t oc[4];
        TMPG_HIP_CHECK(ctx, hipMemcpyAsync(oc, sp.ocnt, 4 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        if (oc[2] > 0) {
            ctx->err = "particles lost: " + std::to_string(oc[2]) + " overflowed the overflow list";
            return TMPG_EOVERFLOW;
        }
        if ((int64_t)oc[0] > sp.ocap / 8) {
            int rc = particles_rebalance(ctx, s);
            if (rc) return rc;
        }
    }
    return TMPG_OK;
}

// the fused push of this context runs the cell-sum kernel (3D, Esirkepov,
// default 8x8x4 tile)
static bool push_cell_sums(const tmpg_ctx* ctx)
{
    const Tiling& tl = ctx->tl;
    if (ctx->g.dim != 3 || ctx->zz || !ctx->use_cell_sums || tl.tx != 8 || tl.ty != 8 || tl.tz != 4) return false;
    return cells_lds(tl.tx + 3, tl.ty + 3, tl.tz + 3, tl.tx + 5, tl.ty + 5, tl.tz + 5, tl.tx * tl.ty * tl.tz).bytes <=
           160 * 1024;
}

int particles_push(tmpg_ctx* ctx, int mode)
{
    const Grid& g = ctx->g;
    const Tiling& tl = ctx->tl;
    if (g.dim == 3 && mode != 2 && !ctx->zz) {
        int rc = fields_primal(ctx);
        if (rc) return rc;
    }
    bool any = false;
    for (int s = 0; s < 2; ++s) any |= ctx->sp[s].n > 0;
    if (!any) return TMPG_OK;
    if (mode != 2) {
        Species& s0 = ctx->sp[0];
        Species& s1 = ctx->sp[1];
        const int blocks = (int)std::min<int64_t>((tl.ntiles + 255) / 256, 4096);
        hipLaunchKernelGGL(k_reset_counts, dim3(blocks), dim3(256), 0, ctx->stream, s0.cnt_b, s1.cnt_b, tl.ntiles,
                           s0.cnt_b ? s0.ocnt : nullptr, s1.cnt_b ? s1.ocnt : nullptr);
        TMPG_HIP_CHECK(ctx, hipGetLastError());
    }
    if (mode == 0) {
        // the two species share the tile's staged fields: one launch
        TiledArgs a;
        a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
        a.pk = ctx->pk;
        a.tl = tl;
        // staged fields: 3D primal nodes -1 .. T+1; 2D and the zigzag
        // scheme's staggered fields -2 .. T+1
        const bool prim = g.dim == 3 && !ctx->zz;
        a.fo = prim ? 1 : 2;
        a.FX = tl.tx + (prim ? 3 : 4);
        a.FY = tl.ty + (prim ? 3 : 4);
        a.FZ = (g.dim == 3) ? tl.tz + (prim ? 3 : 4) : 1;
        a.JX = tl.tx + 5;
        a.JY = tl.ty + 5;
        a.JZ = (g.dim == 3) ? tl.tz + 5 : 1;
        a.c = ctx->cfg.c;
        a.cinv = 1.0f / ctx->cfg.c;
        a.flags = push_flags(ctx);
        a.remap = (tl.ntiles % 8 == 0) ? 1 : 0;
        for (int q = 0; q < 6; ++q) a.f6[q] = prim ? ctx->prim[q] : ctx->f[EX + q];
        for (int q = 0; q < 3; ++q) a.j3[q] = ctx->f[JX + q];
        for (int s = 0; s < 2; ++s) a.sp[s] = spec_args(ctx, s);
        // a species without particles still needs valid count arrays
        static_assert(sizeof(TiledArgs) < 4096, "kernel argument block");
        const size_t lds = sizeof(float) * ((6 * (size_t)a.FX * a.FY * a.FZ + 1) & ~(size_t)1) +
                           sizeof(double) * 3 * (size_t)a.JX * a.JY * a.JZ;
        bool ch = ctx->sp[0].has_ch || ctx->sp[1].has_ch;
        bool id = ctx->sp[0].has_id || ctx->sp[1].has_id;
        if (ctx->sp[0].has_ch != ctx->sp[1].has_ch || ctx->sp[0].has_id != ctx->sp[1].has_id) {
            // both species must carry the same optional arrays in one launch
            ctx->err = "species differ in weights/ids";
            return TMPG_ESTATE;
        }
        const int ncell_tile = tl.tx * tl.ty * tl.tz;
        const CellsLds CL = cells_lds(a.FX, a.FY, a.FZ, a.JX, a.JY, a.JZ, ncell_tile);
        // the cell-sum kernel is compiled for the default 8x8x4 tile; other
        // tilings (explicit, or clamped on grids under 8x8x4 cells) take the
        // general tiled kernel
        const bool cells = push_cell_sums(ctx);
        (void)ncell_tile;
        const bool vay = ctx->cfg.pusher == 1;
        // order 3 only: there a direct deposit (about 300 LDS atomics per
        // particle) costs more than draining the records in passes; at
        // order 2 the direct deposits of the excess particles are cheaper
        const bool dense = ctx->cfg.order == 3 &&
                           ctx->sp[0].n + ctx->sp[1].n > (int64_t)CELLS_R * 9 / 10 * tl.ntiles;
        if (ctx->push_events) {
            TMPG_HIP_CHECK(ctx, hipEventRecord(ctx->push_events[0], ctx->stream));
            ctx->push_recorded = true;
        }
        dispatch(g.dim, push_order(ctx), ch, id, [&](auto D, auto O, auto C, auto I) {
            constexpr int Dv = decltype(D)::value, Ov = decltype(O)::value;
            constexpr bool Cv = decltype(C)::value, Iv = decltype(I)::value;
            if constexpr (Ov == 0) {
                auto k = k_push_tiled<Dv, Ov, Cv, Iv>;
                (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                hipLaunchKernelGGL(k, dim3((unsigned)tl.ntiles), dim3(512), lds, ctx->stream, a);
            } else if (Dv == 3 && cells) {
                // tiles that hold more particles than deposit records on
                // average get the kernel that drains its records in passes
                // Vay runs get their own instantiations, so that the Boris
                // ones keep their register budget
                auto k = vay ? k_push_cells<Ov, Cv, Iv, false, true, 8, 8, 4>
                             : k_push_cells<Ov, Cv, Iv, false, false, 8, 8, 4>;
                if constexpr (Ov == 3)
                    if (dense)
                        k = vay ? k_push_cells<Ov, Cv, Iv, true, true, 8, 8, 4>
                                : k_push_cells<Ov, Cv, Iv, true, false, 8, 8, 4>;
                (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)CL.bytes);
                // persistent workgroups: as many as are resident at once
                int per_cu = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, CELLS_T, CL.bytes) != hipSuccess ||
                    per_cu < 1) {
                    (void)hipGetLastError();
                    per_cu = 1;
                }
                const int64_t grid = std::min<int64_t>(tl.ntiles, (int64_t)ctx->ncu * per_cu);
                TiledArgs ac = a;
                ac.remap = (grid % 8 == 0 && tl.ntiles % 8 == 0) ? 1 : 0;
                hipLaunchKernelGGL(k, dim3((unsigned)grid), dim3(CELLS_T), CL.bytes, ctx->stream, ac, CL);
            } else {
                auto k = k_push_tiled<Dv, Ov, Cv, Iv>;
                (void)hipFuncSetAttribute((const void*)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
                hipLaunchKernelGGL(k, dim3((unsigned)tl.ntiles), dim3(512), lds, ctx->stream, a);
            }
        });
        TMPG_HIP_CHECK(ctx, hipGetLastError());
        if (ctx->push_events) TMPG_HIP_CHECK(ctx, hipEventRecord(ctx->push_events[1], ctx->stream));
        // overflow list of the previous step
        for (int s = 0; s < 2; ++s) {
            Species& sp = ctx->sp[s];
            if (!sp.cnt_a) continue;
            FlatArgs f = flat_args(ctx, s);
            f.in = sp.oa;
            f.lcount = sp.ocnt;
            f.lcap = sp.ocap;
            f.nthreads = sp.ocap;
            dispatch(g.dim, push_order(ctx), sp.has_ch, sp.has_id, [&](auto D, auto O, auto C, auto I) {
                launch_flat_mode<decltype(D)::value, decltype(O)::value, decltype(C)::value, decltype(I)::value>(
                    f, 0, ctx->stream);
            });
            TMPG_HIP_CHECK(ctx, hipGetLastError());
        }
    } else {
        for (int s = 0; s < 2; ++s) {
            Species& sp = ctx->sp[s];
            if (!sp.cnt_a) continue;
            for (int part = 0; part < 2; ++part) {
                FlatArgs f = flat_args(ctx, s);
                if (part == 0) {
                    f.in = sp.a; f.cnt_in = sp.cnt_a; f.tcap_in = sp.tcap; f.nthreads = sp.slots;
                    f.scratch_off = 0;
                } else {
                    f.in = sp.oa; f.lcount = sp.ocnt; f.lcap = sp.ocap; f.nthreads = sp.ocap;
                    f.scratch_off = sp.slots;
                }
                dispatch(g.dim, push_order(ctx), sp.has_ch, sp.has_id, [&](auto D, auto O, auto C, auto I) {
                    launch_flat_mode<decltype(D)::value, decltype(O)::value, decltype(C)::value,
                                     decltype(I)::value>(f, mode, ctx->stream);
                });
                TMPG_HIP_CHECK(ctx, hipGetLastError());
            }
        }
    }
    if (mode != 2) {
        int32_t* oc[2] = {nullptr, nullptr};
        for (int s = 0; s < 2; ++s) {
            Species& sp = ctx->sp[s];
            if (!sp.cnt_a) continue;
            swap_buffers(sp);
            oc[s] = sp.ocnt;
        }
        // the new overflow lists' counts move to slot 0
        if (oc[0] || oc[1]) {
            hipLaunchKernelGGL(k_promote_lists, dim3(1), dim3(64), 0, ctx->stream, oc[0], oc[1]);
            TMPG_HIP_CHECK(ctx, hipGetLastError());
        }
    }
    return TMPG_OK;
}

}  // namespace tmpg

// the per-phase cycle sums of the cell-sum push (profiling build only)
extern "C" int tmpg_phase_cycles(unsigned long long cycles[8], unsigned long long counts[8], int reset)
{
    if (!cycles || !counts) return TMPG_EINVAL;
#ifdef TMPG_PHASE_STAMPS
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpyFromSymbol(cycles, HIP_SYMBOL(tmpg::g_phase_cycles), sizeof(unsigned long long) * 8) != hipSuccess ||
        hipMemcpyFromSymbol(counts, HIP_SYMBOL(tmpg::g_phase_count), sizeof(unsigned long long) * 8) != hipSuccess)
        return TMPG_EHIP;
    if (reset) {
        const unsigned long long z[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (hipMemcpyToSymbol(HIP_SYMBOL(tmpg::g_phase_cycles), z, sizeof(z)) != hipSuccess ||
            hipMemcpyToSymbol(HIP_SYMBOL(tmpg::g_phase_count), z, sizeof(z)) != hipSuccess)
            return TMPG_EHIP;
    }
    return TMPG_OK;
#else
    (void)reset;
    return TMPG_ESTATE;     // built without phase stamps
#endif
}
