// tmpg_internal.h -- library context and launch helpers (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/tmpg.h"

namespace tmpg {

// Grid in HBM: the periodic interior only, one float array per component,
// x fastest: index i + nx*(j + ny*k).  The boundary layout (Fortran arrays
// with ghosts) exists only in tmpg_set_*/tmpg_get_*.
enum { EX = 0, EY, EZ, BX, BY, BZ, JX, JY, JZ, NF };

struct Grid {
    int dim, nx, ny, nz;
    int64_t n;  // cells
};

// One SoA particle buffer: position as a packed integer cell
// (i | j << bx | k << (bx+by)) plus FP32 offsets in [0,1), momentum
// u = gamma*beta, optional weight ch and id.
struct PBuf {
    int32_t* cell = nullptr;
    float *ox = nullptr, *oy = nullptr, *oz = nullptr;
    float *u = nullptr, *v = nullptr, *w = nullptr;
    float* ch = nullptr;
    int32_t* id = nullptr;
};

// Particles of one species in HBM, binned by cell tile: tile t owns slots
// [t*tcap, t*tcap + cnt[t]) of buffer a.  The push reads a and writes b
// (each particle into the segment of the tile it lands in), then the two
// swap.  A particle that finds its tile's segment full goes to the overflow
// list, which the next push serves with the general (flat) kernel.
struct Species {
    int64_t n = 0;          // live particles
    bool has_ch = false, has_id = false;
    int32_t tcap = 0;       // slots per tile
    int64_t slots = 0;      // ntiles * tcap
    PBuf a, b;
    int32_t* cnt_a = nullptr;   // per tile
    int32_t* cnt_b = nullptr;
    int64_t ocap = 0;           // overflow list capacity
    PBuf oa, ob;
    int32_t* ocnt = nullptr;    // [0] entries of oa, [1] of ob, [2] lost particles
    // move/deposit split: old cell, offsets and displacement of the last
    // move, indexed by the particle's slot in b (overflow: slots + index)
    int32_t* cell_old = nullptr;
    float *ox_old = nullptr, *oy_old = nullptr, *oz_old = nullptr;
    float *dx = nullptr, *dy = nullptr, *dz = nullptr;
};

struct PackInfo {
    int bx, by;            // bit widths of i and j in the packed cell
    int mask_x, mask_y;
};

__host__ __device__ inline int32_t pack_cell(const PackInfo& p, int i, int j, int k)
{
    return (int32_t)((uint32_t)i | ((uint32_t)j << p.bx) | ((uint32_t)k << (p.bx + p.by)));
}
__host__ __device__ inline void unpack_cell(const PackInfo& p, int32_t c, int& i, int& j, int& k)
{
    uint32_t u = (uint32_t)c;
    i = (int)(u & (uint32_t)p.mask_x);
    j = (int)((u >> p.bx) & (uint32_t)p.mask_y);
    k = (int)(u >> (p.bx + p.by));
}

// Tiles of the sorted push: tx*ty*tz cells.  Tile index = ti + ntx*(tj + nty*tk).
struct Tiling {
    int tx, ty, tz;
    int ntx, nty, ntz;
    int64_t ntiles;
};

__host__ __device__ inline int64_t tile_of(const Tiling& t, int i, int j, int k)
{
    return (int64_t)(i / t.tx) + (int64_t)t.ntx * ((j / t.ty) + (int64_t)t.nty * (k / t.tz));
}

}  // namespace tmpg

struct tmpg_ctx {
    tmpg_config cfg;
    tmpg::Grid g;
    tmpg::PackInfo pk;
    tmpg::Tiling tl;
    int h, hz, mx, my, mz;  // reference padded layout of the boundary arrays
    hipStream_t stream = nullptr;
    float* f[tmpg::NF] = {};
    float* prim[6] = {};   // 3D primal-averaged E,B for the gather
    float* tmp[3] = {nullptr, nullptr, nullptr};  // filter scratch, one grid per current component
    tmpg::Species sp[2];
    bool move_pending = false;
    bool use_cell_sums = true;  // 3D: per-cell register sums of the deposit
    bool zz = false;            // the original scheme (cfg.scheme == TMPG_SCHEME_ZIGZAG)
    tmpg_model* denoiser = nullptr;  // replaces the filter passes when set
    // diagnostics (FP64)
    double* rho = nullptr;
    double* rho_prev = nullptr;
    double* rho_work = nullptr;  // second FP64 grid of the Gauss diagnostic
    double* red = nullptr;     // reduction scratch
    bool track_cont = false;
    bool have_rho_prev = false;
    int32_t* flags = nullptr;  // device flags: [0] overflow
    int32_t* hostcnt = nullptr;   // pinned: overflow counters copied after each tmpg_step
    hipEvent_t cnt_ev = nullptr;
    bool cnt_pending = false;
    int steps_since_check = 0;
    hipEvent_t ev[8] = {};
    // per-phase timing of tmpg_step: a ring of kProfRing steps x 8 events;
    // a step's elapsed times are folded into prof_sum when its slot is
    // reused or tmpg_last_timings reads them, so the event count stays fixed
    static constexpr int kProfRing = 64;
    std::vector<hipEvent_t> prof;  // kProfRing * 8 events, created at the first tmpg_step
    int prof_head = 0;             // slot of the next step
    int prof_unread = 0;           // recorded steps not yet folded into prof_sum
    double prof_sum[6] = {};       // folded phase times (ms), tmpg_last_timings order
    int prof_steps = 0;            // steps folded into prof_sum
    std::vector<char> prof_push;   // per slot: the push kernel's two events were recorded
    hipEvent_t* push_events = nullptr;  // set during tmpg_step: recorded around the tiled push kernel
    bool push_recorded = false;         // the push recorded push_events in this step
    int sort_every = 4;                 // tmpg_step sorts each tile's particles by cell every this many steps
    int ncu = 256;                      // compute units of the context's device
    int steps_since_sort = 0;
    std::string err;
};

#define TMPG_HIP_CHECK(ctx, call)                                                          \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);               \
            return TMPG_EHIP;                                                              \
        }                                                                                  \
    } while (0)

// kernel launchers implemented in the .hip files
namespace tmpg {
int fields_bhalf(tmpg_ctx* ctx);
int fields_efull(tmpg_ctx* ctx, bool add_j);
int fields_add_current(tmpg_ctx* ctx);
int fields_filter(tmpg_ctx* ctx, int npasses);
int fields_primal(tmpg_ctx* ctx);
int fields_zero_j(tmpg_ctx* ctx);
// mode 0: fused push+deposit, 1: move only, 2: deposit of the last move
int particles_push(tmpg_ctx* ctx, int mode);
// binning: (re)build the tile segments of species s from flat arrays of n
// particles (device pointers); allocates the buffers
int particles_bin(tmpg_ctx* ctx, int s, const PBuf& flat, int64_t n, bool has_ch, bool has_id);
// compact copy of all live particles of species s into flat (device), n out
int particles_flatten(tmpg_ctx* ctx, int s, PBuf& flat, int64_t* n);
int particles_rebalance(tmpg_ctx* ctx, int s);
int alloc_pbuf(tmpg_ctx* ctx, PBuf& b, int64_t n, bool has_ch, bool has_id);
void free_pbuf(PBuf& b);
void free_species(Species& s);
int particles_load_uniform(tmpg_ctx* ctx, int s, int ppc, float drift, float theta, uint64_t seed);
// apply fn to every live particle slot: kernels over the tiled layout use this
int particles_check(tmpg_ctx* ctx);
int particles_sort_tiles(tmpg_ctx* ctx);
int diag_energies(tmpg_ctx* ctx, double out[4]);
int diag_rho(tmpg_ctx* ctx, double* rho_dev);
int diag_continuity(tmpg_ctx* ctx, double* maxres, double* rhonorm);
// Gauss's law residual with rho filtered by npasses binomial passes; rho_a
// and rho_b are two FP64 grids of scratch, red 4 doubles of reductions;
// zero_by_kernel clears them with a kernel instead of a memset
int diag_gauss(tmpg_ctx* ctx, double* rho_a, double* rho_b, double* red, int npasses, bool zero_by_kernel,
               double out[4]);
// the same on the library's static per-GPU scratch: allocates nothing
int diag_gauss_noalloc(tmpg_ctx* ctx, int npasses, double out[4]);
// device address of that scratch on the current device
int noalloc_scratch(tmpg_ctx* ctx, double** out);
int denoise_apply(tmpg_ctx* ctx, tmpg_model* m);
}  // namespace tmpg
