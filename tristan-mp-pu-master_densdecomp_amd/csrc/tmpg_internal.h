// tmpg_internal.h -- library context and launch helpers (not part of the ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

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

// Particles of one species in HBM: SoA, position as a packed integer cell
// (i | j << bx | k << (bx+by)) plus FP32 offsets in [0,1), momentum
// u = gamma*beta.  Stored in tile order: tile t owns slots
// [t*tcap, t*tcap + count[t]) of every array (see tiles below).
struct Species {
    int64_t n = 0;        // live particles
    int64_t cap = 0;      // slots per buffer (ntiles * tcap)
    bool has_ch = false;
    int32_t* cell = nullptr;
    float *ox = nullptr, *oy = nullptr, *oz = nullptr;
    float *u = nullptr, *v = nullptr, *w = nullptr;
    float* ch = nullptr;
    int32_t* id = nullptr;
    // second buffer: the sorted push writes here, then the two swap
    int32_t* cell2 = nullptr;
    float *ox2 = nullptr, *oy2 = nullptr, *oz2 = nullptr;
    float *u2 = nullptr, *v2 = nullptr, *w2 = nullptr;
    float* ch2 = nullptr;
    int32_t* id2 = nullptr;
    // per-tile counts of the current buffer and of the buffer being written
    int32_t* cnt = nullptr;
    int32_t* cnt2 = nullptr;
    int32_t tcap = 0;      // slots per tile
    // move/deposit split: old positions and displacements of the last move
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
    float* tmp = nullptr;  // filter scratch (one component)
    tmpg::Species sp[2];
    bool move_pending = false;
    // diagnostics (FP64)
    double* rho = nullptr;
    double* rho_prev = nullptr;
    double* red = nullptr;     // reduction scratch
    bool track_cont = false;
    bool have_rho_prev = false;
    int32_t* flags = nullptr;  // device flags: [0] overflow
    hipEvent_t ev[8] = {};
    double last_ms[5] = {0, 0, 0, 0, 0};
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
int particles_load_uniform(tmpg_ctx* ctx, int s, int ppc, float drift, float theta, uint64_t seed);
int diag_energies(tmpg_ctx* ctx, double out[4]);
int diag_rho(tmpg_ctx* ctx, double* rho_dev);
int diag_continuity(tmpg_ctx* ctx, double* maxres, double* rhonorm);
}  // namespace tmpg
