// particles.hip -- gather + Boris push + Esirkepov deposit, one thread per
// particle, reading fields from HBM/L2 and depositing with global atomics.
//
// Reference: move_particles / mover_{1,2,3}ord
// (code/particles_movedeposit.F90:63-88, 356-1271) and deposit_particles /
// densdecomp_{1,2,3}ord (code/particles_movedeposit.F90:1281-2051,
// code/particles.F90:678-1358).  The deposit uses the stored old position and
// the push's displacement (the reference recomputes the old position from the
// new momentum, x0 = x - u c/gamma; the two agree in exact arithmetic).
// This kernel is the general path: it serves every particle in the flat
// layout and the overflow list of the tiled push.
#include "pic_math.h"
#include "tmpg_internal.h"

namespace tmpg {
namespace {

struct PushArgs {
    int nx, ny, nz;
    PackInfo pk;
    float c, cinv, qm, q;
    int quirk2;
    const float* f6[6];   // gather sources: primal E,B (3D) or staggered E,B (2D)
    float* j3[3];
    // particles
    int64_t n;
    int32_t* cell;
    float *ox, *oy, *oz, *u, *v, *w;
    const float* ch;
    // split move / deposit scratch
    int32_t* cell_old;
    float *ox_old, *oy_old, *oz_old, *dx, *dy, *dz;
};

__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <int DIM, int ORD>
__device__ __forceinline__ void gather(const PushArgs& a, int i, int j, int k, float ox, float oy, float oz,
                                       float* E)
{
    float Spx[4], Sdx[4], Spy[4], Sdy[4], Spz[4], Sdz[4];
    int dix, diy, diz;
    const bool qk = a.quirk2 != 0;
    gather_weights<ORD>(ox, qk, Spx, Sdx, dix);
    gather_weights<ORD>(oy, qk, Spy, Sdy, diy);
    // node range of nonzero weights: order 1 uses nodes 0..1 only
    constexpr int N0 = (ORD == 1) ? 1 : 0, N1 = (ORD == 1) ? 3 : 4;
#pragma unroll
    for (int q = 0; q < 6; ++q) E[q] = 0.0f;
    if (DIM == 3) {
        gather_weights<ORD>(oz, qk, Spz, Sdz, diz);
        for (int c = N0; c < N1; ++c) {
            int kk = wrapn(k - 1 + c, a.nz);
            for (int b = N0; b < N1; ++b) {
                int jj = wrapn(j - 1 + b, a.ny);
                const int64_t row = (int64_t)a.nx * (jj + (int64_t)a.ny * kk);
                float acc[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int t = N0; t < N1; ++t) {
                    int ii = wrapn(i - 1 + t, a.nx);
#pragma unroll
                    for (int q = 0; q < 6; ++q) acc[q] += a.f6[q][row + ii] * Spx[t];
                }
                float wyz = Spy[b] * Spz[c];
#pragma unroll
                for (int q = 0; q < 6; ++q) E[q] += acc[q] * wyz;
            }
        }
    } else {
        // staggered arrays with matching primal/dual weights
        // (code/particles_movedeposit.F90:1155-1170)
        for (int b = N0; b < N1; ++b) {
            int jp = wrapn(j - 1 + b, a.ny), jd = wrapn(j + diy - 1 + b, a.ny);
            const int64_t rp = (int64_t)a.nx * jp, rd = (int64_t)a.nx * jd;
#pragma unroll
            for (int t = N0; t < N1; ++t) {
                int ip = wrapn(i - 1 + t, a.nx), id = wrapn(i + dix - 1 + t, a.nx);
                E[0] += a.f6[0][rp + id] * Sdx[t] * Spy[b];
                E[1] += a.f6[1][rd + ip] * Spx[t] * Sdy[b];
                E[2] += a.f6[2][rp + ip] * Spx[t] * Spy[b];
                E[3] += a.f6[3][rd + ip] * Spx[t] * Sdy[b];
                E[4] += a.f6[4][rp + id] * Sdx[t] * Spy[b];
                E[5] += a.f6[5][rd + id] * Sdx[t] * Sdy[b];
            }
        }
    }
}

// Esirkepov deposit of one particle: old cell (i,j,k), old offsets o*,
// displacement d*, charge q.  cur holds -J*dt in the reference's units
// (positive sign convention of code/particles.F90:789-811): the x current at
// node a is q * sum_{a'<=a} dSx(a') * (Ay Az + dSy dSz / 12) with
// A = S1 + dS/2, the reference's S1S1 + dS S1/2 + S1 dS/2 + dS dS/3.
template <int DIM, int ORD>
__device__ __forceinline__ void deposit(const PushArgs& a, int i, int j, int k, float ox, float oy, float oz,
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
    float* JXp = a.j3[0];
    float* JYp = a.j3[1];
    float* JZp = a.j3[2];
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
        int ii[W], jj[W], kk[W];
#pragma unroll
        for (int t = 0; t < W; ++t) {
            ii[t] = wrapn(i + bx + t, a.nx);
            jj[t] = wrapn(j + by + t, a.ny);
            kk[t] = wrapn(k + bz + t, a.nz);
        }
#pragma unroll
        for (int c = 0; c < W; ++c) {
#pragma unroll
            for (int b = 0; b < W; ++b) {
                const int64_t row = (int64_t)a.nx * (jj[b] + (int64_t)a.ny * kk[c]);
                const float wyz = Ay[b] * Az[c] + twelfth * dSy[b] * dSz[c];
                const float wxz = Az[c];  // Jy row factor pieces
#pragma unroll
                for (int t = 0; t < W; ++t) {
                    // Jx: prefix along x, last node carries the telescoped zero
                    if (t < W - 1) {
                        float vx = Cx[t] * wyz;
                        if (vx != 0.0f) atomicAdd(JXp + row + ii[t], vx);
                    }
                    // Jy: prefix along y
                    if (b < W - 1) {
                        float vy = Cy[b] * (Ax[t] * wxz + twelfth * dSx[t] * dSz[c]);
                        if (vy != 0.0f) atomicAdd(JYp + row + ii[t], vy);
                    }
                    // Jz: prefix along z
                    if (c < W - 1) {
                        float vz = Cz[c] * (Ax[t] * Ay[b] + twelfth * dSx[t] * dSy[b]);
                        if (vz != 0.0f) atomicAdd(JZp + row + ii[t], vz);
                    }
                }
            }
        }
    } else {
        // 2D: Jx, Jy by prefix sums, Jz directly from the z displacement
        // (code/particles.F90:822-848)
        const float qdz = -q * dz;
        int ii[W], jj[W];
#pragma unroll
        for (int t = 0; t < W; ++t) {
            ii[t] = wrapn(i + bx + t, a.nx);
            jj[t] = wrapn(j + by + t, a.ny);
        }
#pragma unroll
        for (int b = 0; b < W; ++b) {
            const int64_t row = (int64_t)a.nx * jj[b];
#pragma unroll
            for (int t = 0; t < W; ++t) {
                if (t < W - 1) {
                    float vx = Cx[t] * Ay[b];
                    if (vx != 0.0f) atomicAdd(JXp + row + ii[t], vx);
                }
                if (b < W - 1) {
                    float vy = Cy[b] * Ax[t];
                    if (vy != 0.0f) atomicAdd(JYp + row + ii[t], vy);
                }
                float vz = qdz * (Ax[t] * Ay[b] + twelfth * dSx[t] * dSy[b]);
                if (vz != 0.0f) atomicAdd(JZp + row + ii[t], vz);
            }
        }
    }
}

// MODE 0: fused push + deposit, 1: push only (keeps old position and
// displacement), 2: deposit of the last push
template <int DIM, int ORD, int MODE, bool HAS_CH>
__global__ void __launch_bounds__(256) k_push(PushArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    int i, j, k;
    float ox, oy, oz = 0.0f, dx, dy, dz;
    if (MODE == 2) {
        unpack_cell(a.pk, a.cell_old[p], i, j, k);
        ox = a.ox_old[p];
        oy = a.oy_old[p];
        if (DIM == 3) oz = a.oz_old[p];
        dx = a.dx[p];
        dy = a.dy[p];
        dz = a.dz[p];
    } else {
        const int32_t cl = a.cell[p];
        unpack_cell(a.pk, cl, i, j, k);
        ox = a.ox[p];
        oy = a.oy[p];
        if (DIM == 3) oz = a.oz[p];
        float u = a.u[p], v = a.v[p], w = a.w[p];
        float E[6];
        gather<DIM, ORD>(a, i, j, k, ox, oy, oz, E);
        const float se = 0.5f * a.qm, sb = 0.5f * a.qm * a.cinv;
        boris(a.c, a.cinv, se * E[0], se * E[1], se * E[2], sb * E[3], sb * E[4], sb * E[5], u, v, w, dx, dy,
              dz);
        a.u[p] = u;
        a.v[p] = v;
        a.w[p] = w;
        // new position
        float nx_ = ox, ny_ = oy, nz_ = oz;
        int ni = wrapn(i + advance_offset(nx_, dx), a.nx);
        int nj = wrapn(j + advance_offset(ny_, dy), a.ny);
        int nk = k;
        if (DIM == 3) nk = wrapn(k + advance_offset(nz_, dz), a.nz);
        a.cell[p] = pack_cell(a.pk, ni, nj, nk);
        a.ox[p] = nx_;
        a.oy[p] = ny_;
        if (DIM == 3) a.oz[p] = nz_;
        if (MODE == 1) {
            a.cell_old[p] = cl;
            a.ox_old[p] = ox;
            a.oy_old[p] = oy;
            if (DIM == 3) a.oz_old[p] = oz;
            a.dx[p] = dx;
            a.dy[p] = dy;
            a.dz[p] = dz;
            return;
        }
    }
    const float q = HAS_CH ? a.q * a.ch[p] : a.q;
    deposit<DIM, ORD>(a, i, j, k, ox, oy, oz, dx, dy, dz, q);
}

template <int DIM, int ORD, int MODE>
void launch_push(const PushArgs& a, bool has_ch, hipStream_t st)
{
    dim3 grid((unsigned)((a.n + 255) / 256));
    if (has_ch)
        hipLaunchKernelGGL((k_push<DIM, ORD, MODE, true>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_push<DIM, ORD, MODE, false>), grid, dim3(256), 0, st, a);
}

template <int DIM, int ORD>
void launch_mode(const PushArgs& a, int mode, bool has_ch, hipStream_t st)
{
    if (mode == 0) launch_push<DIM, ORD, 0>(a, has_ch, st);
    else if (mode == 1) launch_push<DIM, ORD, 1>(a, has_ch, st);
    else launch_push<DIM, ORD, 2>(a, has_ch, st);
}

}  // namespace

int particles_push(tmpg_ctx* ctx, int mode)
{
    const Grid& g = ctx->g;
    if (g.dim == 3 && mode != 2) {
        int rc = fields_primal(ctx);
        if (rc) return rc;
    }
    for (int s = 0; s < 2; ++s) {
        Species& sp = ctx->sp[s];
        if (sp.n == 0) continue;
        PushArgs a;
        a.nx = g.nx; a.ny = g.ny; a.nz = g.nz;
        a.pk = ctx->pk;
        a.c = ctx->cfg.c;
        a.cinv = 1.0f / ctx->cfg.c;
        a.qm = ctx->cfg.qm[s];
        a.q = ctx->cfg.q[s];
        a.quirk2 = (ctx->cfg.compat & TMPG_Q_GATHER2) ? 1 : 0;
        for (int q = 0; q < 6; ++q) a.f6[q] = (g.dim == 3) ? ctx->prim[q] : ctx->f[EX + q];
        for (int q = 0; q < 3; ++q) a.j3[q] = ctx->f[JX + q];
        a.n = sp.n;
        a.cell = sp.cell; a.ox = sp.ox; a.oy = sp.oy; a.oz = sp.oz;
        a.u = sp.u; a.v = sp.v; a.w = sp.w; a.ch = sp.ch;
        a.cell_old = sp.cell_old; a.ox_old = sp.ox_old; a.oy_old = sp.oy_old; a.oz_old = sp.oz_old;
        a.dx = sp.dx; a.dy = sp.dy; a.dz = sp.dz;
        const int key = g.dim * 10 + ctx->cfg.order;
        switch (key) {
        case 21: launch_mode<2, 1>(a, mode, sp.has_ch, ctx->stream); break;
        case 22: launch_mode<2, 2>(a, mode, sp.has_ch, ctx->stream); break;
        case 23: launch_mode<2, 3>(a, mode, sp.has_ch, ctx->stream); break;
        case 31: launch_mode<3, 1>(a, mode, sp.has_ch, ctx->stream); break;
        case 32: launch_mode<3, 2>(a, mode, sp.has_ch, ctx->stream); break;
        case 33: launch_mode<3, 3>(a, mode, sp.has_ch, ctx->stream); break;
        default: ctx->err = "unsupported dim/order"; return TMPG_EINVAL;
        }
        TMPG_HIP_CHECK(ctx, hipGetLastError());
    }
    return TMPG_OK;
}

}  // namespace tmpg
