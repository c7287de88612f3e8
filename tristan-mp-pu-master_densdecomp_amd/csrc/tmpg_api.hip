// tmpg_api.hip -- the C ABI of include/tmpg.h: context, boundary layout
// conversions, the step sequence and the input-file reader.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "tmpg_internal.h"

using namespace tmpg;

namespace {

std::string g_create_err;

int bits_for(int n)
{
    int b = 0;
    while ((1 << b) < n) ++b;
    return b;
}

template <typename T>
int dalloc(tmpg_ctx* ctx, T** p, int64_t n)
{
    *p = nullptr;
    if (n <= 0) return TMPG_OK;
    hipError_t e = hipMalloc((void**)p, (size_t)n * sizeof(T));
    if (e != hipSuccess) {
        ctx->err = "hipMalloc(" + std::to_string((size_t)n * sizeof(T)) + " B): " + hipGetErrorString(e);
        *p = nullptr;
        return TMPG_ENOMEM;
    }
    return TMPG_OK;
}

void dfree(void* p)
{
    if (p) (void)hipFree(p);
}

int alloc_split_scratch(tmpg_ctx* ctx, Species& s)
{
    // indexed like the particle's slot in the tile buffers, then the overflow list
    if (s.dx) return TMPG_OK;
    const int64_t cap = s.slots + s.ocap;
    int rc = 0;
    rc |= dalloc(ctx, &s.cell_old, cap);
    rc |= dalloc(ctx, &s.ox_old, cap);
    rc |= dalloc(ctx, &s.oy_old, cap);
    if (ctx->g.dim == 3 || ctx->zz) rc |= dalloc(ctx, &s.oz_old, cap);
    rc |= dalloc(ctx, &s.dx, cap);
    rc |= dalloc(ctx, &s.dy, cap);
    rc |= dalloc(ctx, &s.dz, cap);
    return rc ? TMPG_ENOMEM : TMPG_OK;
}

// boundary (Fortran padded, ghosts included) <-> interior
void pad_to_interior(const tmpg_ctx* ctx, const float* src, std::vector<float>& dst)
{
    const Grid& g = ctx->g;
    dst.resize((size_t)g.n);
    const int k0 = (g.dim == 3) ? ctx->hz : 0;
    for (int k = 0; k < g.nz; ++k)
        for (int j = 0; j < g.ny; ++j) {
            const float* s = src + (size_t)(ctx->h) + (size_t)ctx->mx * ((size_t)(j + ctx->h) + (size_t)ctx->my * (k + k0));
            std::memcpy(&dst[(size_t)g.nx * (j + (size_t)g.ny * k)], s, (size_t)g.nx * sizeof(float));
        }
}

// interior -> padded array; every ghost cell holds its periodic image
void interior_to_pad(const tmpg_ctx* ctx, const std::vector<float>& src, float* dst)
{
    const Grid& g = ctx->g;
    const int k0 = (g.dim == 3) ? ctx->hz : 0;
    auto wrap = [](int a, int n) { a %= n; return a < 0 ? a + n : a; };
    for (int kk = 0; kk < ctx->mz; ++kk) {
        int k = (g.dim == 3) ? wrap(kk - k0, g.nz) : 0;
        for (int jj = 0; jj < ctx->my; ++jj) {
            int j = wrap(jj - ctx->h, g.ny);
            const float* row = &src[(size_t)g.nx * (j + (size_t)g.ny * k)];
            float* d = dst + (size_t)ctx->mx * (jj + (size_t)ctx->my * kk);
            for (int ii = 0; ii < ctx->mx; ++ii) d[ii] = row[wrap(ii - ctx->h, g.nx)];
        }
    }
}

int set_grid(tmpg_ctx* ctx, int first, int count, const float* const* arrays)
{
    std::vector<float> buf;
    for (int q = 0; q < count; ++q) {
        if (!arrays[q]) { ctx->err = "null array"; return TMPG_EINVAL; }
        pad_to_interior(ctx, arrays[q], buf);
        TMPG_HIP_CHECK(ctx, hipMemcpyAsync(ctx->f[first + q], buf.data(), buf.size() * sizeof(float),
                                           hipMemcpyHostToDevice, ctx->stream));
        TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    }
    return TMPG_OK;
}

int get_grid(tmpg_ctx* ctx, int first, int count, float* const* arrays)
{
    std::vector<float> buf((size_t)ctx->g.n);
    for (int q = 0; q < count; ++q) {
        if (!arrays[q]) { ctx->err = "null array"; return TMPG_EINVAL; }
        TMPG_HIP_CHECK(ctx, hipMemcpyAsync(buf.data(), ctx->f[first + q], buf.size() * sizeof(float),
                                           hipMemcpyDeviceToHost, ctx->stream));
        TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        interior_to_pad(ctx, buf, arrays[q]);
    }
    return TMPG_OK;
}

bool valid_ctx(const tmpg_ctx* ctx) { return ctx != nullptr; }

}  // namespace

extern "C" {

const char* tmpg_last_error(const tmpg_ctx* ctx)
{
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int tmpg_create(const tmpg_config* cfg, tmpg_ctx** out)
{
    if (!cfg || !out) { g_create_err = "null argument"; return TMPG_EINVAL; }
    *out = nullptr;
    if (cfg->dim != 2 && cfg->dim != 3) { g_create_err = "dim must be 2 or 3"; return TMPG_EINVAL; }
    if (cfg->order < 1 || cfg->order > 3) { g_create_err = "order must be 1, 2 or 3"; return TMPG_EINVAL; }
    if (cfg->pusher < 0 || cfg->pusher > 1) { g_create_err = "pusher must be 0 (Boris) or 1 (Vay)"; return TMPG_EINVAL; }
    if (cfg->scheme != TMPG_SCHEME_ESIRKEPOV && cfg->scheme != TMPG_SCHEME_ZIGZAG) {
        g_create_err = "scheme must be 0 (Esirkepov) or 1 (zigzag)";
        return TMPG_EINVAL;
    }
    if (cfg->scheme == TMPG_SCHEME_ZIGZAG && (cfg->order != 1 || cfg->pusher != 0)) {
        // the -Dzzag build: first-order ghost layout, the Boris push of its mover
        g_create_err = "the zigzag scheme needs order = 1 and the Boris pusher";
        return TMPG_EINVAL;
    }
    const int nz = cfg->dim == 3 ? cfg->nz : 1;
    if (cfg->nx < 4 || cfg->ny < 4 || nz < 1 || (cfg->dim == 3 && nz < 4)) {
        g_create_err = "grid too small (need at least 4 cells per periodic axis)";
        return TMPG_EINVAL;
    }
    const int bx = bits_for(cfg->nx), by = bits_for(cfg->ny), bz = bits_for(nz);
    if (bx + by + bz > 31) { g_create_err = "grid too large for the 31-bit packed cell index"; return TMPG_EINVAL; }
    if (!(cfg->c > 0.0f && cfg->c < 1.0f)) { g_create_err = "c must be in (0,1)"; return TMPG_EINVAL; }
    tmpg_ctx* ctx = new (std::nothrow) tmpg_ctx();
    if (!ctx) { g_create_err = "out of host memory"; return TMPG_ENOMEM; }
    ctx->cfg = *cfg;
    ctx->cfg.nz = nz;
    ctx->zz = cfg->scheme == TMPG_SCHEME_ZIGZAG;
    ctx->sort_every = cfg->sort_every == 0 ? 4 : cfg->sort_every;
    ctx->g.dim = cfg->dim;
    ctx->g.nx = cfg->nx;
    ctx->g.ny = cfg->ny;
    ctx->g.nz = nz;
    ctx->g.n = (int64_t)cfg->nx * cfg->ny * nz;
    ctx->pk.bx = bx;
    ctx->pk.by = by;
    ctx->pk.mask_x = (1 << bx) - 1;
    ctx->pk.mask_y = (1 << by) - 1;
    const int ng = cfg->order == 1 ? 5 : 7;
    const int ngz = cfg->dim == 2 ? 5 : ng;
    ctx->h = ng / 2;
    ctx->hz = ngz / 2;
    ctx->mx = cfg->nx + ng;
    ctx->my = cfg->ny + ng;
    ctx->mz = cfg->dim == 3 ? nz + ngz : 1;
    {
        // cell tiles of the sorted push (default 8x8x4 in 3D; 16x16 in 2D,
        // or 16x8 when 16x16 tiles would not give every CU two workgroups)
        int ncu = 256;
        {
            int dev = 0;
            hipDeviceProp_t prop;
            if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
                ncu = prop.multiProcessorCount;
            else
                (void)hipGetLastError();
        }
        if (ncu > 0) ctx->ncu = ncu;
        int tx = cfg->tile[0] > 0 ? cfg->tile[0] : (cfg->dim == 3 ? 8 : 16);
        int ty = cfg->tile[1] > 0 ? cfg->tile[1] : (cfg->dim == 3 ? 8 : 16);
        if (cfg->dim == 2 && cfg->tile[1] <= 0 &&
            (int64_t)((cfg->nx + 15) / 16) * ((cfg->ny + 15) / 16) < 2 * (int64_t)ncu)
            ty = 8;
        int tz = cfg->dim == 3 ? (cfg->tile[2] > 0 ? cfg->tile[2] : 4) : 1;
        tx = std::min(tx, cfg->nx);
        ty = std::min(ty, cfg->ny);
        tz = std::min(tz, nz);
        const size_t lds = sizeof(float) * (6 * (size_t)(tx + 4) * (ty + 4) * (cfg->dim == 3 ? tz + 4 : 1) + 1) +
                           sizeof(double) * 3 * (size_t)(tx + 5) * (ty + 5) * (cfg->dim == 3 ? tz + 5 : 1);
        if (lds > 160 * 1024) {
            g_create_err = "tile too large for the 160 KiB LDS of a workgroup";
            delete ctx;
            return TMPG_EINVAL;
        }
        Tiling tl;
        tl.tx = tx; tl.ty = ty; tl.tz = tz;
        tl.ntx = (cfg->nx + tx - 1) / tx;
        tl.nty = (cfg->ny + ty - 1) / ty;
        tl.ntz = (nz + tz - 1) / tz;
        tl.ntiles = (int64_t)tl.ntx * tl.nty * tl.ntz;
        ctx->tl = tl;
    }
    hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_create_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete ctx;
        return TMPG_EHIP;
    }
    int rc = 0;
    for (int q = 0; q < NF; ++q) rc |= dalloc(ctx, &ctx->f[q], ctx->g.n);
    for (int q = 0; q < 3; ++q) rc |= dalloc(ctx, &ctx->tmp[q], ctx->g.n);
    if (cfg->dim == 3)
        for (int q = 0; q < 6; ++q) rc |= dalloc(ctx, &ctx->prim[q], ctx->g.n);
    rc |= dalloc(ctx, &ctx->red, 16);   // [0,4) energies, [4,6) continuity, [8,12) Gauss
    rc |= dalloc(ctx, &ctx->flags, 16);
    if (rc) {
        g_create_err = ctx->err;
        tmpg_destroy(ctx);
        return TMPG_ENOMEM;
    }
    e = hipSuccess;
    for (int q = 0; q < NF && e == hipSuccess; ++q) e = hipMemsetAsync(ctx->f[q], 0, ctx->g.n * sizeof(float), ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(ctx->flags, 0, 16 * sizeof(int32_t), ctx->stream);
    for (auto& ev : ctx->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e != hipSuccess) {
        g_create_err = std::string("tmpg_create: ") + hipGetErrorString(e);
        tmpg_destroy(ctx);
        return TMPG_EHIP;
    }
    for (int s = 0; s < 2; ++s) {
        // empty, binned species (the push needs valid tile counts for both)
        if (particles_bin(ctx, s, PBuf(), 0, false, cfg->track_ids != 0)) {
            g_create_err = ctx->err;
            tmpg_destroy(ctx);
            return TMPG_ENOMEM;
        }
    }
    e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        g_create_err = std::string("hipStreamSynchronize: ") + hipGetErrorString(e);
        tmpg_destroy(ctx);
        return TMPG_EHIP;
    }
    {
        // set-up of the allocation-free entry point's scratch on this device
        double* sp = nullptr;
        if (noalloc_scratch(ctx, &sp)) {
            g_create_err = ctx->err;
            tmpg_destroy(ctx);
            return TMPG_EHIP;
        }
    }
    *out = ctx;
    return TMPG_OK;
}

int tmpg_destroy(tmpg_ctx* ctx)
{
    if (!ctx) return TMPG_OK;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto& p : ctx->f) dfree(p);
    for (auto& p : ctx->prim) dfree(p);
    for (auto& p : ctx->tmp) dfree(p);
    dfree(ctx->rho);
    dfree(ctx->rho_prev);
    dfree(ctx->rho_work);
    dfree(ctx->red);
    dfree(ctx->flags);
    for (auto& s : ctx->sp) free_species(s);
    for (auto& ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto& ev : ctx->prof) (void)hipEventDestroy(ev);
    if (ctx->cnt_ev) (void)hipEventDestroy(ctx->cnt_ev);
    if (ctx->hostcnt) (void)hipHostFree(ctx->hostcnt);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return TMPG_OK;
}

int tmpg_sync(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->steps_since_check = 0;
    return particles_check(ctx);
}

int tmpg_set_fields(tmpg_ctx* ctx, const float* ex, const float* ey, const float* ez, const float* bx,
                    const float* by, const float* bz)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    const float* a[6] = {ex, ey, ez, bx, by, bz};
    return set_grid(ctx, EX, 6, a);
}

int tmpg_get_fields(tmpg_ctx* ctx, float* ex, float* ey, float* ez, float* bx, float* by, float* bz)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    float* a[6] = {ex, ey, ez, bx, by, bz};
    return get_grid(ctx, EX, 6, a);
}

int tmpg_set_currents(tmpg_ctx* ctx, const float* jx, const float* jy, const float* jz)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    const float* a[3] = {jx, jy, jz};
    return set_grid(ctx, JX, 3, a);
}

int tmpg_get_currents(tmpg_ctx* ctx, float* jx, float* jy, float* jz)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    float* a[3] = {jx, jy, jz};
    return get_grid(ctx, JX, 3, a);
}

// particles from host arrays in the reference's absolute coordinates: they
// replace the species' particles, or are appended to them
static int put_particles(tmpg_ctx* ctx, int species, int64_t n, const float* x, const float* y, const float* z,
                         const float* u, const float* v, const float* w, const float* ch, const int32_t* id,
                         bool append)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    if (species < 0 || species > 1 || n < 0) { ctx->err = "bad species or count"; return TMPG_EINVAL; }
    if (n > 0 && (!x || !y || !u || !v || !w || (ctx->g.dim == 3 && !z))) {
        ctx->err = "null particle array";
        return TMPG_EINVAL;
    }
    Species& sp = ctx->sp[species];
    const int64_t n0 = append ? sp.n : 0;
    if (n0 + n > INT32_MAX && ctx->cfg.track_ids) { ctx->err = "too many particles for int32 ids"; return TMPG_EINVAL; }
    bool need_ch = append && sp.has_ch;
    bool new_ch = false;
    if (ch)
        for (int64_t p = 0; p < n; ++p)
            if (ch[p] != 1.0f) { new_ch = true; break; }
    need_ch = need_ch || new_ch;
    const bool ids = ctx->cfg.track_ids != 0;
    std::vector<int32_t> cell((size_t)n);
    std::vector<float> o[3];
    for (auto& a : o) a.resize((size_t)n);
    const Grid& g = ctx->g;
    const double base_xy = ctx->h + 1, base_z = ctx->hz + 1;
    for (int64_t p = 0; p < n; ++p) {
        // 2D keeps z only in the zigzag scheme, as an offset in its one cell
        const bool zin = g.dim == 3 || (ctx->zz && z != nullptr);
        double r[3] = {(double)x[p] - base_xy, (double)y[p] - base_xy, zin ? (double)z[p] - base_z : 0.5};
        const int nn[3] = {g.nx, g.ny, g.nz};
        int c[3];
        for (int a = 0; a < 3; ++a) {
            if (!std::isfinite(r[a])) { ctx->err = "non-finite particle coordinate"; return TMPG_EINVAL; }
            double fl = std::floor(r[a]);
            double off = r[a] - fl;
            long long ci = (long long)fl;
            ci %= nn[a];
            if (ci < 0) ci += nn[a];
            float of = (float)off;
            if (of >= 1.0f) { of = 0.0f; ci = (ci + 1) % nn[a]; }
            c[a] = (int)ci;
            o[a][p] = of;
        }
        if (g.dim == 2) c[2] = 0;
        cell[p] = pack_cell(ctx->pk, c[0], c[1], c[2]);
    }
    // the species' particles so far, compacted (appending only)
    PBuf old;
    int64_t nold = 0;
    int rc = TMPG_OK;
    if (append) {
        rc = particles_flatten(ctx, species, old, &nold);
        if (rc) { free_pbuf(old); return rc; }
    }
    PBuf flat;
    rc = alloc_pbuf(ctx, flat, std::max<int64_t>(nold + n, 1), need_ch, ids);
    if (rc) { free_pbuf(old); return rc; }
    auto up = [&](void* d, const void* src, size_t bytes) -> int {
        if (!bytes) return TMPG_OK;
        TMPG_HIP_CHECK(ctx, hipMemcpyAsync(d, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return TMPG_OK;
    };
    auto dd = [&](void* d, const void* src, size_t bytes) -> int {
        if (!bytes) return TMPG_OK;
        TMPG_HIP_CHECK(ctx, hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        return TMPG_OK;
    };
    const size_t fb0 = (size_t)nold * sizeof(float), ib0 = (size_t)nold * sizeof(int32_t);
    if (nold > 0) {
        rc |= dd(flat.cell, old.cell, ib0);
        rc |= dd(flat.ox, old.ox, fb0);
        rc |= dd(flat.oy, old.oy, fb0);
        if (flat.oz) rc |= dd(flat.oz, old.oz, fb0);
        rc |= dd(flat.u, old.u, fb0);
        rc |= dd(flat.v, old.v, fb0);
        rc |= dd(flat.w, old.w, fb0);
        if (ids) rc |= dd(flat.id, old.id, ib0);
    }
    std::vector<float> ones;
    if (need_ch) {
        if (old.ch) rc |= dd(flat.ch, old.ch, fb0);
        if (!old.ch || !new_ch) ones.assign((size_t)std::max(nold, n), 1.0f);
        if (!old.ch) rc |= up(flat.ch, ones.data(), fb0);
    }
    const size_t fb = (size_t)n * sizeof(float);
    rc |= up(flat.cell + nold, cell.data(), (size_t)n * sizeof(int32_t));
    rc |= up(flat.ox + nold, o[0].data(), fb);
    rc |= up(flat.oy + nold, o[1].data(), fb);
    if (flat.oz) rc |= up(flat.oz + nold, o[2].data(), fb);
    rc |= up(flat.u + nold, u, fb);
    rc |= up(flat.v + nold, v, fb);
    rc |= up(flat.w + nold, w, fb);
    if (need_ch) rc |= up(flat.ch + nold, new_ch ? ch : ones.data(), fb);
    std::vector<int32_t> tmp;
    if (ids) {
        const int32_t* src = id;
        if (!src) {
            // ids continue after the particles already there
            tmp.resize((size_t)n);
            for (int64_t p = 0; p < n; ++p) tmp[(size_t)p] = (int32_t)(nold + p);
            src = tmp.data();
        }
        rc |= up(flat.id + nold, src, (size_t)n * sizeof(int32_t));
    }
    (void)hipStreamSynchronize(ctx->stream);
    free_pbuf(old);
    if (!rc) rc = particles_bin(ctx, species, flat, nold + n, need_ch, ids);
    (void)hipStreamSynchronize(ctx->stream);
    free_pbuf(flat);
    ctx->move_pending = false;
    ctx->have_rho_prev = false;
    return rc;
}

int tmpg_set_particles(tmpg_ctx* ctx, int species, int64_t n, const float* x, const float* y, const float* z,
                       const float* u, const float* v, const float* w, const float* ch, const int32_t* id)
{
    return put_particles(ctx, species, n, x, y, z, u, v, w, ch, id, false);
}

int tmpg_add_particles(tmpg_ctx* ctx, int species, int64_t n, const float* x, const float* y, const float* z,
                       const float* u, const float* v, const float* w, const float* ch, const int32_t* id)
{
    return put_particles(ctx, species, n, x, y, z, u, v, w, ch, id, true);
}

int tmpg_count_particles(tmpg_ctx* ctx, int species, int64_t* n)
{
    if (!valid_ctx(ctx) || species < 0 || species > 1 || !n) return TMPG_EINVAL;
    *n = ctx->sp[species].n;
    return TMPG_OK;
}

int tmpg_sort_particles(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    if (ctx->move_pending) {
        ctx->err = "tmpg_sort_particles: a tmpg_move awaits its tmpg_deposit";
        return TMPG_ESTATE;
    }
    ctx->steps_since_sort = 0;
    return particles_sort_tiles(ctx);
}

int tmpg_get_particles(tmpg_ctx* ctx, int species, int64_t n_capacity, int64_t* n, float* x, float* y, float* z,
                       float* u, float* v, float* w, float* ch, int32_t* id)
{
    if (!valid_ctx(ctx) || species < 0 || species > 1) return TMPG_EINVAL;
    Species& sp = ctx->sp[species];
    PBuf flat;
    int64_t tot = 0;
    int rc = particles_flatten(ctx, species, flat, &tot);
    if (rc) { free_pbuf(flat); return rc; }
    const int64_t m = std::min<int64_t>(n_capacity, tot);
    if (n) *n = m;
    const Grid& g = ctx->g;
    std::vector<int32_t> cell((size_t)std::max<int64_t>(m, 0));
    std::vector<float> o[3];
    for (auto& a : o) a.resize(cell.size());
    const size_t fb = (size_t)std::max<int64_t>(m, 0) * sizeof(float);
    auto down = [&](void* d, const void* src, size_t bytes) -> int {
        if (!bytes) return TMPG_OK;
        TMPG_HIP_CHECK(ctx, hipMemcpyAsync(d, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        return TMPG_OK;
    };
    if (m > 0) {
        rc |= down(cell.data(), flat.cell, (size_t)m * sizeof(int32_t));
        rc |= down(o[0].data(), flat.ox, fb);
        rc |= down(o[1].data(), flat.oy, fb);
        if (flat.oz) rc |= down(o[2].data(), flat.oz, fb);
        if (u) rc |= down(u, flat.u, fb);
        if (v) rc |= down(v, flat.v, fb);
        if (w) rc |= down(w, flat.w, fb);
        if (id && sp.has_id) rc |= down(id, flat.id, (size_t)m * sizeof(int32_t));
        if (ch && sp.has_ch) rc |= down(ch, flat.ch, fb);
    }
    hipError_t e = hipStreamSynchronize(ctx->stream);
    free_pbuf(flat);
    if (rc) return rc;
    if (e != hipSuccess) { ctx->err = hipGetErrorString(e); return TMPG_EHIP; }
    if (ch && !sp.has_ch) std::fill(ch, ch + std::max<int64_t>(m, 0), 1.0f);
    const double bxy = ctx->h + 1, bz = ctx->hz + 1;
    for (int64_t p = 0; p < m; ++p) {
        int i, j, k;
        unpack_cell(ctx->pk, cell[(size_t)p], i, j, k);
        if (x) x[p] = (float)(bxy + i + (double)o[0][(size_t)p]);
        if (y) y[p] = (float)(bxy + j + (double)o[1][(size_t)p]);
        if (z) z[p] = (float)(g.dim == 3 ? bz + k + (double)o[2][(size_t)p]
                                         : bz + (ctx->zz ? (double)o[2][(size_t)p] : 0.5));
    }
    return TMPG_OK;
}

int tmpg_advance_b_half(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    return fields_bhalf(ctx);
}

int tmpg_advance_e(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    return fields_efull(ctx, false);
}

int tmpg_reset_currents(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    return fields_zero_j(ctx);
}

static int ensure_rho(tmpg_ctx* ctx)
{
    int rc = 0;
    if (!ctx->rho) rc |= dalloc(ctx, &ctx->rho, ctx->g.n);
    if (!ctx->rho_prev) rc |= dalloc(ctx, &ctx->rho_prev, ctx->g.n);
    return rc ? TMPG_ENOMEM : TMPG_OK;
}

static int pre_push(tmpg_ctx* ctx)
{
    if (!ctx->track_cont) return TMPG_OK;
    int rc = ensure_rho(ctx);
    if (rc) return rc;
    rc = diag_rho(ctx, ctx->rho_prev);
    ctx->have_rho_prev = (rc == TMPG_OK);
    return rc;
}

int tmpg_move(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    for (auto& s : ctx->sp)
        if (s.n > 0) {
            int rc = alloc_split_scratch(ctx, s);
            if (rc) return rc;
        }
    int rc = pre_push(ctx);
    if (rc) return rc;
    rc = particles_push(ctx, 1);
    if (rc == TMPG_OK) ctx->move_pending = true;
    return rc;
}

int tmpg_deposit(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    if (!ctx->move_pending) { ctx->err = "tmpg_deposit: no tmpg_move to deposit"; return TMPG_ESTATE; }
    int rc = particles_push(ctx, 2);
    if (rc == TMPG_OK) ctx->move_pending = false;
    return rc;
}

int tmpg_push_deposit(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    int rc = pre_push(ctx);
    if (rc) return rc;
    ctx->move_pending = false;
    return particles_push(ctx, 0);
}

int tmpg_fold_currents(tmpg_ctx* ctx) { return valid_ctx(ctx) ? TMPG_OK : TMPG_EINVAL; }

int tmpg_filter(tmpg_ctx* ctx, int npasses)
{
    if (!valid_ctx(ctx) || npasses < 0) return TMPG_EINVAL;
    return fields_filter(ctx, npasses);
}

int tmpg_add_current(tmpg_ctx* ctx)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    return fields_add_current(ctx);
}

// elapsed times of the phases of the tmpg_step recorded in ring slot `slot`
// (its events complete), added to ctx->prof_sum
static int prof_fold(tmpg_ctx* ctx, int slot)
{
    hipEvent_t* ev = &ctx->prof[(size_t)slot * 8];
    TMPG_HIP_CHECK(ctx, hipEventSynchronize(ev[5]));
    float t[6] = {0, 0, 0, 0, 0, 0};
    TMPG_HIP_CHECK(ctx, hipEventElapsedTime(&t[0], ev[0], ev[1]));
    TMPG_HIP_CHECK(ctx, hipEventElapsedTime(&t[1], ev[1], ev[2]));
    TMPG_HIP_CHECK(ctx, hipEventElapsedTime(&t[2], ev[2], ev[3]));
    TMPG_HIP_CHECK(ctx, hipEventElapsedTime(&t[3], ev[3], ev[4]));
    TMPG_HIP_CHECK(ctx, hipEventElapsedTime(&t[4], ev[4], ev[5]));
    // the push kernel's events are recorded only when it ran (no particles:
    // none)
    if (ctx->prof_push[(size_t)slot]) TMPG_HIP_CHECK(ctx, hipEventElapsedTime(&t[5], ev[6], ev[7]));
    double* ms = ctx->prof_sum;
    ms[0] += (double)t[0] + t[2] + t[4];
    ms[1] += t[1];
    ms[2] += t[3];
    ms[4] += (double)t[0] + t[1] + t[2] + t[3] + t[4];
    ms[5] += t[5];
    ctx->prof_steps++;
    return TMPG_OK;
}

int tmpg_step(tmpg_ctx* ctx, int nsteps)
{
    if (!valid_ctx(ctx) || nsteps < 0) return TMPG_EINVAL;
    int rc = 0;
    if (!ctx->hostcnt) {
        TMPG_HIP_CHECK(ctx, hipHostMalloc((void**)&ctx->hostcnt, 8 * sizeof(int32_t), hipHostMallocDefault));
        TMPG_HIP_CHECK(ctx, hipEventCreateWithFlags(&ctx->cnt_ev, hipEventDisableTiming));
    }
    auto record = [&](hipEvent_t e) -> int {
        TMPG_HIP_CHECK(ctx, hipEventRecord(e, ctx->stream));
        return TMPG_OK;
    };
    for (int s = 0; s < nsteps && rc == 0; ++s) {
        if (ctx->cnt_pending) {
            // overflow counters copied at an earlier step, read without a
            // wait: a filling overflow list re-bins now rather than at the
            // next 32-step check
            if (hipEventQuery(ctx->cnt_ev) == hipSuccess) {
                ctx->cnt_pending = false;
                bool check = false;
                for (int q = 0; q < 2; ++q)
                    check |= ctx->sp[q].ocnt && (ctx->hostcnt[4 * q + 2] > 0 ||
                                                 (int64_t)ctx->hostcnt[4 * q] > ctx->sp[q].ocap / 8);
                if (check) {
                    ctx->steps_since_check = 0;
                    rc = particles_check(ctx);
                    if (rc) break;
                }
            } else {
                (void)hipGetLastError();   // not ready yet: no sticky status
            }
        }
        if (++ctx->steps_since_check >= 32) {
            // lost particles are an error; a filling overflow list re-bins
            ctx->steps_since_check = 0;
            rc = particles_check(ctx);
            if (rc) break;
        }
        // per-phase HIP events on the context's stream, read by
        // tmpg_last_timings; the ring slot of the step kProfRing steps back is
        // folded into the running sums before it is reused
        while (ctx->prof.size() < (size_t)tmpg_ctx::kProfRing * 8) {
            hipEvent_t e;
            TMPG_HIP_CHECK(ctx, hipEventCreate(&e));
            ctx->prof.push_back(e);
        }
        ctx->prof_push.resize(tmpg_ctx::kProfRing, 0);
        if (ctx->prof_unread == tmpg_ctx::kProfRing) {
            rc = prof_fold(ctx, ctx->prof_head);
            if (rc) break;
            ctx->prof_unread--;
        }
        const int slot = ctx->prof_head;
        hipEvent_t* ev = &ctx->prof[(size_t)slot * 8];
        // code/tristanmainloop.F90:117-242 with the ghost fills and the
        // current fold absorbed into periodic indexing, reset_currents moved
        // before the push and add_current fused into the E update
        TMPG_HIP_CHECK(ctx, hipEventRecord(ev[0], ctx->stream));
        rc = fields_bhalf(ctx);
        if (!rc) rc = fields_zero_j(ctx);
        if (!rc) rc = record(ev[1]);
        if (!rc && ctx->sort_every > 0 && ++ctx->steps_since_sort >= ctx->sort_every) {
            ctx->steps_since_sort = 0;
            rc = particles_sort_tiles(ctx);
        }
        if (!rc) rc = pre_push(ctx);
        ctx->push_events = ev + 6;   // around the tiled push kernel itself
        ctx->push_recorded = false;
        if (!rc) rc = particles_push(ctx, 0);
        ctx->push_events = nullptr;
        if (!rc) rc = record(ev[2]);
        if (!rc) rc = fields_bhalf(ctx);
        if (!rc) rc = record(ev[3]);
        if (!rc) rc = ctx->denoiser ? denoise_apply(ctx, ctx->denoiser) : fields_filter(ctx, ctx->cfg.ntimes);
        if (!rc) rc = record(ev[4]);
        if (!rc) rc = fields_efull(ctx, true);
        if (!rc) rc = record(ev[5]);
        if (rc) break;    // a failed step leaves its slot unclaimed
        ctx->prof_push[(size_t)slot] = ctx->push_recorded ? 1 : 0;
        ctx->prof_head = (slot + 1) % tmpg_ctx::kProfRing;
        ctx->prof_unread++;
        if (!rc && !ctx->cnt_pending && (ctx->steps_since_check & 3) == 0) {
            for (int q = 0; q < 2; ++q) {
                if (ctx->sp[q].ocnt)
                    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(ctx->hostcnt + 4 * q, ctx->sp[q].ocnt, 4 * sizeof(int32_t),
                                                       hipMemcpyDeviceToHost, ctx->stream));
            }
            TMPG_HIP_CHECK(ctx, hipEventRecord(ctx->cnt_ev, ctx->stream));
            ctx->cnt_pending = true;
        }
    }
    ctx->move_pending = false;
    return rc;
}

int tmpg_last_timings(tmpg_ctx* ctx, double ms[6])
{
    if (!valid_ctx(ctx) || !ms) return TMPG_EINVAL;
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    // fold the unread steps, oldest first
    const int R = tmpg_ctx::kProfRing;
    for (; ctx->prof_unread > 0; --ctx->prof_unread) {
        const int rc = prof_fold(ctx, ((ctx->prof_head - ctx->prof_unread) % R + R) % R);
        if (rc) return rc;
    }
    for (int q = 0; q < 6; ++q) ms[q] = ctx->prof_sum[q];
    ms[3] = (double)ctx->prof_steps;
    for (auto& v : ctx->prof_sum) v = 0.0;
    ctx->prof_steps = 0;
    return TMPG_OK;
}

int tmpg_track_continuity(tmpg_ctx* ctx, int on)
{
    if (!valid_ctx(ctx)) return TMPG_EINVAL;
    ctx->track_cont = on != 0;
    ctx->have_rho_prev = false;
    return on ? ensure_rho(ctx) : TMPG_OK;
}

int tmpg_charge_density(tmpg_ctx* ctx, double* rho)
{
    if (!valid_ctx(ctx) || !rho) return TMPG_EINVAL;
    int rc = ensure_rho(ctx);
    if (rc) return rc;
    rc = diag_rho(ctx, ctx->rho);
    if (rc) return rc;
    TMPG_HIP_CHECK(ctx, hipMemcpyAsync(rho, ctx->rho, ctx->g.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    TMPG_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return TMPG_OK;
}

int tmpg_diag(tmpg_ctx* ctx, double out[8])
{
    if (!valid_ctx(ctx) || !out) return TMPG_EINVAL;
    double e[4];
    int rc = diag_energies(ctx, e);
    if (rc) return rc;
    out[0] = e[0];
    out[1] = e[1];
    out[2] = e[2];
    out[3] = e[3];
    out[4] = (double)ctx->sp[0].n;
    out[5] = (double)ctx->sp[1].n;
    out[6] = -1.0;
    out[7] = 0.0;
    if (ctx->track_cont && ctx->have_rho_prev) {
        rc = diag_continuity(ctx, &out[6], &out[7]);
        if (rc) return rc;
    }
    return TMPG_OK;
}

int tmpg_gauss_residual(tmpg_ctx* ctx, double out[4])
{
    if (!valid_ctx(ctx) || !out) return TMPG_EINVAL;
    int rc = ensure_rho(ctx);
    if (!rc && !ctx->rho_work) rc = dalloc(ctx, &ctx->rho_work, ctx->g.n);
    if (rc) return rc ? TMPG_ENOMEM : TMPG_OK;
    int np = ctx->cfg.ntimes;
    if (ctx->denoiser) {
        rc = tmpg_model_get_base_passes(ctx->denoiser, &np);
        if (rc) { ctx->err = "tmpg_gauss_residual: denoiser base passes"; return rc; }
    }
    return diag_gauss(ctx, ctx->rho, ctx->rho_work, ctx->red + 8, np, false, out);
}

int tmpg_gauss_residual_noalloc(tmpg_ctx* ctx, double out[4])
{
    if (!valid_ctx(ctx) || !out) return TMPG_EINVAL;
    int np = ctx->cfg.ntimes;
    if (ctx->denoiser) {
        const int rc = tmpg_model_get_base_passes(ctx->denoiser, &np);
        if (rc) { ctx->err = "tmpg_gauss_residual_noalloc: denoiser base passes"; return rc; }
    }
    return diag_gauss_noalloc(ctx, np, out);
}

int tmpg_load_uniform(tmpg_ctx* ctx, int species, int ppc, float drift_gb, float theta, uint64_t seed)
{
    if (!valid_ctx(ctx) || species < 0 || species > 1 || ppc < 0 || theta < 0.0f) return TMPG_EINVAL;
    Species& sp = ctx->sp[species];
    if (sp.has_ch) { ctx->err = "tmpg_load_uniform: species has per-particle weights"; return TMPG_ESTATE; }
    ctx->have_rho_prev = false;
    return particles_load_uniform(ctx, species, ppc, drift_gb, theta, seed);
}

// ------------------------------------------------------------------ input
// The reference's input format (code/par.c): "<block>" lines open a block,
// "key = value" lines set a key, "#" starts a comment.
struct tmpg_input {
    std::map<std::string, std::string> kv;  // "block/key" -> value
};

static std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

int tmpg_input_open(const char* path, tmpg_input** out)
{
    if (!path || !out) return TMPG_EINVAL;
    std::ifstream in(path);
    if (!in) { g_create_err = std::string("cannot open input file ") + path; return TMPG_EINVAL; }
    tmpg_input* t = new (std::nothrow) tmpg_input();
    if (!t) return TMPG_ENOMEM;
    std::string line, block;
    while (std::getline(in, line)) {
        size_t h = line.find('#');
        if (h != std::string::npos) line = line.substr(0, h);
        line = trim(line);
        if (line.empty()) continue;
        if (line.front() == '<') {
            size_t e = line.find('>');
            block = trim(line.substr(1, e == std::string::npos ? std::string::npos : e - 1));
            continue;
        }
        size_t eq = line.find('=');
        if (eq == std::string::npos) continue;
        t->kv[block + "/" + trim(line.substr(0, eq))] = trim(line.substr(eq + 1));
    }
    *out = t;
    return TMPG_OK;
}

int tmpg_input_close(tmpg_input* in)
{
    delete in;
    return TMPG_OK;
}

int tmpg_input_get_double(const tmpg_input* in, const char* block, const char* key, double def, double* val)
{
    if (!in || !block || !key || !val) return TMPG_EINVAL;
    auto it = in->kv.find(std::string(block) + "/" + key);
    if (it == in->kv.end()) { *val = def; return TMPG_OK; }
    char* end = nullptr;
    double d = std::strtod(it->second.c_str(), &end);
    if (end == it->second.c_str()) return TMPG_EINVAL;
    *val = d;
    return TMPG_OK;
}

int tmpg_input_get_int(const tmpg_input* in, const char* block, const char* key, int def, int* val)
{
    double d;
    int rc = tmpg_input_get_double(in, block, key, (double)def, &d);
    if (rc) return rc;
    *val = (int)std::lround(d);
    return TMPG_OK;
}

}  // extern "C"
