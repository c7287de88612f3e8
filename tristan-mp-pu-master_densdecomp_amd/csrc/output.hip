// output.hip -- the reference's field output file flds.tot.NNN (HDF5).
//
// output_tot (code/output.F90:2494-3875) writes, per output lap, one HDF5
// file of FP32 datasets sampled from the padded arrays at the nodes
// i = istep, 2 istep, ... <= mx (likewise j, k; in 2D k = 1): with one
// domain its istart..ifinish loop (:2645-2700) starts at 1 and ends at mx,
// so a dataset holds mx/istep x my/istep x mz/istep values, the reference's
// mx0/istep with mx0 counting the ghost cells (code/fields.F90:186).
// Datasets, in the file order of :2604-2628:
//   dens, densi   meanq_fld_cur('tdens' / 'idens') (:5229-5482): each
//                 particle adds ch to every node of the box of +-2 cells
//                 around its cell (+-0 in z in 2D), clamped to the padded
//                 array; the ghost sums are folded onto their periodic
//                 interior images like exchange_current
//                 (code/fieldboundaries.F90:1768-2189, one domain; the
//                 ghosts keep their own sums) and every node is divided by
//                 its clamped box volume
//   v3x .. v3zi   the same sums of u/gamma ch over those of ch: the mean
//                 3-velocity of all particles / of the ions, 0 where no
//                 particle reaches (no v3z, v3zi in 2D with sigma = 0 and
//                 pcosthmult = 0, the 15-dataset list)
//   ex .. bz      interpolfld (:4995-5127) at the integer point: neighbour
//                 averages of the staggered fields, with the node clamped
//                 to 2..mx-1 (2..my-1, 2..mz-1; k = 1 in 2D)
//   jx, jy, jz    curx, cury, curz at the nodes
// Arrays are Fortran order (x fastest), so the HDF5 dimensions are
// (mz/istep, my/istep, mx/istep).  The sums run in FP64 (the reference's
// are FP32), over per-cell sums swept by separable box sums, which add the
// same terms as the reference's per-particle boxes.
//
// The HDF5 C library is opened at run time (dlopen), so libtmpg.so does not
// depend on it.
#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tmpg_internal.h"

namespace {

// the HDF5 C API entry points used here (HDF5 >= 1.10: hid_t is int64_t)
using hid_t = int64_t;
using herr_t = int;
using hsize_t = unsigned long long;
struct H5 {
    void* lib = nullptr;
    herr_t (*open)() = nullptr;
    hid_t (*fcreate)(const char*, unsigned, hid_t, hid_t) = nullptr;
    herr_t (*fclose)(hid_t) = nullptr;
    hid_t (*screate_simple)(int, const hsize_t*, const hsize_t*) = nullptr;
    herr_t (*sclose)(hid_t) = nullptr;
    hid_t (*dcreate2)(hid_t, const char*, hid_t, hid_t, hid_t, hid_t, hid_t) = nullptr;
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void*) = nullptr;
    herr_t (*dclose)(hid_t) = nullptr;
    hid_t* native_float = nullptr;   // H5T_NATIVE_FLOAT_g, set by H5open
};
constexpr unsigned kH5FAccTrunc = 0x0002u;   // H5F_ACC_TRUNC
constexpr hid_t kH5PDefault = 0, kH5SAll = 0;

template <class F>
void bind(void* lib, F& fn, const char* name)
{
    fn = reinterpret_cast<F>(dlsym(lib, name));
}

bool load_h5(H5& h, std::string& err)
{
    const char* env = std::getenv("TMPG_HDF5_LIB");
    const char* names[] = {env, "libhdf5.so", "/opt/conda/lib/libhdf5.so"};
    for (const char* n : names) {
        if (!n || !*n) continue;
        h.lib = dlopen(n, RTLD_NOW | RTLD_LOCAL);
        if (h.lib) break;
    }
    if (!h.lib) {
        err = "the HDF5 C library was not found (TMPG_HDF5_LIB names its libhdf5.so)";
        return false;
    }
    bind(h.lib, h.open, "H5open");
    bind(h.lib, h.fcreate, "H5Fcreate");
    bind(h.lib, h.fclose, "H5Fclose");
    bind(h.lib, h.screate_simple, "H5Screate_simple");
    bind(h.lib, h.sclose, "H5Sclose");
    bind(h.lib, h.dcreate2, "H5Dcreate2");
    bind(h.lib, h.dwrite, "H5Dwrite");
    bind(h.lib, h.dclose, "H5Dclose");
    h.native_float = reinterpret_cast<hid_t*>(dlsym(h.lib, "H5T_NATIVE_FLOAT_g"));
    if (!h.open || !h.fcreate || !h.fclose || !h.screate_simple || !h.sclose || !h.dcreate2 || !h.dwrite ||
        !h.dclose || !h.native_float) {
        err = "the HDF5 C library lacks an entry point of the 1.10 API";
        return false;
    }
    if (h.open() < 0) {
        err = "H5open failed";
        return false;
    }
    return true;
}

// the padded grid of the reference (1-based indices in the comments)
struct Pad {
    int dim, mx, my, mz, ng, ngz;
    int64_t at(int i, int j, int k) const   // 1-based
    {
        return (int64_t)(i - 1) + (int64_t)mx * ((j - 1) + (int64_t)my * (k - 1));
    }
    int64_t cells() const { return (int64_t)mx * my * mz; }
};

// the offset of the first cell of line l along an axis
int64_t line_base(const Pad& p, int axis, int64_t l)
{
    if (axis == 0) return l * p.mx;
    if (axis == 1) return (l % p.mx) + (l / p.mx) * (int64_t)p.mx * p.my;
    return l;
}

// separable box sum of half-width r along one axis, the window clamped to
// the array (meanq_fld_cur's lx1 = max(i-idx,1), lx2 = min(i+idx,mx))
void box_axis(const Pad& p, int axis, int r, std::vector<double>& a, std::vector<double>& tmp)
{
    if (r == 0) return;
    const int n = axis == 0 ? p.mx : axis == 1 ? p.my : p.mz;
    const int64_t stride = axis == 0 ? 1 : axis == 1 ? p.mx : (int64_t)p.mx * p.my;
    tmp.assign(a.size(), 0.0);
    const int64_t lines = p.cells() / n;
    for (int64_t l = 0; l < lines; ++l) {
        const int64_t base = line_base(p, axis, l);
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int q = std::max(i - r, 0); q <= std::min(i + r, n - 1); ++q) s += a[(size_t)(base + q * stride)];
            tmp[(size_t)(base + i * stride)] = s;
        }
    }
    a.swap(tmp);
}

// exchange_current's periodic fold on one domain: the upper ghosts
// n-h..n are added to h+1..ng, the lower ghosts 1..h to n-ng+1..n-h-1
// (x, then y, then z in 3D); the ghosts keep their values
void fold(const Pad& p, std::vector<double>& a)
{
    std::vector<double> up, lo;
    for (int axis = 0; axis < p.dim; ++axis) {
        const int n = axis == 0 ? p.mx : axis == 1 ? p.my : p.mz;
        const int gh = axis == 2 ? p.ngz : p.ng;
        const int h = gh / 2;
        const int64_t stride = axis == 0 ? 1 : axis == 1 ? p.mx : (int64_t)p.mx * p.my;
        const int64_t lines = p.cells() / n;
        up.resize((size_t)h + 1);
        lo.resize((size_t)h);
        for (int64_t l = 0; l < lines; ++l) {
            const int64_t b = line_base(p, axis, l);
            auto A = [&](int i) -> double& { return a[(size_t)(b + (int64_t)(i - 1) * stride)]; };
            for (int q = 0; q <= h; ++q) up[(size_t)q] = A(n - h + q);
            for (int q = 0; q < h; ++q) lo[(size_t)q] = A(1 + q);
            for (int q = 0; q <= h; ++q) A(h + 1 + q) += up[(size_t)q];
            for (int q = 0; q < h; ++q) A(n - gh + 1 + q) += lo[(size_t)q];
        }
    }
}

}  // namespace

extern "C" int tmpg_flds_tot_write(const char* path, const tmpg_flds_tot_state* st, char* errbuf, int errlen)
{
    std::string err;
    auto fail = [&](int rc) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, (size_t)errlen, "%s", err.c_str());
        return rc;
    };
    if (!path || !st || (st->dim != 2 && st->dim != 3) || st->order < 1 || st->order > 3 || st->istep < 1 ||
        st->nx < 1 || st->ny < 1 || (st->dim == 3 && st->nz < 1) || st->ni < 0 || st->ne < 0) {
        err = "tmpg_flds_tot_write: bad arguments";
        return fail(TMPG_EINVAL);
    }
    const float* f[9] = {st->ex, st->ey, st->ez, st->bx, st->by, st->bz, st->jx, st->jy, st->jz};
    for (const float* a : f)
        if (!a) {
            err = "tmpg_flds_tot_write: a field array is NULL";
            return fail(TMPG_EINVAL);
        }
    if ((st->ni > 0 && !st->ions) || (st->ne > 0 && !st->lecs)) {
        err = "tmpg_flds_tot_write: a particle array is NULL";
        return fail(TMPG_EINVAL);
    }
    const int dim = st->dim;
    Pad p;
    p.dim = dim;
    p.ng = st->order == 1 ? 5 : 7;
    p.ngz = dim == 2 ? 5 : p.ng;
    p.mx = st->nx + p.ng;
    p.my = st->ny + p.ng;
    p.mz = dim == 3 ? st->nz + p.ngz : 1;
    const int is = st->istep;
    const int ox = p.mx / is, oy = p.my / is, oz = dim == 3 ? p.mz / is : 1;
    if (ox < 1 || oy < 1 || oz < 1) {
        err = "tmpg_flds_tot_write: istep exceeds the padded grid";
        return fail(TMPG_EINVAL);
    }
    // every particle's cell must lie in the padded grid
    for (int s = 0; s < 2; ++s) {
        const int64_t n = s == 0 ? st->ni : st->ne;
        const float* q = s == 0 ? st->ions : st->lecs;
        for (int64_t t = 0; t < n; ++t) {
            const float x = q[t], y = q[n + t], z = dim == 3 ? q[2 * n + t] : 1.0f;
            if (!(x >= 1.0f && x < (float)(p.mx + 1) && y >= 1.0f && y < (float)(p.my + 1) && z >= 1.0f &&
                  z < (float)(p.mz + 1))) {
                err = "tmpg_flds_tot_write: a particle lies outside the padded grid";
                return fail(TMPG_EINVAL);
            }
        }
    }

    H5 h5;
    if (!load_h5(h5, err)) return fail(TMPG_EINVAL);
    const hid_t file = h5.fcreate(path, kH5FAccTrunc, kH5PDefault, kH5PDefault);
    if (file < 0) {
        err = std::string("H5Fcreate failed: ") + path;
        return fail(TMPG_EINVAL);
    }

    const int64_t nc = p.cells();
    std::vector<double> num, den, tmp;
    std::vector<float> buf((size_t)ox * oy * oz);
    const int rx = 2, ry = 2, rz = dim == 3 ? 2 : 0;   // idx, idy, idz (code/output.F90:189-194)
    // dataset d < 8: 0 dens, 1 densi, then v3x, v3xi, v3y, v3yi, v3z, v3zi
    auto moment = [&](int d) {
        const bool ions_only = d & 1;
        const bool ratio = d >= 2;
        const int comp = (d - 2) / 2;   // 0 u, 1 v, 2 w
        num.assign((size_t)nc, 0.0);
        if (ratio) den.assign((size_t)nc, 0.0);
        for (int s = 0; s < (ions_only ? 1 : 2); ++s) {
            const int64_t n = s == 0 ? st->ni : st->ne;
            const float* q = s == 0 ? st->ions : st->lecs;
            for (int64_t t = 0; t < n; ++t) {
                const float u = q[3 * n + t], v = q[4 * n + t], w = q[5 * n + t], ch = q[6 * n + t];
                const int64_t c = p.at((int)q[t], (int)q[n + t], dim == 3 ? (int)q[2 * n + t] : 1);
                if (ratio) {
                    const float gi = 1.0f / std::sqrt(1.0f + u * u + v * v + w * w);
                    const float uc = comp == 0 ? u : comp == 1 ? v : w;
                    num[(size_t)c] += (double)(uc * gi * ch);
                    den[(size_t)c] += (double)ch;
                } else {
                    num[(size_t)c] += (double)ch;
                }
            }
        }
        box_axis(p, 0, rx, num, tmp);
        box_axis(p, 1, ry, num, tmp);
        box_axis(p, 2, rz, num, tmp);
        fold(p, num);
        if (ratio) {
            box_axis(p, 0, rx, den, tmp);
            box_axis(p, 1, ry, den, tmp);
            box_axis(p, 2, rz, den, tmp);
            fold(p, den);
        }
        for (int k = 1; k <= p.mz; ++k)
            for (int j = 1; j <= p.my; ++j)
                for (int i = 1; i <= p.mx; ++i) {
                    const size_t c = (size_t)p.at(i, j, k);
                    if (ratio) {
                        num[c] = den[c] != 0.0 ? num[c] / den[c] : 0.0;
                    } else {
                        const int vx = std::min(i + rx, p.mx) - std::max(i - rx, 1) + 1;
                        const int vy = std::min(j + ry, p.my) - std::max(j - ry, 1) + 1;
                        const int vz = std::min(k + rz, p.mz) - std::max(k - rz, 1) + 1;
                        num[c] /= (double)vx * vy * vz;
                    }
                }
    };
    // interpolfld at the integer point (i, j, k): dx = dy = dz = 0 there,
    // its linear terms vanish and the neighbour averages remain
    const int64_t sx = 1, sy = p.mx, sz = dim == 3 ? (int64_t)p.mx * p.my : 0;
    auto node_field = [&](int d, int i, int j, int k) -> float {
        i = i == 1 ? 2 : i == p.mx ? p.mx - 1 : i;
        j = j == 1 ? 2 : j == p.my ? p.my - 1 : j;
        k = dim == 3 ? (k == 1 ? 2 : k == p.mz ? p.mz - 1 : k) : 1;
        const int64_t l = p.at(i, j, k);
        const float* a = f[d];
        switch (d) {
        case 0: return (a[l] + a[l - sx]) * 0.5f;
        case 1: return (a[l] + a[l - sy]) * 0.5f;
        case 2: return (a[l] + a[l - sz]) * 0.5f;
        case 3: return ((a[l] + a[l - sz]) + (a[l - sy] + a[l - sy - sz])) * 0.25f;
        case 4: return ((a[l] + a[l - sx]) + (a[l - sz] + a[l - sz - sx])) * 0.25f;
        default: return ((a[l] + a[l - sy]) + (a[l - sx] + a[l - sx - sy])) * 0.25f;
        }
    };

    const bool v3z = !(dim == 2 && st->in_plane);
    static const char* const names[17] = {"dens", "densi", "v3x", "v3xi", "v3y", "v3yi", "v3z", "v3zi", "ex",
                                          "ey",   "ez",    "bx",  "by",   "bz",  "jx",   "jy",  "jz"};
    const hsize_t dims[3] = {(hsize_t)oz, (hsize_t)oy, (hsize_t)ox};
    int rc = TMPG_OK;
    for (int d = 0; d < 17 && rc == TMPG_OK; ++d) {
        if (!v3z && (d == 6 || d == 7)) continue;
        if (d < 8) moment(d);
        for (int c = 0; c < oz; ++c)
            for (int b = 0; b < oy; ++b)
                for (int a = 0; a < ox; ++a) {
                    const int i = (a + 1) * is, j = (b + 1) * is, k = dim == 3 ? (c + 1) * is : 1;
                    float val;
                    if (d < 8) val = (float)num[(size_t)p.at(i, j, k)];
                    else if (d < 14) val = node_field(d - 8, i, j, k);
                    else val = f[d - 8][p.at(i, j, k)];
                    buf[(size_t)a + (size_t)ox * (b + (size_t)oy * c)] = val;
                }
        const hid_t space = h5.screate_simple(3, dims, nullptr);
        const hid_t ds = space >= 0 ? h5.dcreate2(file, names[d], *h5.native_float, space, kH5PDefault,
                                                  kH5PDefault, kH5PDefault)
                                    : -1;
        if (ds < 0 || h5.dwrite(ds, *h5.native_float, kH5SAll, kH5SAll, kH5PDefault, buf.data()) < 0) {
            err = std::string("the HDF5 write of ") + names[d] + " failed";
            rc = TMPG_EINVAL;
        }
        if (ds >= 0) h5.dclose(ds);
        if (space >= 0) h5.sclose(space);
    }
    if (h5.fclose(file) < 0 && rc == TMPG_OK) {
        err = "H5Fclose failed";
        rc = TMPG_EINVAL;
    }
    return rc == TMPG_OK ? TMPG_OK : fail(rc);
}

extern "C" int tmpg_write_flds_tot(tmpg_ctx* ctx, const char* path, int istep, int in_plane)
{
    if (!ctx || !path || istep < 1) return TMPG_EINVAL;
    const size_t np = (size_t)ctx->mx * ctx->my * ctx->mz;
    std::vector<float> g(9 * np);
    float* a[9];
    for (int q = 0; q < 9; ++q) a[q] = g.data() + q * np;
    int rc = tmpg_get_fields(ctx, a[0], a[1], a[2], a[3], a[4], a[5]);
    if (!rc) rc = tmpg_get_currents(ctx, a[6], a[7], a[8]);
    if (rc) return rc;
    std::vector<float> prt[2];
    int64_t cnt[2] = {0, 0};
    for (int s = 0; s < 2 && !rc; ++s) {
        int64_t n = 0;
        rc = tmpg_count_particles(ctx, s, &n);
        if (rc) break;
        prt[s].resize((size_t)7 * (size_t)std::max<int64_t>(n, 1));
        float* q = prt[s].data();
        rc = tmpg_get_particles(ctx, s, n, &cnt[s], q, q + n, q + 2 * n, q + 3 * n, q + 4 * n, q + 5 * n,
                                q + 6 * n, nullptr);
        if (!rc && cnt[s] != n) {
            ctx->err = "tmpg_write_flds_tot: the particle count changed during the copy";
            rc = TMPG_ESTATE;
        }
    }
    if (rc) return rc;
    tmpg_flds_tot_state st{};
    st.dim = ctx->g.dim;
    st.order = ctx->cfg.order;
    st.nx = ctx->g.nx;
    st.ny = ctx->g.ny;
    st.nz = ctx->g.nz;
    st.istep = istep;
    st.in_plane = in_plane;
    st.ex = a[0];
    st.ey = a[1];
    st.ez = a[2];
    st.bx = a[3];
    st.by = a[4];
    st.bz = a[5];
    st.jx = a[6];
    st.jy = a[7];
    st.jz = a[8];
    st.ni = cnt[0];
    st.ne = cnt[1];
    st.ions = prt[0].data();
    st.lecs = prt[1].data();
    char eb[256] = {0};
    rc = tmpg_flds_tot_write(path, &st, eb, (int)sizeof eb);
    if (rc) ctx->err = eb;
    return rc;
}
