// loader.hip -- the reference's initial-plasma generator on the host.
//
// tmpg_inject_plasma_region follows inject_plasma_region
// (code/particles.F90:2549-2938) for one periodic domain: the region's
// clamping to the interior, the particle count (Poisson below 10 expected
// particles, code/aux.F90:109-134), uniform positions, and for each ion an
// electron at the same place, both with momenta from maxwell_dist
// (code/particles.F90:2176-2273: a tabulated Maxwell-Juttner energy,
// init_maxw_table at :2126-2163, isotropic angles, Zenitani's flipping for
// the drift).  Random numbers come from the reference's Park-Miller
// generator random() (code/aux.F90:82-99) on the caller's seed, in the
// reference's order of draws, so a run seeded like the reference
// (123457 + rank, code/communications.F90:224-231) places its particles at
// the same FP32 positions; momenta agree to the last bits of the math
// library's cos/sin/exp/sqrt.  FP32 where the reference computes in real,
// FP64 where it does (the seed, the Poisson product, pi).
#include <cmath>
#include <cstdint>
#include <vector>

#include "tmpg_internal.h"

namespace {

#pragma clang fp contract(off)

constexpr int kPdfSz = 1000;   // pdf_sz, code/particles.F90:45

// random(): code/aux.F90:82-99
inline float park_miller(double& seed)
{
    seed = std::fmod(16807.0 * seed, 2147483647.0);
    return (float)(seed / 2147483648.0);
}

// poisson(): code/aux.F90:109-134
float poisson(float numps, double& seed)
{
    const double L = std::exp(-(double)numps);
    float k = 0.0f;
    double p = 1.0;
    while (p >= L) {
        k = k + 1.0f;
        const float u = park_miller(seed);
        p = p * (double)u;
    }
    return k - 1.0f;
}

struct MaxwTable {
    float gamma[kPdfSz];
    float pdf[kPdfSz + 1];
};

// init_maxw_table: the integrated Maxwell-Juttner PDF over kinetic energy
// gamma - 1 in [0, 20 delgam] (3D spread; in 2D with pcosthmult = 0 the
// in-plane form)
void init_maxw_table(float delgam, bool three_d, MaxwTable& t)
{
    const float maxg = delgam * 20.0f + 1.0f;
    float func[kPdfSz];
    for (int i = 0; i < kPdfSz; ++i) {
        const float step = (maxg - 1.0f) / (float)(kPdfSz - 1);
        t.gamma[i] = step * (float)i;
        const float g = t.gamma[i];
        const float e = std::exp(-g / delgam);
        if (three_d) {
            const float r = std::sqrt(g * (g + 2.0f));
            const float a = (g + 1.0f) * r;
            func[i] = a * e;
        } else {
            func[i] = (g + 1.0f) * e;
        }
    }
    t.pdf[0] = 0.0f;
    for (int i = 0; i < kPdfSz; ++i) {
        float s = 0.0f;
        for (int k = 0; k <= i; ++k) s = s + func[k];
        t.pdf[i + 1] = s;
    }
    const float norm = t.pdf[kPdfSz - 1];
    for (int i = 0; i <= kPdfSz; ++i) t.pdf[i] = t.pdf[i] / norm;
}

// maxwell_dist: momenta (gamma*beta) of one particle
void maxwell_dist(float gamma0, float cd, double& seed, const MaxwTable& t, bool three_d, int pcosthmult,
                  float sigma, float& u, float& v, float& w)
{
    const double pi = (double)3.1415927f;   // code/particles.F90:292, a default-real literal
    const float gamma0mag = std::fabs(gamma0);
    const float gm = std::fmax(std::fabs(gamma0), 1.0f);
    const float bd = std::sqrt(1.0f - 1.0f / (gm * gm));
    const float beta_drift = std::copysign(bd, gamma0);
    float rannum = park_miller(seed);
    if (rannum == 1.0f) rannum = park_miller(seed);
    float gam = 0.0f;
    for (int i = 1;; ++i) {
        bool done = false;
        if (i == kPdfSz) {
            gam = t.gamma[kPdfSz - 1];
            done = true;
        }
        if (rannum >= t.pdf[i - 1] && rannum < t.pdf[i]) {
            const float dg = t.gamma[i] - t.gamma[i - 1];
            const float dp = t.pdf[i] - t.pdf[i - 1];
            const float slope = dg / dp;
            const float off = rannum - t.pdf[i - 1];
            const float inc = slope * off;
            gam = t.gamma[i - 1] + inc;
            done = true;
        }
        if (done) break;
    }
    float pcosth;
    if (three_d) {
        pcosth = 2.0f * park_miller(seed) - 1.0f;
    } else {
        pcosth = (2.0f * park_miller(seed) - 1.0f) * (float)pcosthmult;
        if (sigma != 0.0f) pcosth = 2.0f * park_miller(seed) - 1.0f;
    }
    const float r2 = park_miller(seed) * 2.0f;
    const float pphi = (float)((double)r2 * pi);
    const float psinth = std::sqrt(1.0f - pcosth * pcosth);
    const float v0a = cd * std::sqrt(gam * (gam + 2.0f));
    const float v0t = v0a / (1.0f + gam);
    const float vs = v0t * psinth;
    const float ut1 = vs * std::cos(pphi);
    const float vt1 = vs * std::sin(pphi);
    const float wt1 = v0t * pcosth;
    float ptx = (1.0f + gam) * ut1;
    const float pty = (1.0f + gam) * vt1;
    const float ptz = (1.0f + gam) * wt1;
    const float x7 = park_miller(seed);
    const float lhs = (-beta_drift * ut1) / cd;
    if (lhs > x7) ptx = -ptx;
    const float cb = cd * beta_drift;
    const float drift = cb * (gam + 1.0f);
    const float px1 = (ptx + drift) * gamma0mag;
    u = px1 / cd;
    v = pty / cd;
    w = ptz / cd;
}

// the reference's clamp of a region [a1, a2] to this domain's interior
// [lo, hi) (code/particles.F90:2640-2700, one domain: no offsets)
void clamp_region(float a1, float a2, int lo, int hi_edge, float& mn, float& mx)
{
    mn = (float)lo;
    mx = (float)lo;
    if (a1 < (float)lo) mn = (float)lo;
    if (a2 < (float)lo) mx = (float)lo;
    if (a1 >= (float)hi_edge) mn = (float)hi_edge;
    if (a2 >= (float)hi_edge) mx = (float)hi_edge;
    if (a1 >= (float)lo && a1 < (float)hi_edge) mn = a1;
    if (a2 >= (float)lo && a2 < (float)hi_edge) mx = a2;
}

}  // namespace

extern "C" int tmpg_inject_plasma_region(const tmpg_plasma_region* r, double* dseed, int64_t cap, int64_t* n,
                                         float* ions, float* lecs)
{
    if (!r || !dseed || !n || (r->dim != 2 && r->dim != 3) || r->order < 1 || r->order > 3 ||
        r->nx < 1 || r->ny < 1 || (r->dim == 3 && r->nz < 1) || !(r->c > 0.0f) || r->direction < 1 ||
        r->direction > 3 || cap < 0)
        return TMPG_EINVAL;
    const int upsamp_e = r->upsamp_e > 0 ? r->upsamp_e : 1;
    const int ng = r->order == 1 ? 5 : 7;
    const int ngz = r->dim == 2 ? 5 : ng;
    const int h = ng / 2, hz = ngz / 2;
    const int mx = r->nx + ng, my = r->ny + ng, mz = r->dim == 3 ? r->nz + ngz : 1;
    const bool three_d = r->dim == 3 || r->pcosthmult != 0;
    // gamma_drift_in below 1 in magnitude is v/c
    float gd = r->gamma_drift;
    if (std::fabs(gd) < 1.0f) gd = std::copysign(std::sqrt(1.0f / (1.0f - gd * gd)), gd);
    MaxwTable ti, te;
    init_maxw_table(r->delgam_i, three_d, ti);
    init_maxw_table(r->delgam_e, three_d, te);
    float minx, maxx, miny, maxy, minz, maxz;
    clamp_region(r->x1, r->x2, h + 1, mx - h, minx, maxx);
    clamp_region(r->y1, r->y2, h + 1, my - h, miny, maxy);
    if (r->dim == 3) {
        clamp_region(r->z1, r->z2, hz + 1, mz - hz, minz, maxz);
    } else {
        minz = (float)(hz + 1);
        maxz = minz + 1.0f;
    }
    const float dx = maxx - minx, dy = maxy - miny, dz = maxz - minz;
    double seed = *dseed;
    const float half = 0.5f * r->ppc;
    float numps = half * dz;
    numps = numps * dy;
    numps = numps * dx;
    if (numps < 10.0f) numps = numps != 0.0f ? poisson(numps, seed) : 0.0f;
    else numps = std::ceil(numps);
    const int64_t np = (int64_t)numps;
    *n = np;
    if (np > cap || (np > 0 && (!ions || !lecs))) return TMPG_EOVERFLOW;   // seed left unchanged
    const int64_t ne = np * upsamp_e;
    if (ne > cap) { *n = ne; return TMPG_EOVERFLOW; }
    auto swap_dir = [&](float& u, float& v, float& w) {
        if (r->direction == 2) std::swap(u, v);
        if (r->direction == 3) std::swap(u, w);
    };
    int64_t le = 0;
    for (int64_t k = 0; k < np; ++k) {
        const float rx = park_miller(seed);
        const float tx = dx * rx;
        const float x = minx + tx;
        const float ry = park_miller(seed);
        const float ty = dy * ry;
        const float y = miny + ty;
        const float rz = park_miller(seed);
        const float tz = dz * rz;
        const float z = minz + tz;
        float u, v, w;
        maxwell_dist(gd, r->c, seed, ti, three_d, r->pcosthmult, r->sigma, u, v, w);
        swap_dir(u, v, w);
        float* o = ions;
        o[0 * cap + k] = x; o[1 * cap + k] = y; o[2 * cap + k] = z;
        o[3 * cap + k] = u; o[4 * cap + k] = v; o[5 * cap + k] = w;
        o[6 * cap + k] = r->weight;
        for (int e = 0; e < upsamp_e; ++e, ++le) {
            maxwell_dist(gd, r->c, seed, te, three_d, r->pcosthmult, r->sigma, u, v, w);
            swap_dir(u, v, w);
            float* q = lecs;
            q[0 * cap + le] = x; q[1 * cap + le] = y; q[2 * cap + le] = z;
            q[3 * cap + le] = u; q[4 * cap + le] = v; q[5 * cap + le] = w;
            q[6 * cap + le] = r->weight / (float)upsamp_e;
        }
    }
    *dseed = seed;
    return TMPG_OK;
}
