// pic_math.h -- per-particle arithmetic of the push and the deposit (FP32).
//
// Positions are an integer cell plus an offset o in [0,1).  A node n of a
// shape window is the cell (particle cell + n).  All weights are the
// reference's B-splines (code/particles_movedeposit.F90:447-468, 709-790,
// 1035-1128; code/particles.F90:738-773, 925-1020, 1172-1277) written in
// local polynomial pieces:
//   order 1: piece X in [m, m+1)     nodes m, m+1        (knots at integers)
//   order 2: piece X in [m-1/2,m+1/2) nodes m-1, m, m+1  (knots at half-integers)
//   order 3: piece X in [m, m+1)     nodes m-1 .. m+2    (knots at integers)
// The Esirkepov deposit needs dS = S(new) - S(old).  Instead of subtracting
// two O(1) numbers, dS is evaluated as (length) x (divided difference) on
// each polynomial piece the move crosses (at most one knot, since
// |displacement| < 1 cell), which keeps its relative error at FP32 round-off
// even for slow particles.
#pragma once
#include <hip/hip_runtime.h>

namespace tmpg {

template <int ORD> struct Win;                 // deposit window per axis
template <> struct Win<1> { static constexpr int W = 3; };
template <> struct Win<2> { static constexpr int W = 4; };
template <> struct Win<3> { static constexpr int W = 5; };

// ---------------------------------------------------------------- weights
// node weights of a particle at local coordinate xi of piece m; writes the
// weights of nodes (m + first .. m + first + cnt - 1) into w[0..cnt-1].
template <int ORD> struct Piece;

template <> struct Piece<1> {
    static constexpr int first = 0, cnt = 2;
    __device__ __forceinline__ static void w(float xi, float* o) {
        o[0] = 1.0f - xi;
        o[1] = xi;
    }
    // dS over a segment [a,b] inside the piece, d = b - a computed accurately
    __device__ __forceinline__ static void dw(float a, float b, float d, float* o) {
        (void)a; (void)b;
        o[0] = -d;
        o[1] = d;
    }
};

template <> struct Piece<2> {
    static constexpr int first = -1, cnt = 3;
    __device__ __forceinline__ static void w(float xi, float* o) {
        float l = 0.5f - xi, r = 0.5f + xi;
        o[0] = 0.5f * l * l;
        o[1] = 0.75f - xi * xi;
        o[2] = 0.5f * r * r;
    }
    __device__ __forceinline__ static void dw(float a, float b, float d, float* o) {
        float s = a + b;
        o[0] = -0.5f * d * (1.0f - s);
        o[1] = -d * s;
        o[2] = 0.5f * d * (1.0f + s);
    }
};

template <> struct Piece<3> {
    static constexpr int first = -1, cnt = 4;
    __device__ __forceinline__ static void w(float xi, float* o) {
        float r = 1.0f - xi;
        const float sixth = 1.0f / 6.0f, twoth = 2.0f / 3.0f;
        o[0] = sixth * r * r * r;
        o[1] = twoth - xi * xi + 0.5f * xi * xi * xi;
        o[2] = twoth - r * r + 0.5f * r * r * r;
        o[3] = sixth * xi * xi * xi;
    }
    __device__ __forceinline__ static void dw(float a, float b, float d, float* o) {
        const float sixth = 1.0f / 6.0f;
        float A = 1.0f - a, B = 1.0f - b;
        float q = a * a + a * b + b * b;
        float Q = A * A + A * B + B * B;
        o[0] = -d * Q * sixth;
        o[1] = d * (0.5f * q - (a + b));
        o[2] = -d * (0.5f * Q - (A + B));
        o[3] = d * q * sixth;
    }
};

// piece containing the local position X (relative to the particle's cell).
// Order 2 pieces are split at half-integers and X = m + 1/2 belongs to the
// lower piece, as in the reference's `d <= half` branch.
template <int ORD> __device__ __forceinline__ int piece_of(float X)
{
    return (ORD == 2) ? (int)ceilf(X - 0.5f) : (int)floorf(X);
}
// the knot between two adjacent pieces ma < mb
template <int ORD> __device__ __forceinline__ float knot_between(int ma, int mb)
{
    int m = ma > mb ? ma : mb;
    return (ORD == 2) ? (float)m - 0.5f : (float)m;
}

// -------------------------------------------------- one axis of a deposit
// Given old offset o in [0,1) and displacement dl (|dl| < 1), fill, on the
// window of Win<ORD>::W nodes starting at node `base` (relative to the old
// cell):  S1 = old weights, dS = new - old weights.  Returns base.
template <int ORD>
__device__ __forceinline__ int axis_deposit(float o, float dl, float* S1, float* dS)
{
    constexpr int W = Win<ORD>::W;
    using P = Piece<ORD>;
    if constexpr (ORD == 3) {
        // branch-free form of the general case below: the old piece is the
        // cell itself (m0 = 0), the new one m1 in {-1, 0, 1}; the window
        // starts at node -2 when m1 = -1, else at -1 (shift = 1 or 0)
        float w[4], t0[4], t1[4];
        P::w(o, w);
        const int m1 = (int)floorf(o + dl);
        const float K = (m1 > 0) ? 1.0f : 0.0f;
        const float b0 = (m1 == 0) ? o + dl : K;
        const float d0 = (m1 == 0) ? dl : K - o;
        P::dw(o, b0, d0, t0);
        const float d2 = dl - d0;              // zero without a crossing
        const float a1 = K - (float)m1;
        P::dw(a1, a1 + d2, d2, t1);
        const bool sh = m1 < 0;
#pragma unroll
        for (int q = 0; q < W; ++q) {
            const float wq = sh ? (q >= 1 ? w[q - 1] : 0.0f) : (q <= 3 ? w[q] : 0.0f);
            const float tq = sh ? (q >= 1 ? t0[q - 1] : 0.0f) : (q <= 3 ? t0[q] : 0.0f);
            // piece m1 = +1 covers nodes 0..3 (slots 1..4 of a window from -1),
            // m1 = -1 nodes -2..1 (slots 0..3 of a window from -2)
            const float uq = sh ? (q <= 3 ? t1[q] : 0.0f) : (q >= 1 ? t1[q - 1] : 0.0f);
            S1[q] = wq;
            dS[q] = tq + uq;
        }
        return sh ? -2 : -1;
    }
    const float X1 = o + dl;
    const int m0 = piece_of<ORD>(o);
    const int m1 = piece_of<ORD>(X1);
    // window: from the first node of the lower piece; W nodes hold both
    // supports because the two pieces are adjacent
    const int base = (m0 < m1 ? m0 : m1) + P::first;
#pragma unroll
    for (int k = 0; k < W; ++k) { S1[k] = 0.0f; dS[k] = 0.0f; }
    float t[P::cnt];
    // old weights
    {
        float xi = o - (float)m0;  // exact: o and m0 within a factor 2 or m0 = 0
        P::w(xi, t);
        const int off = m0 + P::first - base;
#pragma unroll
        for (int k = 0; k < P::cnt; ++k) {
#pragma unroll
            for (int q = 0; q < W; ++q)
                if (q == off + k) S1[q] = t[k];
        }
    }
    if (m0 == m1) {
        float a = o - (float)m0;
        float b = a + dl;
        P::dw(a, b, dl, t);
        const int off = m0 + P::first - base;
#pragma unroll
        for (int k = 0; k < P::cnt; ++k)
#pragma unroll
            for (int q = 0; q < W; ++q)
                if (q == off + k) dS[q] += t[k];
    } else {
        // knot between the two pieces
        const float K = knot_between<ORD>(m0, m1);
        const float d1 = K - o;            // first segment length
        const float d2 = dl - d1;          // second segment length
        {
            float a = o - (float)m0;
            float b = K - (float)m0;
            P::dw(a, b, d1, t);
            const int off = m0 + P::first - base;
#pragma unroll
            for (int k = 0; k < P::cnt; ++k)
#pragma unroll
                for (int q = 0; q < W; ++q)
                    if (q == off + k) dS[q] += t[k];
        }
        {
            float a = K - (float)m1;
            float b = a + d2;
            P::dw(a, b, d2, t);
            const int off = m1 + P::first - base;
#pragma unroll
            for (int k = 0; k < P::cnt; ++k)
#pragma unroll
                for (int q = 0; q < W; ++q)
                    if (q == off + k) dS[q] += t[k];
        }
    }
    return base;
}

// ------------------------------------------------------- gather weights
// weights of a particle at offset X in [0,1) on the 4-slot frame of nodes
// -1..2; returns the piece.  For X in [0,1) the piece is floor(X) = 0 for
// orders 1 and 3 and ceil(X - 1/2) in {0, 1} for order 2, so the placement
// is a fixed slot assignment (orders 1, 3) or one select (order 2).
template <int ORD>
__device__ __forceinline__ int frame_weights(float X, float* S)
{
    using P = Piece<ORD>;
    float t[P::cnt];
    if constexpr (ORD == 2) {
        const bool up = X > 0.5f;          // piece_of<2>(X) == 1
        P::w(X - (up ? 1.0f : 0.0f), t);
        S[0] = up ? 0.0f : t[0];
        S[1] = up ? t[0] : t[1];
        S[2] = up ? t[1] : t[2];
        S[3] = up ? t[2] : 0.0f;
        return up ? 1 : 0;
    } else {
        P::w(X, t);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int k = q - 1 - P::first;     // slot q holds node q - 1 = first + k
            S[q] = (k >= 0 && k < P::cnt) ? t[k] : 0.0f;
        }
        return 0;
    }
}

// Weights of nodes -1..2 relative to cell i for a particle at offset o in
// [0,1), primal (Sp, at the cell nodes) and dual (Sd, relative to cell
// id = i + did, did = 0 if o >= 1/2 else -1, at offset o -+ 1/2), as the
// movers compute them.  For order 2 with `quirk` the primal weights are
// masked to the dual's slot range, as the reference's loops do
// (code/particles_movedeposit.F90:709-832).
template <int ORD>
__device__ __forceinline__ void gather_weights(float o, bool quirk, float* Sp, float* Sd, int& did)
{
    frame_weights<ORD>(o, Sp);
    // dual: id = i if o >= 1/2 else i - 1
    const bool hi = o >= 0.5f;
    did = hi ? 0 : -1;
    const float od = hi ? o - 0.5f : o + 0.5f;
    const int md = frame_weights<ORD>(od, Sd);
    if (ORD == 2 && quirk) {
        // the reference loops over the dual's slots: nodes (md-1 .. md+1)
        // of the 6-slot frame, applied with the primal's cell origin
        if (md == 0) Sp[3] = 0.0f;
        else Sp[0] = 0.0f;
    }
}

// ------------------------------------------------------------------ Boris
// code/particles_movedeposit.F90:1172-1177, 1224-1267.  e, b are the
// gathered fields already scaled: e = 0.5*qm*E, b = 0.5*qm*B/c.  u in units
// of c (gamma*beta).  Returns the displacement in cells.
__device__ __forceinline__ void boris(float c, float cinv, float ex0, float ey0, float ez0,
                                      float bx0, float by0, float bz0,
                                      float& u, float& v, float& w,
                                      float& dx, float& dy, float& dz)
{
    float u0 = c * u + ex0;
    float v0 = c * v + ey0;
    float w0 = c * w + ez0;
    float g = c * __frsqrt_rn(c * c + u0 * u0 + v0 * v0 + w0 * w0);
    bx0 *= g; by0 *= g; bz0 *= g;
    float f = 2.0f / (1.0f + bx0 * bx0 + by0 * by0 + bz0 * bz0);
    float u1 = (u0 + v0 * bz0 - w0 * by0) * f;
    float v1 = (v0 + w0 * bx0 - u0 * bz0) * f;
    float w1 = (w0 + u0 * by0 - v0 * bx0) * f;
    u0 = u0 + v1 * bz0 - w1 * by0 + ex0;
    v0 = v0 + w1 * bx0 - u1 * bz0 + ey0;
    w0 = w0 + u1 * by0 - v1 * bx0 + ez0;
    u = u0 * cinv;
    v = v0 * cinv;
    w = w0 * cinv;
    g = c * __frsqrt_rn(c * c + u0 * u0 + v0 * v0 + w0 * w0);
    dx = u * g * c;
    dy = v * g * c;
    dz = w * g * c;
}

// ------------------------------------------------------------------ Vay
// code/particles_movedeposit.F90:1199-1223 (under -Dvay; the same block in
// every mover), then the common update of :1251-1267.  Same inputs and
// outputs as boris().
__device__ __forceinline__ void vay(float c, float cinv, float ex0, float ey0, float ez0,
                                    float bx0, float by0, float bz0,
                                    float& u, float& v, float& w,
                                    float& dx, float& dy, float& dz)
{
    float g = 1.0f / sqrtf(1.0f + u * u + v * v + w * w);
    const float vx0 = c * u * g, vy0 = c * v * g, vz0 = c * w * g;
    const float u1 = c * u + 2.0f * ex0 + vy0 * bz0 - vz0 * by0;
    const float v1 = c * v + 2.0f * ey0 + vz0 * bx0 - vx0 * bz0;
    const float w1 = c * w + 2.0f * ez0 + vx0 * by0 - vy0 * bx0;
    const float ustar = cinv * (u1 * bx0 + v1 * by0 + w1 * bz0);
    const float b2 = bx0 * bx0 + by0 * by0 + bz0 * bz0;
    const float sig = cinv * cinv * (c * c + u1 * u1 + v1 * v1 + w1 * w1) - b2;
    g = 1.0f / sqrtf(0.5f * (sig + sqrtf(sig * sig + 4.0f * (b2 + ustar * ustar))));
    const float tx = bx0 * g, ty = by0 * g, tz = bz0 * g;
    const float f = 1.0f / (1.0f + tx * tx + ty * ty + tz * tz);
    const float ut = u1 * tx + v1 * ty + w1 * tz;
    const float u0 = f * (u1 + ut * tx + v1 * tz - w1 * ty);
    const float v0 = f * (v1 + ut * ty + w1 * tx - u1 * tz);
    const float w0 = f * (w1 + ut * tz + u1 * ty - v1 * tx);
    u = u0 * cinv;
    v = v0 * cinv;
    w = w0 * cinv;
    g = c * __frsqrt_rn(c * c + u0 * u0 + v0 * v0 + w0 * w0);
    dx = u * g * c;
    dy = v * g * c;
    dz = w * g * c;
}

// advance offset o by dl; returns the cell shift (-1, 0, +1) and the new
// offset in [0,1)
__device__ __forceinline__ int advance_offset(float& o, float dl)
{
    // all three cases evaluated and selected (no divergent branches)
    const float t = o + dl;
    const bool up = t >= 1.0f, down = t < 0.0f;
    const float ru = dl - (1.0f - o);
    const float rd = 1.0f - ((-dl) - o);
    const float vu = ru < 0.0f ? 0.0f : ru;
    const float vd = rd >= 1.0f ? 0.99999994f : rd;
    o = up ? vu : (down ? vd : t);
    return (int)up - (int)down;
}

}  // namespace tmpg
