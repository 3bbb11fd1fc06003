// Device view of a prepared fill geometry and the per-pair FP64 moment functions, shared by the
// fill kernels and the entry kernels of the compressed operators (one copy per translation unit:
// everything here is inline or a template). Semantics: src/pcbqs/core/static_kernels.hpp:53-136
// (static_pair, pair_kind).
#pragma once
#include "pq_common.h"
#include <cmath>

namespace pq {


// Device view of one prepared geometry.
struct GeomDev {
    int nt;
    const double *V, *L, *tau, *un, *area, *P3, *z, *cent, *radius;
    const double *qp[3];
    const double *qw[3];
    int nq[3];
    // FP32 far records, 2 float4 per triangle: (x0,y0,x1,y1), (x2,y2,area,radius) with the far
    // points taken relative to the triangle's own centroid.
    const float4 *far32;
};


// ---- classification ----------------------------------------------------------------------

// The reference rule (static_kernels.hpp:130-136): every product rounded, then summed. It is
// exactly symmetric in its two triangles (differences are squared, sums commute), so entry (r,c)
// and (c,r) classify alike without choosing the canonical order.
static __device__ __noinline__ int exact_kind(double ax, double ay, double az, double arad, double bx, double by,
                                       double bz, double brad, double near, int all_near, int outer) {
#pragma clang fp contract(off)
    const double dx = ax - bx, dy = ay - by;
    const double d = az - bz;
    const double dist = __builtin_sqrt((dx * dx + dy * dy) + d * d);
    const double size = arad + brad;
    if (!all_near && !(dist <= near * size)) return -1;
    return dist <= size ? outer : 0;
}

// ---- per-pair moment functions -----------------------------------------------------------

// 1/sqrt(x) in FP64: the hardware estimate refined by two Newton steps y <- y + y (1 - x y^2) / 2,
// each of which squares the relative error; the result is within a few ulp of the correctly
// rounded value, at about half the instructions of a correctly rounded square root and division.
__device__ __forceinline__ double inv_sqrt64(double x) {
    double y = __builtin_amdgcn_rsq(x);
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        const double e = __builtin_fma(-(x * y), y, 1.0);
        y = __builtin_fma(0.5 * y, e, y);
    }
    return y;
}

// log(a / b) for positive normal a and b without the division a / b and without a general log:
// a = ma 2^ea, b = mb 2^eb with ma, mb in [1/2, 1); one mantissa is doubled so that ma / mb lies in
// [1/sqrt2, sqrt2]; then log(ma / mb) = 2 atanh(s), s = (ma - mb) / (ma + mb) (ma - mb exact by
// Sterbenz's lemma, |s| <= 3 - 2 sqrt2 < 0.172), summed as 2 s (1 + s^2/3 + ... + s^18/19)
// (truncation below 2^-55 relative), plus (ea - eb) ln 2 with ln 2 split into a head whose low bits
// are zero (so (ea - eb) times it is exact) and its tail. Within 2 ulp of log(a / b) over ratios
// e^-40 .. e^40 and near 1; 38 FP64 instructions against 87 for log(a / b).
__device__ __forceinline__ double log_ratio(double a, double b) {
    constexpr double SQRT2 = 1.41421356237309504880;
    constexpr double LN2_HI = 6.93147180369123816490e-01, LN2_LO = 1.90821492927058770002e-10;
    double ma = __builtin_amdgcn_frexp_mant(a), mb = __builtin_amdgcn_frexp_mant(b);
    int e = __builtin_amdgcn_frexp_exp(a) - __builtin_amdgcn_frexp_exp(b);
    if (ma > SQRT2 * mb) { mb *= 2.0; e += 1; }
    else if (SQRT2 * ma < mb) { ma *= 2.0; e -= 1; }
    const double s = (ma - mb) / (ma + mb);
    const double z = s * s;
    double p = 1.0 / 19;
    p = __builtin_fma(p, z, 1.0 / 17);
    p = __builtin_fma(p, z, 1.0 / 15); p = __builtin_fma(p, z, 1.0 / 13); p = __builtin_fma(p, z, 1.0 / 11);
    p = __builtin_fma(p, z, 1.0 / 9); p = __builtin_fma(p, z, 1.0 / 7); p = __builtin_fma(p, z, 1.0 / 5);
    p = __builtin_fma(p, z, 1.0 / 3);
    const double t = 2.0 * s;
    const double r = __builtin_fma(t * z, p, t);
    const double de = (double)e;
    return __builtin_fma(de, LN2_HI, __builtin_fma(de, LN2_LO, r));
}

// FP64 far 3x3 product rule, reference arithmetic order (static_kernels.hpp:57-72); w / sqrt(r^2)
// taken as w * inv_sqrt64(r^2).
template <bool SCALAR>
__device__ __forceinline__ void far64(const double *pa, const double *pb, double w, double d, double *m) {
#pragma unroll
    for (int j = 0; j < 6; ++j) m[j] = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double x = pa[2 * i], y = pa[2 * i + 1];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double sx = pb[2 * j], sy = pb[2 * j + 1];
            const double dx = x - sx, dy = y - sy;
            const double gg = w * inv_sqrt64(dx * dx + dy * dy + d * d);
            m[0] += gg;
            if (!SCALAR) {
                m[1] += gg * x; m[2] += gg * y; m[3] += gg * sx; m[4] += gg * sy;
                m[5] += gg * (x * sx + y * sy);
            }
        }
    }
}

// The closed-form inner integrals (Wilton, Rao et al., IEEE TAP 1984) of 1/R and (r'-rho)/R over
// source triangle q at observation point (x, y), height offset d (static_kernels.hpp:78-97).
// Edge e runs from vertex e to vertex e + 1, so its end-point distances are the vertex distances
// R_e and R_{e+1}: three square roots per point, not six (nor a seventh for r0, whose only use is
// the guard, taken on r0^2).
template <bool SCALAR>
__device__ __forceinline__ void inner_integrals(const GeomDev &g, int q, double x, double y, double d,
                                                double &I1, double &IRx, double &IRy) {
    const double ad = fabs(d), d2 = d * d;
    I1 = 0; IRx = 0; IRy = 0;
    const double *Vq = g.V + 6 * (size_t)q, *Tq = g.tau + 6 * (size_t)q, *Nq = g.un + 6 * (size_t)q;
    const double *Lq = g.L + 3 * (size_t)q;
    double vx[3], vy[3], R[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        vx[k] = Vq[2 * k] - x;
        vy[k] = Vq[2 * k + 1] - y;
        R[k] = sqrt(vx[k] * vx[k] + vy[k] * vy[k] + d2);
    }
#pragma unroll
    for (int e = 0; e < 3; ++e) {
        const int f = e == 2 ? 0 : e + 1;
        const double sm = vx[e] * Tq[2 * e] + vy[e] * Tq[2 * e + 1];
        const double sp = sm + Lq[e];
        const double t0 = vx[e] * Nq[2 * e] + vy[e] * Nq[2 * e + 1];
        const double r02 = t0 * t0 + d2;
        const double rm = R[e], rp = R[f];
        // asinh(s / r0) = log(A(s)) - log(r0) with A(s) = |s| + R for s >= 0 and r0^2 / (|s| + R)
        // for s < 0 (the same number without cancellation); sp > sm, so the difference of the two
        // asinh is log(num / den) with one of three (num, den) pairs
        double lg = 0.0;
        const double g0 = 1e-14 * Lq[e];
        if (r02 > g0 * g0) {
            const double bp = fabs(sp) + rp, bm = fabs(sm) + rm;
            const bool pneg = sp < 0, mneg = sm < 0;
            const double num = pneg ? bm : (mneg ? bp * bm : bp);
            const double den = pneg ? bp : (mneg ? r02 : bm);
            lg = log_ratio(num, den);
        }
        double term = t0 * lg;
        if (ad > 0) {
            // atan2(y1, x1) - atan2(y2, x2) with x1, x2 > 0: both angles lie in (-pi/2, pi/2), so the
            // difference lies in (-pi, pi) and is the angle of (x1 x2 + y1 y2, y1 x2 - y2 x1)
            const double y1 = t0 * sp, x1 = r02 + ad * rp, y2 = t0 * sm, x2 = r02 + ad * rm;
            term -= ad * atan2(y1 * x2 - y2 * x1, x1 * x2 + y1 * y2);
        }
        I1 += term;
        if (!SCALAR) {
            const double er = .5 * (r02 * lg + sp * rp - sm * rm);
            IRx += er * Nq[2 * e]; IRy += er * Nq[2 * e + 1];
        }
    }
}

// One observation point's contribution to the six moments of canonical pair (a, b).
// LOCAL: first moments about each triangle's own centroid (FP32 planes' frame).
template <bool SCALAR, bool LOCAL>
__device__ __forceinline__ void near_point(const GeomDev &g, int a, int b, int level, int j, double *m) {
    const int nq = g.nq[level];
    const double *pt = g.qp[level] + 2 * ((size_t)a * nq + j);
    const double x = pt[0], y = pt[1];
    const double d = g.z[a] - g.z[b];
    double I1, IRx, IRy;
    inner_integrals<SCALAR>(g, b, x, y, d, I1, IRx, IRy);
    const double w = g.area[a] * g.qw[level][j];
    m[0] = w * I1;
    if (SCALAR) return;
    double xo = x, yo = y, xs = x, ys = y;
    if (LOCAL) {
        xo = x - g.cent[2 * a]; yo = y - g.cent[2 * a + 1];
        xs = x - g.cent[2 * b]; ys = y - g.cent[2 * b + 1];
    }
    const double sx = IRx + xs * I1, sy = IRy + ys * I1;
    m[1] = w * xo * I1; m[2] = w * yo * I1;
    m[3] = w * sx; m[4] = w * sy; m[5] = w * (xo * sx + yo * sy);
}

// Full FP64 moments of one ordered pair (reference static_pair), for explicit pair lists.
template <bool SCALAR>
__device__ void pair64(const GeomDev &g, int p, int q, double d, int k, double *m) {
    if (k < 0) {
        far64<SCALAR>(g.P3 + 6 * (size_t)p, g.P3 + 6 * (size_t)q, g.area[p] * g.area[q] / 9, d, m);
        return;
    }
    for (int j = 0; j < 6; ++j) m[j] = 0;
    const int nq = g.nq[k];
    for (int j = 0; j < nq; ++j) {
        const double *pt = g.qp[k] + 2 * ((size_t)p * nq + j);
        const double x = pt[0], y = pt[1];
        double I1, IRx, IRy;
        inner_integrals<SCALAR>(g, q, x, y, d, I1, IRx, IRy);
        const double w = g.area[p] * g.qw[k][j];
        m[0] += w * I1;
        if (!SCALAR) {
            const double sx = IRx + x * I1, sy = IRy + y * I1;
            m[1] += w * x * I1; m[2] += w * y * I1;
            m[3] += w * sx; m[4] += w * sy; m[5] += w * (x * sx + y * sy);
        }
    }
}

}  // namespace pq

// A prepared geometry on one device (pq_fill_create).
struct pq_fill {
    int device = 0;
    int nt = 0;
    pq::GeomDev g{};
    pq::DevBuf geom, far32, lists[2], counts, flag;
    unsigned long long cap = 0;
    hipStream_t stream = nullptr, copy_stream = nullptr;
    hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
    unsigned long long *host_stat = nullptr;   // pinned: the device fill's list counts and flag
    double far_ms = 0, near_ms = 0;
    long long near_pairs[2] = {0, 0};
    int cus = 256;
};
