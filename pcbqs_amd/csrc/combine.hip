// RWG Galerkin combination of triangle-pair moments into the edge inductance matrix L.
//
// Semantics: src/pcbqs/core/static_kernels.hpp:254-310 (static_combine) as called by
// src/pcbqs/static.py:103-128 (edge_inductance):
//   L[m,n] = f s_m s_n sum_{i,j} a_mi a_nj [Mrr(p,q) - v_nj.Mr(p,q) - v_mi.Mr(q,p) + v_mi.v_nj M0(p,q)]
// with p = tri[i][m], q = tri[j][n]. Each entry is evaluated once, from its canonical pair
// (min(m,n), max(m,n)), and stored to both positions, so L is bitwise symmetric. RWG edges are
// numbered in triangle order, so a 64-edge block touches a compact set of triangles and the plane
// gathers of a tile hit L2.
//
// FP32 planes are in each triangle's local frame (first moments about the row/column
// triangle's centroid c): with u = c_p - v, the pair term becomes
//   Mrr'(p,q) + u_nj.Mr'(p,q) + u_mi.Mr'(q,p) + (u_mi.u_nj) M0(p,q),
// every product of the size of the result (no cancellation of |c|^2 terms).
#include "pq_common.h"
#include <type_traits>

namespace pq {

// The combination by 64 x 64 tiles of edges, one canonical tile pair (R <= C) per workgroup,
// both plane orientations read along rows. Entry (m, n), m in R, n in C, splits into
//   A(m, n) = sum_{i,j} a_mi a_nj [Mrr(p,q) - v_nj.Mr(p,q) + (v_mi.v_nj) M0(p,q)]   (planes at [p][q])
//   B(n, m) = sum_{i,j} a_mi a_nj [- v_mi.Mr(q,p)]                                   (planes at [q][p])
// (local frame: + u_nj.Mr(p,q) and + u_mi.Mr(q,p), u = c - v). Pass A runs a wave per row m with
// lanes over n; pass B a wave per row n with lanes over m, so that in both the row triangle is
// wave-uniform and the lanes read along a plane row. Both land in LDS; each entry is then formed
// once, f s_m s_n (A + B), and written to L[m][n] and L[n][m] (bitwise symmetric). The earlier
// row-per-wave kernel read Mr(q,p) with 64 lanes on 64 rows, 31x its algorithmic bytes.
template <class PT, bool LOCAL, class OT>
__global__ __launch_bounds__(256) void combine_tile_kernel(const PT *M0, const PT *Mx, const PT *My, const PT *Mrr,
                                                          long long ld, const double *cent, const int *tri,
                                                          const double *alpha, const double *vertex,
                                                          const double *scale, long long ne, double factor,
                                                          OT *L, long long ldl) {
    // FP64 arithmetic on the FP32 plane values: the two halves of a pair's term (and the weighted
    // sums over an edge's + and - triangles) partly cancel, so they are not rounded to FP32 apart
    using AT = double;
    __shared__ AT tA[64][65], tB[64][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long id = blockIdx.x;
    long long ti = (long long)((sqrt(8.0 * (double)id + 1.0) - 1.0) * 0.5);
    while ((ti + 1) * (ti + 2) / 2 <= id) ++ti;
    while (ti * (ti + 1) / 2 > id) --ti;
    const long long tj = id - ti * (ti + 1) / 2;
    const long long R0 = tj * 64, C0 = ti * 64;       // R0 <= C0
    const bool diag = R0 == C0;
    // one edge's data: triangles, alpha, and v (or u = c - v in the local frame)
    struct Edge {
        int t[2];
        double a[2], x[2], y[2];
    };
    auto load = [&](long long e) {
        Edge d;
        const long long ee = e < ne ? e : ne - 1;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            d.t[i] = tri[i * ne + ee];
            d.a[i] = alpha[i * ne + ee];
            d.x[i] = vertex[2 * (i * ne + ee)];
            d.y[i] = vertex[2 * (i * ne + ee) + 1];
            if (LOCAL) {
                d.x[i] = cent[2 * (size_t)d.t[i]] - d.x[i];
                d.y[i] = cent[2 * (size_t)d.t[i] + 1] - d.y[i];
            }
        }
        return d;
    };
    // pass A: rows m in R (wave w: 16 rows), lanes n in C
    {
        const Edge n = load(C0 + lane);
        for (int k = 0; k < 16; ++k) {
            const int rm = 16 * w + k;
            const Edge m = load(R0 + rm);
            AT s = 0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const size_t pq = (size_t)m.t[i] * ld + n.t[j];
                    if (LOCAL) {
                        s += m.a[i] * n.a[j] * ((double)Mrr[pq] + n.x[j] * (double)Mx[pq] + n.y[j] * (double)My[pq] +
                                                (m.x[i] * n.x[j] + m.y[i] * n.y[j]) * (double)M0[pq]);
                    } else {
                        s += m.a[i] * n.a[j] * ((double)Mrr[pq] - n.x[j] * (double)Mx[pq] - n.y[j] * (double)My[pq] +
                                                (m.x[i] * n.x[j] + m.y[i] * n.y[j]) * (double)M0[pq]);
                    }
                }
            tA[rm][lane] = s;
        }
    }
    // pass B: rows n in C (wave w: 16 rows), lanes m in R
    {
        const Edge m = load(R0 + lane);
        for (int k = 0; k < 16; ++k) {
            const int cn = 16 * w + k;
            const Edge n = load(C0 + cn);
            AT s = 0;
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const size_t qp = (size_t)n.t[j] * ld + m.t[i];
                    if (LOCAL) {
                        s += m.a[i] * n.a[j] * (m.x[i] * (double)Mx[qp] + m.y[i] * (double)My[qp]);
                    } else {
                        s += m.a[i] * n.a[j] * (-m.x[i] * (double)Mx[qp] - m.y[i] * (double)My[qp]);
                    }
                }
            tB[cn][lane] = s;
        }
    }
    __syncthreads();
    // L[m][n] (rows m, lanes n) and L[n][m] (rows n, lanes m), each from the canonical (m, n)
    for (int k = 0; k < 16; ++k) {
        const int rm = 16 * w + k;
        const long long mm = R0 + rm, nn = C0 + lane;
        if (mm < ne && nn < ne && (!diag || lane >= rm)) {
            const AT v = tA[rm][lane] + tB[lane][rm];
            L[mm * ldl + nn] = (OT)(factor * scale[mm] * scale[nn] * (double)v);
        }
    }
    for (int k = 0; k < 16; ++k) {
        const int cn = 16 * w + k;
        const long long nn = C0 + cn, mm = R0 + lane;
        if (mm < ne && nn < ne && (!diag || lane < cn)) {
            const AT v = tA[lane][cn] + tB[cn][lane];
            L[nn * ldl + mm] = (OT)(factor * scale[mm] * scale[nn] * (double)v);
        }
    }
}

// P[a][b] = f / (A_a A_b) sum_{p in node a} sum_{q in node b} M0[p][q]: the potential coefficients
// of the charge nodes (src/pcbqs/mqs.py:532-556, _set_potential: P = W^T M0 W / (4 pi eps0 er) with
// W[t, node(t)] = 1 / area(node)). Node a's triangles are tris[ptr[a] .. ptr[a+1]). Every entry is
// summed in its canonical order (min node outer), so P is bitwise symmetric. One wave per row,
// lanes over 64 consecutive columns (coalesced P stores).
template <class PT, class OT>
__global__ __launch_bounds__(256) void potential_kernel(const PT *__restrict__ M0, long long ld,
                                                       const int *__restrict__ ptr, const int *__restrict__ tris,
                                                       const double *__restrict__ area, long long n, double f,
                                                       OT *__restrict__ P, long long ldp) {
    const int lane = threadIdx.x & 63;
    const long long a = (long long)blockIdx.y * 4 + (threadIdx.x >> 6);
    const long long b = (long long)blockIdx.x * 64 + lane;
    if (a >= n || b >= n) return;
    const long long lo = a < b ? a : b, hi = a < b ? b : a;
    double s = 0;
    for (int i = ptr[lo]; i < ptr[lo + 1]; ++i) {
        const size_t p = (size_t)tris[i] * ld;
        for (int j = ptr[hi]; j < ptr[hi + 1]; ++j) s += (double)M0[p + tris[j]];
    }
    P[a * ldp + b] = (OT)(f * s / (area[a] * area[b]));
}

}  // namespace pq

using namespace pq;

extern "C" int pq_potential_device(int precision, const void *M0, int64_t nt, int64_t ld, int64_t n,
                                   const int32_t *node_ptr, const int32_t *node_tris, const double *node_area,
                                   double factor, int out_precision, void *P, int64_t ldp, void *stream) {
    return guard([&] {
        PQ_REQUIRE(M0 && node_ptr && node_tris && node_area && P, "pq_potential_device: null argument");
        PQ_REQUIRE(precision == 32 || precision == 64, "pq_potential_device: precision must be 32 or 64");
        PQ_REQUIRE(out_precision == 32 || out_precision == 64, "pq_potential_device: out_precision 32 or 64");
        PQ_REQUIRE(nt > 0 && n > 0 && n <= nt, "pq_potential_device: invalid sizes");
        if (ld == 0) ld = nt;
        if (ldp == 0) ldp = n;
        PQ_REQUIRE(ld >= nt && ldp >= n, "pq_potential_device: leading dimension too small");
        dim3 grid((unsigned)((n + 63) / 64), (unsigned)((n + 3) / 4)), block(256);
        hipStream_t s = (hipStream_t)stream;
#define PQ_POT(PT, OT)                                                                                     \
    hipLaunchKernelGGL((potential_kernel<PT, OT>), grid, block, 0, s, (const PT *)M0, (long long)ld, node_ptr, \
                       node_tris, node_area, (long long)n, factor, (OT *)P, (long long)ldp)
        if (precision == 32) {
            if (out_precision == 32) PQ_POT(float, float); else PQ_POT(float, double);
        } else {
            if (out_precision == 32) PQ_POT(double, float); else PQ_POT(double, double);
        }
#undef PQ_POT
        PQ_HIP(hipGetLastError());
    });
}

extern "C" int pq_combine_device(int precision, const void *planes, int64_t nt, int64_t ld, const double *cent,
                                 const int32_t *tri, const double *alpha, const double *vertex,
                                 const double *scale, int64_t ne, double factor, int out_precision, void *L,
                                 int64_t ldl, void *stream) {
    return guard([&] {
        PQ_REQUIRE(planes && tri && alpha && vertex && scale && L, "pq_combine_device: null argument");
        PQ_REQUIRE(precision == 32 || precision == 64, "pq_combine_device: precision must be 32 or 64");
        PQ_REQUIRE(out_precision == 32 || out_precision == 64, "pq_combine_device: out_precision 32 or 64");
        PQ_REQUIRE(precision == 64 || cent, "pq_combine_device: FP32 (local-frame) planes need centroids");
        PQ_REQUIRE(nt > 0 && ne >= 0, "pq_combine_device: invalid sizes");
        if (ld == 0) ld = nt;
        PQ_REQUIRE(ld >= nt, "pq_combine_device: leading dimension below the triangle count");
        if (ldl == 0) ldl = ne;
        PQ_REQUIRE(ldl >= ne, "pq_combine_device: L leading dimension below the edge count");
        if (ne == 0) return;
        hipStream_t s = (hipStream_t)stream;
        const size_t plane = (size_t)nt * ld;
        const long long T = (ne + 63) / 64;
        dim3 grid((unsigned)(T * (T + 1) / 2)), block(256);
        if (precision == 32) {
            const float *P = (const float *)planes;
            if (out_precision == 32)
                hipLaunchKernelGGL((combine_tile_kernel<float, true, float>), grid, block, 0, s, P, P + plane,
                                   P + 2 * plane, P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale,
                                   (long long)ne, factor, (float *)L, (long long)ldl);
            else
                hipLaunchKernelGGL((combine_tile_kernel<float, true, double>), grid, block, 0, s, P, P + plane,
                                   P + 2 * plane, P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale,
                                   (long long)ne, factor, (double *)L, (long long)ldl);
        } else {
            const double *P = (const double *)planes;
            if (out_precision == 32)
                hipLaunchKernelGGL((combine_tile_kernel<double, false, float>), grid, block, 0, s, P, P + plane,
                                   P + 2 * plane, P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale,
                                   (long long)ne, factor, (float *)L, (long long)ldl);
            else
                hipLaunchKernelGGL((combine_tile_kernel<double, false, double>), grid, block, 0, s, P, P + plane,
                                   P + 2 * plane, P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale,
                                   (long long)ne, factor, (double *)L, (long long)ldl);
        }
        PQ_HIP(hipGetLastError());
    });
}
