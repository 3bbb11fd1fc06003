// RWG Galerkin combination of triangle-pair moments into the edge inductance matrix L.
//
// Semantics: src/pcbqs/core/static_kernels.hpp:254-310 (static_combine) as called by
// src/pcbqs/static.py:103-128 (edge_inductance):
//   L[m,n] = f s_m s_n sum_{i,j} a_mi a_nj [Mrr(p,q) - v_nj.Mr(p,q) - v_mi.Mr(q,p) + v_mi.v_nj M0(p,q)]
// with p = tri[i][m], q = tri[j][n]. Each entry is evaluated from its canonical pair
// (min(m,n), max(m,n)) so L is bitwise symmetric. One wave per row of L (row data wave-uniform),
// lanes over 64 consecutive columns (coalesced L stores); plane gathers hit L2 because RWG
// edges are numbered in triangle order, so neighbouring rows/columns touch neighbouring
// triangles.
//
// FP32 planes are in each triangle's local frame (first moments about the row/column
// triangle's centroid c): with u = c_p - v, the pair term becomes
//   Mrr'(p,q) + u_nj.Mr'(p,q) + u_mi.Mr'(q,p) + (u_mi.u_nj) M0(p,q),
// every product of the size of the result (no cancellation of |c|^2 terms).
#include "pq_common.h"

namespace pq {

template <class PT, bool LOCAL, class OT>
__global__ __launch_bounds__(256) void combine_kernel(const PT *M0, const PT *Mx, const PT *My, const PT *Mrr,
                                                     long long ld, const double *cent, const int *tri,
                                                     const double *alpha, const double *vertex,
                                                     const double *scale, long long ne, double factor,
                                                     OT *L, long long ldl) {
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const long long c = (long long)blockIdx.x * 64 + lane;
    const bool cvalid = c < ne;
    const long long cc = cvalid ? c : ne - 1;
    const long long rbase = (long long)blockIdx.y * 16 + wave * 4;
    for (int k = 0; k < 4; ++k) {
        const long long r = rbase + k;
        if (r >= ne) break;
        const long long a = r <= cc ? r : cc, b = r <= cc ? cc : r;
        int ta[2], tb[2];
        double al[2], bl[2], va[2][2], vb[2][2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            ta[i] = tri[i * ne + a]; tb[i] = tri[i * ne + b];
            al[i] = alpha[i * ne + a]; bl[i] = alpha[i * ne + b];
            va[i][0] = vertex[2 * (i * ne + a)]; va[i][1] = vertex[2 * (i * ne + a) + 1];
            vb[i][0] = vertex[2 * (i * ne + b)]; vb[i][1] = vertex[2 * (i * ne + b) + 1];
            if (LOCAL) {
                va[i][0] = cent[2 * (size_t)ta[i]] - va[i][0]; va[i][1] = cent[2 * (size_t)ta[i] + 1] - va[i][1];
                vb[i][0] = cent[2 * (size_t)tb[i]] - vb[i][0]; vb[i][1] = cent[2 * (size_t)tb[i] + 1] - vb[i][1];
            }
        }
        double sum = 0;
        if (LOCAL) {
            float s32 = 0;
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const size_t p = (size_t)ta[i];
                const float umx = (float)va[i][0], umy = (float)va[i][1];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const size_t q = (size_t)tb[j];
                    const float unx = (float)vb[j][0], uny = (float)vb[j][1];
                    const size_t pq = p * ld + q, qp = q * ld + p;
                    const float t = (float)Mrr[pq] + unx * (float)Mx[pq] + uny * (float)My[pq] +
                                    umx * (float)Mx[qp] + umy * (float)My[qp] + (umx * unx + umy * uny) * (float)M0[pq];
                    s32 += (float)(al[i] * bl[j]) * t;
                }
            }
            sum = s32;
        } else {
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                const size_t p = (size_t)ta[i];
                const double vmx = va[i][0], vmy = va[i][1];
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const size_t q = (size_t)tb[j];
                    const double vnx = vb[j][0], vny = vb[j][1];
                    const size_t pq = p * ld + q, qp = q * ld + p;
                    sum += al[i] * bl[j] * ((double)Mrr[pq] - vnx * (double)Mx[pq] - vny * (double)My[pq] -
                                            vmx * (double)Mx[qp] - vmy * (double)My[qp] +
                                            (vmx * vnx + vmy * vny) * (double)M0[pq]);
                }
            }
        }
        if (cvalid) L[r * ldl + c] = (OT)(factor * scale[a] * scale[b] * sum);
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
        dim3 grid((unsigned)((ne + 63) / 64), (unsigned)((ne + 15) / 16)), block(256);
        if (precision == 32) {
            const float *P = (const float *)planes;
            if (out_precision == 32)
                hipLaunchKernelGGL((combine_kernel<float, true, float>), grid, block, 0, s, P, P + plane, P + 2 * plane,
                                   P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale, (long long)ne, factor,
                                   (float *)L, (long long)ldl);
            else
                hipLaunchKernelGGL((combine_kernel<float, true, double>), grid, block, 0, s, P, P + plane,
                                   P + 2 * plane, P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale,
                                   (long long)ne, factor, (double *)L, (long long)ldl);
        } else {
            const double *P = (const double *)planes;
            if (out_precision == 32)
                hipLaunchKernelGGL((combine_kernel<double, false, float>), grid, block, 0, s, P, P + plane,
                                   P + 2 * plane, P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale,
                                   (long long)ne, factor, (float *)L, (long long)ldl);
            else
                hipLaunchKernelGGL((combine_kernel<double, false, double>), grid, block, 0, s, P, P + plane,
                                   P + 2 * plane, P + 3 * plane, (long long)ld, cent, tri, alpha, vertex, scale,
                                   (long long)ne, factor, (double *)L, (long long)ldl);
        }
        PQ_HIP(hipGetLastError());
    });
}
