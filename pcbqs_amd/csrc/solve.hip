// Loop-space port solve on gfx950: projection of the branch inductance onto the loop basis,
// the loop operator Z = Rl + j w Ll applied to many right-hand sides, block-Jacobi, and a
// lockstep restarted GMRES driving them.
//
// Semantics: the dense loop matrices of src/pcbqs/mqs.py:639-665 (Rl = C^T R C, Ll = C^T L C),
// the port elimination of src/pcbqs/mqs_ports.py:52-69 (X = -Z_ff^-1 Z_fs, Zp = Z_ss + Z_sf X)
// solved iteratively as src/pcbqs/mqs_iterative.py:109-170 (_lockstep_gmres: right-preconditioned,
// restarted, every step one operator product for all columns) with the true-residual acceptance
// of src/pcbqs/peec_iterative.py:100-112.
//
// Layout: a block vector of P complex columns is an (n, 2P) row-major FP64 array, real parts in
// columns 0..P-1 and imaginary parts in P..2P-1. Ll is FP32 (n, ldl) row-major; every product
// accumulates in FP64. Rl is CSR (FP64 values). Reductions are two-stage with a fixed order, so
// results do not depend on scheduling.
#include "pq_common.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <vector>

namespace pq {

// ---- projection: Ll = C^T Lb C ------------------------------------------------------------

// U[k, j] = sum_t s_t Lb[k, row_t] over the branches t of loop j (C in CSC: colptr, rows, signs).
__global__ __launch_bounds__(256) void proj_left_kernel(const float *Lb, long long ldb, int nbr,
                                                       const int *colptr, const int *rows, const float *sign,
                                                       int nl, float *U, long long ldu) {
    const int k = blockIdx.x;
    const float *Lrow = Lb + (size_t)k * ldb;
    for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        double acc = 0;
        for (int t = colptr[j]; t < colptr[j + 1]; ++t) {
            const int r = rows[t];
            if (r < nbr) acc += (double)sign[t] * (double)Lrow[r];
        }
        U[(size_t)k * ldu + j] = (float)acc;
    }
}

// Ll[i, j] = sum_t s_t U[row_t, j] over the branches t of loop i.
__global__ __launch_bounds__(256) void proj_right_kernel(const float *U, long long ldu, int nbr, const int *colptr,
                                                        const int *rows, const float *sign, int nl, float *Ll,
                                                        long long ldl) {
    const int i = blockIdx.x;
    const int t0 = colptr[i], t1 = colptr[i + 1];
    for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        double acc = 0;
        for (int t = t0; t < t1; ++t) {
            const int r = rows[t];
            if (r < nbr) acc += (double)sign[t] * (double)U[(size_t)r * ldu + j];
        }
        Ll[(size_t)i * ldl + j] = (float)acc;
    }
}

// ---- operator products --------------------------------------------------------------------

// W = Rl X on rows [0, n) restricted to columns < n (CSR rows of the full loop matrix); NR = 2P.
template <int NR>
__global__ __launch_bounds__(256) void spmv_kernel(const int *ptr, const int *col, const double *val, int n,
                                                  const double *X, double *W) {
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    for (int t = ptr[row] + lane; t < ptr[row + 1]; t += 64) {
        const int c = col[t];
        if (c >= n) continue;
        const double v = val[t];
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += v * X[(size_t)c * NR + r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        double s = acc[r];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
        if (lane == 0) W[(size_t)row * NR + r] = s;
    }
}

// Y = Ll[0:n, 0:n] X (FP32 matrix, FP64 accumulation), then W_re -= w Y_im, W_im += w Y_re.
// A block holds RPB rows; the X chunk of 256 columns is staged in LDS and shared by them.
template <int NR>
__global__ __launch_bounds__(256) void gemv_kernel(const float *Ll, long long ldl, int n, const double *X,
                                                  double omega, double *W) {
    constexpr int P = NR / 2;
    constexpr int RPW = NR >= 16 ? 1 : 16 / NR;   // rows per wave
    constexpr int RPB = 4 * RPW;
    __shared__ double xs[256 * NR];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = blockIdx.x * RPB + wave * RPW;
    double acc[RPW][NR];
#pragma unroll
    for (int a = 0; a < RPW; ++a)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[a][r] = 0;
    for (int c0 = 0; c0 < n; c0 += 256) {
        const int cn = min(256, n - c0);
        __syncthreads();
        for (int e = threadIdx.x; e < cn * NR; e += 256) xs[e] = X[(size_t)c0 * NR + e];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int cl = q * 64 + lane;
            if (cl >= cn) continue;
            double xv[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) xv[r] = xs[cl * NR + r];
#pragma unroll
            for (int a = 0; a < RPW; ++a) {
                const int row = row0 + a;
                const double m = row < n ? (double)Ll[(size_t)row * ldl + c0 + cl] : 0.0;
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[a][r] += m * xv[r];
            }
        }
    }
#pragma unroll
    for (int a = 0; a < RPW; ++a) {
        const int row = row0 + a;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            double s = acc[a][r];
            for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off);
            acc[a][r] = s;
        }
        if (lane == 0 && row < n) {
#pragma unroll
            for (int p = 0; p < P; ++p) {
                W[(size_t)row * NR + p] -= omega * acc[a][P + p];
                W[(size_t)row * NR + P + p] += omega * acc[a][p];
            }
        }
    }
}

// ---- block-Jacobi --------------------------------------------------------------------------

// Inverse of each diagonal block Z_bb = Rl_bb + j w Ll_bb (size <= BS) by in-place Gauss-Jordan
// with partial pivoting in LDS, complex FP32 (a preconditioner); stored as (nblk, BS, BS)
// complex (re, im interleaved). A padded block is the identity.
constexpr int BS = 64;

__global__ __launch_bounds__(256) void bj_setup_kernel(const float *Ll, long long ldl, const int *ptr,
                                                      const int *col, const double *val, int n, double omega,
                                                      float *inv, int *singular) {
    __shared__ float ar[BS][BS + 1], ai[BS][BS + 1];
    __shared__ int perm[BS];
    __shared__ int piv;
    const int b = blockIdx.x, r0 = b * BS, m = min(BS, n - r0);
    for (int e = threadIdx.x; e < BS * BS; e += blockDim.x) {
        const int i = e / BS, j = e % BS;
        float re = 0, im = 0;
        if (i < m && j < m) im = (float)(omega * (double)Ll[(size_t)(r0 + i) * ldl + r0 + j]);
        else if (i == j) re = 1;
        ar[i][j] = re; ai[i][j] = im;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x)
        for (int t = ptr[r0 + i]; t < ptr[r0 + i + 1]; ++t) {
            const int c = col[t] - r0;
            if (c >= 0 && c < m) ar[i][c] += (float)val[t];
        }
    __syncthreads();
    for (int k = 0; k < BS; ++k) {
        if (threadIdx.x == 0) {
            int best = k;
            float bv = ar[k][k] * ar[k][k] + ai[k][k] * ai[k][k];
            for (int i = k + 1; i < BS; ++i) {
                const float v = ar[i][k] * ar[i][k] + ai[i][k] * ai[i][k];
                if (v > bv) { bv = v; best = i; }
            }
            piv = best;
            perm[k] = best;
            if (!(bv > 0)) atomicOr(singular, 1);
        }
        __syncthreads();
        const int p = piv;
        if (p != k)
            for (int j = threadIdx.x; j < BS; j += blockDim.x) {
                const float tr = ar[k][j], ti = ai[k][j];
                ar[k][j] = ar[p][j]; ai[k][j] = ai[p][j];
                ar[p][j] = tr; ai[p][j] = ti;
            }
        __syncthreads();
        const float dr = ar[k][k], di = ai[k][k], dd = dr * dr + di * di;
        const float ir = dd > 0 ? dr / dd : 0.f, ii = dd > 0 ? -di / dd : 0.f;   // 1 / a_kk
        __syncthreads();
        if (threadIdx.x == 0) { ar[k][k] = 1; ai[k][k] = 0; }
        __syncthreads();
        for (int j = threadIdx.x; j < BS; j += blockDim.x) {
            const float xr = ar[k][j], xi = ai[k][j];
            ar[k][j] = xr * ir - xi * ii;
            ai[k][j] = xr * ii + xi * ir;
        }
        __syncthreads();
        // column k of the other rows: the factors, then a_ik := 0 before the row update
        __shared__ float fr_[BS], fi_[BS];
        for (int i = threadIdx.x; i < BS; i += blockDim.x) {
            fr_[i] = ar[i][k]; fi_[i] = ai[i][k];
            if (i != k) { ar[i][k] = 0; ai[i][k] = 0; }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < BS * BS; e += blockDim.x) {
            const int i = e / BS, j = e % BS;
            if (i == k) continue;
            const float fr = fr_[i], fi = fi_[i];
            const float xr = ar[k][j], xi = ai[k][j];
            ar[i][j] -= fr * xr - fi * xi;
            ai[i][j] -= fr * xi + fi * xr;
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
    for (int k = BS - 1; k >= 0; --k) {
        const int p = perm[k];
        if (p != k)
            for (int i = threadIdx.x; i < BS; i += blockDim.x) {
                const float tr = ar[i][k], ti = ai[i][k];
                ar[i][k] = ar[i][p]; ai[i][k] = ai[i][p];
                ar[i][p] = tr; ai[i][p] = ti;
            }
        __syncthreads();
    }
    float *out = inv + (size_t)b * BS * BS * 2;
    for (int e = threadIdx.x; e < BS * BS; e += blockDim.x) {
        const int i = e / BS, j = e % BS;
        out[2 * e] = ar[i][j];
        out[2 * e + 1] = ai[i][j];
    }
}

// Y_b = Zinv_b X_b per block (complex), NR = 2P.
template <int NR>
__global__ __launch_bounds__(256) void bj_apply_kernel(const float *inv, int n, const double *X, double *Y) {
    constexpr int P = NR / 2;
    __shared__ double xs[BS * NR];
    const int b = blockIdx.x, r0 = b * BS, m = min(BS, n - r0);
    for (int e = threadIdx.x; e < BS * NR; e += blockDim.x) xs[e] = (e / NR) < m ? X[(size_t)r0 * NR + e] : 0.0;
    __syncthreads();
    const float *Zi = inv + (size_t)b * BS * BS * 2;
    for (int e = threadIdx.x; e < m * P; e += blockDim.x) {
        const int i = e / P, p = e % P;
        double yr = 0, yi = 0;
        for (int j = 0; j < m; ++j) {
            const double zr = (double)Zi[2 * (i * BS + j)], zi = (double)Zi[2 * (i * BS + j) + 1];
            const double xr = xs[j * NR + p], xi = xs[j * NR + P + p];
            yr += zr * xr - zi * xi;
            yi += zr * xi + zi * xr;
        }
        Y[(size_t)(r0 + i) * NR + p] = yr;
        Y[(size_t)(r0 + i) * NR + P + p] = yi;
    }
}

// ---- Krylov vector kernels -----------------------------------------------------------------

constexpr int RED_BLOCKS = 256;

// partial[blk][i][p] = sum over the block's rows of conj(V_i[n,p]) W[n,p], i < nv (complex).
__global__ __launch_bounds__(256) void dots_partial_kernel(const double *V, size_t vstride, int nv, const double *W,
                                                          int n, int P, double *partial) {
    __shared__ double sr[256], si[256];
    const int NR = 2 * P;
    const int i = blockIdx.y;
    const double *Vi = V + (size_t)i * vstride;
    for (int p = 0; p < P; ++p) {
        double re = 0, im = 0;
        for (int row = blockIdx.x * 256 + threadIdx.x; row < n; row += gridDim.x * 256) {
            const double vr = Vi[(size_t)row * NR + p], vi = Vi[(size_t)row * NR + P + p];
            const double wr = W[(size_t)row * NR + p], wi = W[(size_t)row * NR + P + p];
            re += vr * wr + vi * wi;
            im += vr * wi - vi * wr;
        }
        sr[threadIdx.x] = re; si[threadIdx.x] = im;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) { sr[threadIdx.x] += sr[threadIdx.x + s]; si[threadIdx.x] += si[threadIdx.x + s]; }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            double *o = partial + ((size_t)blockIdx.x * nv + i) * P * 2 + 2 * p;
            o[0] = sr[0]; o[1] = si[0];
        }
        __syncthreads();
    }
}

// h[i][p] = sum over blocks of the partials, in block order.
__global__ void dots_reduce_kernel(const double *partial, int nblk, int nv, int P, double *h) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nv * P * 2) return;
    double s = 0;
    for (int b = 0; b < nblk; ++b) s += partial[(size_t)b * nv * P * 2 + e];
    h[e] = s;
}

// W[n,p] -= sum_i V_i[n,p] h[i][p] (complex); with sign -1 it adds.
__global__ __launch_bounds__(256) void vec_update_kernel(const double *V, size_t vstride, int nv, const double *h,
                                                        double sign, int n, int P, double *W) {
    const int NR = 2 * P;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (size_t)n * P) return;
    const int row = (int)(e / P), p = (int)(e % P);
    double ar = 0, ai = 0;
    for (int i = 0; i < nv; ++i) {
        const double hr = h[(i * P + p) * 2], hi = h[(i * P + p) * 2 + 1];
        const double vr = V[(size_t)i * vstride + (size_t)row * NR + p];
        const double vi = V[(size_t)i * vstride + (size_t)row * NR + P + p];
        ar += vr * hr - vi * hi;
        ai += vr * hi + vi * hr;
    }
    W[(size_t)row * NR + p] -= sign * ar;
    W[(size_t)row * NR + P + p] -= sign * ai;
}

// Y[n,p] = X[n,p] * s[p] (real per-column scale).
__global__ void vec_scale_kernel(const double *X, const double *s, int n, int P, double *Y) {
    const int NR = 2 * P;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * NR) return;
    const int c = (int)(e % NR);
    Y[e] = X[e] * s[c % P];
}

// Y = X - Y (elementwise)
__global__ void vec_sub_kernel(const double *X, size_t count, double *Y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) Y[e] = X[e] - Y[e];
}

// B[:, p] = -(Rl[0:n, sc_p] + j w Ll[0:n, sc_p]) (the port loops' columns, free rows).
__global__ void rhs_kernel(const float *Ll, long long ldl, const int *ptr, const int *col, const double *val,
                           int n, const int *sc, int P, double omega, double *B) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * P) return;
    const int row = e / P, p = e % P;
    double re = 0;
    for (int t = ptr[row]; t < ptr[row + 1]; ++t)
        if (col[t] == sc[p]) re += val[t];
    B[(size_t)row * 2 * P + p] = -re;
    B[(size_t)row * 2 * P + P + p] = -omega * (double)Ll[(size_t)row * ldl + sc[p]];
}

// Zsf X: out[s][q] = sum_i (Rl[sc_s, i] + j w Ll[sc_s, i]) X[i, q] over free i < n (complex).
__global__ __launch_bounds__(256) void port_reduce_kernel(const float *Ll, long long ldl, const int *ptr,
                                                         const int *col, const double *val, int n, const int *sc,
                                                         int P, double omega, const double *X, double *out) {
    __shared__ double sr[256], si[256];
    const int s = blockIdx.x / P, q = blockIdx.x % P;
    const int NR = 2 * P;
    const int row = sc[s];
    double re = 0, im = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double l = omega * (double)Ll[(size_t)row * ldl + i];
        const double xr = X[(size_t)i * NR + q], xi = X[(size_t)i * NR + P + q];
        re += -l * xi;
        im += l * xr;
    }
    for (int t = ptr[row] + threadIdx.x; t < ptr[row + 1]; t += 256) {
        const int c = col[t];
        if (c >= n) continue;
        re += val[t] * X[(size_t)c * NR + q];
        im += val[t] * X[(size_t)c * NR + P + q];
    }
    sr[threadIdx.x] = re; si[threadIdx.x] = im;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) { sr[threadIdx.x] += sr[threadIdx.x + st]; si[threadIdx.x] += si[threadIdx.x + st]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[2 * blockIdx.x] = sr[0]; out[2 * blockIdx.x + 1] = si[0]; }
}

}  // namespace pq

using namespace pq;

// ---- the loop system handle ---------------------------------------------------------------

struct pq_loops {
    int device = 0;
    int nl = 0, nf = 0, P = 0;
    long long ldl = 0;
    DevBuf Ll, rptr, rcol, rval, sc, inv, singular;
    DevBuf V, W, Z, X, B, R, partial, h, scal;
    hipStream_t stream = nullptr;
    int nblk = 0;
    double omega_bj = -1;
    // the last solve
    int iterations = 0, restarts = 0;
    double residual = 0, seconds = 0, matvec_ms = 0;
    long long matvecs = 0;
};

namespace {

template <class F> void dispatch_nr(int NR, F &&f) {
    switch (NR) {
    case 2: f(std::integral_constant<int, 2>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 6: f(std::integral_constant<int, 6>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    case 16: f(std::integral_constant<int, 16>()); break;
    case 32: f(std::integral_constant<int, 32>()); break;
    default: throw Error(PQ_EINVAL, "port count must be 1, 2, 3, 4, 8 or 16");
    }
}

// W = (Rl + j w Ll) X on the free block.
void apply_operator(pq_loops *L, double omega, const double *X, double *W, int P) {
    const int NR = 2 * P, n = L->nf;
    dispatch_nr(NR, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        hipLaunchKernelGGL(spmv_kernel<N>, dim3((n + 3) / 4), dim3(256), 0, L->stream, L->rptr.as<int>(),
                           L->rcol.as<int>(), L->rval.as<double>(), n, X, W);
        constexpr int RPB = 4 * (N >= 16 ? 1 : 16 / N);
        hipLaunchKernelGGL(gemv_kernel<N>, dim3((n + RPB - 1) / RPB), dim3(256), 0, L->stream, L->Ll.as<float>(),
                           L->ldl, n, X, omega, W);
    });
    PQ_HIP(hipGetLastError());
    ++L->matvecs;
}

void apply_precond(pq_loops *L, int kind, const double *X, double *Y, int P) {
    const int n = L->nf, NR = 2 * P;
    if (kind == 0) {
        PQ_HIP(hipMemcpyAsync(Y, X, (size_t)n * NR * 8, hipMemcpyDeviceToDevice, L->stream));
        return;
    }
    dispatch_nr(NR, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        hipLaunchKernelGGL(bj_apply_kernel<N>, dim3(L->nblk), dim3(256), 0, L->stream, L->inv.as<float>(), n, X, Y);
    });
    PQ_HIP(hipGetLastError());
}

// h[i][p] for i < nv: conj(V_i) . W, synchronously copied to host.
void dots(pq_loops *L, const double *V, size_t vstride, int nv, const double *W, int P, double *h_host) {
    const int nblk = std::min(RED_BLOCKS, (L->nf + 255) / 256);
    hipLaunchKernelGGL(dots_partial_kernel, dim3(nblk, nv), dim3(256), 0, L->stream, V, vstride, nv, W, L->nf, P,
                       L->partial.as<double>());
    hipLaunchKernelGGL(dots_reduce_kernel, dim3((nv * P * 2 + 255) / 256), dim3(256), 0, L->stream,
                       L->partial.as<double>(), nblk, nv, P, L->h.as<double>());
    PQ_HIP(hipGetLastError());
    if (h_host) {
        PQ_HIP(hipMemcpyAsync(h_host, L->h.p, (size_t)nv * P * 2 * 8, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
    }
}

void norms(pq_loops *L, const double *W, int P, std::vector<double> &out) {
    std::vector<double> h(2 * P);
    dots(L, W, 0, 1, W, P, h.data());
    out.resize(P);
    for (int p = 0; p < P; ++p) out[p] = std::sqrt(std::max(h[2 * p], 0.0));
}

void scale_cols(pq_loops *L, const double *X, const std::vector<double> &s, int P, double *Y) {
    PQ_HIP(hipMemcpyAsync(L->scal.p, s.data(), P * 8, hipMemcpyHostToDevice, L->stream));
    const size_t count = (size_t)L->nf * 2 * P;
    hipLaunchKernelGGL(vec_scale_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream, X,
                       L->scal.as<double>(), L->nf, P, Y);
    PQ_HIP(hipGetLastError());
}

void update(pq_loops *L, const double *V, size_t vstride, int nv, const double *h_host, double sign, int P,
            double *W) {
    PQ_HIP(hipMemcpyAsync(L->h.p, h_host, (size_t)nv * P * 2 * 8, hipMemcpyHostToDevice, L->stream));
    const size_t count = (size_t)L->nf * P;
    hipLaunchKernelGGL(vec_update_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream, V, vstride,
                       nv, L->h.as<double>(), sign, L->nf, P, W);
    PQ_HIP(hipGetLastError());
}

using cplx = std::complex<double>;

}  // namespace

extern "C" {

int pq_loops_create(int device, int n_loops, int n_free, const int32_t *c_colptr, const int32_t *c_rows,
                    const float *c_sign, int64_t c_nnz, const float *Lb_dev, int64_t ldb, int n_branch_rows,
                    const int32_t *r_ptr, const int32_t *r_col, const double *r_val, pq_loops **out) {
    return guard([&] {
        PQ_REQUIRE(out && c_colptr && c_rows && c_sign && Lb_dev && r_ptr && r_col && r_val,
                   "pq_loops_create: null argument");
        PQ_REQUIRE(n_loops > 0 && n_free >= 0 && n_free <= n_loops && n_branch_rows > 0 && ldb >= n_branch_rows,
                   "pq_loops_create: invalid sizes");
        PQ_REQUIRE(c_colptr[0] == 0 && c_colptr[n_loops] == c_nnz, "pq_loops_create: inconsistent loop basis");
        for (int64_t t = 0; t < c_nnz; ++t) PQ_REQUIRE(c_rows[t] >= 0, "pq_loops_create: negative branch index");
        PQ_REQUIRE(r_ptr[0] == 0, "pq_loops_create: invalid resistance rows");
        for (int i = 0; i < n_loops; ++i)
            PQ_REQUIRE(r_ptr[i + 1] >= r_ptr[i], "pq_loops_create: resistance rows out of order");
        for (int32_t t = 0; t < r_ptr[n_loops]; ++t)
            PQ_REQUIRE(r_col[t] >= 0 && r_col[t] < n_loops, "pq_loops_create: resistance column out of range");
        DeviceScope scope(device);
        auto *L = new pq_loops();
        try {
            L->device = device; L->nl = n_loops; L->nf = n_free;
            L->ldl = (n_loops + 63) / 64 * 64;
            PQ_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
            DevBuf cp, cr, cs, U;
            cp.reserve((n_loops + 1) * 4); cr.reserve(std::max<int64_t>(c_nnz, 1) * 4); cs.reserve(std::max<int64_t>(c_nnz, 1) * 4);
            PQ_HIP(hipMemcpy(cp.p, c_colptr, (n_loops + 1) * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(cr.p, c_rows, c_nnz * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(cs.p, c_sign, c_nnz * 4, hipMemcpyHostToDevice));
            const long long ldu = L->ldl;
            U.reserve((size_t)n_branch_rows * ldu * 4);
            L->Ll.reserve((size_t)n_loops * L->ldl * 4);
            hipLaunchKernelGGL(proj_left_kernel, dim3(n_branch_rows), dim3(256), 0, L->stream, Lb_dev, (long long)ldb,
                               n_branch_rows, cp.as<int>(), cr.as<int>(), cs.as<float>(), n_loops, U.as<float>(), ldu);
            hipLaunchKernelGGL(proj_right_kernel, dim3(n_loops), dim3(256), 0, L->stream, U.as<float>(), ldu,
                               n_branch_rows, cp.as<int>(), cr.as<int>(), cs.as<float>(), n_loops, L->Ll.as<float>(),
                               L->ldl);
            PQ_HIP(hipGetLastError());
            const int rnnz = r_ptr[n_loops];
            L->rptr.reserve((n_loops + 1) * 4); L->rcol.reserve(std::max(rnnz, 1) * 4); L->rval.reserve(std::max(rnnz, 1) * 8);
            PQ_HIP(hipMemcpyAsync(L->rptr.p, r_ptr, (n_loops + 1) * 4, hipMemcpyHostToDevice, L->stream));
            PQ_HIP(hipMemcpyAsync(L->rcol.p, r_col, (size_t)rnnz * 4, hipMemcpyHostToDevice, L->stream));
            PQ_HIP(hipMemcpyAsync(L->rval.p, r_val, (size_t)rnnz * 8, hipMemcpyHostToDevice, L->stream));
            PQ_HIP(hipStreamSynchronize(L->stream));
            L->singular.reserve(4);
        } catch (...) {
            if (L->stream) (void)hipStreamDestroy(L->stream);
            delete L;
            throw;
        }
        *out = L;
    });
}

int pq_loops_destroy(pq_loops *L) {
    if (!L) return PQ_OK;
    return guard([&] {
        DeviceScope scope(L->device);
        if (L->stream) (void)hipStreamSynchronize(L->stream);
        for (DevBuf *b : {&L->Ll, &L->rptr, &L->rcol, &L->rval, &L->sc, &L->inv, &L->singular, &L->V, &L->W, &L->Z,
                          &L->X, &L->B, &L->R, &L->partial, &L->h, &L->scal})
            b->release();
        if (L->stream) (void)hipStreamDestroy(L->stream);
        delete L;
    });
}

int pq_loops_matrix(pq_loops *L, void **Ll_dev, int64_t *ldl) {
    return guard([&] {
        PQ_REQUIRE(L, "pq_loops_matrix: null handle");
        if (Ll_dev) *Ll_dev = L->Ll.p;
        if (ldl) *ldl = L->ldl;
    });
}

// Port impedances at one angular frequency: X = -Z_ff^-1 Z_fs by GMRES, Zp = Z_ss + Z_sf X.
int pq_loops_solve_ports(pq_loops *L, double omega, int P, const int32_t *port_cols, int precond, double tol,
                         int restart, int maxiter, double *Zp, double *X_host, int64_t *iterations,
                         double *residual) {
    return guard([&] {
        PQ_REQUIRE(L && port_cols && Zp, "pq_loops_solve_ports: null argument");
        PQ_REQUIRE(P >= 1 && tol > 0 && tol < 1 && restart >= 1 && maxiter >= 1, "pq_loops_solve_ports: invalid options");
        PQ_REQUIRE(precond == 0 || precond == 1, "pq_loops_solve_ports: preconditioner 0 (none) or 1 (block Jacobi)");
        for (int p = 0; p < P; ++p)
            PQ_REQUIRE(port_cols[p] >= L->nf && port_cols[p] < L->nl, "pq_loops_solve_ports: port loop out of range");
        DeviceScope scope(L->device);
        const auto started = std::chrono::steady_clock::now();
        const int n = L->nf, NR = 2 * P, m = std::min(restart, std::max(n, 1));
        const size_t vsz = (size_t)n * NR;
        L->matvecs = 0;
        L->sc.reserve(P * 4);
        PQ_HIP(hipMemcpyAsync(L->sc.p, port_cols, P * 4, hipMemcpyHostToDevice, L->stream));
        L->V.reserve((m + 1) * vsz * 8);
        L->W.reserve(vsz * 8); L->Z.reserve(vsz * 8); L->X.reserve(vsz * 8); L->B.reserve(vsz * 8); L->R.reserve(vsz * 8);
        L->partial.reserve((size_t)RED_BLOCKS * (m + 2) * P * 2 * 8);
        L->h.reserve((size_t)(m + 2) * P * 2 * 8);
        L->scal.reserve(P * 8);
        double *V = L->V.as<double>(), *W = L->W.as<double>(), *Z = L->Z.as<double>(), *X = L->X.as<double>();
        double *B = L->B.as<double>(), *R = L->R.as<double>();
        if (n > 0) {
            hipLaunchKernelGGL(rhs_kernel, dim3((n * P + 255) / 256), dim3(256), 0, L->stream, L->Ll.as<float>(), L->ldl,
                               L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, L->sc.as<int>(), P, omega, B);
            PQ_HIP(hipGetLastError());
            if (precond == 1 && L->omega_bj != omega) {
                L->nblk = (n + BS - 1) / BS;
                L->inv.reserve((size_t)L->nblk * BS * BS * 2 * 4);
                PQ_HIP(hipMemsetAsync(L->singular.p, 0, 4, L->stream));
                hipLaunchKernelGGL(bj_setup_kernel, dim3(L->nblk), dim3(256), 0, L->stream, L->Ll.as<float>(), L->ldl,
                                   L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, omega,
                                   L->inv.as<float>(), L->singular.as<int>());
                PQ_HIP(hipGetLastError());
                int sing = 0;
                PQ_HIP(hipMemcpyAsync(&sing, L->singular.p, 4, hipMemcpyDeviceToHost, L->stream));
                PQ_HIP(hipStreamSynchronize(L->stream));
                PQ_REQUIRE(!sing, "block-Jacobi: a singular diagonal block");
                L->omega_bj = omega;
            }
            PQ_HIP(hipMemsetAsync(X, 0, vsz * 8, L->stream));
            std::vector<double> bnorm;
            norms(L, B, P, bnorm);
            for (auto &b : bnorm) b = std::max(b, 1e-300);
            // R = B (x0 = 0)
            PQ_HIP(hipMemcpyAsync(R, B, vsz * 8, hipMemcpyDeviceToDevice, L->stream));
            int its = 0, restarts = 0;
            std::vector<double> res(P, 1.0);
            std::vector<cplx> H((size_t)(m + 1) * m * P), g((size_t)(m + 1) * P), cs((size_t)m * P), sn((size_t)m * P);
            std::vector<double> hh((size_t)(m + 2) * P * 2), hh2((size_t)(m + 2) * P * 2);
            while (true) {
                std::vector<double> beta;
                norms(L, R, P, beta);
                bool done = true;
                for (int p = 0; p < P; ++p) { res[p] = beta[p] / bnorm[p]; done = done && res[p] <= tol; }
                if (done || its >= maxiter) break;
                std::vector<double> inv_beta(P);
                for (int p = 0; p < P; ++p) inv_beta[p] = beta[p] > 0 ? 1.0 / beta[p] : 0.0;
                scale_cols(L, R, inv_beta, P, V);
                std::fill(g.begin(), g.end(), cplx(0));
                for (int p = 0; p < P; ++p) g[p] = beta[p];
                int j = 0;
                std::vector<bool> conv(P, false);
                for (; j < m && its < maxiter; ++j, ++its) {
                    apply_precond(L, precond, V + j * vsz, Z, P);
                    apply_operator(L, omega, Z, W, P);   // W = Rl Z first, then += j w Ll Z
                    // classical Gram-Schmidt, twice
                    dots(L, V, vsz, j + 1, W, P, hh.data());
                    update(L, V, vsz, j + 1, hh.data(), 1.0, P, W);
                    dots(L, V, vsz, j + 1, W, P, hh2.data());
                    update(L, V, vsz, j + 1, hh2.data(), 1.0, P, W);
                    std::vector<double> hn;
                    norms(L, W, P, hn);
                    std::vector<double> inv(P);
                    for (int p = 0; p < P; ++p) inv[p] = hn[p] > 0 ? 1.0 / hn[p] : 0.0;
                    scale_cols(L, W, inv, P, V + (j + 1) * vsz);
                    bool all = true;
                    for (int p = 0; p < P; ++p) {
                        auto Hc = [&](int i) -> cplx & { return H[((size_t)i * m + j) * P + p]; };
                        for (int i = 0; i <= j; ++i)
                            Hc(i) = cplx(hh[(i * P + p) * 2] + hh2[(i * P + p) * 2],
                                         hh[(i * P + p) * 2 + 1] + hh2[(i * P + p) * 2 + 1]);
                        Hc(j + 1) = hn[p];
                        for (int i = 0; i < j; ++i) {
                            const cplx a = Hc(i), b = Hc(i + 1);
                            Hc(i) = std::conj(cs[i * P + p]) * a + std::conj(sn[i * P + p]) * b;
                            Hc(i + 1) = -sn[i * P + p] * a + cs[i * P + p] * b;
                        }
                        const cplx a = Hc(j), b = Hc(j + 1);
                        double den = std::sqrt(std::norm(a) + std::norm(b));
                        if (!(den > 0)) den = 1;
                        cs[j * P + p] = a / den;
                        sn[j * P + p] = b / den;
                        Hc(j) = std::conj(cs[j * P + p]) * a + std::conj(sn[j * P + p]) * b;
                        Hc(j + 1) = 0;
                        g[(j + 1) * P + p] = -sn[j * P + p] * g[j * P + p];
                        g[j * P + p] = std::conj(cs[j * P + p]) * g[j * P + p];
                        res[p] = std::abs(g[(j + 1) * P + p]) / bnorm[p];
                        if (res[p] > tol) all = false;
                    }
                    if (all) { ++j; ++its; break; }
                }
                // y from the triangular systems; X += M^-1 (V y)
                std::vector<double> y((size_t)j * P * 2);
                for (int p = 0; p < P; ++p) {
                    std::vector<cplx> yy(j);
                    for (int i = j - 1; i >= 0; --i) {
                        cplx s = g[i * P + p];
                        for (int k = i + 1; k < j; ++k) s -= H[((size_t)i * m + k) * P + p] * yy[k];
                        const cplx d = H[((size_t)i * m + i) * P + p];
                        yy[i] = std::abs(d) > 0 ? s / d : cplx(0);
                    }
                    for (int i = 0; i < j; ++i) { y[(i * P + p) * 2] = yy[i].real(); y[(i * P + p) * 2 + 1] = yy[i].imag(); }
                }
                PQ_HIP(hipMemsetAsync(W, 0, vsz * 8, L->stream));
                if (j > 0) update(L, V, vsz, j, y.data(), -1.0, P, W);           // W = V y
                apply_precond(L, precond, W, Z, P);
                {
                    std::vector<double> one((size_t)P * 2, 0.0);
                    for (int p = 0; p < P; ++p) one[2 * p] = 1.0;
                    update(L, Z, 0, 1, one.data(), -1.0, P, X);                     // X += Z
                }
                // true residual R = B - A X
                apply_operator(L, omega, X, R, P);
                hipLaunchKernelGGL(vec_sub_kernel, dim3((unsigned)((vsz + 255) / 256)), dim3(256), 0, L->stream, B, vsz, R);
                PQ_HIP(hipGetLastError());
                ++restarts;
            }
            L->iterations = its; L->restarts = restarts;
            double worst = 0;
            for (double r : res) worst = std::max(worst, r);
            L->residual = worst;
            if (iterations) *iterations = its;
            if (residual) *residual = worst;
            if (worst > std::max(5 * tol, 1e-13))
                throw Error(PQ_ECONVERGE, "GMRES did not converge: relative residual " + std::to_string(worst) +
                                              " after " + std::to_string(its) + " iterations");
        }
        // Zp = Z_ss + Z_sf X
        DevBuf red;
        red.reserve((size_t)P * P * 2 * 8);
        if (n > 0) {
            hipLaunchKernelGGL(port_reduce_kernel, dim3(P * P), dim3(256), 0, L->stream, L->Ll.as<float>(), L->ldl,
                               L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, L->sc.as<int>(), P, omega, X,
                               red.as<double>());
            PQ_HIP(hipGetLastError());
        } else {
            PQ_HIP(hipMemsetAsync(red.p, 0, (size_t)P * P * 2 * 8, L->stream));
        }
        std::vector<double> zsf((size_t)P * P * 2);
        PQ_HIP(hipMemcpyAsync(zsf.data(), red.p, zsf.size() * 8, hipMemcpyDeviceToHost, L->stream));
        // Z_ss from the stored rows
        std::vector<float> llss((size_t)P * P);
        for (int s = 0; s < P; ++s)
            for (int q = 0; q < P; ++q)
                PQ_HIP(hipMemcpyAsync(&llss[s * P + q], L->Ll.as<float>() + (size_t)port_cols[s] * L->ldl + port_cols[q], 4,
                                      hipMemcpyDeviceToHost, L->stream));
        std::vector<int> rp(L->nl + 1);
        PQ_HIP(hipMemcpyAsync(rp.data(), L->rptr.p, (L->nl + 1) * 4, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
        for (int s = 0; s < P; ++s) {
            const int row = port_cols[s];
            const int cnt = rp[row + 1] - rp[row];
            std::vector<int> cols(cnt);
            std::vector<double> vals(cnt);
            PQ_HIP(hipMemcpy(cols.data(), L->rcol.as<int>() + rp[row], cnt * 4, hipMemcpyDeviceToHost));
            PQ_HIP(hipMemcpy(vals.data(), L->rval.as<double>() + rp[row], cnt * 8, hipMemcpyDeviceToHost));
            for (int q = 0; q < P; ++q) {
                double re = 0;
                for (int t = 0; t < cnt; ++t) if (cols[t] == port_cols[q]) re += vals[t];
                Zp[2 * (s * P + q)] = re + zsf[2 * (s * P + q)];
                Zp[2 * (s * P + q) + 1] = omega * (double)llss[s * P + q] + zsf[2 * (s * P + q) + 1];
            }
        }
        if (X_host && n > 0) PQ_HIP(hipMemcpy(X_host, X, vsz * 8, hipMemcpyDeviceToHost));
        L->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    });
}

int pq_loops_last_solve(pq_loops *L, int64_t *iterations, int64_t *matvecs, double *residual, double *seconds) {
    return guard([&] {
        PQ_REQUIRE(L, "pq_loops_last_solve: null handle");
        if (iterations) *iterations = L->iterations;
        if (matvecs) *matvecs = L->matvecs;
        if (residual) *residual = L->residual;
        if (seconds) *seconds = L->seconds;
    });
}

}  // extern "C"
