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
// columns 0..P-1 and imaginary parts in P..2P-1. Ll is kept twice, (n, ldl) row-major: in FP64
// (the system: right-hand sides, residuals, port reduction, the FP64 fallback factor) and rounded
// to FP32 (the Krylov products, the preconditioners and the FP32 factor); every product
// accumulates in FP64. Rl is CSR (FP64 values). Reductions are two-stage with a fixed order, so
// results do not depend on scheduling.
#include "pq_common.h"
#include "ldl.h"
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <limits>
#include <map>
#include <tuple>
#include <mutex>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
#include <type_traits>
#include <vector>

namespace pq {

using cplx = std::complex<double>;

// ---- projection: Ll = C^T Lb C ------------------------------------------------------------

// U[k, j] = sum_t s_t Lb[k, row_t] over the branches t of loop j (C in CSC: colptr, rows, signs),
// PROJ_ROWS rows k per workgroup: each loop's branch list is read once for all of them (it was
// read once per row, the kernel's dominant traffic), and the rows' gathers of a branch are
// independent loads in flight together.
constexpr int PROJ_ROWS = 8;
template <class TB>
__global__ __launch_bounds__(256) void proj_left_kernel(const TB *Lb, long long ldb, int nbr,
                                                       const int *colptr, const int *rows, const float *sign,
                                                       int nl, double *U, long long ldu) {
    const int k0 = blockIdx.x * PROJ_ROWS;
    const TB *Lr[PROJ_ROWS];
#pragma unroll
    for (int i = 0; i < PROJ_ROWS; ++i) Lr[i] = Lb + (size_t)min(k0 + i, nbr - 1) * ldb;
    for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        double acc[PROJ_ROWS];
#pragma unroll
        for (int i = 0; i < PROJ_ROWS; ++i) acc[i] = 0;
        for (int t = colptr[j]; t < colptr[j + 1]; ++t) {
            const int r = rows[t];
            if (r >= nbr) continue;
            const double sg = sign[t];
#pragma unroll
            for (int i = 0; i < PROJ_ROWS; ++i) acc[i] += sg * (double)Lr[i][r];
        }
#pragma unroll
        for (int i = 0; i < PROJ_ROWS; ++i)
            if (k0 + i < nbr) U[(size_t)(k0 + i) * ldu + j] = acc[i];
    }
}

__global__ __launch_bounds__(256) void proj_right_kernel(const double *U, long long ldu, int nbr, const int *colptr,
                                                        const int *rows, const float *sign, int nl, double *Ll,
                                                        long long ldl) {
    const int i = blockIdx.x;
    const int t0 = colptr[i], t1 = colptr[i + 1];
    for (int j = threadIdx.x; j < nl; j += blockDim.x) {
        double acc = 0;
        for (int t = t0; t < t1; ++t) {
            const int r = rows[t];
            if (r < nbr) acc += (double)sign[t] * U[(size_t)r * ldu + j];
        }
        Ll[(size_t)i * ldl + j] = acc;
    }
}

// ---- operator products --------------------------------------------------------------------

// W = Rl X on rows [0, n) restricted to columns < n (CSR rows of the full loop matrix); NR = 2P.
// One wave per row, lanes over its nonzeros (all NR columns each), then a butterfly per column.
template <int NR>
__global__ __launch_bounds__(256) void spmv_kernel(const int *__restrict__ ptr, const int *__restrict__ col,
                                                  const double *__restrict__ val, int n, const double *__restrict__ X,
                                                  double *__restrict__ W) {
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
    for (int r = 0; r < NR; ++r)
        for (int off = 32; off > 0; off >>= 1) acc[r] += __shfl_xor(acc[r], off);
    if (lane < NR) {
        double v = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r)
            if (r == lane) v = acc[r];
        W[(size_t)row * NR + lane] = v;
    }
}

// ---- the projection through the spanning tree ----
// A loop is its closing branch plus the tree path between the branch's ends: C e_j = e_cot(j) +
// p(a_j) - p(b_j), p(x) the signed branch set of the path from node x to its root (the part
// above the two paths' meeting node cancels). So with S[x] = sum over that path of s_y Lb[upb y]
// (prefix sums down the tree, one level at a time), (Lb C)[k][j] = Lb[k][cot j] + S[a_j][k] -
// S[b_j][k]: two row reads per loop instead of one gather per branch of the loop. The same again
// on the rows of U = Lb C gives Ll = C^T U. All sums in FP64; U and Ll are stored in FP64 (Lb in
// FP32 or FP64).

// S[x][k] = S[parent x][k] + s_x Row[upb x][k] for the nodes x of one level (rows past nvalid
// carry nothing); k along the lanes, Row read along its rows (Lb is symmetric).
template <class TR>
__global__ __launch_bounds__(256) void tree_level_kernel(const TR *__restrict__ Row, long long ldr, int nvalid,
                                                        int width, const int *__restrict__ nodes, int count,
                                                        const int *__restrict__ parent, const int *__restrict__ upb,
                                                        const float *__restrict__ ups, double *__restrict__ S,
                                                        long long lds) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= width) return;
    for (int e = blockIdx.y; e < count; e += gridDim.y) {
        const int x = nodes[e], b = upb[x];
        double v = S[(size_t)parent[x] * lds + k];
        if (b >= 0 && b < nvalid) v += (double)ups[x] * (double)Row[(size_t)b * ldr + k];
        S[(size_t)x * lds + k] = v;
    }
}

// U[k][j] = Lb[cot j][k] + S[a_j][k] - S[b_j][k] by 64 x 64 tiles through LDS: rows of S and Lb
// read along k, rows of U written along j.
template <class TB>
__global__ __launch_bounds__(256) void tree_u_kernel(const TB *__restrict__ Lb, long long ldb, int nbr,
                                                    const double *__restrict__ S, long long lds,
                                                    const int *__restrict__ cot, const int *__restrict__ ea,
                                                    const int *__restrict__ eb, int nl, double *__restrict__ U,
                                                    long long ldu) {
    __shared__ double t[64][65];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int k0 = blockIdx.x * 64, j0 = blockIdx.y * 64, k = k0 + lane;
    for (int q = 0; q < 16; ++q) {
        const int jl = 16 * w + q, j = j0 + jl;
        double v = 0.0;
        if (j < nl && k < nbr) {
            const int c = cot[j];
            v = S[(size_t)ea[j] * lds + k] - S[(size_t)eb[j] * lds + k];
            if (c >= 0 && c < nbr) v += (double)Lb[(size_t)c * ldb + k];
        }
        t[jl][lane] = v;
    }
    __syncthreads();
    for (int q = 0; q < 16; ++q) {
        const int kl = 16 * w + q, kk = k0 + kl, j = j0 + lane;
        if (kk < nbr && j < nl) U[(size_t)kk * ldu + j] = t[lane][kl];
    }
}

// Ll[i][j] = U[cot i][j] + S2[a_i][j] - S2[b_i][j] (rows i, lanes along j).
__global__ __launch_bounds__(256) void tree_ll_kernel(const double *__restrict__ U, long long ldu, int nbr,
                                                     const double *__restrict__ S2, long long lds,
                                                     const int *__restrict__ cot, const int *__restrict__ ea,
                                                     const int *__restrict__ eb, int nl, double *__restrict__ Ll,
                                                     long long ldl) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= nl) return;
    for (int i = blockIdx.y; i < nl; i += gridDim.y) {
        const int c = cot[i];
        double d = S2[(size_t)ea[i] * lds + j] - S2[(size_t)eb[i] * lds + j];
        if (c >= 0 && c < nbr) d += U[(size_t)c * ldu + j];
        Ll[(size_t)i * ldl + j] = d;
    }
}

// Ll := (Ll + Ll^T) / 2 in FP64 (the loop matrices' symmetric part, as the reference takes it:
// src/pcbqs/mqs_iterative.py:244-257), and its FP32 copy L32 rounded from it. Entries (i, j) and
// (j, i) then agree bitwise in both, which lets the products read columns of Ll as rows. One
// thread per unordered pair (i <= j), so no two threads touch one entry.
__global__ void symmetrize_kernel(double *Ll, float *L32, long long ldl, int n) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long i = e / n, j = e % n;
    if (i >= n || j < i) return;
    const double v = 0.5 * (Ll[i * ldl + j] + Ll[j * ldl + i]);
    Ll[i * ldl + j] = v;
    Ll[j * ldl + i] = v;
    L32[i * ldl + j] = (float)v;
    L32[j * ldl + i] = (float)v;
}

// Panel products: part[i][s][r] = sum over the rows c of chunk s of M[c][i] X[c][r], M (ncols, ld)
// FP32 row-major, X (ncols, NR) FP64, i < nrows. One thread per i (blockDim.x rows per block), so
// each row segment of M is consecutive floats (coalesced), 16 of them in flight per lane; X[c] is
// the same for the whole wave and comes through scalar loads, which keeps the LDS return path out
// of the inner loop. With the symmetric Ll (rows read as columns) this is Ll X. M is streamed once
// per product, so its loads are non-temporal: they do not displace the vectors and bases that the
// rest of the iteration re-reads.
constexpr int GEMV_SPLIT = 32, GEMV_ROWS = 256, GEMV_CHUNK = 640;
template <int NR, class T = float>
__global__ __launch_bounds__(256) void panel_kernel(const T *__restrict__ M, long long ld, int nrows, int ncols,
                                                   const double *__restrict__ X, int chunk,
                                                   double *__restrict__ part) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    const int c0 = blockIdx.y * chunk, c1 = min(ncols, c0 + chunk);
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    const T *col = M + min(row, nrows - 1);
    int c = c0;
    for (; c + 16 <= c1; c += 16) {
        T m[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) m[u] = __builtin_nontemporal_load(col + (size_t)(c + u) * ld);
        const double *x = X + (size_t)c * NR;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double mu = m[u];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += mu * x[u * NR + r];
        }
    }
    for (; c < c1; ++c) {
        const double mv = __builtin_nontemporal_load(col + (size_t)c * ld);
        const double *x = X + (size_t)c * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += mv * x[r];
    }
    if (row < nrows) {
        double *o = part + ((size_t)row * gridDim.y + blockIdx.y) * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) o[r] = acc[r];
    }
}

// W_re -= w sum_s part_im, W_im += w sum_s part_re (chunks summed in order); w in device memory, so
// a captured graph stays valid from one frequency to the next.
template <int NR>
__global__ void gemv_epilogue_kernel(const double *part, int nsplit, int n, const double *omega_dev, double *W) {
    const double omega = *omega_dev;
    constexpr int P = NR / 2;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * P) return;
    const int row = e / P, p = e % P;
    double yr = 0, yi = 0;
    for (int q = 0; q < nsplit; ++q) {
        yr += part[((size_t)row * nsplit + q) * NR + p];
        yi += part[((size_t)row * nsplit + q) * NR + P + p];
    }
    W[(size_t)row * NR + p] -= omega * yi;
    W[(size_t)row * NR + P + p] += omega * yr;
}

// One wave's share of row `row` of Rl X (NR real columns): lanes over the row's CSR entries (port
// columns >= n skipped), each lane's NR-vector of its entries' products in a[]. The first 128
// entries have every load issued up front (an absent or port entry gets weight 0 and row 0's x);
// longer rows continue in a loop.
template <int NR>
__device__ __forceinline__ void rl_row_lanes(const int *__restrict__ ptr, const int *__restrict__ col,
                                             const double *__restrict__ val, int n, const double *__restrict__ X,
                                             int row, int lane, double *a) {
    const int t0 = ptr[row], t1 = ptr[row + 1];
    {
        const int ta = t0 + lane, tb = ta + 64;
        const int ca = ta < t1 ? col[ta] : n, cb = tb < t1 ? col[tb] : n;
        const double va = ca < n ? val[ta] : 0.0, vb = cb < n ? val[tb] : 0.0;
        const double *xa = X + (size_t)(ca < n ? ca : 0) * NR, *xb = X + (size_t)(cb < n ? cb : 0) * NR;
        double ya[NR], yb[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            ya[r] = xa[r];
            yb[r] = xb[r];
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            a[r] += va * ya[r];
            a[r] += vb * yb[r];
        }
    }
    // longer rows (up to ~1,100 entries at 30k RWG: the wave taking one sets the kernel's tail) in
    // groups of 128, each lane's two entries' loads issued before their products
    for (int t = t0 + 128 + lane; t < t1; t += 128) {
        const int tb = t + 64;
        const int ca = col[t], cb = tb < t1 ? col[tb] : n;
        const double va = ca < n ? val[t] : 0.0, vb = cb < n ? val[tb] : 0.0;
        const double *xa = X + (size_t)(ca < n ? ca : 0) * NR, *xb = X + (size_t)(cb < n ? cb : 0) * NR;
        double ya[NR], yb[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            ya[r] = xa[r];
            yb[r] = xb[r];
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            a[r] += va * ya[r];
            a[r] += vb * yb[r];
        }
    }
}

// The wave's sum of the lanes' NR-vectors by recursive halving (each step a lane keeps half of its
// vector and adds its partner's copy of that half; once a single value is left, plain xor steps):
// afterwards lane r * 64 / NR holds column r's sum in a[0], in a fixed order.
template <int NR>
__device__ __forceinline__ void wave_column_sums(double *a, int lane) {
#pragma unroll
    for (int off = 32, half = NR / 2; half >= 1; off >>= 1, half >>= 1) {
        const bool upper = (lane & off) != 0;
#pragma unroll
        for (int i = 0; i < half; ++i) {
            const double keep = upper ? a[i + half] : a[i];
            const double send = upper ? a[i] : a[i + half];
            a[i] = keep + __shfl_xor(send, off);
        }
    }
#pragma unroll
    for (int off = 32 / NR; off >= 1; off >>= 1) a[0] += __shfl_xor(a[0], off);
}

// The operator's epilogue fused with the resistive part: W = Rl X + j w (sum of the nsplit chunk
// partials of Ll X), for NR a power of two. One wave per row: lanes over the row's CSR entries and
// over the partials, then one recursive-halving sum.
template <int NR>
__global__ __launch_bounds__(256) void operator_epilogue_kernel(const int *__restrict__ ptr,
                                                               const int *__restrict__ col,
                                                               const double *__restrict__ val, int n,
                                                               const double *__restrict__ X,
                                                               const double *__restrict__ part, int nsplit,
                                                               const double *__restrict__ omega_dev,
                                                               double *__restrict__ W) {
    static_assert((NR & (NR - 1)) == 0 && NR >= 2 && NR <= 32, "NR a power of two");
    constexpr int P = NR / 2;
    const int lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    double a[NR], b[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) a[r] = b[r] = 0;
    for (int q = lane; q < nsplit; q += 64) {
        const double *pq = part + ((size_t)row * nsplit + q) * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) b[r] += pq[r];
    }
    rl_row_lanes<NR>(ptr, col, val, n, X, row, lane, a);
    const double omega = *omega_dev;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        const double re = a[p] - omega * b[P + p], im = a[P + p] + omega * b[p];
        a[p] = re;
        a[P + p] = im;
    }
    wave_column_sums<NR>(a, lane);
    constexpr int S = 64 / NR;
    if (lane % S == 0) W[(size_t)row * NR + lane / S] = a[0];
}

// The four-wave panel product of one 64-row block (bx) and column chunk (by) of nsplit: the body of
// panel4_kernel.
template <int NR, class T>
__device__ __forceinline__ void panel4_block(const T *__restrict__ M, long long ld, int nrows, int ncols,
                                             const double *__restrict__ X, int chunk, int bx, int by, int nsplit,
                                             double *__restrict__ part, double *red) {
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row = bx * 64 + lane;
    const int c0 = by * chunk, c1 = min(ncols, c0 + chunk);
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    const T *col = M + min(row, nrows - 1);
    int c = c0 + 16 * w;
    for (; c + 16 <= c1; c += 64) {
        T m[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) m[u] = __builtin_nontemporal_load(col + (size_t)(c + u) * ld);
        const double *x = X + (size_t)c * NR;
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double mu = m[u];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += mu * x[u * NR + r];
        }
    }
    for (int cc = c; cc < c1; ++cc) {        // the chunk's one partial group (this wave's, if any)
        const double mv = __builtin_nontemporal_load(col + (size_t)cc * ld);
        const double *x = X + (size_t)cc * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += mv * x[r];
    }
    if (w >= 2) {
#pragma unroll
        for (int r = 0; r < NR; ++r) red[((w - 2) * 64 + lane) * NR + r] = acc[r];
    }
    __syncthreads();
    if (w < 2) {
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += red[(w * 64 + lane) * NR + r];
    }
    __syncthreads();
    if (w == 1) {
#pragma unroll
        for (int r = 0; r < NR; ++r) red[lane * NR + r] = acc[r];
    }
    __syncthreads();
    if (w == 0 && row < nrows) {
        double *o = part + ((size_t)row * nsplit + by) * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) o[r] = acc[r] + red[lane * NR + r];
    }
}

// panel_kernel with a 256-thread block on 64 rows: its four waves take the chunk's groups of 16
// columns in turn, and their sums are combined in LDS in a fixed order, (w0 + w2) + (w1 + w3), so
// the chip fills with a quarter of the chunk partials that one wave per row block needs (fewer
// partials for the epilogue to read). Same partial layout as panel_kernel.
template <int NR, class T = float>
__global__ __launch_bounds__(256) void panel4_kernel(const T *__restrict__ M, long long ld, int nrows, int ncols,
                                                    const double *__restrict__ X, int chunk,
                                                    double *__restrict__ part) {
    __shared__ double red[2 * 64 * NR];
    panel4_block<NR, T>(M, ld, nrows, ncols, X, chunk, blockIdx.x, blockIdx.y, gridDim.y, part, red);
}

// ---- the symmetric operator product ----------------------------------------------------------
//
// Ll is bitwise symmetric (symmetrise_kernel) and stored whole, but this product reads only the
// 256 x 256 tiles on and above the diagonal: half the bytes of panel_kernel. In an off-diagonal
// tile (I, J), I < J, with entries Ll[r][c] (r in row block I, c in column block J):
//   direct      y[r] += sum_c Ll[r][c] x[c]   thread per row, x[c] wave-uniform (scalar loads);
//   transposed  y[c] += sum_r Ll[r][c] x[r]   the tile passes through LDS 64 columns at a time;
//                                             lane per column, each wave sums 64 rows with x[r]
//                                             wave-uniform, and the four waves' sums are added
//                                             in wave order.
// A diagonal tile does the direct product only. Output row block K receives exactly one partial
// per block index s: the direct one of tile (K, s) for s >= K and the transposed one of tile
// (s, K) for s < K, so the epilogue adds nb = ceil(n / 256) partials in
// a fixed order. While one 64-column panel is reduced, the next one is already loading.
constexpr int SYM_T = 256, SYM_SUB = 64, SYM_PITCH = SYM_SUB + 1;

template <bool FULL>
__device__ __forceinline__ void sym_load(const float *colp, long long ld, int cs, int n, float (&m)[SYM_SUB]) {
#pragma unroll
    for (int u = 0; u < SYM_SUB; ++u) m[u] = colp[(size_t)(FULL ? cs + u : min(cs + u, n - 1)) * ld];
}

template <int NR, bool FULL>
__device__ __forceinline__ void sym_direct(const float (&m)[SYM_SUB], const double *__restrict__ X, int cs, int n,
                                           double (&acc)[NR]) {
#pragma unroll
    for (int u = 0; u < SYM_SUB; ++u) {
        if (!FULL && cs + u >= n) break;
        const double mu = m[u];
        const double *x = X + (size_t)(cs + u) * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += mu * x[r];
    }
}

template <int NR>
__global__ __launch_bounds__(256, 2) void sym_panel_kernel(const float *__restrict__ M, long long ld, int n, int nb,
                                                           const double *__restrict__ X, double *__restrict__ part) {
    constexpr int TD = SYM_T * SYM_PITCH / 2 + 1, ZD = 4 * SYM_SUB * (NR + 1);
    __shared__ double lds[TD > ZD ? TD : ZD];
    float *T = reinterpret_cast<float *>(lds);
    // blocks [0, nb (nb - 1) / 2) are the off-diagonal tiles, column block by column block; then
    // the nb diagonal tiles
    const int ntri = nb * (nb - 1) / 2, t = blockIdx.x;
    int I, J;
    if (t < ntri) {
        J = (int)((1.0 + sqrt(1.0 + 8.0 * (double)t)) * 0.5);
        while (J * (J - 1) / 2 > t) --J;
        while ((J + 1) * J / 2 <= t) ++J;
        I = t - J * (J - 1) / 2;
    } else {
        I = J = t - ntri;
    }
    const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int r0 = I * SYM_T, c0 = J * SYM_T, cend = min(n, c0 + SYM_T);
    const int row = r0 + tid;
    const float *colp = M + min(row, n - 1);          // Ll[row][c] = colp[c * ld]
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    if (I == J) {
        int c = c0;
        for (; c + 16 <= cend; c += 16) {
            float m[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) m[u] = colp[(size_t)(c + u) * ld];
            const double *x = X + (size_t)c * NR;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const double mu = m[u];
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r] += mu * x[u * NR + r];
            }
        }
        for (; c < cend; ++c) {
            const double mv = colp[(size_t)c * ld];
            const double *x = X + (size_t)c * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += mv * x[r];
        }
        if (row < n) {
            double *o = part + ((size_t)row * nb + J) * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) o[r] = acc[r];
        }
        return;
    }
    // off-diagonal: row block I < nb - 1 is full (256 rows < n); column block J may be ragged
    float m[SYM_SUB];
    if (c0 + SYM_SUB <= n) sym_load<true>(colp, ld, c0, n, m);
    else sym_load<false>(colp, ld, c0, n, m);
    const double *xr = X + (size_t)(r0 + w * 64) * NR;   // this wave's rows, for the transposed sums
    double *Z = lds;
#pragma unroll 1
    for (int cs = c0; cs < cend; cs += SYM_SUB) {
        if (cs + SYM_SUB <= n) sym_direct<NR, true>(m, X, cs, n, acc);
        else sym_direct<NR, false>(m, X, cs, n, acc);
#pragma unroll
        for (int u = 0; u < SYM_SUB; ++u) T[tid * SYM_PITCH + u] = m[u];
        __syncthreads();
        const int cn = cs + SYM_SUB;
        if (cn < cend) {
            if (cn + SYM_SUB <= n) sym_load<true>(colp, ld, cn, n, m);
            else sym_load<false>(colp, ld, cn, n, m);
        }
        double z[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) z[r] = 0;
        const float *Tw = T + w * 64 * SYM_PITCH + lane;
#pragma unroll 8
        for (int i = 0; i < 64; ++i) {
            const double tv = Tw[i * SYM_PITCH];
            const double *x = xr + (size_t)i * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) z[r] += tv * x[r];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < NR; ++r) Z[(w * 64 + lane) * (NR + 1) + r] = z[r];
        __syncthreads();
        for (int e = tid; e < 64 * NR; e += 256) {
            const int cl = e / NR, r = e % NR;
            if (cs + cl < n)
                part[((size_t)(cs + cl) * nb + I) * NR + r] =
                    Z[cl * (NR + 1) + r] + Z[(64 + cl) * (NR + 1) + r] + Z[(128 + cl) * (NR + 1) + r] +
                    Z[(192 + cl) * (NR + 1) + r];
        }
        __syncthreads();
    }
    double *o = part + ((size_t)row * nb + J) * NR;
#pragma unroll
    for (int r = 0; r < NR; ++r) o[r] = acc[r];
}

// ---- block-Jacobi --------------------------------------------------------------------------

// Inverse of each diagonal block Z_bb = Rl_bb + j w Ll_bb (size <= BS) by in-place Gauss-Jordan
// with partial pivoting, complex FP32 (a preconditioner); stored as (nblk, BS, BS) complex
// (re, im interleaved). A padded block is the identity. 256 loops per block: against 64-loop
// blocks the plain solve of the 30k-RWG board takes 375 iterations instead of 644, the two-level
// one 53 instead of 86 (128-loop blocks: 421 and 64; 512-loop blocks: 283 and 45, but eight times
// the per-frequency set-up). A 256 x 256 complex block (512 KB) does not fit in LDS, so one
// workgroup per block factors it in place in the output buffer, where it stays in the L2 of its
// XCD: wave 0 picks the pivot (largest |a_ik|, the first on ties), the pivot row and the column of
// factors go through LDS, and the rank-1 update is spread over the workgroup's 1024 threads.
constexpr int BS = 256;

__global__ __launch_bounds__(1024) void bj_setup_kernel(const float *Ll, long long ldl, const int *ptr,
                                                       const int *col, const double *val, int n, double omega,
                                                       float *inv, int *perm_g, int *singular) {
    __shared__ float fr_[BS], fi_[BS], pr_[BS], pi_[BS];
    __shared__ int piv;
    const int b = blockIdx.x, r0 = b * BS, m = min(BS, n - r0);
    float2 *a = reinterpret_cast<float2 *>(inv) + (size_t)b * BS * BS;
    int *perm = perm_g + (size_t)b * BS;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int e = tid; e < BS * BS; e += nt) {
        const int i = e / BS, j = e % BS;
        float re = 0, im = 0;
        if (i < m && j < m) im = (float)(omega * (double)Ll[(size_t)(r0 + i) * ldl + r0 + j]);
        else if (i == j) re = 1;
        a[e] = make_float2(re, im);
    }
    __syncthreads();
    for (int i = tid; i < m; i += nt)
        for (int t = ptr[r0 + i]; t < ptr[r0 + i + 1]; ++t) {
            const int c = col[t] - r0;
            if (c >= 0 && c < m) a[(size_t)i * BS + c].x += (float)val[t];
        }
    __syncthreads();
    for (int k = 0; k < BS; ++k) {
        if (tid < 64) {
            float bv = -1;
            int best = k;
            for (int i = k + tid; i < BS; i += 64) {
                const float2 v = a[(size_t)i * BS + k];
                const float q = v.x * v.x + v.y * v.y;
                if (q > bv) { bv = q; best = i; }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const float ov = __shfl_xor(bv, off);
                const int oi = __shfl_xor(best, off);
                if (ov > bv || (ov == bv && oi < best)) { bv = ov; best = oi; }
            }
            if (tid == 0) {
                piv = best;
                perm[k] = best;
                if (!(bv > 0)) atomicOr(singular, 1);
            }
        }
        __syncthreads();
        // rows k and p exchanged: old row k moves to row p, the pivot row to LDS
        const int p = piv;
        for (int j = tid; j < BS; j += nt) {
            const float2 rk = a[(size_t)k * BS + j], rp = a[(size_t)p * BS + j];
            if (p != k) a[(size_t)p * BS + j] = rk;
            pr_[j] = rp.x; pi_[j] = rp.y;
        }
        __syncthreads();
        const float dr = pr_[k], di = pi_[k], dd = dr * dr + di * di;
        const float ir = dd > 0 ? dr / dd : 0.f, ii = dd > 0 ? -di / dd : 0.f;   // 1 / a_kk
        __syncthreads();
        if (tid == 0) { pr_[k] = 1; pi_[k] = 0; }
        __syncthreads();
        for (int j = tid; j < BS; j += nt) {
            const float xr = pr_[j], xi = pi_[j];
            const float yr = xr * ir - xi * ii, yi = xr * ii + xi * ir;
            pr_[j] = yr; pi_[j] = yi;
            a[(size_t)k * BS + j] = make_float2(yr, yi);
        }
        // column k of the other rows: the factors, then a_ik := 0 before the row update
        for (int i = tid; i < BS; i += nt) {
            if (i == k) continue;
            const float2 f = a[(size_t)i * BS + k];
            fr_[i] = f.x; fi_[i] = f.y;
            a[(size_t)i * BS + k] = make_float2(0.f, 0.f);
        }
        __syncthreads();
        for (int e = tid; e < BS * BS; e += nt) {
            const int i = e / BS, j = e % BS;
            if (i == k) continue;
            const float fr = fr_[i], fi = fi_[i];
            const float xr = pr_[j], xi = pi_[j];
            float2 v = a[e];
            v.x -= fr * xr - fi * xi;
            v.y -= fr * xi + fi * xr;
            a[e] = v;
        }
        __syncthreads();
    }
    // undo the row interchanges as column interchanges, last first
    for (int k = BS - 1; k >= 0; --k) {
        const int p = perm[k];
        if (p != k)
            for (int i = tid; i < BS; i += nt) {
                const float2 t = a[(size_t)i * BS + k];
                a[(size_t)i * BS + k] = a[(size_t)i * BS + p];
                a[(size_t)i * BS + p] = t;
            }
        __syncthreads();
    }
}

// The block-Jacobi inverses from each block's modes. Z_b = R_b + j w L_b with R_b symmetric
// positive definite and L_b symmetric: with the generalised eigenvectors V_b (L_b V = R_b V Lam,
// V^T R_b V = I), Z_b^-1 = V_b diag(1 / (1 + j w lam)) V_b^T at every frequency. So one batched
// eigensolve per handle (rocSOLVER dsygvd) replaces the per-frequency Gauss-Jordan inversion, and a
// frequency's inverses are one batched complex GEMM (rocBLAS cgemm) into the same layout.

// R_b and L_b (FP64, BS x BS) of every diagonal block, grid (BS rows, nblk): L from the FP64 loop
// matrix's rows, R from the resistance's CSR rows; beyond a short last block's m rows and columns,
// R = I and L = 0 (modes with lam = 0: the inverse is the identity there, as the padded block is).
__global__ __launch_bounds__(256) void bj_pencil_kernel(const double *__restrict__ Ll64, long long ldl,
                                                       const int *__restrict__ ptr, const int *__restrict__ col,
                                                       const double *__restrict__ val, int n, double *Lb,
                                                       double *Rb) {
    const int i = blockIdx.x, b = blockIdx.y, r0 = b * BS, m = min(BS, n - r0);
    double *lrow = Lb + ((size_t)b * BS + i) * BS, *rrow = Rb + ((size_t)b * BS + i) * BS;
    for (int j = threadIdx.x; j < BS; j += blockDim.x) {
        lrow[j] = (i < m && j < m) ? Ll64[(size_t)(r0 + i) * ldl + r0 + j] : 0.0;
        rrow[j] = (i >= m && i == j) ? 1.0 : 0.0;
    }
    __syncthreads();
    if (i < m)
        for (int t = ptr[r0 + i] + threadIdx.x; t < ptr[r0 + i + 1]; t += blockDim.x) {   // distinct columns
            const int c = col[t] - r0;
            if (c >= 0 && c < m) rrow[c] = val[t];
        }
}

// The eigenvectors (FP64, column-major) as the FP32 complex right operand of the inverses' product.
__global__ void bj_modes_kernel(const double *V, size_t count, float2 *Vc) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) Vc[e] = make_float2((float)V[e], 0.f);
}

// A_b = V_b diag(1 / (1 + j w lam_b)) (column-major: entry e = i + a BS of block b scales by mode a).
__global__ void bj_scale_kernel(const float2 *Vc, const double *lam, size_t count, double omega, float2 *A) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const size_t b = e / ((size_t)BS * BS), a = (e / BS) % BS;
    const double wl = omega * lam[b * BS + a], den = 1.0 + wl * wl;
    const float v = Vc[e].x;
    A[e] = make_float2((float)(v / den), (float)(-v * wl / den));
}

// Y = Z_b^-1 X (+ A) on every 256-loop block (Z_b^-1 FP32 complex), sixteen lanes per row: lane s of a row reads the
// row's columns s, s + 16, ... (a wave's load is four rows' 128 contiguous bytes, no lane repeating
// another's), accumulates every right-hand side against X staged in LDS by right-hand side (conflict-
// free across the sixteen columns), and the sixteen partial sums meet by a fixed butterfly. A
// workgroup takes 16 rows of block blockIdx.x from row 16 blockIdx.y.
template <int NR>
__global__ __launch_bounds__(256) void bj_apply16_kernel(const float *inv, int n, const double *X, const double *A,
                                                        double *Y) {
    constexpr int P = NR / 2;
    __shared__ double xs[NR * BS];
    const int b = blockIdx.x, r0 = b * BS, m = min(BS, n - r0);
    for (int e = threadIdx.x; e < BS * NR; e += blockDim.x) {
        const int j = e / NR, q = e % NR;
        xs[q * BS + j] = j < m ? X[(size_t)r0 * NR + e] : 0.0;
    }
    __syncthreads();
    const float2 *Zi = reinterpret_cast<const float2 *>(inv) + (size_t)b * BS * BS;
    const int sp = threadIdx.x & 15, i = blockIdx.y * 16 + (threadIdx.x >> 4);
    const bool live = i < m;
    double yr[P], yi[P];
#pragma unroll
    for (int p = 0; p < P; ++p) yr[p] = yi[p] = 0;
    if (live) {
        const float2 *zrow = Zi + (size_t)i * BS;
#pragma unroll 4
        for (int j = sp; j < m; j += 16) {
            const float2 z = zrow[j];
            const double zr = z.x, zi = z.y;
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const double xr = xs[p * BS + j], xi = xs[(P + p) * BS + j];
                yr[p] += zr * xr - zi * xi;
                yi[p] += zr * xi + zi * xr;
            }
        }
    }
#pragma unroll
    for (int o = 1; o < 16; o <<= 1)
#pragma unroll
        for (int p = 0; p < P; ++p) {
            yr[p] += __shfl_xor(yr[p], o);
            yi[p] += __shfl_xor(yi[p], o);
        }
    if (live && sp < P) {                             // lane sp writes right-hand side sp
        double vr = yr[0], vi = yi[0];
#pragma unroll
        for (int p = 1; p < P; ++p)
            if (sp == p) { vr = yr[p]; vi = yi[p]; }
        if (A) {
            vr += A[(size_t)(r0 + i) * NR + sp];
            vi += A[(size_t)(r0 + i) * NR + P + sp];
        }
        Y[(size_t)(r0 + i) * NR + sp] = vr;
        Y[(size_t)(r0 + i) * NR + P + sp] = vi;
    }
}

// ---- Krylov vector kernels -----------------------------------------------------------------

constexpr int RED_BLOCKS = 256;

// partial[i][blk][p] = sum over the block's rows of conj(V_i[n,p]) W[n,p], i < nv (complex): one
// row per thread for all P columns, then one LDS tree over the block for the 2P sums together.
template <int P>
__global__ __launch_bounds__(256) void dots_partial_kernel(const double *__restrict__ V, size_t vstride, int nv,
                                                          const double *__restrict__ W, int n,
                                                          double *__restrict__ partial) {
    constexpr int NR = 2 * P;
    __shared__ double s[NR][256];
    const int i = blockIdx.y;
    const double *Vi = V + (size_t)i * vstride;
    double re[P], im[P];
#pragma unroll
    for (int p = 0; p < P; ++p) re[p] = im[p] = 0;
    for (int row = blockIdx.x * 256 + threadIdx.x; row < n; row += gridDim.x * 256) {
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double vr = Vi[(size_t)row * NR + p], vi = Vi[(size_t)row * NR + P + p];
            const double wr = W[(size_t)row * NR + p], wi = W[(size_t)row * NR + P + p];
            re[p] += vr * wr + vi * wi;
            im[p] += vr * wi - vi * wr;
        }
    }
#pragma unroll
    for (int p = 0; p < P; ++p) { s[2 * p][threadIdx.x] = re[p]; s[2 * p + 1][threadIdx.x] = im[p]; }
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st)
#pragma unroll
            for (int q = 0; q < NR; ++q) s[q][threadIdx.x] += s[q][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x < NR) partial[((size_t)i * gridDim.x + blockIdx.x) * NR + threadIdx.x] = s[threadIdx.x][0];
}

// h[i][q] = sum over the nblk block partials of vector i (one wave per output; lanes over blocks,
// then a fixed butterfly); also into h_host (pinned, coherent host memory) when given.
__global__ __launch_bounds__(256) void dots_reduce_kernel(const double *__restrict__ partial, int nblk, int nv,
                                                         int NR, double *__restrict__ h, double *__restrict__ h_host) {
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= nv * NR) return;
    const int i = e / NR, q = e % NR;
    double s = 0;
    for (int b = lane; b < nblk; b += 64) s += partial[((size_t)i * nblk + b) * NR + q];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (lane == 0) {
        h[e] = s;
        if (h_host) h_host[e] = s;
    }
}

// W[n,p] -= sum_i V_i[n,p] h[i][p] (complex); with sign -1 it adds. Four adjacent lanes share an
// output (row, p) and take every fourth vector; two butterfly steps combine them in a fixed order.
__global__ __launch_bounds__(256) void vec_update_kernel(const double *V, size_t vstride, int nv, const double *h,
                                                        double sign, int n, int P, double *W) {
    const int NR = 2 * P;
    const size_t e = ((size_t)blockIdx.x * 256 + threadIdx.x) >> 2;
    const int part = threadIdx.x & 3;
    const bool live = e < (size_t)n * P;
    const int row = live ? (int)(e / P) : 0, p = live ? (int)(e % P) : 0;
    double ar = 0, ai = 0;
    if (live) {
        for (int i = part; i < nv; i += 4) {
            const double hr = h[(i * P + p) * 2], hi = h[(i * P + p) * 2 + 1];
            const double vr = V[(size_t)i * vstride + (size_t)row * NR + p];
            const double vi = V[(size_t)i * vstride + (size_t)row * NR + P + p];
            ar += vr * hr - vi * hi;
            ai += vr * hi + vi * hr;
        }
    }
    ar += __shfl_xor(ar, 1); ai += __shfl_xor(ai, 1);
    ar += __shfl_xor(ar, 2); ai += __shfl_xor(ai, 2);
    if (live && part == 0) {
        W[(size_t)row * NR + p] -= sign * ar;
        W[(size_t)row * NR + P + p] -= sign * ai;
    }
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

// B[:, p] = -(Rl[0:n, sc_p] + j w Ll[0:n, sc_p]) (the port loops' columns, free rows; Ll FP64).
__global__ void rhs_kernel(const double *Ll, long long ldl, const int *ptr, const int *col, const double *val,
                           int n, const int *sc, int P, double omega, double *B) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * P) return;
    const int row = e / P, p = e % P;
    double re = 0;
    for (int t = ptr[row]; t < ptr[row + 1]; ++t)
        if (col[t] == sc[p]) re += val[t];
    B[(size_t)row * 2 * P + p] = -re;
    B[(size_t)row * 2 * P + P + p] = -omega * Ll[(size_t)row * ldl + sc[p]];
}

// Zsf X: out[s][q] = sum_i (Rl[sc_s, i] + j w Ll[sc_s, i]) X[i, q] over free i < n (complex; Ll FP64).
__global__ __launch_bounds__(256) void port_reduce_kernel(const double *Ll, long long ldl, const int *ptr,
                                                         const int *col, const double *val, int n, const int *sc,
                                                         int P, double omega, const double *X, double *out) {
    __shared__ double sr[256], si[256];
    const int s = blockIdx.x / P, q = blockIdx.x % P;
    const int NR = 2 * P;
    const int row = sc[s];
    double re = 0, im = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double l = omega * Ll[(size_t)row * ldl + i];
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


// ---- two-level (coarse space) preconditioner -------------------------------------------------

// The second Gram-Schmidt pass and the normalisation in one. h holds, per port p, the
// coefficients h_i = conj(V_i) . W for i < nv and then (|W_p|^2, 0), as the dots leave them for
// the nv + 1 vectors V_0..V_{nv-1}, W. Then W_p := (W_p - sum_i V_i h_i) / sqrt(d_p) with
// d_p = |W_p|^2 - sum_i |h_i|^2 (Pythagoras: the V_i are orthonormal), and nrm[p] = (d_p, 0), the
// squared Hessenberg subdiagonal (also written to nrm_host, pinned host memory); d_p <= 0 zeroes
// the column. Four adjacent lanes share an output
// as in vec_update_kernel; the sums are combined by the same fixed butterfly in every row.
__global__ __launch_bounds__(256) void vec_update_norm_kernel(const double *V, size_t vstride, int nv, const double *h,
                                                             int n, int P, double *W, double *nrm, double *nrm_host) {
    const int NR = 2 * P;
    const size_t e = ((size_t)blockIdx.x * 256 + threadIdx.x) >> 2;
    const int part = threadIdx.x & 3;
    const bool live = e < (size_t)n * P;
    const int row = live ? (int)(e / P) : 0, p = live ? (int)(e % P) : 0;
    double ar = 0, ai = 0, hh = 0;
    if (live) {
        for (int i = part; i < nv; i += 4) {
            const double hr = h[(i * P + p) * 2], hi = h[(i * P + p) * 2 + 1];
            const double vr = V[(size_t)i * vstride + (size_t)row * NR + p];
            const double vi = V[(size_t)i * vstride + (size_t)row * NR + P + p];
            ar += vr * hr - vi * hi;
            ai += vr * hi + vi * hr;
            hh += hr * hr + hi * hi;
        }
    }
    ar += __shfl_xor(ar, 1); ai += __shfl_xor(ai, 1); hh += __shfl_xor(hh, 1);
    ar += __shfl_xor(ar, 2); ai += __shfl_xor(ai, 2); hh += __shfl_xor(hh, 2);
    if (live && part == 0) {
        const double d = h[((size_t)nv * P + p) * 2] - hh;
        const double wr = W[(size_t)row * NR + p] - ar, wi = W[(size_t)row * NR + P + p] - ai;
        const double s = d > 0 ? sqrt(d) : 0.0;
        W[(size_t)row * NR + p] = d > 0 ? wr / s : 0.0;
        W[(size_t)row * NR + P + p] = d > 0 ? wi / s : 0.0;
        if (row == 0) {
            nrm[2 * p] = d;
            nrm[2 * p + 1] = 0.0;
            nrm_host[2 * p] = d;
            nrm_host[2 * p + 1] = 0.0;
        }
    }
}

// Y[n,p] = X[n,p] / sqrt(nrm[p]) with nrm[p] = (|X_p|^2, 0) as dots_reduce leaves it (0 -> 0).
__global__ void vec_scale_norm_kernel(const double *X, const double *nrm, int n, int P, double *Y) {
    const int NR = 2 * P;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * NR) return;
    const int c = (int)(e % NR) % P;
    const double s = nrm[2 * c];
    Y[e] = s > 0 ? X[e] / sqrt(s) : 0.0;
}

// Restriction partials: partial[chunk][c][r] = sum over the chunk's rows i of Phi[i][c] V[i][r],
// Phi (n, k) FP32 row-major with row stride kp (a multiple of 64 floats, so that every 64-column
// segment is two whole cache lines), V (n, NR) FP64. A block covers 64 coarse columns (the lanes, so each
// row segment of Phi is one coalesced 256-byte read, sixteen in flight per lane) and `rows` rows
// split over its four waves; a row of V is wave-uniform (scalar loads). The four waves' sums are
// added in wave order.
constexpr int cr_rows(int NR) { return NR >= 8 ? 4096 / NR : 512; }
template <int NR, class T>
__global__ __launch_bounds__(256) void coarse_restrict_kernel(const T *__restrict__ Phi, int k, int kp,
                                                             const double *__restrict__ V, int n, int rows,
                                                             double scale, double *__restrict__ partial) {
    constexpr int U = 16;
    __shared__ double red[4 * 64 * NR];
    const int i0 = blockIdx.x * rows, m = min(rows, n - i0);
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.y * 64 + lane;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    const T *col = Phi + (size_t)i0 * kp + min(c, k - 1);
    const double *v = V + (size_t)i0 * NR;
    int i = w;
    for (; i + 4 * (U - 1) < m; i += 4 * U) {
        T f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) f[u] = col[(size_t)(i + 4 * u) * kp];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double fu = (float)f[u];
            const double *vr = v + (size_t)(i + 4 * u) * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += fu * vr[r];
        }
    }
    for (; i < m; i += 4) {
        const double f = (float)col[(size_t)i * kp];
        const double *vr = v + (size_t)i * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[r] += f * vr[r];
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) red[(w * 64 + lane) * NR + r] = acc[r] * scale;
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * NR; e += 256) {
        const int cl = e / NR, r = e % NR;
        if (blockIdx.y * 64 + cl >= k) continue;
        const double s = red[(0 * 64 + cl) * NR + r] + red[(1 * 64 + cl) * NR + r] + red[(2 * 64 + cl) * NR + r] +
                         red[(3 * 64 + cl) * NR + r];
        partial[((size_t)blockIdx.x * k + blockIdx.y * 64 + cl) * NR + r] = s;
    }
}

// The same restriction partials with four consecutive coarse columns per lane (one 8-byte FP16 or
// 16-byte FP32 load per row instead of a 2- or 4-byte one), a block over 256 columns and `rows` rows
// split over its four waves as above. Each sum runs over the same rows in the same order and the
// four waves' sums are added in wave order, so the partials equal coarse_restrict_kernel's
// bitwise. The waves' sums meet in LDS one right-hand side at a time and leave as contiguous
// (column, rhs) rows. NR <= 16.
template <int NR, class T>
__global__ __launch_bounds__(256) void coarse_restrict4_kernel(const T *__restrict__ Phi, int k, int kp,
                                                              const double *__restrict__ V, int n, int rows,
                                                              double scale, double *__restrict__ partial) {
    static_assert(NR <= 16, "four-column restriction: at most 16 right-hand sides");
    constexpr int U = 16;
    typedef T T4 __attribute__((ext_vector_type(4)));
    __shared__ double red[3][256];
    __shared__ double outs[256 * NR];
    const int i0 = blockIdx.x * rows, m = min(rows, n - i0);
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int cb = blockIdx.y * 256, c0 = cb + 4 * lane;
    double acc[4][NR];
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < NR; ++r) acc[j][r] = 0;
    // kp is a multiple of 64, so a lane whose first column lies inside the padded row has all four
    const T *col = Phi + (size_t)i0 * kp + (c0 < kp ? c0 : 0);
    const double *v = V + (size_t)i0 * NR;
    int i = w;
    for (; i + 4 * (U - 1) < m; i += 4 * U) {
        T4 f[U];
#pragma unroll
        for (int u = 0; u < U; ++u) f[u] = *reinterpret_cast<const T4 *>(col + (size_t)(i + 4 * u) * kp);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double *vr = v + (size_t)(i + 4 * u) * NR;
            const double f0 = (float)f[u].x, f1 = (float)f[u].y, f2 = (float)f[u].z, f3 = (float)f[u].w;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double x = vr[r];
                acc[0][r] += f0 * x; acc[1][r] += f1 * x; acc[2][r] += f2 * x; acc[3][r] += f3 * x;
            }
        }
    }
    for (; i < m; i += 4) {
        const T4 f = *reinterpret_cast<const T4 *>(col + (size_t)i * kp);
        const double *vr = v + (size_t)i * NR;
        const double f0 = (float)f.x, f1 = (float)f.y, f2 = (float)f.z, f3 = (float)f.w;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            const double x = vr[r];
            acc[0][r] += f0 * x; acc[1][r] += f1 * x; acc[2][r] += f2 * x; acc[3][r] += f3 * x;
        }
    }
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if (w > 0)
#pragma unroll
            for (int j = 0; j < 4; ++j) red[w - 1][4 * lane + j] = acc[j][r] * scale;
        __syncthreads();
        if (w == 0)
#pragma unroll
            for (int j = 0; j < 4; ++j)
                outs[(4 * lane + j) * NR + r] =
                    acc[j][r] * scale + red[0][4 * lane + j] + red[1][4 * lane + j] + red[2][4 * lane + j];
        __syncthreads();
    }
    const int ncol = min(256, k - cb);
    for (int e = threadIdx.x; e < ncol * NR; e += 256)
        partial[((size_t)blockIdx.x * k + cb) * NR + e] = outs[e];
}

// t[e] = sum over the row-chunk partials, in chunk order (e = c * NR + r).
__global__ void coarse_sum_kernel(const double *partial, int nblk, int total, double *t) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    double s = 0;
    for (int q = 0; q < nblk; ++q) s += partial[(size_t)q * total + e];
    t[e] = s;
}

// Diagonal coarse matrix: c[a][p] = D[a] t[a][p] with t summed from the chunk partials (t and
// c (k, NR), real parts in columns 0..P-1; D complex interleaved); one wave per output (a, p),
// lanes over the chunks, then a fixed butterfly.
__global__ __launch_bounds__(256) void coarse_sum_diag_kernel(const double *partial, int nblk, int k, int P,
                                                             const double *D, double *cvec) {
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= k * P) return;
    const int NR = 2 * P, a = e / P, p = e % P;
    double tr = 0, ti = 0;
    for (int q = lane; q < nblk; q += 64) {
        tr += partial[((size_t)q * k + a) * NR + p];
        ti += partial[((size_t)q * k + a) * NR + P + p];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        tr += __shfl_xor(tr, off);
        ti += __shfl_xor(ti, off);
    }
    if (lane == 0) {
        const double dr = D[2 * a], di = D[2 * a + 1];
        cvec[(size_t)a * NR + p] = dr * tr - di * ti;
        cvec[(size_t)a * NR + P + p] = dr * ti + di * tr;
    }
}

// c = Ainv t (complex; t and c (k, NR) with real parts in columns 0..P-1). One wave per output
// (a, p), lanes over b.
__global__ __launch_bounds__(256) void coarse_solve_kernel(const double *t, const double *Ainv, int k, int P,
                                                          double *cvec) {
    const int NR = 2 * P;
    const int lane = threadIdx.x & 63;
    const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (e >= k * P) return;
    const int a = e / P, p = e % P;
    double yr = 0, yi = 0;
    for (int b = lane; b < k; b += 64) {
        const double tr = t[b * NR + p], ti = t[b * NR + P + p];
        const double ar = Ainv[2 * ((size_t)a * k + b)], ai = Ainv[2 * ((size_t)a * k + b) + 1];
        yr += ar * tr - ai * ti;
        yi += ar * ti + ai * tr;
    }
    for (int off = 32; off > 0; off >>= 1) { yr += __shfl_down(yr, off); yi += __shfl_down(yi, off); }
    if (lane == 0) { cvec[(size_t)a * NR + p] = yr; cvec[(size_t)a * NR + P + p] = yi; }
}

// The prolongation: the three transposed bases (k, ldt) Phi^T, (Rl Phi)^T, (Ll Phi)^T (FP32, or FP16
// scaled by s0, s1, s2) read in one pass against the coarse vector c (k, NR). A 256-thread block
// takes 64 rows (a row per lane) and a chunk of coarse columns (blockIdx.y); its four waves take the
// chunk's groups of U columns in turn (3 U loads in flight per lane), c wave-uniform (scalar
// loads), FP64 sums. The waves' sums are combined in LDS in a fixed order, (w0 + w2) + (w1 + w3),
// so one partial per block: chunk partials PY = Phi c and PA = (Rl + j w Ll) Phi c (real parts in
// columns 0..P-1). Four waves per row block fill the chip with a quarter of the chunk partials that
// one wave per row block needs.
template <int NR, int U, class T>
__global__ __launch_bounds__(256) void prolong_kernel(const T *__restrict__ B0, const T *__restrict__ B1,
                                                     const T *__restrict__ B2, int n, long long ldt, int k,
                                                     const double *__restrict__ C, int chunk,
                                                     const double *__restrict__ omega_dev, double s0, double s1,
                                                     double s2, double *__restrict__ PY, double *__restrict__ PA) {
    constexpr int P = NR / 2, V = 3 * NR;
    __shared__ double red[2 * 64 * V];
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int row = blockIdx.x * 64 + lane;
    const int c0 = blockIdx.y * chunk, c1 = min(k, c0 + chunk);
    double acc[V];
#pragma unroll
    for (int v = 0; v < V; ++v) acc[v] = 0;
    const size_t rr = (size_t)min(row, n - 1);
    int c = c0 + w * U;
    for (; c + U <= c1; c += 4 * U) {
        T f0[U], f1[U], f2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            f0[u] = B0[(size_t)(c + u) * ldt + rr];
            f1[u] = B1[(size_t)(c + u) * ldt + rr];
            f2[u] = B2[(size_t)(c + u) * ldt + rr];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double a = (float)f0[u], b = (float)f1[u], d = (float)f2[u];
            const double *x = C + (size_t)(c + u) * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                acc[r] += a * x[r];
                acc[NR + r] += b * x[r];
                acc[2 * NR + r] += d * x[r];
            }
        }
    }
    for (int cc = c; cc < c1; ++cc) {          // the chunk's one partial group (this wave's, if any)
        const double a = (float)B0[(size_t)cc * ldt + rr], b = (float)B1[(size_t)cc * ldt + rr];
        const double d = (float)B2[(size_t)cc * ldt + rr];
        const double *x = C + (size_t)cc * NR;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            acc[r] += a * x[r];
            acc[NR + r] += b * x[r];
            acc[2 * NR + r] += d * x[r];
        }
    }
    if (w >= 2) {
#pragma unroll
        for (int v = 0; v < V; ++v) red[((w - 2) * 64 + lane) * V + v] = acc[v];
    }
    __syncthreads();
    if (w < 2) {
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] += red[(w * 64 + lane) * V + v];
    }
    __syncthreads();
    if (w == 1) {
#pragma unroll
        for (int v = 0; v < V; ++v) red[lane * V + v] = acc[v];
    }
    __syncthreads();
    if (w == 0 && row < n) {
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] += red[lane * V + v];
        const double omega = *omega_dev;
        double *oy = PY + ((size_t)blockIdx.y * n + row) * NR, *oa = PA + ((size_t)blockIdx.y * n + row) * NR;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const double rre = acc[NR + p] * s1, rim = acc[NR + P + p] * s1;
            const double lre = acc[2 * NR + p] * s2, lim = acc[2 * NR + P + p] * s2;
            oa[p] = rre - omega * lim;
            oa[P + p] = rim + omega * lre;
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) oy[r] = acc[r] * s0;
    }
}

// R = V - sum PA, Y = sum PY over the nsplit chunk partials of prolong_kernel, in chunk order.
__global__ void prolong_epilogue_kernel(const double *__restrict__ PY, const double *__restrict__ PA, int nsplit,
                                        size_t count, const double *__restrict__ V, double *__restrict__ R,
                                        double *__restrict__ Y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    double y = 0, a = 0;
    int q = 0;
    for (; q + 4 <= nsplit; q += 4) {
        const double y0 = PY[(size_t)q * count + e], y1 = PY[(size_t)(q + 1) * count + e];
        const double y2 = PY[(size_t)(q + 2) * count + e], y3 = PY[(size_t)(q + 3) * count + e];
        const double a0 = PA[(size_t)q * count + e], a1 = PA[(size_t)(q + 1) * count + e];
        const double a2 = PA[(size_t)(q + 2) * count + e], a3 = PA[(size_t)(q + 3) * count + e];
        y += y0; y += y1; y += y2; y += y3;
        a += a0; a += a1; a += a2; a += a3;
    }
    for (; q < nsplit; ++q) {
        y += PY[(size_t)q * count + e];
        a += PA[(size_t)q * count + e];
    }
    R[e] = V[e] - a;
    Y[e] = y;
}

// *mx = max |X[e]| over count entries (the bit patterns of non-negative floats order as unsigned
// integers, so an integer maximum is exact and independent of order).
__global__ void absmax_kernel(const float *X, size_t count, unsigned *mx) {
    unsigned m = 0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        m = max(m, __float_as_uint(fabsf(X[e])));
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(mx, m);
}

// Out = In / scale rounded to FP16, for scale = max |In| read from *mx (1 where that is 0).
__global__ void to_half_kernel(const float *In, size_t count, const unsigned *mx, _Float16 *Out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float m = __uint_as_float(*mx);
    Out[e] = (_Float16)(m > 0.f ? In[e] / m : In[e]);
}

// Out (k, n) = In (n, k)^T, FP32; In has row stride ld >= k.
__global__ void transpose_kernel(const float *In, int n, int k, int ld, float *Out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * k) return;
    const size_t i = e / k, c = e % k;
    Out[c * n + i] = In[i * ld + c];
}

// ---- the coarse space's device diagonalisation ----

// Phi (n, k) FP64 rounded to FP32: Out (n, kp) with zero padding columns, and Phi overwritten by the
// rounded values (the basis every later product and Galerkin matrix uses).
__global__ void round_basis_kernel(double *Phi, int n, int k, int kp, float *Out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * kp) return;
    const size_t i = e / kp;
    const int c = (int)(e % kp);
    float v = 0.f;
    if (c < k) {
        v = (float)Phi[i * k + c];
        Phi[i * k + c] = (double)v;
    }
    Out[e] = v;
}

// A := (A + A^T) / 2 for a k x k matrix (one thread per unordered pair).
__global__ void symmetrize64_kernel(double *A, int k) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i = e / k, j = e % k;
    if (i >= (size_t)k || j <= i) return;
    const double v = 0.5 * (A[i * k + j] + A[j * k + i]);
    A[i * k + j] = v;
    A[j * k + i] = v;
}

// W (k, kk) row-major: column j is eigenvector first + j of V (an eigensolver's column-major k x k
// eigenvectors) divided by the square root of its eigenvalue s[first + j].
__global__ void whiten_kernel(const double *V, const double *s, int k, int first, int kk, double *W) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)k * kk) return;
    const size_t i = e / kk;
    const int j = (int)(e % kk);
    W[e] = V[(size_t)(first + j) * k + i] / sqrt(s[first + j]);
}

// Out[c][i] = (float) In[i][c]: FP64 (n, k) -> FP32 transposed (k, ldo), ldo >= n, the padding
// columns i >= n zero (a multiple of 64 floats keeps every 64-row segment on whole cache lines).
__global__ void transpose_round_kernel(const double *In, int n, int k, long long ldo, float *Out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)ldo * k) return;
    const size_t c = e / ldo, i = e % ldo;
    Out[e] = i < (size_t)n ? (float)In[i * k + c] : 0.f;
}

// Phi (n, k) FP64 row-major from k CSC columns (zero beforehand): a thread per column.
__global__ void csc_scatter_kernel(const int *__restrict__ colptr, const int *__restrict__ rows,
                                   const double *__restrict__ vals, int k, double *__restrict__ Phi) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    for (int t = colptr[c]; t < colptr[c + 1]; ++t) Phi[(size_t)rows[t] * k + c] = vals[t];
}

// A := (double)(float)A over count entries: the FP32 basis every later product and Galerkin matrix
// uses, kept in FP64 storage.
__global__ void round_inplace_kernel(double *A, size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) A[e] = (double)(float)A[e];
}

// Y (n, k) = Rl X on the free block (X, Y FP64 row-major, k <= 4096 columns): a workgroup per row,
// each thread holding up to 16 of the row's output columns (c = threadIdx.x + 256 q), so that every
// nonzero of the row issues 16 independent loads of X's row; each output is the nonzeros' sum in
// CSR order.
__global__ __launch_bounds__(256) void csr_rows_kernel(const int *__restrict__ ptr, const int *__restrict__ col,
                                                       const double *__restrict__ val, int n,
                                                       const double *__restrict__ X, int k,
                                                       double *__restrict__ Y) {
    constexpr int Q = 16;
    const int row = blockIdx.x;
    double acc[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] = 0;
    for (int t = ptr[row]; t < ptr[row + 1]; ++t) {
        const int j = col[t];
        if (j >= n) continue;
        const double v = val[t];
        const double *xr = X + (size_t)j * k;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int c = threadIdx.x + 256 * q;
            if (c < k) acc[q] += v * xr[c];
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int c = threadIdx.x + 256 * q;
        if (c < k) Y[(size_t)row * k + c] = acc[q];
    }
}

// Partial sums for strided_sum: part[b] = sum over the entries e = b, b + nb, ... (< count) of
// A[e * stride] or its square, each workgroup's 256 threads combined by a fixed LDS tree.
__global__ __launch_bounds__(256) void strided_partial_kernel(const double *__restrict__ A, size_t count,
                                                              size_t stride, int square, double *__restrict__ part) {
    __shared__ double red[256];
    double acc = 0;
    const size_t nthreads = (size_t)gridDim.x * 256;
    for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < count; e += nthreads) {
        const double v = A[e * stride];
        acc += square ? v * v : v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// out[0] = sum over e < count of A[e * stride] (square = 0) or of its square (square = 1): one
// 1024-thread workgroup, each thread's strided partial then a fixed LDS tree (deterministic).
__global__ __launch_bounds__(1024) void strided_sum_kernel(const double *__restrict__ A, size_t count, size_t stride,
                                                           int square, double *__restrict__ out) {
    __shared__ double part[1024];
    double acc = 0;
    for (size_t e = threadIdx.x; e < count; e += 1024) {
        const double v = A[e * stride];
        acc += square ? v * v : v;
    }
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = part[0];
}

// A (k x k, column-major) := the identity.
__global__ void identity_kernel(double *A, int k) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)k * k) return;
    A[e] = (e / k == e % k) ? 1.0 : 0.0;
}

// Y += X
__global__ void vec_add_kernel(const double *X, size_t count, double *Y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) Y[e] += X[e];
}


// ---- several boards' direct solves in one captured launch ----------------------------------
__global__ void fill_int_kernel(int *a, int count, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) a[i] = v;
}

// A batch holds nz matrices Z = Rl + j w Ll of boards with at most nB free loops each, every one
// padded to nB (the identity beyond the board's own n: the factor of blockdiag(Z, I) is Z's factor
// and the identity), so that the L D L^T kernels of one frequency batch serve uneven boards.
struct BatchMat {
    const float *Ll32;        // the board's FP32 loop matrix (ldl), for the FP32 factor
    const double *Ll64;       // its FP64 loop matrix (ldl): right-hand sides, residuals, ports
    long long ldl;
    const int *rptr, *rcol;   // Rl, CSR over the board's loops
    const double *rval;
    const int *sc;            // the board's port loop columns (P of them)
    int n, P;                 // free loops, ports
    double omega;
};

// F (planes Fr, Fi, lower triangle) = Rl + j w Ll on the board's n loops, the identity beyond.
__global__ __launch_bounds__(256) void batch_form_kernel(const BatchMat *mats, int nB, float *F, long long ldf,
                                                        long long sF) {
    const BatchMat m = mats[blockIdx.z];
    float *Fr = F + blockIdx.z * sF, *Fi = Fr + (size_t)nB * ldf;
    const float w = (float)m.omega;
    for (int row = blockIdx.y; row < nB; row += gridDim.y)
        for (int c = blockIdx.x * 256 + threadIdx.x; c <= row; c += gridDim.x * 256) {
            const bool in = row < m.n;
            Fr[(size_t)row * ldf + c] = in ? 0.f : (c == row ? 1.f : 0.f);
            Fi[(size_t)row * ldf + c] = in ? w * m.Ll32[(size_t)row * m.ldl + c] : 0.f;
        }
}

__global__ void batch_form_r_kernel(const BatchMat *mats, int nB, float *F, long long ldf, long long sF) {
    const BatchMat m = mats[blockIdx.z];
    float *Fr = F + blockIdx.z * sF;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= m.n) return;
    for (int t = m.rptr[row]; t < m.rptr[row + 1]; ++t) {
        const int c = m.rcol[t];
        if (c <= row) Fr[(size_t)row * ldf + c] += (float)m.rval[t];
    }
}

// B[z] (nB, 2P) = -Z_fs of the board's ports (zero rows beyond n, zero columns beyond its P).
__global__ void batch_rhs_kernel(const BatchMat *mats, int nB, int P, double *B) {
    const BatchMat m = mats[blockIdx.z];
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nB * P) return;
    const int row = e / P, p = e % P;
    double re = 0, im = 0;
    if (row < m.n && p < m.P) {
        const int sc = m.sc[p];
        for (int t = m.rptr[row]; t < m.rptr[row + 1]; ++t)
            if (m.rcol[t] == sc) re += m.rval[t];
        im = m.omega * m.Ll64[(size_t)row * m.ldl + sc];
    }
    double *b = B + (size_t)blockIdx.z * nB * 2 * P;
    b[(size_t)row * 2 * P + p] = -re;
    b[(size_t)row * 2 * P + P + p] = -im;
}

// W = B - (Rl + j w Ll64) X on the board's rows (zero beyond n): a thread per row, Ll64's row read
// as its column (symmetric), X's rows wave-uniform.
template <int NR>
__global__ __launch_bounds__(256) void batch_residual_kernel(const BatchMat *mats, int nB, const double *X,
                                                            const double *B, double *W) {
    constexpr int P = NR / 2;
    const BatchMat m = mats[blockIdx.z];
    const size_t off = (size_t)blockIdx.z * nB * NR;
    X += off; B += off; W += off;
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= nB) return;
    double acc[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) acc[r] = 0;
    if (row < m.n) {
        const double *col = m.Ll64 + row;
        for (int c = 0; c < m.n; ++c) {
            const double l = col[(size_t)c * m.ldl];
            const double *x = X + (size_t)c * NR;
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] += l * x[r];
        }
    }
    double rx[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) rx[r] = 0;
    if (row < m.n)
        for (int t = m.rptr[row]; t < m.rptr[row + 1]; ++t) {
            const int c = m.rcol[t];
            if (c >= m.n) continue;
            const double v = m.rval[t];
#pragma unroll
            for (int r = 0; r < NR; ++r) rx[r] += v * X[(size_t)c * NR + r];
        }
    const double w = m.omega;
#pragma unroll
    for (int p = 0; p < P; ++p) {
        // (Rl + j w L)(xr + j xi) = (Rl xr - w L xi) + j (Rl xi + w L xr)
        const double zr = rx[p] - w * acc[P + p], zi = rx[P + p] + w * acc[p];
        const bool in = row < m.n;
        W[(size_t)row * NR + p] = in ? B[(size_t)row * NR + p] - zr : 0.0;
        W[(size_t)row * NR + P + p] = in ? B[(size_t)row * NR + P + p] - zi : 0.0;
    }
}

// Per matrix: the largest over its ports of |V_p| / bn_p (bn null: the squared norms |V_p|^2 into
// norm2 instead). One workgroup per matrix, a fixed-order reduction.
__global__ __launch_bounds__(256) void batch_colnorm_kernel(const BatchMat *mats, int nB, int P, const double *V,
                                                           const double *bn2, double *norm2, double *res) {
    __shared__ double part[256];
    const BatchMat m = mats[blockIdx.x];
    const int NR = 2 * P;
    V += (size_t)blockIdx.x * nB * NR;
    double worst = 0;
    for (int p = 0; p < m.P; ++p) {
        double s = 0;
        for (int row = threadIdx.x; row < m.n; row += 256) {
            const double a = V[(size_t)row * NR + p], b = V[(size_t)row * NR + P + p];
            s += a * a + b * b;
        }
        part[threadIdx.x] = s;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (threadIdx.x < st) part[threadIdx.x] += part[threadIdx.x + st];
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            if (norm2) norm2[(size_t)blockIdx.x * P + p] = part[0];
            else {
                const double r = sqrt(part[0]) / fmax(sqrt(bn2[(size_t)blockIdx.x * P + p]), 1e-300);
                worst = (isnan(r) || isnan(worst)) ? r : fmax(worst, r);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && res) res[blockIdx.x] = worst;
}

// The refinement's stopping rule per matrix (src/pcbqs/kernels.py:169-178): stop once the residual
// is at most tol, after max_steps corrections, or once a correction no longer halved it; a matrix
// still active takes the next correction. state: [0] active, [1] corrections; last: the previous
// residual.
__global__ void batch_control_kernel(int nz, const double *res, double *last, int *state, int step, double tol,
                                     int max_steps) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz || !state[2 * z]) return;
    const double w = res[z];
    if (!(w > tol) || step == max_steps || !(w <= 0.5 * last[z])) {
        state[2 * z] = 0;
        return;
    }
    last[z] = w;
    state[2 * z + 1] += 1;
}

__global__ void batch_init_kernel(int nz, double *last, int *state) {
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz) return;
    last[z] = __builtin_inf();
    state[2 * z] = 1;
    state[2 * z + 1] = 0;
}

// X += D for the matrices still active.
__global__ void batch_add_kernel(int nB, int NR, const int *state, const double *D, double *X) {
    const size_t per = (size_t)nB * NR;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= per || !state[2 * blockIdx.z]) return;
    X[blockIdx.z * per + e] += D[blockIdx.z * per + e];
}

// Zp[z] (P x P complex) = Z_ss + Z_sf X, one workgroup per (matrix, s, q).
__global__ __launch_bounds__(256) void batch_port_kernel(const BatchMat *mats, int nB, int P, const double *X,
                                                        double *Zp) {
    __shared__ double sr[256], si[256];
    const BatchMat m = mats[blockIdx.z];
    const int s = blockIdx.x / P, q = blockIdx.x % P;
    const int NR = 2 * P;
    X += (size_t)blockIdx.z * nB * NR;
    double re = 0, im = 0;
    if (s < m.P && q < m.P) {
        const int row = m.sc[s];
        for (int i = threadIdx.x; i < m.n; i += 256) {
            const double l = m.omega * m.Ll64[(size_t)row * m.ldl + i];
            const double xr = X[(size_t)i * NR + q], xi = X[(size_t)i * NR + P + q];
            re += -l * xi;
            im += l * xr;
        }
        for (int t = m.rptr[row] + threadIdx.x; t < m.rptr[row + 1]; t += 256) {
            const int c = m.rcol[t];
            if (c == m.sc[q]) re += m.rval[t];
            if (c >= m.n) continue;
            re += m.rval[t] * X[(size_t)c * NR + q];
            im += m.rval[t] * X[(size_t)c * NR + P + q];
        }
        if (threadIdx.x == 0) im += m.omega * m.Ll64[(size_t)row * m.ldl + m.sc[q]];
    }
    sr[threadIdx.x] = re; si[threadIdx.x] = im;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) { sr[threadIdx.x] += sr[threadIdx.x + st]; si[threadIdx.x] += si[threadIdx.x + st]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double *o = Zp + ((size_t)blockIdx.z * P * P + blockIdx.x) * 2;
        o[0] = sr[0];
        o[1] = si[0];
    }
}


// ---- the charge (PEEC) solve's dense system ---------------------------------------------------

// E[i][j] = (P[a_i][a_j] - P[a_i][b_j] - P[b_i][a_j] + P[b_i][b_j]) / pscale: T^T P T / pscale for
// the tree branches' incidence columns e_a - e_b.
__global__ __launch_bounds__(256) void peec_electro_kernel(const double *P, long long ldp, int nt, const int *ta,
                                                          const int *tb, double inv, double *E, long long lde) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)nt * nt) return;
    const int i = (int)(e / nt), j = (int)(e % nt);
    const int ai = ta[i], bi = tb[i], aj = ta[j], bj = tb[j];
    const double v = (P[(size_t)ai * ldp + aj] - P[(size_t)ai * ldp + bj]) - (P[(size_t)bi * ldp + aj] - P[(size_t)bi * ldp + bj]);
    E[(size_t)i * lde + j] = v * inv;
}

// S (column-major n x n complex) = Z = Rl + j w Ll over the first n loops of the handle (Z is
// symmetric: S[j][i] = Z[i][j]); then |Z|'s largest entry into *mx (bit pattern of a nonnegative
// double, compared as an integer).
__global__ __launch_bounds__(256) void peec_form_kernel(const double *Ll, long long ldl, int n, double omega, double2 *S) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const size_t i = e / n, j = e % n;
    S[j * n + i] = make_double2(0.0, omega * Ll[i * ldl + j]);
}

__global__ void peec_form_r_kernel(const int *ptr, const int *col, const double *val, int n, double2 *S) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int t = ptr[i]; t < ptr[i + 1]; ++t)
        if (col[t] < n) S[(size_t)col[t] * n + i].x += val[t];
}

__global__ __launch_bounds__(256) void peec_absmax_kernel(const double2 *S, size_t count, unsigned long long *mx) {
    __shared__ double part[256];
    double m = 0;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x) {
        const double2 z = S[e];
        m = fmax(m, sqrt(z.x * z.x + z.y * z.y));
    }
    part[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) part[threadIdx.x] = fmax(part[threadIdx.x], part[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicMax(mx, (unsigned long long)__double_as_longlong(part[0]));
}

// The scaled system (src/pcbqs/kernels.py:63-78): columns j < nl divided by zscale, the others
// multiplied by j w / pscale; E (the tree block's T^T P T / pscale) added where i, j >= nl.
__global__ __launch_bounds__(256) void peec_scale_kernel(double2 *S, int n, int nl, double zscale, double omega,
                                                        double pscale, const double *E, long long lde) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * n) return;
    const size_t j = e / n, i = e % n;                // column-major
    double2 z = S[e];
    if ((int)j < nl) {
        z.x /= zscale; z.y /= zscale;
    } else {
        const double f = omega / pscale;               // (x + j y) j f = -f y + j f x
        z = make_double2(-f * z.y, f * z.x);
        if ((int)i >= nl) z.x += E[(i - nl) * lde + (j - nl)];
    }
    S[e] = z;
}

}  // namespace pq

using namespace pq;

// ---- the loop system handle ---------------------------------------------------------------

struct pq_loops {
    int device = 0;
    int nl = 0, nf = 0, P = 0;
    long long ldl = 0;
    DevBuf Ll, Ll64, rptr, rcol, rval, sc, inv, bjperm, singular;   // Ll FP32, Ll64 FP64 (same ldl)
    DevBuf V, W, Z, X, B, R, partial, h, scal;
    hipStream_t stream = nullptr;
    int nblk = 0;
    double omega_bj = -1;
    int bj_modal = 0;             // block-Jacobi from the blocks' modes: 0 not tried, 1 ready, -1 unusable
    DevBuf bjV, bjlam, bjA;       // modes (FP32 complex, column-major per block), eigenvalues, scaled modes
    // coarse space: Phi (nf, k), Rl Phi, Ll Phi (FP64), the real k x k Galerkin matrices on the host
    int kc = 0, kp = 0;           // coarse columns; row stride of Phi
    long long ldt = 0;            // row stride of the transposed bases Phi^T, (Rl Phi)^T, (Ll Phi)^T
    bool cdiag = false;           // basis diagonalises the Galerkin matrices
    DevBuf Phi, Phit, RPhi, LPhi, Ainv, cvec, cpart, ctot, T1, T2, gpart, ppart;   // Phit, RPhi, LPhi transposed (k, ldt)
    // FP16 copies of the four bases for the solve's restriction and prolongation, each divided by
    // its largest magnitude (hscale: Phi, Phi^T, (Rl Phi)^T, (Ll Phi)^T), when coarse_half is set
    DevBuf Phih, Phith, RPhih, LPhih, hmax;
    double hscale[4] = {1, 1, 1, 1};
    bool coarse_half = false;
    std::vector<double> Rc, Lc;
    double omega_c = -1;
    DevBuf Hd;                    // Arnoldi coefficients per step (device)
    double *Hh = nullptr;         // their pinned host copy
    size_t Hh_bytes = 0;
    // the last solve
    int iterations = 0, restarts = 0;
    double residual = 0, seconds = 0, matvec_ms = 0;
    long long matvecs = 0;
    DevBuf omega_dev;             // the solve's angular frequency, read by the kernels
    int graphs_on = -1;           // capture the Arnoldi steps: -1 = PCBQS_AMD_GRAPHS ("1": on)
    // Arnoldi steps captured as graphs (one per step index), valid while `graph_key` holds
    std::vector<hipGraphExec_t> graphs;
    std::vector<const void *> graph_key;
    // autotuned column-chunk width of the operator product per column count (first use per NR)
    std::map<int, int> gemv_chunk, gemv_rows;
    std::map<int, float> gemv_us;
    // autotuned coarse kernels per column count (rows per restriction block, prolongation chunks
    // and loads in flight)
    std::map<int, int> crows, psplit, pdepth;
    // the direct solve's L D L^T factor (planes of the factor, block inverses, panel products) and
    // the frequency it was formed at
    DevBuf Fldl, Xldl, Wldl;
    hipStream_t aux = nullptr;    // the factorisation's second stream and its two events
    hipEvent_t ev_ready = nullptr, ev_done = nullptr;
    long long ldf = 0;
    double omega_ldl = -1;
    double factor_ms = 0;
    // factors of several frequencies formed and factored together (pq_loops_factor_batch): matrix
    // b at the frequency batch_omega[b], strides as LdlFactor's
    DevBuf Fbat, Xbat, Wbat;
    std::vector<double> batch_omega;
    double batch_ms = 0;
    std::vector<int> batch_bad;   // per matrix of the batch: a zero or non-finite FP32 pivot
    // the FP64 fallback of the direct solve (an LU with partial pivoting of Z_ff in FP64 through
    // rocSOLVER, loaded on first use): frequencies of the last direct call it solved
    int fallbacks = 0;
    void *blas = nullptr;         // rocblas_handle
    DevBuf gpart64;               // the FP64 product's partials (apart from the captured steps' buffers)
};

namespace {

template <class F> void dispatch_nr(int NR, F &&f) {
    switch (NR) {
    case 2: f(std::integral_constant<int, 2>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 6: f(std::integral_constant<int, 6>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    case 10: f(std::integral_constant<int, 10>()); break;
    case 12: f(std::integral_constant<int, 12>()); break;
    case 14: f(std::integral_constant<int, 14>()); break;
    case 16: f(std::integral_constant<int, 16>()); break;
    case 32: f(std::integral_constant<int, 32>()); break;
    default: throw Error(PQ_EINVAL, "port count must be 1 to 8 or 16");
    }
}

// W = (Rl + j w Ll) X on the free block.
int default_chunk(int n) { return std::min(GEMV_CHUNK, std::max(64, (n + GEMV_SPLIT - 1) / GEMV_SPLIT)); }

// The operator product's launch shape: chunk 0 is the symmetric kernel; chunk > 0 the panel
// kernel with `rows` rows per block and `chunk` columns per partial.
struct OpShape {
    int chunk, rows;
};

OpShape gemv_shape(const pq_loops *L, int NR) {
    auto it = L->gemv_chunk.find(NR);
    if (it == L->gemv_chunk.end()) return {default_chunk(L->nf), GEMV_ROWS};
    auto r = L->gemv_rows.find(NR);
    return {it->second, r != L->gemv_rows.end() ? r->second : GEMV_ROWS};
}

int gemv_chunk(const pq_loops *L, int NR) { return gemv_shape(L, NR).chunk; }

// Number of partial sums the operator product writes per row: the column chunks of panel_kernel,
// or the 256-row blocks of sym_panel_kernel (chunk 0).
int operator_splits(int n, int chunk) {
    return chunk > 0 ? (n + chunk - 1) / chunk : (n + SYM_T - 1) / SYM_T;
}

// part = the Ll X partials (operator_splits of them) on the stream.
template <int N> void operator_partials(const pq_loops *L, OpShape sh, const double *X, double *part) {
    const int n = L->nf;
    if (sh.chunk > 0 && sh.rows == 64) {
        hipLaunchKernelGGL(panel4_kernel<N>, dim3((n + 63) / 64, (n + sh.chunk - 1) / sh.chunk), dim3(256), 0,
                           L->stream, L->Ll.as<float>(), L->ldl, n, n, X, sh.chunk, part);
    } else if (sh.chunk > 0) {
        hipLaunchKernelGGL(panel_kernel<N>, dim3((n + sh.rows - 1) / sh.rows, (n + sh.chunk - 1) / sh.chunk),
                           dim3(sh.rows), 0, L->stream, L->Ll.as<float>(), L->ldl, n, n, X, sh.chunk, part);
    } else {
        if constexpr (N <= 16) {
            const int nb = (n + SYM_T - 1) / SYM_T;
            hipLaunchKernelGGL(sym_panel_kernel<N>, dim3(nb * (nb + 1) / 2), dim3(256), 0, L->stream,
                               L->Ll.as<float>(), L->ldl, n, nb, X, part);
        } else {
            throw Error(PQ_EINVAL, "symmetric operator product: at most 16 real columns");
        }
    }
}

// PCBQS_AMD_OPERATOR pins the operator product's kernel: "panel" (1) or "symmetric" (2); unset
// or anything else (0) lets the autotuner choose.
int operator_mode() {
    const char *v = std::getenv("PCBQS_AMD_OPERATOR");
    if (!v) return 0;
    if (!std::strcmp(v, "panel")) return 1;
    if (!std::strcmp(v, "symmetric")) return 2;
    return 0;
}

// W = (Rl + j w Ll) X with the Ll X partials in `part` (operator_splits(n, chunk) of them).
template <int N> void operator_launch(const pq_loops *L, OpShape sh, const double *X, double *W, double *part,
                                      const double *omega_dev) {
    const int n = L->nf, nsplit = operator_splits(n, sh.chunk);
    operator_partials<N>(L, sh, X, part);
    if constexpr ((N & (N - 1)) == 0) {
        hipLaunchKernelGGL(operator_epilogue_kernel<N>, dim3((n + 3) / 4), dim3(256), 0, L->stream, L->rptr.as<int>(),
                           L->rcol.as<int>(), L->rval.as<double>(), n, X, part, nsplit, omega_dev, W);
    } else {
        hipLaunchKernelGGL(spmv_kernel<N>, dim3((n + 3) / 4), dim3(256), 0, L->stream, L->rptr.as<int>(),
                           L->rcol.as<int>(), L->rval.as<double>(), n, X, W);
        hipLaunchKernelGGL(gemv_epilogue_kernel<N>, dim3((n * (N / 2) + 255) / 256), dim3(256), 0, L->stream, part,
                           nsplit, n, omega_dev, W);
    }
}

// The autotuner of the operator product. Candidates: the symmetric upper-triangle kernel (chunk 0,
// up to 16 real columns), panel_kernel over the whole matrix with 128 or 256 rows per block and a
// column-chunk width, and panel4_kernel (rows 64: four waves per 64-row block) with a chunk width;
// together they set the grid, the loads in flight and the partial sums the epilogue reads. Each
// candidate runs once to warm up and three times under HIP events on an input of zeros (the whole
// product: partials, resistive part and epilogue), each time after a cache-evicting fill; the
// fastest is kept for this handle and column count.
// Autotuned launch shapes are shared by every handle of the process with the same device and
// problem shape: the first handle to need one tunes it (under the lock) and later handles take its
// choice, so that two solvers of one board in one process sum in the same order and give the same
// bits (the chunk and row-block choices set the order of the partial sums).
std::mutex tune_mutex;
std::map<std::tuple<int, int, int, int>, std::tuple<int, int, float>> gemv_cache;        // (dev, n, NR, mode)
std::map<std::tuple<int, int, int, int, int>, std::tuple<int, int, int>> coarse_cache;  // (dev, n, k, NR, half)

void tune_gemv(pq_loops *L, int NR) {
    const int n = L->nf;
    const int mode = operator_mode();
    if (L->gemv_chunk.count(NR) || n == 0) return;
    // small systems keep the default shape unless the symmetric kernel is pinned
    if (n < 1024 && !(mode == 2 && NR <= 16)) return;
    std::lock_guard<std::mutex> lock(tune_mutex);
    const auto key = std::make_tuple(L->device, n, NR, mode);
    if (auto hit = gemv_cache.find(key); hit != gemv_cache.end()) {
        L->gemv_chunk[NR] = std::get<0>(hit->second);
        L->gemv_rows[NR] = std::get<1>(hit->second);
        L->gemv_us[NR] = std::get<2>(hit->second);
        return;
    }
    std::vector<OpShape> cand;
    if (NR <= 16 && mode != 1) cand.push_back({0, SYM_T});
    if (mode != 2 || cand.empty()) {
        for (int rows : {256, 128})
            for (int c : {128, 192, 256, 320, 448, 640, 960})
                if ((n + c - 1) / c <= 128 && c < n) cand.push_back({c, rows});
        for (int c : {256, 320, 448, 640, 960, 1280, 1920})      // four waves per 64-row block
            if ((n + c - 1) / c <= 128 && c < n) cand.push_back({c, 64});
    }
    if (cand.empty()) return;
    int maxsplit = 0;
    for (auto c : cand) maxsplit = std::max(maxsplit, operator_splits(n, c.chunk));
    DevBuf X, W, part, om, flush;
    constexpr size_t FLUSH = (size_t)512 << 20;
    flush.reserve(FLUSH);
    X.reserve((size_t)n * NR * 8);
    W.reserve((size_t)n * NR * 8);
    om.reserve(8);
    part.reserve((size_t)maxsplit * n * NR * 8);
    PQ_HIP(hipMemsetAsync(X.p, 0, (size_t)n * NR * 8, L->stream));
    PQ_HIP(hipMemsetAsync(om.p, 0, 8, L->stream));
    hipEvent_t e0, e1;
    PQ_HIP(hipEventCreate(&e0));
    PQ_HIP(hipEventCreate(&e1));
    OpShape best = cand[0];
    float best_us = 1e30f;
    for (auto c : cand) {
        auto launch = [&]() {
            dispatch_nr(NR, [&](auto nr) {
                constexpr int N = decltype(nr)::value;
                operator_launch<N>(L, c, X.as<double>(), W.as<double>(), part.as<double>(), om.as<double>());
            });
        };
        // each timed product starts with Ll out of the caches (a 512 MiB fill in between), as in a
        // solve, where the Krylov vectors and preconditioner data pass between two products: timed
        // back to back, part of Ll stays in the 256 MB last-level cache and favours other shapes
        launch();
        float total = 0;
        for (int r = 0; r < 3; ++r) {
            PQ_HIP(hipMemsetAsync(flush.p, r, FLUSH, L->stream));
            PQ_HIP(hipEventRecord(e0, L->stream));
            launch();
            PQ_HIP(hipEventRecord(e1, L->stream));
            PQ_HIP(hipEventSynchronize(e1));
            float ms = 0;
            PQ_HIP(hipEventElapsedTime(&ms, e0, e1));
            total += ms;
        }
        const float us = total * 1000.f / 3.f;
        if (us < best_us) { best_us = us; best = c; }
    }
    PQ_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    L->gemv_chunk[NR] = best.chunk;
    L->gemv_rows[NR] = best.rows;
    L->gemv_us[NR] = best_us;
    gemv_cache[key] = std::make_tuple(best.chunk, best.rows, best_us);
}

void apply_operator(pq_loops *L, const double *X, double *W, int P) {
    const int NR = 2 * P, n = L->nf;
    const OpShape sh = gemv_shape(L, NR);
    L->gpart.reserve((size_t)operator_splits(n, sh.chunk) * n * NR * 8);
    dispatch_nr(NR, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        operator_launch<N>(L, sh, X, W, L->gpart.as<double>(), L->omega_dev.as<double>());
    });
    PQ_HIP(hipGetLastError());
}

// W = (Rl + j w Ll64) X with the FP64 loop matrix: the system's own product, for right-hand
// sides' residuals (GMRES restarts and acceptance, the direct solve's refinement). The four-wave
// panel product over FP64 rows and the epilogue; twice the FP32 product's bytes.
void apply_operator64(pq_loops *L, const double *X, double *W, int P) {
    const int NR = 2 * P, n = L->nf;
    // a shape fixed by n alone (four-wave 64-row blocks, about eight column chunks), never the
    // autotuner's: the FP64 product sets the residuals, right-hand sides and refinements, whose
    // summation order must not depend on which shape won a timing race
    const OpShape sh = {std::max(256, (n + 8 * 64 - 1) / (8 * 64) * 64), 64};
    const int nsplit = operator_splits(n, sh.chunk);
    L->gpart64.reserve((size_t)nsplit * n * NR * 8);
    double *part = L->gpart64.as<double>();
    const double *omega_dev = L->omega_dev.as<double>();
    dispatch_nr(NR, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        if (sh.rows == 64)
            hipLaunchKernelGGL((panel4_kernel<N, double>), dim3((n + 63) / 64, (n + sh.chunk - 1) / sh.chunk),
                               dim3(256), 0, L->stream, L->Ll64.as<double>(), L->ldl, n, n, X, sh.chunk, part);
        else
            hipLaunchKernelGGL((panel_kernel<N, double>), dim3((n + sh.rows - 1) / sh.rows, (n + sh.chunk - 1) / sh.chunk),
                               dim3(sh.rows), 0, L->stream, L->Ll64.as<double>(), L->ldl, n, n, X, sh.chunk, part);
        if constexpr ((N & (N - 1)) == 0) {
            hipLaunchKernelGGL(operator_epilogue_kernel<N>, dim3((n + 3) / 4), dim3(256), 0, L->stream,
                               L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, X, part, nsplit,
                               omega_dev, W);
        } else {
            hipLaunchKernelGGL(spmv_kernel<N>, dim3((n + 3) / 4), dim3(256), 0, L->stream, L->rptr.as<int>(),
                               L->rcol.as<int>(), L->rval.as<double>(), n, X, W);
            hipLaunchKernelGGL(gemv_epilogue_kernel<N>, dim3((n * (N / 2) + 255) / 256), dim3(256), 0, L->stream,
                               part, nsplit, n, omega_dev, W);
        }
    });
    PQ_HIP(hipGetLastError());
}

void apply_precond(pq_loops *L, int kind, const double *X, double *Y, int P);

void apply_bj(pq_loops *L, const double *X, double *Y, int P, const double *add = nullptr);

// Column chunks of the prolongation (64-row blocks of four waves) before the autotuner has run:
// enough blocks to fill the chip (about 1024), at most 16, at least 64 columns per chunk.
int prolong_splits(int n, int k) {
    const int rb = (n + 63) / 64;
    return std::max(1, std::min({16, (1024 + rb - 1) / rb, (k + 63) / 64}));
}

int restrict_rows(const pq_loops *L, int NR) {
    auto it = L->crows.find(NR);
    return it != L->crows.end() ? it->second : cr_rows(NR);
}

int prolong_split(const pq_loops *L, int NR) {
    auto it = L->psplit.find(NR);
    return it != L->psplit.end() ? it->second : prolong_splits(L->nf, L->kc);
}

int prolong_depth(const pq_loops *L, int NR) {
    auto it = L->pdepth.find(NR);
    return it != L->pdepth.end() ? it->second : 8;
}

// Launch the restriction Phi^T X into `rows`-row chunk partials; returns the chunk count.
// half: read the FP16 copy of Phi (and scale by its largest magnitude) instead of Phi.
template <int N> int restrict_launch(const pq_loops *L, const float *Phi, int k, const double *X, int rows,
                                     double *part, bool half = false) {
    const int n = L->nf, nchunk = (n + rows - 1) / rows;
    if constexpr (N <= 16) {
        if (half)
            hipLaunchKernelGGL((coarse_restrict4_kernel<N, _Float16>), dim3(nchunk, (k + 255) / 256), dim3(256), 0,
                               L->stream, L->Phih.as<_Float16>(), k, L->kp, X, n, rows, L->hscale[0], part);
        else
            hipLaunchKernelGGL((coarse_restrict4_kernel<N, float>), dim3(nchunk, (k + 255) / 256), dim3(256), 0,
                               L->stream, Phi, k, L->kp, X, n, rows, 1.0, part);
        return nchunk;
    }
    if (half)
        hipLaunchKernelGGL((coarse_restrict_kernel<N, _Float16>), dim3(nchunk, (k + 63) / 64), dim3(256), 0, L->stream,
                           L->Phih.as<_Float16>(), k, L->kp, X, n, rows, L->hscale[0], part);
    else
        hipLaunchKernelGGL((coarse_restrict_kernel<N, float>), dim3(nchunk, (k + 63) / 64), dim3(256), 0, L->stream,
                           Phi, k, L->kp, X, n, rows, 1.0, part);
    return nchunk;
}

// Launch the prolongation of L->cvec in about `ps` column chunks, `depth` (8 or 16) columns of each
// basis in flight per lane; returns the chunk count.
template <int N> int prolong_launch(const pq_loops *L, int ps, int depth, const double *omega_dev, double *PY,
                                    double *PA) {
    const int n = L->nf, k = L->kc, pc = (k + ps - 1) / ps, nsplit = (k + pc - 1) / pc;
    const dim3 grid((n + 63) / 64, nsplit);
    const double *sc = L->hscale;
    if (L->coarse_half) {
        const _Float16 *b0 = L->Phith.as<_Float16>(), *b1 = L->RPhih.as<_Float16>(), *b2 = L->LPhih.as<_Float16>();
        if (depth == 16)
            hipLaunchKernelGGL((prolong_kernel<N, 16, _Float16>), grid, dim3(256), 0, L->stream, b0, b1, b2, n, L->ldt, k,
                               L->cvec.as<double>(), pc, omega_dev, sc[1], sc[2], sc[3], PY, PA);
        else
            hipLaunchKernelGGL((prolong_kernel<N, 8, _Float16>), grid, dim3(256), 0, L->stream, b0, b1, b2, n, L->ldt, k,
                               L->cvec.as<double>(), pc, omega_dev, sc[1], sc[2], sc[3], PY, PA);
        return nsplit;
    }
    const float *b0 = L->Phit.as<float>(), *b1 = L->RPhi.as<float>(), *b2 = L->LPhi.as<float>();
    if (depth == 16)
        hipLaunchKernelGGL((prolong_kernel<N, 16, float>), grid, dim3(256), 0, L->stream, b0, b1, b2, n, L->ldt, k,
                           L->cvec.as<double>(), pc, omega_dev, 1.0, 1.0, 1.0, PY, PA);
    else
        hipLaunchKernelGGL((prolong_kernel<N, 8, float>), grid, dim3(256), 0, L->stream, b0, b1, b2, n, L->ldt, k,
                           L->cvec.as<double>(), pc, omega_dev, 1.0, 1.0, 1.0, PY, PA);
    return nsplit;
}

// The coarse kernels' autotuner (first use per column count and coarse space): rows per block of
// the restriction with its chunk sum (more blocks against more partial sums) and column chunks of
// the prolongation, each candidate run once to warm up and three times under HIP events on this
// handle's bases with inputs of zeros; the fastest is kept. Before each timed run a 512 MiB fill
// evicts the bases from the caches (the last-level cache holds 256 MB, as much as the bases of a
// large coarse space), as the operator product does between two applications in a solve.
void tune_coarse(pq_loops *L, int NR) {
    const int n = L->nf, k = L->kc;
    if (k == 0 || n < 1024 || L->crows.count(NR)) return;
    std::lock_guard<std::mutex> lock(tune_mutex);
    const auto key = std::make_tuple(L->device, n, k, NR, (int)L->coarse_half);
    if (auto hit = coarse_cache.find(key); hit != coarse_cache.end()) {
        L->crows[NR] = std::get<0>(hit->second);
        L->psplit[NR] = std::get<1>(hit->second);
        L->pdepth[NR] = std::get<2>(hit->second);
        return;
    }
    const std::vector<int> rows_c = {128, 256, 512, 1024};
    std::vector<int> split_c;
    for (int ps : {1, 2, 3, 4, 6, 8, 12, 16})
        if (ps == 1 || ps <= (k + 63) / 64) split_c.push_back(ps);
    if (split_c.empty()) split_c.push_back(1);
    DevBuf X, part, PY, PA, om, flush, cv, dg;
    constexpr size_t FLUSH = (size_t)512 << 20;
    flush.reserve(FLUSH);
    cv.reserve((size_t)k * NR * 8);
    dg.reserve((size_t)k * 2 * 8);
    PQ_HIP(hipMemsetAsync(dg.p, 0, (size_t)k * 2 * 8, L->stream));
    X.reserve((size_t)n * NR * 8);
    part.reserve((size_t)((n + 127) / 128) * k * NR * 8);
    PY.reserve((size_t)split_c.back() * n * NR * 8);
    PA.reserve((size_t)split_c.back() * n * NR * 8);
    om.reserve(8);
    L->cvec.reserve((size_t)k * NR * 8);
    PQ_HIP(hipMemsetAsync(X.p, 0, (size_t)n * NR * 8, L->stream));
    PQ_HIP(hipMemsetAsync(L->cvec.p, 0, (size_t)k * NR * 8, L->stream));
    PQ_HIP(hipMemsetAsync(om.p, 0, 8, L->stream));
    hipEvent_t e0, e1;
    PQ_HIP(hipEventCreate(&e0));
    PQ_HIP(hipEventCreate(&e1));
    auto timed = [&](auto &&launch) {
        launch();
        float total = 0;
        for (int r = 0; r < 3; ++r) {
            PQ_HIP(hipMemsetAsync(flush.p, r, FLUSH, L->stream));
            PQ_HIP(hipEventRecord(e0, L->stream));
            launch();
            PQ_HIP(hipEventRecord(e1, L->stream));
            PQ_HIP(hipEventSynchronize(e1));
            float ms = 0;
            PQ_HIP(hipEventElapsedTime(&ms, e0, e1));
            total += ms;
        }
        return total / 3.f;
    };
    int best_rows = rows_c[0], best_split = split_c[0], best_depth = 8;
    float best = 1e30f;
    dispatch_nr(NR, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        for (int rows : rows_c) {
            const float t = timed([&] {
                const int nchunk = restrict_launch<N>(L, L->Phi.as<float>(), k, X.as<double>(), rows, part.as<double>(),
                                                      L->coarse_half);
                hipLaunchKernelGGL(coarse_sum_diag_kernel, dim3((k * (N / 2) + 3) / 4), dim3(256), 0, L->stream,
                                   part.as<double>(), nchunk, k, N / 2, dg.as<double>(), cv.as<double>());
            });
            if (t < best) { best = t; best_rows = rows; }
        }
        best = 1e30f;
        for (int depth : {8, 16})
            for (int ps : split_c) {
                const float t = timed([&] {
                    prolong_launch<N>(L, ps, depth, om.as<double>(), PY.as<double>(), PA.as<double>());
                });
                if (t < best) { best = t; best_split = ps; best_depth = depth; }
            }
    });
    PQ_HIP(hipGetLastError());
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    L->crows[NR] = best_rows;
    L->psplit[NR] = best_split;
    L->pdepth[NR] = best_depth;
    coarse_cache[key] = std::make_tuple(best_rows, best_split, best_depth);
}

// c = Ac^-1 Phi^T X, then T1 = X - A Phi c and T2 = Phi c.
void coarse_correction(pq_loops *L, const double *X, int P) {
    const int n = L->nf, NR = 2 * P, k = L->kc;
    dispatch_nr(NR, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        const int nchunk = restrict_launch<N>(L, L->Phi.as<float>(), k, X, restrict_rows(L, N), L->cpart.as<double>(),
                                              L->coarse_half);
        if (L->cdiag) {
            hipLaunchKernelGGL(coarse_sum_diag_kernel, dim3((k * P + 3) / 4), dim3(256), 0, L->stream,
                               L->cpart.as<double>(), nchunk, k, P, L->Ainv.as<double>(), L->cvec.as<double>());
        } else {
            hipLaunchKernelGGL(coarse_sum_kernel, dim3((k * N + 255) / 256), dim3(256), 0, L->stream,
                               L->cpart.as<double>(), nchunk, k * N, L->ctot.as<double>());
            hipLaunchKernelGGL(coarse_solve_kernel, dim3((k * P + 3) / 4), dim3(256), 0, L->stream,
                               L->ctot.as<double>(), L->Ainv.as<double>(), k, P, L->cvec.as<double>());
        }
        // prolongation: the three bases in one pass, then the chunk sums
        const size_t count = (size_t)n * N;
        double *PY = L->ppart.as<double>();
        const int ps = prolong_split(L, N);
        double *PA = PY + (size_t)ps * count;
        const int nsplit = prolong_launch<N>(L, ps, prolong_depth(L, N), L->omega_dev.as<double>(), PY, PA);
        hipLaunchKernelGGL(prolong_epilogue_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream, PY,
                           PA, nsplit, count, X, L->T1.as<double>(), L->T2.as<double>());
    });
    PQ_HIP(hipGetLastError());
}

void apply_precond(pq_loops *L, int kind, const double *X, double *Y, int P) {
    const int n = L->nf, NR = 2 * P;
    if (kind == 2) {
        // deflated two-level: c = Ac^-1 Phi^T x; y = Phi c + BJ(x - A Phi c)
        coarse_correction(L, X, P);
        apply_bj(L, L->T1.as<double>(), Y, P, L->T2.as<double>());
        return;
    }
    if (kind == 0) {
        PQ_HIP(hipMemcpyAsync(Y, X, (size_t)n * NR * 8, hipMemcpyDeviceToDevice, L->stream));
        return;
    }
    apply_bj(L, X, Y, P);
}

void apply_bj(pq_loops *L, const double *X, double *Y, int P, const double *add) {
    dispatch_nr(2 * P, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        hipLaunchKernelGGL(bj_apply16_kernel<N>, dim3(L->nblk, BS / 16), dim3(256), 0, L->stream,
                           L->inv.as<float>(), L->nf, X, add, Y);
    });
    PQ_HIP(hipGetLastError());
}

// Inverse of a complex k x k matrix (row-major, interleaved re/im) by Gauss-Jordan with partial
// pivoting; a pivot below eps * max|a| is dropped (its row and column left zero): the coarse
// problem's near-null directions are then not corrected rather than amplified.
void small_inverse(std::vector<cplx> &a, int k, std::vector<cplx> &inv) {
    double amax = 0;
    for (auto &v : a) amax = std::max(amax, std::abs(v));
    const double eps = 1e-12 * std::max(amax, 1e-300);
    std::vector<int> perm(k);
    std::vector<bool> dropped(k, false);
    for (int c = 0; c < k; ++c) {
        int best = c;
        double bv = std::abs(a[(size_t)c * k + c]);
        for (int i = c + 1; i < k; ++i)
            if (std::abs(a[(size_t)i * k + c]) > bv) { bv = std::abs(a[(size_t)i * k + c]); best = i; }
        perm[c] = best;
        if (best != c)
            for (int j = 0; j < k; ++j) std::swap(a[(size_t)c * k + j], a[(size_t)best * k + j]);
        if (bv <= eps) {
            dropped[c] = true;
            for (int j = 0; j < k; ++j) a[(size_t)c * k + j] = 0;
            for (int i = 0; i < k; ++i) a[(size_t)i * k + c] = 0;
            continue;
        }
        const cplx piv = a[(size_t)c * k + c];
        a[(size_t)c * k + c] = 1;
        for (int j = 0; j < k; ++j) a[(size_t)c * k + j] /= piv;
        for (int i = 0; i < k; ++i) {
            if (i == c) continue;
            const cplx f = a[(size_t)i * k + c];
            if (f == cplx(0)) continue;
            a[(size_t)i * k + c] = 0;
            for (int j = 0; j < k; ++j) a[(size_t)i * k + j] -= f * a[(size_t)c * k + j];
        }
    }
    for (int c = k - 1; c >= 0; --c)
        if (perm[c] != c)
            for (int i = 0; i < k; ++i) std::swap(a[(size_t)i * k + c], a[(size_t)i * k + perm[c]]);
    inv = a;
}

void coarse_at(pq_loops *L, double omega) {
    if (L->omega_c == omega) return;
    const int k = L->kc;
    if (L->cdiag) {
        // a basis that diagonalises both Galerkin matrices: Ac^-1 = diag(1 / (Rc_aa + j w Lc_aa))
        std::vector<double> d(2 * (size_t)k, 0.0);
        double amax = 0;
        for (int a = 0; a < k; ++a)
            amax = std::max(amax, std::abs(cplx(L->Rc[(size_t)a * k + a], omega * L->Lc[(size_t)a * k + a])));
        for (int a = 0; a < k; ++a) {
            const cplx z(L->Rc[(size_t)a * k + a], omega * L->Lc[(size_t)a * k + a]);
            if (std::abs(z) <= 1e-12 * amax) continue;
            const cplx v = 1.0 / z;
            d[2 * a] = v.real(); d[2 * a + 1] = v.imag();
        }
        L->Ainv.reserve(d.size() * 8);
        PQ_HIP(hipMemcpyAsync(L->Ainv.p, d.data(), d.size() * 8, hipMemcpyHostToDevice, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
        L->omega_c = omega;
        return;
    }
    std::vector<cplx> a((size_t)k * k), inv;
    for (size_t e = 0; e < a.size(); ++e) a[e] = cplx(L->Rc[e], omega * L->Lc[e]);
    small_inverse(a, k, inv);
    std::vector<double> flat(2 * inv.size());
    for (size_t e = 0; e < inv.size(); ++e) { flat[2 * e] = inv[e].real(); flat[2 * e + 1] = inv[e].imag(); }
    L->Ainv.reserve(flat.size() * 8);
    PQ_HIP(hipMemcpyAsync(L->Ainv.p, flat.data(), flat.size() * 8, hipMemcpyHostToDevice, L->stream));
    PQ_HIP(hipStreamSynchronize(L->stream));
    L->omega_c = omega;
}

// The solve's angular frequency into device memory (the stream is idle between solves).
void set_omega(pq_loops *L, double omega) {
    L->omega_dev.reserve(8);
    PQ_HIP(hipMemcpy(L->omega_dev.p, &omega, 8, hipMemcpyHostToDevice));
}

void drop_graphs(pq_loops *L) {
    for (auto &g : L->graphs)
        if (g) (void)hipGraphExecDestroy(g);
    L->graphs.clear();
    L->graph_key.clear();
}

// Graphs of the Arnoldi steps replay fixed kernel arguments: they stay valid while every buffer
// address, the column count, the preconditioner, the coarse space and the tuned chunk are those
// they were captured with.
void check_graphs(pq_loops *L, int P, int precond, int m) {
    auto cv = [](long long v) { return reinterpret_cast<const void *>(static_cast<intptr_t>(v)); };
    std::vector<const void *> key = {cv(P), cv(precond), cv(m), cv(L->nf), cv(L->kc), cv(L->cdiag),
                                     cv(gemv_chunk(L, 2 * P)), cv(gemv_shape(L, 2 * P).rows), L->V.p, L->W.p, L->Z.p, L->partial.p, L->Hd.p,
                                     L->inv.p, L->T1.p, L->T2.p, L->cpart.p, L->cvec.p, L->ppart.p, L->gpart.p,
                                     L->ctot.p, L->Phi.p, L->Phit.p, L->RPhi.p, L->LPhi.p, L->Ainv.p,
                                     L->Phih.p, L->Phith.p, L->RPhih.p, L->LPhih.p, cv(L->coarse_half),
                                     L->omega_dev.p, L->Ll.p, L->rptr.p, L->Hh};
    if (key != L->graph_key) {
        drop_graphs(L);
        L->graph_key = key;
    }
    if ((int)L->graphs.size() < m) L->graphs.resize(m, nullptr);
}

// The relative residual estimate above which GMRES steps use the FP64 loop matrix
// (PCBQS_AMD_RELAX; 1 or more: never).
double relax_threshold() {
    const char *v = std::getenv("PCBQS_AMD_RELAX");
    const double f = v ? std::atof(v) : 0.01;
    return f > 0 ? f : 0.01;
}

double inner_tol_factor() {
    const char *v = std::getenv("PCBQS_AMD_INNER_TOL");
    const double f = v ? std::atof(v) : 1.0;
    return f > 0 && f <= 1 ? f : 1.0;
}

bool graphs_enabled() {
    const char *v = std::getenv("PCBQS_AMD_GRAPHS");
    return v && v[0] == '1';
}

// The coarse matrix at omega and the coarse work buffers for P columns.
void prepare_coarse(pq_loops *L, double omega, int P) {
    const int n = L->nf, NR = 2 * P;
    coarse_at(L, omega);
    tune_coarse(L, NR);
    const int rows = restrict_rows(L, NR), nchunk = (n + rows - 1) / rows;
    L->cpart.reserve((size_t)nchunk * L->kc * NR * 8);
    L->cvec.reserve((size_t)L->kc * P * 2 * 8);
    L->ppart.reserve((size_t)2 * prolong_split(L, NR) * n * NR * 8);
    L->ctot.reserve((size_t)L->kc * NR * 8);
    L->T1.reserve((size_t)n * NR * 8); L->T2.reserve((size_t)n * NR * 8);
}

// The free-loop right-hand sides B = -Z_fs (n_free, 2P) of the port columns into L->B.
void port_rhs(pq_loops *L, double omega, int P, const int32_t *port_cols) {
    const int n = L->nf;
    L->sc.reserve(P * 4);
    L->B.reserve((size_t)n * 2 * P * 8);
    PQ_HIP(hipMemcpyAsync(L->sc.p, port_cols, P * 4, hipMemcpyHostToDevice, L->stream));
    hipLaunchKernelGGL(rhs_kernel, dim3((n * P + 255) / 256), dim3(256), 0, L->stream, L->Ll64.as<double>(), L->ldl,
                       L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, L->sc.as<int>(), P, omega,
                       L->B.as<double>());
    PQ_HIP(hipGetLastError());
}

// h[i][p] for i < nv: conj(V_i) . W into device memory (partials, then their fixed-order sum).
void dots_launch(pq_loops *L, const double *V, size_t vstride, int nv, const double *W, int P, double *h,
                 double *h_host = nullptr) {
    const int nblk = std::min(RED_BLOCKS, (L->nf + 255) / 256);
    dispatch_nr(2 * P, [&](auto nr) {
        constexpr int N = decltype(nr)::value;
        hipLaunchKernelGGL(dots_partial_kernel<N / 2>, dim3(nblk, nv), dim3(256), 0, L->stream, V, vstride, nv, W,
                           L->nf, L->partial.as<double>());
    });
    hipLaunchKernelGGL(dots_reduce_kernel, dim3((nv * 2 * P + 3) / 4), dim3(256), 0, L->stream, L->partial.as<double>(),
                       nblk, nv, 2 * P, h, h_host);
    PQ_HIP(hipGetLastError());
}

// h[i][p] for i < nv: conj(V_i) . W, synchronously copied to host.
void dots(pq_loops *L, const double *V, size_t vstride, int nv, const double *W, int P, double *h_host) {
    dots_launch(L, V, vstride, nv, W, P, L->h.as<double>());
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
    const size_t count = (size_t)L->nf * P * 4;
    hipLaunchKernelGGL(vec_update_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream, V, vstride,
                       nv, L->h.as<double>(), sign, L->nf, P, W);
    PQ_HIP(hipGetLastError());
}


// Zp = Z_ss + Z_sf X: the port loops' own block from Rl and the stored Ll rows, and the coupling
// to the free loops' solution X (n, 2P) on the device (src/pcbqs/mqs_ports.py:52-69).
void port_impedance(pq_loops *L, double omega, int P, const int32_t *port_cols, const double *X, double *Zp) {
    const int n = L->nf;
    DevBuf red;
    red.reserve((size_t)P * P * 2 * 8);
    if (n > 0) {
        hipLaunchKernelGGL(port_reduce_kernel, dim3(P * P), dim3(256), 0, L->stream, L->Ll64.as<double>(), L->ldl,
                           L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, L->sc.as<int>(), P, omega, X,
                           red.as<double>());
        PQ_HIP(hipGetLastError());
    } else {
        PQ_HIP(hipMemsetAsync(red.p, 0, (size_t)P * P * 2 * 8, L->stream));
    }
    std::vector<double> zsf((size_t)P * P * 2);
    PQ_HIP(hipMemcpyAsync(zsf.data(), red.p, zsf.size() * 8, hipMemcpyDeviceToHost, L->stream));
    // Z_ss from the stored rows
    std::vector<double> llss((size_t)P * P);
    for (int s = 0; s < P; ++s)
        for (int q = 0; q < P; ++q)
            PQ_HIP(hipMemcpyAsync(&llss[s * P + q], L->Ll64.as<double>() + (size_t)port_cols[s] * L->ldl + port_cols[q],
                                  8, hipMemcpyDeviceToHost, L->stream));
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
            Zp[2 * (s * P + q) + 1] = omega * llss[s * P + q] + zsf[2 * (s * P + q) + 1];
        }
    }
}

// The handle's resistance rows (CSR over loops, FP64) and the flag buffer; synchronises.
void upload_resistance(pq_loops *L, int n_loops, const int32_t *r_ptr, const int32_t *r_col, const double *r_val) {
    const int rnnz = r_ptr[n_loops];
    L->rptr.reserve((n_loops + 1) * 4); L->rcol.reserve(std::max(rnnz, 1) * 4); L->rval.reserve(std::max(rnnz, 1) * 8);
    PQ_HIP(hipMemcpyAsync(L->rptr.p, r_ptr, (n_loops + 1) * 4, hipMemcpyHostToDevice, L->stream));
    PQ_HIP(hipMemcpyAsync(L->rcol.p, r_col, (size_t)rnnz * 4, hipMemcpyHostToDevice, L->stream));
    PQ_HIP(hipMemcpyAsync(L->rval.p, r_val, (size_t)rnnz * 8, hipMemcpyHostToDevice, L->stream));
    PQ_HIP(hipStreamSynchronize(L->stream));
    L->singular.reserve(4);
}

// The factor of matrix b of the last batch (pq_loops_factor_batch) as a single factor.
LdlFactor batch_view(pq_loops *L, int b) {
    const int n = L->nf, nblk = (n + LDL_BLOCK - 1) / LDL_BLOCK;
    const long long ldf = L->ldf;
    const size_t sF = (size_t)2 * n * ldf, sX = (size_t)2 * nblk * LDL_BLOCK * LDL_BLOCK,
                 sW = (size_t)4 * n * 2 * LDL_BLOCK, wsz = (size_t)n * 2 * LDL_BLOCK;
    float *F = L->Fbat.as<float>() + sF * b, *X = L->Xbat.as<float>() + sX * b, *W = L->Wbat.as<float>() + sW * b;
    return LdlFactor{F, F + (size_t)n * ldf, ldf, n, X, X + (size_t)nblk * LDL_BLOCK * LDL_BLOCK,
                     {W, W + 2 * wsz}, {W + wsz, W + 3 * wsz}, L->aux, L->ev_ready, L->ev_done};
}

void check_resistance(int n_loops, const int32_t *r_ptr, const int32_t *r_col) {
    PQ_REQUIRE(r_ptr[0] == 0, "pq_loops_create: invalid resistance rows");
    for (int i = 0; i < n_loops; ++i)
        PQ_REQUIRE(r_ptr[i + 1] >= r_ptr[i], "pq_loops_create: resistance rows out of order");
    for (int32_t t = 0; t < r_ptr[n_loops]; ++t)
        PQ_REQUIRE(r_col[t] >= 0 && r_col[t] < n_loops, "pq_loops_create: resistance column out of range");
}

// ---- the direct solve's FP64 fallback ----------------------------------------------------
// Where the FP32 factor cannot bring a frequency's residual to the reference's level (a zero or
// non-finite FP32 pivot, or refinement that stops converging above 1e-10), that frequency is
// solved again with an FP64 LU with partial pivoting of Z_ff = Rl + j w Ll64, as MixedLU
// refactors in double (src/pcbqs/kernels.py:168-181). rocSOLVER's zgetrf/zgetrs do the LU; they
// are loaded on first use, so a process that never falls back never loads them.

// Z (n, n) complex FP64 = (0, w Ll64) over the free block, then Rl's free entries added (Z is
// symmetric, so its row-major layout is also the column-major one the solver reads).
__global__ void fallback_form_kernel(const double *Ll, long long ldl, int n, double omega, double2 *Z) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * n) return;
    const long long i = e / n, j = e % n;
    Z[e] = make_double2(0.0, omega * Ll[i * ldl + j]);
}

__global__ void fallback_resistance_kernel(const int *ptr, const int *col, const double *val, int n, double2 *Z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int t = ptr[i]; t < ptr[i + 1]; ++t)
        if (col[t] < n) Z[(size_t)i * n + col[t]].x += val[t];
}

// (n, 2P) split layout <-> (n, P) complex column-major
__global__ void fallback_pack_kernel(const double *B, int n, int P, double2 *Bc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * P) return;
    const int i = e / P, p = e % P;
    Bc[(size_t)p * n + i] = make_double2(B[(size_t)i * 2 * P + p], B[(size_t)i * 2 * P + P + p]);
}

__global__ void fallback_unpack_kernel(const double2 *Bc, int n, int P, double *X) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n * P) return;
    const int i = e / P, p = e % P;
    const double2 v = Bc[(size_t)p * n + i];
    X[(size_t)i * 2 * P + p] = v.x;
    X[(size_t)i * 2 * P + P + p] = v.y;
}

// rocBLAS / rocSOLVER entry points this file uses: the FP64 fallback LU, and the coarse space's
// device diagonalisation (dgemm, dsyevd).
struct RocLib {
    decltype(&rocblas_create_handle) create = nullptr;
    decltype(&rocblas_destroy_handle) destroy = nullptr;
    decltype(&rocblas_set_stream) set_stream = nullptr;
    decltype(&rocsolver_zgetrf) getrf = nullptr;
    decltype(&rocsolver_zgetrs) getrs = nullptr;
    decltype(&rocblas_dgemm) dgemm = nullptr;
    decltype(&rocsolver_dsyevd) syevd = nullptr;
    decltype(&rocsolver_dpotrf) potrf = nullptr;
    decltype(&rocblas_dtrsm) trsm = nullptr;
    decltype(&rocsolver_dsygvd_strided_batched) sygvd_batched = nullptr;
    decltype(&rocblas_cgemm_strided_batched) cgemm_batched = nullptr;
    std::string error;
};

const RocLib &roc_lib() {
    static RocLib lu;
    static std::once_flag once;
    std::call_once(once, [] {
        void *blas = dlopen("librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
        if (!blas) blas = dlopen("librocblas.so", RTLD_NOW | RTLD_LOCAL);
        void *solver = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
        if (!solver) solver = dlopen("librocsolver.so", RTLD_NOW | RTLD_LOCAL);
        if (!blas || !solver) {
            lu.error = std::string("cannot load rocBLAS/rocSOLVER: ") + dlerror();
            return;
        }
        lu.create = (decltype(lu.create))dlsym(blas, "rocblas_create_handle");
        lu.destroy = (decltype(lu.destroy))dlsym(blas, "rocblas_destroy_handle");
        lu.set_stream = (decltype(lu.set_stream))dlsym(blas, "rocblas_set_stream");
        lu.getrf = (decltype(lu.getrf))dlsym(solver, "rocsolver_zgetrf");
        lu.getrs = (decltype(lu.getrs))dlsym(solver, "rocsolver_zgetrs");
        lu.dgemm = (decltype(lu.dgemm))dlsym(blas, "rocblas_dgemm");
        lu.syevd = (decltype(lu.syevd))dlsym(solver, "rocsolver_dsyevd");
        lu.potrf = (decltype(lu.potrf))dlsym(solver, "rocsolver_dpotrf");
        lu.trsm = (decltype(lu.trsm))dlsym(blas, "rocblas_dtrsm");
        lu.sygvd_batched = (decltype(lu.sygvd_batched))dlsym(solver, "rocsolver_dsygvd_strided_batched");
        lu.cgemm_batched = (decltype(lu.cgemm_batched))dlsym(blas, "rocblas_cgemm_strided_batched");
        if (!lu.create || !lu.destroy || !lu.set_stream || !lu.getrf || !lu.getrs || !lu.dgemm || !lu.syevd ||
            !lu.potrf || !lu.trsm || !lu.sygvd_batched || !lu.cgemm_batched)
            lu.error = "rocBLAS/rocSOLVER entry points missing";
    });
    if (!lu.error.empty()) throw Error(PQ_EDEVICE, lu.error);
    return lu;
}

void blas_check(rocblas_status st, const char *what) {
    if (st != rocblas_status_success)
        throw Error(PQ_EDEVICE, std::string(what) + " failed (rocblas status " + std::to_string((int)st) + ")");
}

// The handle's rocBLAS handle on its stream (created on first use).
rocblas_handle blas_handle(pq_loops *L) {
    const RocLib &lu = roc_lib();
    if (!L->blas) {
        rocblas_handle h = nullptr;
        blas_check(lu.create(&h), "rocblas_create_handle");
        L->blas = h;
    }
    rocblas_handle h = (rocblas_handle)L->blas;
    blas_check(lu.set_stream(h, L->stream), "rocblas_set_stream");
    return h;
}

// Row-major C (m x n) = op(A) op(B) in FP64, K the inner dimension; lda, ldb, ldc the stored row
// lengths. (rocBLAS is column-major: a row-major matrix is its transpose there, so C^T = op(B)^T
// op(A)^T is computed.)
void gemm_rm(pq_loops *L, bool ta, bool tb, int m, int n, int K, const double *A, int lda, const double *B, int ldb,
             double *C, int ldc) {
    const double one = 1.0, zero = 0.0;
    blas_check(roc_lib().dgemm(blas_handle(L), tb ? rocblas_operation_transpose : rocblas_operation_none,
                               ta ? rocblas_operation_transpose : rocblas_operation_none, n, m, K, &one, B, ldb, A,
                               lda, &zero, C, ldc),
               "rocblas_dgemm");
}

// X (n, 2P) = Z_ff^-1 B at omega by the FP64 LU; the matrix and pivots live only for the call.
void fallback_solve(pq_loops *L, double omega, int P, const double *B, double *X) {
    const RocLib &lu = roc_lib();
    const int n = L->nf;
    rocblas_handle h = blas_handle(L);
    DevBuf Z, piv, info, Bc;
    Z.reserve((size_t)n * n * 16);
    piv.reserve((size_t)n * 4);
    info.reserve(4);
    Bc.reserve((size_t)n * P * 16);
    const long long cnt = (long long)n * n;
    hipLaunchKernelGGL(fallback_form_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, L->stream,
                       L->Ll64.as<double>(), L->ldl, n, omega, Z.as<double2>());
    hipLaunchKernelGGL(fallback_resistance_kernel, dim3((n + 255) / 256), dim3(256), 0, L->stream, L->rptr.as<int>(),
                       L->rcol.as<int>(), L->rval.as<double>(), n, Z.as<double2>());
    hipLaunchKernelGGL(fallback_pack_kernel, dim3((n * P + 255) / 256), dim3(256), 0, L->stream, B, n, P,
                       Bc.as<double2>());
    PQ_HIP(hipGetLastError());
    auto *A = reinterpret_cast<rocblas_double_complex *>(Z.p);
    blas_check(lu.getrf(h, n, n, A, n, piv.as<rocblas_int>(), info.as<rocblas_int>()), "rocsolver_zgetrf");
    rocblas_int singular = 0;
    PQ_HIP(hipMemcpyAsync(&singular, info.p, 4, hipMemcpyDeviceToHost, L->stream));
    PQ_HIP(hipStreamSynchronize(L->stream));
    if (singular) throw Error(PQ_ECONVERGE, "direct solve: Z_ff is singular in FP64 (zero pivot " +
                                                std::to_string(singular) + ")");
    blas_check(lu.getrs(h, rocblas_operation_none, n, P, A, n, piv.as<rocblas_int>(),
                        reinterpret_cast<rocblas_double_complex *>(Bc.p), n),
               "rocsolver_zgetrs");
    hipLaunchKernelGGL(fallback_unpack_kernel, dim3((n * P + 255) / 256), dim3(256), 0, L->stream, Bc.as<double2>(),
                       n, P, X);
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipStreamSynchronize(L->stream));       // Z, the pivots and Bc are freed on return
}

// The largest column's relative residual |B - Z X| / |B| (FP64 system) into *worst; W is work.
double relative_residual64(pq_loops *L, const double *B, const double *X, double *W, int P,
                           const std::vector<double> &bnorm) {
    const size_t count = (size_t)L->nf * 2 * P;
    apply_operator64(L, X, W, P);
    ++L->matvecs;
    hipLaunchKernelGGL(vec_sub_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream, B, count, W);
    PQ_HIP(hipGetLastError());
    std::vector<double> rn;
    norms(L, W, P, rn);
    double worst = 0;
    for (int p = 0; p < P; ++p) {
        const double r = rn[p] / bnorm[p];
        worst = std::isnan(r) || std::isnan(worst) ? std::numeric_limits<double>::quiet_NaN() : std::max(worst, r);
    }
    return worst;
}

// FP16 copies of the bases for the solve's restriction and prolongation, each scaled by its largest
// magnitude (the Galerkin matrices used the FP32 bases; PCBQS_AMD_COARSE_FP16=0 keeps the solve on
// the FP32 bases). The coarse correction is a preconditioner: its rounding moves the iterates, not
// the solution, which is accepted on the true FP64 residual.
void make_half_bases(pq_loops *L) {
    const int n = L->nf, k = L->kc, kp = L->kp;
    const char *hv = std::getenv("PCBQS_AMD_COARSE_FP16");
    L->coarse_half = !(hv && hv[0] == '0');
    if (!L->coarse_half) return;
    L->Phih.reserve((size_t)n * kp * 2); L->Phith.reserve((size_t)L->ldt * k * 2);
    L->RPhih.reserve((size_t)L->ldt * k * 2); L->LPhih.reserve((size_t)L->ldt * k * 2);
    L->hmax.reserve(4 * 4);
    PQ_HIP(hipMemsetAsync(L->hmax.p, 0, 16, L->stream));
    const float *src[4] = {L->Phi.as<float>(), L->Phit.as<float>(), L->RPhi.as<float>(), L->LPhi.as<float>()};
    _Float16 *dst[4] = {L->Phih.as<_Float16>(), L->Phith.as<_Float16>(), L->RPhih.as<_Float16>(),
                        L->LPhih.as<_Float16>()};
    for (int q = 0; q < 4; ++q) {
        const size_t cnt = q == 0 ? (size_t)n * kp : (size_t)L->ldt * k;
        hipLaunchKernelGGL(absmax_kernel, dim3(1024), dim3(256), 0, L->stream, src[q], cnt, L->hmax.as<unsigned>() + q);
        hipLaunchKernelGGL(to_half_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, L->stream, src[q], cnt,
                           L->hmax.as<unsigned>() + q, dst[q]);
    }
    PQ_HIP(hipGetLastError());
    unsigned mx[4];
    PQ_HIP(hipMemcpyAsync(mx, L->hmax.p, 16, hipMemcpyDeviceToHost, L->stream));
    PQ_HIP(hipStreamSynchronize(L->stream));
    for (int q = 0; q < 4; ++q) {
        float m;
        std::memcpy(&m, &mx[q], 4);
        L->hscale[q] = m > 0 ? m : 1.0;
    }
}

// Rl X and Ll X (FP64 (n, k) row-major each) of the k columns of X (FP64 (n, k)): Rl X by CSR rows,
// Ll X as one FP64 GEMM on the FP64 loop matrix (symmetric, so its rows serve as its columns).
void basis_products(pq_loops *L, const double *X64, int k, double *RX, double *LX) {
    const int n = L->nf;
    hipLaunchKernelGGL(csr_rows_kernel, dim3(n), dim3(256), 0, L->stream, L->rptr.as<int>(), L->rcol.as<int>(),
                       L->rval.as<double>(), n, X64, k, RX);
    PQ_HIP(hipGetLastError());
    gemm_rm(L, false, false, n, k, n, L->Ll64.as<double>(), (int)L->ldl, X64, k, LX, k);
}

// A scalar from a device sum (strided_sum_kernel), synchronising the handle's stream.
double device_sum(pq_loops *L, const double *A, size_t count, size_t stride, bool square) {
    constexpr int NB = 512;                                // partials: a fixed count, a fixed order
    DevBuf out, part;
    out.reserve(8);
    part.reserve(NB * 8);
    hipLaunchKernelGGL(strided_partial_kernel, dim3(NB), dim3(256), 0, L->stream, A, count, stride, square ? 1 : 0,
                       part.as<double>());
    hipLaunchKernelGGL(strided_sum_kernel, dim3(1), dim3(1024), 0, L->stream, part.as<double>(), (size_t)NB, (size_t)1,
                       0, out.as<double>());
    PQ_HIP(hipGetLastError());
    double v = 0;
    PQ_HIP(hipMemcpyAsync(&v, out.p, 8, hipMemcpyDeviceToHost, L->stream));
    PQ_HIP(hipStreamSynchronize(L->stream));
    return v;
}

// Eigen-decomposition of a symmetric k x k FP64 matrix A in place on the device (rocSOLVER dsyevd):
// ascending eigenvalues into s (host), eigenvectors as A's columns (column-major).
void device_eigh(pq_loops *L, double *A, int k, std::vector<double> &s) {
    DevBuf D, E, info;
    D.reserve((size_t)k * 8); E.reserve((size_t)k * 8); info.reserve(4);
    blas_check(roc_lib().syevd(blas_handle(L), rocblas_evect_original, rocblas_fill_upper, k, A, k, D.as<double>(),
                               E.as<double>(), info.as<rocblas_int>()),
               "rocsolver_dsyevd");
    rocblas_int bad = 0;
    s.resize(k);
    PQ_HIP(hipMemcpyAsync(&bad, info.p, 4, hipMemcpyDeviceToHost, L->stream));
    PQ_HIP(hipMemcpyAsync(s.data(), D.p, (size_t)k * 8, hipMemcpyDeviceToHost, L->stream));
    PQ_HIP(hipStreamSynchronize(L->stream));
    PQ_REQUIRE(bad == 0, "coarse space: the eigensolver did not converge");
}


// The coarse space of the span of P64 (n_free, k) FP64 row-major on the device, re-expressed in the
// generalised eigenbasis of its Galerkin pair (Rc, Lc) = (P^T Rl P, P^T Ll64 P), so that every
// frequency's coarse solve is a diagonal scaling. Two routes to the pair's eigenbasis:
//  1 (Cholesky): Rc = U^T U (rocSOLVER dpotrf), U^-1 by a triangular solve of the identity,
//    M = U^-T Lc U^-1 = Q Lam Q^T (one dsyevd), T = U^-1 Q; taken when dpotrf succeeds and
//    1 / ||U^-1||_F^2 > drop * trace(Rc): since trace(Rc^-1) = ||U^-1||_F^2 <= k / lambda_min and
//    trace(Rc) >= lambda_max, that guarantees lambda_min > drop * lambda_max, i.e. the eigenvector
//    route below would drop nothing;
//  2 (eigenvectors): Rc = V S V^T (dsyevd), directions with S below drop * max S removed,
//    W = V S^-1/2, W^T Lc W = Q Lam Q^T (dsyevd), T = W Q.
// Either way Psi = P T has Psi^T Rl Psi = I and Psi^T Ll Psi = Lam. P64 is rounded to FP32 first
// (the basis every product uses). *path_out: the route taken.
void coarse_diagonalise(pq_loops *L, DevBuf &P64, int k, double drop, int *k_out, int *path_out) {
    const int n = L->nf;
    L->kc = 0;
    L->omega_c = -1;
    L->crows.clear(); L->psplit.clear(); L->pdepth.clear();
    DevBuf RX, LX, Rc, Lc, W, M, T, Y;
    {
        const size_t cnt = (size_t)n * k;
        hipLaunchKernelGGL(round_inplace_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, L->stream,
                           P64.as<double>(), cnt);
        PQ_HIP(hipGetLastError());
    }
    RX.reserve((size_t)n * k * 8); LX.reserve((size_t)n * k * 8);
    basis_products(L, P64.as<double>(), k, RX.as<double>(), LX.as<double>());
    Rc.reserve((size_t)k * k * 8); Lc.reserve((size_t)k * k * 8);
    gemm_rm(L, true, false, k, k, n, P64.as<double>(), k, RX.as<double>(), k, Rc.as<double>(), k);
    gemm_rm(L, true, false, k, k, n, P64.as<double>(), k, LX.as<double>(), k, Lc.as<double>(), k);
    const unsigned kb = (unsigned)(((size_t)k * k + 255) / 256);
    hipLaunchKernelGGL(symmetrize64_kernel, dim3(kb), dim3(256), 0, L->stream, Rc.as<double>(), k);
    hipLaunchKernelGGL(symmetrize64_kernel, dim3(kb), dim3(256), 0, L->stream, Lc.as<double>(), k);
    PQ_HIP(hipGetLastError());
    std::vector<double> lam;
    int kk = 0;
    const RocLib &lib = roc_lib();
    rocblas_handle h = blas_handle(L);
    {   // route 1: Cholesky of Rc
        const double trace = device_sum(L, Rc.as<double>(), (size_t)k, (size_t)k + 1, false);
        DevBuf U, Ui, info;
        U.reserve((size_t)k * k * 8); Ui.reserve((size_t)k * k * 8); info.reserve(4);
        PQ_HIP(hipMemcpyAsync(U.p, Rc.p, (size_t)k * k * 8, hipMemcpyDeviceToDevice, L->stream));
        blas_check(lib.potrf(h, rocblas_fill_upper, k, U.as<double>(), k, info.as<rocblas_int>()), "rocsolver_dpotrf");
        rocblas_int bad = 0;
        PQ_HIP(hipMemcpyAsync(&bad, info.p, 4, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
        bool ok = bad == 0 && trace > 0;
        if (ok) {
            hipLaunchKernelGGL(identity_kernel, dim3(kb), dim3(256), 0, L->stream, Ui.as<double>(), k);
            PQ_HIP(hipGetLastError());
            const double one = 1.0;
            blas_check(lib.trsm(h, rocblas_side_left, rocblas_fill_upper, rocblas_operation_none,
                                rocblas_diagonal_non_unit, k, k, &one, U.as<double>(), k, Ui.as<double>(), k),
                       "rocblas_dtrsm");                                   // Ui = U^-1 (column-major)
            const double inv_trace = device_sum(L, Ui.as<double>(), (size_t)k * k, 1, true);
            ok = std::isfinite(inv_trace) && inv_trace > 0 && 1.0 / inv_trace > drop * trace;
        }
        if (ok) {
            // Ui column-major read row-major is Ui^T: M = Ui^T Lc Ui, T = Ui Q (row-major k x k)
            T.reserve((size_t)k * k * 8); M.reserve((size_t)k * k * 8);
            gemm_rm(L, false, true, k, k, k, Lc.as<double>(), k, Ui.as<double>(), k, T.as<double>(), k);  // Lc Ui
            gemm_rm(L, false, false, k, k, k, Ui.as<double>(), k, T.as<double>(), k, M.as<double>(), k);  // Ui^T (.)
            hipLaunchKernelGGL(symmetrize64_kernel, dim3(kb), dim3(256), 0, L->stream, M.as<double>(), k);
            PQ_HIP(hipGetLastError());
            device_eigh(L, M.as<double>(), k, lam);                     // M's columns: Q
            gemm_rm(L, true, true, k, k, k, Ui.as<double>(), k, M.as<double>(), k, T.as<double>(), k);
            kk = k;
            *path_out = 1;
        }
    }
    if (!kk) {   // route 2: eigenvectors of Rc
        std::vector<double> sv;
        device_eigh(L, Rc.as<double>(), k, sv);                        // Rc's columns: V
        const double smax = sv.empty() ? 0.0 : sv.back();
        PQ_REQUIRE(smax > 0, "coarse space: the Galerkin resistance is not positive");
        int first = 0;
        while (first < k && !(sv[first] > drop * smax)) ++first;
        kk = k - first;
        W.reserve((size_t)k * kk * 8);
        {
            DevBuf sd;
            sd.reserve((size_t)k * 8);
            PQ_HIP(hipMemcpyAsync(sd.p, sv.data(), (size_t)k * 8, hipMemcpyHostToDevice, L->stream));
            hipLaunchKernelGGL(whiten_kernel, dim3((unsigned)(((size_t)k * kk + 255) / 256)), dim3(256), 0, L->stream,
                               Rc.as<double>(), sd.as<double>(), k, first, kk, W.as<double>());
            PQ_HIP(hipGetLastError());
            PQ_HIP(hipStreamSynchronize(L->stream));
        }
        T.reserve((size_t)k * kk * 8);
        M.reserve((size_t)kk * kk * 8);
        gemm_rm(L, false, false, k, kk, k, Lc.as<double>(), k, W.as<double>(), kk, T.as<double>(), kk);   // Lc W
        gemm_rm(L, true, false, kk, kk, k, W.as<double>(), kk, T.as<double>(), kk, M.as<double>(), kk);    // W^T Lc W
        hipLaunchKernelGGL(symmetrize64_kernel, dim3((unsigned)(((size_t)kk * kk + 255) / 256)), dim3(256), 0,
                           L->stream, M.as<double>(), kk);
        PQ_HIP(hipGetLastError());
        device_eigh(L, M.as<double>(), kk, lam);                       // M's columns: Q
        // T = W Q (M holds Q column-major: as a row-major array it is Q^T)
        gemm_rm(L, false, true, k, kk, kk, W.as<double>(), kk, M.as<double>(), kk, T.as<double>(), kk);
        *path_out = 2;
    }
    // Psi = Phi T, Rl Psi = (Rl Phi) T, Ll Psi = (Ll Phi) T
    Y.reserve((size_t)n * kk * 8);
    const int kpp = (kk + 63) / 64 * 64;
    L->kc = kk;
    L->kp = kpp;
    L->ldt = (n + 63) / 64 * 64;
    L->Phi.reserve((size_t)n * kpp * 4); L->Phit.reserve((size_t)L->ldt * kk * 4);
    L->RPhi.reserve((size_t)L->ldt * kk * 4); L->LPhi.reserve((size_t)L->ldt * kk * 4);
    const unsigned nb = (unsigned)(((size_t)L->ldt * kk + 255) / 256);
    gemm_rm(L, false, false, n, kk, k, P64.as<double>(), k, T.as<double>(), kk, Y.as<double>(), kk);
    hipLaunchKernelGGL(round_basis_kernel, dim3((unsigned)(((size_t)n * kpp + 255) / 256)), dim3(256), 0, L->stream,
                       Y.as<double>(), n, kk, kpp, L->Phi.as<float>());
    hipLaunchKernelGGL(transpose_round_kernel, dim3(nb), dim3(256), 0, L->stream, Y.as<double>(), n, kk, L->ldt,
                       L->Phit.as<float>());
    gemm_rm(L, false, false, n, kk, k, RX.as<double>(), k, T.as<double>(), kk, Y.as<double>(), kk);
    hipLaunchKernelGGL(transpose_round_kernel, dim3(nb), dim3(256), 0, L->stream, Y.as<double>(), n, kk, L->ldt,
                       L->RPhi.as<float>());
    gemm_rm(L, false, false, n, kk, k, LX.as<double>(), k, T.as<double>(), kk, Y.as<double>(), kk);
    hipLaunchKernelGGL(transpose_round_kernel, dim3(nb), dim3(256), 0, L->stream, Y.as<double>(), n, kk, L->ldt,
                       L->LPhi.as<float>());
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipStreamSynchronize(L->stream));
    // the Galerkin pair in the new basis: I and Lam (their off-diagonal parts are rounding)
    L->Rc.assign((size_t)kk * kk, 0.0);
    L->Lc.assign((size_t)kk * kk, 0.0);
    for (int a = 0; a < kk; ++a) {
        L->Rc[(size_t)a * kk + a] = 1.0;
        L->Lc[(size_t)a * kk + a] = lam[a];
    }
    L->cdiag = true;
    make_half_bases(L);
    *k_out = kk;
}


// The block-Jacobi inverses at omega into L->inv: from the blocks' modes (computed once per handle:
// the pencils R_b, L_b in FP64, rocSOLVER's batched dsygvd) by one batched cgemm, or, where a block's
// R_b is not positive definite, by the per-frequency Gauss-Jordan inversion with partial pivoting.
void block_jacobi_modes(pq_loops *L) {
    const int n = L->nf;
    L->nblk = (n + BS - 1) / BS;
    const int nb = L->nblk;
    const size_t per = (size_t)BS * BS, count = per * nb;
    const char *mv = std::getenv("PCBQS_AMD_BJ_MODAL");
    if (n > 0 && L->bj_modal == 0 && !(mv && mv[0] == '0')) {
        const RocLib &lib = roc_lib();
        DevBuf Lb, Rb, E, info;
        Lb.reserve(count * 8); Rb.reserve(count * 8); E.reserve((size_t)nb * BS * 8); info.reserve((size_t)nb * 4);
        L->bjlam.reserve((size_t)nb * BS * 8);
        hipLaunchKernelGGL(bj_pencil_kernel, dim3(BS, nb), dim3(256), 0, L->stream, L->Ll64.as<double>(), L->ldl,
                           L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, Lb.as<double>(),
                           Rb.as<double>());
        PQ_HIP(hipGetLastError());
        blas_check(lib.sygvd_batched(blas_handle(L), rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, BS,
                                     Lb.as<double>(), BS, (rocblas_stride)per, Rb.as<double>(), BS,
                                     (rocblas_stride)per, L->bjlam.as<double>(), BS, E.as<double>(), BS,
                                     info.as<rocblas_int>(), nb),
                   "rocsolver_dsygvd_strided_batched");
        std::vector<rocblas_int> bad(nb);
        PQ_HIP(hipMemcpyAsync(bad.data(), info.p, (size_t)nb * 4, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
        L->bj_modal = std::all_of(bad.begin(), bad.end(), [](rocblas_int v) { return v == 0; }) ? 1 : -1;
        if (L->bj_modal == 1) {
            L->bjV.reserve(count * 8);
            L->bjA.reserve(count * 8);
            hipLaunchKernelGGL(bj_modes_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream,
                               Lb.as<double>(), count, L->bjV.as<float2>());
            PQ_HIP(hipGetLastError());
            PQ_HIP(hipStreamSynchronize(L->stream));      // the pencils are freed on return
        }
    }
}

void block_jacobi_setup(pq_loops *L, double omega) {
    const int n = L->nf;
    L->nblk = (n + BS - 1) / BS;
    const int nb = L->nblk;
    const size_t per = (size_t)BS * BS, count = per * nb;
    L->inv.reserve(count * 2 * 4);
    block_jacobi_modes(L);
    if (L->bj_modal == 1) {
        hipLaunchKernelGGL(bj_scale_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream,
                           L->bjV.as<float2>(), L->bjlam.as<double>(), count, omega, L->bjA.as<float2>());
        PQ_HIP(hipGetLastError());
        // inv_b = A_b V_b^T (symmetric, so its column-major and row-major layouts agree)
        const rocblas_float_complex one(1.f, 0.f), zero(0.f, 0.f);
        blas_check(roc_lib().cgemm_batched(blas_handle(L), rocblas_operation_none, rocblas_operation_transpose, BS, BS,
                                           BS, &one, L->bjA.as<rocblas_float_complex>(), BS, (rocblas_stride)per,
                                           L->bjV.as<rocblas_float_complex>(), BS, (rocblas_stride)per, &zero,
                                           L->inv.as<rocblas_float_complex>(), BS, (rocblas_stride)per, nb),
                   "rocblas_cgemm_strided_batched");
        return;
    }
    PQ_HIP(hipMemsetAsync(L->singular.p, 0, 4, L->stream));
    L->bjperm.reserve((size_t)nb * BS * 4);
    hipLaunchKernelGGL(bj_setup_kernel, dim3(nb), dim3(1024), 0, L->stream, L->Ll.as<float>(), L->ldl,
                       L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, omega, L->inv.as<float>(),
                       L->bjperm.as<int>(), L->singular.as<int>());
    PQ_HIP(hipGetLastError());
    int sing = 0;
    PQ_HIP(hipMemcpyAsync(&sing, L->singular.p, 4, hipMemcpyDeviceToHost, L->stream));
    PQ_HIP(hipStreamSynchronize(L->stream));
    PQ_REQUIRE(!sing, "block-Jacobi: a singular diagonal block");
}

}  // namespace

extern "C" {

int pq_loops_create(int device, int n_loops, int n_free, const int32_t *c_colptr, const int32_t *c_rows,
                    const float *c_sign, int64_t c_nnz, const void *Lb_dev, int lb_precision, int64_t ldb,
                    int n_branch_rows, const int32_t *r_ptr, const int32_t *r_col, const double *r_val,
                    pq_loops **out) {
    return guard([&] {
        PQ_REQUIRE(out && c_colptr && c_rows && c_sign && Lb_dev && r_ptr && r_col && r_val,
                   "pq_loops_create: null argument");
        PQ_REQUIRE(lb_precision == 32 || lb_precision == 64, "pq_loops_create: Lb precision must be 32 or 64");
        PQ_REQUIRE(n_loops > 0 && n_free >= 0 && n_free <= n_loops && n_branch_rows > 0 && ldb >= n_branch_rows,
                   "pq_loops_create: invalid sizes");
        PQ_REQUIRE(c_colptr[0] == 0 && c_colptr[n_loops] == c_nnz, "pq_loops_create: inconsistent loop basis");
        for (int64_t t = 0; t < c_nnz; ++t) PQ_REQUIRE(c_rows[t] >= 0, "pq_loops_create: negative branch index");
        check_resistance(n_loops, r_ptr, r_col);
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
            U.reserve((size_t)n_branch_rows * ldu * 8);
            L->Ll.reserve((size_t)n_loops * L->ldl * 4);
            L->Ll64.reserve((size_t)n_loops * L->ldl * 8);
            const dim3 lgrid((n_branch_rows + PROJ_ROWS - 1) / PROJ_ROWS);
            if (lb_precision == 64)
                hipLaunchKernelGGL(proj_left_kernel<double>, lgrid, dim3(256), 0, L->stream, (const double *)Lb_dev,
                                   (long long)ldb, n_branch_rows, cp.as<int>(), cr.as<int>(), cs.as<float>(), n_loops,
                                   U.as<double>(), ldu);
            else
                hipLaunchKernelGGL(proj_left_kernel<float>, lgrid, dim3(256), 0, L->stream, (const float *)Lb_dev,
                                   (long long)ldb, n_branch_rows, cp.as<int>(), cr.as<int>(), cs.as<float>(), n_loops,
                                   U.as<double>(), ldu);
            hipLaunchKernelGGL(proj_right_kernel, dim3(n_loops), dim3(256), 0, L->stream, U.as<double>(), ldu,
                               n_branch_rows, cp.as<int>(), cr.as<int>(), cs.as<float>(), n_loops, L->Ll64.as<double>(),
                               L->ldl);
            {
                const long long cnt = (long long)n_loops * n_loops;
                hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, L->stream,
                                   L->Ll64.as<double>(), L->Ll.as<float>(), L->ldl, n_loops);
            }
            PQ_HIP(hipGetLastError());
            upload_resistance(L, n_loops, r_ptr, r_col, r_val);
        } catch (...) {
            if (L->stream) (void)hipStreamDestroy(L->stream);
            delete L;
            throw;
        }
        *out = L;
    });
}

int pq_loops_create_tree(int device, int n_loops, int n_free, const pq_loop_tree *tr, const void *Lb_dev,
                         int lb_precision, int64_t ldb, int n_branch_rows, const int32_t *r_ptr, const int32_t *r_col,
                         const double *r_val, pq_loops **out) {
    return guard([&] {
        PQ_REQUIRE(out && tr && Lb_dev && r_ptr && r_col && r_val && tr->parent && tr->up_branch && tr->up_sign &&
                       tr->level_ptr && tr->level_nodes && tr->closing && tr->end_a && tr->end_b,
                   "pq_loops_create_tree: null argument");
        PQ_REQUIRE(lb_precision == 32 || lb_precision == 64, "pq_loops_create_tree: Lb precision must be 32 or 64");
        const int nn = tr->n_nodes, nlev = tr->n_levels;
        PQ_REQUIRE(n_loops > 0 && n_free >= 0 && n_free <= n_loops && n_branch_rows > 0 && ldb >= n_branch_rows &&
                       nn > 0 && nlev >= 0,
                   "pq_loops_create_tree: invalid sizes");
        // every listed node hangs under a root or a node of an earlier level, listed once
        std::vector<char> done(nn, 0);
        for (int x = 0; x < nn; ++x) {
            PQ_REQUIRE(tr->parent[x] >= -1 && tr->parent[x] < nn, "pq_loops_create_tree: parent out of range");
            if (tr->parent[x] < 0) done[x] = 1;
        }
        PQ_REQUIRE(tr->level_ptr[0] == 0, "pq_loops_create_tree: invalid levels");
        for (int d = 0; d < nlev; ++d) {
            PQ_REQUIRE(tr->level_ptr[d + 1] >= tr->level_ptr[d], "pq_loops_create_tree: levels out of order");
            for (int e = tr->level_ptr[d]; e < tr->level_ptr[d + 1]; ++e) {
                const int x = tr->level_nodes[e];
                PQ_REQUIRE(x >= 0 && x < nn && tr->parent[x] >= 0 && done[tr->parent[x]] == 1 && !done[x],
                           "pq_loops_create_tree: a level lists a node before its parent, twice, or a root");
            }
            for (int e = tr->level_ptr[d]; e < tr->level_ptr[d + 1]; ++e) done[tr->level_nodes[e]] = 1;
        }
        for (int j = 0; j < n_loops; ++j)
            PQ_REQUIRE(tr->end_a[j] >= 0 && tr->end_a[j] < nn && tr->end_b[j] >= 0 && tr->end_b[j] < nn,
                       "pq_loops_create_tree: loop end out of range");
        check_resistance(n_loops, r_ptr, r_col);
        DeviceScope scope(device);
        auto *L = new pq_loops();
        try {
            L->device = device; L->nl = n_loops; L->nf = n_free;
            L->ldl = (n_loops + 63) / 64 * 64;
            PQ_HIP(hipStreamCreateWithFlags(&L->stream, hipStreamNonBlocking));
            const int nbr = n_branch_rows, nlv = tr->level_ptr[nlev];
            DevBuf par, upb, ups, lev, cot, ea, eb, S, U;
            par.reserve((size_t)nn * 4); upb.reserve((size_t)nn * 4); ups.reserve((size_t)nn * 4);
            lev.reserve((size_t)std::max(nlv, 1) * 4);
            cot.reserve((size_t)n_loops * 4); ea.reserve((size_t)n_loops * 4); eb.reserve((size_t)n_loops * 4);
            PQ_HIP(hipMemcpy(par.p, tr->parent, (size_t)nn * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(upb.p, tr->up_branch, (size_t)nn * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(ups.p, tr->up_sign, (size_t)nn * 4, hipMemcpyHostToDevice));
            if (nlv) PQ_HIP(hipMemcpy(lev.p, tr->level_nodes, (size_t)nlv * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(cot.p, tr->closing, (size_t)n_loops * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(ea.p, tr->end_a, (size_t)n_loops * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(eb.p, tr->end_b, (size_t)n_loops * 4, hipMemcpyHostToDevice));
            auto levels = [&](auto *Row, long long ldr, int width, double *Sd, long long lds) {
                using TR = std::remove_const_t<std::remove_pointer_t<decltype(Row)>>;
                for (int d = 0; d < nlev; ++d) {
                    const int c = tr->level_ptr[d + 1] - tr->level_ptr[d];
                    if (!c) continue;
                    hipLaunchKernelGGL(tree_level_kernel<TR>, dim3((width + 255) / 256, std::min(c, 65535)), dim3(256),
                                       0, L->stream, Row, ldr, nbr, width, lev.as<int>() + tr->level_ptr[d], c,
                                       par.as<int>(), upb.as<int>(), ups.as<float>(), Sd, lds);
                }
            };
            // U = Lb C (nbr x n_loops, FP64, row stride ldu)
            const long long ldu = L->ldl, lds1 = (nbr + 63) / 64 * 64;
            S.reserve((size_t)nn * lds1 * 8);
            PQ_HIP(hipMemsetAsync(S.p, 0, (size_t)nn * lds1 * 8, L->stream));
            U.reserve((size_t)nbr * ldu * 8);
            const dim3 ugrid((nbr + 63) / 64, (n_loops + 63) / 64);
            if (lb_precision == 64) {
                const double *Lb = (const double *)Lb_dev;
                levels(Lb, ldb, nbr, S.as<double>(), lds1);
                hipLaunchKernelGGL(tree_u_kernel<double>, ugrid, dim3(256), 0, L->stream, Lb, (long long)ldb, nbr,
                                   S.as<double>(), lds1, cot.as<int>(), ea.as<int>(), eb.as<int>(), n_loops,
                                   U.as<double>(), ldu);
            } else {
                const float *Lb = (const float *)Lb_dev;
                levels(Lb, ldb, nbr, S.as<double>(), lds1);
                hipLaunchKernelGGL(tree_u_kernel<float>, ugrid, dim3(256), 0, L->stream, Lb, (long long)ldb, nbr,
                                   S.as<double>(), lds1, cot.as<int>(), ea.as<int>(), eb.as<int>(), n_loops,
                                   U.as<double>(), ldu);
            }
            PQ_HIP(hipStreamSynchronize(L->stream));
            S.release();
            // Ll = C^T U, the same prefix sums down the tree on the rows of U
            const long long lds2 = L->ldl;
            S.reserve((size_t)nn * lds2 * 8);
            PQ_HIP(hipMemsetAsync(S.p, 0, (size_t)nn * lds2 * 8, L->stream));
            levels((const double *)U.as<double>(), ldu, n_loops, S.as<double>(), lds2);
            L->Ll.reserve((size_t)n_loops * L->ldl * 4);
            L->Ll64.reserve((size_t)n_loops * L->ldl * 8);
            hipLaunchKernelGGL(tree_ll_kernel, dim3((n_loops + 255) / 256, std::min(n_loops, 65535)), dim3(256), 0,
                               L->stream, U.as<double>(), ldu, nbr, S.as<double>(), lds2, cot.as<int>(), ea.as<int>(),
                               eb.as<int>(), n_loops, L->Ll64.as<double>(), L->ldl);
            {
                const long long cnt = (long long)n_loops * n_loops;
                hipLaunchKernelGGL(symmetrize_kernel, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, L->stream,
                                   L->Ll64.as<double>(), L->Ll.as<float>(), L->ldl, n_loops);
            }
            PQ_HIP(hipGetLastError());
            upload_resistance(L, n_loops, r_ptr, r_col, r_val);
        } catch (...) {
            if (L->stream) {
                (void)hipStreamSynchronize(L->stream);
                (void)hipStreamDestroy(L->stream);
            }
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
        for (DevBuf *b : {&L->Ll, &L->Ll64, &L->gpart64, &L->rptr, &L->rcol, &L->rval, &L->sc, &L->inv, &L->bjperm, &L->singular, &L->bjV, &L->bjlam, &L->bjA, &L->V, &L->W, &L->Z,
                          &L->X, &L->B, &L->R, &L->partial, &L->h, &L->scal, &L->Phi, &L->RPhi, &L->LPhi, &L->Ainv,
                          &L->cvec, &L->cpart, &L->ctot, &L->T1, &L->T2, &L->gpart, &L->Phit, &L->ppart})
            b->release();
        drop_graphs(L);
        L->Hd.release();
        if (L->Hh) (void)hipHostFree(L->Hh);
        if (L->aux) (void)hipStreamSynchronize(L->aux);
        if (L->ev_ready) (void)hipEventDestroy(L->ev_ready);
        if (L->ev_done) (void)hipEventDestroy(L->ev_done);
        if (L->aux) (void)hipStreamDestroy(L->aux);
        if (L->blas) (void)roc_lib().destroy((rocblas_handle)L->blas);
        if (L->stream) (void)hipStreamDestroy(L->stream);
        delete L;
    });
}

int pq_loops_matrix(pq_loops *L, int precision, void **Ll_dev, int64_t *ldl) {
    return guard([&] {
        PQ_REQUIRE(L, "pq_loops_matrix: null handle");
        PQ_REQUIRE(precision == 32 || precision == 64, "pq_loops_matrix: precision must be 32 or 64");
        if (Ll_dev) *Ll_dev = precision == 64 ? L->Ll64.p : L->Ll.p;
        if (ldl) *ldl = L->ldl;
    });
}

// Port impedances at one angular frequency: X = -Z_ff^-1 Z_fs by GMRES, Zp = Z_ss + Z_sf X.
int pq_loops_solve_ports(pq_loops *L, double omega, int P, const int32_t *port_cols, int precond, double tol,
                         int restart, int maxiter, double *Zp, double *X_host, const double *X0_host,
                         int64_t *iterations, double *residual) {
    return guard([&] {
        PQ_REQUIRE(L && port_cols && Zp, "pq_loops_solve_ports: null argument");
        PQ_REQUIRE(P >= 1 && tol > 0 && tol < 1 && restart >= 1 && maxiter >= 1, "pq_loops_solve_ports: invalid options");
        PQ_REQUIRE(precond >= 0 && precond <= 2, "pq_loops_solve_ports: preconditioner 0 (none), 1 (block Jacobi) or 2 (two-level)");
        PQ_REQUIRE(precond != 2 || L->kc > 0, "pq_loops_solve_ports: two-level preconditioner without a coarse space");
        for (int p = 0; p < P; ++p)
            PQ_REQUIRE(port_cols[p] >= L->nf && port_cols[p] < L->nl, "pq_loops_solve_ports: port loop out of range");
        DeviceScope scope(L->device);
        tune_gemv(L, 2 * P);
        set_omega(L, omega);
        {   // the operator product's partial sums, sized before any capture
            const int nf = std::max(L->nf, 1);
            L->gpart.reserve((size_t)operator_splits(nf, gemv_chunk(L, 2 * P)) * nf * 2 * P * 8);
        }
        const auto started = std::chrono::steady_clock::now();
        const int n = L->nf, NR = 2 * P, m = std::min(restart, std::max(n, 1));
        const size_t vsz = (size_t)n * NR;
        L->matvecs = 0;
        L->sc.reserve(P * 4);
        PQ_HIP(hipMemcpyAsync(L->sc.p, port_cols, P * 4, hipMemcpyHostToDevice, L->stream));
        L->V.reserve((m + 1) * vsz * 8);
        L->W.reserve(vsz * 8); L->Z.reserve(vsz * 8); L->X.reserve(vsz * 8); L->B.reserve(vsz * 8); L->R.reserve(vsz * 8);
        L->partial.reserve((size_t)RED_BLOCKS * (m + 2) * P * 2 * 8);
        L->Hd.reserve(8);
        L->h.reserve((size_t)(m + 2) * P * 2 * 8);
        L->scal.reserve(P * 8);
        double *V = L->V.as<double>(), *W = L->W.as<double>(), *Z = L->Z.as<double>(), *X = L->X.as<double>();
        double *B = L->B.as<double>(), *R = L->R.as<double>();
        if (n > 0) {
            hipLaunchKernelGGL(rhs_kernel, dim3((n * P + 255) / 256), dim3(256), 0, L->stream, L->Ll64.as<double>(), L->ldl,
                               L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, L->sc.as<int>(), P, omega, B);
            PQ_HIP(hipGetLastError());
            if (precond == 2) prepare_coarse(L, omega, P);
            if (precond >= 1 && L->omega_bj != omega) {
                block_jacobi_setup(L, omega);
                L->omega_bj = omega;
            }
            std::vector<double> bnorm;
            norms(L, B, P, bnorm);
            for (auto &b : bnorm) b = std::max(b, 1e-300);
            if (X0_host) {                                                     // R = B - A x0
                PQ_HIP(hipMemcpyAsync(X, X0_host, vsz * 8, hipMemcpyHostToDevice, L->stream));
                apply_operator64(L, X, R, P);
                ++L->matvecs;
                hipLaunchKernelGGL(vec_sub_kernel, dim3((unsigned)((vsz + 255) / 256)), dim3(256), 0, L->stream, B,
                                   vsz, R);
                PQ_HIP(hipGetLastError());
            } else {                                                           // x0 = 0
                PQ_HIP(hipMemsetAsync(X, 0, vsz * 8, L->stream));
                PQ_HIP(hipMemcpyAsync(R, B, vsz * 8, hipMemcpyDeviceToDevice, L->stream));
            }
            // Per Arnoldi step j the device writes, at Hd + j * hstride: the two Gram-Schmidt
            // passes (m+1 coefficients each) and the new vector's squared norm; the kernels that
            // produce them also store them at Hh + j * hstride (pinned, coherent host memory), read
            // once the step's event has fired. The host works one step behind, so the device never
            // waits for the Givens rotations.
            const size_t hstride = (size_t)(2 * (m + 1) + 1) * P * 2;
            L->Hd.reserve((size_t)m * hstride * 8);
            const size_t hbytes = (size_t)m * hstride * 8;
            if (L->Hh_bytes < hbytes) {
                if (L->Hh) (void)hipHostFree(L->Hh);
                L->Hh = nullptr;
                PQ_HIP(hipHostMalloc((void **)&L->Hh, hbytes, hipHostMallocCoherent));
                L->Hh_bytes = hbytes;
            }
            // The entries of step j's host copy that its kernels write: the first pass over j + 1
            // vectors, the second over j + 2 (its last entry |w|^2), the squared norms.
            const size_t pass = (size_t)(m + 1) * P * 2;
            auto step_slots = [&](int j, auto &&f) {
                volatile double *g = L->Hh + j * hstride;
                for (size_t e = 0; e < (size_t)(j + 1) * P * 2; ++e) f(g[e]);
                for (size_t e = 0; e < (size_t)(j + 2) * P * 2; ++e) f(g[pass + e]);
                for (size_t e = 0; e < (size_t)P * 2; ++e) f(g[2 * pass + e]);
            };
            double *Hd = L->Hd.as<double>();
            auto dots_to = [&](const double *Vb, size_t vs, int nv, const double *Wv, double *hout, double *hhost) {
                dots_launch(L, Vb, vs, nv, Wv, P, hout, hhost);
            };
            auto update_dev = [&](const double *Vb, size_t vs, int nv, const double *hdev, double sign, double *Wv) {
                const size_t count = (size_t)n * P * 4;
                hipLaunchKernelGGL(vec_update_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, L->stream, Vb,
                                   vs, nv, hdev, sign, n, P, Wv);
            };
            // Relaxed products (inexact Krylov: Simoncini & Szyld 2003, Bouras & Fraysse 2005): a
            // product's error enters the solution weighted by that step's coefficient, which is large
            // while the residual is large and small once it has fallen. So the steps of a cycle use
            // the FP64 Ll while the last known residual estimate exceeds `relax` (relative), and the
            // FP32 copy (half the bytes) after; the true residual at the cycle's end then meets the
            // FP64 system's tolerance without a refinement cycle.
            const double relax = relax_threshold();
            double estimate = 1.0;                 // the latest relative residual estimate (max over ports)
            auto step_body = [&](int j, bool exact) {
                double *h1 = Hd + j * hstride, *h2 = h1 + (size_t)(m + 1) * P * 2, *hn = h2 + (size_t)(m + 1) * P * 2;
                // the host copies, written by the same kernels (pinned, coherent memory)
                double *g1 = L->Hh + j * hstride, *g2 = g1 + (size_t)(m + 1) * P * 2, *gn = g2 + (size_t)(m + 1) * P * 2;
                // the new vector is built in place in its slot V_{j+1}: CGS2, the second pass
                // also giving |w|^2, and the last update normalising it
                double *Wj = V + (j + 1) * vsz;
                apply_precond(L, precond, V + j * vsz, Z, P);
                if (exact) apply_operator64(L, Z, Wj, P);
                else apply_operator(L, Z, Wj, P);
                dots_to(V, vsz, j + 1, Wj, h1, g1);
                update_dev(V, vsz, j + 1, h1, 1.0, Wj);
                dots_to(V, vsz, j + 2, Wj, h2, g2);
                hipLaunchKernelGGL(vec_update_norm_kernel, dim3((unsigned)(((size_t)n * P * 4 + 255) / 256)), dim3(256), 0,
                                   L->stream, V, vsz, j + 1, h2, n, P, Wj, hn, gn);
                PQ_HIP(hipGetLastError());
            };
            // Each step's kernels are captured once as a graph (step j's shapes and offsets are
            // fixed) and replayed: one launch per step instead of about fifteen.
            const bool use_graphs = L->graphs_on < 0 ? graphs_enabled() : L->graphs_on != 0;
            if (use_graphs) check_graphs(L, P, precond, m);
            auto enqueue_step = [&](int j) {
                step_slots(j, [](volatile double &v) { v = std::numeric_limits<double>::quiet_NaN(); });
                const bool exact = estimate > relax;
                if (exact) {
                    step_body(j, true);            // FP64 steps run eagerly (the graphs hold FP32 steps)
                } else if (use_graphs) {
                    if (!L->graphs[j]) {
                        hipGraph_t graph = nullptr;
                        PQ_HIP(hipStreamBeginCapture(L->stream, hipStreamCaptureModeThreadLocal));
                        try {
                            step_body(j, false);
                        } catch (...) {
                            (void)hipStreamEndCapture(L->stream, &graph);
                            if (graph) (void)hipGraphDestroy(graph);
                            throw;
                        }
                        PQ_HIP(hipStreamEndCapture(L->stream, &graph));
                        const hipError_t e = hipGraphInstantiate(&L->graphs[j], graph, nullptr, nullptr, 0);
                        (void)hipGraphDestroy(graph);
                        PQ_HIP(e);
                    }
                    PQ_HIP(hipGraphLaunch(L->graphs[j], L->stream));
                } else {
                    step_body(j, false);
                }
                ++L->matvecs;
            };
            // Step j's coefficients reach the host as the kernels store them: each entry was set to
            // NaN before the step was enqueued and every aligned 8-byte store lands whole, so the
            // host waits until no entry is NaN; no event or fence orders the stores. Once the stream
            // has gone idle, entries that are still NaN are the computed values (a diverged solve,
            // caught by the residual test); a failed stream raises its error.
            auto wait_step = [&](int j) {
                bool idle = false;
                for (unsigned long spins = 1;; ++spins) {
                    bool ready = true;
                    step_slots(j, [&](volatile double &v) { ready = ready && !std::isnan((double)v); });
                    if (ready || idle) return;
                    if ((spins & 255) != 0) continue;
                    const hipError_t e = hipStreamQuery(L->stream);
                    if (e == hipErrorNotReady) continue;
                    PQ_HIP(e);
                    idle = true;
                }
            };
            // A cycle stops once its residual estimate reaches PCBQS_AMD_INNER_TOL times tol (1 by
            // default; the FP64 true residual then decides whether another cycle runs).
            const double inner_tol = tol * inner_tol_factor();
            int its = 0, restarts = 0;
            std::vector<double> res(P, 1.0);
            std::vector<cplx> H((size_t)(m + 1) * m * P), g((size_t)(m + 1) * P), cs((size_t)m * P), sn((size_t)m * P);
            while (true) {
                std::vector<double> beta;
                norms(L, R, P, beta);
                bool done = true;
                estimate = 0;
                for (int p = 0; p < P; ++p) {
                    res[p] = beta[p] / bnorm[p];
                    done = done && res[p] <= tol;
                    estimate = std::max(estimate, res[p]);
                }
                if (done || its >= maxiter) break;
                std::vector<double> inv_beta(P);
                for (int p = 0; p < P; ++p) inv_beta[p] = beta[p] > 0 ? 1.0 / beta[p] : 0.0;
                scale_cols(L, R, inv_beta, P, V);
                std::fill(g.begin(), g.end(), cplx(0));
                for (int p = 0; p < P; ++p) g[p] = beta[p];
                const int steps = std::min(m, maxiter - its);
                int used = 0;                      // Arnoldi steps whose rotations are applied
                bool converged = false;
                auto process = [&](int j) {
                    wait_step(j);
                    const double *h1 = L->Hh + j * hstride, *h2 = h1 + (size_t)(m + 1) * P * 2;
                    const double *hn = h2 + (size_t)(m + 1) * P * 2;
                    bool all = true;
                    for (int p = 0; p < P; ++p) {
                        auto Hc = [&](int i) -> cplx & { return H[((size_t)i * m + j) * P + p]; };
                        for (int i = 0; i <= j; ++i)
                            Hc(i) = cplx(h1[(i * P + p) * 2] + h2[(i * P + p) * 2],
                                         h1[(i * P + p) * 2 + 1] + h2[(i * P + p) * 2 + 1]);
                        Hc(j + 1) = std::sqrt(std::max(hn[2 * p], 0.0));
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
                        if (res[p] > inner_tol) all = false;
                    }
                    estimate = 0;
                    for (int p = 0; p < P; ++p) estimate = std::max(estimate, res[p]);
                    used = j + 1;
                    return all;
                };
                for (int j = 0; j < steps; ++j) {
                    enqueue_step(j);
                    if (j > 0 && process(j - 1)) { converged = true; break; }
                }
                if (!converged && used < steps) process(steps - 1);
                const int jn = used;
                its += jn;
                // y from the triangular systems; X += M^-1 (V y)
                std::vector<double> y((size_t)jn * P * 2);
                for (int p = 0; p < P; ++p) {
                    std::vector<cplx> yy(jn);
                    for (int i = jn - 1; i >= 0; --i) {
                        cplx sacc = g[i * P + p];
                        for (int k = i + 1; k < jn; ++k) sacc -= H[((size_t)i * m + k) * P + p] * yy[k];
                        const cplx d = H[((size_t)i * m + i) * P + p];
                        yy[i] = std::abs(d) > 0 ? sacc / d : cplx(0);
                    }
                    for (int i = 0; i < jn; ++i) { y[(i * P + p) * 2] = yy[i].real(); y[(i * P + p) * 2 + 1] = yy[i].imag(); }
                }
                PQ_HIP(hipMemsetAsync(W, 0, vsz * 8, L->stream));
                if (jn > 0) update(L, V, vsz, jn, y.data(), -1.0, P, W);          // W = V y
                apply_precond(L, precond, W, Z, P);
                {
                    std::vector<double> one((size_t)P * 2, 0.0);
                    for (int p = 0; p < P; ++p) one[2 * p] = 1.0;
                    update(L, Z, 0, 1, one.data(), -1.0, P, X);                    // X += Z
                }
                apply_operator64(L, X, R, P);                             // true residual, FP64 system
                ++L->matvecs;
                hipLaunchKernelGGL(vec_sub_kernel, dim3((unsigned)((vsz + 255) / 256)), dim3(256), 0, L->stream, B, vsz, R);
                PQ_HIP(hipGetLastError());
                ++restarts;
            }
            PQ_HIP(hipStreamSynchronize(L->stream));
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
        port_impedance(L, omega, P, port_cols, X, Zp);
        if (X_host && n > 0) PQ_HIP(hipMemcpy(X_host, X, vsz * 8, hipMemcpyDeviceToHost));
        L->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    });
}

// The direct port solve at one frequency: Z_ff = Rl + j w Ll formed in FP32 on the device and
// factored as L D L^T without pivoting (both parts of Z_ff are symmetric positive definite), then
// X = -Z_ff^-1 Z_fs refined in FP64 as src/pcbqs/kernels.py:125-181 (MixedLU): the FP32 solution is
// corrected by its residual, taken with the FP64 operator product, until the largest column's
// relative residual is at most tol, for at most max_steps corrections, stopping early once a
// correction no longer halves it; a result above 1e-10 then raises PQ_ECONVERGE. The factor is
// kept for the next call at the same frequency.
// Form and factor Z_ff at nw frequencies together (one launch per panel step for all of them);
// later direct solves at these frequencies use these factors. *ms: device time of the batch.
int pq_loops_factor_batch(pq_loops *L, int nw, const double *omegas, double *ms) {
    return guard([&] {
        PQ_REQUIRE(L && omegas && nw >= 1, "pq_loops_factor_batch: invalid arguments");
        for (int b = 0; b < nw; ++b)
            PQ_REQUIRE(std::isfinite(omegas[b]) && omegas[b] >= 0, "pq_loops_factor_batch: invalid frequency");
        DeviceScope scope(L->device);
        const int n = L->nf;
        L->batch_omega.clear();
        L->batch_ms = 0;
        if (n <= 0) return;
        const int nblk = (n + LDL_BLOCK - 1) / LDL_BLOCK;
        const long long ldf = (long long)nblk * LDL_BLOCK;
        const size_t sF = (size_t)2 * n * ldf, sX = (size_t)2 * nblk * LDL_BLOCK * LDL_BLOCK,
                     sW = (size_t)4 * n * 2 * LDL_BLOCK;
        L->ldf = ldf;
        L->Fbat.reserve(sF * nw * 4);
        L->Xbat.reserve(sX * nw * 4);
        L->Wbat.reserve(sW * nw * 4);
        if (!L->aux) {
            ldl_create_aux(&L->aux);
            PQ_HIP(hipEventCreateWithFlags(&L->ev_ready, hipEventDisableTiming));
            PQ_HIP(hipEventCreateWithFlags(&L->ev_done, hipEventDisableTiming));
        }
        L->batch_omega.assign(omegas, omegas + nw);
        hipEvent_t e0, e1;
        PQ_HIP(hipEventCreate(&e0));
        PQ_HIP(hipEventCreate(&e1));
        PQ_HIP(hipEventRecord(e0, L->stream));
        for (int b = 0; b < nw; ++b)
            ldl_form(L->stream, batch_view(L, b), L->Ll.as<float>(), L->ldl, L->rptr.as<int>(), L->rcol.as<int>(),
                     L->rval.as<double>(), omegas[b]);
        LdlFactor f = batch_view(L, 0);
        f.batch = nw;
        f.sF = (long long)sF; f.sX = (long long)sX; f.sW = (long long)sW;
        // one flag per matrix: a zero or non-finite FP32 pivot sends that frequency to the FP64
        // fallback in the direct solves that use this batch
        L->singular.reserve((size_t)nw * 4);
        PQ_HIP(hipMemsetAsync(L->singular.p, 0, (size_t)nw * 4, L->stream));
        ldl_factor(L->stream, f, L->singular.as<int>());
        PQ_HIP(hipEventRecord(e1, L->stream));
        L->batch_bad.assign(nw, 0);
        PQ_HIP(hipMemcpyAsync(L->batch_bad.data(), L->singular.p, (size_t)nw * 4, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
        float t = 0;
        (void)hipEventElapsedTime(&t, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        L->batch_ms = t;
        if (ms) *ms = t;
    });
}

int pq_loops_solve_direct(pq_loops *L, double omega, int P, const int32_t *port_cols, double tol, int max_steps,
                          double *Zp, double *X_host, int64_t *steps, double *residual) {
    return guard([&] {
        PQ_REQUIRE(L && port_cols && Zp, "pq_loops_solve_direct: null argument");
        PQ_REQUIRE((P >= 1 && P <= 8) || P == 16, "pq_loops_solve_direct: 1 to 8 or 16 ports");
        PQ_REQUIRE(tol > 0 && tol < 1 && max_steps >= 0, "pq_loops_solve_direct: invalid options");
        for (int p = 0; p < P; ++p)
            PQ_REQUIRE(port_cols[p] >= L->nf && port_cols[p] < L->nl, "pq_loops_solve_direct: port loop out of range");
        DeviceScope scope(L->device);
        const auto started = std::chrono::steady_clock::now();
        const int n = L->nf, NR = 2 * P;
        const size_t vsz = (size_t)n * NR;
        L->matvecs = 0;
        L->sc.reserve(P * 4);
        PQ_HIP(hipMemcpyAsync(L->sc.p, port_cols, P * 4, hipMemcpyHostToDevice, L->stream));
        L->V.reserve(2 * vsz * 8);
        L->W.reserve(vsz * 8); L->Z.reserve(vsz * 8); L->X.reserve(vsz * 8); L->B.reserve(vsz * 8);
        L->partial.reserve((size_t)RED_BLOCKS * P * 2 * 8);
        L->h.reserve((size_t)P * 2 * 8);
        double *X = L->X.as<double>(), *B = L->B.as<double>(), *W = L->W.as<double>(), *D = L->Z.as<double>();
        double *Y = L->V.as<double>(), *Ys = Y + vsz;
        int corrections = 0;
        double worst = 0;
        if (n > 0) {
            tune_gemv(L, NR);
            set_omega(L, omega);
            hipLaunchKernelGGL(rhs_kernel, dim3((n * P + 255) / 256), dim3(256), 0, L->stream, L->Ll64.as<double>(), L->ldl,
                               L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, L->sc.as<int>(), P, omega, B);
            PQ_HIP(hipGetLastError());
            const int nblk = (n + LDL_BLOCK - 1) / LDL_BLOCK;
            L->ldf = (long long)nblk * LDL_BLOCK;
            L->Fldl.reserve((size_t)2 * n * L->ldf * 4);
            L->Xldl.reserve((size_t)2 * nblk * LDL_BLOCK * LDL_BLOCK * 4);
            L->Wldl.reserve((size_t)4 * n * 2 * LDL_BLOCK * 4);
            if (!L->aux) {
                ldl_create_aux(&L->aux);
                PQ_HIP(hipEventCreateWithFlags(&L->ev_ready, hipEventDisableTiming));
                PQ_HIP(hipEventCreateWithFlags(&L->ev_done, hipEventDisableTiming));
            }
            const size_t wsz = (size_t)n * 2 * LDL_BLOCK;
            float *Wb = L->Wldl.as<float>();
            LdlFactor f{L->Fldl.as<float>(), L->Fldl.as<float>() + (size_t)n * L->ldf, L->ldf, n,
                        L->Xldl.as<float>(), L->Xldl.as<float>() + (size_t)nblk * LDL_BLOCK * LDL_BLOCK,
                        {Wb, Wb + 2 * wsz}, {Wb + wsz, Wb + 3 * wsz}, L->aux, L->ev_ready, L->ev_done};
            // a factor of this frequency from the last batch, if there is one
            bool bad_pivot = false;
            const auto hit = std::find(L->batch_omega.begin(), L->batch_omega.end(), omega);
            if (hit != L->batch_omega.end()) {
                const int b = (int)(hit - L->batch_omega.begin());
                f = batch_view(L, b);
                bad_pivot = L->batch_bad[b] != 0;
                L->factor_ms = L->batch_ms / (double)L->batch_omega.size();
            } else if (L->omega_ldl != omega) {
                L->omega_ldl = -1;
                hipEvent_t e0, e1;
                PQ_HIP(hipEventCreate(&e0));
                PQ_HIP(hipEventCreate(&e1));
                PQ_HIP(hipEventRecord(e0, L->stream));
                ldl_form(L->stream, f, L->Ll.as<float>(), L->ldl, L->rptr.as<int>(), L->rcol.as<int>(),
                         L->rval.as<double>(), omega);
                PQ_HIP(hipMemsetAsync(L->singular.p, 0, 4, L->stream));
                ldl_factor(L->stream, f, L->singular.as<int>());
                PQ_HIP(hipEventRecord(e1, L->stream));
                int bad = 0;
                PQ_HIP(hipMemcpyAsync(&bad, L->singular.p, 4, hipMemcpyDeviceToHost, L->stream));
                PQ_HIP(hipStreamSynchronize(L->stream));
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                (void)hipEventDestroy(e0);
                (void)hipEventDestroy(e1);
                bad_pivot = bad != 0;
                L->factor_ms = ms;
                L->omega_ldl = bad_pivot ? -1 : omega;      // a failed factor is not kept
            }
            std::vector<double> bnorm;
            norms(L, B, P, bnorm);
            for (auto &b : bnorm) b = std::max(b, 1e-300);
            const size_t count = vsz;
            if (!bad_pivot) {
                ldl_solve(L->stream, f, P, B, Y, Ys, X);
                double last = std::numeric_limits<double>::infinity();
                for (int step = 0;; ++step) {
                    worst = relative_residual64(L, B, X, W, P, bnorm);        // W = B - Z X (FP64 system)
                    if (!(worst > tol) || step == max_steps) break;
                    if (!(worst <= 0.5 * last)) break;                         // no longer converging
                    last = worst;
                    ldl_solve(L->stream, f, P, W, Y, Ys, D);                    // X += Z^-1 (B - Z X)
                    hipLaunchKernelGGL(vec_add_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                                       L->stream, D, count, X);
                    PQ_HIP(hipGetLastError());
                    ++corrections;
                }
            }
            L->fallbacks = 0;
            if (bad_pivot || !(worst <= std::max(tol, 1e-10))) {             // FP64 LU, as MixedLU
                fallback_solve(L, omega, P, B, X);
                worst = relative_residual64(L, B, X, W, P, bnorm);
                L->fallbacks = 1;
            }
        } else {
            L->fallbacks = 0;
        }
        L->iterations = corrections;
        L->restarts = 0;
        L->residual = worst;
        if (steps) *steps = corrections;
        if (residual) *residual = worst;
        if (!(worst <= std::max(tol, 1e-10)))
            throw Error(PQ_ECONVERGE, "direct solve: relative residual " + std::to_string(worst) +
                                          " after the FP64 fallback factor");
        port_impedance(L, omega, P, port_cols, X, Zp);
        if (X_host && n > 0) PQ_HIP(hipMemcpy(X_host, X, vsz * 8, hipMemcpyDeviceToHost));
        L->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    });
}

// pq_loops_solve_direct at nw frequencies at once: the factorisations batched
// (pq_loops_factor_batch), then the refinement in lockstep: every frequency's solves with the
// factor run in the same launches, the residual products one frequency at a time, and each
// frequency keeps its own stopping rule (a frequency that has stopped takes no more corrections),
// so every frequency gets exactly the operations of its own pq_loops_solve_direct.
int pq_loops_solve_direct_batch(pq_loops *L, int nw, const double *omegas, int P, const int32_t *port_cols,
                                double tol, int max_steps, double *Zp, int64_t *steps, double *residual) {
    return guard([&] {
        PQ_REQUIRE(L && omegas && port_cols && Zp && steps && residual && nw >= 1,
                   "pq_loops_solve_direct_batch: null argument");
        PQ_REQUIRE((P >= 1 && P <= 8) || P == 16, "pq_loops_solve_direct_batch: 1 to 8 or 16 ports");
        PQ_REQUIRE(tol > 0 && tol < 1 && max_steps >= 0, "pq_loops_solve_direct_batch: invalid options");
        for (int p = 0; p < P; ++p)
            PQ_REQUIRE(port_cols[p] >= L->nf && port_cols[p] < L->nl,
                       "pq_loops_solve_direct_batch: port loop out of range");
        const int n = L->nf, NR = 2 * P;
        DeviceScope scope(L->device);
        const auto started = std::chrono::steady_clock::now();
        L->sc.reserve(P * 4);
        PQ_HIP(hipMemcpyAsync(L->sc.p, port_cols, P * 4, hipMemcpyHostToDevice, L->stream));
        if (n <= 0) {
            for (int b = 0; b < nw; ++b) {
                port_impedance(L, omegas[b], P, port_cols, nullptr, Zp + (size_t)b * P * P * 2);
                steps[b] = 0;
                residual[b] = 0;
            }
            return;
        }
        const int rc = pq_loops_factor_batch(L, nw, omegas, nullptr);
        if (rc != PQ_OK) throw Error(rc, pq_last_error());
        const size_t vsz = (size_t)n * NR;
        DevBuf Bv, Xv, Wv, Dv, Yv, Ysv;
        for (DevBuf *d : {&Bv, &Xv, &Wv, &Dv, &Yv, &Ysv}) d->reserve(vsz * nw * 8);
        L->partial.reserve((size_t)RED_BLOCKS * P * 2 * 8);
        L->h.reserve((size_t)P * 2 * 8);
        double *B = Bv.as<double>(), *X = Xv.as<double>(), *W = Wv.as<double>(), *D = Dv.as<double>();
        tune_gemv(L, NR);
        LdlFactor f = batch_view(L, 0);
        const int nblk = (n + LDL_BLOCK - 1) / LDL_BLOCK;
        f.batch = nw;
        f.sF = (long long)2 * n * L->ldf;
        f.sX = (long long)2 * nblk * LDL_BLOCK * LDL_BLOCK;
        f.sW = (long long)4 * n * 2 * LDL_BLOCK;
        std::vector<std::vector<double>> bnorm(nw);
        std::vector<double> rn, worst(nw, 0.0), last(nw, std::numeric_limits<double>::infinity());
        std::vector<int> corr(nw, 0);
        std::vector<char> active(nw, 1);
        for (int b = 0; b < nw; ++b)
            if (L->batch_bad[b]) active[b] = 0;          // no usable FP32 factor: the FP64 fallback below
        for (int b = 0; b < nw; ++b) {
            hipLaunchKernelGGL(rhs_kernel, dim3((n * P + 255) / 256), dim3(256), 0, L->stream, L->Ll64.as<double>(),
                               L->ldl, L->rptr.as<int>(), L->rcol.as<int>(), L->rval.as<double>(), n, L->sc.as<int>(), P,
                               omegas[b], B + b * vsz);
            PQ_HIP(hipGetLastError());
            norms(L, B + b * vsz, P, bnorm[b]);
            for (auto &v : bnorm[b]) v = std::max(v, 1e-300);
        }
        ldl_solve(L->stream, f, P, B, Yv.as<double>(), Ysv.as<double>(), X);
        L->matvecs = 0;
        for (int step = 0;; ++step) {
            bool any = false;
            for (int b = 0; b < nw; ++b) {
                if (!active[b]) continue;
                PQ_HIP(hipStreamSynchronize(L->stream));                 // earlier products have read omega
                set_omega(L, omegas[b]);
                const double w = relative_residual64(L, B + b * vsz, X + b * vsz, W + b * vsz, P, bnorm[b]);
                worst[b] = w;
                if (!(w > tol) || step == max_steps || !(w <= 0.5 * last[b])) {
                    active[b] = 0;
                    continue;
                }
                last[b] = w;
                any = true;
            }
            if (!any) break;
            ldl_solve(L->stream, f, P, W, Yv.as<double>(), Ysv.as<double>(), D);   // every system at once
            for (int b = 0; b < nw; ++b)
                if (active[b]) {
                    hipLaunchKernelGGL(vec_add_kernel, dim3((unsigned)((vsz + 255) / 256)), dim3(256), 0, L->stream,
                                       D + b * vsz, vsz, X + b * vsz);
                    ++corr[b];
                }
            PQ_HIP(hipGetLastError());
        }
        // frequencies the FP32 factor could not bring to the reference's level: the FP64 LU, one
        // frequency at a time (the others keep their refined solutions)
        L->fallbacks = 0;
        for (int b = 0; b < nw; ++b)
            if (L->batch_bad[b] || !(worst[b] <= std::max(tol, 1e-10))) {
                PQ_HIP(hipStreamSynchronize(L->stream));
                set_omega(L, omegas[b]);
                fallback_solve(L, omegas[b], P, B + b * vsz, X + b * vsz);
                worst[b] = relative_residual64(L, B + b * vsz, X + b * vsz, W + b * vsz, P, bnorm[b]);
                ++L->fallbacks;
            }
        double wmax = 0;
        for (int b = 0; b < nw; ++b) {
            steps[b] = corr[b];
            residual[b] = worst[b];
            wmax = std::isnan(worst[b]) ? worst[b] : std::max(wmax, worst[b]);
        }
        if (!(wmax <= std::max(tol, 1e-10)))
            throw Error(PQ_ECONVERGE, "direct solve: relative residual " + std::to_string(wmax) +
                                          " after the FP64 fallback factor");
        for (int b = 0; b < nw; ++b)
            port_impedance(L, omegas[b], P, port_cols, X + b * vsz, Zp + (size_t)b * P * P * 2);
        L->iterations = *std::max_element(corr.begin(), corr.end());
        L->restarts = 0;
        L->residual = wmax;
        L->factor_ms = L->batch_ms / nw;
        L->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - started).count();
    });
}

// Direct port solves of several boards at n_omega frequencies each, every batch of boards in one
// captured graph launch. Boards are taken largest first and grouped while their free-loop counts
// stay within 3/4 of the group's largest (padding: the identity beyond a board's own loops) and
// the group's factors fit in about 24 GB (boards of up to 8 ports and of 9 to 16 in separate
// groups, so that the solves carry no more columns than needed). Per group the graph holds: the nz =
// boards x frequencies
// matrices formed in FP32 and factored as L D L^T in the same launches; the right-hand sides;
// the solve; GRAPH_STEPS refinement steps against each board's FP64 system with every matrix's own
// stopping rule (kernels.py:169-178) kept on the device; the port reduction. Matrices the graph
// leaves above max(tol, 1e-10) (or with a zero FP32 pivot) are solved again one by one by
// pq_loops_solve_direct (more corrections, and the FP64 LU where those stall). Zp: per board in
// order, (n_omega, P_b, P_b) interleaved complex; steps, residual: (count, n_omega); stats[0] the
// graph launches, stats[1] the matrices solved again one by one.
constexpr int GRAPH_STEPS = 2;

int pq_loops_solve_direct_multi(int count, pq_loops *const *handles, const int *Ps, const int32_t *const *port_cols,
                                int nw, const double *omegas, double tol, int max_steps, double *Zp, int64_t *steps,
                                double *residual, int64_t *stats) {
    return guard([&] {
        PQ_REQUIRE(count >= 1 && handles && Ps && port_cols && nw >= 1 && omegas && Zp && steps && residual,
                   "pq_loops_solve_direct_multi: null argument");
        PQ_REQUIRE(tol > 0 && tol < 1 && max_steps >= 0, "pq_loops_solve_direct_multi: invalid options");
        const int device = handles[0]->device;
        for (int b = 0; b < count; ++b) {
            const pq_loops *L = handles[b];
            PQ_REQUIRE(L && L->device == device, "pq_loops_solve_direct_multi: handles on one device");
            PQ_REQUIRE((Ps[b] >= 1 && Ps[b] <= 8) || Ps[b] == 16, "pq_loops_solve_direct_multi: 1 to 8 or 16 ports");
            for (int p = 0; p < Ps[b]; ++p)
                PQ_REQUIRE(port_cols[b][p] >= L->nf && port_cols[b][p] < L->nl,
                           "pq_loops_solve_direct_multi: port loop out of range");
        }
        for (int f = 0; f < nw; ++f)
            PQ_REQUIRE(std::isfinite(omegas[f]) && omegas[f] >= 0, "pq_loops_solve_direct_multi: invalid frequency");
        DeviceScope scope(device);
        std::vector<size_t> zoff(count + 1, 0);              // each board's offset into Zp
        for (int b = 0; b < count; ++b) zoff[b + 1] = zoff[b] + (size_t)nw * Ps[b] * Ps[b] * 2;
        std::vector<int> order(count);
        for (int b = 0; b < count; ++b) order[b] = b;
        // boards of up to 8 ports and of 9 to 16 (16 device columns) apart, each set largest first
        auto wide = [&](int b) { return Ps[b] > 8; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return wide(a) != wide(b) ? wide(a) : handles[a]->nf > handles[b]->nf;
        });
        int64_t launches = 0, redone = 0;
        hipStream_t st = nullptr;
        PQ_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { if (s) { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); } }
        } sguard{st};
        const size_t budget = (size_t)24 << 30;
        // the groups' buffers, grown as needed and reused from group to group
        DevBuf dm, F, Xi, Wb, B, X, W, D, Y, Ys, bn2, res, last, state, bad, zp, scs;
        for (size_t g0 = 0; g0 < order.size();) {
            const int nmax = std::max(handles[order[g0]]->nf, 1);
            const int nblk = (nmax + LDL_BLOCK - 1) / LDL_BLOCK, nB = nblk * LDL_BLOCK;
            const long long ldf = nB;
            const size_t sF = (size_t)2 * nB * ldf, sX = (size_t)2 * nblk * LDL_BLOCK * LDL_BLOCK,
                         sW = (size_t)4 * nB * 2 * LDL_BLOCK;
            const size_t per_matrix = (sF + sX + sW) * 4 + (size_t)6 * nB * 32 * 8;
            size_t g1 = g0 + 1;
            while (g1 < order.size() && wide(order[g1]) == wide(order[g0]) && 4 * handles[order[g1]]->nf >= 3 * nmax &&
                   (g1 - g0 + 1) * nw * per_matrix <= budget && (g1 - g0 + 1) * nw <= 65535)
                ++g1;
            const int nb_boards = (int)(g1 - g0), nz = nb_boards * nw;
            int P = 1;
            for (size_t g = g0; g < g1; ++g) P = std::max(P, Ps[order[g]]);
            if (P > 8) P = 16;
            const int NR = 2 * P;
            // the group's matrices: board-major, frequencies within
            std::vector<BatchMat> mats(nz);
            scs.reserve((size_t)nb_boards * 16 * 4);
            {
                std::vector<int32_t> sch((size_t)nb_boards * 16, 0);
                for (int i = 0; i < nb_boards; ++i)
                    for (int p = 0; p < Ps[order[g0 + i]]; ++p) sch[(size_t)i * 16 + p] = port_cols[order[g0 + i]][p];
                PQ_HIP(hipMemcpyAsync(scs.p, sch.data(), sch.size() * 4, hipMemcpyHostToDevice, st));
                PQ_HIP(hipStreamSynchronize(st));
            }
            for (int i = 0; i < nb_boards; ++i) {
                const pq_loops *L = handles[order[g0 + i]];
                for (int f = 0; f < nw; ++f) {
                    BatchMat &m = mats[(size_t)i * nw + f];
                    m.Ll32 = L->Ll.as<float>(); m.Ll64 = L->Ll64.as<double>(); m.ldl = L->ldl;
                    m.rptr = L->rptr.as<int>(); m.rcol = L->rcol.as<int>(); m.rval = L->rval.as<double>();
                    m.sc = scs.as<int>() + (size_t)i * 16;
                    m.n = L->nf; m.P = Ps[order[g0 + i]]; m.omega = omegas[f];
                }
            }
            dm.reserve(sizeof(BatchMat) * nz);
            PQ_HIP(hipMemcpyAsync(dm.p, mats.data(), sizeof(BatchMat) * nz, hipMemcpyHostToDevice, st));
            F.reserve(sF * nz * 4); Xi.reserve(sX * nz * 4); Wb.reserve(sW * nz * 4);
            const size_t vs = (size_t)nB * NR;
            for (DevBuf *v : {&B, &X, &W, &D, &Y, &Ys}) v->reserve(vs * nz * 8);
            bn2.reserve((size_t)nz * P * 8); res.reserve((size_t)nz * 8); last.reserve((size_t)nz * 8);
            state.reserve((size_t)nz * 2 * 4); bad.reserve((size_t)nz * 4); zp.reserve((size_t)nz * P * P * 2 * 8);
            PQ_HIP(hipStreamSynchronize(st));
            const size_t wsz = (size_t)nB * 2 * LDL_BLOCK;
            float *Wf = Wb.as<float>();
            LdlFactor fct{F.as<float>(), F.as<float>() + (size_t)nB * ldf, ldf, nB, Xi.as<float>(),
                          Xi.as<float>() + (size_t)nblk * LDL_BLOCK * LDL_BLOCK, {Wf, Wf + 2 * wsz},
                          {Wf + wsz, Wf + 3 * wsz}, nullptr, nullptr, nullptr};
            fct.batch = nz;
            fct.sF = (long long)sF; fct.sX = (long long)sX; fct.sW = (long long)sW;
            const BatchMat *dmat = dm.as<BatchMat>();
            const int S = std::min(GRAPH_STEPS, max_steps);
            auto body = [&]() {
                const unsigned zb = (unsigned)((nz + 255) / 256);
                hipLaunchKernelGGL(batch_init_kernel, dim3(zb), dim3(256), 0, st, nz, last.as<double>(), state.as<int>());
                hipLaunchKernelGGL(fill_int_kernel, dim3(zb), dim3(256), 0, st, bad.as<int>(), nz, 0);
                hipLaunchKernelGGL(batch_form_kernel, dim3((nB + 255) / 256, std::min(nB, 4096), nz), dim3(256), 0, st,
                                   dmat, nB, F.as<float>(), ldf, (long long)sF);
                hipLaunchKernelGGL(batch_form_r_kernel, dim3((nB + 255) / 256, 1, nz), dim3(256), 0, st, dmat, nB,
                                   F.as<float>(), ldf, (long long)sF);
                ldl_factor(st, fct, bad.as<int>());
                hipLaunchKernelGGL(batch_rhs_kernel, dim3((nB * P + 255) / 256, 1, nz), dim3(256), 0, st, dmat, nB, P,
                                   B.as<double>());
                hipLaunchKernelGGL(batch_colnorm_kernel, dim3(nz), dim3(256), 0, st, dmat, nB, P, B.as<double>(),
                                   nullptr, bn2.as<double>(), nullptr);
                ldl_solve(st, fct, P, B.as<double>(), Y.as<double>(), Ys.as<double>(), X.as<double>());
                for (int step = 0; step <= S; ++step) {
                    dispatch_nr(NR, [&](auto nr) {
                        constexpr int N = decltype(nr)::value;
                        hipLaunchKernelGGL(batch_residual_kernel<N>, dim3((nB + 255) / 256, 1, nz), dim3(256), 0, st,
                                           dmat, nB, X.as<double>(), B.as<double>(), W.as<double>());
                    });
                    hipLaunchKernelGGL(batch_colnorm_kernel, dim3(nz), dim3(256), 0, st, dmat, nB, P, W.as<double>(),
                                       bn2.as<double>(), nullptr, res.as<double>());
                    hipLaunchKernelGGL(batch_control_kernel, dim3(zb), dim3(256), 0, st, nz, res.as<double>(),
                                       last.as<double>(), state.as<int>(), step, tol, S);
                    if (step == S) break;
                    ldl_solve(st, fct, P, W.as<double>(), Y.as<double>(), Ys.as<double>(), D.as<double>());
                    hipLaunchKernelGGL(batch_add_kernel, dim3((unsigned)((vs + 255) / 256), 1, nz), dim3(256), 0, st,
                                       nB, NR, state.as<int>(), D.as<double>(), X.as<double>());
                }
                hipLaunchKernelGGL(batch_port_kernel, dim3(P * P, 1, nz), dim3(256), 0, st, dmat, nB, P, X.as<double>(),
                                   zp.as<double>());
                PQ_HIP(hipGetLastError());
            };
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            PQ_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                body();
            } catch (...) {
                (void)hipStreamEndCapture(st, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            PQ_HIP(hipStreamEndCapture(st, &graph));
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            PQ_HIP(ie);
            const hipError_t le = hipGraphLaunch(exec, st);
            const hipError_t se = le == hipSuccess ? hipStreamSynchronize(st) : le;
            (void)hipGraphExecDestroy(exec);
            PQ_HIP(se);
            ++launches;
            std::vector<double> zh((size_t)nz * P * P * 2), rh(nz);
            std::vector<int> sh((size_t)nz * 2), bh(nz);
            PQ_HIP(hipMemcpy(zh.data(), zp.p, zh.size() * 8, hipMemcpyDeviceToHost));
            PQ_HIP(hipMemcpy(rh.data(), res.p, rh.size() * 8, hipMemcpyDeviceToHost));
            PQ_HIP(hipMemcpy(sh.data(), state.p, sh.size() * 4, hipMemcpyDeviceToHost));
            PQ_HIP(hipMemcpy(bh.data(), bad.p, bh.size() * 4, hipMemcpyDeviceToHost));
            for (int i = 0; i < nb_boards; ++i) {
                const int b = order[g0 + i], Pb = Ps[b];
                for (int f = 0; f < nw; ++f) {
                    const int z = i * nw + f;
                    double *out = Zp + zoff[b] + (size_t)f * Pb * Pb * 2;
                    if (bh[z] || !(rh[z] <= std::max(tol, 1e-10))) {
                        // not brought to the reference's level in the graph: one by one, with
                        // the full refinement and the FP64 fallback
                        int64_t st1 = 0;
                        double r1 = 0;
                        const int rc = pq_loops_solve_direct(handles[b], omegas[f], Pb, port_cols[b], tol, max_steps,
                                                             out, nullptr, &st1, &r1);
                        if (rc != PQ_OK) throw Error(rc, pq_last_error());
                        steps[(size_t)b * nw + f] = st1;
                        residual[(size_t)b * nw + f] = r1;
                        ++redone;
                        continue;
                    }
                    for (int s2 = 0; s2 < Pb; ++s2)
                        for (int q = 0; q < Pb; ++q) {
                            out[2 * (s2 * Pb + q)] = zh[((size_t)z * P * P + s2 * P + q) * 2];
                            out[2 * (s2 * Pb + q) + 1] = zh[((size_t)z * P * P + s2 * P + q) * 2 + 1];
                        }
                    steps[(size_t)b * nw + f] = sh[2 * z + 1];
                    residual[(size_t)b * nw + f] = rh[z];
                }
            }
            g0 = g1;
        }
        if (stats) { stats[0] = launches; stats[1] = redone; }
    });
}

// T^T P T / pscale of the tree branches (ends ta, tb: host int32, nt of them) into E_dev (nt x nt
// FP64, row stride lde), P_dev the node potential matrix (n_nodes square, row stride ldp): the charge
// solve's electrostatic block (src/pcbqs/peec.py:178-180).
int pq_peec_electro(int device, const double *P_dev, int64_t ldp, int n_nodes, int nt, const int32_t *ta,
                    const int32_t *tb, double pscale, double *E_dev, int64_t lde) {
    return guard([&] {
        PQ_REQUIRE(P_dev && E_dev && (nt == 0 || (ta && tb)), "pq_peec_electro: null argument");
        PQ_REQUIRE(nt >= 0 && n_nodes > 0 && ldp >= n_nodes && lde >= nt && pscale > 0,
                   "pq_peec_electro: invalid sizes");
        for (int i = 0; i < nt; ++i)
            PQ_REQUIRE(ta[i] >= 0 && ta[i] < n_nodes && tb[i] >= 0 && tb[i] < n_nodes,
                       "pq_peec_electro: tree branch end out of range");
        if (nt == 0) return;
        DeviceScope scope(device);
        DevBuf a, b;
        a.reserve((size_t)nt * 4); b.reserve((size_t)nt * 4);
        PQ_HIP(hipMemcpy(a.p, ta, (size_t)nt * 4, hipMemcpyHostToDevice));
        PQ_HIP(hipMemcpy(b.p, tb, (size_t)nt * 4, hipMemcpyHostToDevice));
        hipLaunchKernelGGL(peec_electro_kernel, dim3((unsigned)(((size_t)nt * nt + 255) / 256)), dim3(256), 0, nullptr,
                           P_dev, (long long)ldp, nt, a.as<int>(), b.as<int>(), 1.0 / pscale, E_dev, (long long)lde);
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipDeviceSynchronize());
    });
}

// The charge solve at one frequency (src/pcbqs/peec.py:236-265 with kernels.py:63-78): on the
// handle's n = n_free unknowns (nl loop currents, then the n - nl tree coordinates), Z = Rl + j w
// Ll (FP64), zscale = max |Z|, the system S = Z / zscale on the first nl columns and j w Z / pscale
// on the others, with E (n - nl square, device) added to the tree block; S X = B / zscale solved by
// an FP64 LU with partial pivoting (rocSOLVER zgetrf / zgetrs), as the reference's
// lu_factor_rows / lu_solve_rows. B, X: (n, nrhs) complex host, interleaved, row-major.
int pq_peec_solve(pq_loops *L, int nl, const double *E_dev, int64_t lde, double pscale, double omega, int nrhs,
                  const double *B_host, double *X_host, double *zscale_out) {
    return guard([&] {
        PQ_REQUIRE(L && B_host && X_host && (E_dev || L->nf == nl), "pq_peec_solve: null argument");
        const int n = L->nf;
        PQ_REQUIRE(nl >= 0 && nl <= n && nrhs >= 1 && pscale > 0 && omega > 0 && std::isfinite(omega) &&
                       lde >= n - nl, "pq_peec_solve: invalid arguments");
        PQ_REQUIRE(n <= 40000, "pq_peec_solve: at most 40000 unknowns (the dense system)");
        DeviceScope scope(L->device);
        if (n == 0) return;
        const RocLib &lu = roc_lib();
        rocblas_handle h = blas_handle(L);
        DevBuf S, piv, info, Bc, mx;
        S.reserve((size_t)n * n * 16);
        piv.reserve((size_t)n * 4); info.reserve(4); mx.reserve(8);
        Bc.reserve((size_t)n * nrhs * 16);
        const size_t cnt = (size_t)n * n;
        const unsigned g = (unsigned)((cnt + 255) / 256);
        hipLaunchKernelGGL(peec_form_kernel, dim3(g), dim3(256), 0, L->stream, L->Ll64.as<double>(), L->ldl, n, omega,
                           S.as<double2>());
        hipLaunchKernelGGL(peec_form_r_kernel, dim3((n + 255) / 256), dim3(256), 0, L->stream, L->rptr.as<int>(),
                           L->rcol.as<int>(), L->rval.as<double>(), n, S.as<double2>());
        PQ_HIP(hipMemsetAsync(mx.p, 0, 8, L->stream));
        hipLaunchKernelGGL(peec_absmax_kernel, dim3(1024), dim3(256), 0, L->stream, S.as<double2>(), cnt,
                           mx.as<unsigned long long>());
        PQ_HIP(hipGetLastError());
        double zscale = 0;
        PQ_HIP(hipMemcpyAsync(&zscale, mx.p, 8, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
        zscale = std::max(zscale, std::numeric_limits<double>::min());
        hipLaunchKernelGGL(peec_scale_kernel, dim3(g), dim3(256), 0, L->stream, S.as<double2>(), n, nl, zscale, omega,
                           pscale, E_dev, (long long)lde);
        PQ_HIP(hipGetLastError());
        {   // B / zscale, column-major
            std::vector<double> bc((size_t)n * nrhs * 2);
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < nrhs; ++c) {
                    bc[2 * ((size_t)c * n + i)] = B_host[2 * ((size_t)i * nrhs + c)] / zscale;
                    bc[2 * ((size_t)c * n + i) + 1] = B_host[2 * ((size_t)i * nrhs + c) + 1] / zscale;
                }
            PQ_HIP(hipMemcpyAsync(Bc.p, bc.data(), bc.size() * 8, hipMemcpyHostToDevice, L->stream));
            auto *A = reinterpret_cast<rocblas_double_complex *>(S.p);
            blas_check(lu.getrf(h, n, n, A, n, piv.as<rocblas_int>(), info.as<rocblas_int>()), "rocsolver_zgetrf");
            rocblas_int bad = 0;
            PQ_HIP(hipMemcpyAsync(&bad, info.p, 4, hipMemcpyDeviceToHost, L->stream));
            PQ_HIP(hipStreamSynchronize(L->stream));
            PQ_REQUIRE(bad == 0, "pq_peec_solve: the charge system is singular");
            blas_check(lu.getrs(h, rocblas_operation_none, n, nrhs, A, n, piv.as<rocblas_int>(),
                                reinterpret_cast<rocblas_double_complex *>(Bc.p), n),
                       "rocsolver_zgetrs");
            PQ_HIP(hipMemcpyAsync(bc.data(), Bc.p, bc.size() * 8, hipMemcpyDeviceToHost, L->stream));
            PQ_HIP(hipStreamSynchronize(L->stream));
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < nrhs; ++c) {
                    X_host[2 * ((size_t)i * nrhs + c)] = bc[2 * ((size_t)c * n + i)];
                    X_host[2 * ((size_t)i * nrhs + c) + 1] = bc[2 * ((size_t)c * n + i) + 1];
                }
        }
        if (zscale_out) *zscale_out = zscale;
    });
}

// Device time (ms) of the factorisation (forming Z_ff and its L D L^T) that the last direct solve used.
int pq_loops_factor_time(pq_loops *L, double *ms) {
    return guard([&] {
        PQ_REQUIRE(L && ms, "pq_loops_factor_time: null argument");
        *ms = L->factor_ms;
    });
}

// The coarse space of the two-level preconditioner: Phi (n_free, k) FP64 host, row-major, rounded
// to FP32 once (the basis every later product uses). Rl Phi, Ll Phi (FP64 loop matrix) and the
// Galerkin matrices Phi^T Rl Phi, Phi^T Ll Phi are formed on the device in FP64; each frequency
// then only inverts the k x k coarse matrix.
int pq_loops_set_coarse(pq_loops *L, int k, const double *Phi, int diagonal) {
    return guard([&] {
        PQ_REQUIRE(L && (k == 0 || Phi), "pq_loops_set_coarse: null argument");
        PQ_REQUIRE(k >= 0 && k <= 4096, "pq_loops_set_coarse: 0 <= k <= 4096");
        DeviceScope scope(L->device);
        L->kc = k;
        L->cdiag = diagonal != 0;
        L->omega_c = -1;
        L->crows.clear();
        L->psplit.clear();
        L->pdepth.clear();
        if (k == 0) return;
        const int n = L->nf;
        L->kp = (k + 63) / 64 * 64;
        const int kp = L->kp;
        L->ldt = (n + 63) / 64 * 64;
        L->Phi.reserve((size_t)n * kp * 4); L->Phit.reserve((size_t)L->ldt * k * 4);
        L->RPhi.reserve((size_t)L->ldt * k * 4); L->LPhi.reserve((size_t)L->ldt * k * 4);
        DevBuf P64, RX, LX, G;
        P64.reserve((size_t)n * k * 8); RX.reserve((size_t)n * k * 8); LX.reserve((size_t)n * k * 8);
        G.reserve((size_t)k * k * 8);
        PQ_HIP(hipMemcpyAsync(P64.p, Phi, (size_t)n * k * 8, hipMemcpyHostToDevice, L->stream));
        // the FP32 basis (P64 rounded to it in place), its transpose, then Rl Phi and Ll Phi (FP64
        // loop matrix) and the Galerkin matrices Phi^T Rl Phi, Phi^T Ll Phi in FP64
        hipLaunchKernelGGL(round_basis_kernel, dim3((unsigned)(((size_t)n * kp + 255) / 256)), dim3(256), 0, L->stream,
                           P64.as<double>(), n, k, kp, L->Phi.as<float>());
        const unsigned nb = (unsigned)(((size_t)L->ldt * k + 255) / 256);
        hipLaunchKernelGGL(transpose_round_kernel, dim3(nb), dim3(256), 0, L->stream, P64.as<double>(), n, k, L->ldt,
                           L->Phit.as<float>());
        PQ_HIP(hipGetLastError());
        basis_products(L, P64.as<double>(), k, RX.as<double>(), LX.as<double>());
        hipLaunchKernelGGL(transpose_round_kernel, dim3(nb), dim3(256), 0, L->stream, RX.as<double>(), n, k, L->ldt,
                           L->RPhi.as<float>());
        hipLaunchKernelGGL(transpose_round_kernel, dim3(nb), dim3(256), 0, L->stream, LX.as<double>(), n, k, L->ldt,
                           L->LPhi.as<float>());
        PQ_HIP(hipGetLastError());
        const unsigned kb = (unsigned)(((size_t)k * k + 255) / 256);
        for (int which = 0; which < 2; ++which) {
            std::vector<double> &Gh = which ? L->Lc : L->Rc;
            Gh.assign((size_t)k * k, 0.0);
            gemm_rm(L, true, false, k, k, n, P64.as<double>(), k, (which ? LX : RX).as<double>(), k, G.as<double>(), k);
            // the Galerkin matrices of symmetric operators: average away the rounding asymmetry
            hipLaunchKernelGGL(symmetrize64_kernel, dim3(kb), dim3(256), 0, L->stream, G.as<double>(), k);
            PQ_HIP(hipGetLastError());
            PQ_HIP(hipMemcpyAsync(Gh.data(), G.p, (size_t)k * k * 8, hipMemcpyDeviceToHost, L->stream));
            PQ_HIP(hipStreamSynchronize(L->stream));
        }
        make_half_bases(L);
    });
}

// The coarse space of Phi (n_free, k) FP64 host re-expressed on the device in the generalised
// eigenbasis of its Galerkin pair (coarse_diagonalise). *k_out receives the kept dimension.
int pq_loops_set_coarse_diagonal(pq_loops *L, int k, const double *Phi, double drop, int *k_out) {
    return guard([&] {
        PQ_REQUIRE(L && Phi && k_out, "pq_loops_set_coarse_diagonal: null argument");
        PQ_REQUIRE(k >= 1 && k <= 4096, "pq_loops_set_coarse_diagonal: 1 <= k <= 4096");
        PQ_REQUIRE(drop >= 0 && drop < 1, "pq_loops_set_coarse_diagonal: 0 <= drop < 1");
        DeviceScope scope(L->device);
        const int n = L->nf;
        PQ_REQUIRE(n > 0, "pq_loops_set_coarse_diagonal: no free loops");
        DevBuf P64;
        P64.reserve((size_t)n * k * 8);
        PQ_HIP(hipMemcpyAsync(P64.p, Phi, (size_t)n * k * 8, hipMemcpyHostToDevice, L->stream));
        int path = 0;
        coarse_diagonalise(L, P64, k, drop, k_out, &path);
    });
}

int pq_loops_set_coarse_csc(pq_loops *L, int k, const int32_t *colptr, const int32_t *rows, const double *vals,
                            double drop, int *k_out, int *path_out) {
    return guard([&] {
        PQ_REQUIRE(L && colptr && k_out && path_out, "pq_loops_set_coarse_csc: null argument");
        PQ_REQUIRE(k >= 1 && k <= 4096, "pq_loops_set_coarse_csc: 1 <= k <= 4096");
        PQ_REQUIRE(drop >= 0 && drop < 1, "pq_loops_set_coarse_csc: 0 <= drop < 1");
        const int n = L->nf;
        PQ_REQUIRE(n > 0, "pq_loops_set_coarse_csc: no free loops");
        PQ_REQUIRE(colptr[0] == 0, "pq_loops_set_coarse_csc: colptr[0] must be 0");
        for (int c = 0; c < k; ++c)
            PQ_REQUIRE(colptr[c + 1] >= colptr[c], "pq_loops_set_coarse_csc: column pointers out of order");
        const int64_t nnz = colptr[k];
        PQ_REQUIRE(nnz == 0 || (rows && vals), "pq_loops_set_coarse_csc: null entries");
        for (int64_t t = 0; t < nnz; ++t)
            PQ_REQUIRE(rows[t] >= 0 && rows[t] < n, "pq_loops_set_coarse_csc: row index out of range");
        DeviceScope scope(L->device);
        DevBuf P64, cp, rw, vl;
        P64.reserve((size_t)n * k * 8);
        cp.reserve((size_t)(k + 1) * 4);
        rw.reserve((size_t)std::max<int64_t>(nnz, 1) * 4);
        vl.reserve((size_t)std::max<int64_t>(nnz, 1) * 8);
        PQ_HIP(hipMemsetAsync(P64.p, 0, (size_t)n * k * 8, L->stream));
        PQ_HIP(hipMemcpyAsync(cp.p, colptr, (size_t)(k + 1) * 4, hipMemcpyHostToDevice, L->stream));
        if (nnz) {
            PQ_HIP(hipMemcpyAsync(rw.p, rows, (size_t)nnz * 4, hipMemcpyHostToDevice, L->stream));
            PQ_HIP(hipMemcpyAsync(vl.p, vals, (size_t)nnz * 8, hipMemcpyHostToDevice, L->stream));
        }
        hipLaunchKernelGGL(csc_scatter_kernel, dim3((k + 255) / 256), dim3(256), 0, L->stream, cp.as<int>(),
                           rw.as<int>(), vl.as<double>(), k, P64.as<double>());
        PQ_HIP(hipGetLastError());
        coarse_diagonalise(L, P64, k, drop, k_out, path_out);
    });
}

int pq_loops_prepare_block_jacobi(pq_loops *L, int *modal_out) {
    return guard([&] {
        PQ_REQUIRE(L, "pq_loops_prepare_block_jacobi: null handle");
        DeviceScope scope(L->device);
        block_jacobi_modes(L);
        if (modal_out) *modal_out = L->bj_modal;
    });
}

int pq_loops_rhs(pq_loops *L, double omega, int P, const int32_t *port_cols, double *B_host) {
    return guard([&] {
        PQ_REQUIRE(L && port_cols && B_host, "pq_loops_rhs: null argument");
        PQ_REQUIRE(P >= 1, "pq_loops_rhs: P >= 1");
        for (int p = 0; p < P; ++p)
            PQ_REQUIRE(port_cols[p] >= L->nf && port_cols[p] < L->nl, "pq_loops_rhs: port loop out of range");
        DeviceScope scope(L->device);
        if (L->nf == 0) return;
        port_rhs(L, omega, P, port_cols);
        PQ_HIP(hipMemcpyAsync(B_host, L->B.p, (size_t)L->nf * 2 * P * 8, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
    });
}

int pq_loops_operator(pq_loops *L, double omega, int P, int precision, const double *X_host, double *W_host) {
    return guard([&] {
        PQ_REQUIRE(L && X_host && W_host, "pq_loops_operator: null argument");
        PQ_REQUIRE(P >= 1, "pq_loops_operator: P >= 1");
        PQ_REQUIRE(precision == 32 || precision == 64, "pq_loops_operator: precision must be 32 or 64");
        dispatch_nr(2 * P, [](auto) {});
        DeviceScope scope(L->device);
        const int n = L->nf;
        if (n == 0) return;
        tune_gemv(L, 2 * P);
        set_omega(L, omega);
        const size_t bytes = (size_t)n * 2 * P * 8;
        DevBuf X, W;
        X.reserve(bytes);
        W.reserve(bytes);
        PQ_HIP(hipMemcpyAsync(X.p, X_host, bytes, hipMemcpyHostToDevice, L->stream));
        if (precision == 64) apply_operator64(L, X.as<double>(), W.as<double>(), P);
        else apply_operator(L, X.as<double>(), W.as<double>(), P);
        PQ_HIP(hipMemcpyAsync(W_host, W.p, bytes, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
    });
}

int pq_loops_coarse_solution(pq_loops *L, double omega, int P, const int32_t *port_cols, double *Xc_host) {
    return guard([&] {
        PQ_REQUIRE(L && port_cols && Xc_host, "pq_loops_coarse_solution: null argument");
        PQ_REQUIRE(P >= 1, "pq_loops_coarse_solution: P >= 1");
        PQ_REQUIRE(L->kc > 0, "pq_loops_coarse_solution: no coarse space set");
        for (int p = 0; p < P; ++p)
            PQ_REQUIRE(port_cols[p] >= L->nf && port_cols[p] < L->nl,
                       "pq_loops_coarse_solution: port loop out of range");
        dispatch_nr(2 * P, [](auto) {});
        DeviceScope scope(L->device);
        const int n = L->nf;
        if (n == 0) return;
        port_rhs(L, omega, P, port_cols);
        prepare_coarse(L, omega, P);
        set_omega(L, omega);
        coarse_correction(L, L->B.as<double>(), P);
        PQ_HIP(hipMemcpyAsync(Xc_host, L->T2.p, (size_t)n * 2 * P * 8, hipMemcpyDeviceToHost, L->stream));
        PQ_HIP(hipStreamSynchronize(L->stream));
    });
}

int pq_loops_coarse_matrices(pq_loops *L, double *Rc, double *Lc) {
    return guard([&] {
        PQ_REQUIRE(L && Rc && Lc, "pq_loops_coarse_matrices: null argument");
        PQ_REQUIRE(L->kc > 0, "pq_loops_coarse_matrices: no coarse space set");
        std::copy(L->Rc.begin(), L->Rc.end(), Rc);
        std::copy(L->Lc.begin(), L->Lc.end(), Lc);
    });
}

int pq_loops_use_graphs(pq_loops *L, int enable) {
    return guard([&] {
        PQ_REQUIRE(L, "pq_loops_use_graphs: null handle");
        L->graphs_on = enable < 0 ? -1 : (enable != 0);
        if (!enable) {
            DeviceScope scope(L->device);
            drop_graphs(L);
        }
    });
}

int pq_loops_tuning(pq_loops *L, int NR, int *chunk, double *microseconds) {
    return guard([&] {
        PQ_REQUIRE(L && chunk, "pq_loops_tuning: null argument");
        DeviceScope scope(L->device);
        tune_gemv(L, NR);
        *chunk = gemv_chunk(L, NR);
        auto it = L->gemv_us.find(NR);
        if (microseconds) *microseconds = it != L->gemv_us.end() ? it->second : 0.0;
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

int pq_loops_last_fallbacks(pq_loops *L, int64_t *count) {
    return guard([&] {
        PQ_REQUIRE(L && count, "pq_loops_last_fallbacks: null argument");
        *count = L->fallbacks;
    });
}

}  // extern "C"
