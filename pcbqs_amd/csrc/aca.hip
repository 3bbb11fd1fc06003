// Construction of the compressed RWG inductance operator on the device: entry blocks, checked ACA
// for the admissible blocks and dense leaves, with every factor kept in HBM and assembled into the
// product handle of hmatrix.hip. No dense matrix of the operator exists at any point.
//
// Semantics:
//   entries      canonical_pair and rwg_entry / rwg_block (src/pcbqs/core/compressed_kernels.hpp:9-39,
//                73-108): each triangle pair integrated as (min, max), the first moments swapped
//                back (p > q) or averaged (p == q); the four support terms combined with rounded
//                products and summed in (sa, sb) order; scaled by mu0 / 4 pi.
//   aca          aca (src/pcbqs/core/hmatrix_build.hpp:143-270): rank cap min(128, 0.7 m n / (m + n)),
//                16 check rows drawn by a partial Fisher-Yates shuffle from a SplitMix64 stream
//                seeded by the block's identity, pivot row, column pivot at the row's first largest
//                residual, a second independent check set before accepting, the next pivot row the
//                unused check row of largest residual (else the unused row of largest |U| in the
//                last column). Residual norms are summed in a fixed tree order instead of the
//                compiled loops' sequential order; entries are the device's FP64 pair integrals.
//   dense leaves dense_blocks (hmatrix_build.hpp:329-352), diagonal leaves symmetrised
//                (hmatrix.py:204-206).
//   The tree, admissibility and the level loop are host logic (pcbqs_amd/hmatrix.py).
//
// Device design: one 256-thread workgroup per admissible block runs the whole ACA (no host round
// trip per rank): the pivot row is written straight into the next V row, the column into the next
// U row (U is kept rank-major, U[l m + i], so both the row's and the column's U V corrections read
// coalesced rows), the check rows' residuals stay in a per-block HBM workspace updated in place,
// and the block-wide argmax and norms are LDS tree reductions. The ACA's sequential dependence is
// per block; a tree level supplies hundreds to thousands of independent blocks to fill the chip.
// A block whose check rows outgrow their workspace (more than nine second check sets) stops with
// a status that makes the host run it again with room for the worst case, so results do not
// depend on the workspace sizing. Converged factors are packed into chunks (U transposed to the
// product's row-major layout); the product handle gathers them in the final block order.
#include "pq_common.h"
#include "pair.h"
#include "hmatrix.h"

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <vector>

namespace pq {
namespace {

constexpr int ACA_RMAX = 128;          // the reference's rank cap
constexpr int ACA_CHECK = 16;          // rows per check set
constexpr int ACA_CHECK_FIRST = 160;   // check-row room of the first attempt: the first set and nine more
// Launch shapes: general blocks (every pair classified, the analytic path compiled in) in 256-thread
// workgroups; blocks shown far throughout in 256 threads, or 1024 from ACA_WIDE rows plus columns on
// (the few large blocks of the top tree levels, which would otherwise run one wave per SIMD).
constexpr int ACA_WIDE = 1536;
enum AcaShape { ACA_GENERAL = 0, ACA_FAR = 1, ACA_FAR_WIDE = 2 };


struct EntryArgs {
    GeomDev g;
    const int *sup;        // (ne, 2): plus and minus triangle
    const double *vx;      // (ne, 2, 2): their free vertices
    const double *al;      // (ne, 2): +-1 / (2 area)
    double scale, near;
    int outer, all_near;
};

// Moments of the ordered pair (p, q) from its canonical pair (compressed_kernels.hpp:9-21).
// FAR: the caller has shown that every pair of the block is far under the near rule (the block's
// triangle centroids are farther apart than near times the largest radii), so the far rule is
// applied without classifying, and the analytic path, which needs several times the registers, is
// not compiled in.
template <bool FAR>
__device__ __forceinline__ void canonical64(const EntryArgs &E, int p, int q, double *m) {
    const GeomDev &g = E.g;
    const int a = min(p, q), b = max(p, q);
    const double d = g.z[a] - g.z[b];
    if (FAR) {
        far64<false>(g.P3 + 6 * (size_t)a, g.P3 + 6 * (size_t)b, g.area[a] * g.area[b] / 9, d, m);
    } else {
        const int k = exact_kind(g.cent[2 * a], g.cent[2 * a + 1], g.z[a], g.radius[a], g.cent[2 * b],
                                 g.cent[2 * b + 1], g.z[b], g.radius[b], E.near, E.all_near, E.outer);
        pair64<false>(g, a, b, d, k, m);
    }
    if (p > q) {
        double t = m[1]; m[1] = m[3]; m[3] = t;
        t = m[2]; m[2] = m[4]; m[4] = t;
    } else if (p == q) {
        const double x = (m[1] + m[3]) / 2, y = (m[2] + m[4]) / 2;
        m[1] = m[3] = x;
        m[2] = m[4] = y;
    }
}

// Entry (ea, eb) of the RWG inductance (compressed_kernels.hpp:23-39), products rounded.
template <bool FAR>
__device__ double rwg_entry64(const EntryArgs &E, int ea, int eb) {
#pragma clang fp contract(off)
    double total = 0.0;
#pragma nounroll
    for (int sa = 0; sa < 2; ++sa) {
        const int p = E.sup[2 * ea + sa];
        const double vpx = E.vx[4 * ea + 2 * sa], vpy = E.vx[4 * ea + 2 * sa + 1], wa = E.al[2 * ea + sa];
#pragma nounroll
        for (int sb = 0; sb < 2; ++sb) {
            const int q = E.sup[2 * eb + sb];
            const double vqx = E.vx[4 * eb + 2 * sb], vqy = E.vx[4 * eb + 2 * sb + 1], wb = E.al[2 * eb + sb];
            double m[6];
            canonical64<FAR>(E, p, q, m);
            const double val = m[5] - (vqx * m[1] + vqy * m[2]) - (vpx * m[3] + vpy * m[4]) +
                               (vpx * vqx + vpy * vqy) * m[0];
            total += (wa * wb) * val;
        }
    }
    return E.scale * total;
}

__global__ __launch_bounds__(256) void entries_kernel(EntryArgs E, const int *rows, int nr, const int *cols, int nc,
                                                      double *out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)nr * nc) return;
    const int i = (int)(t / nc), j = (int)(t % nc);
    out[t] = rwg_entry64<false>(E, rows[i], cols[j]);
}

// ---- the ACA -----------------------------------------------------------------------------------

struct AcaBlock {
    long long a_off, b_off;    // into the batch's row and column index arrays
    long long u_off, v_off, r_off, i_off;   // workspace: U cap x m, V cap x n, R chk x n; chk + m ints
    int m, n, cap, chk;
};

struct AcaOut {
    int status;                // 1 converged, 0 not, -1 check rows beyond the workspace
    int rank;
    long long entries;
};

// SplitMix64 (Steele, Lea & Flood, OOPSLA 2014).
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long &s) {
    unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// k <= 16 distinct indices of [0, m): a partial Fisher-Yates shuffle of 0..m-1 whose touched
// positions (at most 2k) are kept in a small map in LDS (key, val: 2k entries) instead of an m-long
// array. One thread draws.
__device__ void draw(unsigned long long &state, int m, int k, int *out, int *key, int *val) {
    int used = 0;
    auto get = [&](int p) {
        for (int t = 0; t < used; ++t)
            if (key[t] == p) return val[t];
        return p;
    };
    auto put = [&](int p, int v) {
        for (int t = 0; t < used; ++t)
            if (key[t] == p) { val[t] = v; return; }
        key[used] = p; val[used] = v; ++used;
    };
    for (int i = 0; i < k; ++i) {
        const unsigned long long r = splitmix64(state);
        const int j = i + (int)(r % (unsigned long long)(m - i));
        const int vi = get(i), vj = get(j);
        put(i, vj);
        put(j, vi);
        out[i] = vj;
    }
}

template <int T>
__device__ double block_sum(double v, double *red) {
    const int tid = threadIdx.x;
    red[tid] = v;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// The largest v over the workgroup, the smallest index among equal values (the compiled loops'
// first maximum); threads without a candidate pass (-1, INT_MAX).
template <int T>
__device__ void block_argmax(double &v, int &idx, double *red, int *redi) {
    const int tid = threadIdx.x;
    red[tid] = v;
    redi[tid] = idx;
    __syncthreads();
    for (int s = T / 2; s > 0; s >>= 1) {
        if (tid < s) {
            const double v2 = red[tid + s];
            const int i2 = redi[tid + s];
            if (v2 > red[tid] || (v2 == red[tid] && i2 < redi[tid])) {
                red[tid] = v2;
                redi[tid] = i2;
            }
        }
        __syncthreads();
    }
    v = red[0];
    idx = redi[0];
    __syncthreads();
}

__device__ __forceinline__ double wave_sum(double s) {
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    return s;
}

template <bool FAR, int T>
__global__ __launch_bounds__(T) void aca_kernel(EntryArgs E, const int *rows, const int *cols, const AcaBlock *blocks,
                                               double *ws, int *iws, AcaOut *out, double tol) {
#pragma clang fp contract(off)
    __shared__ double red[T];
    __shared__ int redi[T];
    __shared__ double s_u[ACA_RMAX];
    __shared__ int s_int[3];                     // check count, draw status, next pivot row
    __shared__ int s_key[2 * ACA_CHECK], s_val[2 * ACA_CHECK];
    const AcaBlock B = blocks[blockIdx.x];
    const int *a = rows + B.a_off, *b = cols + B.b_off;
    const int m = B.m, n = B.n, cap = B.cap;
    double *U = ws + B.u_off, *V = ws + B.v_off, *R = ws + B.r_off;
    int *chk = iws + B.i_off, *used = chk + B.chk;
    const int tid = threadIdx.x, wave = tid / 64, lane = tid % 64;
    for (int k = tid; k < m; k += T) used[k] = 0;
    unsigned long long state = 0;                // thread 0's stream
    if (tid == 0) {
        state = 0x5DEECE66DULL ^ ((unsigned long long)a[0] * 0x100000001B3ULL) ^ ((unsigned long long)b[0] << 21) ^
                ((unsigned long long)m << 42) ^ (unsigned long long)n;
        draw(state, m, min(m, ACA_CHECK), chk, s_key, s_val);
        s_int[0] = min(m, ACA_CHECK);
    }
    __syncthreads();
    int nchk = s_int[0];
    long long entries = (long long)nchk * n;
    double part = 0;
    for (long long t = tid; t < (long long)nchk * n; t += T) {
        const double x = rwg_entry64<FAR>(E, a[chk[t / n]], b[t % n]);
        R[t] = x;
        part += x * x;
    }
    double ref2 = block_sum<T>(part, red);
    int rank = 0, i = 0, status = 0;
    for (int it = 0; it < cap; ++it) {
        // the pivot row's residual, into V's next row
        if (tid < rank) s_u[tid] = U[(long long)tid * m + i];
        __syncthreads();
        double *vr = V + (long long)rank * n;
        const int ea = a[i];
        double best = -1.0;
        int jb = INT_MAX;
        for (int j = tid; j < n; j += T) {
            double s = 0.0;
            for (int l = 0; l < rank; ++l) s += s_u[l] * V[(long long)l * n + j];
            const double v = rwg_entry64<FAR>(E, ea, b[j]) - s;
            vr[j] = v;
            if (fabs(v) > best) { best = fabs(v); jb = j; }
        }
        entries += n;
        block_argmax<T>(best, jb, red, redi);
        const double pivot = vr[jb];
        if (fabs(pivot) > DBL_MIN) {
            if (tid < rank) s_u[tid] = V[(long long)tid * n + jb];
            __syncthreads();
            double *uc = U + (long long)rank * m;
            const int eb = b[jb];
            for (int k = tid; k < m; k += T) {
                double s = 0.0;
                for (int l = 0; l < rank; ++l) s += U[(long long)l * m + k] * s_u[l];
                uc[k] = rwg_entry64<FAR>(E, a[k], eb) - s;
            }
            entries += m;
            for (int j = tid; j < n; j += T) vr[j] = vr[j] / pivot;
            __syncthreads();
            for (long long t = tid; t < (long long)nchk * n; t += T)
                R[t] -= uc[chk[t / n]] * vr[t % n];
            ++rank;
        }
        if (tid == 0) used[i] = 1;
        part = 0;
        for (long long t = tid; t < (long long)nchk * n; t += T) part += R[t] * R[t];
        const double res2 = block_sum<T>(part, red);
        if (sqrt(res2) <= tol * fmax(sqrt(ref2), DBL_MIN)) {
            // a second independent check set before accepting
            const int k = min(m, ACA_CHECK);
            if (tid == 0) {
                s_int[1] = nchk + k > B.chk ? -1 : 0;
                if (s_int[1] == 0) draw(state, m, k, chk + nchk, s_key, s_val);
            }
            __syncthreads();
            if (s_int[1] < 0) { status = -1; break; }
            double pe = 0, pf = 0;
            for (long long t = tid; t < (long long)k * n; t += T) {
                const int row = chk[nchk + t / n], j = (int)(t % n);
                const double x = rwg_entry64<FAR>(E, a[row], b[j]);
                double s = 0.0;
                for (int l = 0; l < rank; ++l) s += U[(long long)l * m + row] * V[(long long)l * n + j];
                const double e = x - s;
                R[(long long)nchk * n + t] = e;
                pe += e * e;
                pf += x * x;
            }
            entries += (long long)k * n;
            const double e2 = block_sum<T>(pe, red), f2 = block_sum<T>(pf, red);
            if (sqrt(e2) <= tol * fmax(sqrt(f2), DBL_MIN)) { status = 1; break; }
            nchk += k;
            ref2 += f2;
        }
        // next pivot row: the unused check row of largest residual (first in check order)
        double wbest = -1.0;
        int wc = INT_MAX;
        for (int c = wave; c < nchk; c += T / 64) {
            if (used[chk[c]]) continue;
            double s = 0.0;
            for (int j = lane; j < n; j += 64) s += R[(long long)c * n + j] * R[(long long)c * n + j];
            s = wave_sum(s);
            if (s > wbest) { wbest = s; wc = c; }
        }
        if (lane == 0) { red[wave] = wbest; redi[wave] = wc; }
        __syncthreads();
        if (tid == 0) {
            double bb = -1.0;
            int bc = INT_MAX;
            for (int w = 0; w < T / 64; ++w)
                if (redi[w] != INT_MAX && (bc == INT_MAX || red[w] > bb || (red[w] == bb && redi[w] < bc))) {
                    bb = red[w];
                    bc = redi[w];
                }
            s_int[2] = bc == INT_MAX ? -1 : chk[bc];
        }
        __syncthreads();
        int next = s_int[2];
        __syncthreads();
        if (next < 0) {
            // else the unused row of largest |U| in the last column (the first unused row at rank 0)
            double vb = -1.0;
            int kb = INT_MAX;
            for (int k = tid; k < m; k += T) {
                if (used[k]) continue;
                const double v = rank ? fabs(U[(long long)(rank - 1) * m + k]) : 0.0;
                if (v > vb) { vb = v; kb = k; }
            }
            block_argmax<T>(vb, kb, red, redi);
            next = kb == INT_MAX ? -1 : kb;
        }
        if (next < 0) break;
        i = next;
    }
    if (tid == 0) out[blockIdx.x] = AcaOut{status, rank, entries};
}

// Converged factors into a chunk: U^T (m x r, row-major), then V (r x n).
__global__ __launch_bounds__(256) void aca_pack_kernel(const AcaBlock *blocks, const int *which, const int *ranks,
                                                       const long long *dst_off, const double *ws, double *dst) {
    const AcaBlock B = blocks[which[blockIdx.x]];
    const int r = ranks[blockIdx.x];
    const double *U = ws + B.u_off, *V = ws + B.v_off;
    double *du = dst + dst_off[blockIdx.x], *dv = du + (long long)B.m * r;
    for (long long t = threadIdx.x; t < (long long)B.m * r; t += blockDim.x)
        du[t] = U[(t % r) * B.m + t / r];
    for (long long t = threadIdx.x; t < (long long)r * B.n; t += blockDim.x) dv[t] = V[t];
}

// ---- dense leaves ------------------------------------------------------------------------------

struct DenseBlock {
    long long a_off, b_off, dst;
    int m, n, same;
};

__global__ __launch_bounds__(256) void dense_kernel(EntryArgs E, const int *rows, const int *cols,
                                                    const DenseBlock *blocks, double *dst) {
    const DenseBlock B = blocks[blockIdx.x];
    const int *a = rows + B.a_off, *b = cols + B.b_off;
    double *o = dst + B.dst;
    const long long total = (long long)B.m * B.n;
    for (long long t = threadIdx.x; t < total; t += blockDim.x) o[t] = rwg_entry64<false>(E, a[t / B.n], b[t % B.n]);
    if (!B.same) return;
    __syncthreads();
    // (block + block^T) / 2, each unordered pair by one thread
    for (long long t = threadIdx.x; t < total; t += blockDim.x) {
        const int i = (int)(t / B.n), j = (int)(t % B.n);
        if (i >= j) continue;
        const double x = (o[t] + o[(long long)j * B.n + i]) / 2;
        o[t] = x;
        o[(long long)j * B.n + i] = x;
    }
}

// ---- gathering kept blocks into the product's factor buffer ----------------------------------

struct Piece {
    const double *src;
    long long dst, count;
};

__global__ __launch_bounds__(256) void gather_kernel(const Piece *pieces, double *dst) {
    const Piece P = pieces[blockIdx.x];
    for (long long t = threadIdx.x; t < P.count; t += blockDim.x) dst[P.dst + t] = P.src[t];
}

}  // namespace
}  // namespace pq

using namespace pq;

struct pq_entries {
    struct Kept {
        std::vector<int> a, b;
        int rank = -1;                           // -1: dense
        bool same = false;
        size_t chunk = 0;
        long long off = 0;
    };
    int device = 0;
    long long ne = 0;
    pq_fill *f = nullptr;
    EntryArgs args{};
    std::vector<int> sup;                        // host copies for the far test: supports,
    std::vector<double> cent, z, radius;         // centroids, heights and radii
    DevBuf dsup, vx, al;
    hipStream_t stream = nullptr;
    std::vector<Kept> kept;
    std::vector<DevBuf> chunks;
    long long entries = 0, reruns = 0;
    double aca_ms = 0, dense_ms = 0;
};

namespace {

template <class T> void upload(DevBuf &d, const T *src, size_t count, hipStream_t s) {
    d.reserve(std::max<size_t>(count * sizeof(T), 16));
    if (count) PQ_HIP(hipMemcpyAsync(d.p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
}

void check_lists(const pq_entries *e, int64_t nblocks, const int64_t *row_ptr, const int32_t *rows,
                 const int64_t *col_ptr, const int32_t *cols, const char *who) {
    PQ_REQUIRE(nblocks >= 0 && (nblocks == 0 || (row_ptr && rows && col_ptr && cols)), who);
    for (int64_t q = 0; q < nblocks; ++q) {
        PQ_REQUIRE(row_ptr[q + 1] > row_ptr[q] && col_ptr[q + 1] > col_ptr[q] &&
                       row_ptr[q + 1] - row_ptr[q] < (1LL << 31) && col_ptr[q + 1] - col_ptr[q] < (1LL << 31),
                   "hmatrix blocks: nonempty index vectors required");
    }
    if (nblocks) PQ_REQUIRE(row_ptr[0] == 0 && col_ptr[0] == 0, "hmatrix blocks: offsets must start at 0");
    for (int64_t t = 0; t < (nblocks ? row_ptr[nblocks] : 0); ++t)
        PQ_REQUIRE(rows[t] >= 0 && rows[t] < e->ne, "hmatrix blocks: edge index out of range");
    for (int64_t t = 0; t < (nblocks ? col_ptr[nblocks] : 0); ++t)
        PQ_REQUIRE(cols[t] >= 0 && cols[t] < e->ne, "hmatrix blocks: edge index out of range");
}

// The launch shape of block q: far throughout when the boxes of its row and column triangles'
// centroids (x, y, z) are farther apart than near times the sum of the largest radii on each side,
// with a 1e-12 margin over the rounding of the pair rule (dist <= near (r_p + r_q): near).
int aca_shape(const pq_entries *e, const int32_t *r, long long m, const int32_t *c, long long n) {
    if (e->args.all_near) return ACA_GENERAL;
    auto box = [&](const int32_t *ids, long long k, double lo[3], double hi[3], double &rmax) {
        for (int d = 0; d < 3; ++d) { lo[d] = HUGE_VAL; hi[d] = -HUGE_VAL; }
        rmax = 0;
        for (long long i = 0; i < k; ++i)
            for (int s = 0; s < 2; ++s) {
                const int t = e->sup[2 * (size_t)ids[i] + s];
                const double p[3] = {e->cent[2 * (size_t)t], e->cent[2 * (size_t)t + 1], e->z[(size_t)t]};
                for (int d = 0; d < 3; ++d) { lo[d] = std::min(lo[d], p[d]); hi[d] = std::max(hi[d], p[d]); }
                rmax = std::max(rmax, e->radius[(size_t)t]);
            }
    };
    double la[3], ha[3], lb[3], hb[3], ra, rb;
    box(r, m, la, ha, ra);
    box(c, n, lb, hb, rb);
    double g2 = 0;
    for (int d = 0; d < 3; ++d) {
        const double gap = std::max(0.0, std::max(la[d] - hb[d], lb[d] - ha[d]));
        g2 += gap * gap;
    }
    if (!(std::sqrt(g2) > e->args.near * (ra + rb) * (1 + 1e-12))) return ACA_GENERAL;
    return m + n >= ACA_WIDE ? ACA_FAR_WIDE : ACA_FAR;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    PQ_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

constexpr size_t ACA_WORKSPACE_BYTES = size_t(4) << 30;   // per launch; larger levels go in several

// One launch over blocks `sel` of the lists, with `chk` check-row room per block (0: the first
// attempt's). Returns each block's result.
std::vector<AcaOut> aca_launch(pq_entries *e, const std::vector<int> &sel, const int64_t *row_ptr,
                               const int64_t *col_ptr, const DevBuf &drows, const DevBuf &dcols,
                               const std::vector<int> &caps, bool worst, double tol, int shape,
                               std::vector<AcaBlock> &desc, DevBuf &ws, DevBuf &iws) {
    desc.assign(sel.size(), AcaBlock{});
    long long dw = 0, iw = 0;
    for (size_t t = 0; t < sel.size(); ++t) {
        const int q = sel[t];
        AcaBlock &B = desc[t];
        B.m = (int)(row_ptr[q + 1] - row_ptr[q]);
        B.n = (int)(col_ptr[q + 1] - col_ptr[q]);
        B.cap = caps[q];
        B.chk = worst ? ACA_CHECK * (B.cap + 1) : std::min(ACA_CHECK_FIRST, ACA_CHECK * (B.cap + 1));
        B.a_off = row_ptr[q];
        B.b_off = col_ptr[q];
        B.u_off = dw; dw += (long long)B.cap * B.m;
        B.v_off = dw; dw += (long long)B.cap * B.n;
        B.r_off = dw; dw += (long long)B.chk * B.n;
        B.i_off = iw; iw += B.chk + B.m;
    }
    ws.reserve(std::max<size_t>((size_t)dw * 8, 16));
    iws.reserve(std::max<size_t>((size_t)iw * 4, 16));
    DevBuf ddesc, dout;
    upload(ddesc, desc.data(), desc.size(), e->stream);
    dout.reserve(sel.size() * sizeof(AcaOut));
#define PQ_ACA(FAR, T)                                                                                   \
    hipLaunchKernelGGL((aca_kernel<FAR, T>), dim3((unsigned)sel.size()), dim3(T), 0, e->stream, e->args,  \
                       drows.as<int>(), dcols.as<int>(), ddesc.as<AcaBlock>(), ws.as<double>(), iws.as<int>(), \
                       dout.as<AcaOut>(), tol)
    if (shape == ACA_FAR_WIDE) PQ_ACA(true, 1024);
    else if (shape == ACA_FAR) PQ_ACA(true, 256);
    else PQ_ACA(false, 256);
#undef PQ_ACA
    PQ_HIP(hipGetLastError());
    std::vector<AcaOut> res(sel.size());
    PQ_HIP(hipMemcpyAsync(res.data(), dout.p, res.size() * sizeof(AcaOut), hipMemcpyDeviceToHost, e->stream));
    PQ_HIP(hipStreamSynchronize(e->stream));
    return res;
}

// Packs the converged blocks of one launch into a new chunk and records them.
void aca_keep(pq_entries *e, const std::vector<int> &sel, const std::vector<AcaOut> &res,
              const std::vector<AcaBlock> &desc, const DevBuf &ws, const int64_t *row_ptr, const int32_t *rows,
              const int64_t *col_ptr, const int32_t *cols, int64_t *ids, int64_t *ranks) {
    std::vector<int> which, rk;
    std::vector<long long> off;
    long long total = 0;
    for (size_t t = 0; t < sel.size(); ++t) {
        if (res[t].status != 1) continue;
        which.push_back((int)t);
        rk.push_back(res[t].rank);
        off.push_back(total);
        total += (long long)res[t].rank * (desc[t].m + desc[t].n);
    }
    if (which.empty()) return;
    e->chunks.emplace_back();
    DevBuf &chunk = e->chunks.back();
    chunk.reserve(std::max<size_t>((size_t)total * 8, 16));
    DevBuf ddesc, dwhich, drk, doff;
    upload(ddesc, desc.data(), desc.size(), e->stream);
    upload(dwhich, which.data(), which.size(), e->stream);
    upload(drk, rk.data(), rk.size(), e->stream);
    upload(doff, off.data(), off.size(), e->stream);
    hipLaunchKernelGGL(aca_pack_kernel, dim3((unsigned)which.size()), dim3(256), 0, e->stream,
                       ddesc.as<AcaBlock>(), dwhich.as<int>(), drk.as<int>(), doff.as<long long>(), ws.as<double>(),
                       chunk.as<double>());
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipStreamSynchronize(e->stream));
    for (size_t w = 0; w < which.size(); ++w) {
        const int q = sel[(size_t)which[w]];
        pq_entries::Kept k;
        k.a.assign(rows + row_ptr[q], rows + row_ptr[q + 1]);
        k.b.assign(cols + col_ptr[q], cols + col_ptr[q + 1]);
        k.rank = rk[w];
        k.same = false;
        k.chunk = e->chunks.size() - 1;
        k.off = off[w];
        ids[q] = (int64_t)e->kept.size();
        ranks[q] = rk[w];
        e->kept.push_back(std::move(k));
    }
}

}  // namespace

extern "C" {

int pq_rwg_entries_create(pq_fill *f, int64_t ne, const int32_t *support, const double *vertices,
                          const double *alpha, double scale, double near, int outer_levels, pq_entries **out) {
    return guard([&] {
        PQ_REQUIRE(f && out && (ne == 0 || (support && vertices && alpha)), "pq_rwg_entries_create: null argument");
        PQ_REQUIRE(ne >= 0 && ne < (1LL << 31), "pq_rwg_entries_create: invalid edge count");
        PQ_REQUIRE(!std::isnan(near) && near >= 0, "near must be nonnegative (or infinity)");
        PQ_REQUIRE(outer_levels >= 0 && outer_levels <= 2 && f->g.nq[outer_levels] == 6 * (1 << (2 * outer_levels)),
                   "pq_rwg_entries_create: outer level was not prepared");
        PQ_REQUIRE(std::isfinite(scale), "pq_rwg_entries_create: scale must be finite");
        for (int64_t t = 0; t < 2 * ne; ++t)
            PQ_REQUIRE(support[t] >= 0 && support[t] < f->nt, "pq_rwg_entries_create: triangle index out of range");
        DeviceScope scope(f->device);
        auto *e = new pq_entries();
        try {
            e->device = f->device;
            e->ne = ne;
            e->f = f;
            PQ_HIP(hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
            upload(e->dsup, support, (size_t)(2 * ne), e->stream);
            e->sup.assign(support, support + 2 * ne);
            const size_t nt = (size_t)f->nt;
            e->cent.resize(2 * nt);
            e->z.resize(nt);
            e->radius.resize(nt);
            PQ_HIP(hipMemcpy(e->cent.data(), f->g.cent, 2 * nt * 8, hipMemcpyDeviceToHost));
            PQ_HIP(hipMemcpy(e->z.data(), f->g.z, nt * 8, hipMemcpyDeviceToHost));
            PQ_HIP(hipMemcpy(e->radius.data(), f->g.radius, nt * 8, hipMemcpyDeviceToHost));
            upload(e->vx, vertices, (size_t)(4 * ne), e->stream);
            upload(e->al, alpha, (size_t)(2 * ne), e->stream);
            PQ_HIP(hipStreamSynchronize(e->stream));
            EntryArgs &A = e->args;
            A.g = f->g;
            A.sup = e->dsup.as<int>();
            A.vx = e->vx.as<double>();
            A.al = e->al.as<double>();
            A.scale = scale;
            A.all_near = std::isinf(near) ? 1 : 0;
            A.near = A.all_near ? 0.0 : near;
            A.outer = outer_levels;
        } catch (...) {
            if (e->stream) (void)hipStreamDestroy(e->stream);
            delete e;
            throw;
        }
        *out = e;
    });
}

int pq_entries_destroy(pq_entries *e) {
    if (!e) return PQ_OK;
    return guard([&] {
        DeviceScope scope(e->device);
        if (e->stream) {
            (void)hipStreamSynchronize(e->stream);
            (void)hipStreamDestroy(e->stream);
        }
        delete e;
    });
}

int pq_entries_block(pq_entries *e, const int32_t *rows, int64_t nr, const int32_t *cols, int64_t nc, double *out) {
    return guard([&] {
        PQ_REQUIRE(e && nr >= 0 && nc >= 0 && (nr * nc == 0 || (rows && cols && out)), "pq_entries_block: invalid argument");
        PQ_REQUIRE(nr < (1LL << 31) && nc < (1LL << 31), "pq_entries_block: block too large");
        for (int64_t t = 0; t < nr; ++t) PQ_REQUIRE(rows[t] >= 0 && rows[t] < e->ne, "pq_entries_block: edge index out of range");
        for (int64_t t = 0; t < nc; ++t) PQ_REQUIRE(cols[t] >= 0 && cols[t] < e->ne, "pq_entries_block: edge index out of range");
        if (!nr || !nc) return;
        DeviceScope scope(e->device);
        DevBuf dr, dc, dout;
        upload(dr, rows, (size_t)nr, e->stream);
        upload(dc, cols, (size_t)nc, e->stream);
        dout.reserve((size_t)(nr * nc) * 8);
        const long long total = nr * nc;
        hipLaunchKernelGGL(entries_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, e->stream, e->args,
                           dr.as<int>(), (int)nr, dc.as<int>(), (int)nc, dout.as<double>());
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipMemcpyAsync(out, dout.p, (size_t)total * 8, hipMemcpyDeviceToHost, e->stream));
        PQ_HIP(hipStreamSynchronize(e->stream));
    });
}

int pq_aca_blocks(pq_entries *e, int64_t nblocks, const int64_t *row_ptr, const int32_t *rows, const int64_t *col_ptr,
                  const int32_t *cols, double tol, int64_t *ids, int64_t *ranks, int64_t *entries) {
    return guard([&] {
        PQ_REQUIRE(e && (nblocks == 0 || (ids && ranks)), "pq_aca_blocks: null argument");
        PQ_REQUIRE(tol > 0 && tol < 1, "aca_blocks: tolerance in (0, 1) required");
        check_lists(e, nblocks, row_ptr, rows, col_ptr, cols, "pq_aca_blocks: null argument");
        long long count = 0;
        if (entries) *entries = 0;
        if (!nblocks) return;
        DeviceScope scope(e->device);
        std::vector<int> caps((size_t)nblocks);
        std::vector<int> todo[3];
        for (int64_t q = 0; q < nblocks; ++q) {
            ids[q] = -1;
            ranks[q] = -1;
            const long long mi = row_ptr[q + 1] - row_ptr[q], ni = col_ptr[q + 1] - col_ptr[q];
            const double m = (double)mi, n = (double)ni;
            const long long cap = std::min(128LL, (long long)((0.7 * m) * n / (m + n)));
            caps[(size_t)q] = (int)std::max(0LL, cap);
            if (cap >= 1) todo[aca_shape(e, rows + row_ptr[q], mi, cols + col_ptr[q], ni)].push_back((int)q);
        }
        DevBuf drows, dcols, ws, iws;
        upload(drows, rows, (size_t)row_ptr[nblocks], e->stream);
        upload(dcols, cols, (size_t)col_ptr[nblocks], e->stream);
        hipEvent_t e0, e1;
        PQ_HIP(hipEventCreate(&e0));
        PQ_HIP(hipEventCreate(&e1));
        PQ_HIP(hipEventRecord(e0, e->stream));
        for (int shape = ACA_FAR_WIDE; shape >= ACA_GENERAL; --shape)
        for (int pass = 0; pass < 2 && !todo[shape].empty(); ++pass) {
            const bool worst = pass == 1;
            std::vector<int> &work = todo[shape];
            std::vector<int> again;
            size_t start = 0;
            while (start < work.size()) {
                // as many blocks as fit the workspace bound (always at least one)
                size_t stop = start;
                size_t bytes = 0;
                while (stop < work.size()) {
                    const int q = work[stop];
                    const long long m = row_ptr[q + 1] - row_ptr[q], n = col_ptr[q + 1] - col_ptr[q];
                    const long long c = caps[(size_t)q];
                    const long long chk = worst ? ACA_CHECK * (c + 1) : std::min<long long>(ACA_CHECK_FIRST, ACA_CHECK * (c + 1));
                    const size_t add = (size_t)(c * (m + n) + chk * n) * 8 + (size_t)(chk + m) * 4;
                    if (stop > start && bytes + add > ACA_WORKSPACE_BYTES) break;
                    bytes += add;
                    ++stop;
                }
                std::vector<int> sel(work.begin() + (long)start, work.begin() + (long)stop);
                std::vector<AcaBlock> desc;
                std::vector<AcaOut> res = aca_launch(e, sel, row_ptr, col_ptr, drows, dcols, caps, worst, tol, shape,
                                                     desc, ws, iws);
                for (size_t t = 0; t < sel.size(); ++t) {
                    count += res[t].entries;
                    if (res[t].status < 0) again.push_back(sel[t]);
                }
                aca_keep(e, sel, res, desc, ws, row_ptr, rows, col_ptr, cols, ids, ranks);
                start = stop;
            }
            e->reruns += (long long)again.size();
            work.swap(again);
        }
        for (const auto &left : todo)
            PQ_REQUIRE(left.empty(), "pq_aca_blocks: check rows beyond the worst-case workspace");
        PQ_HIP(hipEventRecord(e1, e->stream));
        PQ_HIP(hipEventSynchronize(e1));
        e->aca_ms += elapsed(e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        e->entries += count;
        if (entries) *entries = count;
    });
}

int pq_dense_blocks(pq_entries *e, int64_t nblocks, const int64_t *row_ptr, const int32_t *rows, const int64_t *col_ptr,
                    const int32_t *cols, const int32_t *same, int64_t *ids) {
    return guard([&] {
        PQ_REQUIRE(e && (nblocks == 0 || (same && ids)), "pq_dense_blocks: null argument");
        check_lists(e, nblocks, row_ptr, rows, col_ptr, cols, "pq_dense_blocks: null argument");
        if (!nblocks) return;
        std::vector<DenseBlock> desc((size_t)nblocks);
        long long total = 0;
        for (int64_t q = 0; q < nblocks; ++q) {
            DenseBlock &B = desc[(size_t)q];
            B.m = (int)(row_ptr[q + 1] - row_ptr[q]);
            B.n = (int)(col_ptr[q + 1] - col_ptr[q]);
            B.same = same[q] != 0;
            PQ_REQUIRE(!B.same || B.m == B.n, "pq_dense_blocks: a diagonal block must be square");
            B.a_off = row_ptr[q];
            B.b_off = col_ptr[q];
            B.dst = total;
            total += (long long)B.m * B.n;
        }
        DeviceScope scope(e->device);
        DevBuf drows, dcols, ddesc;
        upload(drows, rows, (size_t)row_ptr[nblocks], e->stream);
        upload(dcols, cols, (size_t)col_ptr[nblocks], e->stream);
        upload(ddesc, desc.data(), desc.size(), e->stream);
        e->chunks.emplace_back();
        DevBuf &chunk = e->chunks.back();
        chunk.reserve(std::max<size_t>((size_t)total * 8, 16));
        hipEvent_t e0, e1;
        PQ_HIP(hipEventCreate(&e0));
        PQ_HIP(hipEventCreate(&e1));
        PQ_HIP(hipEventRecord(e0, e->stream));
        hipLaunchKernelGGL(dense_kernel, dim3((unsigned)nblocks), dim3(256), 0, e->stream, e->args, drows.as<int>(),
                           dcols.as<int>(), ddesc.as<DenseBlock>(), chunk.as<double>());
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipEventRecord(e1, e->stream));
        PQ_HIP(hipEventSynchronize(e1));
        e->dense_ms += elapsed(e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (int64_t q = 0; q < nblocks; ++q) {
            pq_entries::Kept k;
            k.a.assign(rows + row_ptr[q], rows + row_ptr[q + 1]);
            k.b.assign(cols + col_ptr[q], cols + col_ptr[q + 1]);
            k.rank = -1;
            k.same = desc[(size_t)q].same != 0;
            k.chunk = e->chunks.size() - 1;
            k.off = desc[(size_t)q].dst;
            ids[q] = (int64_t)e->kept.size();
            e->kept.push_back(std::move(k));
        }
        e->entries += total;
    });
}

int pq_entries_factors(pq_entries *e, int64_t id, double *u, double *v) {
    return guard([&] {
        PQ_REQUIRE(e && id >= 0 && id < (int64_t)e->kept.size() && u, "pq_entries_factors: invalid argument");
        const pq_entries::Kept &k = e->kept[(size_t)id];
        const long long m = (long long)k.a.size(), n = (long long)k.b.size();
        DeviceScope scope(e->device);
        const double *src = e->chunks[k.chunk].as<double>() + k.off;
        if (k.rank < 0) {
            PQ_HIP(hipMemcpy(u, src, (size_t)(m * n) * 8, hipMemcpyDeviceToHost));
        } else {
            PQ_REQUIRE(v, "pq_entries_factors: null V for a low-rank block");
            PQ_HIP(hipMemcpy(u, src, (size_t)(m * k.rank) * 8, hipMemcpyDeviceToHost));
            PQ_HIP(hipMemcpy(v, src + m * k.rank, (size_t)(k.rank * n) * 8, hipMemcpyDeviceToHost));
        }
    });
}

int pq_entries_hmatrix(pq_entries *e, int64_t nblocks, const int64_t *ids, pq_hmatrix **out) {
    return guard([&] {
        PQ_REQUIRE(e && out && (nblocks == 0 || ids), "pq_entries_hmatrix: null argument");
        std::vector<int64_t> meta((size_t)nblocks * 6);
        std::vector<int32_t> idx;
        std::vector<Piece> pieces;
        long long fo = 0;
        for (int64_t q = 0; q < nblocks; ++q) {
            PQ_REQUIRE(ids[q] >= 0 && ids[q] < (int64_t)e->kept.size(), "pq_entries_hmatrix: unknown block id");
            const pq_entries::Kept &k = e->kept[(size_t)ids[q]];
            const long long m = (long long)k.a.size(), n = (long long)k.b.size();
            int64_t *row = meta.data() + 6 * q;
            row[0] = (int64_t)idx.size();
            row[1] = m;
            idx.insert(idx.end(), k.a.begin(), k.a.end());
            row[2] = (int64_t)idx.size();
            row[3] = n;
            idx.insert(idx.end(), k.b.begin(), k.b.end());
            row[4] = k.rank;
            row[5] = k.same ? 1 : 0;
            const long long count = k.rank < 0 ? m * n : (long long)k.rank * (m + n);
            if (count) pieces.push_back(Piece{e->chunks[k.chunk].as<double>() + k.off, fo, count});
            fo += count;
        }
        *out = hmatrix_make(e->device, e->ne, nblocks, meta.data(), idx.data(), (int64_t)idx.size(), fo, false,
                            [&](void *dst) {
                                DevBuf dp;
                                upload(dp, pieces.data(), pieces.size(), e->stream);
                                for (size_t p0 = 0; p0 < pieces.size(); p0 += 65535) {
                                    const size_t cnt = std::min<size_t>(65535, pieces.size() - p0);
                                    hipLaunchKernelGGL(gather_kernel, dim3((unsigned)cnt), dim3(256), 0, e->stream,
                                                       dp.as<Piece>() + p0, static_cast<double *>(dst));
                                    PQ_HIP(hipGetLastError());
                                }
                                PQ_HIP(hipStreamSynchronize(e->stream));
                            });
    });
}

int pq_entries_stats(pq_entries *e, int64_t *entries, int64_t *reruns, int64_t *bytes, double *ms) {
    return guard([&] {
        PQ_REQUIRE(e, "pq_entries_stats: null argument");
        if (entries) *entries = e->entries;
        if (reruns) *reruns = e->reruns;
        if (bytes) {
            size_t b = 0;
            for (const DevBuf &c : e->chunks) b += c.bytes;
            *bytes = (int64_t)b;
        }
        if (ms) { ms[0] = e->aca_ms; ms[1] = e->dense_ms; }
    });
}

}  // extern "C"
