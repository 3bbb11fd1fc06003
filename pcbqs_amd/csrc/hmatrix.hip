// Product of a block-compressed (H-matrix) operator with a block of vectors, Y = A X.
//
// Semantics: HMatrix._matmat and the compiled product it dispatches to
// (src/pcbqs/hmatrix.py:223-297, src/pcbqs/core/hmatrix_kernels.hpp:20-101). A is given as a list
// of blocks (a, b, U, V, same): rows a and columns b of A hold U (dense, |a| x |b|) or U V (low
// rank, |a| x r times r x |b|); a block that is not `same` also stands for its transpose at rows b
// and columns a (A is transpose-symmetric), so it contributes U^T X[a] (dense) or V^T (U^T X[a])
// to rows b. The sum is deterministic in two phases: every block writes its contribution rows
// into slots of its own (forward rows, then transposed rows, slots numbered in block order), and
// every output row then adds its slots in a fixed order (a wave per row: each lane its share of
// the slots in slot order, then a butterfly over the lanes). Results therefore do not depend on
// the schedule, and agree with the compiled reference to rounding (the same terms, summed in a
// different order).
//
// On the device: one workgroup per block (dense and low-rank blocks in separate launches, the
// low-rank kernel holding V X[b] and U^T X[a] in LDS); each of the block's matrix-vector products
// gives a thread one output row and all columns of the current column group, splitting the inner
// dimension over thread groups (reduced through LDS in a fixed order) when the block has fewer
// than 256 output rows. X and Y are row-major (n x k); k is processed in groups of up to 8 real or
// 4 complex columns, so the slot buffer holds one group.
#include "pq_common.h"
#include "hmatrix.h"

#include <algorithm>
#include <vector>

namespace pq {

namespace {

constexpr int HM_THREADS = 256;
constexpr int HM_RMAX = 256;        // largest rank of a low-rank block (the LDS for V X[b] and U^T X[a])

struct cd {
    double re, im;
};
__device__ __forceinline__ void hm_zero(double &a) { a = 0.0; }
__device__ __forceinline__ void hm_zero(cd &a) { a.re = 0.0; a.im = 0.0; }
__device__ __forceinline__ void hm_madd(double &acc, double m, double v) { acc = fma(m, v, acc); }
__device__ __forceinline__ void hm_madd(cd &acc, const cd &m, const cd &v) {
    acc.re += m.re * v.re - m.im * v.im;
    acc.im += m.re * v.im + m.im * v.re;
}
__device__ __forceinline__ void hm_add(double &acc, double v) { acc += v; }
__device__ __forceinline__ void hm_add(cd &acc, const cd &v) {
    acc.re += v.re;
    acc.im += v.im;
}

// Block record on the device: a_off, la, b_off, lb, rank (-1: dense), same, fwd_slot, tr_slot,
// u_off, v_off (element offsets into the factor buffer).
struct HBlock {
    long long a_off, la, b_off, lb, rank, same, fwd, tr, u_off, v_off;
};

// out[o] = sum_q M(o, q) vec(q) for o < O, q < K, each output a KC-column row. ROW: M(o, q) =
// M[o ld + q] (a thread walks its own row); otherwise M(o, q) = M[q ld + o] (lanes along a row).
template <class T, int KC, bool ROW, class VecF, class OutF>
__device__ void hm_gemv(const T *__restrict__ M, long long ld, int O, int K, VecF vec, OutF out, T (*part)[KC]) {
    const int tid = threadIdx.x;
    for (int o0 = 0; o0 < O; o0 += HM_THREADS) {
        const int Oc = min(HM_THREADS, O - o0);
        const int G = HM_THREADS / Oc;                // thread groups splitting K (uniform)
        // adjacent lanes take adjacent outputs, so that the lanes of a wave share their gathered rows
        const int o = tid % Oc, g = tid / Oc;
        const bool act = g < G;
        T acc[KC];
#pragma unroll
        for (int c = 0; c < KC; ++c) hm_zero(acc[c]);
        if (act) {
#pragma unroll 4
            for (int q = g; q < K; q += G) {
                const T m = ROW ? M[(size_t)(o0 + o) * ld + q] : M[(size_t)q * ld + o0 + o];
                T v[KC];
                vec(q, v);
#pragma unroll
                for (int c = 0; c < KC; ++c) hm_madd(acc[c], m, v[c]);
            }
        }
        if (G == 1) {
            if (act) out(o0 + o, acc);
        } else {
            if (act) {
#pragma unroll
                for (int c = 0; c < KC; ++c) part[g * Oc + o][c] = acc[c];
            }
            __syncthreads();
            if (tid < Oc) {
                T r[KC];
#pragma unroll
                for (int c = 0; c < KC; ++c) r[c] = part[tid][c];
                for (int gg = 1; gg < G; ++gg) {
#pragma unroll
                    for (int c = 0; c < KC; ++c) hm_add(r[c], part[gg * Oc + tid][c]);
                }
                out(o0 + tid, r);
            }
            __syncthreads();
        }
    }
}

// Gathered rows of X (columns c0 .. c0 + kc - 1, zero beyond kc).
template <class T, int KC> struct XRows {
    const T *x;
    const int *ids;
    long long k;
    int c0, kc;
    __device__ void operator()(int q, T (&v)[KC]) const {
        const T *row = x + (size_t)ids[q] * k + c0;
#pragma unroll
        for (int c = 0; c < KC; ++c) {
            if (c < kc) v[c] = row[c];
            else hm_zero(v[c]);
        }
    }
};

// Gathered rows of X staged in LDS once per block (HM_XS_BYTES), so that the block's products
// read them from there instead of through an index load and a dependent global load each.
constexpr int HM_XS_BYTES = 16384;

template <class T, int KC> __device__ bool hm_stage(T (*xs)[KC], int rows, int cap, const XRows<T, KC> &X) {
    if (rows > cap) return false;                     // uniform over the workgroup
    __syncthreads();                                  // the previous product has finished reading xs
    for (int q = threadIdx.x; q < rows; q += HM_THREADS) {
        T v[KC];
        X(q, v);
#pragma unroll
        for (int c = 0; c < KC; ++c) xs[q][c] = v[c];
    }
    __syncthreads();
    return true;
}

template <class T, int KC> struct SlotRows {
    T *s;
    long long slot0;
    __device__ void operator()(int o, const T (&a)[KC]) const {
        T *dst = s + (size_t)(slot0 + o) * KC;
#pragma unroll
        for (int c = 0; c < KC; ++c) dst[c] = a[c];
    }
};

template <class T, int KC> struct LdsRows {
    T (*t)[KC];
    __device__ void operator()(int q, T (&v)[KC]) const {
#pragma unroll
        for (int c = 0; c < KC; ++c) v[c] = t[q][c];
    }
};

template <class T, int KC> struct LdsStore {
    T (*t)[KC];
    __device__ void operator()(int o, const T (&a)[KC]) const {
#pragma unroll
        for (int c = 0; c < KC; ++c) t[o][c] = a[c];
    }
};

// The kernels' LDS, sized per handle at launch (dynamic shared memory): the split-K partials
// (256 rows), then for low-rank blocks V X[b] and U^T X[a] (the handle's largest rank each), then
// the staged rows of X (the handle's largest block side, at most HM_XS_BYTES).
extern __shared__ __align__(16) unsigned char hm_lds[];

// Phase 1, dense blocks: forward rows U X[b]; unless `same`, transposed rows U^T X[a].
template <class T, int KC>
__global__ __launch_bounds__(HM_THREADS) void hm_dense_kernel(const HBlock *__restrict__ blocks,
                                                             const int *__restrict__ list, const int *__restrict__ idx,
                                                             const T *__restrict__ F, const T *__restrict__ x,
                                                             long long k, int c0, int kc, T *__restrict__ s,
                                                             int xs_rows) {
    T (*part)[KC] = reinterpret_cast<T (*)[KC]>(hm_lds);
    T (*xs)[KC] = part + HM_THREADS;
    const HBlock b = blocks[list[blockIdx.x]];
    const T *u = F + b.u_off;
    const int la = (int)b.la, lb = (int)b.lb;
    XRows<T, KC> xa{x, idx + b.a_off, k, c0, kc}, xb{x, idx + b.b_off, k, c0, kc};
    const LdsRows<T, KC> xl{xs};
    if (hm_stage(xs, lb, xs_rows, xb)) hm_gemv<T, KC, true>(u, lb, la, lb, xl, SlotRows<T, KC>{s, b.fwd}, part);
    else hm_gemv<T, KC, true>(u, lb, la, lb, xb, SlotRows<T, KC>{s, b.fwd}, part);
    if (!b.same) {
        if (hm_stage(xs, la, xs_rows, xa)) hm_gemv<T, KC, false>(u, lb, lb, la, xl, SlotRows<T, KC>{s, b.tr}, part);
        else hm_gemv<T, KC, false>(u, lb, lb, la, xa, SlotRows<T, KC>{s, b.tr}, part);
    }
}

// Phase 1, low-rank blocks: t = V X[b], q = U^T X[a] (LDS), then forward rows U t and, unless
// `same`, transposed rows V^T q.
template <class T, int KC>
__global__ __launch_bounds__(HM_THREADS) void hm_lowrank_kernel(const HBlock *__restrict__ blocks,
                                                               const int *__restrict__ list, const int *__restrict__ idx,
                                                               const T *__restrict__ F, const T *__restrict__ x,
                                                               long long k, int c0, int kc, T *__restrict__ s,
                                                               int rmax, int xs_rows) {
    T (*part)[KC] = reinterpret_cast<T (*)[KC]>(hm_lds);
    T (*tl)[KC] = part + HM_THREADS;
    T (*ql)[KC] = tl + rmax;
    T (*xs)[KC] = ql + rmax;
    const HBlock b = blocks[list[blockIdx.x]];
    const T *u = F + b.u_off, *v = F + b.v_off;
    const int la = (int)b.la, lb = (int)b.lb, r = (int)b.rank;
    XRows<T, KC> xa{x, idx + b.a_off, k, c0, kc}, xb{x, idx + b.b_off, k, c0, kc};
    const LdsRows<T, KC> xl{xs};
    if (hm_stage(xs, lb, xs_rows, xb)) hm_gemv<T, KC, true>(v, lb, r, lb, xl, LdsStore<T, KC>{tl}, part);
    else hm_gemv<T, KC, true>(v, lb, r, lb, xb, LdsStore<T, KC>{tl}, part);
    if (!b.same) {
        if (hm_stage(xs, la, xs_rows, xa)) hm_gemv<T, KC, false>(u, r, r, la, xl, LdsStore<T, KC>{ql}, part);
        else hm_gemv<T, KC, false>(u, r, r, la, xa, LdsStore<T, KC>{ql}, part);
    }
    __syncthreads();
    hm_gemv<T, KC, true>(u, r, la, r, LdsRows<T, KC>{tl}, SlotRows<T, KC>{s, b.fwd}, part);
    if (!b.same) hm_gemv<T, KC, false>(v, lb, lb, r, LdsRows<T, KC>{ql}, SlotRows<T, KC>{s, b.tr}, part);
}

__device__ __forceinline__ double hm_shfl_xor(double v, int m) { return __shfl_xor(v, m); }
__device__ __forceinline__ cd hm_shfl_xor(const cd &v, int m) { return cd{__shfl_xor(v.re, m), __shfl_xor(v.im, m)}; }

// Phase 2: Y[row] = the sum of the row's slots, one wave per row: lane l adds slots l, l + 64, ...
// in slot order, then a fixed butterfly adds the lanes (deterministic, schedule-independent).
template <class T, int KC>
__global__ __launch_bounds__(HM_THREADS) void hm_rows_kernel(const long long *__restrict__ rp,
                                                            const long long *__restrict__ sl, const T *__restrict__ s,
                                                            long long n, long long k, int c0, int kc,
                                                            T *__restrict__ y) {
    const long long row = (long long)blockIdx.x * (HM_THREADS / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;                          // whole waves leave together
    T acc[KC];
#pragma unroll
    for (int c = 0; c < KC; ++c) hm_zero(acc[c]);
    const long long p1 = rp[row + 1];
    for (long long p = rp[row] + lane; p < p1; p += 64) {
        const T *src = s + (size_t)sl[p] * KC;
#pragma unroll
        for (int c = 0; c < KC; ++c) hm_add(acc[c], src[c]);
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
#pragma unroll
        for (int c = 0; c < KC; ++c) hm_add(acc[c], hm_shfl_xor(acc[c], m));
    }
    if (lane == 0) {
        T *dst = y + (size_t)row * k + c0;
#pragma unroll
        for (int c = 0; c < KC; ++c)
            if (c < kc) dst[c] = acc[c];
    }
}

}  // namespace

}  // namespace pq

struct pq_hmatrix {
    int device = 0;
    long long n = 0, nslots = 0, n_factor = 0;
    bool complex_ = false;
    int n_dense = 0, n_lowrank = 0;
    long long rmax = 0, side_dense = 0, side_lr = 0;  // largest rank; largest block side of each kind
    hipStream_t stream = nullptr;
    pq::DevBuf blocks, idx, rp, sl, F, dense_list, lr_list, s, x, y;
    long long xk = 0;                              // columns the x / y buffers were last sized for
};

namespace pq {
namespace {

template <class T, int KC> void hm_apply(pq_hmatrix *h, long long k, const void *x, void *y, hipStream_t st) {
    const T *xd = static_cast<const T *>(x);
    T *yd = static_cast<T *>(y);
    const int cap = HM_XS_BYTES / (KC * (int)sizeof(T));
    const int xs_dense = (int)std::min<long long>(h->side_dense, cap), xs_lr = (int)std::min<long long>(h->side_lr, cap);
    const size_t row = KC * sizeof(T);
    const size_t lds_dense = (HM_THREADS + xs_dense) * row, lds_lr = (HM_THREADS + 2 * h->rmax + xs_lr) * row;
    for (long long c0 = 0; c0 < k; c0 += KC) {
        const int kc = (int)std::min<long long>(KC, k - c0);
        if (h->n_dense)
            hipLaunchKernelGGL((hm_dense_kernel<T, KC>), dim3(h->n_dense), dim3(HM_THREADS), lds_dense, st,
                               h->blocks.as<HBlock>(), h->dense_list.as<int>(), h->idx.as<int>(), h->F.as<T>(), xd, k,
                               (int)c0, kc, h->s.as<T>(), xs_dense);
        if (h->n_lowrank)
            hipLaunchKernelGGL((hm_lowrank_kernel<T, KC>), dim3(h->n_lowrank), dim3(HM_THREADS), lds_lr, st,
                               h->blocks.as<HBlock>(), h->lr_list.as<int>(), h->idx.as<int>(), h->F.as<T>(), xd, k,
                               (int)c0, kc, h->s.as<T>(), (int)h->rmax, xs_lr);
        if (h->n)
            hipLaunchKernelGGL((hm_rows_kernel<T, KC>), dim3((unsigned)((h->n + 3) / 4)),
                               dim3(HM_THREADS), 0, st, h->rp.as<long long>(), h->sl.as<long long>(), h->s.as<T>(),
                               h->n, k, (int)c0, kc, yd);
        PQ_HIP(hipGetLastError());
    }
}

// The column group is the smallest of 2, 4, 8 real (1, 2, 4 complex) columns that holds k, or the
// largest: fewer columns per group, fewer registers and no work on padding columns.
void apply_device(pq_hmatrix *h, long long k, const void *x, void *y, hipStream_t st) {
    if (h->complex_) {
        if (k <= 1) hm_apply<cd, 1>(h, k, x, y, st);
        else if (k <= 2) hm_apply<cd, 2>(h, k, x, y, st);
        else hm_apply<cd, 4>(h, k, x, y, st);
    } else {
        if (k <= 2) hm_apply<double, 2>(h, k, x, y, st);
        else if (k <= 4) hm_apply<double, 4>(h, k, x, y, st);
        else hm_apply<double, 8>(h, k, x, y, st);
    }
}

}  // namespace
}  // namespace pq

using namespace pq;

pq_hmatrix *pq::hmatrix_make(int device, int64_t n, int64_t n_blocks, const int64_t *meta, const int32_t *idx,
                             int64_t n_idx, int64_t n_factor, bool is_complex,
                             const std::function<void(void *)> &upload) {
    PQ_REQUIRE(n >= 0 && n < (1LL << 31) && n_blocks >= 0 && n_blocks < (1LL << 31) && n_idx >= 0 &&
                   n_factor >= 0,
               "pq_hmatrix_create: invalid sizes");
    std::vector<HBlock> hb((size_t)n_blocks);
    std::vector<int> dense, lowrank;
    long long so = 0, fo = 0;
    for (long long q = 0; q < n_blocks; ++q) {
        const int64_t *m = meta + 6 * q;
        HBlock &b = hb[(size_t)q];
        b.a_off = m[0]; b.la = m[1]; b.b_off = m[2]; b.lb = m[3]; b.rank = m[4]; b.same = m[5] != 0;
        PQ_REQUIRE(b.la >= 0 && b.lb >= 0 && b.a_off >= 0 && b.b_off >= 0 && b.a_off + b.la <= n_idx &&
                       b.b_off + b.lb <= n_idx && b.la < (1LL << 31) && b.lb < (1LL << 31),
                   "pq_hmatrix_create: block index range outside the index array");
        PQ_REQUIRE(b.rank >= -1 && b.rank <= HM_RMAX,
                   "pq_hmatrix_create: block rank must be -1 (dense) or 0..256");
        b.fwd = so;
        so += b.la;
        b.tr = b.same ? -1 : so;
        if (!b.same) so += b.lb;
        b.u_off = fo;
        if (b.rank < 0) {
            fo += b.la * b.lb;
            b.v_off = -1;
            dense.push_back((int)q);
        } else {
            fo += b.la * b.rank;
            b.v_off = fo;
            fo += b.rank * b.lb;
            lowrank.push_back((int)q);
        }
    }
    PQ_REQUIRE(fo == n_factor, "pq_hmatrix_create: factor array length does not match the blocks");
    for (int64_t t = 0; t < n_idx; ++t)
        PQ_REQUIRE(idx[t] >= 0 && idx[t] < n, "pq_hmatrix_create: row or column index out of range");
    // slots of each output row, in slot order (a stable counting sort of the slots' targets)
    std::vector<long long> rp((size_t)n + 1, 0), sl((size_t)so);
    for (const HBlock &b : hb) {
        for (long long i = 0; i < b.la; ++i) ++rp[(size_t)idx[b.a_off + i] + 1];
        if (!b.same)
            for (long long j = 0; j < b.lb; ++j) ++rp[(size_t)idx[b.b_off + j] + 1];
    }
    for (long long r = 0; r < n; ++r) rp[(size_t)r + 1] += rp[(size_t)r];
    {
        std::vector<long long> next(rp.begin(), rp.end() - 1);
        for (const HBlock &b : hb) {
            for (long long i = 0; i < b.la; ++i) sl[(size_t)next[(size_t)idx[b.a_off + i]]++] = b.fwd + i;
            if (!b.same)
                for (long long j = 0; j < b.lb; ++j) sl[(size_t)next[(size_t)idx[b.b_off + j]]++] = b.tr + j;
        }
    }
    DeviceScope scope(device);
    auto *h = new pq_hmatrix();
    try {
        h->device = device; h->n = n; h->nslots = so; h->n_factor = n_factor; h->complex_ = is_complex;
        h->n_dense = (int)dense.size(); h->n_lowrank = (int)lowrank.size();
        for (const HBlock &b : hb) {
            const long long side = std::max(b.la, b.lb);
            if (b.rank < 0) h->side_dense = std::max(h->side_dense, side);
            else {
                h->side_lr = std::max(h->side_lr, side);
                h->rmax = std::max(h->rmax, b.rank);
            }
        }
        const size_t es = h->complex_ ? 16 : 8;
        PQ_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        auto up = [&](DevBuf &d, const void *src, size_t bytes) {
            d.reserve(std::max<size_t>(bytes, 16));
            if (bytes) PQ_HIP(hipMemcpy(d.p, src, bytes, hipMemcpyHostToDevice));
        };
        up(h->blocks, hb.data(), hb.size() * sizeof(HBlock));
        up(h->idx, idx, (size_t)n_idx * 4);
        up(h->rp, rp.data(), rp.size() * 8);
        up(h->sl, sl.data(), sl.size() * 8);
        h->F.reserve(std::max<size_t>((size_t)n_factor * es, 16));
        if (n_factor) upload(h->F.p);
        up(h->dense_list, dense.data(), dense.size() * 4);
        up(h->lr_list, lowrank.data(), lowrank.size() * 4);
        h->s.reserve(std::max<size_t>((size_t)so * 8 * 8, 16));   // one column group: 8 real or 4 complex
    } catch (...) {
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        throw;
    }
    return h;
}

extern "C" {

int pq_hmatrix_create(int device, int64_t n, int64_t n_blocks, const int64_t *meta, const int32_t *idx, int64_t n_idx,
                      const void *factors, int64_t n_factor, int is_complex, pq_hmatrix **out) {
    return guard([&] {
        PQ_REQUIRE(out && (n_blocks == 0 || (meta && idx && factors)), "pq_hmatrix_create: null argument");
        const size_t es = is_complex ? 16 : 8;
        *out = hmatrix_make(device, n, n_blocks, meta, idx, n_idx, n_factor, is_complex != 0, [&](void *dst) {
            PQ_HIP(hipMemcpy(dst, factors, (size_t)n_factor * es, hipMemcpyHostToDevice));
        });
    });
}

int pq_hmatrix_destroy(pq_hmatrix *h) {
    if (!h) return PQ_OK;
    return guard([&] {
        DeviceScope scope(h->device);
        if (h->stream) {
            (void)hipStreamSynchronize(h->stream);
            (void)hipStreamDestroy(h->stream);
        }
        delete h;
    });
}

int pq_hmatrix_info(pq_hmatrix *h, int64_t info[5]) {
    return guard([&] {
        PQ_REQUIRE(h && info, "pq_hmatrix_info: null argument");
        info[0] = h->n;
        info[1] = h->nslots;
        info[2] = h->n_factor * (h->complex_ ? 16 : 8);
        info[3] = h->n_dense;
        info[4] = h->n_lowrank;
    });
}

int pq_hmatrix_apply_device(pq_hmatrix *h, int64_t k, const void *x_dev, void *y_dev, void *stream) {
    return guard([&] {
        PQ_REQUIRE(h && k >= 0 && (k == 0 || h->n == 0 || (x_dev && y_dev)), "pq_hmatrix_apply_device: invalid argument");
        if (!k || !h->n) return;
        DeviceScope scope(h->device);
        apply_device(h, k, x_dev, y_dev, stream ? static_cast<hipStream_t>(stream) : h->stream);
    });
}

int pq_hmatrix_apply(pq_hmatrix *h, int64_t k, const void *x, void *y) {
    return guard([&] {
        PQ_REQUIRE(h && k >= 0 && (k == 0 || h->n == 0 || (x && y)), "pq_hmatrix_apply: invalid argument");
        if (!k || !h->n) return;
        DeviceScope scope(h->device);
        const size_t bytes = (size_t)h->n * (size_t)k * (h->complex_ ? 16 : 8);
        h->x.reserve(bytes);
        h->y.reserve(bytes);
        h->xk = k;
        PQ_HIP(hipMemcpyAsync(h->x.p, x, bytes, hipMemcpyHostToDevice, h->stream));
        apply_device(h, k, h->x.p, h->y.p, h->stream);
        PQ_HIP(hipMemcpyAsync(y, h->y.p, bytes, hipMemcpyDeviceToHost, h->stream));
        PQ_HIP(hipStreamSynchronize(h->stream));
    });
}

int pq_hmatrix_time(pq_hmatrix *h, int reps, double *ms) {
    return guard([&] {
        PQ_REQUIRE(h && ms && reps > 0, "pq_hmatrix_time: invalid argument");
        PQ_REQUIRE(h->xk > 0, "pq_hmatrix_time: call pq_hmatrix_apply first (the timed products reuse its vectors)");
        DeviceScope scope(h->device);
        hipEvent_t e0, e1;
        PQ_HIP(hipEventCreate(&e0));
        PQ_HIP(hipEventCreate(&e1));
        PQ_HIP(hipEventRecord(e0, h->stream));
        for (int r = 0; r < reps; ++r) apply_device(h, h->xk, h->x.p, h->y.p, h->stream);
        PQ_HIP(hipEventRecord(e1, h->stream));
        PQ_HIP(hipEventSynchronize(e1));
        float t = 0.f;
        PQ_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = t / reps;
    });
}

}  // extern "C"
