// Triangle-pair 1/R moment fill for gfx950.
//
// Semantics follow src/pcbqs/core/static_kernels.hpp:53-252 (static_pair, pair_kind,
// static_fill) and the device row-strip contract of src/pcbqs/static_cupy.py:127-197.
//
// Work split:
//   far32_tile_kernel  the device-resident FP32 fill: 512-thread blocks over the 64 x 64 tiles of
//                the upper triangle. Row and column triangles' records sit in LDS; each unordered
//                pair is integrated once in its canonical order; forward entries are stored from
//                registers and reverse entries are staged through LDS so that both are coalesced
//                row segments. Moments are FP32 in each triangle's local frame.
//   far32_kernel / far64_kernel  row-strip fills (the host fill_symmetric path and the FP64
//                parity layout): four rows per wave, the 64 lanes over consecutive columns, every
//                entry from its canonical pair (min, max), so reciprocal entries are bitwise equal.
//   Pairs the near rule keeps analytic are not integrated by the far kernels: they go to
//                block-local LDS lists, flushed with one global atomic per block and level.
//   near kernel  FP64 closed-form source integrals, one thread per near pair looping over the
//                pair's observation points, summed sequentially in point order (the reference's
//                order); both entries stored through the L2 (scattered partial lines merge there).
// Classification is the reference rule evaluated in FP64 without contraction; an FP32
// prefilter with a 1e-5 relative margin settles the clearly-far pairs without it.
#include "pq_common.h"
#include "pair.h"
#include <mutex>
#include <cstdlib>
#include <cmath>
#include <vector>
#include <algorithm>

namespace pq {

constexpr int WAVE = 64;
constexpr int FAR_ROWS_PER_WAVE = 4;
constexpr int FAR_ROWS_PER_BLOCK = 4 * FAR_ROWS_PER_WAVE;  // 4 waves per 256-thread block

struct NearLists {
    unsigned long long *list[2];   // packed (row << 32) | col entries
    unsigned long long *count;     // [0],[1]: list lengths; [2]: analytic pairs with row <= col
    unsigned long long cap;
};


// Entry (r, c) of the planes from the canonical pair's moments (static_cupy_kernels.py:106-115
// contract; diagonal: the mean of the two first moments, static_kernels.hpp:230-236).
template <class T, bool SCALAR>
__device__ __forceinline__ void store_entry(T *out, size_t plane, size_t i, int r, int c, const double *m,
                                            unsigned *flag) {
    bool ok = isfinite(m[0]);
    // planes are written once and read by a later kernel or copy: streaming (non-temporal) stores
    __builtin_nontemporal_store((T)m[0], out + i);
    if (!SCALAR) {
        double mx, my;
        if (r == c) { mx = (m[1] + m[3]) / 2; my = (m[2] + m[4]) / 2; }
        else if (r < c) { mx = m[1]; my = m[2]; }
        else { mx = m[3]; my = m[4]; }
        __builtin_nontemporal_store((T)mx, out + plane + i);
        __builtin_nontemporal_store((T)my, out + 2 * plane + i);
        __builtin_nontemporal_store((T)m[5], out + 3 * plane + i);
        ok = ok && isfinite(mx) && isfinite(my) && isfinite(m[5]);
    }
    if (!ok) atomicOr(flag, 1u);
}

// ---- near-list append, aggregated per wave ------------------------------------------------

__device__ __forceinline__ void wave_append(bool want, unsigned long long *list, unsigned long long *count,
                                            unsigned long long cap, unsigned long long entry) {
    const unsigned long long mask = __ballot(want);
    if (!mask) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    const unsigned long long rank = __popcll(mask & ((1ull << lane) - 1ull));
    if (want && base + rank < cap) list[base + rank] = entry;
}

// ---- far kernels ---------------------------------------------------------------------------

struct FarArgs {
    int r0, rows, n;        // rows [r0, r0+rows) of the n x n matrix; out row 0 is row r0
    long long ld;           // elements between rows of a plane (>= n)
    double near;
    float near32;
    int all_near, outer, mirror;
    size_t plane;           // elements between planes
    NearLists lists;
    unsigned *flag;
};

// Classification of entry (r, c): -1 far, else the analytic level. An FP32 prefilter (the
// differences taken in FP64, then rounded) settles the clearly-far pairs; its relative error is
// below 1e-6, far inside the 1e-5 margin, so it never decides a pair the exact rule would not.
struct TriVals {
    double x, y, z, rad;
};

__device__ __forceinline__ int classify(const TriVals &r, const TriVals &c, const FarArgs &A) {
    const float dx = (float)(r.x - c.x), dy = (float)(r.y - c.y), dz = (float)(r.z - c.z);
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float thr = A.near32 * ((float)r.rad + (float)c.rad);
    const bool certain_far = !A.all_near && d2 > thr * thr * 1.00001f;
    if (certain_far) return -1;
    return exact_kind(r.x, r.y, r.z, r.rad, c.x, c.y, c.z, c.rad, A.near, A.all_near, A.outer);
}

__device__ __forceinline__ TriVals tri_vals(const GeomDev &g, int t) {
    return TriVals{g.cent[2 * t], g.cent[2 * t + 1], g.z[t], g.radius[t]};
}

// Near entries are collected per block in LDS (cheap LDS atomics, one per wave and list) and
// reserved in the global lists once per block and list: device-scope atomics on three shared
// counters, issued per wave from every XCD, had serialized the whole fill.
constexpr int NEAR_CAP = 512;   // LDS entries per list and block; overflow goes straight to global

struct BlockNear {
    unsigned long long *list[2];  // LDS lists
    unsigned *count;              // LDS: [0],[1] list lengths (may exceed NEAR_CAP), [2] analytic, [3],[4] bases
    unsigned long long *gbase;    // LDS: reserved global offsets of the two lists
};

__device__ __forceinline__ void block_near_init(const BlockNear &B) {
    if (threadIdx.x < 5) B.count[threadIdx.x] = 0;
}

__device__ __forceinline__ void block_append(const BlockNear &B, const FarArgs &A, int which, bool want,
                                             unsigned long long entry) {
    const unsigned long long mask = __ballot(want);
    if (!mask) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)mask) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(B.count + which, (unsigned)__popcll(mask));
    base = __shfl(base, leader);
    const unsigned slot = base + (unsigned)__popcll(mask & ((1ull << lane) - 1ull));
    if (want && slot < NEAR_CAP) B.list[which][slot] = entry;
    // overflow of the LDS list (dense near rules): straight to the global list
    wave_append(want && slot >= NEAR_CAP, A.lists.list[which], A.lists.count + which, A.lists.cap, entry);
}

__device__ __forceinline__ void block_near(const BlockNear &B, const FarArgs &A, bool valid, int kind, int r, int c) {
    const bool near_entry = valid && kind >= 0 && (r <= c || !A.mirror);
    const unsigned long long e = ((unsigned long long)(unsigned)r << 32) | (unsigned)c;
    const bool l1 = kind > 0;
    block_append(B, A, 0, near_entry && !l1, e);
    block_append(B, A, 1, near_entry && l1, e);
    const unsigned long long am = __ballot(A.mirror ? near_entry : (valid && kind >= 0 && r <= c));
    if (am && __lane_id() == __ffsll((long long)am) - 1) atomicAdd(B.count + 2, (unsigned)__popcll(am));
}

// After a __syncthreads(): reserve global space once per list, then copy the LDS lists out.
__device__ __forceinline__ void block_near_flush(const BlockNear &B, const FarArgs &A) {
    if (threadIdx.x < 2) {
        const unsigned n = min(B.count[threadIdx.x], (unsigned)NEAR_CAP);
        B.gbase[threadIdx.x] = n ? atomicAdd(A.lists.count + threadIdx.x, (unsigned long long)n) : 0ull;
    } else if (threadIdx.x == 2 && B.count[2]) {
        atomicAdd(A.lists.count + 2, (unsigned long long)B.count[2]);
    }
    __syncthreads();
    for (int which = 0; which < 2; ++which) {
        const unsigned n = min(B.count[which], (unsigned)NEAR_CAP);
        for (unsigned i = threadIdx.x; i < n; i += blockDim.x) {
            const unsigned long long at = B.gbase[which] + i;
            if (at < A.lists.cap) A.lists.list[which][at] = B.list[which][i];
        }
    }
}

// FP32 local-frame far kernel.
template <bool SCALAR>
__global__ __launch_bounds__(256) void far32_kernel(GeomDev g, FarArgs A, float *out) {
    __shared__ unsigned long long nlist[2][NEAR_CAP], nbase[2];
    __shared__ unsigned ncount[5];
    const BlockNear B{{nlist[0], nlist[1]}, ncount, nbase};
    block_near_init(B);
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.x * WAVE + lane;
    const bool cvalid = c < A.n;
    const int cc = cvalid ? c : A.n - 1;
    const float4 c0 = g.far32[2 * cc], c1 = g.far32[2 * cc + 1];
    const TriVals cv = tri_vals(g, cc);
    __builtin_amdgcn_s_waitcnt(0);
    const int rbase = blockIdx.y * FAR_ROWS_PER_BLOCK + wave * FAR_ROWS_PER_WAVE;
    for (int k = 0; k < FAR_ROWS_PER_WAVE; ++k) {
        const int rl = rbase + k;
        if (rl >= A.rows) break;                       // wave-uniform: the wave leaves together
        const bool rvalid = true;
        const int r = A.r0 + rl;                       // wave-uniform: scalar loads below
        const float4 q0 = g.far32[2 * r], q1 = g.far32[2 * r + 1];
        const TriVals rv = tri_vals(g, r);
        const bool rfirst = r <= c;
        const int kind = classify(rv, cv, A);
        block_near(B, A, cvalid && rvalid, kind, r, c);
        // canonical pair (a, b) = (min, max): frame offset c_a - c_b, rounded once
        const float ox = (float)(rv.x - cv.x), oy = (float)(rv.y - cv.y);
        const float dcx = rfirst ? ox : -ox, dcy = rfirst ? oy : -oy;
        const float dz = (float)(rv.z - cv.z);
        const float dz2 = dz * dz;
        const float4 a0 = rfirst ? q0 : c0, a1 = rfirst ? q1 : c1;
        const float4 b0 = rfirst ? c0 : q0, b1 = rfirst ? c1 : q1;
        const float ax[3] = {a0.x, a0.z, a1.x}, ay[3] = {a0.y, a0.w, a1.y};
        const float bx[3] = {b0.x, b0.z, b1.x}, by[3] = {b0.y, b0.w, b1.y};
        const float w = a1.z * b1.z * (1.0f / 9.0f);
        float G[3] = {0, 0, 0}, H[3] = {0, 0, 0}, Tx[3] = {0, 0, 0}, Ty[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const float px = ax[i] + dcx, py = ay[i] + dcy;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const float dx = px - bx[j], dy = py - by[j];
                const float r2 = dx * dx + dy * dy + dz2;
                const float gv = w * __builtin_amdgcn_rsqf(r2);
                G[i] += gv;
                if (!SCALAR) {
                    H[j] += gv;
                    Tx[i] += gv * bx[j]; Ty[i] += gv * by[j];
                }
            }
        }
        if (!cvalid || !rvalid || kind >= 0) continue;
        const size_t idx = (size_t)rl * A.ld + c;
        const float m0 = G[0] + G[1] + G[2];
        bool ok = isfinite(m0);
        out[idx] = m0;
        if (!SCALAR) {
            const float rox = G[0] * ax[0] + G[1] * ax[1] + G[2] * ax[2];
            const float roy = G[0] * ay[0] + G[1] * ay[1] + G[2] * ay[2];
            const float rsx = H[0] * bx[0] + H[1] * bx[1] + H[2] * bx[2];
            const float rsy = H[0] * by[0] + H[1] * by[1] + H[2] * by[2];
            const float rr = ax[0] * Tx[0] + ay[0] * Ty[0] + ax[1] * Tx[1] + ay[1] * Ty[1] + ax[2] * Tx[2] +
                             ay[2] * Ty[2];
            const float mx = rfirst ? rox : rsx, my = rfirst ? roy : rsy;
            out[A.plane + idx] = mx;
            out[2 * A.plane + idx] = my;
            out[3 * A.plane + idx] = rr;
            ok = ok && isfinite(mx) && isfinite(my) && isfinite(rr);
        }
        if (!ok) atomicOr(A.flag, 1u);
    }
    __syncthreads();
    block_near_flush(B, A);
}

// FP32 local-frame far kernel over upper-triangular 64 x 64 tiles: each unordered pair is
// integrated once, its row r < column c order is canonical without any selection, the forward
// entries (r, c) are stored straight from registers and the reverse entries (c, r) are staged in
// LDS (4 planes x 64 x 65 floats; the odd row pitch keeps the column-wise writes conflict-free)
// and stored as coalesced rows once the tile is done. An entry that must not be stored from here
// (an analytic pair, the lower half of a diagonal tile, beyond the matrix) is marked with a NaN
// in the M0 stage; a computed non-finite moment still sets the error flag.
constexpr int TILE = 64;

__device__ __forceinline__ void tile_coords(unsigned t, int nb, int &br, int &bc) {
    // row-major enumeration of the upper-triangular tiles: row br holds nb - br tiles
    const double N = 2.0 * nb + 1.0;
    int b = (int)floor((N - sqrt(N * N - 8.0 * (double)t)) * 0.5);
    b = b < 0 ? 0 : (b >= nb ? nb - 1 : b);
    auto first = [nb](int row) { return (long long)row * nb - (long long)row * (row - 1) / 2; };
    while (b > 0 && first(b) > (long long)t) --b;
    while (b + 1 < nb && first(b + 1) <= (long long)t) ++b;
    br = b;
    bc = b + (int)((long long)t - first(b));
}

template <bool SCALAR>
__global__ __launch_bounds__(512) void far32_tile_kernel(GeomDev g, FarArgs A, float *out, int nb) {
    constexpr int NP = SCALAR ? 1 : 4;
    constexpr int WAVES = 8, HALF = TILE / 2, ROWS = HALF / WAVES, PITCH = HALF + 1;
    __shared__ float stage[NP * TILE * PITCH];
    __shared__ float4 rowf[2 * TILE], colf[2 * TILE];
    __shared__ TriVals rowd[TILE], cold[TILE];
    __shared__ unsigned long long nlist[2][NEAR_CAP], nbase[2];
    __shared__ unsigned ncount[5];
    const BlockNear B{{nlist[0], nlist[1]}, ncount, nbase};
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int br, bc;
    tile_coords(blockIdx.x, nb, br, bc);
    const int R0 = br * TILE, C0 = bc * TILE;
    // The tile's 64 row and 64 column triangles go through LDS: the row loop below then holds no
    // global load, so no wait on the (in-order) vector memory counter ever covers the stores of
    // earlier rows.
    if (threadIdx.x < 2 * TILE) {
        const int i = threadIdx.x & (TILE - 1);
        const bool row = threadIdx.x < TILE;
        const int t = min((row ? R0 : C0) + i, A.n - 1);
        float4 *f = row ? rowf : colf;
        f[2 * i] = g.far32[2 * t];
        f[2 * i + 1] = g.far32[2 * t + 1];
        (row ? rowd : cold)[i] = tri_vals(g, t);
    }
    block_near_init(B);
    __syncthreads();
    const int c = C0 + lane;
    const bool cvalid = c < A.n;
    const float4 c0 = colf[2 * lane], c1 = colf[2 * lane + 1];
    const float bx[3] = {c0.x, c0.z, c1.x}, by[3] = {c0.y, c0.w, c1.y};
    const float barea = c1.z;
    const TriVals cv = cold[lane];
    for (int half = 0; half < 2; ++half) {
#pragma unroll 2
        for (int k = 0; k < ROWS; ++k) {
            const int hl = wave * ROWS + k;            // row within the half
            const int rl = half * HALF + hl;
            const int r = R0 + rl;                     // wave-uniform
            const bool valid = cvalid && r < A.n && r <= c;
            const float4 q0 = rowf[2 * rl], q1 = rowf[2 * rl + 1];
            const TriVals rv = rowd[rl];
            const int kind = valid ? classify(rv, cv, A) : -1;
            block_near(B, A, valid, kind, r, c);
            const float dcx = (float)(rv.x - cv.x), dcy = (float)(rv.y - cv.y), dz = (float)(rv.z - cv.z);
            const float dz2 = dz * dz;
            const float ax[3] = {q0.x, q0.z, q1.x}, ay[3] = {q0.y, q0.w, q1.y};
            const float w = q1.z * barea * (1.0f / 9.0f);
            float G[3] = {0, 0, 0}, H[3] = {0, 0, 0}, Tx[3] = {0, 0, 0}, Ty[3] = {0, 0, 0};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const float px = ax[i] + dcx, py = ay[i] + dcy;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    const float dx = px - bx[j], dy = py - by[j];
                    const float r2 = dx * dx + dy * dy + dz2;
                    const float gv = w * __builtin_amdgcn_rsqf(r2);
                    G[i] += gv;
                    if (!SCALAR) {
                        H[j] += gv;
                        Tx[i] += gv * bx[j]; Ty[i] += gv * by[j];
                    }
                }
            }
            const bool store = valid && kind < 0;
            const float m0 = G[0] + G[1] + G[2];
            float rox = 0, roy = 0, rsx = 0, rsy = 0, rr = 0;
            if (!SCALAR) {
                rox = G[0] * ax[0] + G[1] * ax[1] + G[2] * ax[2];
                roy = G[0] * ay[0] + G[1] * ay[1] + G[2] * ay[2];
                rsx = H[0] * bx[0] + H[1] * bx[1] + H[2] * bx[2];
                rsy = H[0] * by[0] + H[1] * by[1] + H[2] * by[2];
                rr = ax[0] * Tx[0] + ay[0] * Ty[0] + ax[1] * Tx[1] + ay[1] * Ty[1] + ax[2] * Tx[2] + ay[2] * Ty[2];
            }
            if (store) {
                // the planes are written once and read by a later kernel: streaming (non-temporal) stores
                const size_t idx = (size_t)r * A.ld + c;
                __builtin_nontemporal_store(m0, out + idx);
                if (!SCALAR) {
                    __builtin_nontemporal_store(rox, out + A.plane + idx);
                    __builtin_nontemporal_store(roy, out + 2 * A.plane + idx);
                    __builtin_nontemporal_store(rr, out + 3 * A.plane + idx);
                }
                const bool ok = isfinite(m0) && (SCALAR || (isfinite(rox) && isfinite(roy) && isfinite(rsx) &&
                                                            isfinite(rsy) && isfinite(rr)));
                if (!ok) atomicOr(A.flag, 1u);
            }
            // reverse entry (c, r): staged at [c - C0][r - R0 - half * 32]
            const int sidx = lane * PITCH + hl;
            stage[sidx] = (store && r != c) ? m0 : __builtin_nanf("");
            if (!SCALAR) {
                stage[TILE * PITCH + sidx] = rsx;
                stage[2 * TILE * PITCH + sidx] = rsy;
                stage[3 * TILE * PITCH + sidx] = rr;
            }
        }
        __syncthreads();
        // the staged half: 64 matrix rows C0.. of 32 columns, two rows per wave instruction
        const int sub = lane >> 5, col = lane & 31;
        for (int k = 0; k < TILE / (2 * WAVES); ++k) {
            const int cl = (wave * (TILE / (2 * WAVES)) + k) * 2 + sub;
            const int sidx = cl * PITCH + col;
            const float m0 = stage[sidx];
            if (!isnan(m0)) {
                const size_t idx = (size_t)(C0 + cl) * A.ld + R0 + half * HALF + col;
                __builtin_nontemporal_store(m0, out + idx);
                if (!SCALAR) {
                    __builtin_nontemporal_store(stage[TILE * PITCH + sidx], out + A.plane + idx);
                    __builtin_nontemporal_store(stage[2 * TILE * PITCH + sidx], out + 2 * A.plane + idx);
                    __builtin_nontemporal_store(stage[3 * TILE * PITCH + sidx], out + 3 * A.plane + idx);
                }
            }
        }
        __syncthreads();
    }
    block_near_flush(B, A);
}

// FP64 reference-layout far kernel by upper-triangular 64 x 64 tiles (the whole matrix, both
// entries of every pair): each unordered far pair r < c is integrated once with far64's arithmetic.
// Its entry (r, c) is stored along row r (the lanes run over c); the reverse entry (c, r) is staged
// in LDS, 16 tile rows at a time, and stored along rows c too. Near pairs go to the block's near
// lists as in the FP32 tile kernel (the near kernels write both of their entries).
template <bool SCALAR>
__global__ __launch_bounds__(512) void far64_tile_kernel(GeomDev g, FarArgs A, double *out, int nb) {
    constexpr int NP = SCALAR ? 1 : 4;
    constexpr int WAVES = 8, PART = 16, PARTS = TILE / PART, ROWS = PART / WAVES, PITCH = PART + 1;
    constexpr int RW = 7;                                  // far points (6) and area per triangle
    __shared__ double stage[NP * TILE * PITCH];
    __shared__ double rowp[TILE * RW], colp[TILE * RW];
    __shared__ TriVals rowd[TILE], cold[TILE];
    __shared__ unsigned long long nlist[2][NEAR_CAP], nbase[2];
    __shared__ unsigned ncount[5];
    const BlockNear B{{nlist[0], nlist[1]}, ncount, nbase};
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    int br, bc;
    tile_coords(blockIdx.x, nb, br, bc);
    const int R0 = br * TILE, C0 = bc * TILE;
    if (threadIdx.x < 2 * TILE) {
        const int i = threadIdx.x & (TILE - 1);
        const bool row = threadIdx.x < TILE;
        const int t = min((row ? R0 : C0) + i, A.n - 1);
        double *p = (row ? rowp : colp) + i * RW;
#pragma unroll
        for (int q = 0; q < 6; ++q) p[q] = g.P3[6 * (size_t)t + q];
        p[6] = g.area[t];
        (row ? rowd : cold)[i] = tri_vals(g, t);
    }
    block_near_init(B);
    __syncthreads();
    const int c = C0 + lane;
    const bool cvalid = c < A.n;
    double pc[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) pc[q] = colp[lane * RW + q];
    const double areac = colp[lane * RW + 6];
    const TriVals cv = cold[lane];
    for (int part = 0; part < PARTS; ++part) {
#pragma unroll
        for (int k = 0; k < ROWS; ++k) {
            const int hl = wave * ROWS + k;                // row within the part
            const int rl = part * PART + hl;
            const int r = R0 + rl;                         // wave-uniform
            const bool valid = cvalid && r < A.n && r <= c;
            const TriVals rv = rowd[rl];
            const int kind = valid ? classify(rv, cv, A) : -1;
            block_near(B, A, valid, kind, r, c);
            const bool store = valid && kind < 0;
            double pr[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) pr[q] = rowp[rl * RW + q];
            double m[6];
            far64<SCALAR>(pr, pc, rowp[rl * RW + 6] * areac / 9, rv.z - cv.z, m);
            if (store) {
                store_entry<double, SCALAR>(out, A.plane, (size_t)r * A.ld + c, r, c, m, A.flag);
                if (!SCALAR && !(isfinite(m[3]) && isfinite(m[4]))) atomicOr(A.flag, 1u);
            }
            // reverse entry (c, r): staged at [c - C0][r - R0 - part * PART]
            const int sidx = lane * PITCH + hl;
            stage[sidx] = (store && r != c) ? m[0] : __builtin_nan("");
            if (!SCALAR) {
                stage[TILE * PITCH + sidx] = m[3];
                stage[2 * TILE * PITCH + sidx] = m[4];
                stage[3 * TILE * PITCH + sidx] = m[5];
            }
        }
        __syncthreads();
        // the staged part: 64 matrix rows C0.. of PART columns, four rows per wave instruction
        const int sub = lane / PART, col = lane % PART;
        for (int k = 0; k < TILE / (4 * WAVES); ++k) {
            const int cl = (wave * (TILE / (4 * WAVES)) + k) * 4 + sub;
            const int sidx = cl * PITCH + col;
            const double m0 = stage[sidx];
            if (!isnan(m0)) {
                const size_t idx = (size_t)(C0 + cl) * A.ld + R0 + part * PART + col;
                __builtin_nontemporal_store(m0, out + idx);
                if (!SCALAR) {
                    __builtin_nontemporal_store(stage[TILE * PITCH + sidx], out + A.plane + idx);
                    __builtin_nontemporal_store(stage[2 * TILE * PITCH + sidx], out + 2 * A.plane + idx);
                    __builtin_nontemporal_store(stage[3 * TILE * PITCH + sidx], out + 3 * A.plane + idx);
                }
            }
        }
        __syncthreads();
    }
    block_near_flush(B, A);
}

// FP64 reference-layout far kernel over row strips (the strips of fill_symmetric's host planes):
// every entry of the strip from its canonical pair.
template <bool SCALAR>
__global__ __launch_bounds__(256) void far64_kernel(GeomDev g, FarArgs A, double *out) {
    __shared__ unsigned long long nlist[2][NEAR_CAP], nbase[2];
    __shared__ unsigned ncount[5];
    const BlockNear B{{nlist[0], nlist[1]}, ncount, nbase};
    block_near_init(B);
    __syncthreads();
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int c = blockIdx.x * WAVE + lane;
    const bool cvalid = c < A.n;
    const int cc = cvalid ? c : A.n - 1;
    double pc[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) pc[i] = g.P3[6 * (size_t)cc + i];
    const double areac = g.area[cc];
    const TriVals cv = tri_vals(g, cc);
    // The row loop stores; settle the column loads here so that no wait inside the loop (the
    // vector memory counter is in order) has to cover the stores of earlier rows.
    __builtin_amdgcn_s_waitcnt(0);
    const int rbase = blockIdx.y * FAR_ROWS_PER_BLOCK + wave * FAR_ROWS_PER_WAVE;
    for (int k = 0; k < FAR_ROWS_PER_WAVE; ++k) {
        const int rl = rbase + k;
        if (rl >= A.rows) break;                       // wave-uniform: the wave leaves together
        const bool rvalid = true;
        const int r = A.r0 + rl;
        double pr[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) pr[i] = g.P3[6 * (size_t)r + i];
        const double arear = g.area[r];
        const TriVals rv = tri_vals(g, r);
        const bool rfirst = r <= c;
        const int kind = classify(rv, cv, A);
        block_near(B, A, cvalid && rvalid, kind, r, c);
        double pa[6], pb[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) { pa[i] = rfirst ? pr[i] : pc[i]; pb[i] = rfirst ? pc[i] : pr[i]; }
        const double w = (rfirst ? arear : areac) * (rfirst ? areac : arear) / 9;
        const double d = rfirst ? rv.z - cv.z : cv.z - rv.z;
        double m[6];
        far64<SCALAR>(pa, pb, w, d, m);
        if (!cvalid || !rvalid || kind >= 0) continue;
        store_entry<double, SCALAR>(out, A.plane, (size_t)rl * A.ld + c, r, c, m, A.flag);
    }
    __syncthreads();
    block_near_flush(B, A);
}

// ---- near kernel -----------------------------------------------------------------------------

// Entry (r, c) as store_entry, with ordinary (write-back) stores: the near entries are scattered
// (a pair's reverse entry lies down a column), and cached stores let the L2 merge the partial lines
// that nearby pairs write instead of sending every 4- or 8-byte entry to HBM on its own.
template <class T, bool SCALAR>
__device__ __forceinline__ void store_entry_cached(T *out, size_t plane, size_t i, int r, int c, const double *m,
                                                   unsigned *flag) {
    bool ok = isfinite(m[0]);
    out[i] = (T)m[0];
    if (!SCALAR) {
        double mx, my;
        if (r == c) { mx = (m[1] + m[3]) / 2; my = (m[2] + m[4]) / 2; }
        else if (r < c) { mx = m[1]; my = m[2]; }
        else { mx = m[3]; my = m[4]; }
        out[plane + i] = (T)mx;
        out[2 * plane + i] = (T)my;
        out[3 * plane + i] = (T)m[5];
        ok = ok && isfinite(mx) && isfinite(my) && isfinite(m[5]);
    }
    if (!ok) atomicOr(flag, 1u);
}

// One thread per near pair of a list: the pair's observation points in order, each point's
// rounded moments added to the pair's sums in that order (the reference's sequential sum), then
// both entries stored. No barrier and no exchange between threads: the first layout (a thread
// per (pair, point), the sums through LDS between two block barriers) spent most of its time
// waiting on the list and geometry loads of every iteration.
template <class T, bool SCALAR, bool LOCAL>
__global__ __launch_bounds__(256) void near_kernel(GeomDev g, const unsigned long long *list,
                                                  const unsigned long long *count, unsigned long long cap,
                                                  int level, int r0, long long ld, int mirror, T *out,
                                                  size_t plane, unsigned *flag) {
    const unsigned long long total = min(*count, cap);
    const int nq = g.nq[level];
    for (unsigned long long idx = (unsigned long long)blockIdx.x * 256 + threadIdx.x; idx < total;
         idx += (unsigned long long)gridDim.x * 256) {
        const unsigned long long e = list[idx];
        const int r = (int)(e >> 32), c = (int)(e & 0xffffffffu);
        const int a = min(r, c), b = max(r, c);
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < nq; ++j) {
            double m[6] = {0, 0, 0, 0, 0, 0};
            near_point<SCALAR, LOCAL>(g, a, b, level, j, m);
#pragma unroll
            for (int k = 0; k < 6; ++k) s[k] = __dadd_rn(s[k], m[k]);
        }
        store_entry_cached<T, SCALAR>(out, plane, (size_t)(r - r0) * ld + c, r, c, s, flag);
        if (mirror && r != c) store_entry_cached<T, SCALAR>(out, plane, (size_t)(c - r0) * ld + r, c, r, s, flag);
    }
}

// ---- explicit pair list ---------------------------------------------------------------------

template <bool SCALAR>
__global__ __launch_bounds__(128) void pairs_kernel(GeomDev g, const int *P, const int *Q, const double *D,
                                                   const int *K, const int *order, long long n, double *out) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long long i = order[t];
    double m[6];
    pair64<SCALAR>(g, P[i], Q[i], D[i], K[i], m);
    const int width = SCALAR ? 1 : 6;
    for (int k = 0; k < width; ++k) out[i * width + k] = m[k];
}

// The kind buckets of an explicit pair list on the device (pq_pairs_noalloc): cursor[k + 1] starts
// at bucket k's first slot (set by the host's counts), and each pair takes the next slot of its
// bucket. The order within a bucket depends on the atomics' order, but every pair's moments are
// computed from that pair alone and stored at its own index, so the output does not.
__global__ void bucket_init_kernel(unsigned *cursor, unsigned s0, unsigned s1, unsigned s2, unsigned s3) {
    if (threadIdx.x == 0) {
        cursor[0] = s0; cursor[1] = s1; cursor[2] = s2; cursor[3] = s3;
    }
}

__global__ void bucket_scatter_kernel(const int *K, long long n, unsigned *cursor, int *order) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned slot = atomicAdd(&cursor[K[i] + 1], 1u);
    order[slot] = (int)i;
}

// FP32 far records: far points relative to the centroid, area, radius.
__global__ void far32_records(GeomDev g, float4 *rec) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nt) return;
    const double cx = g.cent[2 * t], cy = g.cent[2 * t + 1];
    const double *p = g.P3 + 6 * (size_t)t;
    rec[2 * t] = make_float4((float)(p[0] - cx), (float)(p[1] - cy), (float)(p[2] - cx), (float)(p[3] - cy));
    rec[2 * t + 1] = make_float4((float)(p[4] - cx), (float)(p[5] - cy), (float)g.area[t], (float)g.radius[t]);
}

}  // namespace pq

// ---- handle ---------------------------------------------------------------------------------

using namespace pq;

// ---- the allocation-free explicit pair entry (pq_pairs_noalloc) -------------------------------
// Its staging buffers live in this file-scope device array instead of device memory allocated per
// call: p, q, kind and the kind order (4 x n int32), d (n FP64), the moments (n x 6 FP64) and four
// bucket cursors, each part 256-byte aligned. Every device the library runs on holds its own copy
// until the process ends. Host code must not pass the array's name to a kernel or a copy: its
// device address comes from hipGetSymbolAddress, taken once per device (at pq_fill_create, with the
// device current) and kept in scratch_ptr.
constexpr long long PAIRS_NOALLOC_MAX = 1LL << 20;
constexpr size_t PAIRS_SCRATCH_BYTES = (size_t)PAIRS_NOALLOC_MAX * (4 * 4 + 8 + 6 * 8) + 8 * 256;
static __device__ __align__(256) unsigned char scratch[PAIRS_SCRATCH_BYTES];

namespace {

constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mutex;
void *scratch_ptr[SCRATCH_DEVICES] = {};

// The device address of `scratch` on the current device `device` (taken once per device).
void *scratch_setup(int device) {
    PQ_REQUIRE(device >= 0 && device < SCRATCH_DEVICES, "pq_pairs_noalloc: device ordinal beyond 63");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device]) {
        void *p = nullptr;
        PQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        PQ_REQUIRE(p, "pq_pairs_noalloc: no device address for the scratch array");
        scratch_ptr[device] = p;
    }
    return scratch_ptr[device];
}

size_t arr_bytes(long long n) { return (size_t)n * sizeof(double); }

void launch_far(pq_fill *f, int precision, bool scalar, const FarArgs &A, void *out, hipStream_t s) {
    dim3 grid((A.n + WAVE - 1) / WAVE, (A.rows + FAR_ROWS_PER_BLOCK - 1) / FAR_ROWS_PER_BLOCK);
    if (precision == 32 && A.mirror && A.r0 == 0 && A.rows == A.n) {
        const int nb = (A.n + TILE - 1) / TILE;
        dim3 tiles((unsigned)((long long)nb * (nb + 1) / 2));
        if (scalar) hipLaunchKernelGGL(far32_tile_kernel<true>, tiles, dim3(512), 0, s, f->g, A, (float *)out, nb);
        else hipLaunchKernelGGL(far32_tile_kernel<false>, tiles, dim3(512), 0, s, f->g, A, (float *)out, nb);
    } else if (precision == 32) {
        if (scalar) hipLaunchKernelGGL(far32_kernel<true>, grid, dim3(256), 0, s, f->g, A, (float *)out);
        else hipLaunchKernelGGL(far32_kernel<false>, grid, dim3(256), 0, s, f->g, A, (float *)out);
    } else if (A.mirror && A.r0 == 0 && A.rows == A.n && std::getenv("PCBQS_AMD_FAR64_STRIPS") == nullptr) {
        const int nb = (A.n + TILE - 1) / TILE;
        dim3 tiles((unsigned)((long long)nb * (nb + 1) / 2));
        if (scalar) hipLaunchKernelGGL(far64_tile_kernel<true>, tiles, dim3(512), 0, s, f->g, A, (double *)out, nb);
        else hipLaunchKernelGGL(far64_tile_kernel<false>, tiles, dim3(512), 0, s, f->g, A, (double *)out, nb);
    } else {
        if (scalar) hipLaunchKernelGGL(far64_kernel<true>, grid, dim3(256), 0, s, f->g, A, (double *)out);
        else hipLaunchKernelGGL(far64_kernel<false>, grid, dim3(256), 0, s, f->g, A, (double *)out);
    }
    PQ_HIP(hipGetLastError());
}

void launch_near(pq_fill *f, int precision, bool scalar, int which, int level, const FarArgs &A, void *out,
                 hipStream_t s) {
    dim3 grid(f->cus * 8), block(256);
    const unsigned long long *list = f->lists[which].as<unsigned long long>();
    const unsigned long long *count = f->counts.as<unsigned long long>() + which;
#define PQ_NEAR(T, S, LOC)                                                                           \
    hipLaunchKernelGGL((near_kernel<T, S, LOC>), grid, block, 0, s, f->g, list, count, A.lists.cap, level, \
                       A.r0, A.ld, A.mirror, (T *)out, A.plane, A.flag)
    if (precision == 32) {
        if (scalar) PQ_NEAR(float, true, true); else PQ_NEAR(float, false, true);
    } else {
        if (scalar) PQ_NEAR(double, true, false); else PQ_NEAR(double, false, false);
    }
#undef PQ_NEAR
    PQ_HIP(hipGetLastError());
}

void check_rule(double near, int outer, pq_fill *f) {
    PQ_REQUIRE(!std::isnan(near) && near >= 0, "near must be nonnegative (or infinity)");
    PQ_REQUIRE(outer >= 0 && outer <= 2, "outer_levels must be 0, 1 or 2");
    PQ_REQUIRE(f->g.nq[outer] == 6 * (1 << (2 * outer)), "outer level was not prepared");
}

FarArgs far_args(pq_fill *f, double near, int outer, int r0, int rows, int mirror, size_t plane, long long ld) {
    FarArgs A{};
    A.r0 = r0; A.rows = rows; A.n = f->nt; A.ld = ld;
    A.all_near = std::isinf(near) ? 1 : 0;
    A.near = A.all_near ? 0.0 : near;
    A.near32 = (float)A.near;
    A.outer = outer; A.mirror = mirror; A.plane = plane;
    A.lists.list[0] = f->lists[0].as<unsigned long long>();
    A.lists.list[1] = f->lists[1].as<unsigned long long>();
    A.lists.count = f->counts.as<unsigned long long>();
    A.lists.cap = f->cap;
    A.flag = f->flag.as<unsigned>();
    return A;
}

void reserve_lists(pq_fill *f, unsigned long long cap) {
    if (cap <= f->cap) return;
    f->lists[0].reserve(cap * 8);
    f->lists[1].reserve(cap * 8);
    f->cap = cap;
}

}  // namespace

extern "C" {

int pq_fill_create(const pq_geometry *g, int device, pq_fill **out) {
    return guard([&] {
        PQ_REQUIRE(g && out, "pq_fill_create: null argument");
        PQ_REQUIRE(g->nt > 0, "pq_fill_create: empty geometry");
        for (int k = 0; k < 3; ++k)
            PQ_REQUIRE(g->qp[k] && g->qw[k] && g->nq[k] > 0 && g->nq[k] <= 256,
                       "pq_fill_create: invalid observation quadrature");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
            throw Error(PQ_EDEVICE, "no HIP device available");
        PQ_REQUIRE(device >= 0 && device < ndev, "pq_fill_create: invalid device ordinal");
        DeviceScope scope(device);
        scratch_setup(device);                         // once per device, with the device current
        auto *f = new pq_fill();
        try {
            f->device = device;
            f->nt = g->nt;
            const long long nt = g->nt;
            hipDeviceProp_t prop;
            PQ_HIP(hipGetDeviceProperties(&prop, device));
            f->cus = prop.multiProcessorCount;
            // one buffer for every FP64 array
            const long long sizes[] = {nt * 6, nt * 3, nt * 6, nt * 6, nt, nt * 6, nt, nt * 2, nt,
                                       nt * 2LL * g->nq[0], g->nq[0], nt * 2LL * g->nq[1], g->nq[1],
                                       nt * 2LL * g->nq[2], g->nq[2]};
            const double *src[] = {g->V, g->L, g->tau, g->un, g->area, g->P3, g->z, g->cent, g->radius,
                                   g->qp[0], g->qw[0], g->qp[1], g->qw[1], g->qp[2], g->qw[2]};
            size_t total = 0;
            for (long long s : sizes) total += (arr_bytes(s) + 255) / 256 * 256;
            f->geom.reserve(total);
            const double *dst[15];
            size_t off = 0;
            for (int i = 0; i < 15; ++i) {
                PQ_REQUIRE(src[i], "pq_fill_create: missing geometry array");
                double *d = (double *)((char *)f->geom.p + off);
                PQ_HIP(hipMemcpy(d, src[i], arr_bytes(sizes[i]), hipMemcpyHostToDevice));
                dst[i] = d;
                off += (arr_bytes(sizes[i]) + 255) / 256 * 256;
            }
            GeomDev &G = f->g;
            G.nt = g->nt;
            G.V = dst[0]; G.L = dst[1]; G.tau = dst[2]; G.un = dst[3]; G.area = dst[4]; G.P3 = dst[5];
            G.z = dst[6]; G.cent = dst[7]; G.radius = dst[8];
            for (int k = 0; k < 3; ++k) { G.qp[k] = dst[9 + 2 * k]; G.qw[k] = dst[10 + 2 * k]; G.nq[k] = g->nq[k]; }
            f->far32.reserve((size_t)nt * 2 * sizeof(float4));
            G.far32 = f->far32.as<float4>();
            hipLaunchKernelGGL(far32_records, dim3((nt + 255) / 256), dim3(256), 0, 0, G, f->far32.as<float4>());
            PQ_HIP(hipGetLastError());
            // [0..3] the list counters, [4] the device fill's non-finite flag: one clear, one copy back
            f->counts.reserve(5 * sizeof(unsigned long long));
            PQ_HIP(hipHostMalloc((void **)&f->host_stat, 5 * sizeof(unsigned long long), hipHostMallocDefault));
            f->flag.reserve(sizeof(unsigned));
            PQ_HIP(hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking));
            PQ_HIP(hipStreamCreateWithFlags(&f->copy_stream, hipStreamNonBlocking));
            for (auto &e : f->ev) PQ_HIP(hipEventCreate(&e));
            PQ_HIP(hipDeviceSynchronize());
        } catch (...) {
            pq_fill_destroy(f);
            throw;
        }
        *out = f;
    });
}

int pq_fill_destroy(pq_fill *f) {
    if (!f) return PQ_OK;
    return guard([&] {
        DeviceScope scope(f->device);
        if (f->stream) (void)hipStreamSynchronize(f->stream);
        if (f->copy_stream) (void)hipStreamSynchronize(f->copy_stream);
        for (auto e : f->ev) if (e) (void)hipEventDestroy(e);
        if (f->stream) (void)hipStreamDestroy(f->stream);
        if (f->copy_stream) (void)hipStreamDestroy(f->copy_stream);
        f->geom.release(); f->far32.release(); f->lists[0].release(); f->lists[1].release();
        f->counts.release(); f->flag.release();
        if (f->host_stat) (void)hipHostFree(f->host_stat);
        delete f;
    });
}

int pq_fill_device(pq_fill *f, double near, int outer, int scalar_only, int precision, void *planes,
                   int64_t ld, int64_t *analytic, void *stream_) {
    return guard([&] {
        PQ_REQUIRE(f && planes, "pq_fill_device: null argument");
        PQ_REQUIRE(precision == 32 || precision == 64, "pq_fill_device: precision must be 32 or 64");
        check_rule(near, outer, f);
        DeviceScope scope(f->device);
        hipStream_t s = stream_ ? (hipStream_t)stream_ : f->stream;
        const long long n = f->nt;
        if (ld == 0) ld = n;
        PQ_REQUIRE(ld >= n, "pq_fill_device: leading dimension below the triangle count");
        const size_t plane = (size_t)n * ld;
        reserve_lists(f, std::max<unsigned long long>(1ull << 20, 64ull * n));
        for (int attempt = 0; attempt < 2; ++attempt) {
            FarArgs A = far_args(f, near, outer, 0, (int)n, 1, plane, ld);
            A.flag = (unsigned *)(f->counts.as<unsigned long long>() + 4);
            PQ_HIP(hipMemsetAsync(f->counts.p, 0, 5 * sizeof(unsigned long long), s));
            PQ_HIP(hipEventRecord(f->ev[0], s));
            launch_far(f, precision, scalar_only, A, planes, s);
            PQ_HIP(hipEventRecord(f->ev[1], s));
            launch_near(f, precision, scalar_only, 0, 0, A, planes, s);
            if (outer > 0) launch_near(f, precision, scalar_only, 1, outer, A, planes, s);
            PQ_HIP(hipEventRecord(f->ev[2], s));
            PQ_HIP(hipMemcpyAsync(f->host_stat, f->counts.p, 5 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
            PQ_HIP(hipStreamSynchronize(s));
            const unsigned long long *counts = f->host_stat;
            const unsigned flag = (unsigned)counts[4];
            const unsigned long long need = std::max(counts[0], counts[1]);
            if (need > f->cap) {             // lists overflowed: grow them and fill again
                PQ_REQUIRE(attempt == 0, "pq_fill_device: near lists overflowed twice");
                reserve_lists(f, need);
                continue;
            }
            float t0 = 0, t1 = 0;
            PQ_HIP(hipEventElapsedTime(&t0, f->ev[0], f->ev[1]));
            PQ_HIP(hipEventElapsedTime(&t1, f->ev[1], f->ev[2]));
            f->far_ms = t0; f->near_ms = t1;
            f->near_pairs[0] = (long long)counts[0]; f->near_pairs[1] = (long long)counts[1];
            if (analytic) *analytic = (int64_t)counts[2];
            if (flag) throw Error(PQ_ENONFINITE, "non-finite triangle-pair moments from fill backend");
            break;
        }
    });
}

int pq_fill_last_timing(pq_fill *f, double *far_ms, double *near_ms, int64_t near_pairs[2]) {
    return guard([&] {
        PQ_REQUIRE(f, "pq_fill_last_timing: null handle");
        if (far_ms) *far_ms = f->far_ms;
        if (near_ms) *near_ms = f->near_ms;
        if (near_pairs) { near_pairs[0] = f->near_pairs[0]; near_pairs[1] = f->near_pairs[1]; }
    });
}

// Row strips on the device, each copied into the caller's host rows (static_cupy.py:127-197).
int pq_fill_symmetric(pq_fill *f, double near, int outer, int scalar_only, double *const planes[4],
                      int64_t *analytic) {
    return guard([&] {
        PQ_REQUIRE(f && planes && planes[0], "pq_fill_symmetric: null argument");
        const int width = scalar_only ? 1 : 4;
        for (int k = 0; k < width; ++k) PQ_REQUIRE(planes[k], "pq_fill_symmetric: missing output plane");
        check_rule(near, outer, f);
        DeviceScope scope(f->device);
        const long long n = f->nt;
        const long long rows = std::max(1LL, std::min(n, (1LL << 22) / n));
        const size_t plane = (size_t)rows * n;
        DevBuf strip[2];
        strip[0].reserve(plane * width * sizeof(double));
        strip[1].reserve(plane * width * sizeof(double));
        reserve_lists(f, std::max<unsigned long long>(plane, 1024));
        hipEvent_t done[2], copied[2];
        for (int i = 0; i < 2; ++i) { PQ_HIP(hipEventCreate(&done[i])); PQ_HIP(hipEventCreate(&copied[i])); }
        DevBuf counts;
        const long long nstrips = (n + rows - 1) / rows;
        counts.reserve(4 * sizeof(unsigned long long) * nstrips);
        PQ_HIP(hipMemsetAsync(counts.p, 0, 4 * sizeof(unsigned long long) * nstrips, f->stream));
        PQ_HIP(hipMemsetAsync(f->flag.p, 0, sizeof(unsigned), f->stream));
        bool used[2] = {false, false};
        unsigned long long analytic_total = 0;
        // The strip near lists are shared by both buffers: serialize strips on one stream and
        // copy each finished strip on the copy stream while the next one is integrated.
        for (long long s = 0; s < nstrips; ++s) {
            const int b = (int)(s & 1);
            const long long r0 = s * rows, m = std::min(rows, n - r0);
            if (used[b]) PQ_HIP(hipStreamWaitEvent(f->stream, copied[b], 0));
            FarArgs A = far_args(f, near, outer, (int)r0, (int)m, 0, plane, n);
            A.lists.count = counts.as<unsigned long long>() + 4 * s;
            launch_far(f, 64, scalar_only, A, strip[b].p, f->stream);
            const unsigned long long *cnt = A.lists.count;
            // near kernels read this strip's counters
            {
                const int nq0 = f->g.nq[0];
                (void)nq0;
                dim3 grid(f->cus * 8), block(256);
                for (int which = 0; which < (outer > 0 ? 2 : 1); ++which) {
                    const int level = which ? outer : 0;
#define PQ_NEAR_STRIP(S)                                                                                    \
    hipLaunchKernelGGL((near_kernel<double, S, false>), grid, block, 0, f->stream, f->g,                      \
                       f->lists[which].as<unsigned long long>(), cnt + which, A.lists.cap, level, (int)r0,     \
                       (long long)n, 0, strip[b].as<double>(), plane, A.flag)
                    if (scalar_only) PQ_NEAR_STRIP(true); else PQ_NEAR_STRIP(false);
#undef PQ_NEAR_STRIP
                    PQ_HIP(hipGetLastError());
                }
            }
            PQ_HIP(hipEventRecord(done[b], f->stream));
            PQ_HIP(hipStreamWaitEvent(f->copy_stream, done[b], 0));
            for (int k = 0; k < width; ++k)
                PQ_HIP(hipMemcpyAsync(planes[k] + (size_t)r0 * n, strip[b].as<double>() + (size_t)k * plane,
                                      (size_t)m * n * sizeof(double), hipMemcpyDeviceToHost, f->copy_stream));
            PQ_HIP(hipEventRecord(copied[b], f->copy_stream));
            used[b] = true;
        }
        PQ_HIP(hipStreamSynchronize(f->stream));
        PQ_HIP(hipStreamSynchronize(f->copy_stream));
        std::vector<unsigned long long> host(4 * nstrips);
        PQ_HIP(hipMemcpy(host.data(), counts.p, host.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        unsigned flag = 0;
        PQ_HIP(hipMemcpy(&flag, f->flag.p, sizeof(flag), hipMemcpyDeviceToHost));
        for (long long s = 0; s < nstrips; ++s) analytic_total += host[4 * s + 2];
        for (int i = 0; i < 2; ++i) { (void)hipEventDestroy(done[i]); (void)hipEventDestroy(copied[i]); }
        strip[0].release(); strip[1].release(); counts.release();
        if (analytic) *analytic = (int64_t)analytic_total;
        if (flag) throw Error(PQ_ENONFINITE, "non-finite triangle-pair moments from fill backend");
    });
}

int pq_pairs(pq_fill *f, const int32_t *p, const int32_t *q, const double *d, const int32_t *kind, int64_t n,
             int scalar_only, double *out) {
    return guard([&] {
        PQ_REQUIRE(f && (n == 0 || (p && q && d && kind && out)), "pq_pairs: null argument");
        PQ_REQUIRE(n >= 0, "pq_pairs: negative pair count");
        if (n == 0) return;
        // validate on the host (static_kernels.hpp:156-159) and sort by kind so far and analytic
        // pairs do not share wavefronts (static_cupy.py:117-120)
        std::vector<int> order((size_t)n);
        std::vector<long long> bucket(4, 0);
        for (int64_t i = 0; i < n; ++i) {
            PQ_REQUIRE(p[i] >= 0 && p[i] < f->nt && q[i] >= 0 && q[i] < f->nt && kind[i] >= -1 && kind[i] <= 2 &&
                           std::isfinite(d[i]),
                       "pq_pairs: invalid triangle index, kind or offset");
            if (kind[i] >= 0)
                PQ_REQUIRE(f->g.nq[kind[i]] == 6 * (1 << (2 * kind[i])),
                           "pair kinds must select far (-1) or a prepared analytic level");
            ++bucket[kind[i] + 1];
        }
        std::vector<long long> start(4, 0);
        for (int k = 1; k < 4; ++k) start[k] = start[k - 1] + bucket[k - 1];
        for (int64_t i = 0; i < n; ++i) order[(size_t)start[kind[i] + 1]++] = (int)i;
        DeviceScope scope(f->device);
        const int width = scalar_only ? 1 : 6;
        DevBuf buf;
        const size_t bytes = (size_t)n * (4 * 4 + 8) + (size_t)n * width * 8;
        buf.reserve(bytes + 5 * 256);
        char *base = (char *)buf.p;
        auto carve = [&](size_t sz) { char *r = base; base += (sz + 255) / 256 * 256; return r; };
        int *dp = (int *)carve(n * 4), *dq = (int *)carve(n * 4), *dk = (int *)carve(n * 4),
            *dord = (int *)carve(n * 4);
        double *dd = (double *)carve(n * 8), *dout = (double *)carve(n * width * 8);
        hipStream_t s = f->stream;
        PQ_HIP(hipMemcpyAsync(dp, p, n * 4, hipMemcpyHostToDevice, s));
        PQ_HIP(hipMemcpyAsync(dq, q, n * 4, hipMemcpyHostToDevice, s));
        PQ_HIP(hipMemcpyAsync(dk, kind, n * 4, hipMemcpyHostToDevice, s));
        PQ_HIP(hipMemcpyAsync(dord, order.data(), n * 4, hipMemcpyHostToDevice, s));
        PQ_HIP(hipMemcpyAsync(dd, d, n * 8, hipMemcpyHostToDevice, s));
        dim3 grid((unsigned)((n + 127) / 128));
        if (scalar_only) hipLaunchKernelGGL(pairs_kernel<true>, grid, dim3(128), 0, s, f->g, dp, dq, dd, dk, dord, (long long)n, dout);
        else hipLaunchKernelGGL(pairs_kernel<false>, grid, dim3(128), 0, s, f->g, dp, dq, dd, dk, dord, (long long)n, dout);
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipMemcpyAsync(out, dout, (size_t)n * width * 8, hipMemcpyDeviceToHost, s));
        PQ_HIP(hipStreamSynchronize(s));
        buf.release();
    });
}

int pq_pairs_noalloc(pq_fill *f, const int32_t *p, const int32_t *q, const double *d, const int32_t *kind, int64_t n,
                     int scalar_only, double *out) {
    return guard([&] {
        PQ_REQUIRE(f && (n == 0 || (p && q && d && kind && out)), "pq_pairs_noalloc: null argument");
        PQ_REQUIRE(n >= 0, "pq_pairs_noalloc: negative pair count");
        PQ_REQUIRE(n <= PAIRS_NOALLOC_MAX, "pq_pairs_noalloc: at most 1048576 pairs per call");
        if (n == 0) return;
        long long bucket[4] = {0, 0, 0, 0};
        for (int64_t i = 0; i < n; ++i) {
            PQ_REQUIRE(p[i] >= 0 && p[i] < f->nt && q[i] >= 0 && q[i] < f->nt && kind[i] >= -1 && kind[i] <= 2 &&
                           std::isfinite(d[i]),
                       "pq_pairs_noalloc: invalid triangle index, kind or offset");
            if (kind[i] >= 0)
                PQ_REQUIRE(f->g.nq[kind[i]] == 6 * (1 << (2 * kind[i])),
                           "pair kinds must select far (-1) or a prepared analytic level");
            ++bucket[kind[i] + 1];
        }
        DeviceScope scope(f->device);
        char *base = (char *)scratch_setup(f->device);
        auto carve = [&](size_t sz) { char *r = base; base += (sz + 255) / 256 * 256; return r; };
        const int width = scalar_only ? 1 : 6;
        int *dp = (int *)carve(n * 4), *dq = (int *)carve(n * 4), *dk = (int *)carve(n * 4),
            *dord = (int *)carve(n * 4);
        double *dd = (double *)carve(n * 8), *dout = (double *)carve(n * 6 * 8);
        unsigned *cursor = (unsigned *)carve(4 * sizeof(unsigned));
        hipStream_t s = f->stream;
        PQ_HIP(hipMemcpyAsync(dp, p, n * 4, hipMemcpyHostToDevice, s));
        PQ_HIP(hipMemcpyAsync(dq, q, n * 4, hipMemcpyHostToDevice, s));
        PQ_HIP(hipMemcpyAsync(dk, kind, n * 4, hipMemcpyHostToDevice, s));
        PQ_HIP(hipMemcpyAsync(dd, d, n * 8, hipMemcpyHostToDevice, s));
        // far pairs first, then each analytic level (the buckets pq_pairs sorts on the host)
        const unsigned s1 = (unsigned)bucket[0], s2 = s1 + (unsigned)bucket[1], s3 = s2 + (unsigned)bucket[2];
        hipLaunchKernelGGL(bucket_init_kernel, dim3(1), dim3(64), 0, s, cursor, 0u, s1, s2, s3);
        const dim3 grid((unsigned)((n + 127) / 128));
        hipLaunchKernelGGL(bucket_scatter_kernel, grid, dim3(128), 0, s, dk, (long long)n, cursor, dord);
        if (scalar_only) hipLaunchKernelGGL(pairs_kernel<true>, grid, dim3(128), 0, s, f->g, dp, dq, dd, dk, dord, (long long)n, dout);
        else hipLaunchKernelGGL(pairs_kernel<false>, grid, dim3(128), 0, s, f->g, dp, dq, dd, dk, dord, (long long)n, dout);
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipMemcpyAsync(out, dout, (size_t)n * width * 8, hipMemcpyDeviceToHost, s));
        PQ_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
