// The direct port solve's factorisation: Z = Rl + j w Ll on the free loops, complex symmetric with
// symmetric positive definite real and imaginary parts, factored as Z = L D L^T (L unit lower, D
// diagonal) in FP32 without pivoting. For such matrices Gaussian elimination without pivoting is
// stable (growth factor below 3: N. J. Higham, "Factorizing complex symmetric matrices with positive
// definite real and imaginary parts", Math. Comp. 67 (1998)), so no row exchanges and no pivot
// search sit on the critical path, and the symmetric form halves the work of an LU. The port solve
// refines the FP32 solution in FP64 (the residual taken with the FP64 operator product).
//
// Right-looking, 64-wide panels. Per panel k:
//   ldl_diag_kernel   one workgroup factors the 64 x 64 diagonal block in registers and inverts its
//                     unit lower factor (the inverse serves the panel and the solves);
//   ldl_panel_kernel  W = A21 L11^-T and L21 = W D^-1 (MFMA, 64 x 64 tiles down the panel);
//   ldl_update_kernel A22 -= L21 W^T on the lower triangle's 64 x 64 tiles (MFMA FP32, the bulk of
//                     the flops).
// Complex products are four real MFMA products on split real / imaginary planes.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <limits>
#include <string>
#include <type_traits>
#include <vector>

#include "pq_common.h"
#include "ldl.h"

namespace pq {
namespace {

constexpr int LB = LDL_BLOCK;
// An entry of a one-launch sweep's solution not yet posted (a signalling NaN, which no arithmetic
// yields; see the persistent sweeps below).
constexpr unsigned long long LDL_EMPTY = 0x7FF4C0DE5EA7B10CULL;
constexpr int TP = 36;                      // LDS row pitch (floats) of a 32-wide K chunk
typedef float f32x4 __attribute__((ext_vector_type(4)));

// F = Rl + j w Ll, lower triangle: Fi = w Ll rounded to FP32, Fr = 0 (Rl's entries follow).
__global__ __launch_bounds__(256) void ldl_form_kernel(const float *__restrict__ Ll, long long ldl, int n, float omega,
                                                      float *__restrict__ Fr, float *__restrict__ Fi, long long ldf) {
    for (int row = blockIdx.y; row < n; row += gridDim.y)
        for (int c = blockIdx.x * 256 + threadIdx.x; c <= row; c += gridDim.x * 256) {
            Fr[(size_t)row * ldf + c] = 0.f;
            Fi[(size_t)row * ldf + c] = omega * Ll[(size_t)row * ldl + c];
        }
}

__global__ void ldl_form_r_kernel(const int *ptr, const int *col, const double *val, int n, float *Fr, long long ldf) {
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    for (int t = ptr[row]; t < ptr[row + 1]; ++t) {
        const int c = col[t];
        if (c <= row) Fr[(size_t)row * ldf + c] += (float)val[t];
    }
}

// 1 / (dr + j di) = conj(e) / (s |e|^2) with e = d / s, s = max(|dr|, |di|), so that the square
// stays in FP32's range. Hardware reciprocals (1 ulp): the factor is FP32 and the solve refines
// its result in FP64, so correctly rounded divisions (a dependent chain of about a dozen
// instructions each, on the critical path of every elimination step) buy nothing.
__device__ inline void cinv(float dr, float di, float &qr, float &qi) {
    const float rs = __builtin_amdgcn_rcpf(fmaxf(fabsf(dr), fabsf(di)));
    const float er = dr * rs, ei = di * rs;
    const float q = __builtin_amdgcn_rcpf(er * er + ei * ei) * rs;
    qr = er * q;
    qi = -ei * q;
}

// The b x b diagonal block at (k, k) factored in place (L below the diagonal, D on it), and the
// inverse X = L^-1 of its unit lower factor written to Lir / Lii (64 x 64, the identity beyond b).
// The block and X live in LDS (rows past b: the identity). Blocked by 16: for each 16 x 16
// diagonal sub-block S,
//   (a) S = L_S D_S L_S^T and X_S = L_S^-1 in 16 steps by one wave (lane shuffles, no barriers);
//   (b) below it, W = A_21 X_S^T and L_21 = W D_S^-1;
//   (c) the lower triangle below and right of it, A_22 -= L_21 W^T;
// then X's block rows below the first, X_I = -X_S,I (sum_{M < I} L_IM X_M), one block row at a
// time. A zero or non-finite pivot sets bad[matrix] (one flag per matrix of a batch).
__global__ __launch_bounds__(256) void ldl_diag_kernel(float *Fr, float *Fi, long long ldf, int k, int b, float *Lir,
                                                      float *Lii, int *bad, long long sF, long long sX) {
    __shared__ float Ar[LB][LB + 1], Ai[LB][LB + 1], Xr[LB][LB + 1], Xi[LB][LB + 1];
    __shared__ float Wr[LB - 16][17], Wi[LB - 16][17];
    __shared__ float Tr[16][LB + 1], Ti[16][LB + 1];
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // matrix blockIdx.z of a batch
    Lir += blockIdx.z * sX; Lii += blockIdx.z * sX;
    // on the factorisation's critical path beside the bulk trailing update: its waves issue first
    __builtin_amdgcn_s_setprio(3);
    const int t = threadIdx.x;
    for (int e = t; e < LB * LB; e += 256) {
        const int r = e >> 6, c = e & 63;
        float vr = 0.f, vi = 0.f;
        if (r < b && c <= r) {
            vr = Fr[(size_t)(k + r) * ldf + k + c];
            vi = Fi[(size_t)(k + r) * ldf + k + c];
        } else if (r >= b && c == r) {
            vr = 1.f;
        }
        Ar[r][c] = vr;
        Ai[r][c] = vi;
        Xr[r][c] = r == c ? 1.f : 0.f;
        Xi[r][c] = 0.f;
    }
    __syncthreads();
    for (int s0 = 0; s0 < LB; s0 += 16) {
        // (a) in the registers of wave 0: lane l holds row l >> 2 at columns 4 (l & 3) .. + 3 of the
        // sub-block and of X_S; step j exchanges the pivot, column j and row j of X_S by lane
        // shuffles, so the 16 steps need no workgroup barrier
        if (t < 64) {
            const int r = t >> 2, g = t & 3;
            float fr[4], fi[4], gr[4], gi[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                fr[q] = Ar[s0 + r][s0 + 4 * g + q];
                fi[q] = Ai[s0 + r][s0 + 4 * g + q];
                gr[q] = 4 * g + q == r ? 1.f : 0.f;
                gi[q] = 0.f;
            }
            for (int j = 0; j < 16; ++j) {
                const int jq = j & 3, jl = j >> 2;
                float vr = fr[0], vi = fi[0];                 // this lane's entry in column j's slot
#pragma unroll
                for (int q = 1; q < 4; ++q)
                    if (jq == q) { vr = fr[q]; vi = fi[q]; }
                const float dr = __shfl(vr, 4 * j + jl), di = __shfl(vi, 4 * j + jl);
                if (t == 0 && !(isfinite(dr) && isfinite(di) && (dr != 0.f || di != 0.f))) bad[blockIdx.z] = 1;
                const float cr = __shfl(vr, 4 * r + jl), ci = __shfl(vi, 4 * r + jl);   // A[r][j]
                float ur[4], ui[4], wr[4], wi[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    ur[q] = __shfl(vr, 4 * (4 * g + q) + jl);                        // A[c][j]
                    ui[q] = __shfl(vi, 4 * (4 * g + q) + jl);
                    wr[q] = __shfl(gr[q], 4 * j + g);                                 // X[j][c]
                    wi[q] = __shfl(gi[q], 4 * j + g);
                }
                if (r > j) {
                    float qr, qi;
                    cinv(dr, di, qr, qi);
                    const float lr = cr * qr - ci * qi, li = cr * qi + ci * qr;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int c = 4 * g + q;
                        if (c > j && c <= r) {
                            fr[q] -= lr * ur[q] - li * ui[q];
                            fi[q] -= lr * ui[q] + li * ur[q];
                        }
                        if (c <= j) {
                            gr[q] -= lr * wr[q] - li * wi[q];
                            gi[q] -= lr * wi[q] + li * wr[q];
                        }
                        if (c == j) { fr[q] = lr; fi[q] = li; }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                Ar[s0 + r][s0 + 4 * g + q] = fr[q];
                Ai[s0 + r][s0 + 4 * g + q] = fi[q];
                Xr[s0 + r][s0 + 4 * g + q] = gr[q];
                Xi[s0 + r][s0 + 4 * g + q] = gi[q];
            }
        }
        __syncthreads();
        const int rows = LB - s0 - 16;
        if (rows == 0) break;
        // (b) W[r][c] = sum_u A[s0 + 16 + r][s0 + u] X[s0 + c][s0 + u] (X_S holds zeros above its
        // diagonal, so every sum runs over all 16 terms, unrolled)
        for (int e = t; e < rows * 16; e += 256) {
            const int r = e >> 4, c = e & 15;
            float wr = 0.f, wi = 0.f;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const float ar = Ar[s0 + 16 + r][s0 + u], ai = Ai[s0 + 16 + r][s0 + u];
                const float xr = Xr[s0 + c][s0 + u], xi = Xi[s0 + c][s0 + u];
                wr += ar * xr - ai * xi;
                wi += ar * xi + ai * xr;
            }
            Wr[r][c] = wr;
            Wi[r][c] = wi;
        }
        __syncthreads();
        for (int e = t; e < rows * 16; e += 256) {
            const int r = e >> 4, c = e & 15;
            float qr, qi;
            cinv(Ar[s0 + c][s0 + c], Ai[s0 + c][s0 + c], qr, qi);
            const float wr = Wr[r][c], wi = Wi[r][c];
            Ar[s0 + 16 + r][s0 + c] = wr * qr - wi * qi;
            Ai[s0 + 16 + r][s0 + c] = wr * qi + wi * qr;
        }
        __syncthreads();
        // (c)
        for (int e = t; e < rows * rows; e += 256) {
            const int r = e / rows, c = e % rows;
            if (c > r) continue;
            float ur = 0.f, ui = 0.f;
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const float lr = Ar[s0 + 16 + r][s0 + u], li = Ai[s0 + 16 + r][s0 + u];
                const float wr = Wr[c][u], wi = Wi[c][u];
                ur += lr * wr - li * wi;
                ui += lr * wi + li * wr;
            }
            Ar[s0 + 16 + r][s0 + 16 + c] -= ur;
            Ai[s0 + 16 + r][s0 + 16 + c] -= ui;
        }
        __syncthreads();
    }
    // X's block rows I = 1, 2, 3: T = sum_{u < 16 I} L[16 I + r][u] X[u][c] (c < 16 I), then
    // X[16 I + r][c] = -sum_v X[16 I + r][16 I + v] T[v][c] (X is stored with its zeros, so the
    // sums run over whole 16-blocks, unrolled).
    for (int I = 16; I < LB; I += 16) {
        for (int e = t; e < 16 * I; e += 256) {
            const int r = e / I, c = e % I;
            float ur = 0.f, ui = 0.f;
            for (int u0 = 0; u0 < I; u0 += 16)
#pragma unroll
            for (int u = u0; u < u0 + 16; ++u) {
                const float lr = Ar[I + r][u], li = Ai[I + r][u];
                const float xr = Xr[u][c], xi = Xi[u][c];
                ur += lr * xr - li * xi;
                ui += lr * xi + li * xr;
            }
            Tr[r][c] = ur;
            Ti[r][c] = ui;
        }
        __syncthreads();
        for (int e = t; e < 16 * I; e += 256) {
            const int r = e / I, c = e % I;
            float ur = 0.f, ui = 0.f;
#pragma unroll
            for (int v = 0; v < 16; ++v) {
                const float xr = Xr[I + r][I + v], xi = Xi[I + r][I + v];
                ur += xr * Tr[v][c] - xi * Ti[v][c];
                ui += xr * Ti[v][c] + xi * Tr[v][c];
            }
            Xr[I + r][c] = -ur;
            Xi[I + r][c] = -ui;
        }
        __syncthreads();
    }
    for (int e = t; e < LB * LB; e += 256) {
        const int r = e >> 6, c = e & 63;
        if (r < b && c <= r) {
            Fr[(size_t)(k + r) * ldf + k + c] = Ar[r][c];
            Fi[(size_t)(k + r) * ldf + k + c] = Ai[r][c];
        }
        Lir[e] = Xr[r][c];
        Lii[e] = Xi[r][c];
    }
}

// One wave's 32 x 32 part of a 64 x 64 complex tile product P[i][c] = sum_t A[i][t] B[c][t]
// (t < kdim, a multiple of 32): waves 2 x 2 over the tile, each 2 x 2 MFMA 16x16x4 FP32 blocks for the real and the
// imaginary part. The K dimension is staged through LDS in two chunks of 32; in MFMA step s of a
// chunk the four lane groups g = lane >> 4 supply k = 8 g + s to A and B alike, so each lane reads
// its eight k values of a row as two 16-byte LDS loads. Rows at or past arows / brows read zero.
struct TileAcc {
    f32x4 r[2][2], i[2][2];
};

__device__ __attribute__((always_inline)) inline void tile_abt(const float *__restrict__ Ar, const float *__restrict__ Ai, long long lda, int arows,
                                const float *__restrict__ Br, const float *__restrict__ Bi, long long ldb, int brows,
                                int kdim, float (*sA)[LB][TP], float (*sB)[LB][TP], TileAcc &acc) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1, g = lane >> 4, l16 = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            acc.r[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            acc.i[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    // Chunk kc + 32's global loads are issued before chunk kc's MFMAs (registers, then LDS after the
    // chunk's barrier), so their latency overlaps the products.
    float4 pa[2][2], pb[2][2];              // [e iteration][plane]
    auto fetch = [&](int kc) {
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int e = tid + it * 256;   // LB * 8 = 512 float4 slots per plane
            const int r = e >> 3, q = (e & 7) * 4, ra = min(r, arows - 1), rb = min(r, brows - 1);
            // rows past the end load the last valid row (arows, brows >= 1) and are zeroed by value
            pa[it][0] = *reinterpret_cast<const float4 *>(Ar + (size_t)ra * lda + kc + q);
            pa[it][1] = *reinterpret_cast<const float4 *>(Ai + (size_t)ra * lda + kc + q);
            pb[it][0] = *reinterpret_cast<const float4 *>(Br + (size_t)rb * ldb + kc + q);
            pb[it][1] = *reinterpret_cast<const float4 *>(Bi + (size_t)rb * ldb + kc + q);
            if (r >= arows) pa[it][0] = pa[it][1] = zero;
            if (r >= brows) pb[it][0] = pb[it][1] = zero;
        }
    };
    auto stage = [&]() {
#pragma unroll
        for (int it = 0; it < 2; ++it) {
            const int e = tid + it * 256, r = e >> 3, q = (e & 7) * 4;
            *reinterpret_cast<float4 *>(&sA[0][r][q]) = pa[it][0];
            *reinterpret_cast<float4 *>(&sA[1][r][q]) = pa[it][1];
            *reinterpret_cast<float4 *>(&sB[0][r][q]) = pb[it][0];
            *reinterpret_cast<float4 *>(&sB[1][r][q]) = pb[it][1];
        }
    };
    fetch(0);
    for (int kc = 0; kc < kdim; kc += 32) {
        stage();
        __syncthreads();
        if (kc + 32 < kdim) fetch(kc + 32);
        float fa[2][2][8], fb[2][2][8];     // [plane][row or column group][s]
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float4 *ra = reinterpret_cast<const float4 *>(&sA[p][wr * 32 + h * 16 + l16][8 * g]);
                const float4 *rb = reinterpret_cast<const float4 *>(&sB[p][wc * 32 + h * 16 + l16][8 * g]);
                const float4 a0 = ra[0], a1 = ra[1], b0 = rb[0], b1 = rb[1];
                fa[p][h][0] = a0.x; fa[p][h][1] = a0.y; fa[p][h][2] = a0.z; fa[p][h][3] = a0.w;
                fa[p][h][4] = a1.x; fa[p][h][5] = a1.y; fa[p][h][6] = a1.z; fa[p][h][7] = a1.w;
                fb[p][h][0] = b0.x; fb[p][h][1] = b0.y; fb[p][h][2] = b0.z; fb[p][h][3] = b0.w;
                fb[p][h][4] = b1.x; fb[p][h][5] = b1.y; fb[p][h][6] = b1.z; fb[p][h][7] = b1.w;
            }
#pragma unroll
        for (int s = 0; s < 8; ++s)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    acc.r[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[0][a][s], fb[0][b][s], acc.r[a][b], 0, 0, 0);
                    acc.r[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(-fa[1][a][s], fb[1][b][s], acc.r[a][b], 0, 0, 0);
                    acc.i[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[0][a][s], fb[1][b][s], acc.i[a][b], 0, 0, 0);
                    acc.i[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[1][a][s], fb[0][b][s], acc.i[a][b], 0, 0, 0);
                }
        __syncthreads();
    }
}

// Below the diagonal block of panel k, rows from I0 = k + 64 + 64 blockIdx.x: W = A21 X^T with
// X = L11^-1 (so W = A21 L11^-T = L21 D), then L21 = W D^-1 in place of A21.
__global__ __launch_bounds__(256) void ldl_panel_kernel(float *Fr, float *Fi, long long ldf, int n, int k,
                                                       const float *Lir, const float *Lii, float *Wr, float *Wi,
                                                       int ldw, long long sF, long long sX, long long sW) {
    __shared__ float sA[2][LB][TP], sB[2][LB][TP];
    __shared__ float qr[LB], qi[LB];
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // matrix blockIdx.z of a batch
    Lir += blockIdx.z * sX; Lii += blockIdx.z * sX;
    Wr += blockIdx.z * sW; Wi += blockIdx.z * sW;
    __builtin_amdgcn_s_setprio(3);                    // on the critical path, as the diagonal block
    const int I0 = k + LB + LB * blockIdx.x;
    if (threadIdx.x < LB) {
        const size_t dg = (size_t)(k + threadIdx.x) * ldf + k + threadIdx.x;
        cinv(Fr[dg], Fi[dg], qr[threadIdx.x], qi[threadIdx.x]);
    }
    TileAcc acc;
    tile_abt(Fr + (size_t)I0 * ldf + k, Fi + (size_t)I0 * ldf + k, ldf, min(LB, n - I0), Lir, Lii, LB, LB, LB, sA, sB,
             acc);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1, g = lane >> 4, l16 = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = I0 + wr * 32 + a * 16 + 4 * g + q, col = wc * 32 + b * 16 + l16;
                if (row >= n) continue;
                const float vr = acc.r[a][b][q], vi = acc.i[a][b][q];
                Wr[(size_t)row * ldw + col] = vr;
                Wi[(size_t)row * ldw + col] = vi;
                Fr[(size_t)row * ldf + k + col] = vr * qr[col] - vi * qi[col];
                Fi[(size_t)row * ldf + k + col] = vr * qi[col] + vi * qr[col];
            }
}

// A22 -= L21 W^T on the lower triangle from row / column m0, one 64 x 64 tile (ti, tj), tj <= ti,
// per workgroup: L21 = F[rows][k0 .. k0 + kdim), W (n, ldw) rows m0 .. (kdim 64: one panel, 128:
// two). part 1: the first tc tile columns (tj < tc: the next panels' columns); part 2: the rest
// (tj >= tc; blockIdx.x indexes the triangle over ti - tc, tj - tc).
__global__ __launch_bounds__(256) void ldl_update_kernel(float *Fr, float *Fi, long long ldf, int n, int k0, int kdim,
                                                        int m0, const float *Wr, const float *Wi, int ldw, int part,
                                                        int tc, long long sF, long long sW) {
    __shared__ float sA[2][LB][TP], sB[2][LB][TP];
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // matrix blockIdx.z of a batch
    Wr += blockIdx.z * sW; Wi += blockIdx.z * sW;
    const int T = (n - m0 + LB - 1) / LB;
    // Workgroups go to the 8 XCDs in turn (workgroup w to XCD w mod 8), each with its own L2: give
    // each XCD one contiguous run of tiles, so that the tiles it holds at once share their L21 rows
    // and W blocks in its L2 instead of every XCD fetching every operand.
    int id, ti, tj;
    {
        const unsigned nw = gridDim.x, q = nw / 8, r = nw % 8, x = blockIdx.x % 8, j = blockIdx.x / 8;
        id = (int)(x * q + min(x, r) + j);
    }
    if (part == 1) {
        tj = 0;
        while (id >= T - tj) { id -= T - tj; ++tj; }
        ti = tj + id;
    } else {
        // The triangle over ti - tc >= tj - tc in bands of G tile rows, a band's tiles column by
        // column: consecutive tiles share their W block, and a band's G row blocks of L21 stay in
        // the L2 while the band's columns pass.
        constexpr int G = 16;
        const int Tt = T - tc;
        int b = (int)((sqrtf(8.f * (float)id + 1.f) - 1.f) * 0.5f) / G;
        while (b > 0 && (b * G) * (b * G + 1) / 2 > id) --b;
        while ((b + 1) * G < Tt && ((b + 1) * G) * ((b + 1) * G + 1) / 2 <= id) ++b;
        const int r0 = b * G, g = min(G, Tt - r0);
        int o = id - r0 * (r0 + 1) / 2;
        if (o < r0 * g) {
            ti = r0 + o % g;
            tj = o / g;
        } else {
            o -= r0 * g;
            int c = 0;
            while (o >= g - c) { o -= g - c; ++c; }
            ti = r0 + c + o;
            tj = r0 + c;
        }
        ti += tc;
        tj += tc;
    }
    const int I0 = m0 + LB * ti, C0 = m0 + LB * tj;
    TileAcc acc;
    tile_abt(Fr + (size_t)I0 * ldf + k0, Fi + (size_t)I0 * ldf + k0, ldf, min(LB, n - I0), Wr + (size_t)C0 * ldw,
             Wi + (size_t)C0 * ldw, ldw, min(LB, n - C0), kdim, sA, sB, acc);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, wr = w >> 1, wc = w & 1, g = lane >> 4, l16 = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = I0 + wr * 32 + a * 16 + 4 * g + q, col = C0 + wc * 32 + b * 16 + l16;
                if (row >= n || col > row) continue;
                Fr[(size_t)row * ldf + col] -= acc.r[a][b][q];
                Fi[(size_t)row * ldf + col] -= acc.i[a][b][q];
            }
}

// ---- solves: L y = b, z = D^-1 y, L^T x = z, P complex columns (n, 2P) FP64 ----

// Forward step over block k0 .. k0 + bk. Every workgroup stages the block's stored inverse X_k in
// LDS and forms y_k = X_k b_k (workgroup 0 writes it to Ys). Then a workgroup takes 64 rows
// r >= k0 + bk (lane = row) and its four waves the four 16-column quarters of L[r, block]; the
// waves' sums of L[r, t] y_k[t] meet in LDS and wave 0 subtracts them from Y in wave order.
template <int P>
__global__ __launch_bounds__(256) void ldl_forward_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                         long long ldf, const float *__restrict__ Xr,
                                                         const float *__restrict__ Xi, int n, int k0, int bk, double *Y,
                                                         double *Ys, long long sF, long long sX, long long sV) {
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // system blockIdx.z of a batch
    Xr += blockIdx.z * sX; Xi += blockIdx.z * sX;
    Y += blockIdx.z * sV; Ys += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ float xs[2][LB][LB + 1];
    __shared__ double bs[LB][NR], ys[LB][NR], part[3][LB][NR];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // this thread's quarter of its row of L and (wave 0) the row of Y, loaded first: they do not
    // depend on y_k, so their latency overlaps the staging and the small product
    const int r = k0 + bk + blockIdx.x * LB + lane;   // (rows below the block exist only when bk = 64)
    float lre[16], lim[16];
    double yr[P], yi[P];
    if (r < n) {
        const float4 *lr4 = reinterpret_cast<const float4 *>(Fr + (size_t)r * ldf + k0 + 16 * wv);
        const float4 *li4 = reinterpret_cast<const float4 *>(Fi + (size_t)r * ldf + k0 + 16 * wv);
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) {
            const float4 a = lr4[t4], b = li4[t4];
            lre[4 * t4] = a.x; lre[4 * t4 + 1] = a.y; lre[4 * t4 + 2] = a.z; lre[4 * t4 + 3] = a.w;
            lim[4 * t4] = b.x; lim[4 * t4 + 1] = b.y; lim[4 * t4 + 2] = b.z; lim[4 * t4 + 3] = b.w;
        }
        if (wv == 0)
#pragma unroll
            for (int p = 0; p < P; ++p) { yr[p] = Y[(size_t)r * NR + p]; yi[p] = Y[(size_t)r * NR + P + p]; }
    }
    for (int e = tid; e < LB * LB; e += 256) {
        xs[0][e >> 6][e & 63] = Xr[e];
        xs[1][e >> 6][e & 63] = Xi[e];
    }
    for (int e = tid; e < LB * NR; e += 256) {
        const int r = e / NR, q = e % NR;
        bs[r][q] = r < bk ? Y[(size_t)(k0 + r) * NR + q] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < LB * P; e += 256) {
        const int i = e / P, p = e % P;
        double sr = 0, si = 0;
#pragma unroll 16
        for (int t = 0; t < LB; ++t) {                // X_k is stored with its zeros above the diagonal
            const double lr = xs[0][i][t], li = xs[1][i][t];
            sr += lr * bs[t][p] - li * bs[t][P + p];
            si += lr * bs[t][P + p] + li * bs[t][p];
        }
        ys[i][p] = sr;
        ys[i][P + p] = si;
        if (blockIdx.x == 0 && i < bk) {
            Ys[(size_t)(k0 + i) * NR + p] = sr;
            Ys[(size_t)(k0 + i) * NR + P + p] = si;
        }
    }
    __syncthreads();
    double ar[P], ai[P];
#pragma unroll
    for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
    if (r < n) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int t = 16 * wv + u;
            const double xr = lre[u], xi = lim[u];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                ar[p] += xr * ys[t][p] - xi * ys[t][P + p];
                ai[p] += xr * ys[t][P + p] + xi * ys[t][p];
            }
        }
    }
    if (wv > 0)
#pragma unroll
        for (int p = 0; p < P; ++p) { part[wv - 1][lane][p] = ar[p]; part[wv - 1][lane][P + p] = ai[p]; }
    __syncthreads();
    if (wv == 0 && r < n) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int p = 0; p < P; ++p) { ar[p] += part[q][lane][p]; ai[p] += part[q][lane][P + p]; }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            Y[(size_t)r * NR + p] = yr[p] - ar[p];
            Y[(size_t)r * NR + P + p] = yi[p] - ai[p];
        }
    }
}

// z = D^-1 y (FP64 complex division by the FP32 pivots); with `mark`, its entries (the same
// shape) are marked empty for a one-launch backward sweep.
__global__ void ldl_scale_kernel(const float *Fr, const float *Fi, long long ldf, int n, int P, const double *Ys,
                                 double *Z, long long sF, long long sV, double *mark) {
    const int NR = 2 * P;
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;
    Ys += blockIdx.z * sV; Z += blockIdx.z * sV;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * P) return;
    const int r = (int)(e / P), p = (int)(e % P);
    if (mark) {
        mark += blockIdx.z * sV;
        mark[(size_t)r * NR + p] = mark[(size_t)r * NR + P + p] = __longlong_as_double((long long)LDL_EMPTY);
    }
    const double dr = Fr[(size_t)r * ldf + r], di = Fi[(size_t)r * ldf + r];
    const double yr = Ys[(size_t)r * NR + p], yi = Ys[(size_t)r * NR + P + p];
    const double den = dr * dr + di * di;
    Z[(size_t)r * NR + p] = (yr * dr + yi * di) / den;
    Z[(size_t)r * NR + P + p] = (yi * dr - yr * di) / den;
}

// Backward step over block k0 .. k0 + bk: every workgroup stages X_k in LDS and forms
// x_k = X_k^T z_k (workgroup 0 writes it to Xo). Then a workgroup takes 64 rows r < k0 (lane = r)
// and its four waves the four 16-row quarters of the block's rows: L[k0 + t][r] read across lanes,
// the waves' sums of L[k0 + t][r] x_k[t] meet in LDS and wave 0 subtracts them from Z.
template <int P>
__global__ __launch_bounds__(256) void ldl_backward_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                          long long ldf, const float *__restrict__ Xr,
                                                          const float *__restrict__ Xi, int k0, int bk, double *Z,
                                                          double *Xo, long long sF, long long sX, long long sV,
                                                          int r0) {
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // system blockIdx.z of a batch
    Xr += blockIdx.z * sX; Xi += blockIdx.z * sX;
    Z += blockIdx.z * sV; Xo += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ float xsm[2][LB][LB + 1];
    __shared__ double zs[LB][NR], xs[LB][NR], part[3][LB][NR];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    // this thread's 16 entries of column r of the block's rows and (wave 0) the row of Z, loaded
    // first (they do not depend on x_k); columns from r0 (0, or the band's first below a band)
    const int r = r0 + blockIdx.x * LB + lane, t0 = 16 * wv, t1 = min(bk, t0 + 16);
    float lre[16], lim[16];
    double zr[P], zi[P];
    if (r < k0) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const bool in = t0 + u < t1;
            lre[u] = in ? Fr[(size_t)(k0 + t0 + u) * ldf + r] : 0.f;
            lim[u] = in ? Fi[(size_t)(k0 + t0 + u) * ldf + r] : 0.f;
        }
        if (wv == 0)
#pragma unroll
            for (int p = 0; p < P; ++p) { zr[p] = Z[(size_t)r * NR + p]; zi[p] = Z[(size_t)r * NR + P + p]; }
    }
    for (int e = tid; e < LB * LB; e += 256) {
        xsm[0][e >> 6][e & 63] = Xr[e];
        xsm[1][e >> 6][e & 63] = Xi[e];
    }
    for (int e = tid; e < LB * NR; e += 256) {
        const int r = e / NR, q = e % NR;
        zs[r][q] = r < bk ? Z[(size_t)(k0 + r) * NR + q] : 0.0;
    }
    __syncthreads();
    for (int e = tid; e < LB * P; e += 256) {
        const int c = e % LB, p = e / LB;
        double sr = 0, si = 0;
#pragma unroll 16
        for (int t = 0; t < LB; ++t) {                // zeros above X_k's diagonal and past bk in zs
            const double lr = xsm[0][t][c], li = xsm[1][t][c];
            sr += lr * zs[t][p] - li * zs[t][P + p];
            si += lr * zs[t][P + p] + li * zs[t][p];
        }
        xs[c][p] = sr;
        xs[c][P + p] = si;
        if (blockIdx.x == 0 && c < bk) {
            Xo[(size_t)(k0 + c) * NR + p] = sr;
            Xo[(size_t)(k0 + c) * NR + P + p] = si;
        }
    }
    __syncthreads();
    double ar[P], ai[P];
#pragma unroll
    for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
    if (r < k0) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int t = t0 + u;                     // entries past bk are zero, and xs rows past bk too
            const double xr = lre[u], xi = lim[u];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                ar[p] += xr * xs[t][p] - xi * xs[t][P + p];
                ai[p] += xr * xs[t][P + p] + xi * xs[t][p];
            }
        }
    }
    if (wv > 0)
#pragma unroll
        for (int p = 0; p < P; ++p) { part[wv - 1][lane][p] = ar[p]; part[wv - 1][lane][P + p] = ai[p]; }
    __syncthreads();
    if (wv == 0 && r < k0) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int p = 0; p < P; ++p) { ar[p] += part[q][lane][p]; ai[p] += part[q][lane][P + p]; }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            Z[(size_t)r * NR + p] = zr[p] - ar[p];
            Z[(size_t)r * NR + P + p] = zi[p] - ai[p];
        }
    }
}

// ---- persistent sweeps: one launch per sweep ----
//
// The per-block sweeps above pay one dependent launch per 64-row block. Here one launch does a
// whole sweep. Workgroup i owns row block i in the forward sweep (block nb - 1 - i in the backward
// one, so that workgroups start in the order their blocks are solved). It stages its block's
// stored inverse, then takes the earlier blocks' solutions as they appear and subtracts their
// products with its rows in the order and grouping of the per-block kernels, so that both give the
// same values. The solution entries are their own flags: the launch before marks them empty (a
// signalling NaN, which no arithmetic yields), a block posts each entry with an agent-scope store
// and a reader polls them with agent-scope loads until none is marked, so no fence or separate
// flag sits on the chain. A workgroup waits only for blocks of lower workgroup index, which are
// dispatched before it, so every wait ends; each wait is bounded all the same, and entries still
// marked when it runs out are read as the NaN they hold, which the refinement reports as a
// non-finite residual instead of hanging.
constexpr int LDL_SPIN = 1 << 20;                    // polls, each ~1 us, before a wait gives up

__global__ void ldl_mark_empty_kernel(double *V, size_t count) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) V[e] = __longlong_as_double((long long)LDL_EMPTY);
}

__device__ __forceinline__ void ldl_post(double *p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long *>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Waits until the 16 rows of the 64 x NR entries at V that the calling wave uses are posted
// (rows 16 wv .. 16 wv + 15; rows from `rows` on read as zero); lane l keeps entries l + 64 c of
// them. Each wave polls on its own (loads from clamped addresses, none skipped, so they are all in
// flight together), so no workgroup barrier sits in the wait.
template <int NR>
__device__ __forceinline__ void ldl_take_rows(const double *V, int rows, unsigned long long (&v)[(16 * NR + 63) / 64]) {
    constexpr int CW = (16 * NR + 63) / 64;
    const int lane = threadIdx.x & 63, r0 = 16 * (threadIdx.x >> 6);
    const unsigned long long *src = reinterpret_cast<const unsigned long long *>(V);
    for (int i = 0;; ++i) {
        bool all = true;
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            const int e = lane + 64 * c;
            const bool in = e < 16 * NR && r0 + e / NR < rows;
            const unsigned long long x =
                __hip_atomic_load(src + (in ? (size_t)r0 * NR + e : 0), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v[c] = in ? x : 0ull;
            all = all && v[c] != LDL_EMPTY;
        }
        if (__all(all) || i >= LDL_SPIN) break;
        __builtin_amdgcn_s_sleep(1);
    }
}

// Entry e (local row * NR + column) of the wave's 16 rows, read from the lane that keeps it.
template <int NR>
__device__ __forceinline__ double ldl_entry(const unsigned long long (&v)[(16 * NR + 63) / 64], int e) {
    const unsigned long long x = v[e / 64];
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)x, e % 64);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(x >> 32), e % 64);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Forward: row block rb's rows start from B; for each earlier block k whose per-block launch
// reaches them (all, or those within the band), b_r -= L[r, block k] y_k (lane = row, wave = a
// 16-column quarter, the quarters summed in wave order), then y_rb = X_rb b_rb posted to Ys.
template <int P>
__global__ __launch_bounds__(256) void ldl_fwd_persist_kernel(const float *__restrict__ Fr,
                                                             const float *__restrict__ Fi, long long ldf,
                                                             const float *__restrict__ Xr,
                                                             const float *__restrict__ Xi, int n, int band,
                                                             const double *B, double *Ys, long long sF,
                                                             long long sX, long long sV) {
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // system blockIdx.z of a batch
    Xr += blockIdx.z * sX; Xi += blockIdx.z * sX;
    B += blockIdx.z * sV; Ys += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ float xs[2][LB][LB + 1];
    __shared__ double ys[LB][NR], part[2][3][LB][NR];  // partials by parity of the block taken
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int rb = blockIdx.x, R0 = rb * LB, bk = min(LB, n - R0), r = R0 + lane;
    int taken = 0;
    double yr[P], yi[P];
    if (wv == 0 && r < n)
#pragma unroll
        for (int p = 0; p < P; ++p) { yr[p] = B[(size_t)r * NR + p]; yi[p] = B[(size_t)r * NR + P + p]; }
    const float *xr = Xr + (size_t)rb * LB * LB, *xi = Xi + (size_t)rb * LB * LB;
    for (int e = tid; e < LB * LB; e += 256) {
        xs[0][e >> 6][e & 63] = xr[e];
        xs[1][e >> 6][e & 63] = xi[e];
    }
    for (int k = 0; k < rb; ++k) {
        const int k0 = k * LB;                        // a full block: it has rows after it
        const int rows = band > 0 ? min(n - k0 - LB, band) : n - k0 - LB;
        if (rb - k - 1 >= max(1, (rows + LB - 1) / LB)) continue;
        float lre[16] = {}, lim[16] = {};             // loaded before the wait: they do not depend on y_k
        if (r < n) {
            const float4 *lr4 = reinterpret_cast<const float4 *>(Fr + (size_t)r * ldf + k0 + 16 * wv);
            const float4 *li4 = reinterpret_cast<const float4 *>(Fi + (size_t)r * ldf + k0 + 16 * wv);
#pragma unroll
            for (int t4 = 0; t4 < 4; ++t4) {
                const float4 a = lr4[t4], b = li4[t4];
                lre[4 * t4] = a.x; lre[4 * t4 + 1] = a.y; lre[4 * t4 + 2] = a.z; lre[4 * t4 + 3] = a.w;
                lim[4 * t4] = b.x; lim[4 * t4 + 1] = b.y; lim[4 * t4 + 2] = b.z; lim[4 * t4 + 3] = b.w;
            }
        }
        unsigned long long v[(16 * NR + 63) / 64];
        ldl_take_rows<NR>(Ys + (size_t)k0 * NR, LB, v);
        double ar[P], ai[P];
#pragma unroll
        for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {                // y_k[16 wv + u], rows past n with zero L
            const double lr = lre[u], li = lim[u];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const double yre = ldl_entry<NR>(v, u * NR + p), yim = ldl_entry<NR>(v, u * NR + P + p);
                ar[p] += lr * yre - li * yim;
                ai[p] += lr * yim + li * yre;
            }
        }
        double(*pt)[LB][NR] = part[taken & 1];
        if (wv > 0)
#pragma unroll
            for (int p = 0; p < P; ++p) { pt[wv - 1][lane][p] = ar[p]; pt[wv - 1][lane][P + p] = ai[p]; }
        __syncthreads();
        if (wv == 0 && r < n) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int p = 0; p < P; ++p) { ar[p] += pt[q][lane][p]; ai[p] += pt[q][lane][P + p]; }
#pragma unroll
            for (int p = 0; p < P; ++p) { yr[p] = yr[p] - ar[p]; yi[p] = yi[p] - ai[p]; }
        }
        ++taken;
    }
    if (wv == 0)                                      // b_rb, zero past the last row
#pragma unroll
        for (int p = 0; p < P; ++p) {
            ys[lane][p] = lane < bk ? yr[p] : 0.0;
            ys[lane][P + p] = lane < bk ? yi[p] : 0.0;
        }
    __syncthreads();
    for (int e = tid; e < LB * P; e += 256) {
        const int i = e / P, p = e % P;
        double sr = 0, si = 0;
#pragma unroll 16
        for (int t = 0; t < LB; ++t) {
            const double lr = xs[0][i][t], li = xs[1][i][t];
            sr += lr * ys[t][p] - li * ys[t][P + p];
            si += lr * ys[t][P + p] + li * ys[t][p];
        }
        if (i < bk) {
            ldl_post(Ys + (size_t)(R0 + i) * NR + p, sr);
            ldl_post(Ys + (size_t)(R0 + i) * NR + P + p, si);
        }
    }
}

// Backward: column block cb = nb - 1 - blockIdx.x starts from z; for each later block j whose
// per-block launch reaches its columns (all, or those within the band), z_c -= L[block j, c]^T x_j
// (lane = column, wave = a 16-row quarter of block j, summed in wave order), then
// x_cb = X_cb^T z_cb posted to Xo.
template <int P>
__global__ __launch_bounds__(256) void ldl_bwd_persist_kernel(const float *__restrict__ Fr,
                                                             const float *__restrict__ Fi, long long ldf,
                                                             const float *__restrict__ Xr,
                                                             const float *__restrict__ Xi, int n, int band,
                                                             const double *Z, double *Xo, long long sF,
                                                             long long sX, long long sV) {
    const int nb = gridDim.x;
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // system blockIdx.z of a batch
    Xr += blockIdx.z * sX; Xi += blockIdx.z * sX;
    Z += blockIdx.z * sV; Xo += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ float xsm[2][LB][LB + 1];
    __shared__ double xs[LB][NR], part[2][3][LB][NR];  // partials by parity of the block taken
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int cb = nb - 1 - (int)blockIdx.x, C0 = cb * LB, bc = min(LB, n - C0), c = C0 + lane;
    int taken = 0;
    double zr[P], zi[P];
    if (wv == 0 && c < n)
#pragma unroll
        for (int p = 0; p < P; ++p) { zr[p] = Z[(size_t)c * NR + p]; zi[p] = Z[(size_t)c * NR + P + p]; }
    const float *xr = Xr + (size_t)cb * LB * LB, *xi = Xi + (size_t)cb * LB * LB;
    for (int e = tid; e < LB * LB; e += 256) {
        xsm[0][e >> 6][e & 63] = xr[e];
        xsm[1][e >> 6][e & 63] = xi[e];
    }
    for (int j = nb - 1; j > cb; --j) {
        const int k0 = j * LB, bk = min(LB, n - k0);
        const int r0 = band > 0 ? max(0, (k0 - band) / LB * LB) : 0;
        if (C0 < r0) continue;
        const int t0 = 16 * wv;
        float lre[16], lim[16];                       // c < k0: every column of this block is below block j
#pragma unroll
        for (int u = 0; u < 16; ++u) {                // rows past bk: loaded from the last row, then zeroed
            const size_t at = (size_t)(k0 + min(t0 + u, bk - 1)) * ldf + c;
            const float a = Fr[at], b = Fi[at];
            lre[u] = t0 + u < bk ? a : 0.f;
            lim[u] = t0 + u < bk ? b : 0.f;
        }
        unsigned long long v[(16 * NR + 63) / 64];
        ldl_take_rows<NR>(Xo + (size_t)k0 * NR, bk, v);
        double ar[P], ai[P];
#pragma unroll
        for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
#pragma unroll
        for (int u = 0; u < 16; ++u) {                // x_j[t0 + u], zero past bk
            const double lr = lre[u], li = lim[u];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const double xre = ldl_entry<NR>(v, u * NR + p), xim = ldl_entry<NR>(v, u * NR + P + p);
                ar[p] += lr * xre - li * xim;
                ai[p] += lr * xim + li * xre;
            }
        }
        double(*pt)[LB][NR] = part[taken & 1];
        if (wv > 0)
#pragma unroll
            for (int p = 0; p < P; ++p) { pt[wv - 1][lane][p] = ar[p]; pt[wv - 1][lane][P + p] = ai[p]; }
        __syncthreads();
        if (wv == 0 && c < n) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int p = 0; p < P; ++p) { ar[p] += pt[q][lane][p]; ai[p] += pt[q][lane][P + p]; }
#pragma unroll
            for (int p = 0; p < P; ++p) { zr[p] = zr[p] - ar[p]; zi[p] = zi[p] - ai[p]; }
        }
        ++taken;
    }
    if (wv == 0)                                      // z_cb, zero past the last row
#pragma unroll
        for (int p = 0; p < P; ++p) {
            xs[lane][p] = lane < bc ? zr[p] : 0.0;
            xs[lane][P + p] = lane < bc ? zi[p] : 0.0;
        }
    __syncthreads();
    for (int e = tid; e < LB * P; e += 256) {
        const int cc = e % LB, p = e / LB;
        double sr = 0, si = 0;
#pragma unroll 16
        for (int t = 0; t < LB; ++t) {
            const double lr = xsm[0][t][cc], li = xsm[1][t][cc];
            sr += lr * xs[t][p] - li * xs[t][P + p];
            si += lr * xs[t][P + p] + li * xs[t][p];
        }
        if (cc < bc) {
            ldl_post(Xo + (size_t)(C0 + cc) * NR + p, sr);
            ldl_post(Xo + (size_t)(C0 + cc) * NR + P + p, si);
        }
    }
}

// ---- super-block sweeps: SB blocks of 64 rows per sequential launch ----
//
// The per-block sweeps above are a chain of one launch per 64-row block, each paying a dependent
// launch's fixed cost. Here the sweeps go by super-blocks of SB = 8 blocks: one workgroup per
// system solves a super-block's eight diagonal blocks in turn (each block's product with its
// stored inverse, then its update of the super-block's later rows, one barrier apart), and one
// parallel launch then applies the whole super-block to every row beyond it. That pays for a
// batch of many small systems (a workgroup each); one large system is faster with the per-block
// sweeps, whose launches spread each block's columns over many workgroups. Every sum runs in a
// fixed order, so batched and single systems give the same bits.
constexpr int SB = 8;

// Forward, super-block rows s0 .. s0 + 64 nbk: for each block j, y_j = X_j b_j (to Ys and LDS),
// then b_r -= L[r, block j] y_j for the super-block's rows r after block j (a thread per row, the
// block's 64 columns in order).
template <int P>
__global__ __launch_bounds__(256) void ldl_fwd_super_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                           long long ldf, const float *__restrict__ Xr,
                                                           const float *__restrict__ Xi, int n, int s0, int nbk,
                                                           double *Y, double *Ys, long long sF, long long sX,
                                                           long long sV) {
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;     // system blockIdx.z of a batch
    Xr += blockIdx.z * sX; Xi += blockIdx.z * sX;
    Y += blockIdx.z * sV; Ys += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ float xs[2][LB][LB + 1];
    __shared__ double bs[LB][NR], ys[LB][NR];
    const int tid = threadIdx.x;
    const int rend = min(s0 + LB * nbk, n);
    for (int j = 0; j < nbk; ++j) {
        const int k0 = s0 + LB * j, bk = min(LB, n - k0), kb = k0 / LB;
        const float *xr = Xr + (size_t)kb * LB * LB, *xi = Xi + (size_t)kb * LB * LB;
        for (int e = tid; e < LB * LB; e += 256) {
            xs[0][e >> 6][e & 63] = xr[e];
            xs[1][e >> 6][e & 63] = xi[e];
        }
        for (int e = tid; e < LB * NR; e += 256) {
            const int r = e / NR, q = e % NR;
            bs[r][q] = r < bk ? Y[(size_t)(k0 + r) * NR + q] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < LB * P; e += 256) {
            const int i = e / P, p = e % P;
            double sr = 0, si = 0;
#pragma unroll 16
            for (int t = 0; t < LB; ++t) {            // X_j is stored with its zeros above the diagonal
                const double lr = xs[0][i][t], li = xs[1][i][t];
                sr += lr * bs[t][p] - li * bs[t][P + p];
                si += lr * bs[t][P + p] + li * bs[t][p];
            }
            ys[i][p] = sr;
            ys[i][P + p] = si;
            if (i < bk) {
                Ys[(size_t)(k0 + i) * NR + p] = sr;
                Ys[(size_t)(k0 + i) * NR + P + p] = si;
            }
        }
        __syncthreads();
        for (int r = k0 + bk + tid; r < rend; r += 256) {   // rows after a block exist only when bk = 64
            double ar[P], ai[P];
#pragma unroll
            for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
            const float4 *lr4 = reinterpret_cast<const float4 *>(Fr + (size_t)r * ldf + k0);
            const float4 *li4 = reinterpret_cast<const float4 *>(Fi + (size_t)r * ldf + k0);
            // the row's 64 entries in four quarters, each quarter's loads issued together
#pragma unroll 1
            for (int h = 0; h < 4; ++h) {
                float4 a[4], b[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) { a[q] = lr4[4 * h + q]; b[q] = li4[4 * h + q]; }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float lre[4] = {a[q].x, a[q].y, a[q].z, a[q].w}, lim[4] = {b[q].x, b[q].y, b[q].z, b[q].w};
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int t = 16 * h + 4 * q + u;
                        const double xr2 = lre[u], xi2 = lim[u];
#pragma unroll
                        for (int p = 0; p < P; ++p) {
                            ar[p] += xr2 * ys[t][p] - xi2 * ys[t][P + p];
                            ai[p] += xr2 * ys[t][P + p] + xi2 * ys[t][p];
                        }
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                Y[(size_t)r * NR + p] -= ar[p];
                Y[(size_t)r * NR + P + p] -= ai[p];
            }
        }
        __syncthreads();
    }
}

// Forward, rows r >= r0 (= s0 + 64 nbk) after super-block s0: b_r -= sum over the super-block's
// columns of L[r, t] y_t. A workgroup takes 64 rows (lane = row); in each of the super-block's
// blocks its four waves take the four 16-column quarters; the waves' sums meet in LDS and wave 0
// subtracts them in wave order.
template <int P>
__global__ __launch_bounds__(256) void ldl_fwd_bulk_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                          long long ldf, int n, int s0, int nbk, double *Y,
                                                          const double *__restrict__ Ys, long long sF, long long sV) {
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;
    Y += blockIdx.z * sV; Ys += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ double part[3][LB][NR];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = s0 + LB * nbk + blockIdx.x * LB + lane;
    const int rr = min(r, n - 1);                     // rows past n read the last row, their sums unused
    double ar[P], ai[P];
#pragma unroll
    for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
#pragma unroll 2
    for (int j = 0; j < nbk; ++j) {
        const int k0 = s0 + LB * j;                   // every block before the rows below is full
        float lre[16], lim[16];
        const float4 *lr4 = reinterpret_cast<const float4 *>(Fr + (size_t)rr * ldf + k0 + 16 * wv);
        const float4 *li4 = reinterpret_cast<const float4 *>(Fi + (size_t)rr * ldf + k0 + 16 * wv);
#pragma unroll
        for (int t4 = 0; t4 < 4; ++t4) {
            const float4 a = lr4[t4], b = li4[t4];
            lre[4 * t4] = a.x; lre[4 * t4 + 1] = a.y; lre[4 * t4 + 2] = a.z; lre[4 * t4 + 3] = a.w;
            lim[4 * t4] = b.x; lim[4 * t4 + 1] = b.y; lim[4 * t4 + 2] = b.z; lim[4 * t4 + 3] = b.w;
        }
        const double *yb = Ys + (size_t)(k0 + 16 * wv) * NR;   // wave-uniform rows of y
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const double xr = lre[u], xi = lim[u];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const double yr = yb[u * NR + p], yi = yb[u * NR + P + p];
                ar[p] += xr * yr - xi * yi;
                ai[p] += xr * yi + xi * yr;
            }
        }
    }
    if (wv > 0)
#pragma unroll
        for (int p = 0; p < P; ++p) { part[wv - 1][lane][p] = ar[p]; part[wv - 1][lane][P + p] = ai[p]; }
    __syncthreads();
    if (wv == 0 && r < n) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int p = 0; p < P; ++p) { ar[p] += part[q][lane][p]; ai[p] += part[q][lane][P + p]; }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            Y[(size_t)r * NR + p] -= ar[p];
            Y[(size_t)r * NR + P + p] -= ai[p];
        }
    }
}

// Backward, super-block rows s0 .. s0 + 64 nbk, last block first: x_j = X_j^T z_j (to Xo and LDS),
// then z_c -= sum_t L[k0 + t][c] x_t for the super-block's columns c before block j (a thread per
// column, the block's rows in order).
template <int P>
__global__ __launch_bounds__(256) void ldl_bwd_super_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                           long long ldf, const float *__restrict__ Xr,
                                                           const float *__restrict__ Xi, int n, int s0, int nbk,
                                                           double *Z, double *Xo, long long sF, long long sX,
                                                           long long sV) {
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;
    Xr += blockIdx.z * sX; Xi += blockIdx.z * sX;
    Z += blockIdx.z * sV; Xo += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ float xsm[2][LB][LB + 1];
    __shared__ double zs[LB][NR], xs[LB][NR];
    const int tid = threadIdx.x;
    for (int j = nbk - 1; j >= 0; --j) {
        const int k0 = s0 + LB * j, bk = min(LB, n - k0), kb = k0 / LB;
        const float *xr = Xr + (size_t)kb * LB * LB, *xi = Xi + (size_t)kb * LB * LB;
        for (int e = tid; e < LB * LB; e += 256) {
            xsm[0][e >> 6][e & 63] = xr[e];
            xsm[1][e >> 6][e & 63] = xi[e];
        }
        for (int e = tid; e < LB * NR; e += 256) {
            const int r = e / NR, q = e % NR;
            zs[r][q] = r < bk ? Z[(size_t)(k0 + r) * NR + q] : 0.0;
        }
        __syncthreads();
        for (int e = tid; e < LB * P; e += 256) {
            const int c = e % LB, p = e / LB;
            double sr = 0, si = 0;
#pragma unroll 16
            for (int t = 0; t < LB; ++t) {            // zeros above X_j's diagonal and past bk in zs
                const double lr = xsm[0][t][c], li = xsm[1][t][c];
                sr += lr * zs[t][p] - li * zs[t][P + p];
                si += lr * zs[t][P + p] + li * zs[t][p];
            }
            xs[c][p] = sr;
            xs[c][P + p] = si;
            if (c < bk) {
                Xo[(size_t)(k0 + c) * NR + p] = sr;
                Xo[(size_t)(k0 + c) * NR + P + p] = si;
            }
        }
        __syncthreads();
        for (int c = s0 + tid; c < k0; c += 256) {
            double ar[P], ai[P];
#pragma unroll
            for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
            for (int t0 = 0; t0 < bk; t0 += 16) {
                float lre[16], lim[16];
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const bool in = t0 + u < bk;
                    lre[u] = in ? Fr[(size_t)(k0 + t0 + u) * ldf + c] : 0.f;
                    lim[u] = in ? Fi[(size_t)(k0 + t0 + u) * ldf + c] : 0.f;
                }
#pragma unroll
                for (int u = 0; u < 16; ++u) {
                    const int t = t0 + u;
                    const double lr = lre[u], li = lim[u];
#pragma unroll
                    for (int p = 0; p < P; ++p) {
                        ar[p] += lr * xs[t][p] - li * xs[t][P + p];
                        ai[p] += lr * xs[t][P + p] + li * xs[t][p];
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < P; ++p) {
                Z[(size_t)c * NR + p] -= ar[p];
                Z[(size_t)c * NR + P + p] -= ai[p];
            }
        }
        __syncthreads();
    }
}

// Backward, columns r < s0 before super-block s0: z_r -= sum over the super-block's rows t of
// L[t][r] x_t. A workgroup takes 64 columns (lane = r, reads across lanes); in each block its four
// waves take the four 16-row quarters; the waves' sums meet in LDS and wave 0 subtracts them.
template <int P>
__global__ __launch_bounds__(256) void ldl_bwd_bulk_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                          long long ldf, int n, int s0, int nbk, double *Z,
                                                          const double *__restrict__ Xo, long long sF, long long sV) {
    Fr += blockIdx.z * sF; Fi += blockIdx.z * sF;
    Z += blockIdx.z * sV; Xo += blockIdx.z * sV;
    constexpr int NR = 2 * P;
    __shared__ double part[3][LB][NR];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int r = blockIdx.x * LB + lane;             // r < s0 always: s0 is a multiple of 64
    double ar[P], ai[P];
#pragma unroll
    for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
#pragma unroll 2
    for (int j = 0; j < nbk; ++j) {
        const int k0 = s0 + LB * j, bk = min(LB, n - k0), t0 = 16 * wv;
        float lre[16], lim[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const bool in = t0 + u < bk;
            lre[u] = in ? Fr[(size_t)(k0 + t0 + u) * ldf + r] : 0.f;
            lim[u] = in ? Fi[(size_t)(k0 + t0 + u) * ldf + r] : 0.f;
        }
        const double *xb = Xo + (size_t)(k0 + t0) * NR;  // wave-uniform rows of x
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            if (t0 + u >= bk) break;                   // wave-uniform
            const double xr = lre[u], xi = lim[u];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                const double vr = xb[u * NR + p], vi = xb[u * NR + P + p];
                ar[p] += xr * vr - xi * vi;
                ai[p] += xr * vi + xi * vr;
            }
        }
    }
    if (wv > 0)
#pragma unroll
        for (int p = 0; p < P; ++p) { part[wv - 1][lane][p] = ar[p]; part[wv - 1][lane][P + p] = ai[p]; }
    __syncthreads();
    if (wv == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int p = 0; p < P; ++p) { ar[p] += part[q][lane][p]; ai[p] += part[q][lane][P + p]; }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            Z[(size_t)r * NR + p] -= ar[p];
            Z[(size_t)r * NR + P + p] -= ai[p];
        }
    }
}

// ---- the dense symmetric positive definite solve (real A FP64 on the device) ----

// F = A in FP32 (lower triangle), imaginary plane zero.
__global__ __launch_bounds__(256) void spd_form_kernel(const double *__restrict__ A, long long lda, int n,
                                                      float *__restrict__ Fr, float *__restrict__ Fi, long long ldf) {
    for (int row = blockIdx.y; row < n; row += gridDim.y)
        for (int c = blockIdx.x * 256 + threadIdx.x; c <= row; c += gridDim.x * 256) {
            Fr[(size_t)row * ldf + c] = (float)A[(size_t)row * lda + c];
            Fi[(size_t)row * ldf + c] = 0.f;
        }
}

// Chunk partials of Y = A X for symmetric A (FP64): thread per row, reading column `row` of the
// chunk's rows of A (coalesced across threads), X rows wave-uniform; part[(chunk, row)][q].
template <int NR>
__global__ __launch_bounds__(256) void spd_gemv_kernel(const double *__restrict__ A, long long lda, int n,
                                                      const double *__restrict__ X, int chunk,
                                                      double *__restrict__ part) {
    const int row = blockIdx.x * 256 + threadIdx.x;
    const int k0 = blockIdx.y * chunk, k1 = min(n, k0 + chunk);
    const int rr = min(row, n - 1);
    double acc[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) acc[q] = 0;
    for (int k = k0; k < k1; ++k) {
        const double a = A[(size_t)k * lda + rr];
#pragma unroll
        for (int q = 0; q < NR; ++q) acc[q] += a * X[(size_t)k * NR + q];
    }
    if (row < n)
#pragma unroll
        for (int q = 0; q < NR; ++q) part[((size_t)blockIdx.y * n + row) * NR + q] = acc[q];
}

// R = B - sum of the nsplit chunk partials (in chunk order).
__global__ void spd_residual_kernel(const double *__restrict__ part, int nsplit, size_t count,
                                    const double *__restrict__ B, double *__restrict__ R) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    double s = 0;
    for (int q = 0; q < nsplit; ++q) s += part[(size_t)q * count + e];
    R[e] = B[e] - s;
}

// out[q] = sum over rows of V[row][q]^2 (one workgroup per column q, fixed-order tree).
__global__ __launch_bounds__(256) void colnorm2_kernel(const double *__restrict__ V, int n, int NR,
                                                      double *__restrict__ out) {
    __shared__ double s[256];
    const int q = blockIdx.x;
    double a = 0;
    for (int row = threadIdx.x; row < n; row += 256) {
        const double v = V[(size_t)row * NR + q];
        a += v * v;
    }
    s[threadIdx.x] = a;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] += s[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = s[0];
}

__global__ void add_kernel(const double *X, size_t count, double *Y) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) Y[e] += X[e];
}

template <class F> void dispatch_ports(int P, F &&f) {
    switch (P) {
    case 1: f(std::integral_constant<int, 1>{}); break;
    case 2: f(std::integral_constant<int, 2>{}); break;
    case 3: f(std::integral_constant<int, 3>{}); break;
    case 4: f(std::integral_constant<int, 4>{}); break;
    case 5: f(std::integral_constant<int, 5>{}); break;
    case 6: f(std::integral_constant<int, 6>{}); break;
    case 7: f(std::integral_constant<int, 7>{}); break;
    case 8: f(std::integral_constant<int, 8>{}); break;
    case 16: f(std::integral_constant<int, 16>{}); break;
    default: throw Error(PQ_EINVAL, "direct solve: 1 to 8 or 16 columns");
    }
}

// A complex-symmetric matrix given by its lower-triangle entries (unique, row >= column): the
// FP32 planes of the factor (zeroed beforehand).
__global__ __launch_bounds__(256) void csym_scatter_kernel(const int *__restrict__ rows, const int *__restrict__ cols,
                                                          const double *__restrict__ vals, long long nnz,
                                                          float *__restrict__ Fr, float *__restrict__ Fi,
                                                          long long ldf) {
    const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
    if (t >= nnz) return;
    const size_t o = (size_t)rows[t] * ldf + cols[t];
    Fr[o] = (float)vals[2 * t];
    Fi[o] = (float)vals[2 * t + 1];
}

// R = B - A X for P complex columns (rows [P real parts, P imaginary parts]), A in CSR with both
// triangles, complex FP64; one thread per row, the row's entries in CSR order.
__global__ __launch_bounds__(256) void csym_residual_kernel(const int *__restrict__ rp, const int *__restrict__ ci,
                                                           const double *__restrict__ av, int n, int P,
                                                           const double *__restrict__ X,
                                                           const double *__restrict__ B, double *__restrict__ R) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int NR = 2 * P;
    for (int q = 0; q < P; ++q) {
        double re = 0, im = 0;
        for (int e = rp[i]; e < rp[i + 1]; ++e) {
            const double ar = av[2 * e], ai = av[2 * e + 1];
            const double xr = X[(size_t)ci[e] * NR + q], xi = X[(size_t)ci[e] * NR + P + q];
            re += ar * xr - ai * xi;
            im += ar * xi + ai * xr;
        }
        R[(size_t)i * NR + q] = B[(size_t)i * NR + q] - re;
        R[(size_t)i * NR + P + q] = B[(size_t)i * NR + P + q] - im;
    }
}

}  // namespace

void ldl_create_aux(hipStream_t *aux) {
    const char *env = std::getenv("PCBQS_AMD_LDL_CUMASK");
    int dev = 0, cus = 0;
    PQ_HIP(hipGetDevice(&dev));
    PQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (!(env && env[0] == '1') || cus < 64) {
        PQ_HIP(hipStreamCreateWithFlags(aux, hipStreamNonBlocking));
        return;
    }
    std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0xffffffffu);
    if (cus % 32) mask.back() = (1u << (cus % 32)) - 1;
    for (int i = 0; i < 8; ++i) {
        const int cu = 32 * i + i;
        if (cu < cus) mask[(size_t)cu / 32] &= ~(1u << (cu % 32));
    }
    PQ_HIP(hipExtStreamCreateWithCUMask(aux, (uint32_t)mask.size(), mask.data()));
}

void ldl_form(hipStream_t st, const LdlFactor &f, const float *Ll, long long ldl, const int *rptr, const int *rcol,
              const double *rval, double omega) {
    const int n = f.n;
    if (n <= 0) return;
    hipLaunchKernelGGL(ldl_form_kernel, dim3((n + 255) / 256, std::min(n, 65535)), dim3(256), 0, st, Ll, ldl, n,
                       (float)omega, f.Fr, f.Fi, f.ldf);
    hipLaunchKernelGGL(ldl_form_r_kernel, dim3((n + 255) / 256), dim3(256), 0, st, rptr, rcol, rval, n, f.Fr, f.ldf);
    PQ_HIP(hipGetLastError());
}

// Panels are paired into 128-wide super-panels K .. K + 127. On st, per super-panel: the first
// panel (diagonal block, panel into W[:, 0:64]), its update of the second panel's 64 columns only,
// the second panel (diagonal block, panel into W[:, 64:128]); then one K = 128 update of the trailing
// matrix from row / column K + 128 (half the read-modify-write traffic of two 64-wide updates). Of
// that update, the first two tile columns (all the next super-panel reads) run on st; the rest goes
// to f.aux after them and runs beside the next super-panel's panels, and the next first-columns
// update waits for it, since both write those tile columns. W alternates between two buffers, so a
// buffer is rewritten only after the bulk update that read it has been waited for.
void ldl_factor(hipStream_t st, const LdlFactor &f, int *bad) {
    const int n = f.n;
    constexpr int LDW = 2 * LB;
    const unsigned nz = (unsigned)std::max(1, f.batch);
    auto diag = [&](int k) {
        const int b = std::min(LB, n - k);
        hipLaunchKernelGGL(ldl_diag_kernel, dim3(1, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf, k, b,
                           f.Lir + (size_t)(k / LB) * LB * LB, f.Lii + (size_t)(k / LB) * LB * LB, bad, f.sF, f.sX);
    };
    // under a band, rows past the last one that the columns just factored reach are zero in L21 and
    // W: the kernels take that row as their end
    auto reach = [&](int last_col) { return f.band > 0 ? std::min(n, last_col + 1 + f.band) : n; };
    auto panel = [&](int k, float *wr, float *wi) {
        const int ne = reach(k + LB - 1), tiles = (ne - k - LB + LB - 1) / LB;
        if (tiles <= 0) return;
        hipLaunchKernelGGL(ldl_panel_kernel, dim3(tiles, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf, ne, k,
                           f.Lir + (size_t)(k / LB) * LB * LB, f.Lii + (size_t)(k / LB) * LB * LB, wr, wi, LDW, f.sF,
                           f.sX, f.sW);
    };
    if (n <= 0) return;
    diag(0);
    bool bulk = false;                                   // a bulk update is pending on f.aux
    for (int K = 0, s = 0; K + LB < n; K += 2 * LB, ++s) {
        float *wr = f.Wr[s & 1], *wi = f.Wi[s & 1];
        panel(K, wr, wi);
        {   // the first panel's update of the second panel's columns (rows from K + 64)
            const int ne = reach(K + LB - 1), T = (ne - K - LB + LB - 1) / LB;
            if (T > 0)
                hipLaunchKernelGGL(ldl_update_kernel, dim3(T, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf, ne, K, LB,
                                   K + LB, wr, wi, LDW, 1, 1, f.sF, f.sW);
        }
        diag(K + LB);
        if (K + 2 * LB >= n) break;
        panel(K + LB, wr + LB, wi + LB);
        const int ne = reach(K + 2 * LB - 1);
        const int m0 = K + 2 * LB, T = (ne - m0 + LB - 1) / LB, tc = std::min(2, T);
        if (T <= 0) {                                    // a band ending inside the super-panel
            if (bulk) PQ_HIP(hipStreamWaitEvent(st, f.done, 0));
            bulk = false;
            diag(m0);
            continue;
        }
        if (bulk) PQ_HIP(hipStreamWaitEvent(st, f.done, 0));
        hipLaunchKernelGGL(ldl_update_kernel, dim3((unsigned)(tc * T - tc * (tc - 1) / 2), 1, nz), dim3(256), 0, st,
                           f.Fr, f.Fi, f.ldf, ne, K, 2 * LB, m0, wr, wi, LDW, 1, tc, f.sF, f.sW);
        if (!f.aux) {                                    // one stream (a captured batch): the bulk in order
            if (T > tc)
                hipLaunchKernelGGL(ldl_update_kernel, dim3((unsigned)((size_t)(T - tc) * (T - tc + 1) / 2), 1, nz),
                                   dim3(256), 0, st, f.Fr, f.Fi, f.ldf, ne, K, 2 * LB, m0, wr, wi, LDW, 2, tc, f.sF,
                                   f.sW);
            diag(m0);
            continue;
        }
        bulk = T > tc;
        if (bulk) {
            PQ_HIP(hipEventRecord(f.ready, st));
            PQ_HIP(hipStreamWaitEvent(f.aux, f.ready, 0));
            hipLaunchKernelGGL(ldl_update_kernel, dim3((unsigned)((size_t)(T - tc) * (T - tc + 1) / 2), 1, nz), dim3(256),
                               0, f.aux, f.Fr, f.Fi, f.ldf, ne, K, 2 * LB, m0, wr, wi, LDW, 2, tc, f.sF, f.sW);
            PQ_HIP(hipEventRecord(f.done, f.aux));
        }
        diag(m0);
    }
    if (bulk) PQ_HIP(hipStreamWaitEvent(st, f.done, 0));
    PQ_HIP(hipGetLastError());
}

void ldl_solve(hipStream_t st, const LdlFactor &f, int P, const double *B, double *Y, double *Ys, double *X) {
    const int n = f.n, nb = (n + LB - 1) / LB;
    if (n <= 0) return;
    const unsigned nz = (unsigned)std::max(1, f.batch);
    // Super-block sweeps for systems up to 4,096 unknowns (the many small systems of a multi-board
    // batch: a sequential workgroup per system, 2 launches per 512 rows); above, or within a band,
    // one launch per sweep (a large system's blocks spread over many workgroups, which the one
    // sequential workgroup cannot match). PCBQS_AMD_LDL_BLOCK_SWEEPS set (read on every call):
    // one launch per 64-row block instead, with the same values.
    const bool block_env = std::getenv("PCBQS_AMD_LDL_BLOCK_SWEEPS") != nullptr;
    const bool block_sweeps = block_env || n > 4096 || f.band > 0;
    const long long sV = (long long)n * 2 * P;           // the systems' vectors, one after another
    if (block_sweeps && !block_env) {
        const size_t count = (size_t)n * 2 * P * nz;
        hipLaunchKernelGGL(ldl_mark_empty_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, Ys, count);
        dispatch_ports(P, [&](auto pc) {
            constexpr int PP = decltype(pc)::value;
            hipLaunchKernelGGL(ldl_fwd_persist_kernel<PP>, dim3(nb, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf, f.Lir,
                               f.Lii, n, f.band, B, Ys, f.sF, f.sX, sV);
            hipLaunchKernelGGL(ldl_scale_kernel, dim3((unsigned)(((size_t)n * PP + 255) / 256), 1, nz), dim3(256), 0,
                               st, f.Fr, f.Fi, f.ldf, n, PP, Ys, Y, f.sF, sV, X);
            hipLaunchKernelGGL(ldl_bwd_persist_kernel<PP>, dim3(nb, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf, f.Lir,
                               f.Lii, n, f.band, Y, X, f.sF, f.sX, sV);
        });
        PQ_HIP(hipGetLastError());
        return;
    }
    const size_t bytes = (size_t)n * 2 * P * 8 * nz;
    PQ_HIP(hipMemcpyAsync(Y, B, bytes, hipMemcpyDeviceToDevice, st));
    dispatch_ports(P, [&](auto pc) {
        constexpr int PP = decltype(pc)::value;
        if (block_sweeps) {                              // one launch per 64-row block
            for (int kb = 0; kb < nb; ++kb) {
                const int k0 = kb * LB, bk = std::min(LB, n - k0);
                const int rows = f.band > 0 ? std::min(n - k0 - bk, f.band) : n - k0 - bk;
                const int grid = std::max(1, (rows + LB - 1) / LB);
                hipLaunchKernelGGL(ldl_forward_kernel<PP>, dim3(grid, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf,
                                   f.Lir + (size_t)kb * LB * LB, f.Lii + (size_t)kb * LB * LB, n, k0, bk, Y, Ys, f.sF,
                                   f.sX, sV);
            }
        } else {
            for (int kb0 = 0; kb0 < nb; kb0 += SB) {
                const int s0 = kb0 * LB, nbk = std::min(SB, nb - kb0), below = n - s0 - nbk * LB;
                hipLaunchKernelGGL(ldl_fwd_super_kernel<PP>, dim3(1, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf,
                                   f.Lir, f.Lii, n, s0, nbk, Y, Ys, f.sF, f.sX, sV);
                if (below > 0)
                    hipLaunchKernelGGL(ldl_fwd_bulk_kernel<PP>, dim3((below + LB - 1) / LB, 1, nz), dim3(256), 0, st,
                                       f.Fr, f.Fi, f.ldf, n, s0, nbk, Y, Ys, f.sF, sV);
            }
        }
        hipLaunchKernelGGL(ldl_scale_kernel, dim3((unsigned)(((size_t)n * PP + 255) / 256), 1, nz), dim3(256), 0, st,
                           f.Fr, f.Fi, f.ldf, n, PP, Ys, Y, f.sF, sV, nullptr);
        if (block_sweeps) {
            for (int kb = nb - 1; kb >= 0; --kb) {
                const int k0 = kb * LB, bk = std::min(LB, n - k0);
                // columns r < k0 that the block's rows reach: all, or from k0 - band down a band
                const int r0 = f.band > 0 ? std::max(0, (k0 - f.band) / LB * LB) : 0;
                const int grid = std::max(1, (k0 - r0 + LB - 1) / LB);
                hipLaunchKernelGGL(ldl_backward_kernel<PP>, dim3(grid, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf,
                                   f.Lir + (size_t)kb * LB * LB, f.Lii + (size_t)kb * LB * LB, k0, bk, Y, X, f.sF,
                                   f.sX, sV, r0);
            }
        } else {
            for (int kb0 = (nb - 1) / SB * SB; kb0 >= 0; kb0 -= SB) {
                const int s0 = kb0 * LB, nbk = std::min(SB, nb - kb0);
                hipLaunchKernelGGL(ldl_bwd_super_kernel<PP>, dim3(1, 1, nz), dim3(256), 0, st, f.Fr, f.Fi, f.ldf,
                                   f.Lir, f.Lii, n, s0, nbk, Y, X, f.sF, f.sX, sV);
                if (s0 > 0)
                    hipLaunchKernelGGL(ldl_bwd_bulk_kernel<PP>, dim3(s0 / LB, 1, nz), dim3(256), 0, st, f.Fr, f.Fi,
                                       f.ldf, n, s0, nbk, Y, X, f.sF, sV);
            }
        }
    });
    PQ_HIP(hipGetLastError());
}

}  // namespace pq

using namespace pq;

extern "C" {

// X = A^-1 B for a real symmetric positive definite A (n x n FP64 on the device, row stride lda,
// its lower triangle read) and nrhs right-hand sides B (n x nrhs FP64 host, row-major): A factored
// as L D L^T in FP32 on the device (the direct port solve's kernels), the solution refined in FP64
// against A until every column's relative residual is at most tol, at most max_steps corrections,
// stopping once a correction no longer halves it (a result above 1e-10 returns PQ_ECONVERGE). No
// counterpart as an entry point in the reference: it serves capacitance_matrix's solve_symmetric
// (src/pcbqs/peec.py:95-123, kernels.py solve_symmetric).
int pq_spd_solve(int device, int n, const double *A_dev, int64_t lda, int nrhs, const double *B_host, double tol,
                 int max_steps, double *X_host, int64_t *steps, double *residual, double *factor_ms) {
    return guard([&] {
        PQ_REQUIRE(A_dev && B_host && X_host, "pq_spd_solve: null argument");
        PQ_REQUIRE(n >= 1 && lda >= n && nrhs >= 1 && tol > 0 && tol < 1 && max_steps >= 0,
                   "pq_spd_solve: invalid arguments");
        DeviceScope scope(device);
        DevBuf F, Xi, Wb, bad;                       // freed after the streams have drained (below)
        hipStream_t st = nullptr, aux = nullptr;
        hipEvent_t ready = nullptr, done = nullptr, e0 = nullptr, e1 = nullptr;
        struct Cleanup {
            hipStream_t &st, &aux;
            hipEvent_t &ready, &done, &e0, &e1;
            ~Cleanup() {
                if (st) (void)hipStreamSynchronize(st);
                if (aux) (void)hipStreamSynchronize(aux);
                for (hipEvent_t e : {ready, done, e0, e1})
                    if (e) (void)hipEventDestroy(e);
                if (aux) (void)hipStreamDestroy(aux);
                if (st) (void)hipStreamDestroy(st);
            }
        } cleanup{st, aux, ready, done, e0, e1};
        PQ_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        ldl_create_aux(&aux);
        PQ_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        PQ_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        PQ_HIP(hipEventCreate(&e0));
        PQ_HIP(hipEventCreate(&e1));
        const int nblk = (n + LB - 1) / LB;
        const long long ldf = (long long)nblk * LB;
        F.reserve((size_t)2 * n * ldf * 4);
        Xi.reserve((size_t)2 * nblk * LB * LB * 4);
        const size_t wsz = (size_t)n * 2 * LB;
        Wb.reserve(4 * wsz * 4);
        bad.reserve(4);
        float *wb = Wb.as<float>();
        LdlFactor f{F.as<float>(), F.as<float>() + (size_t)n * ldf, ldf, n, Xi.as<float>(),
                    Xi.as<float>() + (size_t)nblk * LB * LB, {wb, wb + 2 * wsz}, {wb + wsz, wb + 3 * wsz}, aux,
                    ready, done};
        PQ_HIP(hipEventRecord(e0, st));
        hipLaunchKernelGGL(spd_form_kernel, dim3((n + 255) / 256, std::min(n, 65535)), dim3(256), 0, st, A_dev, lda, n,
                           f.Fr, f.Fi, f.ldf);
        PQ_HIP(hipMemsetAsync(bad.p, 0, 4, st));
        ldl_factor(st, f, bad.as<int>());
        PQ_HIP(hipEventRecord(e1, st));
        int flag = 0;
        PQ_HIP(hipMemcpyAsync(&flag, bad.p, 4, hipMemcpyDeviceToHost, st));
        PQ_HIP(hipStreamSynchronize(st));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (factor_ms) *factor_ms = ms;
        PQ_REQUIRE(!flag, "pq_spd_solve: a zero or non-finite pivot (A is not positive definite)");
        // the right-hand sides in groups of up to 16, each as the real parts of complex columns
        int64_t worst_steps = 0;
        double worst = 0;
        const int chunk = 256, nsplit = (n + chunk - 1) / chunk;
        for (int c0 = 0; c0 < nrhs; c0 += 16) {
            const int cn = std::min(16, nrhs - c0);
            const int P = cn <= 8 ? cn : 16, NR = 2 * P;
            const size_t count = (size_t)n * NR;
            std::vector<double> hb(count, 0.0);
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < cn; ++q) hb[(size_t)i * NR + q] = B_host[(size_t)i * nrhs + c0 + q];
            DevBuf Bd, X, R, D, Y, Ys, part, nrm;
            for (DevBuf *b : {&Bd, &X, &R, &D, &Y, &Ys}) b->reserve(count * 8);
            part.reserve((size_t)nsplit * count * 8);
            nrm.reserve(NR * 8);
            PQ_HIP(hipMemcpyAsync(Bd.p, hb.data(), count * 8, hipMemcpyHostToDevice, st));
            std::vector<double> bn(NR), rn(NR);
            auto colnorms = [&](const double *V, std::vector<double> &out) {
                hipLaunchKernelGGL(colnorm2_kernel, dim3(NR), dim3(256), 0, st, V, n, NR, nrm.as<double>());
                PQ_HIP(hipMemcpyAsync(out.data(), nrm.p, NR * 8, hipMemcpyDeviceToHost, st));
                PQ_HIP(hipStreamSynchronize(st));
                for (auto &v : out) v = std::sqrt(v);
            };
            auto residual_of = [&](const double *Xv) {
                dispatch_ports(P, [&](auto pc) {
                    constexpr int N = 2 * decltype(pc)::value;
                    hipLaunchKernelGGL(spd_gemv_kernel<N>, dim3((n + 255) / 256, nsplit), dim3(256), 0, st, A_dev, lda,
                                       n, Xv, chunk, part.as<double>());
                });
                hipLaunchKernelGGL(spd_residual_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st,
                                   part.as<double>(), nsplit, count, Bd.as<double>(), R.as<double>());
                PQ_HIP(hipGetLastError());
            };
            colnorms(Bd.as<double>(), bn);
            ldl_solve(st, f, P, Bd.as<double>(), Y.as<double>(), Ys.as<double>(), X.as<double>());
            double last = std::numeric_limits<double>::infinity(), res = 0;
            int corr = 0;
            for (int step = 0;; ++step) {
                residual_of(X.as<double>());
                colnorms(R.as<double>(), rn);
                res = 0;
                for (int q = 0; q < cn; ++q) res = std::max(res, rn[q] / std::max(bn[q], 1e-300));
                if (!(res > tol) || step == max_steps || res > 0.5 * last) break;
                last = res;
                ldl_solve(st, f, P, R.as<double>(), Y.as<double>(), Ys.as<double>(), D.as<double>());
                hipLaunchKernelGGL(add_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, D.as<double>(),
                                   count, X.as<double>());
                PQ_HIP(hipGetLastError());
                ++corr;
            }
            worst = std::max(worst, res);
            worst_steps = std::max<int64_t>(worst_steps, corr);
            std::vector<double> hx(count);
            PQ_HIP(hipMemcpyAsync(hx.data(), X.p, count * 8, hipMemcpyDeviceToHost, st));
            PQ_HIP(hipStreamSynchronize(st));
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < cn; ++q) X_host[(size_t)i * nrhs + c0 + q] = hx[(size_t)i * NR + q];
        }
        if (steps) *steps = worst_steps;
        if (residual) *residual = worst;
        if (worst > std::max(tol, 1e-10))
            throw Error(PQ_ECONVERGE, "pq_spd_solve: the FP32 factor's refinement stalls at relative residual " +
                                          std::to_string(worst));
    });
}

}  // extern "C"

// ---- a complex-symmetric solver on the device from sparse entries ------------------------------

struct pq_csym {
    int device = 0, n = 0;
    hipStream_t st = nullptr, aux = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    DevBuf F, Xi, Wb, rp, ci, av;
    LdlFactor f{};
    double factor_ms = 0;
};

namespace {
void csym_release(pq_csym *h) {
    if (h->st) (void)hipStreamSynchronize(h->st);
    if (h->aux) (void)hipStreamSynchronize(h->aux);
    for (hipEvent_t e : {h->ready, h->done})
        if (e) (void)hipEventDestroy(e);
    if (h->aux) (void)hipStreamDestroy(h->aux);
    if (h->st) (void)hipStreamDestroy(h->st);
    delete h;
}
}  // namespace

extern "C" {

int pq_csym_create(int device, int n, int64_t nnz, const int32_t *rows, const int32_t *cols, const double *vals,
                   pq_csym **out, double *factor_ms, int64_t *bytes) {
    return guard([&] {
        PQ_REQUIRE(out && rows && cols && vals, "pq_csym_create: null argument");
        PQ_REQUIRE(n >= 1 && nnz >= 0, "pq_csym_create: invalid sizes");
        std::vector<long long> key((size_t)nnz);
        std::vector<int> cnt((size_t)n + 1, 0);
        int band = 0;                                     // the largest row - column of an entry
        for (int64_t t = 0; t < nnz; ++t) {
            PQ_REQUIRE(rows[t] >= 0 && rows[t] < n && cols[t] >= 0 && cols[t] <= rows[t],
                       "pq_csym_create: entries must lie in the lower triangle of an n x n matrix");
            band = std::max(band, rows[t] - cols[t]);
            PQ_REQUIRE(std::isfinite(vals[2 * t]) && std::isfinite(vals[2 * t + 1]), "pq_csym_create: entry not finite");
            key[(size_t)t] = (long long)rows[t] * n + cols[t];
            ++cnt[(size_t)rows[t] + 1];
            if (rows[t] != cols[t]) ++cnt[(size_t)cols[t] + 1];
        }
        std::sort(key.begin(), key.end());
        PQ_REQUIRE(std::adjacent_find(key.begin(), key.end()) == key.end(), "pq_csym_create: duplicate entry");
        // both triangles in CSR for the FP64 residual
        for (int i = 0; i < n; ++i) cnt[(size_t)i + 1] += cnt[(size_t)i];
        const int full = cnt[(size_t)n];
        std::vector<int> ci((size_t)std::max(full, 1)), next(cnt.begin(), cnt.end() - 1);
        std::vector<double> av((size_t)2 * std::max(full, 1));
        auto put = [&](int r, int c, double re, double im) {
            const int e = next[(size_t)r]++;
            ci[(size_t)e] = c;
            av[(size_t)2 * e] = re;
            av[(size_t)2 * e + 1] = im;
        };
        for (int64_t t = 0; t < nnz; ++t) {
            put(rows[t], cols[t], vals[2 * t], vals[2 * t + 1]);
            if (rows[t] != cols[t]) put(cols[t], rows[t], vals[2 * t], vals[2 * t + 1]);
        }
        DeviceScope scope(device);
        auto *h = new pq_csym();
        try {
            h->device = device; h->n = n;
            PQ_HIP(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
            ldl_create_aux(&h->aux);
            PQ_HIP(hipEventCreateWithFlags(&h->ready, hipEventDisableTiming));
            PQ_HIP(hipEventCreateWithFlags(&h->done, hipEventDisableTiming));
            const int nblk = (n + LB - 1) / LB;
            const long long ldf = (long long)nblk * LB;
            // the factor is stored dense (n x n complex FP32, even within a band): refuse by name
            // what the device cannot hold instead of failing inside an allocation
            size_t free_bytes = 0, total_bytes = 0;
            PQ_HIP(hipMemGetInfo(&free_bytes, &total_bytes));
            const size_t need = (size_t)2 * n * ldf * 4 + (size_t)2 * nblk * LB * LB * 4 + (size_t)16 * n * 2 * LB;
            if (need > free_bytes)
                throw Error(PQ_ENOMEM, "pq_csym_create: the dense factor of " + std::to_string(n) + " unknowns needs " +
                                           std::to_string(need >> 20) + " MiB, " + std::to_string(free_bytes >> 20) +
                                           " MiB free on the device");
            h->F.reserve((size_t)2 * n * ldf * 4);
            h->Xi.reserve((size_t)2 * nblk * LB * LB * 4);
            const size_t wsz = (size_t)n * 2 * LB;
            h->Wb.reserve(4 * wsz * 4);
            float *wb = h->Wb.as<float>();
            h->f = LdlFactor{h->F.as<float>(), h->F.as<float>() + (size_t)n * ldf, ldf, n, h->Xi.as<float>(),
                             h->Xi.as<float>() + (size_t)nblk * LB * LB, {wb, wb + 2 * wsz}, {wb + wsz, wb + 3 * wsz},
                             h->aux, h->ready, h->done};
            // a banded matrix (the graph Laplacian in a bandwidth-reducing order) factors and sweeps
            // within its band; the dense kernels otherwise
            if (band + 2 * LB < n) h->f.band = std::max(band, 1);
            h->rp.reserve(cnt.size() * 4);
            h->ci.reserve(ci.size() * 4);
            h->av.reserve(av.size() * 8);
            PQ_HIP(hipMemcpy(h->rp.p, cnt.data(), cnt.size() * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(h->ci.p, ci.data(), ci.size() * 4, hipMemcpyHostToDevice));
            PQ_HIP(hipMemcpy(h->av.p, av.data(), av.size() * 8, hipMemcpyHostToDevice));
            DevBuf r, c, v, bad;
            r.reserve((size_t)std::max<int64_t>(nnz, 1) * 4);
            c.reserve((size_t)std::max<int64_t>(nnz, 1) * 4);
            v.reserve((size_t)std::max<int64_t>(nnz, 1) * 16);
            bad.reserve(4);
            if (nnz) {
                PQ_HIP(hipMemcpy(r.p, rows, (size_t)nnz * 4, hipMemcpyHostToDevice));
                PQ_HIP(hipMemcpy(c.p, cols, (size_t)nnz * 4, hipMemcpyHostToDevice));
                PQ_HIP(hipMemcpy(v.p, vals, (size_t)nnz * 16, hipMemcpyHostToDevice));
            }
            hipEvent_t e0, e1;
            PQ_HIP(hipEventCreate(&e0));
            PQ_HIP(hipEventCreate(&e1));
            PQ_HIP(hipEventRecord(e0, h->st));
            PQ_HIP(hipMemsetAsync(h->F.p, 0, (size_t)2 * n * ldf * 4, h->st));
            if (nnz)
                hipLaunchKernelGGL(csym_scatter_kernel, dim3((unsigned)((nnz + 255) / 256)), dim3(256), 0, h->st,
                                   r.as<int>(), c.as<int>(), v.as<double>(), (long long)nnz, h->f.Fr, h->f.Fi, ldf);
            PQ_HIP(hipMemsetAsync(bad.p, 0, 4, h->st));
            ldl_factor(h->st, h->f, bad.as<int>());
            PQ_HIP(hipEventRecord(e1, h->st));
            int flag = 0;
            PQ_HIP(hipMemcpyAsync(&flag, bad.p, 4, hipMemcpyDeviceToHost, h->st));
            PQ_HIP(hipStreamSynchronize(h->st));
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            h->factor_ms = ms;
            PQ_REQUIRE(!flag, "pq_csym_create: a zero or non-finite pivot");
        } catch (...) {
            csym_release(h);
            throw;
        }
        if (factor_ms) *factor_ms = h->factor_ms;
        if (bytes) *bytes = (int64_t)(h->F.bytes + h->Xi.bytes);
        *out = h;
    });
}

int pq_csym_destroy(pq_csym *h) {
    if (!h) return PQ_OK;
    return guard([&] {
        DeviceScope scope(h->device);
        csym_release(h);
    });
}

int pq_csym_solve(pq_csym *h, int nrhs, const double *B, double tol, int max_steps, double *X, int64_t *steps,
                  double *residual) {
    return guard([&] {
        PQ_REQUIRE(h && B && X && nrhs >= 1 && tol > 0 && tol < 1 && max_steps >= 0, "pq_csym_solve: invalid arguments");
        DeviceScope scope(h->device);
        const int n = h->n;
        hipStream_t st = h->st;
        int64_t worst_steps = 0;
        double worst = 0;
        for (int c0 = 0; c0 < nrhs; c0 += 16) {
            const int cn = std::min(16, nrhs - c0);
            const int P = cn <= 8 ? cn : 16, NR = 2 * P;
            const size_t count = (size_t)n * NR;
            std::vector<double> hb(count, 0.0);
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < cn; ++q) {
                    hb[(size_t)i * NR + q] = B[2 * ((size_t)i * nrhs + c0 + q)];
                    hb[(size_t)i * NR + P + q] = B[2 * ((size_t)i * nrhs + c0 + q) + 1];
                }
            DevBuf Bd, Xd, R, D, Y, Ys, nrm;
            for (DevBuf *b : {&Bd, &Xd, &R, &D, &Y, &Ys}) b->reserve(count * 8);
            nrm.reserve(NR * 8);
            PQ_HIP(hipMemcpyAsync(Bd.p, hb.data(), count * 8, hipMemcpyHostToDevice, st));
            std::vector<double> bn(NR), rn(NR);
            auto colnorms = [&](const double *V, std::vector<double> &o) {
                hipLaunchKernelGGL(colnorm2_kernel, dim3(NR), dim3(256), 0, st, V, n, NR, nrm.as<double>());
                PQ_HIP(hipMemcpyAsync(o.data(), nrm.p, NR * 8, hipMemcpyDeviceToHost, st));
                PQ_HIP(hipStreamSynchronize(st));
            };
            auto relres = [&]() {
                double r = 0;
                for (int q = 0; q < cn; ++q)
                    r = std::max(r, std::sqrt((rn[q] + rn[P + q]) / std::max(bn[q] + bn[P + q], 1e-300)));
                return r;
            };
            colnorms(Bd.as<double>(), bn);
            ldl_solve(st, h->f, P, Bd.as<double>(), Y.as<double>(), Ys.as<double>(), Xd.as<double>());
            double last = std::numeric_limits<double>::infinity(), res = 0;
            int corr = 0;
            for (int step = 0;; ++step) {
                hipLaunchKernelGGL(csym_residual_kernel, dim3((n + 255) / 256), dim3(256), 0, st, h->rp.as<int>(),
                                   h->ci.as<int>(), h->av.as<double>(), n, P, Xd.as<double>(), Bd.as<double>(),
                                   R.as<double>());
                PQ_HIP(hipGetLastError());
                colnorms(R.as<double>(), rn);
                res = relres();
                if (!(res > tol) || step == max_steps || res > 0.5 * last) break;
                last = res;
                ldl_solve(st, h->f, P, R.as<double>(), Y.as<double>(), Ys.as<double>(), D.as<double>());
                hipLaunchKernelGGL(add_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, D.as<double>(),
                                   count, Xd.as<double>());
                PQ_HIP(hipGetLastError());
                ++corr;
            }
            worst = std::max(worst, res);
            worst_steps = std::max<int64_t>(worst_steps, corr);
            std::vector<double> hx(count);
            PQ_HIP(hipMemcpyAsync(hx.data(), Xd.p, count * 8, hipMemcpyDeviceToHost, st));
            PQ_HIP(hipStreamSynchronize(st));
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < cn; ++q) {
                    X[2 * ((size_t)i * nrhs + c0 + q)] = hx[(size_t)i * NR + q];
                    X[2 * ((size_t)i * nrhs + c0 + q) + 1] = hx[(size_t)i * NR + P + q];
                }
        }
        if (steps) *steps = worst_steps;
        if (residual) *residual = worst;
        if (worst > std::max(tol, 1e-10))
            throw Error(PQ_ECONVERGE, "pq_csym_solve: the FP32 factor's refinement stalls at relative residual " +
                                          std::to_string(worst));
    });
}

}  // extern "C"
