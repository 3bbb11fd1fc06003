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
#include <type_traits>

#include "pq_common.h"
#include "ldl.h"

namespace pq {
namespace {

constexpr int LB = LDL_BLOCK;
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

// 1 / (dr + j di), scaled so that neither the square nor the quotient leaves FP32's range early.
__device__ inline void cinv(float dr, float di, float &qr, float &qi) {
    const float s = fmaxf(fabsf(dr), fabsf(di));
    const float er = dr / s, ei = di / s, den = s * (er * er + ei * ei);
    qr = er / den;
    qi = -ei / den;
}

// The b x b diagonal block at (k, k) factored in place (L below the diagonal, D on it), and the
// inverse X = L^-1 of its unit lower factor written to Lir / Lii (64 x 64, the identity beyond b).
// Thread t holds row t >> 2 at columns (t & 3) + 4 s, s < 16, of the block and of X in registers.
// Step j: the pivot and row j of X go through LDS, then column j of L (l_ij = a_ij / d_j), then
// every row i > j updates a_ic -= l_ij a_cj (j < c <= i, a_cj the column before scaling) and
// x_ic -= l_ij x_jc (c <= j). A zero or non-finite pivot sets *bad.
__global__ __launch_bounds__(256) void ldl_diag_kernel(float *Fr, float *Fi, long long ldf, int k, int b, float *Lir,
                                                      float *Lii, int *bad) {
    __shared__ float pv[2], lr[LB], li[LB], cr[LB], ci[LB];
    __shared__ float rr[2][LB], ri[2][LB];
    const int t = threadIdx.x, i = t >> 2, c0 = t & 3;
    const bool live = i < b;
    float *rowr = Fr + (size_t)(k + (live ? i : 0)) * ldf + k, *rowi = Fi + (size_t)(k + (live ? i : 0)) * ldf + k;
    float ar[16], ai[16], xr[16], xi[16];
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int c = c0 + 4 * s;
        const bool in = live && c <= i;
        ar[s] = in ? rowr[c] : 0.f;
        ai[s] = in ? rowi[c] : 0.f;
        xr[s] = c == i ? 1.f : 0.f;
        xi[s] = 0.f;
    }
    for (int j = 0; j < b; ++j) {
        const int sj = j >> 2, buf = j & 1;
        const bool colj = c0 == (j & 3);
        if (i == j) {
            if (colj) {
                float vr = 0.f, vi = 0.f;
#pragma unroll
                for (int s = 0; s < 16; ++s)
                    if (s == sj) { vr = ar[s]; vi = ai[s]; }
                pv[0] = vr;
                pv[1] = vi;
            }
#pragma unroll
            for (int s = 0; s < 16; ++s) { rr[buf][c0 + 4 * s] = xr[s]; ri[buf][c0 + 4 * s] = xi[s]; }
        }
        __syncthreads();
        const float dr = pv[0], di = pv[1];
        if (t == 0 && !(isfinite(dr) && isfinite(di) && (dr != 0.f || di != 0.f))) *bad = 1;
        if (live && i > j && colj) {
            float vr = 0.f, vi = 0.f;
#pragma unroll
            for (int s = 0; s < 16; ++s)
                if (s == sj) { vr = ar[s]; vi = ai[s]; }
            float qr, qi;
            cinv(dr, di, qr, qi);
            const float l_r = vr * qr - vi * qi, l_i = vr * qi + vi * qr;
#pragma unroll
            for (int s = 0; s < 16; ++s)
                if (s == sj) { ar[s] = l_r; ai[s] = l_i; }
            lr[i] = l_r; li[i] = l_i;
            cr[i] = vr; ci[i] = vi;
        }
        __syncthreads();
        if (live && i > j) {
            const float l_r = lr[i], l_i = li[i];
#pragma unroll
            for (int s = 0; s < 16; ++s) {
                const int c = c0 + 4 * s;
                if (c > j && c <= i) {
                    ar[s] -= l_r * cr[c] - l_i * ci[c];
                    ai[s] -= l_r * ci[c] + l_i * cr[c];
                }
                if (c <= j) {
                    const float ur = rr[buf][c], ui = ri[buf][c];
                    xr[s] -= l_r * ur - l_i * ui;
                    xi[s] -= l_r * ui + l_i * ur;
                }
            }
        }
    }
#pragma unroll
    for (int s = 0; s < 16; ++s) {
        const int c = c0 + 4 * s;
        if (live && c <= i) { rowr[c] = ar[s]; rowi[c] = ai[s]; }
        Lir[i * LB + c] = xr[s];
        Lii[i * LB + c] = xi[s];
    }
}

// One wave's 32 x 32 part of a 64 x 64 complex tile product P[i][c] = sum_t A[i][t] B[c][t]
// (t < 64): waves 2 x 2 over the tile, each 2 x 2 MFMA 16x16x4 FP32 blocks for the real and the
// imaginary part. The K dimension is staged through LDS in two chunks of 32; in MFMA step s of a
// chunk the four lane groups g = lane >> 4 supply k = 8 g + s to A and B alike, so each lane reads
// its eight k values of a row as two 16-byte LDS loads. Rows at or past arows / brows read zero.
struct TileAcc {
    f32x4 r[2][2], i[2][2];
};

__device__ __attribute__((always_inline)) inline void tile_abt(const float *__restrict__ Ar, const float *__restrict__ Ai, long long lda, int arows,
                                const float *__restrict__ Br, const float *__restrict__ Bi, long long ldb, int brows,
                                float (*sA)[LB][TP], float (*sB)[LB][TP], TileAcc &acc) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1, g = lane >> 4, l16 = lane & 15;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            acc.r[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
            acc.i[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
        }
    const float4 zero = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int kc = 0; kc < LB; kc += 32) {
        for (int e = tid; e < LB * 8; e += 256) {
            // rows past the end load the last valid row (arows, brows >= 1) and are zeroed by value
            const int r = e >> 3, q = (e & 7) * 4, ra = min(r, arows - 1), rb = min(r, brows - 1);
            float4 a0 = *reinterpret_cast<const float4 *>(Ar + (size_t)ra * lda + kc + q);
            float4 a1 = *reinterpret_cast<const float4 *>(Ai + (size_t)ra * lda + kc + q);
            float4 b0 = *reinterpret_cast<const float4 *>(Br + (size_t)rb * ldb + kc + q);
            float4 b1 = *reinterpret_cast<const float4 *>(Bi + (size_t)rb * ldb + kc + q);
            if (r >= arows) a0 = a1 = zero;
            if (r >= brows) b0 = b1 = zero;
            *reinterpret_cast<float4 *>(&sA[0][r][q]) = a0;
            *reinterpret_cast<float4 *>(&sA[1][r][q]) = a1;
            *reinterpret_cast<float4 *>(&sB[0][r][q]) = b0;
            *reinterpret_cast<float4 *>(&sB[1][r][q]) = b1;
        }
        __syncthreads();
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
                                                       const float *Lir, const float *Lii, float *Wr, float *Wi) {
    __shared__ float sA[2][LB][TP], sB[2][LB][TP];
    __shared__ float qr[LB], qi[LB];
    const int I0 = k + LB + LB * blockIdx.x;
    if (threadIdx.x < LB) {
        const size_t dg = (size_t)(k + threadIdx.x) * ldf + k + threadIdx.x;
        cinv(Fr[dg], Fi[dg], qr[threadIdx.x], qi[threadIdx.x]);
    }
    TileAcc acc;
    tile_abt(Fr + (size_t)I0 * ldf + k, Fi + (size_t)I0 * ldf + k, ldf, min(LB, n - I0), Lir, Lii, LB, LB, sA, sB, acc);
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
                Wr[(size_t)row * LB + col] = vr;
                Wi[(size_t)row * LB + col] = vi;
                Fr[(size_t)row * ldf + k + col] = vr * qr[col] - vi * qi[col];
                Fi[(size_t)row * ldf + k + col] = vr * qi[col] + vi * qr[col];
            }
}

// A22 -= L21 W^T on the lower triangle below panel k: one 64 x 64 tile (ti, tj), tj <= ti, per
// workgroup, blockIdx.x = ti (ti + 1) / 2 + tj.
__global__ __launch_bounds__(256) void ldl_update_kernel(float *Fr, float *Fi, long long ldf, int n, int k,
                                                        const float *Wr, const float *Wi) {
    __shared__ float sA[2][LB][TP], sB[2][LB][TP];
    const int id = blockIdx.x;
    int ti = (int)((sqrtf(8.f * (float)id + 1.f) - 1.f) * 0.5f);
    while ((ti + 1) * (ti + 2) / 2 <= id) ++ti;
    while (ti * (ti + 1) / 2 > id) --ti;
    const int tj = id - ti * (ti + 1) / 2;
    const int m0 = k + LB, I0 = m0 + LB * ti, C0 = m0 + LB * tj;
    TileAcc acc;
    tile_abt(Fr + (size_t)I0 * ldf + k, Fi + (size_t)I0 * ldf + k, ldf, min(LB, n - I0), Wr + (size_t)C0 * LB,
             Wi + (size_t)C0 * LB, LB, min(LB, n - C0), sA, sB, acc);
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

// Forward step over block k0 .. k0 + bk: every workgroup forms y_k = X_k b_k (X_k the stored
// inverse of L_kk), workgroup 0 writes it to Ys, and each workgroup subtracts L[r, block] y_k from
// its rows r >= k0 + bk of Y (a thread per row).
template <int P>
__global__ __launch_bounds__(256) void ldl_forward_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                         long long ldf, const float *__restrict__ Xr,
                                                         const float *__restrict__ Xi, int n, int k0, int bk, double *Y,
                                                         double *Ys) {
    constexpr int NR = 2 * P;
    __shared__ double bs[LB][NR], ys[LB][NR];
    for (int e = threadIdx.x; e < LB * NR; e += 256) {
        const int r = e / NR, q = e % NR;
        bs[r][q] = r < bk ? Y[(size_t)(k0 + r) * NR + q] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < LB * P; e += 256) {
        const int i = e / P, p = e % P;
        double sr = 0, si = 0;
        for (int t = 0; t <= i; ++t) {
            const double lr = Xr[i * LB + t], li = Xi[i * LB + t];
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
    for (int r = k0 + bk + blockIdx.x * 256 + threadIdx.x; r < n; r += gridDim.x * 256) {
        double ar[P], ai[P];
#pragma unroll
        for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
        const float4 *lr4 = reinterpret_cast<const float4 *>(Fr + (size_t)r * ldf + k0);
        const float4 *li4 = reinterpret_cast<const float4 *>(Fi + (size_t)r * ldf + k0);
#pragma unroll 2
        for (int t4 = 0; t4 < LB / 4; ++t4) {
            const float4 a = lr4[t4], b = li4[t4];
            const float lre[4] = {a.x, a.y, a.z, a.w}, lim[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = 4 * t4 + u;
                const double xr = lre[u], xi = lim[u];
#pragma unroll
                for (int p = 0; p < P; ++p) {
                    ar[p] += xr * ys[t][p] - xi * ys[t][P + p];
                    ai[p] += xr * ys[t][P + p] + xi * ys[t][p];
                }
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            Y[(size_t)r * NR + p] -= ar[p];
            Y[(size_t)r * NR + P + p] -= ai[p];
        }
    }
}

// z = D^-1 y (FP64 complex division by the FP32 pivots).
__global__ void ldl_scale_kernel(const float *Fr, const float *Fi, long long ldf, int n, int P, const double *Ys,
                                 double *Z) {
    const int NR = 2 * P;
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)n * P) return;
    const int r = (int)(e / P), p = (int)(e % P);
    const double dr = Fr[(size_t)r * ldf + r], di = Fi[(size_t)r * ldf + r];
    const double yr = Ys[(size_t)r * NR + p], yi = Ys[(size_t)r * NR + P + p];
    const double den = dr * dr + di * di;
    Z[(size_t)r * NR + p] = (yr * dr + yi * di) / den;
    Z[(size_t)r * NR + P + p] = (yi * dr - yr * di) / den;
}

// Backward step over block k0 .. k0 + bk: x_k = X_k^T z_k (workgroup 0 writes it to Xo), then each
// workgroup subtracts sum_t L[k0 + t][r] x_k[t] from its rows r < k0 of Z (a thread per row, the
// block's rows read across threads).
template <int P>
__global__ __launch_bounds__(256) void ldl_backward_kernel(const float *__restrict__ Fr, const float *__restrict__ Fi,
                                                          long long ldf, const float *__restrict__ Xr,
                                                          const float *__restrict__ Xi, int k0, int bk, double *Z,
                                                          double *Xo) {
    constexpr int NR = 2 * P;
    __shared__ double zs[LB][NR], xs[LB][NR];
    for (int e = threadIdx.x; e < LB * NR; e += 256) {
        const int r = e / NR, q = e % NR;
        zs[r][q] = r < bk ? Z[(size_t)(k0 + r) * NR + q] : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < LB * P; e += 256) {
        const int c = e / P, p = e % P;
        double sr = 0, si = 0;
        for (int t = c; t < bk; ++t) {
            const double lr = Xr[t * LB + c], li = Xi[t * LB + c];
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
    for (int r = blockIdx.x * 256 + threadIdx.x; r < k0; r += gridDim.x * 256) {
        double ar[P], ai[P];
#pragma unroll
        for (int p = 0; p < P; ++p) ar[p] = ai[p] = 0;
        for (int t = 0; t < bk; ++t) {
            const double xr = Fr[(size_t)(k0 + t) * ldf + r], xi = Fi[(size_t)(k0 + t) * ldf + r];
#pragma unroll
            for (int p = 0; p < P; ++p) {
                ar[p] += xr * xs[t][p] - xi * xs[t][P + p];
                ai[p] += xr * xs[t][P + p] + xi * xs[t][p];
            }
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            Z[(size_t)r * NR + p] -= ar[p];
            Z[(size_t)r * NR + P + p] -= ai[p];
        }
    }
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

}  // namespace

void ldl_form(hipStream_t st, const LdlFactor &f, const float *Ll, long long ldl, const int *rptr, const int *rcol,
              const double *rval, double omega) {
    const int n = f.n;
    if (n <= 0) return;
    hipLaunchKernelGGL(ldl_form_kernel, dim3((n + 255) / 256, std::min(n, 65535)), dim3(256), 0, st, Ll, ldl, n,
                       (float)omega, f.Fr, f.Fi, f.ldf);
    hipLaunchKernelGGL(ldl_form_r_kernel, dim3((n + 255) / 256), dim3(256), 0, st, rptr, rcol, rval, n, f.Fr, f.ldf);
    PQ_HIP(hipGetLastError());
}

void ldl_factor(hipStream_t st, const LdlFactor &f, int *bad) {
    const int n = f.n;
    for (int k = 0; k < n; k += LB) {
        const int b = std::min(LB, n - k);
        float *xr = f.Lir + (size_t)(k / LB) * LB * LB, *xi = f.Lii + (size_t)(k / LB) * LB * LB;
        hipLaunchKernelGGL(ldl_diag_kernel, dim3(1), dim3(256), 0, st, f.Fr, f.Fi, f.ldf, k, b, xr, xi, bad);
        if (k + b < n) {
            const int tiles = (n - k - LB + LB - 1) / LB;
            hipLaunchKernelGGL(ldl_panel_kernel, dim3(tiles), dim3(256), 0, st, f.Fr, f.Fi, f.ldf, n, k, xr, xi, f.Wr,
                               f.Wi);
            hipLaunchKernelGGL(ldl_update_kernel, dim3((unsigned)((size_t)tiles * (tiles + 1) / 2)), dim3(256), 0, st,
                               f.Fr, f.Fi, f.ldf, n, k, f.Wr, f.Wi);
        }
    }
    PQ_HIP(hipGetLastError());
}

void ldl_solve(hipStream_t st, const LdlFactor &f, int P, const double *B, double *Y, double *Ys, double *X) {
    const int n = f.n, nb = (n + LB - 1) / LB;
    if (n <= 0) return;
    const size_t bytes = (size_t)n * 2 * P * 8;
    PQ_HIP(hipMemcpyAsync(Y, B, bytes, hipMemcpyDeviceToDevice, st));
    dispatch_ports(P, [&](auto pc) {
        constexpr int PP = decltype(pc)::value;
        for (int kb = 0; kb < nb; ++kb) {
            const int k0 = kb * LB, bk = std::min(LB, n - k0);
            const int rows = n - k0 - bk, grid = std::max(1, (rows + 255) / 256);
            hipLaunchKernelGGL(ldl_forward_kernel<PP>, dim3(grid), dim3(256), 0, st, f.Fr, f.Fi, f.ldf,
                               f.Lir + (size_t)kb * LB * LB, f.Lii + (size_t)kb * LB * LB, n, k0, bk, Y, Ys);
        }
        hipLaunchKernelGGL(ldl_scale_kernel, dim3((unsigned)(((size_t)n * PP + 255) / 256)), dim3(256), 0, st, f.Fr,
                           f.Fi, f.ldf, n, PP, Ys, Y);
        for (int kb = nb - 1; kb >= 0; --kb) {
            const int k0 = kb * LB, bk = std::min(LB, n - k0), grid = std::max(1, (k0 + 255) / 256);
            hipLaunchKernelGGL(ldl_backward_kernel<PP>, dim3(grid), dim3(256), 0, st, f.Fr, f.Fi, f.ldf,
                               f.Lir + (size_t)kb * LB * LB, f.Lii + (size_t)kb * LB * LB, k0, bk, Y, X);
        }
    });
    PQ_HIP(hipGetLastError());
}

}  // namespace pq
