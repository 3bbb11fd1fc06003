// Kernels of the message-passing network over the branch graph (RWG edges and via segments) that
// refines the coarse-space solution of the loop solve. Dense layers run on the matrix cores:
// gfx950's v_mfma_f32_16x16x32_bf16 (v_mfma_f32_16x16x16_bf16 where the reduction is 16 long),
// operands rounded to BF16 (round to nearest even) as they are staged, products accumulated in
// FP32. Activations live in HBM as FP32 row-major (N, F).
//
// Operand layout (wave64), lane l = 0..63, g = l / 16, q = 0..3:
//   16x16x16: A lane l supplies A[l % 16][4 g + q], B lane l supplies B[4 g + q][l % 16]
//   16x16x32: A lane l supplies A[l % 16][8 g + e], B lane l supplies B[8 g + e][l % 16], e = 0..7
//   D (16 x 16), both:                 lane l holds   D[4 g + q][l % 16]
// A and B take their reduction index from the same (lane group, element) pair, so the product is
// the matrix product whatever order the unit sums that pair's terms in.
// No counterpart in the reference.
#include "pq_common.h"

namespace pq {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef short bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16v8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// D += A B over a reduction of 32: two 8-byte halves per operand (elements 8 g .. 8 g + 7).
__device__ __forceinline__ f32x4 mfma32(const bf16x4 &a0, const bf16x4 &a1, const bf16x4 &b0, const bf16x4 &b1,
                                        const f32x4 &c) {
    const bf16x8 a = {a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
    const bf16x8 b = {b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16v8, a), __builtin_bit_cast(bf16v8, b), c, 0,
                                                   0, 0);
}

// FP32 -> BF16 bits, round to nearest even (NaN kept quiet).
__device__ __forceinline__ unsigned short to_bf16(float x) {
    const unsigned u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (unsigned short)((u >> 16) | 0x40u);
    return (unsigned short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// C (N, M) = act([A1 | A2] (N, K) W (K, M) + bias) (+ res), K = K1 + K2; Zpre (optional) keeps the
// pre-activation. A block covers 64 rows (16 per wave) and all M columns; W is staged once per
// block in LDS as BF16, transposed (M, K) so that each lane's four B values are one 8-byte read.
template <int K1, int K2, int M>
__global__ __launch_bounds__(256) void mpnn_gemm_kernel(int N, const float *__restrict__ A1, int lda1,
                                                       const float *__restrict__ A2, int lda2,
                                                       const float *__restrict__ W, const float *__restrict__ bias,
                                                       const float *__restrict__ res, int ldres, int relu,
                                                       float *__restrict__ C, int ldc, float *__restrict__ Zpre) {
    constexpr int K = K1 + K2, KP = K + 4;   // LDS row pitch (BF16 elements)
    __shared__ unsigned short wt[M * KP];
    for (int e = threadIdx.x; e < K * M; e += 256) {
        const int k = e / M, j = e % M;
        wt[j * KP + k] = to_bf16(W[e]);
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row0 = blockIdx.x * 64 + wave * 16;
    const int arow = min(row0 + (lane & 15), N - 1);
    const int kq = 4 * (lane >> 4);
    f32x4 acc[M / 16];
#pragma unroll
    for (int t = 0; t < M / 16; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto load4 = [&](const float *p) {
        const float4 v = *reinterpret_cast<const float4 *>(p);
        bf16x4 a;
        a[0] = (short)to_bf16(v.x); a[1] = (short)to_bf16(v.y); a[2] = (short)to_bf16(v.z); a[3] = (short)to_bf16(v.w);
        return a;
    };
    if constexpr (K % 32 == 0 && K1 % 32 == 0) {
        const int k8 = 8 * (lane >> 4);
#pragma unroll
        for (int ks = 0; ks < K / 32; ++ks) {
            const int k0 = ks * 32;
            const float *src = k0 < K1 ? A1 + (size_t)arow * lda1 + k0 : A2 + (size_t)arow * lda2 + (k0 - K1);
            const bf16x4 a0 = load4(src + k8), a1 = load4(src + k8 + 4);
#pragma unroll
            for (int t = 0; t < M / 16; ++t) {
                const unsigned short *w = &wt[(t * 16 + (lane & 15)) * KP + k0 + k8];
                acc[t] = mfma32(a0, a1, *reinterpret_cast<const bf16x4 *>(w), *reinterpret_cast<const bf16x4 *>(w + 4),
                                acc[t]);
            }
        }
    } else {
#pragma unroll
        for (int ks = 0; ks < K / 16; ++ks) {
            const int k0 = ks * 16;
            const float *src = k0 < K1 ? A1 + (size_t)arow * lda1 + k0 : A2 + (size_t)arow * lda2 + (k0 - K1);
            const bf16x4 a = load4(src + kq);
#pragma unroll
            for (int t = 0; t < M / 16; ++t) {
                const bf16x4 b = *reinterpret_cast<const bf16x4 *>(&wt[(t * 16 + (lane & 15)) * KP + k0 + kq]);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < M / 16; ++t) {
        const int j = t * 16 + (lane & 15);
        const float bj = bias ? bias[j] : 0.f;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int r = row0 + kq + q;
            if (r >= N) continue;
            const float z = acc[t][q] + bj;
            if (Zpre) Zpre[(size_t)r * M + j] = z;
            float y = relu ? fmaxf(z, 0.f) : z;
            if (res) y += res[(size_t)r * ldres + j];
            C[(size_t)r * ldc + j] = y;
        }
    }
}

// Weight gradient partials: part[blk] = ([A1 | A2]^T G over the block's rows (K, M), then the
// column sums of G (M)). Rows are staged 64 at a time in LDS as BF16, both operands transposed
// (k or j major) so that each lane's four values along the reduction are one 8-byte read; the
// (K/16) x (M/16) output tiles are spread over the four waves. The column sums use the FP32 G.
constexpr int TN_ROWS = 64, TN_PITCH = TN_ROWS + 4;
template <int K1, int K2, int M>
__global__ __launch_bounds__(256) void mpnn_gemm_tn_kernel(int N, int chunk, const float *__restrict__ A1, int lda1,
                                                          const float *__restrict__ A2, int lda2,
                                                          const float *__restrict__ G, int ldg,
                                                          float *__restrict__ part) {
    constexpr int K = K1 + K2, TK = K / 16, TM = M / 16, T = TK * TM, TPW = (T + 3) / 4;
    __shared__ unsigned short at[K * TN_PITCH];
    __shared__ unsigned short gt[M * TN_PITCH];
    __shared__ float colsum[256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n0 = blockIdx.x * chunk, n1 = min(N, n0 + chunk);
    f32x4 acc[TPW];
#pragma unroll
    for (int t = 0; t < TPW; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    float csum = 0.f;                         // column threadIdx.x % M of G (M divides 256)
    const int kq = 4 * (lane >> 4);
    for (int s0 = n0; s0 < n1; s0 += TN_ROWS) {
        const int rows = min(TN_ROWS, n1 - s0);
        for (int e = threadIdx.x; e < TN_ROWS * K; e += 256) {
            const int r = e / K, k = e % K;
            float v = 0.f;
            if (r < rows) v = k < K1 ? A1[(size_t)(s0 + r) * lda1 + k] : A2[(size_t)(s0 + r) * lda2 + (k - K1)];
            at[k * TN_PITCH + r] = to_bf16(v);
        }
        for (int e = threadIdx.x; e < TN_ROWS * M; e += 256) {
            const int r = e / M, j = e % M;
            float v = 0.f;
            if (r < rows) v = G[(size_t)(s0 + r) * ldg + j];
            csum += v;
            gt[j * TN_PITCH + r] = to_bf16(v);
        }
        __syncthreads();
        const int k8 = 8 * (lane >> 4);
#pragma unroll
        for (int ns = 0; ns < TN_ROWS / 32; ++ns) {
#pragma unroll
            for (int t = 0; t < TPW; ++t) {
                const int tile = wave * TPW + t;
                if (tile >= T) break;
                const int kt = tile / TM, mt = tile % TM;
                const unsigned short *pa = &at[(kt * 16 + (lane & 15)) * TN_PITCH + ns * 32 + k8];
                const unsigned short *pg = &gt[(mt * 16 + (lane & 15)) * TN_PITCH + ns * 32 + k8];
                acc[t] = mfma32(*reinterpret_cast<const bf16x4 *>(pa), *reinterpret_cast<const bf16x4 *>(pa + 4),
                                *reinterpret_cast<const bf16x4 *>(pg), *reinterpret_cast<const bf16x4 *>(pg + 4), acc[t]);
            }
        }
        __syncthreads();
    }
    float *o = part + (size_t)blockIdx.x * (K * M + M);
#pragma unroll
    for (int t = 0; t < TPW; ++t) {
        const int tile = wave * TPW + t;
        if (tile >= T) break;
        const int kt = tile / TM, mt = tile % TM;
#pragma unroll
        for (int q = 0; q < 4; ++q) o[(size_t)(kt * 16 + kq + q) * M + mt * 16 + (lane & 15)] = acc[t][q];
    }
    colsum[threadIdx.x] = csum;
    __syncthreads();
    if (threadIdx.x < M) {
        float s = 0.f;
        for (int t = threadIdx.x; t < 256; t += M) s += colsum[t];
        o[K * M + threadIdx.x] = s;
    }
}

// out[e] = sum over the nblk partials of element e, in block order.
// Sum of the nblk partials (rows of width) in two fixed-order passes: slice y of gridDim.y sums
// partials y, y + S, y + 2S, ... into slice[y]; then the slices are summed in order. The same result
// for every schedule.
__global__ void mpnn_reduce_kernel(const float *__restrict__ part, int nblk, int width, float *__restrict__ slice) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= width) return;
    float s = 0.f;
    for (int b = blockIdx.y; b < nblk; b += gridDim.y) s += part[(size_t)b * width + e];
    slice[(size_t)blockIdx.y * width + e] = s;
}

__global__ void mpnn_reduce_slices_kernel(const float *__restrict__ slice, int slices, int width, float *__restrict__ out) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= width) return;
    float s = 0.f;
    for (int y = 0; y < slices; ++y) s += slice[(size_t)y * width + e];
    out[e] = s;
}

// out[i][f] = base[i][f] + base2[i][f] + sum_t w[t] in[col[t]][f] over the CSR row i (64 features;
// base, base2 optional): 16 lanes per node, four features (one float4) each.
__global__ __launch_bounds__(256) void mpnn_aggregate_kernel(int N, const int *__restrict__ ptr,
                                                            const int *__restrict__ col, const float *__restrict__ w,
                                                            const float *__restrict__ in, int ldi,
                                                            const float *__restrict__ base, int ldb,
                                                            const float *__restrict__ base2, int ldb2,
                                                            float *__restrict__ out, int ldo) {
    const int node = blockIdx.x * 16 + (threadIdx.x >> 4);
    const int f = 4 * (threadIdx.x & 15);
    if (node >= N) return;
    float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int t = ptr[node]; t < ptr[node + 1]; ++t) {
        const float wt = w[t];
        const float4 v = *reinterpret_cast<const float4 *>(in + (size_t)col[t] * ldi + f);
        s.x += wt * v.x; s.y += wt * v.y; s.z += wt * v.z; s.w += wt * v.w;
    }
    if (base) {
        const float4 v = *reinterpret_cast<const float4 *>(base + (size_t)node * ldb + f);
        s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
    if (base2) {
        const float4 v = *reinterpret_cast<const float4 *>(base2 + (size_t)node * ldb2 + f);
        s.x += v.x; s.y += v.y; s.z += v.z; s.w += v.w;
    }
    *reinterpret_cast<float4 *>(out + (size_t)node * ldo + f) = s;
}

// out = g where z > 0, else 0 (the ReLU's backward).
__global__ void mpnn_relu_mask_kernel(size_t count, const float *__restrict__ z, const float *__restrict__ g,
                                      float *__restrict__ out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) out[e] = z[e] > 0.f ? g[e] : 0.f;
}

}  // namespace pq

using namespace pq;

namespace {

// The (K1, K2, M) shapes the network uses: input layer (16, 0, 64) (also the output layer's input
// gradient), message-passing layers (64, 64, 64), their input gradient (64, 0, 128) and the output
// layer (64, 0, 16).
template <class F> void dispatch_shape(int k1, int k2, int m, F &&f) {
#define PQ_SHAPE(a, b, c) \
    if (k1 == a && k2 == b && m == c) { f(std::integral_constant<int, a>(), std::integral_constant<int, b>(), std::integral_constant<int, c>()); return; }
    PQ_SHAPE(16, 0, 64)
    PQ_SHAPE(64, 64, 64)
    PQ_SHAPE(64, 0, 128)
    PQ_SHAPE(64, 0, 16)
#undef PQ_SHAPE
    throw Error(PQ_EINVAL, "mpnn: unsupported layer shape");
}

}  // namespace

extern "C" {

int pq_mpnn_gemm(int N, const float *A1, int lda1, int k1, const float *A2, int lda2, int k2, const float *W, int m,
                 const float *bias, const float *res, int ldres, int relu, float *C, int ldc, float *Zpre,
                 void *stream) {
    return guard([&] {
        PQ_REQUIRE(N >= 0 && A1 && W && C && (k2 == 0 || A2), "pq_mpnn_gemm: null argument");
        PQ_REQUIRE(lda1 >= k1 && (k2 == 0 || lda2 >= k2) && ldc >= m && (!res || ldres >= m),
                   "pq_mpnn_gemm: leading dimension too small");
        PQ_REQUIRE(lda1 % 4 == 0 && (k2 == 0 || lda2 % 4 == 0), "pq_mpnn_gemm: leading dimensions must be multiples of 4");
        if (N == 0) return;
        dispatch_shape(k1, k2, m, [&](auto a, auto b, auto c) {
            constexpr int K1 = decltype(a)::value, K2 = decltype(b)::value, M = decltype(c)::value;
            hipLaunchKernelGGL((mpnn_gemm_kernel<K1, K2, M>), dim3((N + 63) / 64), dim3(256), 0, (hipStream_t)stream, N,
                               A1, lda1, A2, lda2, W, bias, res, ldres, relu, C, ldc, Zpre);
        });
        PQ_HIP(hipGetLastError());
    });
}

// Weight-gradient blocks: 128 rows each (two staged steps), at most 1024: a 30k-node graph gets
// 235 workgroups (1,024-row blocks left 30 of the 256 CUs busy: 281 us per layer).
static int tn_blocks(int N) { return std::max(1, std::min(1024, (N + 127) / 128)); }
constexpr int TN_SLICES = 16;                 // first-pass slices of the partials' sum

int pq_mpnn_workspace(int N, int k, int m, size_t *floats) {
    return guard([&] {
        PQ_REQUIRE(N > 0 && k > 0 && m > 0 && floats, "pq_mpnn_workspace: invalid argument");
        *floats = (size_t)(tn_blocks(N) + 1 + TN_SLICES) * ((size_t)k * m + m);
    });
}

int pq_mpnn_gemm_tn(int N, const float *A1, int lda1, int k1, const float *A2, int lda2, int k2, const float *G,
                    int ldg, int m, float *work, size_t work_floats, float *dW, float *db, void *stream) {
    return guard([&] {
        PQ_REQUIRE(N > 0 && A1 && G && work && dW && db && (k2 == 0 || A2), "pq_mpnn_gemm_tn: null argument or N = 0");
        const int nblk = tn_blocks(N);
        const int K = k1 + k2;
        const size_t width = (size_t)K * m + m;
        PQ_REQUIRE(work_floats >= (size_t)(nblk + 1 + TN_SLICES) * width, "pq_mpnn_gemm_tn: workspace too small");
        int chunk = (N + nblk - 1) / nblk;
        chunk = (chunk + TN_ROWS - 1) / TN_ROWS * TN_ROWS;
        dispatch_shape(k1, k2, m, [&](auto a, auto b, auto c) {
            constexpr int K1 = decltype(a)::value, K2 = decltype(b)::value, M = decltype(c)::value;
            hipLaunchKernelGGL((mpnn_gemm_tn_kernel<K1, K2, M>), dim3(nblk), dim3(256), 0, (hipStream_t)stream, N, chunk,
                               A1, lda1, A2, lda2, G, ldg, work);
        });
        // dW then db: the partial of each block is (K * M weights, M column sums)
        const int slices = std::min(TN_SLICES, nblk);
        float *slice = work + (size_t)(nblk + 1) * width;
        hipLaunchKernelGGL(mpnn_reduce_kernel, dim3((unsigned)((width + 255) / 256), (unsigned)slices), dim3(256), 0,
                           (hipStream_t)stream, work, nblk, (int)width, slice);
        hipLaunchKernelGGL(mpnn_reduce_slices_kernel, dim3((unsigned)((width + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, slice, slices, (int)width, work + (size_t)nblk * width);
        PQ_HIP(hipMemcpyAsync(dW, work + (size_t)nblk * width, (size_t)K * m * 4, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
        PQ_HIP(hipMemcpyAsync(db, work + (size_t)nblk * width + (size_t)K * m, (size_t)m * 4, hipMemcpyDeviceToDevice,
                              (hipStream_t)stream));
        PQ_HIP(hipGetLastError());
    });
}

int pq_mpnn_aggregate(int N, const int32_t *ptr, const int32_t *col, const float *w, const float *in, int ldi,
                      const float *base, int ldb, const float *base2, int ldb2, float *out, int ldo, void *stream) {
    return guard([&] {
        PQ_REQUIRE(N >= 0 && ptr && col && w && in && out, "pq_mpnn_aggregate: null argument");
        PQ_REQUIRE(ldi % 4 == 0 && ldo % 4 == 0 && (!base || ldb % 4 == 0) && (!base2 || ldb2 % 4 == 0),
                   "pq_mpnn_aggregate: leading dimensions must be multiples of 4");
        if (N == 0) return;
        hipLaunchKernelGGL(mpnn_aggregate_kernel, dim3((N + 15) / 16), dim3(256), 0, (hipStream_t)stream, N, ptr, col, w,
                           in, ldi, base, ldb, base2, ldb2, out, ldo);
        PQ_HIP(hipGetLastError());
    });
}

int pq_mpnn_relu_mask(size_t count, const float *z, const float *g, float *out, void *stream) {
    return guard([&] {
        PQ_REQUIRE(z && g && out, "pq_mpnn_relu_mask: null argument");
        if (count == 0) return;
        hipLaunchKernelGGL(mpnn_relu_mask_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                           (hipStream_t)stream, count, z, g, out);
        PQ_HIP(hipGetLastError());
    });
}

}  // extern "C"
