// Biot-Savart field of sheet currents and straight filaments at observation points.
//
// Semantics: src/pcbqs/biot.py:24-51 (_h_tris: H = (1 / 4 pi) sum_s w_s J_s x R / |R|^3 over the
// 6-point rule of every triangle, a term dropped where |R|^3 <= 1e-30), biot.py:54-79 (h_sheets:
// the near triangles split 4x twice, which the host does while building the source list) and
// biot.py:82-105 (h_segments: the finite straight filament, H = I (c1 - c2) / (4 pi d) along
// u x perp/d, dropped where d <= 1e-9).
//
// Layout: a sheet source is (x, y, z, then per excitation k: Jx re, Jx im, Jy re, Jy im), the
// current already multiplied by its quadrature weight (area w_q); a segment is (p0, p1, then I re,
// I im per k). H is (npts, 3, K, 2). Each block stages a tile of sources in LDS and every thread
// accumulates two observation points for all K excitations in registers (an N-body sum; VALU
// bound).
#include "pq_common.h"

#include <hipfft/hipfft.h>
#include <vector>

#include <chrono>

namespace pq {

constexpr int BIOT_PPT = 2;   // observation points per thread

template <typename T> __device__ __forceinline__ T inv_cube_root3(T r2);
template <> __device__ __forceinline__ float inv_cube_root3<float>(float r2) {
    const float rs = rsqrtf(r2);
    return rs * rs * rs;
}
template <> __device__ __forceinline__ double inv_cube_root3<double>(double r2) {
    return 1.0 / (r2 * sqrt(r2));
}

// One block: 256 x BIOT_PPT points and the sources [chunk * cs, (chunk + 1) * cs) of grid row
// blockIdx.y; the block's sums go to H + blockIdx.y * (np * 3 * K * 2) (the partial of that chunk).
template <typename T, int K, int TILE>
__global__ __launch_bounds__(256) void biot_sheet_kernel(int ns, int cs, const T *__restrict__ src, int np,
                                                        const T *__restrict__ pts, T *__restrict__ H) {
    constexpr int S = 3 + 4 * K;
    __shared__ T tile[TILE * S];
    const int c0 = blockIdx.y * cs, c1 = min(ns, c0 + cs);
    H += (size_t)blockIdx.y * np * 3 * K * 2;
    T px[BIOT_PPT], py[BIOT_PPT], pz[BIOT_PPT];
    T h[BIOT_PPT][3][K][2];
    const int p0 = (blockIdx.x * 256 + threadIdx.x) * BIOT_PPT;
#pragma unroll
    for (int q = 0; q < BIOT_PPT; ++q) {
        const int p = min(p0 + q, np - 1);
        px[q] = pts[3 * p]; py[q] = pts[3 * p + 1]; pz[q] = pts[3 * p + 2];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < K; ++k) h[q][c][k][0] = h[q][c][k][1] = 0;
    }
    for (int s0 = c0; s0 < c1; s0 += TILE) {
        const int m = min(TILE, c1 - s0);
        __syncthreads();
        for (int e = threadIdx.x; e < m * S; e += 256) tile[e] = src[(size_t)s0 * S + e];
        __syncthreads();
        for (int j = 0; j < m; ++j) {
            const T *sj = tile + j * S;
            const T sx = sj[0], sy = sj[1], sz = sj[2];
#pragma unroll
            for (int q = 0; q < BIOT_PPT; ++q) {
                const T rx = px[q] - sx, ry = py[q] - sy, rz = pz[q] - sz;
                const T r2 = rx * rx + ry * ry + rz * rz;
                const T inv = r2 > (T)1e-20 ? inv_cube_root3<T>(r2) : (T)0;     // |R|^3 > 1e-30
                const T a = rz * inv, b = ry * inv, c = rx * inv;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const T jxr = sj[3 + 4 * k], jxi = sj[4 + 4 * k], jyr = sj[5 + 4 * k], jyi = sj[6 + 4 * k];
                    h[q][0][k][0] += jyr * a; h[q][0][k][1] += jyi * a;
                    h[q][1][k][0] -= jxr * a; h[q][1][k][1] -= jxi * a;
                    h[q][2][k][0] += jxr * b - jyr * c;
                    h[q][2][k][1] += jxi * b - jyi * c;
                }
            }
        }
    }
    const T f = (T)(1.0 / (4.0 * 3.14159265358979323846));
#pragma unroll
    for (int q = 0; q < BIOT_PPT; ++q) {
        const int p = p0 + q;
        if (p >= np) continue;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int k = 0; k < K; ++k) {
                H[(((size_t)p * 3 + c) * K + k) * 2] = f * h[q][c][k][0];
                H[(((size_t)p * 3 + c) * K + k) * 2 + 1] = f * h[q][c][k][1];
            }
    }
}

// H += the field of nseg finite straight filaments, one thread per point.
template <typename T, int K>
__global__ __launch_bounds__(256) void biot_segment_kernel(int nseg, const T *__restrict__ seg, int np,
                                                          const T *__restrict__ pts, T *__restrict__ H) {
    constexpr int S = 6 + 2 * K;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= np) return;
    const T px = pts[3 * p], py = pts[3 * p + 1], pz = pts[3 * p + 2];
    T h[3][K][2];
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < K; ++k) h[c][k][0] = h[c][k][1] = 0;
    for (int g = 0; g < nseg; ++g) {
        const T *sg = seg + (size_t)g * S;
        T ux = sg[3] - sg[0], uy = sg[4] - sg[1], uz = sg[5] - sg[2];
        const T len = sqrt(ux * ux + uy * uy + uz * uz);
        ux /= len; uy /= len; uz /= len;
        const T r0x = px - sg[0], r0y = py - sg[1], r0z = pz - sg[2];
        const T s = r0x * ux + r0y * uy + r0z * uz;
        const T qx = r0x - s * ux, qy = r0y - s * uy, qz = r0z - s * uz;
        const T d = sqrt(qx * qx + qy * qy + qz * qz);
        if (!(d > (T)1e-9)) continue;
        const T c1 = s / sqrt(s * s + d * d), c2 = (s - len) / sqrt((s - len) * (s - len) + d * d);
        const T gain = (c1 - c2) / ((T)(4.0 * 3.14159265358979323846) * d);
        const T nx = qx / d, ny = qy / d, nz = qz / d;
        const T dx = (uy * nz - uz * ny) * gain, dy = (uz * nx - ux * nz) * gain, dz = (ux * ny - uy * nx) * gain;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const T ir = sg[6 + 2 * k], ii = sg[7 + 2 * k];
            h[0][k][0] += dx * ir; h[0][k][1] += dx * ii;
            h[1][k][0] += dy * ir; h[1][k][1] += dy * ii;
            h[2][k][0] += dz * ir; h[2][k][1] += dz * ii;
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int k = 0; k < K; ++k) {
            H[(((size_t)p * 3 + c) * K + k) * 2] += h[c][k][0];
            H[(((size_t)p * 3 + c) * K + k) * 2 + 1] += h[c][k][1];
        }
}

// out[e] = sum over the chunks of part[c][e], in chunk order, as FP64.
template <typename T> __global__ void chunk_sum_kernel(const T *part, int chunks, size_t n, double *out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    T s = 0;
    for (int c = 0; c < chunks; ++c) s += part[(size_t)c * n + e];
    out[e] = (double)s;
}

template <typename T> __global__ void to_precision_kernel(const double *in, size_t n, T *out) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = (T)in[e];
}

}  // namespace pq

using namespace pq;

namespace {

template <class F> void dispatch_k(int K, F &&f) {
    switch (K) {
    case 1: f(std::integral_constant<int, 1>()); break;
    case 2: f(std::integral_constant<int, 2>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    default: throw Error(PQ_EINVAL, "pq_biot_field: K must be 1, 2, 4 or 8 (pad the excitations)");
    }
}

template <typename T>
void biot_run(int K, int64_t ns, const double *src, int64_t nseg, const double *seg, int64_t np, const double *pts,
              double *H, double *kernel_ms) {
    const size_t S = 3 + 4 * (size_t)K, G = 6 + 2 * (size_t)K;
    hipStream_t stream = nullptr;
    PQ_HIP(hipStreamCreate(&stream));
    struct Cleanup {
        hipStream_t s;
        ~Cleanup() { (void)hipStreamDestroy(s); }
    } cleanup{stream};
    DevBuf stage, dsrc, dseg, dpts, dH;
    const size_t nsrc = (size_t)ns * S, nsg = (size_t)nseg * G, npt = (size_t)np * 3, nh = (size_t)np * 3 * K * 2;
    // few points: split the sources over grid rows so that the launch still fills the chip
    const int pblocks = (int)((np + 256 * BIOT_PPT - 1) / (256 * BIOT_PPT));
    const int tile = sizeof(T) == 8 ? 128 : 256;
    int chunks = std::max(1, std::min((2048 + pblocks - 1) / pblocks, (int)((ns + 4 * tile - 1) / (4 * tile))));
    while (chunks > 1 && (size_t)chunks * nh * sizeof(T) > ((size_t)1 << 30)) --chunks;
    const int cs = (int)((ns + chunks - 1) / chunks);
    stage.reserve(std::max({nsrc, nsg, npt, nh, (size_t)1}) * 8);
    dsrc.reserve(std::max(nsrc, (size_t)1) * sizeof(T));
    dseg.reserve(std::max(nsg, (size_t)1) * sizeof(T));
    dpts.reserve(std::max(npt, (size_t)1) * sizeof(T));
    dH.reserve(std::max(nh * chunks, (size_t)1) * sizeof(T));
    auto upload = [&](const double *host, size_t n, DevBuf &dst) {
        if (!n) return;
        PQ_HIP(hipMemcpyAsync(stage.p, host, n * 8, hipMemcpyHostToDevice, stream));
        hipLaunchKernelGGL(to_precision_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream,
                           stage.as<double>(), n, dst.as<T>());
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipStreamSynchronize(stream));
    };
    upload(src, nsrc, dsrc);
    upload(seg, nsg, dseg);
    upload(pts, npt, dpts);
    hipEvent_t e0, e1;
    PQ_HIP(hipEventCreate(&e0));
    PQ_HIP(hipEventCreate(&e1));
    PQ_HIP(hipEventRecord(e0, stream));
    dispatch_k(K, [&](auto kk) {
        constexpr int KK = decltype(kk)::value;
        constexpr int TILE = sizeof(T) == 8 ? 128 : 256;
        hipLaunchKernelGGL((biot_sheet_kernel<T, KK, TILE>), dim3((unsigned)pblocks, (unsigned)chunks), dim3(256), 0,
                           stream, (int)ns, cs, dsrc.as<T>(), (int)np, dpts.as<T>(), dH.as<T>());
        if (nseg > 0)
            hipLaunchKernelGGL((biot_segment_kernel<T, KK>), dim3((unsigned)((np + 255) / 256)), dim3(256), 0, stream,
                               (int)nseg, dseg.as<T>(), (int)np, dpts.as<T>(), dH.as<T>());
    });
    PQ_HIP(hipGetLastError());
    hipLaunchKernelGGL(chunk_sum_kernel<T>, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, stream, dH.as<T>(),
                       chunks, nh, stage.as<double>());
    PQ_HIP(hipEventRecord(e1, stream));
    PQ_HIP(hipGetLastError());
    PQ_HIP(hipMemcpyAsync(H, stage.p, nh * 8, hipMemcpyDeviceToHost, stream));
    PQ_HIP(hipStreamSynchronize(stream));
    float ms = 0;
    PQ_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel_ms) *kernel_ms = ms;
}

}  // namespace

extern "C" {

int pq_biot_field(int device, int precision, int K, int64_t ns, const double *src, int64_t nseg, const double *seg,
                  int64_t np, const double *pts, double *H, double *kernel_ms) {
    return guard([&] {
        PQ_REQUIRE(precision == 32 || precision == 64, "pq_biot_field: precision 32 or 64");
        PQ_REQUIRE(ns >= 0 && nseg >= 0 && np >= 0, "pq_biot_field: negative count");
        PQ_REQUIRE((ns == 0 || src) && (nseg == 0 || seg) && (np == 0 || (pts && H)), "pq_biot_field: null argument");
        PQ_REQUIRE(ns < (int64_t)1 << 31 && nseg < (int64_t)1 << 31 && np < (int64_t)1 << 31,
                   "pq_biot_field: counts must fit in 32 bits");
        dispatch_k(K, [](auto) {});
        if (kernel_ms) *kernel_ms = 0;
        if (np == 0) return;
        DeviceScope scope(device);
        if (precision == 64)
            biot_run<double>(K, ns, src, nseg, seg, np, pts, H, kernel_ms);
        else
            biot_run<float>(K, ns, src, nseg, seg, np, pts, H, kernel_ms);
    });
}

}  // extern "C"

// ---- the near-field grid: sheet currents on a uniform grid convolved with the Biot-Savart kernel ----
//
// Semantics: h_grid (src/pcbqs/biot.py:108-181). Each layer's current density (Jx, Jy on an
// sx x sy grid of spacing dx, dy, deposited by the host) is convolved, linearly (zero padding to
// an fx x fy FFT of at least 3 s - 2 points per axis), with the kernel of the layer's height
// difference dz: G(r) = dx dy / (4 pi |r|^3) over the offsets r = (i dx, j dy, dz), |i| < sx,
// |j| < sy; H = (Jy * Gz, -Jx * Gz, Jx * Gy - Jy * Gx), summed over layers in the spectrum, then
// one inverse FFT per component and excitation and the observation window cropped out of the full
// convolution (every `step`-th point from (ox0, oy0)).

namespace {

__global__ void bg_pad_kernel(const double2 *__restrict__ J, int sx, int sy, int K, int k, int fx, int fy,
                              double2 *__restrict__ out) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)fx * fy) return;
    const int i = (int)(e / fy), j = (int)(e % fy);
    out[e] = (i < sx && j < sy) ? J[((size_t)i * sy + j) * K + k] : make_double2(0.0, 0.0);
}

__global__ void bg_green_kernel(double dx, double dy, double dz, int sx, int sy, int fx, int fy,
                                double2 *__restrict__ Gx, double2 *__restrict__ Gy, double2 *__restrict__ Gz) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)fx * fy) return;
    const int i = (int)(e / fy), j = (int)(e % fy);
    double gx = 0, gy = 0, gz = 0;
    if (i < 2 * sx - 1 && j < 2 * sy - 1) {
        const double rx = (i - (sx - 1)) * dx, ry = (j - (sy - 1)) * dy;
        const double r2 = rx * rx + ry * ry + dz * dz;
        const double g = dx * dy / (4.0 * M_PI * r2 * sqrt(r2));
        gx = rx * g; gy = ry * g; gz = dz * g;
    }
    Gx[e] = make_double2(gx, 0.0);
    Gy[e] = make_double2(gy, 0.0);
    Gz[e] = make_double2(gz, 0.0);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// A += (Jy Gz, -Jx Gz, Jx Gy - Jy Gx) in the spectrum, one excitation.
__global__ void bg_accum_kernel(const double2 *__restrict__ SJx, const double2 *__restrict__ SJy,
                                const double2 *__restrict__ Gx, const double2 *__restrict__ Gy,
                                const double2 *__restrict__ Gz, long long count, double2 *__restrict__ Ax,
                                double2 *__restrict__ Ay, double2 *__restrict__ Az) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= count) return;
    const double2 jx = SJx[e], jy = SJy[e], gx = Gx[e], gy = Gy[e], gz = Gz[e];
    const double2 a = cmul(jy, gz), b = cmul(jx, gz), c = cmul(jx, gy), d = cmul(jy, gx);
    Ax[e] = make_double2(Ax[e].x + a.x, Ax[e].y + a.y);
    Ay[e] = make_double2(Ay[e].x - b.x, Ay[e].y - b.y);
    Az[e] = make_double2(Az[e].x + c.x - d.x, Az[e].y + c.y - d.y);
}

// H[i][j][h][c][k] = A[ox0 + step i][oy0 + step j] / (fx fy) (the inverse FFT is unnormalised).
__global__ void bg_crop_kernel(const double2 *__restrict__ A, int fy, double scale, int ox0, int oy0, int step,
                               int nox, int noy, int nh, int h, int c, int K, int k, double2 *__restrict__ H) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= nox * noy) return;
    const int i = e / noy, j = e % noy;
    const double2 v = A[(size_t)(ox0 + step * i) * fy + oy0 + step * j];
    H[((((size_t)i * noy + j) * nh + h) * 3 + c) * K + k] = make_double2(v.x * scale, v.y * scale);
}

struct FftPlan {
    hipfftHandle h = 0;
    bool made = false;
    ~FftPlan() {
        if (made) (void)hipfftDestroy(h);
    }
};

#define PQ_FFT(call)                                                                                \
    do {                                                                                            \
        hipfftResult r_ = (call);                                                                   \
        if (r_ != HIPFFT_SUCCESS) throw Error(PQ_EDEVICE, std::string(#call) + " failed: " + std::to_string((int)r_)); \
    } while (0)

}  // namespace

extern "C" {

int pq_biot_grid(int device, int nlayers, const double *layer_z, int K, int sx, int sy, const double *J, double dx,
                 double dy, int nh, const double *heights, int fx, int fy, int ox0, int oy0, int step, int nox,
                 int noy, double *H, double *kernel_ms) {
    return guard([&] {
        PQ_REQUIRE(layer_z && J && heights && H, "pq_biot_grid: null argument");
        PQ_REQUIRE(nlayers >= 1 && K >= 1 && sx >= 1 && sy >= 1 && nh >= 1 && nox >= 1 && noy >= 1 && step >= 1,
                   "pq_biot_grid: invalid sizes");
        PQ_REQUIRE(dx > 0 && dy > 0 && std::isfinite(dx) && std::isfinite(dy), "pq_biot_grid: invalid spacing");
        PQ_REQUIRE(fx >= 3 * sx - 2 && fy >= 3 * sy - 2, "pq_biot_grid: the FFT is shorter than the linear convolution");
        PQ_REQUIRE(ox0 >= 0 && oy0 >= 0 && ox0 + (long long)step * (nox - 1) < 3LL * sx - 2 &&
                       oy0 + (long long)step * (noy - 1) < 3LL * sy - 2,
                   "pq_biot_grid: the observation window lies outside the convolution");
        for (int h = 0; h < nh; ++h)
            for (int l = 0; l < nlayers; ++l)
                PQ_REQUIRE(std::isfinite(heights[h]) && heights[h] != layer_z[l],
                           "pq_biot_grid: the field must be taken away from the copper planes");
        DeviceScope scope(device);
        hipStream_t st = nullptr;
        PQ_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() {
                if (s) {
                    (void)hipStreamSynchronize(s);
                    (void)hipStreamDestroy(s);
                }
            }
        } sg{st};
        const long long F = (long long)fx * fy;
        const size_t fb = (size_t)F * sizeof(double2);
        const size_t jcount = (size_t)sx * sy * K;
        DevBuf Jd, S, G, A, Hd;
        Jd.reserve(jcount * 2 * sizeof(double2));
        S.reserve(fb * 2 * K * nlayers);             // the layers' current spectra
        G.reserve(fb * 3);
        A.reserve(fb * 3 * K);
        Hd.reserve((size_t)nox * noy * nh * 3 * K * sizeof(double2));
        FftPlan plan;
        PQ_FFT(hipfftPlan2d(&plan.h, fx, fy, HIPFFT_Z2Z));
        plan.made = true;
        PQ_FFT(hipfftSetStream(plan.h, st));
        hipEvent_t e0, e1;
        PQ_HIP(hipEventCreate(&e0));
        PQ_HIP(hipEventCreate(&e1));
        const unsigned gridF = (unsigned)((F + 255) / 256);
        double2 *Sp = S.as<double2>(), *Gp = G.as<double2>(), *Ap = A.as<double2>();
        PQ_HIP(hipEventRecord(e0, st));
        for (int l = 0; l < nlayers; ++l) {
            PQ_HIP(hipMemcpyAsync(Jd.p, J + (size_t)l * 2 * jcount * 2, jcount * 2 * sizeof(double2),
                                  hipMemcpyHostToDevice, st));
            for (int c = 0; c < 2; ++c)
                for (int k = 0; k < K; ++k) {
                    double2 *dst = Sp + ((size_t)(l * 2 + c) * K + k) * F;
                    hipLaunchKernelGGL(bg_pad_kernel, dim3(gridF), dim3(256), 0, st, Jd.as<double2>() + c * jcount,
                                       sx, sy, K, k, fx, fy, dst);
                    PQ_FFT(hipfftExecZ2Z(plan.h, reinterpret_cast<hipfftDoubleComplex *>(dst),
                                         reinterpret_cast<hipfftDoubleComplex *>(dst), HIPFFT_FORWARD));
                }
        }
        for (int h = 0; h < nh; ++h) {
            PQ_HIP(hipMemsetAsync(Ap, 0, fb * 3 * K, st));
            for (int l = 0; l < nlayers; ++l) {
                hipLaunchKernelGGL(bg_green_kernel, dim3(gridF), dim3(256), 0, st, dx, dy, heights[h] - layer_z[l], sx,
                                   sy, fx, fy, Gp, Gp + F, Gp + 2 * F);
                for (int c = 0; c < 3; ++c)
                    PQ_FFT(hipfftExecZ2Z(plan.h, reinterpret_cast<hipfftDoubleComplex *>(Gp + c * F),
                                         reinterpret_cast<hipfftDoubleComplex *>(Gp + c * F), HIPFFT_FORWARD));
                for (int k = 0; k < K; ++k)
                    hipLaunchKernelGGL(bg_accum_kernel, dim3(gridF), dim3(256), 0, st,
                                       Sp + ((size_t)(l * 2) * K + k) * F, Sp + ((size_t)(l * 2 + 1) * K + k) * F, Gp,
                                       Gp + F, Gp + 2 * F, F, Ap + (size_t)(0 * K + k) * F, Ap + (size_t)(1 * K + k) * F,
                                       Ap + (size_t)(2 * K + k) * F);
            }
            for (int c = 0; c < 3; ++c)
                for (int k = 0; k < K; ++k) {
                    double2 *a = Ap + (size_t)(c * K + k) * F;
                    PQ_FFT(hipfftExecZ2Z(plan.h, reinterpret_cast<hipfftDoubleComplex *>(a),
                                         reinterpret_cast<hipfftDoubleComplex *>(a), HIPFFT_BACKWARD));
                    hipLaunchKernelGGL(bg_crop_kernel, dim3((nox * noy + 255) / 256), dim3(256), 0, st, a, fy,
                                       1.0 / (double)F, ox0, oy0, step, nox, noy, nh, h, c, K, k, Hd.as<double2>());
                }
            PQ_HIP(hipGetLastError());
        }
        PQ_HIP(hipEventRecord(e1, st));
        PQ_HIP(hipMemcpyAsync(H, Hd.p, (size_t)nox * noy * nh * 3 * K * sizeof(double2), hipMemcpyDeviceToHost, st));
        PQ_HIP(hipStreamSynchronize(st));
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (kernel_ms) *kernel_ms = ms;
    });
}

}  // extern "C"
