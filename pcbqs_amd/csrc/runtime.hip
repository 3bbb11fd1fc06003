// Runtime entry points of libpcbqs_hip.so: error state, device memory, copies.
#include "pq_common.h"

namespace pq {
static thread_local std::string last_error;
void set_error(const char *msg) { last_error = msg ? msg : ""; }
}  // namespace pq

using namespace pq;

// The store pattern of the FP32 fill without its arithmetic: one wave per 64 x 64 tile writes its
// 64 rows as 64-float (256 B) segments of every plane, non-temporally, value r - c at (r, c). Timed
// by pq_store_probe as the store ceiling beside the fill's roofline.
__global__ __launch_bounds__(64) void store_probe_kernel(float *out, long long n, long long ld, size_t plane,
                                                        int nplanes) {
    const long long c = (long long)blockIdx.x * 64 + threadIdx.x;
    const long long r0 = (long long)blockIdx.y * 64, r1 = r0 + 64 < n ? r0 + 64 : n;
    if (c >= n) return;
    for (long long r = r0; r < r1; ++r) {
        const float v = (float)(r - c);
        for (int p = 0; p < nplanes; ++p) __builtin_nontemporal_store(v, out + p * plane + r * ld + c);
    }
}

template <class T> __global__ void gather_strided_kernel(const T *src, size_t stride, size_t count, T *dst) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < count) dst[i] = src[i * stride];
}

extern "C" {

const char *pq_last_error(void) { return last_error.c_str(); }

int pq_device_count(int *count) {
    return guard([&] {
        PQ_REQUIRE(count, "pq_device_count: null output");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        if (e != hipSuccess) n = 0;
        *count = n;
    });
}

int pq_malloc(int device, size_t bytes, void **ptr) {
    return guard([&] {
        PQ_REQUIRE(ptr, "pq_malloc: null output");
        DeviceScope scope(device);
        *ptr = nullptr;
        PQ_HIP(hipMalloc(ptr, bytes ? bytes : 1));
    });
}

int pq_free(void *ptr) {
    return guard([&] {
        if (ptr) PQ_HIP(hipFree(ptr));
    });
}

int pq_memcpy_h2d(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if (bytes) PQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    });
}

int pq_memcpy_d2h(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if (bytes) PQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    });
}

int pq_memcpy_d2h_strided(void *dst, const void *src, size_t elem_bytes, size_t stride_bytes, size_t count) {
    return guard([&] {
        PQ_REQUIRE(stride_bytes >= elem_bytes, "pq_memcpy_d2h_strided: stride below the element size");
        PQ_REQUIRE(elem_bytes == 4 || elem_bytes == 8, "pq_memcpy_d2h_strided: elements of 4 or 8 bytes");
        PQ_REQUIRE(stride_bytes % elem_bytes == 0, "pq_memcpy_d2h_strided: stride not a multiple of the element");
        if (!count) return;
        // gathered on the device, then one contiguous copy (a 2-D copy of one-element rows goes row by row)
        DevBuf tmp;
        tmp.reserve(count * elem_bytes);
        const size_t st = stride_bytes / elem_bytes;
        const unsigned grid = (unsigned)((count + 255) / 256);
        if (elem_bytes == 4)
            hipLaunchKernelGGL(gather_strided_kernel<uint32_t>, dim3(grid), dim3(256), 0, 0,
                               static_cast<const uint32_t *>(src), st, count, tmp.as<uint32_t>());
        else
            hipLaunchKernelGGL(gather_strided_kernel<uint64_t>, dim3(grid), dim3(256), 0, 0,
                               static_cast<const uint64_t *>(src), st, count, tmp.as<uint64_t>());
        PQ_HIP(hipGetLastError());
        PQ_HIP(hipMemcpy(dst, tmp.p, count * elem_bytes, hipMemcpyDeviceToHost));
    });
}

int pq_memcpy_d2d(void *dst, const void *src, size_t bytes) {
    return guard([&] {
        if (bytes) PQ_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    });
}

int pq_memset(void *dst, int value, size_t bytes) {
    return guard([&] {
        if (bytes) PQ_HIP(hipMemset(dst, value, bytes));
    });
}

int pq_stream_create(int device, void **stream) {
    return guard([&] {
        PQ_REQUIRE(stream, "pq_stream_create: null output");
        DeviceScope scope(device);
        hipStream_t s = nullptr;
        PQ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        *stream = s;
    });
}

int pq_stream_destroy(void *stream) {
    return guard([&] {
        if (stream) PQ_HIP(hipStreamDestroy((hipStream_t)stream));
    });
}

int pq_stream_synchronize(void *stream) {
    return guard([&] { PQ_HIP(hipStreamSynchronize((hipStream_t)stream)); });
}

int pq_store_probe(void *planes, int64_t n, int64_t ld, int nplanes, double *ms) {
    return guard([&] {
        PQ_REQUIRE(planes && ms, "pq_store_probe: null argument");
        PQ_REQUIRE(n > 0 && ld >= n && nplanes >= 1 && nplanes <= 4, "pq_store_probe: invalid sizes");
        PQ_REQUIRE((n + 63) / 64 <= 65535, "pq_store_probe: more than 65535 row tiles");
        const dim3 grid((unsigned)((n + 63) / 64), (unsigned)((n + 63) / 64));
        const size_t plane = (size_t)n * ld;
        hipEvent_t e0, e1;
        PQ_HIP(hipEventCreate(&e0));
        PQ_HIP(hipEventCreate(&e1));
        hipLaunchKernelGGL(store_probe_kernel, grid, dim3(64), 0, 0, (float *)planes, (long long)n, (long long)ld,
                           plane, nplanes);
        PQ_HIP(hipEventRecord(e0, 0));
        for (int r = 0; r < 3; ++r)
            hipLaunchKernelGGL(store_probe_kernel, grid, dim3(64), 0, 0, (float *)planes, (long long)n, (long long)ld,
                               plane, nplanes);
        PQ_HIP(hipEventRecord(e1, 0));
        PQ_HIP(hipEventSynchronize(e1));
        float t = 0;
        PQ_HIP(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        PQ_HIP(hipGetLastError());
        *ms = t / 3.0;
    });
}

int pq_synchronize(int device) {
    return guard([&] {
        DeviceScope scope(device);
        PQ_HIP(hipDeviceSynchronize());
    });
}

}  // extern "C"
