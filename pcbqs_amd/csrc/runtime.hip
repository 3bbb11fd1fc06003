// Runtime entry points of libpcbqs_hip.so: error state, device memory, copies.
#include "pq_common.h"

namespace pq {
static thread_local std::string last_error;
void set_error(const char *msg) { last_error = msg ? msg : ""; }
}  // namespace pq

using namespace pq;

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

int pq_synchronize(int device) {
    return guard([&] {
        DeviceScope scope(device);
        PQ_HIP(hipDeviceSynchronize());
    });
}

}  // extern "C"
