// Shared host/device helpers of libpcbqs_hip.so: error state, HIP checks.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <stdexcept>

#include "../../include/pcbqs_amd.h"

namespace pq {

// An error crossing the C ABI: a status and a message (kept per thread).
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string &m) : std::runtime_error(m), status(s) {}
};

void set_error(const char *msg);

#define PQ_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::pq::Error(e_ == hipErrorOutOfMemory ? PQ_ENOMEM : PQ_EDEVICE,             \
                              std::string(#call) + ": " + hipGetErrorString(e_));             \
        } while (0)

#define PQ_REQUIRE(cond, msg)                                                                 \
    do {                                                                                      \
        if (!(cond)) throw ::pq::Error(PQ_EINVAL, msg);                                       \
    } while (0)

// Run `body` and translate exceptions into a status code and pq_last_error().
template <class F> int guard(F &&body) {
    try {
        body();
        return PQ_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.status;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return PQ_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return PQ_EDEVICE;
    }
}

// A device allocation, freed when its owner (a handle or a scope) ends. Not copyable.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    ~DevBuf() { release(); }
    void reserve(size_t n) {
        if (n <= bytes) return;
        release();
        PQ_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Switches to a device for the lifetime of the object.
struct DeviceScope {
    int old = 0;
    explicit DeviceScope(int dev) {
        PQ_HIP(hipGetDevice(&old));
        if (old != dev) PQ_HIP(hipSetDevice(dev));
    }
    ~DeviceScope() { (void)hipSetDevice(old); }
};

}  // namespace pq
