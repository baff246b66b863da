// Helpers shared by the C-ABI translation units: error reporting, the
// per-thread stream and stream-ordered device buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

#include "../../include/xcape_hip.h"

namespace xcape {

int fail(int code, const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);
int ensure_device();
hipStream_t thread_stream();
hipError_t zero_counters(unsigned long long *p, int n, hipStream_t s);
void copy_stats(const unsigned long long *h, xcape_stats *stats);
const void *surrogate_device();   // loaded surrogate (device), or nullptr

// Stream-ordered device allocation released on scope exit (hipFreeAsync on
// the same stream, so it never races the work that uses it).
template <typename T> struct DevBuf {
    T *ptr = nullptr;
    hipStream_t s = nullptr;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf()
    {
        if (ptr) (void)hipFreeAsync(ptr, s);
    }
    hipError_t alloc(size_t n, hipStream_t st)
    {
        s = st;
        if (n == 0) n = 1;
        return hipMallocAsync((void **)&ptr, n * sizeof(T), st);
    }
    hipError_t upload(const T *h, size_t n, hipStream_t st)
    {
        if (n == 0) return hipSuccess;
        return hipMemcpyAsync(ptr, h, n * sizeof(T), hipMemcpyHostToDevice, st);
    }
    hipError_t download(T *h, size_t n, hipStream_t st) const
    {
        if (n == 0) return hipSuccess;
        return hipMemcpyAsync(h, ptr, n * sizeof(T), hipMemcpyDeviceToHost, st);
    }
};

}  // namespace xcape
