/*
 * Internal helpers shared by the sqaod-rocm library sources: error type,
 * HIP status checks and small device utilities for gfx950 (wave64).
 */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

namespace sqaod_rocm {

/* Errors cross the C ABI as a non-zero status plus a message retrievable with
 * sqaod_rocm_last_error(); inside the library they are C++ exceptions, as in
 * the reference (sqaodc/common/defines.cpp:32-43 throwError -> runtime_error). */
struct Error : std::runtime_error {
    explicit Error(const std::string &s) : std::runtime_error(s) {}
};

[[noreturn]] inline void throwErrorf(const char *file, int line, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    char out[1200];
    snprintf(out, sizeof(out), "%s:%d %s", file, line, buf);
    throw Error(out);
}

#define SQ_THROW(...) ::sqaod_rocm::throwErrorf(__FILE__, __LINE__, __VA_ARGS__)
#define SQ_THROW_IF(cond, ...) \
    do {                        \
        if (cond)               \
            SQ_THROW(__VA_ARGS__); \
    } while (0)

#define HIP_CHECK(expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            ::sqaod_rocm::throwErrorf(__FILE__, __LINE__, "HIP error %d (%s) in %s", (int)e_, \
                                      hipGetErrorString(e_), #expr);                          \
    } while (0)

inline int divru(int a, int b) { return (a + b - 1) / b; }
inline int roundUp(int a, int b) { return divru(a, b) * b; }

/* Device memory owned by a solver; freed in the destructor. */
template <class T>
struct DevBuf {
    T *ptr = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void allocate(size_t count) {
        if (count == n && ptr)
            return;
        release();
        if (count) {
            HIP_CHECK(hipMalloc((void **)&ptr, count * sizeof(T)));
            n = count;
        }
    }
    void release() {
        if (ptr)
            (void)hipFree(ptr);
        ptr = nullptr;
        n = 0;
    }
    size_t bytes() const { return n * sizeof(T); }
};

/* Pinned host memory (hipHostMalloc; never hipHostRegister). */
template <class T>
struct PinnedBuf {
    T *ptr = nullptr;
    size_t n = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    /* flags: hipHostMallocDefault, or hipHostMallocCoherent | hipHostMallocMapped
     * for a flag the device writes and the host polls */
    void allocate(size_t count, unsigned flags = hipHostMallocDefault) {
        if (count <= n && ptr)
            return;
        release();
        if (count) {
            HIP_CHECK(hipHostMalloc((void **)&ptr, count * sizeof(T), flags));
            n = count;
        }
    }
    void release() {
        if (ptr)
            (void)hipHostFree(ptr);
        ptr = nullptr;
        n = 0;
    }
};

}  // namespace sqaod_rocm
