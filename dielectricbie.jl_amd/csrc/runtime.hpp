// runtime.hpp -- device-buffer and error-handling helpers shared by the C-ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>

#include "internal.hpp"

namespace dbie {

// Checks once per process that a gfx950 device is visible; throws DBIE_NO_DEVICE otherwise.
void ensure_device();

// Device buffer with RAII.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t b) {
        if (b <= bytes) return;
        release();
        if (hipMalloc(&p, b) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            fail(DBIE_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed");
        }
        bytes = b;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Host-to-device copy that completes before returning (host sources may be temporaries).
template <class T>
void upload(DBuf &b, const T *src, size_t n, hipStream_t st) {
    b.ensure(std::max<size_t>(n, 1) * sizeof(T));
    if (n) {
        DBIE_HIP(hipMemcpyAsync(b.p, src, n * sizeof(T), hipMemcpyHostToDevice, st));
        DBIE_HIP(hipStreamSynchronize(st));
    }
}
inline void h2d(void *dst, const void *src, size_t bytes, hipStream_t st) {
    DBIE_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st));
    DBIE_HIP(hipStreamSynchronize(st));
}

// Runs f, mapping exceptions to status codes and the thread-local error message.
template <class F>
dbie_status guard(F &&f) {
    try {
        f();
        return DBIE_OK;
    } catch (const Error &e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return DBIE_OOM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return DBIE_BAD_ARG;
    }
}

}  // namespace dbie
