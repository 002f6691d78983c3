// runtime.hpp -- device-buffer and error-handling helpers shared by the C-ABI sources.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <exception>
#include <new>
#include <stdexcept>
#include <string>

#include "internal.hpp"

namespace dbie {

// Checks once per process that a gfx950 device is visible; throws DBIE_NO_DEVICE otherwise.
void ensure_device();

// Per-thread cache of device blocks for temporaries.  hipFree synchronises the whole device, so a
// temporary freed by one host thread (an operator build, a solve) would stall the solves other threads
// are running; a pooled buffer instead returns its block to its thread's cache, and the thread's next
// temporary of the same size class takes it.  A block is reused only by the thread that released it,
// and every function with pooled temporaries synchronises its stream before returning, so a reused
// block is no longer in use on the device.  Blocks of 8 GiB and more, and blocks beyond 16 GiB cached
// per thread, are freed.  pool_take returns nullptr when hipMalloc fails (after freeing the cache).
void *pool_take(size_t bytes, size_t *got);
void pool_give(void *p, size_t bytes);

// Device buffer with RAII.  pooled = true: blocks come from and return to the thread's cache.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    bool pooled = false;
    DBuf() = default;
    explicit DBuf(bool pool) : pooled(pool) {}
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) {
            // during stack unwinding the buffer's stream may still have work queued: free it (hipFree
            // waits for the device) rather than hand it to a later temporary
            if (pooled && std::uncaught_exceptions() == 0) pool_give(p, bytes);
            else (void)hipFree(p);
        }
        p = nullptr;
        bytes = 0;
    }
    void ensure(size_t b) {
        if (b <= bytes) return;
        release();
        if (pooled) {
            size_t got = 0;
            p = pool_take(b, &got);
            if (!p) fail(DBIE_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed");
            bytes = got;
            return;
        }
        if (hipMalloc(&p, b) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            fail(DBIE_OOM, "hipMalloc of " + std::to_string(b) + " bytes failed");
        }
        bytes = b;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// Pinned host memory (hipHostMalloc) for the small per-iteration device-to-host copies of a Krylov
// solve: copies into it are asynchronous, where copies into pageable memory block the host until the
// device has reached them.  Grows on demand; the caller re-sizes only with no copy in flight.
struct PinnedBuf {
    void *p = nullptr;
    size_t bytes = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() {
        if (p) (void)hipHostFree(p);
    }
    void ensure(size_t b) {
        if (b <= bytes) return;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        if (hipHostMalloc(&p, b, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            p = nullptr;
            fail(DBIE_OOM, "hipHostMalloc of " + std::to_string(b) + " bytes failed");
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
