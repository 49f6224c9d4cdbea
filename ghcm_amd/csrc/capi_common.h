// capi_common.h -- helpers shared by the C-ABI translation units of libghcm: the
// thread-local error message, HIP error propagation and grow-only device / pinned buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <stdexcept>
#include <vector>
#include <thread>

#include <algorithm>
#include <cstddef>

#include "ghcm.h"

namespace ghcm_capi {

// Records a printf-style message for ghcm_last_error() and returns `status`.
int fail(int status, const char* fmt, ...);

// Runs an entry point's body so that no exception crosses the C boundary: a host container
// that cannot grow becomes GHCM_ERR_OUT_OF_MEMORY.
template <class F>
int no_throw(F&& body) noexcept {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(GHCM_ERR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::length_error&) {
        return fail(GHCM_ERR_OUT_OF_MEMORY, "host allocation too large");
    } catch (const std::exception& e) {
        return fail(GHCM_ERR_INVALID_ARG, "%s", e.what());
    }
}
// Validation shared by contexts and servers.
int check_config(const ghcm_config* c);

inline bool valid_kind(int k) { return k >= GHCM_FIELD_COST_F32 && k <= GHCM_FIELD_SPEED_BF16; }
inline size_t kind_bytes(int k) {
    return (k == GHCM_FIELD_SPEED_F16 || k == GHCM_FIELD_SPEED_BF16) ? 2 : 4;
}

// Algorithmic HBM bytes of one tile visit: field read + time read (tile and the 4 halo
// edges) + time write + 40 B of tile metadata.
inline uint32_t bytes_per_visit(uint32_t tw, uint32_t th, int kind) {
    const uint32_t P = tw * th;
    return P * (uint32_t)kind_bytes(kind) + (P + 2 * (tw + th)) * 4 + P * 4 + 40;
}

// The same for a tw x th x td tile of a 3-D solve: field + time (tile and its 6 halo
// faces) + time write + 40 B.
inline uint32_t bytes_per_visit3d(uint32_t tw, uint32_t th, uint32_t td, int kind) {
    const uint32_t P = tw * th * td;
    const uint32_t faces = 2 * (tw * th + th * td + tw * td);
    return P * (uint32_t)kind_bytes(kind) + (P + faces) * 4 + P * 4 + 40;
}

#define HIP_TRY(expr)                                                                        \
    do {                                                                                     \
        hipError_t e_ = (expr);                                                              \
        if (e_ != hipSuccess)                                                                \
            return ::ghcm_capi::fail(GHCM_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

// Waits for `ev` by polling it.  A blocking wait may sleep for tens of microseconds past
// the GPU's signal, which a wave loop polled a few times per solve pays every time.
inline hipError_t spin_wait(hipEvent_t ev) {
    for (;;) {
        const hipError_t q = hipEventQuery(ev);
        if (q != hipErrorNotReady) return q;
        __builtin_ia32_pause();
    }
}

// Records `ev` on `st` and polls it: the stream's work so far has finished on return.
inline hipError_t spin_sync(hipStream_t st, hipEvent_t ev) {
    const hipError_t e = hipEventRecord(ev, st);
    return e != hipSuccess ? e : spin_wait(ev);
}

// Device buffer that only grows; reused across solves.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int reserve(size_t want) {
        if (want <= bytes) return GHCM_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t b = std::max<size_t>(want, 256);
        hipError_t e = hipMalloc(&p, b);
        if (e != hipSuccess) {
            p = nullptr;
            return fail(GHCM_ERR_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        }
        bytes = b;
        return GHCM_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
};

// Pinned host buffer that only grows: mapped into the device's address space and
// coherent, so kernels read and write it directly (`dev` is the device view).
// Host copy into or out of pinned staging.  One core's memcpy bandwidth (about 8-10 GB/s) is
// far below the host's, so copies of 4 MB and more are split into 2 MB-aligned parts over a
// pool of up to 7 worker threads (started on first use, shared by every caller) plus the
// calling thread.
void par_memcpy(void* dst, const void* src, size_t n);

struct PinnedBuf {
    void* p = nullptr;
    void* dev = nullptr;
    size_t bytes = 0;
    int reserve(size_t want) {
        if (want <= bytes) return GHCM_OK;
        release();
        size_t b = std::max<size_t>(want, 256);
        hipError_t e = hipHostMalloc(&p, b, hipHostMallocMapped | hipHostMallocCoherent);
        if (e == hipSuccess) e = hipHostGetDevicePointer(&dev, p, 0);
        if (e != hipSuccess) {
            if (p) (void)hipHostFree(p);
            p = dev = nullptr;
            return fail(GHCM_ERR_OUT_OF_MEMORY, "hipHostMalloc(%zu) failed: %s", b, hipGetErrorString(e));
        }
        // Pinned memory comes back from the runtime with whatever a previous owner left in
        // it; host records are read as tags and progress words, so start from zero.
        std::memset(p, 0, b);
        bytes = b;
        return GHCM_OK;
    }
    void release() {
        if (p) (void)hipHostFree(p);
        p = dev = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return static_cast<T*>(p); }
    template <class T> T* dev_as() const { return static_cast<T*>(dev); }
};

}  // namespace ghcm_capi
