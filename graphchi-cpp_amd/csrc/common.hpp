// Internal helpers shared by every translation unit of libgchi.so (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <string>

#include "gchi.h"

#define GCHI_HD __host__ __device__ __forceinline__
#define GCHI_DEV __device__ __forceinline__

namespace gchi {

constexpr int kWave = 64;  // CDNA wavefront width
// Largest edge count of one graph: edge offsets are 32-bit, and a 32-bit edge
// index plus a grid stride (at most 2^24) must not wrap past 2^32.
constexpr uint64_t kMaxEdges = (1ull << 32) - (1ull << 25);

// ---- error reporting (thread-local last error message) ----
void set_error(const std::string &msg);
int fail(int status, const std::string &msg);
int hip_fail(hipError_t e, const char *what, const char *file, int line);

#define GCHI_HIP(call)                                                         \
    do {                                                                       \
        hipError_t _e = (call);                                                \
        if (_e != hipSuccess) return ::gchi::hip_fail(_e, #call, __FILE__, __LINE__); \
    } while (0)

#define GCHI_CHECK(cond, status, msg)                                          \
    do {                                                                       \
        if (!(cond)) return ::gchi::fail((status), (msg));                     \
    } while (0)

#define GCHI_TRY(expr)                                                         \
    do {                                                                       \
        int _s = (expr);                                                       \
        if (_s != GCHI_OK) return _s;                                          \
    } while (0)

#define GCHI_LAUNCH_CHECK()                                                    \
    do {                                                                       \
        hipError_t _e = hipGetLastError();                                     \
        if (_e != hipSuccess) return ::gchi::hip_fail(_e, "kernel launch", __FILE__, __LINE__); \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

GCHI_HD uint64_t div_up(uint64_t a, uint64_t b) { return (a + b - 1) / b; }
GCHI_HD size_t align_up(size_t a, size_t b) { return (a + b - 1) / b * b; }

// Number of bits needed to represent v (0 -> 0): restates log2Up of
// src/util/radixSort.hpp:77-82 (smallest a with 2^a > v).
GCHI_HD int bits_for(uint64_t v) {
    int a = 0;
    while (a < 64 && (1ull << a) <= v) a++;
    return a;
}

// Carves aligned sub-buffers out of one caller-provided workspace.
struct Carver {
    char *base;
    size_t off = 0;
    size_t cap;
    Carver(void *b, size_t c) : base(static_cast<char *>(b)), cap(c) {}
    template <typename T>
    T *take(size_t count) {
        off = align_up(off, 256);
        T *p = reinterpret_cast<T *>(base ? base + off : nullptr);
        off += count * sizeof(T);
        return p;
    }
    bool ok() const { return off <= cap; }
};

// Size-only carver (for *_workspace queries).
struct Sizer {
    size_t off = 0;
    template <typename T>
    void take(size_t count) {
        off = align_up(off, 256);
        off += count * sizeof(T);
    }
};

// Pointer-returning size-only carver, for layouts shared with Carver.
struct SizerAdapter {
    Sizer z;
    template <typename T>
    T *take(size_t count) {
        z.take<T>(count);
        return nullptr;
    }
};

// ---- device-side wave helpers ----
GCHI_DEV int lane_id() { return __lane_id(); }
GCHI_DEV uint64_t lanemask_lt() { return (1ull << lane_id()) - 1ull; }

template <typename T>
GCHI_DEV T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Agent-scope atomics for cross-workgroup (and cross-XCD) communication.
template <typename T>
GCHI_DEV T ld_acquire_agent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
GCHI_DEV T ld_relaxed_agent(const T *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
GCHI_DEV void st_release_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
GCHI_DEV void st_relaxed_agent(T *p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Fill kernel (used instead of hipMemsetAsync so that capture into a HIP graph
// never records a memset node).
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
void launch_fill_u64(uint64_t *p, uint64_t v, size_t n, hipStream_t s);

// Device-wide exclusive scan of uint32 counts into uint32 or uint64 offsets;
// writes n+1 outputs (the last one is the total).  Workspace from scan_workspace.
size_t scan_workspace(size_t n);
int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *work, size_t work_bytes,
                       hipStream_t s);
int exclusive_scan_u32_to_u64(const uint32_t *in, uint64_t *out, size_t n, void *work,
                              size_t work_bytes, hipStream_t s);

}  // namespace gchi
