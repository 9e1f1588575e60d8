// Runtime plumbing of libgchi.so: status text, thread-local errors, device
// selection, fill kernels and the device-wide exclusive scan used by the graph
// builders (in-degree histogram -> CSC offsets, compaction offsets).
#include "common.hpp"
#include "internal.hpp"

#include <cstdio>
#include <cstring>

namespace gchi {

static thread_local std::string t_last_error;

void set_error(const std::string &msg) { t_last_error = msg; }

int fail(int status, const std::string &msg) {
    t_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
    t_last_error = buf;
    return GCHI_ERR_HIP;
}

// ------------------------------------------------------------------ fills
template <typename T>
__global__ void fill_kernel(T *p, T v, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

template <typename T>
static void launch_fill(T *p, T v, size_t n, hipStream_t s) {
    if (n == 0) return;
    size_t blocks = div_up(n, 256);
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_kernel<T>, dim3((unsigned)blocks), dim3(256), 0, s, p, v, n);
}
void launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) { launch_fill(p, v, n, s); }
void launch_fill_u64(uint64_t *p, uint64_t v, size_t n, hipStream_t s) { launch_fill(p, v, n, s); }

// ------------------------------------------------------------------ scan
// Reduce-then-scan: (1) per-tile sums, (2) one workgroup scans the tile sums,
// (3) each tile scans its elements and adds its base.  Tiles of 4096 counts.
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T>
GCHI_DEV T block_exclusive_scan(T v, T *s_tmp, T *total) {
    // s_tmp: kScanThreads/kWave entries of wave totals
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    T incl = v;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        T n = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += n;
    }
    if (lane == kWave - 1) s_tmp[wave] = incl;
    __syncthreads();
    T wave_base = 0, tot = 0;
    for (int w = 0; w < kScanThreads / kWave; w++) {
        T t = s_tmp[w];
        if (w < wave) wave_base += t;
        tot += t;
    }
    __syncthreads();
    *total = tot;
    return wave_base + incl - v;
}

__global__ void scan_tile_sums(const uint32_t *in, size_t n, uint64_t *tile_sums) {
    __shared__ uint64_t s_tmp[kScanThreads / kWave];
    size_t base = (size_t)blockIdx.x * kScanTile;
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        size_t idx = base + (size_t)i * kScanThreads + threadIdx.x;
        if (idx < n) acc += in[idx];
    }
    uint64_t tot;
    block_exclusive_scan<uint64_t>(acc, s_tmp, &tot);
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = tot;
}

__global__ void scan_tile_bases(uint64_t *tile_sums, size_t ntiles) {
    // single workgroup, sequential over chunks of kScanThreads
    __shared__ uint64_t s_tmp[kScanThreads / kWave];
    uint64_t carry = 0;
    for (size_t c = 0; c < ntiles; c += kScanThreads) {
        size_t idx = c + threadIdx.x;
        uint64_t v = idx < ntiles ? tile_sums[idx] : 0;
        uint64_t tot;
        uint64_t ex = block_exclusive_scan<uint64_t>(v, s_tmp, &tot);
        if (idx < ntiles) tile_sums[idx] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) tile_sums[ntiles] = carry;
}

template <typename OutT>
__global__ void scan_tiles(const uint32_t *in, size_t n, const uint64_t *tile_bases, OutT *out) {
    __shared__ uint64_t s_tmp[kScanThreads / kWave];
    size_t base = (size_t)blockIdx.x * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        size_t idx = base + i;
        v[i] = idx < n ? in[idx] : 0u;
        acc += v[i];
    }
    uint64_t tot;
    uint64_t run = tile_bases[blockIdx.x] + block_exclusive_scan<uint64_t>(acc, s_tmp, &tot);
#pragma unroll
    for (int i = 0; i < kScanItems; i++) {
        size_t idx = base + i;
        if (idx < n) out[idx] = (OutT)run;
        run += v[i];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
        size_t ntiles = gridDim.x;
        out[n] = (OutT)tile_bases[ntiles];
    }
}

size_t scan_workspace(size_t n) {
    size_t ntiles = div_up(n, kScanTile);
    return align_up((ntiles + 1) * sizeof(uint64_t), 256);
}

template <typename OutT>
static int exclusive_scan_impl(const uint32_t *in, OutT *out, size_t n, void *work, size_t work_bytes,
                               hipStream_t s) {
    GCHI_CHECK(out != nullptr, GCHI_ERR_INVALID, "scan: null output");
    if (n == 0) {
        hipLaunchKernelGGL(fill_kernel<OutT>, dim3(1), dim3(1), 0, s, out, (OutT)0, (size_t)1);
        GCHI_LAUNCH_CHECK();
        return GCHI_OK;
    }
    GCHI_CHECK(work_bytes >= scan_workspace(n), GCHI_ERR_WORKSPACE, "scan: workspace too small");
    size_t ntiles = div_up(n, kScanTile);
    uint64_t *sums = reinterpret_cast<uint64_t *>(work);
    hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, n, sums);
    hipLaunchKernelGGL(scan_tile_bases, dim3(1), dim3(kScanThreads), 0, s, sums, ntiles);
    hipLaunchKernelGGL(scan_tiles<OutT>, dim3((unsigned)ntiles), dim3(kScanThreads), 0, s, in, n, sums, out);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

int exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n, void *work, size_t work_bytes,
                       hipStream_t s) {
    return exclusive_scan_impl<uint32_t>(in, out, n, work, work_bytes, s);
}
int exclusive_scan_u32_to_u64(const uint32_t *in, uint64_t *out, size_t n, void *work,
                              size_t work_bytes, hipStream_t s) {
    return exclusive_scan_impl<uint64_t>(in, out, n, work, work_bytes, s);
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_version(void) { return 100; }

extern "C" const char *gchi_strerror(int status) {
    switch (status) {
        case GCHI_OK: return "ok";
        case GCHI_ERR_INVALID: return "invalid argument";
        case GCHI_ERR_HIP: return "HIP runtime error";
        case GCHI_ERR_WORKSPACE: return "workspace too small";
        case GCHI_ERR_UNSUPPORTED: return "unsupported on device";
        case GCHI_ERR_FORMAT: return "malformed input";
        case GCHI_ERR_OVERFLOW: return "result overflows output type";
        case GCHI_ERR_NO_DEVICE: return "no HIP device";
        case GCHI_ERR_IO: return "I/O error";
        case GCHI_ERR_TIMEOUT: return "bounded device wait exceeded";
        default: return "unknown status";
    }
}

extern "C" int gchi_last_error(char *buf, size_t len) {
    if (!buf || len == 0) return GCHI_ERR_INVALID;
    size_t n = t_last_error.size();
    if (n >= len) n = len - 1;
    memcpy(buf, t_last_error.data(), n);
    buf[n] = 0;
    return GCHI_OK;
}

extern "C" int gchi_device_open(int device) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(GCHI_ERR_NO_DEVICE, "no HIP device visible");
    GCHI_CHECK(device >= 0 && device < count, GCHI_ERR_INVALID, "device index out of range");
    GCHI_HIP(hipSetDevice(device));
    GCHI_HIP(hipFree(nullptr));
    return tc_scratch_setup(device);
}

extern "C" int gchi_device_info(int *num_cus, char *arch, size_t arch_len) {
    int dev = 0;
    GCHI_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    GCHI_HIP(hipGetDeviceProperties(&p, dev));
    if (num_cus) *num_cus = p.multiProcessorCount;
    if (arch && arch_len) {
        snprintf(arch, arch_len, "%s", p.gcnArchName);
    }
    return GCHI_OK;
}

extern "C" int gchi_memcpy(void *dst, const void *src, size_t bytes, void *stream) {
    if (bytes == 0) return GCHI_OK;
    GCHI_CHECK(dst && src, GCHI_ERR_INVALID, "memcpy: null pointer");
    hipStream_t s = as_stream(stream);
    GCHI_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
    GCHI_HIP(hipStreamSynchronize(s));
    return GCHI_OK;
}

extern "C" int gchi_malloc(void **d_ptr, size_t bytes) {
    GCHI_CHECK(d_ptr, GCHI_ERR_INVALID, "malloc: null pointer");
    *d_ptr = nullptr;
    GCHI_HIP(hipMalloc(d_ptr, bytes ? bytes : 1));
    return GCHI_OK;
}

extern "C" int gchi_free(void *d_ptr) {
    if (d_ptr) GCHI_HIP(hipFree(d_ptr));
    return GCHI_OK;
}

extern "C" int gchi_host_alloc(void **h_ptr, size_t bytes) {
    GCHI_CHECK(h_ptr, GCHI_ERR_INVALID, "host_alloc: null pointer");
    *h_ptr = nullptr;
    GCHI_HIP(hipHostMalloc(h_ptr, bytes ? bytes : 1, hipHostMallocDefault));
    return GCHI_OK;
}

extern "C" int gchi_host_free(void *h_ptr) {
    if (h_ptr) GCHI_HIP(hipHostFree(h_ptr));
    return GCHI_OK;
}

extern "C" int gchi_synchronize(void *stream) {
    GCHI_HIP(hipStreamSynchronize(as_stream(stream)));
    return GCHI_OK;
}
