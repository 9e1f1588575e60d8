// Synthetic input: R-MAT generation (device and host, bit-identical) and the
// stable self-loop filter of sharder::preprocessing_add_edge
// (src/preprocessing/sharder.hpp:427-446).
#include "common.hpp"
#include "rmat_rng.hpp"

#include <algorithm>
#include <thread>
#include <vector>

namespace gchi {

__global__ void rmat_kernel(uint32_t *src, uint32_t *dst, uint64_t first, uint64_t n, int scale,
                            uint64_t key, RmatThresholds t) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        uint32_t s, d;
        rmat_edge(key, first + i, scale, t, &s, &d);
        src[i] = s;
        dst[i] = d;
    }
}

// ---- stable compaction of non-self-loop pairs ----
constexpr int kCompThreads = 256;
constexpr int kCompItems = 16;
constexpr int kCompTile = kCompThreads * kCompItems;
constexpr int kCompWaves = kCompThreads / kWave;

__global__ void compact_count(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t *tile_counts,
                              uint32_t *maxid) {
    __shared__ uint32_t s_cnt[kCompWaves];
    __shared__ uint32_t s_max[kCompWaves];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint64_t base = (uint64_t)blockIdx.x * kCompTile + (uint64_t)wave * (kCompItems * kWave);
    uint32_t cnt = 0, mx = 0;
#pragma unroll 4
    for (int i = 0; i < kCompItems; i++) {
        uint64_t idx = base + (uint64_t)i * kWave + lane;
        bool keep = false;
        if (idx < n) {
            uint32_t s = src[idx], d = dst[idx];
            keep = s != d;
            if (keep) mx = max(mx, max(s, d));
        }
        cnt += (uint32_t)__popcll(__ballot(keep));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = max(mx, (uint32_t)__shfl_xor((int)mx, o, kWave));
    if (lane == 0) {
        s_cnt[wave] = cnt;
        s_max[wave] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t c = 0, m = 0;
        for (int w = 0; w < kCompWaves; w++) {
            c += s_cnt[w];
            m = max(m, s_max[w]);
        }
        tile_counts[blockIdx.x] = c;
        if (m) atomicMax(maxid, m);
    }
}

__global__ void compact_scatter(const uint32_t *src, const uint32_t *dst, uint64_t n, const uint64_t *tile_off,
                                uint32_t *src_out, uint32_t *dst_out, uint64_t *count_out, uint32_t ntiles) {
    __shared__ uint32_t s_cnt[kCompWaves];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint64_t base = (uint64_t)blockIdx.x * kCompTile + (uint64_t)wave * (kCompItems * kWave);
    uint32_t cnt = 0;
    for (int i = 0; i < kCompItems; i++) {
        uint64_t idx = base + (uint64_t)i * kWave + lane;
        bool keep = idx < n && src[idx] != dst[idx];
        cnt += (uint32_t)__popcll(__ballot(keep));
    }
    if (lane == 0) s_cnt[wave] = cnt;
    __syncthreads();
    uint64_t pos = tile_off[blockIdx.x];
    for (int w = 0; w < wave; w++) pos += s_cnt[w];
    for (int i = 0; i < kCompItems; i++) {
        uint64_t idx = base + (uint64_t)i * kWave + lane;
        uint32_t s = 0, d = 0;
        bool keep = false;
        if (idx < n) {
            s = src[idx];
            d = dst[idx];
            keep = s != d;
        }
        uint64_t m = __ballot(keep);
        if (keep) {
            uint64_t p = pos + (uint64_t)__popcll(m & lanemask_lt());
            src_out[p] = s;
            dst_out[p] = d;
        }
        pos += (uint64_t)__popcll(m);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *count_out = tile_off[ntiles];
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_rmat_generate(uint32_t *d_src, uint32_t *d_dst, uint64_t first_edge, uint64_t nedges,
                                  int scale, uint64_t seed, const double abcd[4], void *stream) {
    GCHI_CHECK(d_src && d_dst && abcd, GCHI_ERR_INVALID, "rmat: null pointer");
    GCHI_CHECK(scale >= 1 && scale <= 32, GCHI_ERR_INVALID, "rmat: scale must be in [1,32]");
    if (nedges == 0) return GCHI_OK;
    RmatThresholds t = rmat_thresholds(abcd);
    uint64_t key = mix64(seed);
    uint64_t blocks = std::min<uint64_t>(div_up(nedges, 256), 1u << 20);
    hipLaunchKernelGGL(rmat_kernel, dim3((unsigned)blocks), dim3(256), 0, as_stream(stream), d_src, d_dst,
                       first_edge, nedges, scale, key, t);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_rmat_generate_host(uint32_t *h_src, uint32_t *h_dst, uint64_t first_edge, uint64_t nedges,
                                       int scale, uint64_t seed, const double abcd[4], int nthreads) {
    GCHI_CHECK(h_src && h_dst && abcd, GCHI_ERR_INVALID, "rmat: null pointer");
    GCHI_CHECK(scale >= 1 && scale <= 32, GCHI_ERR_INVALID, "rmat: scale must be in [1,32]");
    RmatThresholds t = rmat_thresholds(abcd);
    uint64_t key = mix64(seed);
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    if ((uint64_t)nthreads > nedges / 4096 + 1) nthreads = (int)(nedges / 4096 + 1);
    std::vector<std::thread> th;
    for (int w = 0; w < nthreads; w++) {
        uint64_t b = nedges * w / nthreads, e = nedges * (w + 1) / nthreads;
        th.emplace_back([=]() {
            for (uint64_t i = b; i < e; i++) rmat_edge(key, first_edge + i, scale, t, &h_src[i], &h_dst[i]);
        });
    }
    for (auto &x : th) x.join();
    return GCHI_OK;
}

extern "C" int gchi_drop_self_loops_workspace(uint64_t nedges, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    uint64_t ntiles = div_up(nedges, kCompTile);
    Sizer z;
    z.take<uint32_t>(ntiles + 1);
    z.take<uint64_t>(ntiles + 1);
    z.take<char>(scan_workspace(ntiles + 1));
    *bytes = z.off;
    return GCHI_OK;
}

extern "C" int gchi_drop_self_loops(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges,
                                    uint32_t *d_src_out, uint32_t *d_dst_out, uint64_t *d_count,
                                    uint32_t *d_maxid, void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK(d_count && d_maxid, GCHI_ERR_INVALID, "drop_self_loops: null outputs");
    hipStream_t s = as_stream(stream);
    launch_fill_u32(d_maxid, 0u, 1, s);
    if (nedges == 0) {
        launch_fill_u64(d_count, 0ull, 1, s);
        GCHI_LAUNCH_CHECK();
        return GCHI_OK;
    }
    GCHI_CHECK(d_src && d_dst && d_src_out && d_dst_out, GCHI_ERR_INVALID, "drop_self_loops: null pointer");
    GCHI_CHECK(d_src_out != d_src && d_dst_out != d_dst, GCHI_ERR_INVALID, "drop_self_loops: in-place not supported");
    uint64_t ntiles = div_up(nedges, kCompTile);
    GCHI_CHECK(ntiles < (1ull << 31), GCHI_ERR_INVALID, "drop_self_loops: too many edges");
    size_t need = 0;
    gchi_drop_self_loops_workspace(nedges, &need);
    GCHI_CHECK(work_bytes >= need, GCHI_ERR_WORKSPACE, "drop_self_loops: workspace too small");
    Carver c(d_work, work_bytes);
    uint32_t *cnt = c.take<uint32_t>(ntiles + 1);
    uint64_t *off = c.take<uint64_t>(ntiles + 1);
    void *sw = c.take<char>(scan_workspace(ntiles + 1));
    hipLaunchKernelGGL(compact_count, dim3((unsigned)ntiles), dim3(kCompThreads), 0, s, d_src, d_dst, nedges,
                       cnt, d_maxid);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(exclusive_scan_u32_to_u64(cnt, off, ntiles, sw, scan_workspace(ntiles + 1), s));
    hipLaunchKernelGGL(compact_scatter, dim3((unsigned)ntiles), dim3(kCompThreads), 0, s, d_src, d_dst, nedges,
                       off, d_src_out, d_dst_out, d_count, (uint32_t)ntiles);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}
