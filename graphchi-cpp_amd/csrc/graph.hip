// Device graph construction: degree histograms and the in-edge (CSC) view.
//
// The reference keeps every interval's in-edges in its memory shard, sorted by
// src (src/preprocessing/sharder.hpp:640, src/shards/memoryshard.hpp:448-497),
// and per-vertex {indegree, outdegree} in <G>_degs.bin
// (src/engine/auxdata/degree_data.hpp:45-48).  Here the whole graph is one
// resident CSC sorted by (dst, src): in_off[N+1], in_src[E], plus degree arrays.
#include "common.hpp"
#include "internal.hpp"

namespace gchi {

// Degree histograms.  On skewed graphs a few vertices own a large share of the
// edges, and one global counter per such vertex would serialise every atomic
// on it: each workgroup first combines counts in LDS hash tables (LdsCombiner)
// and adds them to the global counts once.
constexpr int kHistThreads = 256;
constexpr int kHistSlots = 2048;

__global__ __launch_bounds__(kHistThreads) void degree_hist(const uint32_t *src, const uint32_t *dst, uint64_t n,
                                                            uint32_t nv, uint32_t *outdeg, uint32_t *indeg,
                                                            uint32_t *err) {
    __shared__ LdsCombiner<kHistSlots> c_out, c_in;
    comb_clear<kHistSlots, kHistThreads>(c_out);
    comb_clear<kHistSlots, kHistThreads>(c_in);
    __syncthreads();
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t s = src[i], d = dst[i];
        if (s >= nv || d >= nv) {
            atomicOr(err, 1u);
            continue;
        }
        if (outdeg) comb_add(c_out, s, outdeg);
        if (indeg) comb_add(c_in, d, indeg);
    }
    __syncthreads();
    if (outdeg) comb_flush<kHistSlots, kHistThreads>(c_out, outdeg);
    if (indeg) comb_flush<kHistSlots, kHistThreads>(c_in, indeg);
}

__global__ void pack_keys(const uint32_t *a, const uint32_t *b, uint64_t n, int b_bits, uint64_t *keys) {
    // key = (a << b_bits) | b
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        keys[i] = ((uint64_t)a[i] << b_bits) | (uint64_t)b[i];
}

__global__ void low_bits_u32(const uint64_t *keys, uint64_t n, int bits, uint32_t *out) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t m = bits >= 64 ? ~0ull : ((1ull << bits) - 1ull);
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        out[i] = (uint32_t)(keys[i] & m);
}

unsigned grid_for(uint64_t n, unsigned threads, unsigned cap) {
    uint64_t b = div_up(n, threads);
    if (b == 0) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

int degrees(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t nv, uint32_t *outdeg, uint32_t *indeg,
            uint32_t *err, hipStream_t s) {
    if (outdeg) launch_fill_u32(outdeg, 0u, nv, s);
    if (indeg) launch_fill_u32(indeg, 0u, nv, s);
    if (n)
        hipLaunchKernelGGL(degree_hist, dim3(grid_for(n, kHistThreads * 64, 2048)), dim3(kHistThreads), 0, s, src,
                           dst, n, nv, outdeg, indeg, err);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

template <typename C>
static void csc_layout(C &c, uint64_t E, uint32_t N, uint64_t **keys, uint32_t **indeg, uint32_t **err, void **sw,
                       size_t *sw_bytes) {
    int vb = bits_for(N ? N - 1 : 0);
    if (vb == 0) vb = 1;
    *keys = c.template take<uint64_t>(E ? E : 1);
    *indeg = c.template take<uint32_t>(N ? N : 1);
    *err = c.template take<uint32_t>(1);
    size_t ss = std::max(sort_workspace_bytes(E, 2 * vb, false), scan_workspace((size_t)N + 1));
    *sw = c.template take<char>(ss);
    *sw_bytes = ss;
}

int csc_build(const uint32_t *src, const uint32_t *dst, uint64_t E, uint32_t N, uint32_t *in_off, uint32_t *in_src,
              uint32_t *outdeg, uint32_t *indeg_out, void *work, size_t work_bytes, hipStream_t s) {
    GCHI_CHECK(E < (1ull << 32), GCHI_ERR_UNSUPPORTED, "csc_build: E must be < 2^32 (uint32 offsets)");
    GCHI_CHECK(in_off && (E == 0 || (src && dst && in_src)), GCHI_ERR_INVALID, "csc_build: null pointer");
    Carver c(work, work_bytes);
    uint64_t *keys;
    uint32_t *indeg, *err;
    void *sw;
    size_t swb;
    csc_layout(c, E, N, &keys, &indeg, &err, &sw, &swb);
    GCHI_CHECK(c.ok(), GCHI_ERR_WORKSPACE, "csc_build: workspace too small");
    uint32_t *ind = indeg_out ? indeg_out : indeg;
    launch_fill_u32(err, 0u, 1, s);
    GCHI_TRY(degrees(src, dst, E, N, outdeg, ind, err, s));
    int vb = bits_for(N ? N - 1 : 0);
    if (vb == 0) vb = 1;
    if (E) {
        hipLaunchKernelGGL(pack_keys, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, dst, src, E, vb, keys);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(radix_sort_u64(keys, nullptr, E, 2 * vb, sw, swb, s));
        hipLaunchKernelGGL(low_bits_u32, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, keys, E, vb, in_src);
        GCHI_LAUNCH_CHECK();
    }
    GCHI_TRY(exclusive_scan_u32(ind, in_off, N, sw, swb, s));
    return GCHI_OK;
}

size_t csc_workspace(uint64_t E, uint32_t N) {
    SizerAdapter a;
    uint64_t *k;
    uint32_t *i, *e;
    void *sw;
    size_t swb;
    csc_layout(a, E, N, &k, &i, &e, &sw, &swb);
    return a.z.off;
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_csc_build_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    *bytes = csc_workspace(nedges, nvertices);
    return GCHI_OK;
}

extern "C" int gchi_csc_build(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                              uint32_t *d_in_off, uint32_t *d_in_src, uint32_t *d_outdeg, uint32_t *d_indeg,
                              void *d_work, size_t work_bytes, void *stream) {
    return csc_build(d_src, d_dst, nedges, nvertices, d_in_off, d_in_src, d_outdeg, d_indeg, d_work, work_bytes,
                     as_stream(stream));
}
