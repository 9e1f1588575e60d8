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

unsigned grid_for(uint64_t n, unsigned threads, unsigned cap) {
    uint64_t b = div_up(n, threads);
    if (b == 0) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// Counting by sorted runs.  Device-wide atomics on random counters run at a few
// billion per second on this part whatever the spread of the ids, so the CSC
// build counts degrees from runs of equal ids in sorted keys instead: the key
// is (dst << sb) | src with sb = vb rounded up to a whole 8-bit digit, the
// first sb/8 radix passes leave the keys ordered by src alone (out-degrees are
// the lengths of its runs), and the remaining passes finish the (dst, src)
// order (in-degrees are the lengths of the dst runs).

// run_start[f] = index of the first key of each run of equal fields f.
__global__ void field_run_starts(const uint64_t *__restrict__ k, uint64_t n, int shift, uint32_t mask, uint32_t nv,
                                 uint32_t *__restrict__ run_start) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t f = (uint32_t)(k[i] >> shift) & mask;
        const bool first = i == 0 || ((uint32_t)(k[i - 1] >> shift) & mask) != f;
        if (first && f < nv) run_start[f] = (uint32_t)i;
    }
}

// cnt[f] = (or +=, with accumulate) the length of the run of field f, written
// at its last key by one thread; fields with no run keep what cnt holds.
// low_out, when given, receives key & low_mask.
__global__ void field_run_lengths(const uint64_t *__restrict__ k, uint64_t n, int shift, uint32_t mask, uint32_t nv,
                                  const uint32_t *__restrict__ run_start, uint32_t *__restrict__ cnt,
                                  uint32_t *__restrict__ low_out, uint32_t low_mask, int accumulate) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t key = k[i];
        const uint32_t f = (uint32_t)(key >> shift) & mask;
        if (low_out) low_out[i] = (uint32_t)key & low_mask;
        const bool last = i + 1 == n || ((uint32_t)(k[i + 1] >> shift) & mask) != f;
        if (last && f < nv) {
            const uint32_t len = (uint32_t)(i + 1) - run_start[f];
            cnt[f] = accumulate ? cnt[f] + len : len;
        }
    }
}

int count_field_runs(const uint64_t *keys, uint64_t n, int shift, uint32_t mask, uint32_t nv, uint32_t *run_start,
                     uint32_t *cnt, uint32_t *low_out, uint32_t low_mask, hipStream_t s, bool accumulate) {
    if (n == 0) return GCHI_OK;
    const unsigned g = grid_for(n, 256, 65536);
    hipLaunchKernelGGL(field_run_starts, dim3(g), dim3(256), 0, s, keys, n, shift, mask, nv, run_start);
    hipLaunchKernelGGL(field_run_lengths, dim3(g), dim3(256), 0, s, keys, n, shift, mask, nv, run_start, cnt, low_out,
                       low_mask, accumulate ? 1 : 0);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

__global__ void pack_csc_keys(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t nv, int sb,
                              uint64_t *keys, uint32_t *err) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t s = src[i], d = dst[i];
        if (s >= nv || d >= nv) atomicOr(err, 1u);
        keys[i] = ((uint64_t)d << sb) | (uint64_t)s;
    }
}

struct CscBits {
    int vb;  // bits of a vertex id
    int sb;  // bits the src field takes in a key (vb rounded up to 8)
};
static CscBits csc_bits(uint32_t N) {
    int vb = bits_for(N ? N - 1 : 0);
    if (vb == 0) vb = 1;
    return CscBits{vb, (vb + 7) / 8 * 8};
}

template <typename C>
static void csc_layout(C &c, uint64_t E, uint32_t N, uint64_t **keys, uint32_t **indeg, uint32_t **run_start,
                       uint32_t **err, void **sw, size_t *sw_bytes) {
    const CscBits cb = csc_bits(N);
    *err = c.template take<uint32_t>(1);  // status word (GCHI_WORK_* bits) first: gchi_work_status reads it
    *keys = c.template take<uint64_t>(E ? E : 1);
    *indeg = c.template take<uint32_t>(N ? N : 1);
    *run_start = c.template take<uint32_t>(N ? N : 1);
    size_t ss = std::max(sort_workspace_bytes(E, std::max(cb.sb, cb.vb), false), scan_workspace((size_t)N + 1));
    *sw = c.template take<char>(ss);
    *sw_bytes = ss;
}

int csc_build(const uint32_t *src, const uint32_t *dst, uint64_t E, uint32_t N, uint32_t *in_off, uint32_t *in_src,
              uint32_t *outdeg, uint32_t *indeg_out, void *work, size_t work_bytes, hipStream_t s) {
    GCHI_CHECK(E <= kMaxEdges, GCHI_ERR_UNSUPPORTED, "csc_build: E must be <= 2^32 - 2^25 (uint32 offsets)");
    GCHI_CHECK(in_off && (E == 0 || (src && dst && in_src)), GCHI_ERR_INVALID, "csc_build: null pointer");
    Carver c(work, work_bytes);
    uint64_t *keys;
    uint32_t *indeg, *run_start, *err;
    void *sw;
    size_t swb;
    csc_layout(c, E, N, &keys, &indeg, &run_start, &err, &sw, &swb);
    GCHI_CHECK(c.ok(), GCHI_ERR_WORKSPACE, "csc_build: workspace too small");
    uint32_t *ind = indeg_out ? indeg_out : indeg;
    const CscBits cb = csc_bits(N);
    const uint32_t vmask = cb.vb >= 32 ? 0xffffffffu : ((1u << cb.vb) - 1u);
    launch_fill_u32(err, 0u, 1, s);
    if (outdeg) launch_fill_u32(outdeg, 0u, N, s);
    launch_fill_u32(ind, 0u, N, s);
    if (E) {
        const unsigned g = grid_for(E, 256, 65536);
        hipLaunchKernelGGL(pack_csc_keys, dim3(g), dim3(256), 0, s, src, dst, E, N, cb.sb, keys, err);
        GCHI_LAUNCH_CHECK();
        // src passes: keys ordered by src; out-degrees from its runs
        uint64_t *by_src = nullptr, *by_dst = nullptr;
        GCHI_TRY(radix_sort_u64_bits(keys, nullptr, nullptr, E, 0, cb.sb, sw, swb, s, &by_src, err));
        if (outdeg) GCHI_TRY(count_field_runs(by_src, E, 0, vmask, N, run_start, outdeg, nullptr, 0u, s));
        // dst passes (stable, so src order within a dst run is kept); the
        // buffer not holding by_src takes the other half of the ping-pong
        uint64_t *other = by_src == keys ? nullptr : keys;
        GCHI_TRY(radix_sort_u64_bits(by_src, other, nullptr, E, cb.sb, cb.sb + cb.vb, sw, swb, s, &by_dst, err));
        GCHI_TRY(count_field_runs(by_dst, E, cb.sb, vmask, N, run_start, ind, in_src, vmask, s));
    }
    GCHI_TRY(exclusive_scan_u32(ind, in_off, N, sw, swb, s));
    return GCHI_OK;
}

size_t csc_workspace(uint64_t E, uint32_t N) {
    SizerAdapter a;
    uint64_t *k;
    uint32_t *i, *r, *e;
    void *sw;
    size_t swb;
    csc_layout(a, E, N, &k, &i, &r, &e, &sw, &swb);
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
