// Adjacency for running a program's own update() on the device in the
// reference's parallel-sliding-windows order (engine/psw_engine.hpp).
//
// The reference stores one value per edge, shared by the edge's two ends: the
// source reaches it among its out-edges (sliding shards, sorted by src then
// dst), the destination among its in-edges (memory shard, sorted by src within
// the destination; src/preprocessing/sharder.hpp:640, src/shards/
// memoryshard.hpp:448-497).  Here the values live in one array in out-edge
// (src, dst) order, and each in-edge carries the index of its value there.
// Both orders come from stable sorts on the original edge index, so the k-th
// copy of a repeated edge is the same edge from both ends.
//
// Updates run in vertex-id order (a vertex sees what lower ids wrote this
// iteration and what higher ids wrote the iteration before): lower[v] counts
// the incident edges whose other end has a smaller id, the number of updates v
// waits for.
#include "common.hpp"
#include "internal.hpp"

namespace gchi {

// key = (a << vb) | b, value = the edge's index in the input
__global__ void psw_pack(const uint32_t *__restrict__ a, const uint32_t *__restrict__ b, uint64_t n, int vb,
                         uint32_t nv, uint64_t *__restrict__ keys, uint32_t *__restrict__ idx, uint32_t *err) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t x = a[i], y = b[i];
        if (x >= nv || y >= nv) atomicOr(err, (uint32_t)GCHI_WORK_BAD_ID);
        keys[i] = ((uint64_t)x << vb) | (uint64_t)y;
        idx[i] = (uint32_t)i;
    }
}

// out-edge p: its dst, its source's out-degree, and where input edge idx[p]
// went (out_eidx, when given: the input index of out-edge p)
__global__ void psw_out_edges(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx, uint64_t n,
                              int vb, uint32_t vmask, uint32_t nv, uint32_t *__restrict__ out_dst,
                              uint32_t *__restrict__ pos_of, uint32_t *__restrict__ outdeg,
                              uint32_t *__restrict__ out_eidx) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += stride) {
        const uint64_t k = keys[p];
        const uint32_t s = (uint32_t)(k >> vb);
        out_dst[p] = (uint32_t)k & vmask;
        pos_of[idx[p]] = (uint32_t)p;
        if (out_eidx) out_eidx[p] = idx[p];
        if (s < nv) atomicAdd(&outdeg[s], 1u);
    }
}

// in-edge q: its src, the index of its value, its destination's in-degree
__global__ void psw_in_edges(const uint64_t *__restrict__ keys, const uint32_t *__restrict__ idx, uint64_t n, int vb,
                             uint32_t vmask, uint32_t nv, const uint32_t *__restrict__ pos_of,
                             uint32_t *__restrict__ in_src, uint32_t *__restrict__ in_eidx,
                             uint32_t *__restrict__ indeg) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += stride) {
        const uint64_t k = keys[q];
        const uint32_t d = (uint32_t)(k >> vb);
        const uint32_t i = idx[q];
        in_src[q] = (uint32_t)k & vmask;
        in_eidx[q] = i < n ? pos_of[i] : 0u;
        if (d < nv) atomicAdd(&indeg[d], 1u);
    }
}

// lower[max(u, v)] += 1 for every edge u-v, u != v
__global__ void psw_lower(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint64_t n, uint32_t nv,
                          uint32_t *__restrict__ lower) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t u = src[i], v = dst[i];
        const uint32_t hi = u > v ? u : v;
        if (u != v && hi < nv) atomicAdd(&lower[hi], 1u);
    }
}

template <typename C>
static void psw_layout(C &c, uint64_t E, uint32_t N, int key_bits, uint32_t **err, uint64_t **keys, uint32_t **idx,
                       uint32_t **pos_of, uint32_t **deg, void **sw, size_t *swb) {
    *err = c.template take<uint32_t>(1);  // status word first: gchi_work_status reads it
    *keys = c.template take<uint64_t>(E ? E : 1);
    *idx = c.template take<uint32_t>(E ? E : 1);
    *pos_of = c.template take<uint32_t>(E ? E : 1);
    *deg = c.template take<uint32_t>((size_t)N + 1);
    const size_t ss = std::max(sort_workspace_bytes(E, key_bits, true), scan_workspace((size_t)N + 1));
    *sw = c.template take<char>(ss);
    *swb = ss;
}

static int psw_key_bits(uint32_t N) {
    int vb = bits_for(N ? N - 1 : 0);
    return vb == 0 ? 1 : vb;
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_psw_graph_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    SizerAdapter a;
    uint32_t *e, *i, *p, *d;
    uint64_t *k;
    void *sw;
    size_t swb;
    psw_layout(a, nedges, nvertices, 2 * psw_key_bits(nvertices), &e, &k, &i, &p, &d, &sw, &swb);
    *bytes = a.z.off;
    return GCHI_OK;
}

extern "C" int gchi_psw_graph_build(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges,
                                    uint32_t nvertices, uint32_t *d_out_off, uint32_t *d_out_dst, uint32_t *d_in_off,
                                    uint32_t *d_in_src, uint32_t *d_in_eidx, uint32_t *d_lower, void *d_work,
                                    size_t work_bytes, void *stream) {
    return gchi_psw_graph_build_ex(d_src, d_dst, nedges, nvertices, d_out_off, d_out_dst, d_in_off, d_in_src, d_in_eidx,
                                   d_lower, nullptr, d_work, work_bytes, stream);
}

extern "C" int gchi_psw_graph_build_ex(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges,
                                       uint32_t nvertices, uint32_t *d_out_off, uint32_t *d_out_dst,
                                       uint32_t *d_in_off, uint32_t *d_in_src, uint32_t *d_in_eidx, uint32_t *d_lower,
                                       uint32_t *d_out_eidx, void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK(nedges <= kMaxEdges, GCHI_ERR_UNSUPPORTED, "psw_graph_build: E must be <= 2^32 - 2^25");
    GCHI_CHECK(nvertices > 0, GCHI_ERR_INVALID, "psw_graph_build: no vertices");
    GCHI_CHECK(d_out_off && d_in_off && d_lower && d_work, GCHI_ERR_INVALID, "psw_graph_build: null pointer");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst && d_out_dst && d_in_src && d_in_eidx), GCHI_ERR_INVALID,
               "psw_graph_build: null edge arrays");
    hipStream_t s = as_stream(stream);
    const uint32_t N = nvertices;
    const int vb = psw_key_bits(N);
    const uint32_t vmask = vb >= 32 ? 0xffffffffu : ((1u << vb) - 1u);
    Carver c(d_work, work_bytes);
    uint32_t *err, *idx, *pos_of, *deg;
    uint64_t *keys;
    void *sw;
    size_t swb;
    psw_layout(c, nedges, N, 2 * vb, &err, &keys, &idx, &pos_of, &deg, &sw, &swb);
    GCHI_CHECK(c.ok(), GCHI_ERR_WORKSPACE, "psw_graph_build: workspace too small");
    launch_fill_u32(err, 0u, 1, s);
    launch_fill_u32(d_lower, 0u, N, s);
    const unsigned g = grid_for(nedges, 256, 65536);
    // out-edges: (src, dst), stable on the input order
    launch_fill_u32(deg, 0u, (size_t)N + 1, s);
    if (nedges) {
        hipLaunchKernelGGL(psw_pack, dim3(g), dim3(256), 0, s, d_src, d_dst, nedges, vb, N, keys, idx, err);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(radix_sort_u64(keys, idx, nedges, 2 * vb, sw, swb, s, err));
        hipLaunchKernelGGL(psw_out_edges, dim3(g), dim3(256), 0, s, keys, idx, nedges, vb, vmask, N, d_out_dst,
                           pos_of, deg, d_out_eidx);
        GCHI_LAUNCH_CHECK();
    }
    GCHI_TRY(exclusive_scan_u32(deg, d_out_off, N, sw, swb, s));
    // in-edges: (dst, src), stable on the input order
    launch_fill_u32(deg, 0u, (size_t)N + 1, s);
    if (nedges) {
        hipLaunchKernelGGL(psw_pack, dim3(g), dim3(256), 0, s, d_dst, d_src, nedges, vb, N, keys, idx, err);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(radix_sort_u64(keys, idx, nedges, 2 * vb, sw, swb, s, err));
        hipLaunchKernelGGL(psw_in_edges, dim3(g), dim3(256), 0, s, keys, idx, nedges, vb, vmask, N, pos_of,
                           d_in_src, d_in_eidx, deg);
        hipLaunchKernelGGL(psw_lower, dim3(g), dim3(256), 0, s, d_src, d_dst, nedges, N, d_lower);
        GCHI_LAUNCH_CHECK();
    }
    GCHI_TRY(exclusive_scan_u32(deg, d_in_off, N, sw, swb, s));
    return GCHI_OK;
}
