// Connected components by minimum label.
//
// Reference: ConnectedComponentsProgram::update (example_apps/connectedcomponents.cpp:79-121)
// propagates min labels over in- and out-edges until no edge label changes
// (:125-138); its fixed point is, for every vertex, the smallest vertex id of
// its weakly connected component.  Two device forms:
//   gchi_cc_minlabel  : lock-free union-find over the edge list.  A root is
//       always hooked under a smaller root with CAS, so every root is the
//       smallest id of its set; a final pass writes each vertex's root.  Result
//       identical to the reference's fixed point, independent of scheduling.
//   gchi_cc_labelprop_sweep : one synchronous min-propagation sweep
//       (label_out[v] = min(label_in[v], label_in[u] for every edge u-v)) with a
//       count of changed labels -- the reference's per-iteration semantics.
#include "common.hpp"
#include "internal.hpp"

namespace gchi {

// Follows parent pointers to a root.  Loads may be stale across workgroups:
// every value parent[x] ever held is an ancestor of x in the same set, so a
// stale walk still ends at a member's ancestor; progress is guaranteed by the
// CAS below, which returns the current value.
GCHI_DEV uint32_t uf_find(uint32_t *parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        uint32_t gp = parent[p];
        if (gp != p) parent[x] = gp;  // path halving; gp is an ancestor of x
        x = p;
        p = gp;
    }
    return x;
}

__global__ void cc_init(uint32_t *parent, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) parent[v] = v;
}

// One union per edge.  With `half` (a symmetric list: every u-v present as both
// (u,v) and (v,u)) only the entries with u < v are used: the other half would
// union the same sets again.
__global__ void cc_hook(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint64_t e,
                        uint32_t n, uint32_t *parent, int half) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t u = __builtin_nontemporal_load(&src[i]);
        uint32_t v = __builtin_nontemporal_load(&dst[i]);
        if (u >= n || v >= n) continue;  // ids beyond N are rejected on the host side
        if (half && u >= v) continue;
        u = uf_find(parent, u);
        v = uf_find(parent, v);
        while (u != v) {
            if (u < v) {
                uint32_t t = u;
                u = v;
                v = t;
            }
            // u > v: hook root u under v
            uint32_t old = atomicCAS(&parent[u], u, v);
            if (old == u) break;
            u = uf_find(parent, old);
            v = uf_find(parent, v);
        }
    }
}

__global__ void cc_compress(uint32_t *parent, uint32_t *labels, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        uint32_t r = v, p = parent[v];
        while (p != r) {
            r = p;
            p = parent[r];
        }
        labels[v] = r;
    }
}

__global__ void cc_copy(const uint32_t *a, uint32_t *b, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) b[v] = a[v];
}

__global__ void cc_prop_edges(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint64_t e,
                              uint32_t n, const uint32_t *__restrict__ lin, uint32_t *lout) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t u = __builtin_nontemporal_load(&src[i]);
        uint32_t v = __builtin_nontemporal_load(&dst[i]);
        if (u >= n || v >= n) continue;
        uint32_t lu = lin[u], lv = lin[v];
        if (lu < lv) atomicMin(&lout[v], lu);
        else if (lv < lu) atomicMin(&lout[u], lv);
    }
}

__global__ void cc_count_changed(const uint32_t *lin, const uint32_t *lout, uint32_t n, uint64_t *changed) {
    uint32_t c = 0;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) c += lin[v] != lout[v];
    c = wave_sum(c);
    if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd((unsigned long long *)changed, (unsigned long long)c);
}

__global__ void cc_iota(uint32_t *l, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) l[v] = v;
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_cc_workspace(uint32_t nvertices, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    Sizer z;
    z.take<uint32_t>(nvertices ? nvertices : 1);
    *bytes = z.off;
    return GCHI_OK;
}

extern "C" int gchi_cc_minlabel_ex(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                                   int flags, uint32_t *d_labels, void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK((flags & ~GCHI_CC_SYMMETRIC) == 0, GCHI_ERR_INVALID, "cc_minlabel: unknown flags");
    GCHI_CHECK(d_labels && nvertices > 0, GCHI_ERR_INVALID, "cc_minlabel: null labels or no vertices");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "cc_minlabel: null edges");
    size_t need = 0;
    gchi_cc_workspace(nvertices, &need);
    GCHI_CHECK(d_work && work_bytes >= need, GCHI_ERR_WORKSPACE, "cc_minlabel: workspace too small");
    hipStream_t s = as_stream(stream);
    uint32_t *parent = static_cast<uint32_t *>(d_work);
    hipLaunchKernelGGL(cc_init, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, s, parent, nvertices);
    if (nedges)
        hipLaunchKernelGGL(cc_hook, dim3(grid_for(nedges, 256, 1u << 16)), dim3(256), 0, s, d_src, d_dst, nedges,
                           nvertices, parent, (flags & GCHI_CC_SYMMETRIC) ? 1 : 0);
    hipLaunchKernelGGL(cc_compress, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, s, parent, d_labels,
                       nvertices);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_cc_minlabel(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                                uint32_t *d_labels, void *d_work, size_t work_bytes, void *stream) {
    return gchi_cc_minlabel_ex(d_src, d_dst, nedges, nvertices, 0, d_labels, d_work, work_bytes, stream);
}

extern "C" int gchi_cc_labelprop_init(uint32_t *d_labels, uint32_t nvertices, void *stream) {
    GCHI_CHECK(d_labels, GCHI_ERR_INVALID, "cc_labelprop_init: null labels");
    hipLaunchKernelGGL(cc_iota, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, as_stream(stream), d_labels,
                       nvertices);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_cc_labelprop_sweep(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges,
                                       uint32_t nvertices, const uint32_t *d_labels_in, uint32_t *d_labels_out,
                                       uint64_t *d_changed, void *stream) {
    GCHI_CHECK(d_labels_in && d_labels_out && d_changed && d_labels_in != d_labels_out, GCHI_ERR_INVALID,
               "cc_labelprop_sweep: bad label buffers");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "cc_labelprop_sweep: null edges");
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(cc_copy, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, s, d_labels_in, d_labels_out,
                       nvertices);
    if (nedges)
        hipLaunchKernelGGL(cc_prop_edges, dim3(grid_for(nedges, 256, 1u << 16)), dim3(256), 0, s, d_src, d_dst,
                           nedges, nvertices, d_labels_in, d_labels_out);
    launch_fill_u64(d_changed, 0ull, 1, s);
    hipLaunchKernelGGL(cc_count_changed, dim3(grid_for(nvertices, 256, 4096)), dim3(256), 0, s, d_labels_in,
                       d_labels_out, nvertices, d_changed);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}
