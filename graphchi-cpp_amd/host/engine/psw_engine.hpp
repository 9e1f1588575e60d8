// The generic device engine: a program's own update() runs on the device for
// every vertex, in the order of the reference's parallel sliding windows
// (graphchi_engine::run / exec_updates, src/engine/graphchi_engine.hpp:396-454,
// 718-992).  Available when the application is compiled with hipcc; the
// program's member functions are then compiled for the device too
// (graphchi_basic_includes.hpp).
//
// Order.  The reference runs an iteration interval by interval, sub-interval by
// sub-interval, and a vertex sees an edge value written this iteration exactly
// when the writer has a smaller id (what src/tests/basic_smoketest.cpp:65
// checks); it sees the previous iteration's value otherwise.  Updates on
// vertices that share no edge commute, so the outcome is that of running the
// updates one after another in id order, and any schedule in which every
// vertex runs after all of its lower-id neighbours gives it exactly.  A pass
// here is one persistent kernel over a queue of ready vertices:
//   * dep[v] starts at the number of incident edges whose other end has a
//     smaller id (gchi_psw_graph_build); vertices at zero start in the queue;
//   * each lane takes the next queue slot, waits until a vertex is written
//     there, runs update() on it, then decrements dep of its higher-id
//     neighbours, queueing those that reach zero.
// The vertex in slot s is queued by an update taken from a slot below s, so the
// lane holding the lowest unfinished slot can always run: the pass ends when
// every slot is taken.  Visibility across XCDs: an update ends with a release
// fence at agent scope, its decrements are relaxed (eight in flight per lane),
// the lane whose decrement reaches zero takes an acquire fence before it
// writes the queue entry with a release store, and the lane that finds the
// entry (polling with relaxed loads) takes an acquire fence before it runs the
// vertex; so every update a vertex waited for happens before it runs.
//
// Memory.  Edge values live in one array in out-edge (src, dst) order; every
// in-edge and out-edge is a graphchi_edge view (adjacent id, value pointer)
// built once in HBM, so vertex.inedge(i) / outedge(i) return pointers into
// them.  Vertex values are an array of VertexDataType.  Both start at zero (the
// reference's engine starts from zero-filled vertex data and, for graphs
// converted without edge values, zero edge data).
//
// Not supported on the device: the scheduler (selective scheduling),
// repeat_updates(), dynamic edge data, and the reference's writing of edge
// values back into the shard files.  The iteration hooks run on the host
// between passes; before_exec_interval / after_exec_interval are called once
// per iteration with the whole vertex range.  log_change() on the device adds
// to one FP64 accumulator per pass (its order of additions is not fixed);
// <base>.deltalog and set_last_iteration() follow the reference, including its
// in-memory mode for graphs of one shard (below).
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "api/graph_objects.hpp"
#include "api/graphchi_context.hpp"
#include "api/vertex_aggregator.hpp"
#include "engine/device.hpp"
#include "logger/logger.hpp"
#include "metrics/metrics.hpp"
#include "shards/shard_reader.hpp"

namespace graphchi {
namespace psw {

constexpr uint32_t kEmpty = 0xffffffffu;
constexpr int kThreads = 256;

// assert() in an update() on the device records the failure (and the largest
// failing line) here; the engine stops with a fatal log after the pass.  A
// device trap would end the process without naming the assertion.
static __device__ unsigned int device_assert_failed;
static __device__ unsigned int device_assert_line;

GCHI_HD inline void check_assert(bool ok, const char *expr, const char *file, int line) {
#if defined(__HIP_DEVICE_COMPILE__)
    (void)expr;
    (void)file;
    if (!ok) {
        atomicOr(&device_assert_failed, 1u);
        atomicMax(&device_assert_line, (unsigned int)line);
    }
#else
    if (!ok) {
        fprintf(stderr, "%s:%d: assertion failed: %s\n", file, line, expr);
        abort();
    }
#endif
}

template <typename VT, typename ET>
struct PassArgs {
    uint32_t n;
    const uint32_t *out_off, *in_off;  // [n+1]
    const uint32_t *out_dst, *in_src;  // [E]
    graphchi_edge<ET> *out_edges, *in_edges;  // [E] views
    VT *vdata;                         // [n]
    uint32_t *dep;                     // [n] updates still awaited this pass
    uint32_t *queue;                   // [n] vertices in the order they became ready
    uint32_t *counters;                // [0] slots taken, [1] slots filled, [2] a wait gave up
    uint32_t spin_limit;               // waits with no vertex queued meanwhile before giving up
};

template <typename T>
__global__ void fill_zero(T *p, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = T();
}

template <typename ET>
__global__ void edge_views(const uint32_t *out_dst, const uint32_t *in_src, const uint32_t *in_eidx, uint64_t E,
                           ET *edata, graphchi_edge<ET> *out_edges, graphchi_edge<ET> *in_edges) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E; e += stride) {
        out_edges[e] = graphchi_edge<ET>(out_dst[e], edata + e);
        in_edges[e] = graphchi_edge<ET>(in_src[e], edata + in_eidx[e]);
    }
}

__global__ void pass_reset(const uint32_t *lower, uint32_t n, uint32_t *dep, uint32_t *queue, uint32_t *counters) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t v = gid; v < n; v += stride) {
        dep[v] = lower[v];
        queue[v] = kEmpty;
    }
    if (gid < 3) counters[gid] = 0u;
}

__global__ void pass_seed(const uint32_t *lower, uint32_t n, uint32_t *queue, uint32_t *counters) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint64_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        if (lower[v] == 0u) {
            const uint32_t s = atomicAdd(&counters[1], 1u);
            __hip_atomic_store(&queue[s], (uint32_t)v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// One decrement of dep[w] for each neighbour id w > v in nb[0, cnt), eight in
// flight; a vertex whose count reaches zero is queued.
template <typename VT, typename ET>
__device__ inline void release_higher(const PassArgs<VT, ET> &a, uint32_t v, const uint32_t *nb, uint32_t cnt) {
    constexpr int kBatch = 8;
    for (uint32_t k0 = 0; k0 < cnt; k0 += kBatch) {
        uint32_t w[kBatch], left[kBatch];
#pragma unroll
        for (int q = 0; q < kBatch; q++) w[q] = k0 + q < cnt ? nb[k0 + q] : 0u;  // 0 is never above v
#pragma unroll
        for (int q = 0; q < kBatch; q++)
            left[q] = w[q] > v ? __hip_atomic_fetch_add(&a.dep[w[q]], 0xffffffffu, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                               : 0u;
#pragma unroll
        for (int q = 0; q < kBatch; q++) {
            if (w[q] > v && left[q] == 1u) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                const uint32_t s = atomicAdd(&a.counters[1], 1u);
                __hip_atomic_store(&a.queue[s], w[q], __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// One pass.  Every lane loops on its own: take a slot, run the vertex queued
// there once it is, repeat; lanes whose vertex is not queued yet wait while the
// others of their wavefront go on.
template <class P, typename VT, typename ET>
__global__ __launch_bounds__(kThreads) void pass_run(P prog, PassArgs<VT, ET> a, graphchi_context *ctx) {
    uint32_t slot = 0, spins = 0, seen = 0;
    bool have = false;
    for (;;) {
        if (!have) {
            slot = atomicAdd(&a.counters[0], 1u);
            have = true;
        }
        if (slot >= a.n) break;
        // polled with a relaxed load: an acquire on every poll invalidates the
        // L2 each time, and thousands of waiting lanes made that 14x slower
        const uint32_t v = __hip_atomic_load(&a.queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (v == kEmpty) {
            const uint32_t filled = __hip_atomic_load(&a.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (filled != seen) {
                seen = filled;
                spins = 0;
            } else if (++spins > a.spin_limit) {
                atomicOr(&a.counters[2], 1u);
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // pairs with the queue entry's release store
        const uint32_t ib = a.in_off[v], ie = a.in_off[v + 1], ob = a.out_off[v], oe = a.out_off[v + 1];
        graphchi_vertex<VT, ET> vx(v, a.in_edges + ib, a.out_edges + ob, (int)(ie - ib), (int)(oe - ob));
        vx.dataptr = a.vdata + v;
        prog.P::update(vx, *ctx);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        release_higher(a, v, a.in_src + ib, ie - ib);
        release_higher(a, v, a.out_dst + ob, oe - ob);
        have = false;
    }
}

// <base>.deltalog: iteration, updates, work, change (graphchi_engine.hpp:593-599)
inline void write_delta_log(const std::string &base, const graphchi_context &ctx, int iteration, uint64_t updates,
                            uint64_t work) {
    if (FILE *df = fopen((base + ".deltalog").c_str(), iteration == 0 ? "w" : "a")) {
        fprintf(df, "%d,%lu,%lu,%lf\n", iteration, (unsigned long)updates, (unsigned long)work, ctx.get_delta());
        fclose(df);
    }
}

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) logstream(LOG_FATAL) << what << ": " << hipGetErrorString(e) << std::endl;
}

inline unsigned grid_for(uint64_t n, unsigned cap) {
    uint64_t b = (n + kThreads - 1) / kThreads;
    if (b == 0) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

// Runs niters iterations of program on the graph stored in base's shards and
// writes the vertex values to <base>.<sizeof VT>B.vout.
template <typename VT, typename ET, class ProgramT>
void run_program(ProgramT &program, const std::string &base, int nshards, uint32_t n, int niters,
                 graphchi_context &ctx, metrics &m, size_t &nedges_out) {
    static_assert(std::is_trivially_copyable<VT>::value && std::is_trivially_copyable<ET>::value,
                  "device vertex programs need trivially copyable vertex and edge data");
    device::open();
    metrics_entry me = m.start_time();
    device::buffer<uint32_t> src, dst;
    const uint64_t E = load_shard_edges_device(base, nshards, src, dst);
    nedges_out = E;
    ctx.nedges = E;
    const uint64_t E1 = E ? E : 1;
    device::buffer<uint32_t> out_off((size_t)n + 1), in_off((size_t)n + 1), out_dst(E1), in_src(E1), in_eidx(E1),
        lower(n);
    {
        size_t wb = 0;
        device::check(gchi_psw_graph_workspace(E, n, &wb), "gchi_psw_graph_workspace");
        device::workspace w(wb);
        device::check(gchi_psw_graph_build(src.get(), dst.get(), E, n, out_off.get(), out_dst.get(), in_off.get(),
                                           in_src.get(), in_eidx.get(), lower.get(), w.get(), wb, nullptr),
                      "gchi_psw_graph_build");
        device::check(gchi_work_status(w.get(), nullptr), "gchi_psw_graph_build");
    }
    src.release();
    dst.release();
    device::buffer<ET> edata(E1);
    device::buffer<VT> vdata(n);
    device::buffer<graphchi_edge<ET>> out_edges(E1), in_edges(E1);
    device::buffer<uint32_t> dep(n), queue(n), counters(4);
    device::buffer<graphchi_context> d_ctx(1);
    device::buffer<double> d_delta(1);
    ET *ed = edata.get();
    VT *vd = vdata.get();
    graphchi_edge<ET> *oe = out_edges.get(), *ie = in_edges.get();
    const uint32_t *od = out_dst.get(), *is = in_src.get(), *ix = in_eidx.get(), *lw = lower.get();
    uint32_t *dp = dep.get(), *qu = queue.get(), *ct = counters.get();
    graphchi_context *dctx = d_ctx.get();
    hipLaunchKernelGGL(fill_zero<ET>, dim3(grid_for(E, 4096)), dim3(kThreads), 0, 0, ed, (uint64_t)E);
    hipLaunchKernelGGL(fill_zero<VT>, dim3(grid_for(n, 4096)), dim3(kThreads), 0, 0, vd, (uint64_t)n);
    if (E)
        hipLaunchKernelGGL(edge_views<ET>, dim3(grid_for(E, 4096)), dim3(kThreads), 0, 0, od, is, ix, (uint64_t)E, ed,
                           oe, ie);
    hip_check(hipGetLastError(), "psw setup");
    m.stop_time(me, "load_shards");

    PassArgs<VT, ET> a{n, out_off.get(), in_off.get(), od, is, oe, ie, vd, dp, qu, ct, 1u << 24};
    int dev = 0, ncu = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        ncu = prop.multiProcessorCount;
    const unsigned grid = grid_for(n, (unsigned)ncu * 8);
    // The reference runs a graph of one shard (and fewer than 2 * 40M vertices)
    // in its in-memory mode (graphchi_engine.hpp:296-301, 465-523): all
    // iterations inside one outer iteration, so log_change() accumulates over
    // the whole run, the exec-interval hooks get (0, N), a last iteration ends
    // the run only when it is above 0, and <base>.deltalog gets one line at the
    // end.  Otherwise every iteration starts from zero change and logs a line.
    const bool inmem = nshards == 1 && n < 2u * 40000000u;
    const vid_t win_en = inmem ? n : (n ? n - 1 : 0);
    uint64_t updates = 0, work = 0;
    int last_done = -1;
    metrics_entry mu = m.start_time();
    if (inmem) ctx.reset_deltas(1);
    for (int iter = 0; iter < niters; iter++) {
        ctx.iteration = iter;
        if (!inmem) ctx.reset_deltas(1);
        program.before_iteration(iter, ctx);
        program.before_exec_interval(0, win_en, ctx);
        graphchi_context dc = ctx;  // the plain fields, for update() on the device
        dc.device_delta = d_delta.get();
        const double zero = 0.0;
        const unsigned int uzero = 0;
        d_ctx.upload(&dc, 1);
        d_delta.upload(&zero, 1);
        hip_check(hipMemcpyToSymbol(HIP_SYMBOL(device_assert_failed), &uzero, 4), "hipMemcpyToSymbol");
        hip_check(hipMemcpyToSymbol(HIP_SYMBOL(device_assert_line), &uzero, 4), "hipMemcpyToSymbol");
        hipLaunchKernelGGL(pass_reset, dim3(grid_for(n, 4096)), dim3(kThreads), 0, 0, lw, n, dp, qu, ct);
        hipLaunchKernelGGL(pass_seed, dim3(grid_for(n, 4096)), dim3(kThreads), 0, 0, lw, n, qu, ct);
        if (n) hipLaunchKernelGGL((pass_run<ProgramT, VT, ET>), dim3(grid), dim3(kThreads), 0, 0, program, a, dctx);
        hip_check(hipGetLastError(), "psw pass launch");
        hip_check(hipDeviceSynchronize(), "psw pass");
        const std::vector<uint32_t> cnt = counters.to_host(3);
        if (cnt[2]) logstream(LOG_FATAL) << "psw pass: a wait for a queued vertex gave up" << std::endl;
        unsigned int failed = 0, line = 0;
        hip_check(hipMemcpyFromSymbol(&failed, HIP_SYMBOL(device_assert_failed), 4), "hipMemcpyFromSymbol");
        hip_check(hipMemcpyFromSymbol(&line, HIP_SYMBOL(device_assert_line), 4), "hipMemcpyFromSymbol");
        if (failed)
            logstream(LOG_FATAL) << "assertion failed in update() on the device (line " << line << "), iteration "
                                 << iter << std::endl;
        double delta = 0.0;
        d_delta.download(&delta, 1);
        ctx.log_change(delta);
        updates += n;
        work += inmem ? E : 2 * E;  // the reference's count: shard edges, or both ends of every edge
        program.after_exec_interval(0, win_en, ctx);
        program.after_iteration(iter, ctx);
        last_done = iter;
        if (!inmem) write_delta_log(base, ctx, iter, updates, work);
        if (inmem ? (ctx.last_iteration > 0 && ctx.last_iteration <= iter)
                  : (ctx.last_iteration >= 0 && iter >= ctx.last_iteration))
            break;
    }
    if (inmem && last_done >= 0) write_delta_log(base, ctx, last_done, updates, work);
    m.stop_time(mu, "execute-updates");
    m.set("updates", (double)updates);
    write_vertex_values<VT>(base, vdata.to_host(n));
}

}  // namespace psw
}  // namespace graphchi
#endif  // __HIPCC__
