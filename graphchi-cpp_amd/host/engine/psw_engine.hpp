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
//   * each wavefront takes the next queue slot, waits until a vertex is
//     written there, runs update() on it (one lane), then decrements dep of
//     its higher-id neighbours (all lanes), queueing those that reach zero.
// The vertex in slot s is queued by an update taken from a slot below s, so the
// wavefront holding the lowest unfinished slot can always run: the pass ends
// when every slot is taken.  Visibility across XCDs: an update ends with a
// release fence at agent scope, its decrements are relaxed (eight in flight
// per lane), the lane whose decrement reaches zero takes an acquire fence
// before it writes the queue entry with a release store, and the wavefront
// that finds the entry (polling with relaxed loads) takes an acquire fence
// before it runs the vertex; so every update a vertex waited for happens
// before it runs.
//
// Memory.  Every edge has two graphchi_edge views built once in HBM, the
// out-edge in (src, dst) order and the in-edge in (dst, src) order; each holds
// the adjacent id, a copy of the value and a pointer to the other view's copy,
// so vertex.inedge(i) / outedge(i) return pointers into them and a read is one
// load next to the vertex's other views.  A vertex with up to 4096 views (64 KB)
// has them copied into its wavefront's LDS for its update() and copied back
// after it.  Vertex values are an array of
// VertexDataType, starting at zero (the reference's engine starts from
// zero-filled vertex data).
//
// Selective scheduling (src/engine/bitset_scheduler.hpp): two bitsets in HBM,
// this iteration's tasks and the next's; update() reaches them through a
// scheduler object built on the device.  As in the reference, a vertex's
// scheduled flag is read when its interval starts (the reference reads it
// when it loads the sub-interval), so with scheduling on, a graph of several
// shards runs one pass per interval; add_task(v, true) then runs v in this
// iteration exactly when v's interval has not started.
//
// Edge values start from the shard files' edge-data blocks (zlib-inflated on
// the host) and, for a graph of several shards, are written back there after
// the run, as the reference's engine commits them; a run on the same shards
// then starts from them.  Not supported on the device: repeat_updates() and
// dynamic edge data.  The iteration hooks run on the host
// between passes; before_exec_interval / after_exec_interval are called once
// per iteration with the whole vertex range.  log_change() on the device adds
// to one FP64 accumulator per pass (its order of additions is not fixed);
// <base>.deltalog and set_last_iteration() follow the reference, including its
// in-memory mode for graphs of one shard (below).
#pragma once
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "api/chifilenames.hpp"
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
constexpr int kWaveLanes = 64;

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

inline void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) logstream(LOG_FATAL) << what << ": " << hipGetErrorString(e) << std::endl;
}

template <typename VT, typename ET>
struct PassArgs {
    uint32_t vlo, vhi;                 // the pass runs vertices [vlo, vhi)
    const uint32_t *out_off, *in_off;  // [n+1]
    const uint32_t *out_dst, *in_src;  // [E]
    graphchi_edge<ET> *out_edges, *in_edges;  // [E] views
    VT *vdata;                         // [n]
    uint32_t *dep;                     // [n] updates still awaited this pass
    const uint8_t *sched;              // [n] 1: update() runs this pass (null: every vertex)
    uint32_t *queue;                   // [vhi - vlo] vertices in the order they became ready
    uint32_t *counters;                // [0] slots taken, [1] slots filled, [2] a wait gave up
    uint32_t spin_limit;               // waits with no vertex queued meanwhile before giving up
};

template <typename T>
__global__ void fill_zero(T *p, uint64_t n) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = T();
}

// The two views of every edge, each the other's twin, holding the edge's
// value from the shard files (val_in in shard order; out-edge p is input edge
// out_eidx[p]).  in_eidx is a permutation: in-edge q is out-edge in_eidx[q].
template <typename ET>
__global__ void edge_views(const uint32_t *out_dst, const uint32_t *in_src, const uint32_t *in_eidx,
                           const uint32_t *out_eidx, const ET *val_in, uint64_t E, graphchi_edge<ET> *out_edges,
                           graphchi_edge<ET> *in_edges) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < E; q += stride) {
        const uint32_t p = in_eidx[q];
        const ET x = val_in[out_eidx[p]];
        in_edges[q] = graphchi_edge<ET>(in_src[q], x, &out_edges[p].value);
        out_edges[p] = graphchi_edge<ET>(out_dst[p], x, &in_edges[q].value);
    }
}

// The edge values back in shard order, for writing them to the shard files.
template <typename ET>
__global__ void edge_values_out(const graphchi_edge<ET> *out_edges, const uint32_t *out_eidx, uint64_t E, ET *val_out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < E; p += stride)
        val_out[out_eidx[p]] = out_edges[p].value;
}

#ifndef GRAPHCHI_DISABLE_COMPRESSION
constexpr int kEdataZlib = 1;  // the edge-data files' form, as filename_shard_edata names them
#else
constexpr int kEdataZlib = 0;
#endif

// dep[v] = the updates v waits for in a pass over [vlo, vhi): its lower-id
// incident edges, less those to vertices below vlo (cross, when given: done in
// an earlier pass)
__global__ void pass_reset(const uint32_t *lower, const uint32_t *cross, uint32_t vlo, uint32_t vhi, uint32_t *dep,
                           uint32_t *queue, uint32_t *counters) {
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint64_t v = (uint64_t)vlo + gid; v < vhi; v += stride) {
        dep[v] = lower[v] - (cross ? cross[v] : 0u);
        queue[v - vlo] = kEmpty;
    }
    if (gid < 3) counters[gid] = 0u;
}

__global__ void pass_seed(const uint32_t *dep, uint32_t vlo, uint32_t vhi, uint32_t *queue, uint32_t *counters) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)vlo + blockIdx.x * blockDim.x + threadIdx.x; v < vhi; v += stride) {
        if (dep[v] == 0u) {
            const uint32_t s = atomicAdd(&counters[1], 1u);
            __hip_atomic_store(&queue[s], (uint32_t)v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// cross[v] for v in [vlo, vhi) = incident edges of v whose other end is below vlo
__global__ void count_cross(const uint32_t *in_off, const uint32_t *in_src, const uint32_t *out_off,
                            const uint32_t *out_dst, uint32_t vlo, uint32_t vhi, uint32_t *cross) {
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint64_t v = (uint64_t)vlo + blockIdx.x * blockDim.x + threadIdx.x; v < vhi; v += stride) {
        const uint32_t lo = vlo;
        uint32_t c = 0;
        for (uint32_t k = in_off[v]; k < in_off[v + 1]; k++) c += in_src[k] < lo ? 1u : 0u;
        for (uint32_t k = out_off[v]; k < out_off[v + 1]; k++) c += out_dst[k] < lo ? 1u : 0u;
        cross[v] = c;
    }
}

// The scheduler update() sees on the device: bitset_scheduler.hpp:38-104 with a
// byte per vertex instead of a bit, so that concurrent updates set flags with
// plain stores.  cur holds this iteration's tasks, next the next iteration's,
// has_new the reference's has_new_tasks.  Built in device memory by a kernel,
// so its virtual functions dispatch to their device code.  With cur null it
// is the reference's non_scheduler (nothing recorded, every vertex scheduled).
class device_scheduler : public ischeduler {
  public:
    GCHI_HD device_scheduler(uint8_t *cur, uint8_t *next, uint32_t *has_new, uint32_t n)
        : cur_(cur), next_(next), has_new_(has_new), n_(n) {}
    GCHI_HD void add_task(vid_t v, bool also_this_iteration = false) override {
        if (!cur_ || v >= n_) return;
        next_[v] = 1;
        if (also_this_iteration) cur_[v] = 1;
        *has_new_ = 1u;
    }
    GCHI_HD void remove_tasks(vid_t from, vid_t to) override {  // [from, to], as dense_bitset::clear_bits
        if (!cur_) return;
        for (uint64_t v = from; v <= to && v < n_; v++) next_[v] = 0;
    }
    GCHI_HD void add_task_to_all() override {
        if (!cur_) return;
        for (uint32_t v = 0; v < n_; v++) cur_[v] = 1;
        *has_new_ = 1u;
    }
    GCHI_HD bool is_scheduled(vid_t v) override { return !cur_ || (v < n_ && cur_[v]); }
    GCHI_HD size_t num_tasks() override {
        size_t c = 0;
        if (cur_)
            for (uint32_t v = 0; v < n_; v++) c += cur_[v];
        return c;
    }

  private:
    uint8_t *cur_, *next_;
    uint32_t *has_new_;
    uint32_t n_;
};

__global__ void make_device_scheduler(device_scheduler *p, uint8_t *cur, uint8_t *next, uint32_t *has_new,
                                      uint32_t n) {
    if (blockIdx.x == 0 && threadIdx.x == 0) new (p) device_scheduler(cur, next, has_new, n);
}

// The pass's scheduled flags, read when it starts (the reference reads them
// when it loads the sub-interval, graphchi_engine.hpp:556-566): snap[v] =
// cur[v] (every vertex when cur is null) for v in [vlo, vhi); stats[0] += the
// scheduled vertices, stats[1] += their in- and out-edges (the reference's
// nupdates and work).
__global__ void snapshot_schedule(const uint8_t *cur, const uint32_t *in_off, const uint32_t *out_off, uint32_t vlo,
                                  uint32_t vhi, uint8_t *snap, unsigned long long *stats) {
    const uint32_t stride = gridDim.x * blockDim.x;
    unsigned long long cnt = 0, edges = 0;
    for (uint64_t v = (uint64_t)vlo + blockIdx.x * blockDim.x + threadIdx.x; v < vhi; v += stride) {
        const uint8_t s = cur ? cur[v] : (uint8_t)1;
        snap[v] = s;
        if (s) {
            cnt++;
            edges += (in_off[v + 1] - in_off[v]) + (out_off[v + 1] - out_off[v]);
        }
    }
    if (cnt) {
        atomicAdd(&stats[0], cnt);
        atomicAdd(&stats[1], edges);
    }
}

// The scheduler the program's hooks see on the host.  The flags stay in HBM:
// the first call after device work reads them back, and the engine writes
// changed ones back before the next device step.
class host_scheduler : public ischeduler {
  public:
    uint8_t *cur = nullptr, *next = nullptr;
    uint32_t *has_new = nullptr;
    uint32_t n = 0;

    // host-and-device only because the base's virtuals are; the device never
    // holds this object (its bodies are empty in the device pass)
    GCHI_HD void add_task(vid_t v, bool also_this_iteration = false) override {
#if !defined(__HIP_DEVICE_COMPILE__)
        if (v >= n) return;
        pull();
        hn_[v] = 1;
        if (also_this_iteration) hc_[v] = 1;
        hf_ = 1u;
        dirty_ = true;
#endif
    }
    GCHI_HD void remove_tasks(vid_t from, vid_t to) override {
#if !defined(__HIP_DEVICE_COMPILE__)
        pull();
        for (uint64_t v = from; v <= to && v < n; v++) hn_[v] = 0;
        dirty_ = true;
#endif
    }
    GCHI_HD void add_task_to_all() override {
#if !defined(__HIP_DEVICE_COMPILE__)
        pull();
        std::fill(hc_.begin(), hc_.end(), (uint8_t)1);
        hf_ = 1u;
        dirty_ = true;
#endif
    }
    GCHI_HD bool is_scheduled(vid_t v) override {
#if !defined(__HIP_DEVICE_COMPILE__)
        pull();
        return v < n && hc_[v];
#else
        return false;
#endif
    }
    GCHI_HD size_t num_tasks() override {
        size_t c = 0;
#if !defined(__HIP_DEVICE_COMPILE__)
        pull();
        for (uint8_t x : hc_) c += x;
#endif
        return c;
    }
    // the engine's side
    void push() {
        if (!dirty_) return;
        hip_check(hipMemcpy(cur, hc_.data(), n, hipMemcpyHostToDevice), "scheduler upload");
        hip_check(hipMemcpy(next, hn_.data(), n, hipMemcpyHostToDevice), "scheduler upload");
        hip_check(hipMemcpy(has_new, &hf_, 4, hipMemcpyHostToDevice), "scheduler upload");
        dirty_ = false;
    }
    void stale() {  // device work may have changed the flags
        push();
        valid_ = false;
    }

  private:
    std::vector<uint8_t> hc_, hn_;
    uint32_t hf_ = 0;
    bool valid_ = false, dirty_ = false;
    void pull() {
        if (valid_) return;
        hc_.resize(n);
        hn_.resize(n);
        hip_check(hipMemcpy(hc_.data(), cur, n, hipMemcpyDeviceToHost), "scheduler download");
        hip_check(hipMemcpy(hn_.data(), next, n, hipMemcpyDeviceToHost), "scheduler download");
        hip_check(hipMemcpy(&hf_, has_new, 4, hipMemcpyDeviceToHost), "scheduler download");
        valid_ = true;
    }
};

// One decrement of dep[w] for each neighbour id w > v in nb[0, cnt) that
// belongs to this pass, spread over the wavefront's lanes with eight in flight
// per lane; a vertex whose count reaches zero is queued.
constexpr int kReleaseBatch = 8;
template <typename VT, typename ET>
__device__ inline void release_higher(const PassArgs<VT, ET> &a, uint32_t v, const uint32_t *nb, uint32_t cnt,
                                      uint32_t lane) {
    for (uint32_t k0 = 0; k0 < cnt; k0 += kWaveLanes * kReleaseBatch) {
        uint32_t w[kReleaseBatch], left[kReleaseBatch];
#pragma unroll
        for (int q = 0; q < kReleaseBatch; q++) {
            const uint32_t k = k0 + (uint32_t)q * kWaveLanes + lane;
            w[q] = k < cnt ? nb[k] : 0u;  // 0 is never above v
            if (w[q] >= a.vhi) w[q] = 0u;  // a later pass's vertex
        }
#pragma unroll
        for (int q = 0; q < kReleaseBatch; q++)
            left[q] = w[q] > v ? __hip_atomic_fetch_add(&a.dep[w[q]], 0xffffffffu, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                               : 0u;
#pragma unroll
        for (int q = 0; q < kReleaseBatch; q++) {
            if (w[q] > v && left[q] == 1u) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
                const uint32_t s = atomicAdd(&a.counters[1], 1u);
                __hip_atomic_store(&a.queue[s], w[q], __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
}

// Loads every 64-byte line of a vertex's views across the wavefront (16 per
// lane in flight) before update() walks them one edge at a time on one lane.
template <typename ET>
__device__ inline void prefetch_views(const graphchi_edge<ET> *e, uint32_t cnt, uint32_t lane) {
    const char *p = reinterpret_cast<const char *>(e);
    const uint64_t lines = ((uint64_t)cnt * sizeof(graphchi_edge<ET>) + 63) / 64;
    uint32_t acc = 0;
    for (uint64_t l0 = 0; l0 < lines; l0 += 16 * kWaveLanes) {
        uint32_t x[16];
#pragma unroll
        for (int j = 0; j < 16; j++) {
            const uint64_t l = l0 + (uint64_t)j * kWaveLanes + lane;
            x[j] = l < lines ? *reinterpret_cast<const uint32_t *>(p + l * 64) : 0u;
        }
#pragma unroll
        for (int j = 0; j < 16; j++) acc += x[j];
    }
    asm volatile("" ::"v"(acc));  // the loads complete here
}

// A vertex's views copied into the wavefront's LDS area for its update(), when
// they fit: update() then walks them at LDS latency.  Twins that point into the
// vertex's own views (a self-loop) are pointed at the copies.
constexpr uint32_t kStageBytes = 65536;  // per wavefront
constexpr int kPassThreads = 128;        // two wavefronts, 128 KB of LDS per workgroup
template <typename ET>
__device__ inline uint32_t stage_capacity() {
    return kStageBytes / (uint32_t)sizeof(graphchi_edge<ET>);
}

template <typename ET>
__device__ inline bool in_views(const ET *t, graphchi_edge<ET> *e, uint32_t cnt, uint32_t *at) {
    const char *b = reinterpret_cast<const char *>(e), *x = reinterpret_cast<const char *>(t);
    if (x < b || x >= b + (uint64_t)cnt * sizeof(graphchi_edge<ET>)) return false;
    *at = (uint32_t)((uint64_t)(x - b) / sizeof(graphchi_edge<ET>));
    return true;
}

template <typename ET>
__device__ inline void stage_in(graphchi_edge<ET> *stage, graphchi_edge<ET> *in, uint32_t nin, graphchi_edge<ET> *out,
                                uint32_t nout, uint32_t lane) {
    for (uint32_t i = lane; i < nin + nout; i += kWaveLanes) {
        graphchi_edge<ET> e = i < nin ? in[i] : out[i - nin];
        uint32_t at;
        if (in_views(e.twin, in, nin, &at)) e.twin = &stage[at].value;
        else if (in_views(e.twin, out, nout, &at)) e.twin = &stage[nin + at].value;
        stage[i] = e;
    }
}

// The copies back to their places after update() (which may have reordered
// them, sort_edges_indirect): each view goes to the slot of its position, and
// the view its twin belongs to is pointed at it.
template <typename ET>
__device__ inline void stage_out(graphchi_edge<ET> *stage, graphchi_edge<ET> *in, uint32_t nin,
                                 graphchi_edge<ET> *out, uint32_t nout, uint32_t lane) {
    for (uint32_t i = lane; i < nin + nout; i += kWaveLanes) {
        graphchi_edge<ET> e = stage[i];
        graphchi_edge<ET> *g = i < nin ? in + i : out + (i - nin);
        uint32_t at;
        const bool self = in_views(e.twin, stage, nin + nout, &at);
        if (self) e.twin = at < nin ? &in[at].value : &out[at - nin].value;
        // a view that moved: the other end's view follows it
        if (!self && e.twin && e.twin != g->twin) graphchi_edge<ET>::view_of(e.twin)->twin = &g->value;
        *g = e;
    }
}

// One pass, one worker per wavefront: lane 0 takes the next slot and waits
// until a vertex is queued there, the wavefront stages its views in LDS, lane 0
// runs update() on it, and the whole wavefront copies the views back and
// decrements the counts of its higher-id neighbours.  A
// wavefront runs one update at a time: with a vertex per lane, a lane whose
// vertex was ready waited for the other lanes' updates (a hub's included),
// since the lanes of a wavefront advance together.  The fences are
// wavefront-wide instructions, so the release after lane 0's update orders its
// stores before every lane's decrements.
template <class P, typename VT, typename ET>
__global__ __launch_bounds__(kPassThreads) void pass_run(P prog, PassArgs<VT, ET> a, graphchi_context *ctx) {
    const uint32_t lane = threadIdx.x & (kWaveLanes - 1);
    const uint32_t nslots = a.vhi - a.vlo;
    HIP_DYNAMIC_SHARED(char, lds_raw)
    graphchi_edge<ET> *stage =
        reinterpret_cast<graphchi_edge<ET> *>(lds_raw + (size_t)(threadIdx.x / kWaveLanes) * kStageBytes);
    uint32_t spins = 0, seen = 0;
    for (;;) {
        uint32_t slot = 0;
        if (lane == 0) slot = atomicAdd(&a.counters[0], 1u);
        slot = __shfl(slot, 0);
        if (slot >= nslots) break;
        // polled with a relaxed load: an acquire on every poll invalidates the
        // L2 each time, and thousands of waiting workers made that 14x slower
        uint32_t v = kEmpty;
        bool gave_up = false;
        for (;;) {
            uint32_t got = kEmpty, stop = 0;
            if (lane == 0) {
                got = __hip_atomic_load(&a.queue[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (got == kEmpty) {
                    const uint32_t filled =
                        __hip_atomic_load(&a.counters[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    if (filled != seen) {
                        seen = filled;
                        spins = 0;
                    } else if (++spins > a.spin_limit) {
                        atomicOr(&a.counters[2], 1u);
                        stop = 1;
                    }
                }
            }
            v = __shfl(got, 0);
            if (v != kEmpty) break;
            if (__shfl(stop, 0)) {
                gave_up = true;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (gave_up) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // pairs with the queue entry's release store
        const uint32_t ib = a.in_off[v], ie = a.in_off[v + 1], ob = a.out_off[v], oe = a.out_off[v + 1];
        const bool run = !a.sched || a.sched[v];
        const uint32_t nin = ie - ib, nout = oe - ob;
        graphchi_edge<ET> *gin = a.in_edges + ib, *gout = a.out_edges + ob;
        const bool staged = run && nin + nout >= kWaveLanes / 4 && nin + nout <= stage_capacity<ET>();
        if (staged) {
            stage_in(stage, gin, nin, gout, nout, lane);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // the copies are in LDS before lane 0 reads them
            __builtin_amdgcn_wave_barrier();
        } else if (run && nin + nout >= kWaveLanes) {
            prefetch_views(gin, nin, lane);
            prefetch_views(gout, nout, lane);
        }
        if (lane == 0 && run) {
            graphchi_vertex<VT, ET> vx(v, staged ? stage : gin, staged ? stage + nin : gout, (int)nin, (int)nout);
            vx.dataptr = a.vdata + v;
            vx.twins_are_views = !staged;  // staged copies are re-linked by stage_out
            prog.P::update(vx, *ctx);
        }
        if (staged) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // lane 0's LDS writes before the others read
            __builtin_amdgcn_wave_barrier();
            stage_out(stage, gin, nin, gout, nout, lane);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        release_higher(a, v, a.in_src + ib, ie - ib, lane);
        release_higher(a, v, a.out_dst + ob, oe - ob, lane);
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

inline unsigned grid_for(uint64_t n, unsigned cap) {
    uint64_t b = (n + kThreads - 1) / kThreads;
    if (b == 0) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

// Runs niters iterations of program on the graph stored in base's shards and
// writes the vertex values to <base>.<sizeof VT>B.vout.
template <typename VT, typename ET, class ProgramT>
void run_program(ProgramT &program, const std::string &base, int nshards, uint32_t n, int niters, bool scheduling,
                 graphchi_context &ctx, metrics &m, size_t &nedges_out) {
    static_assert(std::is_trivially_copyable<VT>::value && std::is_trivially_copyable<ET>::value,
                  "device vertex programs need trivially copyable vertex and edge data");
    device::open();
    metrics_entry me = m.start_time();
    device::buffer<uint32_t> src, dst;
    std::vector<uint64_t> shard_edges;
    const uint64_t E = load_shard_edges_device(base, nshards, src, dst, &shard_edges);
    nedges_out = E;
    ctx.nedges = E;
    const uint64_t E1 = E ? E : 1;
    const uint32_t n1 = n ? n : 1;
    device::buffer<uint32_t> out_off((size_t)n + 1), in_off((size_t)n + 1), out_dst(E1), in_src(E1), in_eidx(E1),
        out_eidx(E1), lower(n1);
    {
        size_t wb = 0;
        device::check(gchi_psw_graph_workspace(E, n, &wb), "gchi_psw_graph_workspace");
        device::workspace w(wb);
        device::check(gchi_psw_graph_build_ex(src.get(), dst.get(), E, n, out_off.get(), out_dst.get(), in_off.get(),
                                              in_src.get(), in_eidx.get(), lower.get(), out_eidx.get(), w.get(), wb,
                                              nullptr),
                      "gchi_psw_graph_build");
        device::check(gchi_work_status(w.get(), nullptr), "gchi_psw_graph_build");
    }
    src.release();
    dst.release();
    // the edge values the shard files hold, in shard order
    // (memory_shard::load, src/shards/memoryshard.hpp:268-300)
    device::buffer<ET> val_in(E1);
    {
        std::vector<ET> h(E1);
        uint64_t at = 0;
        for (size_t p = 0; p < shard_edges.size(); p++) {
            const std::string ed = filename_shard_edata<ET>(base, (int)p, nshards);
            device::check(gchi_edata_read(ed.c_str(), kEdataZlib, (uint32_t)sizeof(ET), shard_edges[p], h.data() + at),
                          ("reading the edge values of shard " + std::to_string(p)).c_str());
            at += shard_edges[p];
        }
        val_in.upload(h.data(), E1);
    }
    device::buffer<VT> vdata(n1);
    device::buffer<graphchi_edge<ET>> out_edges(E1), in_edges(E1);
    device::buffer<uint32_t> dep(n1), queue(n1), counters(4);
    device::buffer<graphchi_context> d_ctx(1);
    device::buffer<double> d_delta(1);
    VT *vd = vdata.get();
    graphchi_edge<ET> *oe = out_edges.get(), *ie = in_edges.get();
    const uint32_t *od = out_dst.get(), *is = in_src.get(), *ix = in_eidx.get(), *lw = lower.get();
    const uint32_t *ox = out_eidx.get();
    const ET *vin = val_in.get();
    const uint32_t *ioff = in_off.get(), *ooff = out_off.get();
    uint32_t *dp = dep.get(), *qu = queue.get(), *ct = counters.get();
    graphchi_context *dctx = d_ctx.get();
    hipLaunchKernelGGL(fill_zero<VT>, dim3(grid_for(n, 4096)), dim3(kThreads), 0, 0, vd, (uint64_t)n);
    if (E)
        hipLaunchKernelGGL(edge_views<ET>, dim3(grid_for(E, 4096)), dim3(kThreads), 0, 0, od, is, ix, ox, vin,
                           (uint64_t)E, oe, ie);
    hip_check(hipGetLastError(), "psw setup");

    // The reference runs a graph of one shard (and fewer than 2 * 40M vertices)
    // in its in-memory mode (graphchi_engine.hpp:296-301, 465-523): all
    // iterations inside one outer iteration, so log_change() accumulates over
    // the whole run, the exec-interval hooks get (0, N), a last iteration ends
    // the run only when it is above 0, and <base>.deltalog gets one line at the
    // end.  Otherwise every iteration starts from zero change and logs a line.
    const bool inmem = nshards == 1 && n < 2u * 40000000u;

    // Passes: one over all vertices, or with scheduling on a graph of several
    // shards, one per interval in order, so that each reads its scheduled
    // flags after the intervals before it ran.  A pass over [vlo, vhi) waits
    // only for neighbours in it: cross[v] counts the others (done earlier).
    struct pass_range {
        uint32_t lo, hi;
    };
    std::vector<pass_range> passes;
    const bool per_interval = scheduling && !inmem;
    device::buffer<uint32_t> cross(per_interval ? n1 : 1);
    uint32_t *crs = per_interval ? cross.get() : nullptr;
    if (per_interval) {
        uint32_t lo = 0;
        for (vid_t end : load_interval_ends(base, nshards)) {
            const uint32_t hi = std::min<uint64_t>((uint64_t)end + 1, n);
            if (hi > lo) passes.push_back({lo, hi});
            lo = std::max(lo, hi);
        }
        if (lo < n) passes.push_back({lo, n});
        for (const pass_range &r : passes)
            hipLaunchKernelGGL(count_cross, dim3(grid_for(r.hi - r.lo, 4096)), dim3(kThreads), 0, 0, ioff, is, ooff, od,
                               r.lo, r.hi, crs);
    } else {
        passes.push_back({0, n});
    }

    // the scheduler (graphchi_engine.hpp:285-291: every vertex scheduled at the start)
    device::buffer<uint8_t> sched_a(scheduling ? n1 : 1), sched_b(scheduling ? n1 : 1), snap(scheduling ? n1 : 1);
    device::buffer<uint32_t> has_new(1);
    device::buffer<unsigned long long> stats(2);
    device::buffer<uint8_t> dsched_mem(sizeof(device_scheduler));
    device_scheduler *dsched = reinterpret_cast<device_scheduler *>(dsched_mem.get());
    uint32_t *hnew = has_new.get();
    uint8_t *snp = snap.get();
    unsigned long long *sts = stats.get();
    host_scheduler hsched;
    ischeduler no_scheduler;
    if (scheduling) {
        hsched.cur = sched_a.get();
        hsched.next = sched_b.get();
        hsched.has_new = has_new.get();
        hsched.n = n;
        const uint32_t one = 1u;
        hip_check(hipMemset(hsched.cur, 1, n), "scheduler init");
        hip_check(hipMemset(hsched.next, 0, n), "scheduler init");
        has_new.upload(&one, 1);
    }
    auto make_dsched = [&]() {
        uint8_t *c = scheduling ? hsched.cur : nullptr, *x = scheduling ? hsched.next : nullptr;
        hipLaunchKernelGGL(make_device_scheduler, dim3(1), dim3(1), 0, 0, dsched, c, x, hnew, n);
        hip_check(hipGetLastError(), "scheduler setup");
    };
    make_dsched();
    ctx.scheduler = scheduling ? static_cast<ischeduler *>(&hsched) : &no_scheduler;
    auto read_has_new = [&]() {
        hsched.stale();
        uint32_t f = 0;
        has_new.download(&f, 1);
        return f != 0;
    };
    auto clear_has_new = [&]() {
        hsched.stale();
        const uint32_t zero = 0;
        has_new.upload(&zero, 1);
    };
    // bitset_scheduler::new_iteration (bitset_scheduler.hpp:50-58)
    auto new_iteration = [&](int iter) {
        if (iter == 0) return;
        hsched.stale();
        std::swap(hsched.cur, hsched.next);
        uint8_t *x = hsched.next;
        hipLaunchKernelGGL(fill_zero<uint8_t>, dim3(grid_for(n, 4096)), dim3(kThreads), 0, 0, x, (uint64_t)n);
        make_dsched();
    };
    m.stop_time(me, "load_shards");

    int dev = 0, ncu = 256;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
        ncu = prop.multiProcessorCount;
    uint64_t updates = 0, work = 0;
    bool started = false;
    metrics_entry mu = m.start_time();
    if (inmem) ctx.reset_deltas(1);

    // one pass over [lo, hi) with the given scheduled flags (null: all); adds
    // its log_change() total to ctx
    auto run_pass = [&](const pass_range &r, const uint8_t *sched) {
        const uint32_t cnt_v = r.hi - r.lo;
        PassArgs<VT, ET> a{r.lo, r.hi, ooff, ioff, od, is, oe, ie, vd, dp, sched, qu, ct, 1u << 24};
        graphchi_context dc = ctx;  // the plain fields, for update() on the device
        dc.device_delta = d_delta.get();
        dc.scheduler = dsched;
        const double zero = 0.0;
        const unsigned int uzero = 0;
        d_ctx.upload(&dc, 1);
        d_delta.upload(&zero, 1);
        hip_check(hipMemcpyToSymbol(HIP_SYMBOL(device_assert_failed), &uzero, 4), "hipMemcpyToSymbol");
        hip_check(hipMemcpyToSymbol(HIP_SYMBOL(device_assert_line), &uzero, 4), "hipMemcpyToSymbol");
        const unsigned g = grid_for(cnt_v, 4096);
        hipLaunchKernelGGL(pass_reset, dim3(g), dim3(kThreads), 0, 0, lw, crs, r.lo, r.hi, dp, qu, ct);
        hipLaunchKernelGGL(pass_seed, dim3(g), dim3(kThreads), 0, 0, dp, r.lo, r.hi, qu, ct);
        // one 64 KB LDS area per wavefront: one workgroup per CU
        static bool lds_set = false;
        if (!lds_set) {
            hip_check(hipFuncSetAttribute((const void *)pass_run<ProgramT, VT, ET>,
                                          hipFuncAttributeMaxDynamicSharedMemorySize,
                                          (kPassThreads / kWaveLanes) * (int)kStageBytes),
                      "psw pass LDS");
            lds_set = true;
        }
        const unsigned pgrid = (unsigned)std::min<uint64_t>(((uint64_t)cnt_v + 1) / 2, (uint64_t)ncu);
        hipLaunchKernelGGL((pass_run<ProgramT, VT, ET>), dim3(pgrid ? pgrid : 1), dim3(kPassThreads),
                           (kPassThreads / kWaveLanes) * kStageBytes, 0, program, a, dctx);
        hip_check(hipGetLastError(), "psw pass launch");
        hip_check(hipDeviceSynchronize(), "psw pass");
        hsched.stale();
        const std::vector<uint32_t> cnt = counters.to_host(3);
        if (cnt[2]) logstream(LOG_FATAL) << "psw pass: a wait for a queued vertex gave up" << std::endl;
        unsigned int failed = 0, line = 0;
        hip_check(hipMemcpyFromSymbol(&failed, HIP_SYMBOL(device_assert_failed), 4), "hipMemcpyFromSymbol");
        hip_check(hipMemcpyFromSymbol(&line, HIP_SYMBOL(device_assert_line), 4), "hipMemcpyFromSymbol");
        if (failed)
            logstream(LOG_FATAL) << "assertion failed in update() on the device (line " << line << "), iteration "
                                 << ctx.iteration << std::endl;
        double delta = 0.0;
        d_delta.download(&delta, 1);
        ctx.log_change(delta);
    };
    // snapshot of this iteration's flags over [lo, hi); returns (vertices, edges)
    auto take_snapshot = [&](const pass_range &r, bool all) {
        hsched.stale();
        const unsigned long long z[2] = {0, 0};
        stats.upload(z, 2);
        const uint8_t *c = all ? nullptr : hsched.cur;
        hipLaunchKernelGGL(snapshot_schedule, dim3(grid_for(r.hi - r.lo, 4096)), dim3(kThreads), 0, 0, c, ioff, ooff,
                           r.lo, r.hi, snp, sts);
        hip_check(hipGetLastError(), "scheduler snapshot");
        std::vector<unsigned long long> st = stats.to_host(2);
        return std::make_pair((uint64_t)st[0], (uint64_t)st[1]);
    };

    for (int iter = 0; iter < niters; iter++) {
        started = true;
        ctx.iteration = iter;
        if (!inmem) ctx.reset_deltas(1);
        program.before_iteration(iter, ctx);
        if (inmem) {
            // graphchi_engine.hpp:465-511
            program.before_exec_interval(0, n, ctx);
            if (scheduling) {
                if (iter > 0 && !read_has_new()) break;
                new_iteration(iter);
                const auto st = take_snapshot(passes[0], iter == 0);
                if (st.first == 0) break;
                clear_has_new();
                updates += st.first;
                work += st.second;
                run_pass(passes[0], snp);
            } else {
                updates += n;
                work += E;
                if (n) run_pass(passes[0], nullptr);
            }
            program.after_exec_interval(0, n, ctx);
            program.after_iteration(iter, ctx);
            if (ctx.last_iteration > 0 && ctx.last_iteration <= iter) break;
            continue;
        }
        // graphchi_engine.hpp:798-960
        if (scheduling) {
            if (!read_has_new()) break;
            clear_has_new();
            new_iteration(iter);
            for (const pass_range &r : passes) {
                program.before_exec_interval(r.lo, r.hi - 1, ctx);
                const auto st = take_snapshot(r, false);
                updates += st.first;
                work += st.second;
                if (st.first) run_pass(r, snp);
                program.after_exec_interval(r.lo, r.hi - 1, ctx);
            }
        } else {
            const vid_t win_en = n ? n - 1 : 0;
            program.before_exec_interval(0, win_en, ctx);
            updates += n;
            work += 2 * E;  // both ends of every edge
            if (n) run_pass(passes[0], nullptr);
            program.after_exec_interval(0, win_en, ctx);
        }
        program.after_iteration(iter, ctx);
        write_delta_log(base, ctx, iter, updates, work);
        if (ctx.last_iteration >= 0 && iter >= ctx.last_iteration) break;
    }
    if (inmem && started) write_delta_log(base, ctx, ctx.iteration, updates, work);
    m.stop_time(mu, "execute-updates");
    m.set("updates", (double)updates);
    m.set("work", (double)work);
    hsched.stale();
    ctx.scheduler = nullptr;
    // With several shards the reference commits the edge values to the shard
    // files as it goes (memory_shard::commit, sliding_shard writes); in its
    // in-memory mode it does not (graphchi_engine.hpp:923-933).
    if (!inmem && E) {
        ET *vout = val_in.get();
        hipLaunchKernelGGL(edge_values_out<ET>, dim3(grid_for(E, 4096)), dim3(kThreads), 0, 0, oe, ox, (uint64_t)E,
                           vout);
        hip_check(hipGetLastError(), "edge values");
        const std::vector<ET> h = val_in.to_host(E);
        uint64_t at = 0;
        for (size_t p = 0; p < shard_edges.size(); p++) {
            const std::string ed = filename_shard_edata<ET>(base, (int)p, nshards);
            device::check(gchi_edata_write(ed.c_str(), kEdataZlib, (uint32_t)sizeof(ET), shard_edges[p], h.data() + at),
                          ("writing the edge values of shard " + std::to_string(p)).c_str());
            at += shard_edges[p];
        }
    }
    write_vertex_values<VT>(base, vdata.to_host(n));
}

}  // namespace psw
}  // namespace graphchi
#endif  // __HIPCC__
