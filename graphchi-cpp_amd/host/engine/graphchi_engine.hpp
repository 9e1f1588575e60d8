// graphchi_engine with the reference's constructor and run() interface
// (src/engine/graphchi_engine.hpp:148, 718-992, 1002-1075), executing on the
// device.
//
// The reference runs any GraphChiProgram through the parallel-sliding-windows
// loop on the CPU.  Here every program runs on the device:
//   * a program with a registered device algorithm runs that algorithm:
//         GCHI_DEVICE_PROGRAM(MyProgram, graphchi::device::min_label_components);
//     the programs of the reference's connectedcomponents and trianglecounting
//     apps (::ConnectedComponentsProgram, ::TriangleCountingProgram) are
//     registered here by name, so those sources compile and run unchanged;
//   * any other program, in an application compiled with hipcc, runs its own
//     update() on the device in the reference's order (engine/psw_engine.hpp);
//   * compiled with a host compiler, an unregistered program stops with a
//     fatal log: there is no CPU fallback.
//
// The graph is read from its shard files once (the adjacency decode of
// memory_shard), uploaded to HBM, and the whole computation runs there; the
// per-vertex results are written to <base>.<sizeof VT>B.vout like the
// reference's vertex-data store.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "api/chifilenames.hpp"
#include "api/graphchi_context.hpp"
#include "api/graphchi_program.hpp"
#include "api/vertex_aggregator.hpp"
#include "engine/device.hpp"
#include "logger/logger.hpp"
#include "metrics/metrics.hpp"
#include "engine/psw_engine.hpp"
#include "shards/shard_reader.hpp"
#include "util/cmdopts.hpp"

struct ConnectedComponentsProgram;  // example_apps/connectedcomponents.cpp:69
struct TriangleCountingProgram;     // example_apps/trianglecounting.cpp:282

namespace graphchi {
namespace device {
// Labels = smallest vertex id of each weakly connected component
// (the fixed point of connectedcomponents.cpp:79-138).
struct min_label_components {};
// Per-vertex incident triangles of a degree-ordered graph (trianglecounting.cpp).
struct triangle_counting {};
}  // namespace device

template <typename ProgramT>
struct device_program {};  // no registration: no ::impl

template <>
struct device_program<::ConnectedComponentsProgram> {
    typedef device::min_label_components impl;
};
template <>
struct device_program<::TriangleCountingProgram> {
    typedef device::triangle_counting impl;
};

template <typename P, typename = void>
struct has_device_program : std::false_type {};
template <typename P>
struct has_device_program<P, typename std::conditional<false, typename device_program<P>::impl, void>::type>
    : std::true_type {};

template <typename VertexDataType, typename EdgeDataType,
          typename svertex_t = graphchi_vertex<VertexDataType, EdgeDataType>>
class graphchi_engine {
  public:
    graphchi_engine(std::string base_filename, int nshards, bool selective_scheduling, metrics &m)
        : base_(base_filename), nshards_(nshards), scheduling_(selective_scheduling), m_(m) {
        nvertices_ = get_num_vertices(base_);
        ctx_.nvertices = nvertices_;
        m_.set("engine", "device");
    }
    virtual ~graphchi_engine() {}

    template <class ProgramT>
    void run(ProgramT &program, int niters) {
        ctx_.num_iterations = niters;
        metrics_entry me = m_.start_time();
        run_dispatch(program, has_device_program<ProgramT>());
        m_.stop_time(me, "runtime");
    }

    size_t num_vertices() const { return nvertices_; }
    size_t num_edges() const { return nedges_; }
    graphchi_context &get_context() { return ctx_; }
    void set_modifies_inedges(bool) {}
    void set_modifies_outedges(bool) {}
    void set_enable_deterministic_parallelism(bool) {}
    void set_membudget_mb(int) {}
    void set_exec_threads(int) {}
    void set_load_threads(int) {}
    void set_disable_vertexdata_storage() {}
    void set_only_adjacency(bool) {}
    void set_reset_vertexdata(bool) {}

  private:
    // The shards' edges, decoded on the device into a graph resident in HBM.
    device::graph load_graph() {
        metrics_entry me = m_.start_time();
        device::buffer<uint32_t> src, dst;
        nedges_ = load_shard_edges_device(base_, nshards_, src, dst);
        ctx_.nedges = nedges_;
        device::graph g(std::move(src), std::move(dst), nedges_, (uint32_t)nvertices_, true);
        m_.stop_time(me, "load_shards");
        return g;
    }

    template <class ProgramT>
    void run_dispatch(ProgramT &program, std::false_type) {
#if defined(__HIPCC__)
        ctx_.last_iteration = -1;
        psw::run_program<VertexDataType, EdgeDataType>(program, base_, nshards_, (uint32_t)nvertices_,
                                                       ctx_.num_iterations, scheduling_, ctx_, m_, nedges_);
#else
        (void)program;
        logstream(LOG_FATAL) << "no device implementation is registered for this program "
                             << "(GCHI_DEVICE_PROGRAM), and its update() runs on the device only in an "
                             << "application compiled with hipcc; this engine has no CPU path" << std::endl;
#endif
    }

    template <class ProgramT>
    void run_dispatch(ProgramT &program, std::true_type) {
        run_impl(program, typename device_program<ProgramT>::impl());
    }

    // The program's update() is replaced by the registered device algorithm;
    // the iteration hooks still run between rounds.  A program whose update()
    // sets a public `converged` flag false when it writes a label (as the
    // reference's ConnectedComponentsProgram does, connectedcomponents.cpp:106-
    // 116) gets that side effect from the round's result.
    template <class P>
    static auto set_converged(P &p, bool v, int) -> decltype(p.converged = v, void()) {
        p.converged = v;
    }
    template <class P>
    static void set_converged(P &, bool, long) {}

    template <class ProgramT>
    struct cc_rounds {
        graphchi_engine *eng;
        ProgramT *program;
        int niters;
        uint64_t updates = 0, work = 0;
        static int hook(void *user, uint32_t round, uint64_t dropped, uint64_t evaluated) {
            cc_rounds *c = static_cast<cc_rounds *>(user);
            graphchi_context &ctx = c->eng->ctx_;
            ctx.iteration = (int)round;
            if (round > 0 && dropped) set_converged(*c->program, false, 0);
            c->updates += evaluated;
            c->work += c->eng->nedges_;  // every interval's edges are loaded each iteration
            c->program->after_iteration((int)round, ctx);
            c->eng->write_delta_log((int)round, c->updates, c->work);
            const bool last = (ctx.last_iteration >= 0 && (int)round >= ctx.last_iteration) ||
                              (int)round + 1 >= c->niters || dropped == 0;
            if (last) return 1;
            ctx.iteration = (int)round + 1;
            c->program->before_iteration((int)round + 1, ctx);
            return 0;
        }
    };

    // connectedcomponents.cpp:79-121 on a three-vertex probe: iteration 0 takes
    // the minimum of the vertex's and its neighbours' ids and writes it to the
    // out-edges; later iterations take the minimum over the edge labels and
    // write it to the edges holding larger ones.  A program registered as
    // min-label components whose update() does not do that is refused.
    template <class ProgramT>
    void probe_min_label(ProgramT &program) {
        ischeduler probe_sched;
        graphchi_context pc;
        pc.scheduler = &probe_sched;
        pc.nvertices = 10;
        uint32_t in_val = 7, out_val = 8, label = 0;
        ProgramT p = program;
        // a fresh vertex per call: edge views copy their values when built
        auto update_at = [&](int iteration) {
            graphchi_edge<EdgeDataType> ein[1], eout[1];
            graphchi_vertex<VertexDataType, EdgeDataType> v(5, ein, eout, 0, 0);
            v.dataptr = reinterpret_cast<VertexDataType *>(&label);
            v.add_inedge(3, reinterpret_cast<EdgeDataType *>(&in_val), false);
            v.add_outedge(9, reinterpret_cast<EdgeDataType *>(&out_val), false);
            pc.iteration = iteration;
            p.update(v, pc);
        };
        update_at(0);
        const bool ok0 = label == 3 && out_val == 3;
        in_val = 1;
        out_val = 4;
        update_at(1);
        const bool ok1 = label == 1 && out_val == 1 && in_val == 1;
        if (!ok0 || !ok1)
            logstream(LOG_FATAL) << "the program registered as min-label connected components does not compute "
                                 << "min-label propagation in its update(); refusing to run the device algorithm"
                                 << std::endl;
    }

    template <class ProgramT>
    void run_impl(ProgramT &program, device::min_label_components) {
        static_assert(sizeof(VertexDataType) == 4, "min-label components writes 4-byte labels");
        probe_min_label(program);
        device::graph g = load_graph();
        const uint32_t n = (uint32_t)nvertices_;
        // the undirected adjacency (every edge in both rows) as a CSR: the
        // update() reads in- and out-edges alike
        const uint64_t e2 = 2 * g.e;
        device::buffer<uint32_t> s2(e2 ? e2 : 1), d2(e2 ? e2 : 1);
        if (g.e) {
            device::check(gchi_memcpy(s2.get(), g.src.get(), 4 * g.e, nullptr), "copy");
            device::check(gchi_memcpy(s2.get() + g.e, g.dst.get(), 4 * g.e, nullptr), "copy");
            device::check(gchi_memcpy(d2.get(), g.dst.get(), 4 * g.e, nullptr), "copy");
            device::check(gchi_memcpy(d2.get() + g.e, g.src.get(), 4 * g.e, nullptr), "copy");
        }
        device::graph sym(std::move(s2), std::move(d2), e2, n, false);
        ctx_.iteration = 0;
        ctx_.last_iteration = -1;
        program.before_iteration(0, ctx_);
        metrics_entry me = m_.start_time();
        device::buffer<uint32_t> labels(nvertices_);
        size_t ws = 0;
        device::check(gchi_cc_scheduled_workspace(n, e2, &ws), "cc workspace");
        device::workspace w(ws);
        cc_rounds<ProgramT> hooks{this, &program, ctx_.num_iterations > 0 ? ctx_.num_iterations : 1};
        uint32_t rounds = 0;
        uint64_t evaluated = 0;
        device::check(gchi_cc_scheduled_ex(sym.in_off.get(), sym.in_src.get(), n, 0, labels.get(), &rounds, &evaluated,
                                           &cc_rounds<ProgramT>::hook, &hooks, w.get(), ws, nullptr),
                      "gchi_cc_scheduled");
        m_.stop_time(me, "execute-updates");
        m_.set("iterations", (double)rounds);
        m_.set("updates", (double)evaluated);
        std::vector<uint32_t> h = labels.to_host(nvertices_);
        std::vector<VertexDataType> out(h.begin(), h.end());
        write_vertex_values<VertexDataType>(base_, out);
    }

    // <base>.deltalog, one line per iteration: iteration, updates, work, delta
    // (src/engine/graphchi_engine.hpp:593-599).
    void write_delta_log(int iteration, uint64_t updates, uint64_t work) {
        FILE *df = fopen((base_ + ".deltalog").c_str(), iteration == 0 ? "w" : "a");
        if (!df) return;
        fprintf(df, "%d,%lu,%lu,%lf\n", iteration, (unsigned long)updates, (unsigned long)work, ctx_.get_delta());
        fclose(df);
    }

    template <class ProgramT>
    void run_impl(ProgramT &program, device::triangle_counting) {
        static_assert(sizeof(VertexDataType) == 4, "triangle counting writes uint32 counts");
        device::graph g = load_graph();
        const uint32_t n = (uint32_t)nvertices_;
        ctx_.iteration = 0;
        program.before_iteration(0, ctx_);
        metrics_entry me = m_.start_time();
        device::buffer<uint32_t> vmap(n), off((size_t)n + 1), adj(g.e ? g.e : 1), pv(n), ovf(1);
        device::buffer<uint64_t> total(1);
        size_t ws = 0;
        device::check(gchi_tc_prepare_workspace(g.e, n, &ws), "tc workspace");
        uint64_t m = 0;
        {
            device::workspace w(ws);
            device::check(gchi_tc_prepare(g.src.get(), g.dst.get(), g.e, n, vmap.get(), off.get(), adj.get(), &m,
                                          w.get(), ws, nullptr),
                          "gchi_tc_prepare");
        }
        device::check(gchi_tc_count_workspace(n, &ws), "tc workspace");
        device::workspace w2(ws);
        device::check(gchi_tc_count(off.get(), adj.get(), n, pv.get(), total.get(), ovf.get(), w2.get(), ws, nullptr),
                      "gchi_tc_count");
        device::check(gchi_synchronize(nullptr), "sync");
        m_.stop_time(me, "execute-updates");
        if (ovf.to_host(1)[0]) logstream(LOG_ERROR) << "a per-vertex triangle count exceeds 32 bits" << std::endl;
        // counts are indexed by the (degree, id) rank of this graph's vertices
        std::vector<uint32_t> h_pv = pv.to_host(n), h_vm = vmap.to_host(n);
        std::vector<VertexDataType> out(n);
        for (uint32_t v = 0; v < n; v++) out[v] = (VertexDataType)h_pv[h_vm[v]];
        ctx_.set_last_iteration(0);
        program.after_iteration(0, ctx_);
        write_vertex_values<VertexDataType>(base_, out);
    }

    std::string base_;
    int nshards_;
    bool scheduling_;  // selective scheduling (the generic device engine's bitset scheduler)
    metrics &m_;
    graphchi_context ctx_;
    size_t nvertices_ = 0, nedges_ = 0;
};

// The reference's dynamic-graph engine (trianglecounting's host engine) has the
// same interface for the programs run here.
template <typename VertexDataType, typename EdgeDataType,
          typename svertex_t = graphchi_vertex<VertexDataType, EdgeDataType>>
class graphchi_dynamicgraph_engine : public graphchi_engine<VertexDataType, EdgeDataType, svertex_t> {
  public:
    using graphchi_engine<VertexDataType, EdgeDataType, svertex_t>::graphchi_engine;
};

}  // namespace graphchi

#define GCHI_DEVICE_PROGRAM(ProgramT, Impl)       \
    template <>                                    \
    struct graphchi::device_program<ProgramT> {    \
        typedef Impl impl;                         \
    };
