// graphchi_engine with the reference's constructor and run() interface
// (src/engine/graphchi_engine.hpp:148, 718-992, 1002-1075), executing on the
// device.
//
// The reference runs any GraphChiProgram through the parallel-sliding-windows
// loop on the CPU.  Here a program runs on the device when a device
// implementation is registered for its type:
//     GCHI_DEVICE_PROGRAM(MyProgram, graphchi::device::min_label_components);
// The programs of the reference's connectedcomponents and trianglecounting apps
// (::ConnectedComponentsProgram, ::TriangleCountingProgram) are registered
// here by name, so those sources compile and run unchanged.  A program without
// a registration stops with a fatal log: there is no CPU fallback.
//
// The graph is read from its shard files once (the adjacency decode of
// memory_shard), uploaded to HBM, and the whole computation runs there; the
// per-vertex results are written to <base>.<sizeof VT>B.vout like the
// reference's vertex-data store.
#pragma once
#include <stdint.h>

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
        : base_(base_filename), nshards_(nshards), m_(m) {
        (void)selective_scheduling;
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
    void run_dispatch(ProgramT &, std::false_type) {
        logstream(LOG_FATAL) << "no device implementation is registered for this program "
                             << "(GCHI_DEVICE_PROGRAM); this engine has no CPU path" << std::endl;
    }

    template <class ProgramT>
    void run_dispatch(ProgramT &program, std::true_type) {
        run_impl(program, typename device_program<ProgramT>::impl());
    }

    template <class ProgramT>
    void run_impl(ProgramT &program, device::min_label_components) {
        static_assert(sizeof(VertexDataType) == 4, "min-label components writes 4-byte labels");
        device::graph g = load_graph();
        ctx_.iteration = 0;
        program.before_iteration(0, ctx_);
        metrics_entry me = m_.start_time();
        device::buffer<uint32_t> labels(nvertices_);
        size_t ws = 0;
        device::check(gchi_cc_workspace((uint32_t)nvertices_, &ws), "cc workspace");
        device::workspace w(ws);
        device::check(gchi_cc_minlabel(g.src.get(), g.dst.get(), g.e, (uint32_t)nvertices_, labels.get(), w.get(), ws,
                                       nullptr),
                      "gchi_cc_minlabel");
        device::check(gchi_synchronize(nullptr), "sync");
        m_.stop_time(me, "execute-updates");
        program.after_iteration(0, ctx_);
        std::vector<uint32_t> h = labels.to_host(nvertices_);
        std::vector<VertexDataType> out(h.begin(), h.end());
        write_vertex_values<VertexDataType>(base_, out);
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
