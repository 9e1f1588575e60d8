// The functional (gather-apply-scatter) API of src/api/functional/functional_api.hpp:54-137
// with the same kernel interface and runners, executed on the device.
//
// A kernel's operator is identified by evaluating its methods on the host:
// reset() must be the identity of plus(), plus() must be FP32 addition,
// op_neighborval() the identity, value_to_neighbor() r / outdeg in FP32,
// initial_value() a constant, and compute_vertexvalue(s) an affine map
// c_v + d*s whose constant c_v is the same for every vertex, or 0 except on a
// seed set (personalized form).  The coefficients are recovered from the
// kernel's own outputs (snapped to the short decimal that reproduces them)
// and every probe is re-checked against the device formula
// (float)(c_v + d*(double)s).  Such kernels -- PageRank in all the reference's
// forms -- run as the device's synchronous pull sweep; any other kernel stops
// with a fatal log (there is no CPU path).
//
// Iteration semantics: iteration 0 sets initial values and scatters,
// iterations 1..niters-1 gather, apply and scatter, so niters iterations are
// niters-1 sweeps.  The synchronous runner (functional_bulksync.hpp:84-113)
// replays its sweeps from one HIP graph; the semi-synchronous runner
// (functional_semisync.hpp:76-110) sweeps the engine's sub-intervals in order,
// each seeing the values earlier ones wrote in the same sweep.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "api/chifilenames.hpp"
#include "api/graphchi_context.hpp"
#include "api/vertex_aggregator.hpp"
#include "engine/device.hpp"
#include "graphchi_types.hpp"
#include "logger/logger.hpp"
#include "metrics/metrics.hpp"
#include "preprocessing/conversions.hpp"
#include "shards/shard_reader.hpp"
#include "util/cmdopts.hpp"

namespace graphchi {

struct vertex_info {
    vid_t vertexid;
    int indegree;
    int outdegree;
};

template <typename FVertexDataType, typename FEdgeDataType>
struct functional_kernel {
    typedef FVertexDataType VertexDataType;
    typedef FEdgeDataType EdgeDataType;
    functional_kernel() {}
    virtual ~functional_kernel() {}
    virtual VertexDataType initial_value(graphchi_context &info, vertex_info &myvertex) = 0;
    virtual VertexDataType reset() = 0;
    virtual EdgeDataType op_neighborval(graphchi_context &info, vertex_info &myvertex, vid_t nbid,
                                        EdgeDataType nbval) = 0;
    virtual EdgeDataType plus(VertexDataType curval, EdgeDataType toadd) = 0;
    virtual VertexDataType compute_vertexvalue(graphchi_context &ginfo, vertex_info &myvertex,
                                               EdgeDataType nbvalsum) = 0;
    virtual EdgeDataType value_to_neighbor(graphchi_context &info, vertex_info &myvertex, vid_t nbid,
                                           VertexDataType myval) = 0;
};

namespace detail {

struct affine_form {
    float init = 0.0f;
    double teleport = 0.0;       // c_v on non-seed vertices
    double seed_teleport = 0.0;  // extra c_v on seed vertices
    double damping = 0.0;
    std::vector<vid_t> seeds;
};

inline double snap_decimal(double x, double scale) {
    double r = std::round(x * scale) / scale;
    return std::fabs(r - x) <= 1e-6 * std::fabs(x) + 1e-12 ? r : x;
}

inline float apply_form(double c, double d, float s) {
    volatile double t = d * (double)s;
    return (float)(c + t);
}

template <class KERNEL>
affine_form identify(graphchi_context &ctx, const std::vector<uint32_t> &indeg, const std::vector<uint32_t> &outdeg) {
    typedef typename KERNEL::VertexDataType VT;
    typedef typename KERNEL::EdgeDataType ET;
    static_assert(std::is_same<VT, float>::value && std::is_same<ET, float>::value,
                  "device functional runner supports float vertex and edge values");
    KERNEL k;
    const size_t n = indeg.size();
    auto info = [&](vid_t v) {
        vertex_info vi;
        vi.vertexid = v;
        vi.indegree = (int)indeg[v];
        vi.outdegree = (int)outdeg[v];
        return vi;
    };
    auto fail = [](const char *what) {
        logstream(LOG_FATAL) << "functional kernel is not a PageRank-form operator (" << what
                             << "); the device engine runs only such kernels" << std::endl;
    };
    const float xs[] = {0.0f, 1.0f, 0.5f, 3.25f, 1e-3f, 123.5f, 7.0e-6f};
    std::vector<vid_t> probe;
    for (vid_t v = 0; v < n && probe.size() < 16; v += (vid_t)std::max<size_t>(1, n / 16)) probe.push_back(v);
    probe.push_back((vid_t)(n - 1));
    const float zero = k.reset();
    for (float x : xs) {
        if (k.plus(zero, x) != x) fail("reset() is not the identity of plus()");
        for (float y : xs)
            if (k.plus(x, y) != x + y) fail("plus() is not addition");
    }
    for (vid_t v : probe) {
        vertex_info vi = info(v);
        for (float x : xs) {
            if (k.op_neighborval(ctx, vi, 0, x) != x) fail("op_neighborval() is not the identity");
            if (outdeg[v] && k.value_to_neighbor(ctx, vi, 0, x) != x / (float)outdeg[v])
                fail("value_to_neighbor() is not r/outdeg");
        }
    }
    affine_form f;
    {
        vertex_info v0 = info(0);
        f.init = k.initial_value(ctx, v0);
        for (vid_t v = 0; v < n; v++) {
            vertex_info vi = info(v);
            if (k.initial_value(ctx, vi) != f.init) fail("initial_value() depends on the vertex");
        }
    }
    // constant terms c_v = compute_vertexvalue(v, 0): one value, or 0 and one value
    std::vector<float> c(n);
    float cmax = 0.0f, cmin = 0.0f;
    for (vid_t v = 0; v < n; v++) {
        vertex_info vi = info(v);
        c[v] = k.compute_vertexvalue(ctx, vi, 0.0f);
        if (v == 0 || c[v] > cmax) cmax = c[v];
        if (v == 0 || c[v] < cmin) cmin = c[v];
    }
    bool seeded = false;
    if (cmin != cmax) {
        for (vid_t v = 0; v < n; v++)
            if (c[v] != 0.0f && c[v] != cmax) fail("compute_vertexvalue() constant term takes more than two values");
        if (cmin != 0.0f) fail("compute_vertexvalue() constant term is not 0 outside the seed set");
        seeded = true;
        for (vid_t v = 0; v < n; v++)
            if (c[v] == cmax) f.seeds.push_back(v);
    }
    const vid_t vc = seeded ? f.seeds[0] : 0;
    vertex_info vci = info(vc);
    const double big = 1048576.0;
    double d = ((double)k.compute_vertexvalue(ctx, vci, (float)big) - (double)cmax) / big;
    // snap d to the decimal (or 1 - decimal) that reproduces it, and c to c' / N or c'
    double cands_d[] = {1.0 - snap_decimal(1.0 - d, 1e6), snap_decimal(d, 1e6), d};
    double cands_c[] = {snap_decimal((double)cmax * (double)n, 1e6) / (double)n, snap_decimal((double)cmax, 1e6),
                        (double)cmax};
    bool found = false;
    for (double dd : cands_d) {
        for (double cc : cands_c) {
            if ((float)cc != cmax) continue;
            bool ok = true;
            for (vid_t v : probe) {
                vertex_info vi = info(seeded ? vc : v);
                for (float s : xs) {
                    float want = k.compute_vertexvalue(ctx, vi, s);
                    if (apply_form(cc, dd, s) != want) ok = false;
                }
            }
            if (ok) {
                f.damping = dd;
                if (seeded) f.seed_teleport = cc;
                else f.teleport = cc;
                found = true;
                break;
            }
        }
        if (found) break;
    }
    if (!found) fail("compute_vertexvalue() is not c + d*s evaluated in double");
    return f;
}

inline void write_seed_bits(const std::vector<vid_t> &seeds, size_t n, device::buffer<uint32_t> &bits) {
    std::vector<uint32_t> w((n + 31) / 32, 0u);
    for (vid_t v : seeds) w[v >> 5] |= 1u << (v & 31);
    bits.resize(w.size());
    bits.upload(w.data(), w.size());
}

// The engine's sub-intervals for a functional kernel: each interval is cut into
// windows of at most maxwindow + 1 vertices, maxwindow = membudget_mb * 2^20 /
// 300 (src/engine/graphchi_engine.hpp:729-732, 862-869; determine_next_window
// returns the window's upper bound for computational edges, :311-313).
// Returns the block starts followed by n.
inline std::vector<uint32_t> semisync_blocks(size_t n, const std::vector<vid_t> &interval_ends, int membudget_mb) {
    const uint64_t maxwindow = (uint64_t)membudget_mb * 1024 * 1024 / 3 / 100;
    std::vector<uint32_t> b;
    uint64_t st = 0;
    for (vid_t en : interval_ends) {
        for (uint64_t sub = st; sub <= en;) {
            b.push_back((uint32_t)sub);
            sub = std::min<uint64_t>(sub + maxwindow, en) + 1;
        }
        st = (uint64_t)en + 1;
    }
    b.push_back((uint32_t)n);
    return b;
}

// Loads the graph, identifies the kernel and runs niters iterations on the
// device: synchronous sweeps replayed from one HIP graph, or semi-synchronous
// block sweeps.
template <class KERNEL>
void run_functional_device(std::string filename, int niters, metrics &m, bool semisync) {
    metrics_entry total = m.start_time();
    int nshards = convert_if_notexists<PairContainer<typename KERNEL::EdgeDataType>>(
        filename, get_option_string("nshards", "auto"));
    const size_t n = get_num_vertices(filename);
    metrics_entry ml = m.start_time();
    device::buffer<uint32_t> src, dst;
    const uint64_t ne = load_shard_edges_device(filename, nshards, src, dst);
    device::graph g(std::move(src), std::move(dst), ne, (uint32_t)n, false);
    m.stop_time(ml, "load_graph");
    std::vector<uint32_t> indeg = g.indeg.to_host(n), outdeg = g.outdeg.to_host(n);
    graphchi_context ctx;
    ctx.nvertices = n;
    ctx.nedges = g.e;
    ctx.num_iterations = niters;
    affine_form f = identify<KERNEL>(ctx, indeg, outdeg);
    logstream(LOG_INFO) << "device operator: r = " << f.teleport << (f.seeds.empty() ? "" : " (+seed term)") << " + "
                        << f.damping << " * sum(r/outdeg), init " << f.init << (semisync ? ", semi-synchronous" : "")
                        << std::endl;
    // launch shape: the rank-space tiled sweep with the largest LDS cache (the
    // fastest shape on the benchmark graph), or the autotuner's pick (autotune
    // 1); the semi-synchronous block sweeps take every shape but that one
    gchi_pr_tune tune{512, 0, 4, 1 << 20};
    if (semisync) tune = gchi_pr_tune{0, 0, 0, 0};
    if (get_option_int("autotune", 0)) {
        double ms = 0;
        device::check(gchi_autotune_pr(&g.dg, 3, &tune, &ms, nullptr, nullptr, 0, nullptr, nullptr), "autotune");
        if (semisync && tune.persistent == 4) tune = gchi_pr_tune{0, 0, 0, 0};
    }
    device::buffer<uint32_t> seed_bits;
    gchi_pr_op op{f.teleport, f.damping, nullptr, nullptr, 0.0};
    if (!f.seeds.empty()) {
        write_seed_bits(f.seeds, n, seed_bits);
        op.d_seed_bits = seed_bits.get();
        op.seed_teleport = f.seed_teleport;
    }
    // the personalized form starts from the restart vector only when asked to
    const bool seed_start = !f.seeds.empty() && f.init == 0.0f;
    const int sweeps = niters > 1 ? niters - 1 : 0;
    device::buffer<float> x0(n), x1(n), r0(n), r1(n);
    device::buffer<double> stats(2 * (size_t)(sweeps ? sweeps : 1));
    metrics_entry mu = m.start_time();
    if (semisync) {
        const std::vector<uint32_t> blocks =
            semisync_blocks(n, load_interval_ends(filename, nshards), get_option_int("membudget_mb", 1024));
        gchi_pr_blocked *bl = nullptr;
        device::check(gchi_pr_blocked_create(&g.dg, blocks.data(), (uint32_t)blocks.size() - 1, &tune, &bl, nullptr),
                      "gchi_pr_blocked_create");
        m.set("semisync_blocks", (double)(blocks.size() - 1));
        device::check(gchi_pr_blocked_init(bl, f.init, seed_start ? seed_bits.get() : nullptr, (float)f.seed_teleport,
                                           x0.get(), r0.get(), nullptr),
                      "gchi_pr_blocked_init");
        float *rs[2] = {r0.get(), r1.get()};
        for (int k = 0; k < sweeps; k++)
            device::check(gchi_pr_blocked_sweep(bl, &op, x0.get(), rs[k & 1], rs[(k + 1) & 1], stats.get() + 2 * k,
                                                nullptr),
                          "gchi_pr_blocked_sweep");
        device::check(gchi_synchronize(nullptr), "sync");
        gchi_pr_blocked_destroy(bl);
    } else {
        gchi_pr_plan *plan = nullptr;
        device::check(gchi_pr_plan_create(&g.dg, &tune, &plan, nullptr), "gchi_pr_plan_create");
        if (seed_start)
            device::check(gchi_pr_init_seeds(plan, seed_bits.get(), (float)f.seed_teleport, x0.get(), r0.get(),
                                             nullptr),
                          "gchi_pr_init_seeds");
        // iterations 1..niters-1 captured once into a HIP graph and replayed
        gchi_pr_exec *ex = nullptr;
        device::check(gchi_pr_exec_create_ex(plan, &op, f.init, seed_start ? 0 : 1, sweeps, x0.get(), x1.get(),
                                             r0.get(), r1.get(), stats.get(), &ex),
                      "gchi_pr_exec_create");
        device::check(gchi_pr_exec_launch(ex, nullptr), "gchi_pr_exec_launch");
        device::check(gchi_synchronize(nullptr), "sync");
        gchi_pr_exec_destroy(ex);
        gchi_pr_plan_destroy(plan);
    }
    m.stop_time(mu, "execute-updates");
    std::vector<float> r = (sweeps & 1 ? r1 : r0).to_host(n);
    if (sweeps) {
        std::vector<double> st = stats.to_host(2 * (size_t)sweeps);
        m.set("last_residual", st[2 * (sweeps - 1)]);
        m.set("rank_sum", st[2 * (sweeps - 1) + 1]);
    }
    write_vertex_values<float>(filename, r);
    m.set("iterations", (double)niters);
    m.stop_time(total, "runtime");
}

}  // namespace detail

template <class KERNEL>
void run_functional_unweighted_synchronous(std::string filename, int niters, metrics &m) {
    detail::run_functional_device<KERNEL>(filename, niters, m, false);
}

template <class KERNEL>
void run_functional_unweighted_semisynchronous(std::string filename, int niters, metrics &m) {
    detail::run_functional_device<KERNEL>(filename, niters, m, true);
}

}  // namespace graphchi
