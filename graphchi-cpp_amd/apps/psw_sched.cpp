// Selective scheduling as a vertex program: min-label propagation that only
// runs vertices a neighbour has scheduled.  Each update takes the minimum of
// its label and the labels on its edges (stored as label + 1, 0 meaning
// unset), rewrites the edges that differ and schedules the neighbour across
// each: for the next iteration, and also for this one when the two ids differ
// in parity (add_task(v, true)).  Vertices with id % 7 == 3 drop their own
// next-iteration task in odd iterations (remove_tasks).  The hooks use the
// scheduler on the host too: before iteration 2 they schedule vertex 0, after
// iteration 1 they drop the tasks of vertices 5-9, and after every iteration
// they print num_tasks().  A vertex keeps its label and a hash of the labels
// and iterations of every update it ran, so its value records which
// iterations ran it.  Written against the GraphChi program API only: the same
// source builds against this package (hipcc: update() runs on the device) and
// against the reference's engine (the oracle's recipe), whose vertex values,
// deltalog (updates and work per iteration) and printed task counts are the
// expected ones.
//
// usage: psw_sched file <edge list> [niters 40] [nshards auto]
#include <stdint.h>
#include <stdio.h>

#include <string>

#include "graphchi_basic_includes.hpp"

using namespace graphchi;

typedef uint64_t VertexDataType;  // trace << 32 | label
typedef uint32_t EdgeDataType;    // label + 1, 0: unset

static inline uint32_t sched_mix(uint32_t a, uint32_t b) {
    uint32_t h = (a ^ 0x85EBCA6Bu) * 0x9E3779B1u + b;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    return h;
}

struct PswSchedProgram : public GraphChiProgram<VertexDataType, EdgeDataType> {
    void update(graphchi_vertex<VertexDataType, EdgeDataType> &vertex, graphchi_context &gcontext) {
        const uint32_t id = vertex.id();
        const VertexDataType old = vertex.get_data();
        const uint32_t label = gcontext.iteration == 0 ? id : (uint32_t)old;
        uint32_t trace = gcontext.iteration == 0 ? 0u : (uint32_t)(old >> 32);
        uint32_t m = label;
        for (int i = 0; i < vertex.num_edges(); i++) {
            const uint32_t x = vertex.edge(i)->get_data();
            if (x != 0 && x - 1 < m) m = x - 1;
        }
        trace = sched_mix(trace, (uint32_t)gcontext.iteration * 0x10001u + m);
        for (int i = 0; i < vertex.num_edges(); i++) {
            graphchi_edge<EdgeDataType> *e = vertex.edge(i);
            if (e->get_data() != m + 1) {
                e->set_data(m + 1);
                const vid_t nb = e->vertex_id();
                if (nb != id) gcontext.scheduler->add_task(nb, ((nb ^ id) & 1u) != 0);
            }
        }
        if (id % 7 == 3 && (gcontext.iteration & 1)) gcontext.scheduler->remove_tasks(id, id);
        vertex.set_data(((VertexDataType)trace << 32) | m);
    }
    void before_iteration(int iteration, graphchi_context &gcontext) {
        if (iteration == 2 && gcontext.nvertices > 0) gcontext.scheduler->add_task(0);
    }
    void after_iteration(int iteration, graphchi_context &gcontext) {
        if (iteration == 1) gcontext.scheduler->remove_tasks(5, 9);
        printf("psw_sched: iteration %d tasks %lu\n", iteration, (unsigned long)gcontext.scheduler->num_tasks());
        fflush(stdout);
    }
    void before_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
    void after_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
};

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    metrics m("psw_sched");
    std::string filename = get_option_string("file");
    int niters = get_option_int("niters", 40);
    int nshards = convert_if_notexists<EdgeDataType>(filename, get_option_string("nshards", "auto"));
    PswSchedProgram program;
    graphchi_engine<VertexDataType, EdgeDataType> engine(filename, nshards, true, m);
    engine.run(program, niters);
    metrics_report(m);
    return 0;
}
