// A vertex program whose result depends on which iteration's value each
// update sees on each edge: a vertex folds its own value and order-free
// summaries (sum, xor of hashes) of its in-edge and out-edge values into a
// 32-bit hash, writes a hash of it and the neighbour's id to its out-edges and
// to the in-edges from higher ids, and keeps it as its value.  An update that
// saw an edge value from the wrong iteration changes every later hash it
// reaches.  The summaries do not depend on the order of a vertex's edges: the
// reference's engine appends out-edges from several shards loaded in parallel,
// so that order is not fixed when a graph has more than one shard.  Written
// against the GraphChi program API only, so the same source builds against this
// package (hipcc: update() runs on the device) and against the reference's
// engine (the oracle's recipe), whose vertex values are the expected ones.
//
// usage: psw_probe file <edge list> [niters 4] [nshards auto]
#include <stdint.h>

#include <string>

#include "graphchi_basic_includes.hpp"

using namespace graphchi;

typedef uint32_t VertexDataType;
typedef uint32_t EdgeDataType;

static inline uint32_t probe_mix(uint32_t a, uint32_t b) {
    uint32_t h = (a * 0x9E3779B1u) ^ (b + 0x7F4A7C15u + (a << 6) + (a >> 2));
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}

struct PswProbeProgram : public GraphChiProgram<VertexDataType, EdgeDataType> {
    void update(graphchi_vertex<VertexDataType, EdgeDataType> &vertex, graphchi_context &gcontext) {
        const uint32_t id = vertex.id();
        if (gcontext.iteration == 0) {
            for (int i = 0; i < vertex.num_outedges(); i++)
                vertex.outedge(i)->set_data(probe_mix(id, vertex.outedge(i)->vertex_id()));
            vertex.set_data(id);
            return;
        }
        uint32_t s_in = 0, x_in = 0, s_out = 0, x_out = 0;
        for (int i = 0; i < vertex.num_inedges(); i++) {
            const uint32_t v = vertex.inedge(i)->get_data();
            s_in += v;
            x_in ^= probe_mix(v, 1u);
        }
        for (int i = 0; i < vertex.num_outedges(); i++) {
            const uint32_t v = vertex.outedge(i)->get_data();
            s_out += v;
            x_out ^= probe_mix(v, 2u);
        }
        const uint32_t acc = probe_mix(probe_mix(vertex.get_data(), s_in ^ x_out), x_in + s_out);
        for (int i = 0; i < vertex.num_outedges(); i++)
            vertex.outedge(i)->set_data(probe_mix(acc, vertex.outedge(i)->vertex_id()));
        for (int i = 0; i < vertex.num_inedges(); i++)
            if (vertex.inedge(i)->vertex_id() > id)
                vertex.inedge(i)->set_data(probe_mix(acc ^ 0x5bd1e995u, vertex.inedge(i)->vertex_id()));
        vertex.set_data(acc);
    }
    void before_iteration(int iteration, graphchi_context &gcontext) {}
    void after_iteration(int iteration, graphchi_context &gcontext) {}
    void before_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
    void after_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
};

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    metrics m("psw_probe");
    std::string filename = get_option_string("file");
    int niters = get_option_int("niters", 4);
    int nshards = convert_if_notexists<EdgeDataType>(filename, get_option_string("nshards", "auto"));
    PswProbeProgram program;
    graphchi_engine<VertexDataType, EdgeDataType> engine(filename, nshards, false, m);
    engine.run(program, niters);
    metrics_report(m);
    return 0;
}
