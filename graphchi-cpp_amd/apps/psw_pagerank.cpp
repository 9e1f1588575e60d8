// PageRank as a vertex program with values on the edges: each vertex sums its
// in-edge values, takes 0.15 + 0.85 * sum, and writes rank / out-degree to its
// out-edges; log_change() records the change of every rank, and the run stops
// once an iteration's total change is below `tol` (set_last_iteration).  The
// edge values it reads are those of the reference's update order, so the
// ranks depend on that order.  Written against the GraphChi program API only:
// the same source builds against this package (hipcc: update() runs on the
// device) and against the reference's engine (the oracle's recipe).
//
// usage: psw_pagerank file <edge list> [niters 30] [tol 1e-3] [nshards auto]
#include <cmath>
#include <string>

#include "graphchi_basic_includes.hpp"

using namespace graphchi;

typedef float VertexDataType;
typedef float EdgeDataType;

struct PswPagerankProgram : public GraphChiProgram<VertexDataType, EdgeDataType> {
    double tol = 1e-3;
    int iterations_run = 0;

    void update(graphchi_vertex<VertexDataType, EdgeDataType> &vertex, graphchi_context &gcontext) {
        if (gcontext.iteration == 0) {
            for (int i = 0; i < vertex.num_outedges(); i++) vertex.outedge(i)->set_data(1.0 / vertex.num_outedges());
            vertex.set_data(0.15f);
            return;
        }
        float sum = 0.0f;
        for (int i = 0; i < vertex.num_inedges(); i++) sum += vertex.inedge(i)->get_data();
        const float rank = 0.15 + 0.85 * sum;
        if (vertex.num_outedges() > 0) {
            const float share = rank / vertex.num_outedges();
            for (int i = 0; i < vertex.num_outedges(); i++) vertex.outedge(i)->set_data(share);
        }
        gcontext.log_change(std::fabs(rank - vertex.get_data()));
        vertex.set_data(rank);
    }
    void before_iteration(int iteration, graphchi_context &gcontext) {}
    void after_iteration(int iteration, graphchi_context &gcontext) {
        iterations_run = iteration + 1;
        if (iteration > 0 && gcontext.get_delta() < tol) gcontext.set_last_iteration(iteration);
    }
    void before_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
    void after_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
};

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    metrics m("psw_pagerank");
    std::string filename = get_option_string("file");
    int niters = get_option_int("niters", 30);
    int nshards = convert_if_notexists<EdgeDataType>(filename, get_option_string("nshards", "auto"));
    PswPagerankProgram program;
    program.tol = get_option_float("tol", 1e-3);
    graphchi_engine<VertexDataType, EdgeDataType> engine(filename, nshards, false, m);
    engine.run(program, niters);
    logstream(LOG_INFO) << "iterations run: " << program.iterations_run << std::endl;
    metrics_report(m);
    return 0;
}
