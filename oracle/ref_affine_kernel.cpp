// Test harness (oracle side only): runs the reference's own synchronous
// functional engine (run_functional_unweighted_synchronous,
// src/api/functional/functional_api.hpp:123-137) with an affine kernel that is
// not shipped as an app:
//   variant 0: r = 0.15 + 0.85*s            (the pagerank.cpp formula)
//   variant 1: r = (v in S ? 0.15/|S| : 0) + 0.85*s   (personalized; S read
//              from the text file named by the "seeds" option, one id a line)
// Compiled against the reference headers by oracle/Makefile; results land in
// <file>.4B.vout like every functional app.
#define GRAPHCHI_DISABLE_COMPRESSION
#include <fstream>
#include <set>
#include <string>

#include "graphchi_basic_includes.hpp"
#include "api/functional/functional_api.hpp"

using namespace graphchi;

static int g_variant = 0;
static std::set<vid_t> g_seeds;

struct affine_kernel : public functional_kernel<float, float> {
    float initial_value(graphchi_context &info, vertex_info &myvertex) { return 1.0; }
    float reset() { return 0.0; }
    float op_neighborval(graphchi_context &info, vertex_info &myvertex, vid_t nbid, float nbval) { return nbval; }
    float plus(float curval, float toadd) { return curval + toadd; }
    float compute_vertexvalue(graphchi_context &ginfo, vertex_info &myvertex, float nbvalsum) {
        double restart = 0.15;
        if (g_variant == 1)
            restart = g_seeds.count(myvertex.vertexid) ? 0.15 / (double)g_seeds.size() : 0.0;
        return restart + (1 - 0.15) * nbvalsum;
    }
    float value_to_neighbor(graphchi_context &info, vertex_info &myvertex, vid_t nbid, float myval) {
        return myval / myvertex.outdegree;
    }
};

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    metrics m("affine-kernel");
    g_variant = get_option_int("variant", 0);
    if (g_variant == 1) {
        std::ifstream in(get_option_string("seeds").c_str());
        vid_t v;
        while (in >> v) g_seeds.insert(v);
        assert(!g_seeds.empty());
    }
    run_functional_unweighted_synchronous<affine_kernel>(get_option_string("file"), get_option_int("niters", 4), m);
    metrics_report(m);
    return 0;
}
