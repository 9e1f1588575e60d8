// PageRank through the functional API on the device.
//   pagerank file <edgelist> [niters 10] [form normalised|unnormalised] [top 20]
// form normalised:   r = 0.15/N + 0.85 * sum(r_u/outdeg_u)   (pagerank_functional)
// form unnormalised: r = 0.15 + 0.85 * sum(r_u/outdeg_u)     (pagerank.cpp)
#include <string>

#include "api/functional/functional_api.hpp"
#include "graphchi_basic_includes.hpp"
#include "util/toplist.hpp"

using namespace graphchi;

static bool g_normalised = true;

struct pr_kernel : public functional_kernel<float, float> {
    float initial_value(graphchi_context &, vertex_info &) { return 1.0f; }
    float reset() { return 0.0f; }
    float op_neighborval(graphchi_context &, vertex_info &, vid_t, float nbval) { return nbval; }
    float plus(float acc, float x) { return acc + x; }
    float compute_vertexvalue(graphchi_context &ctx, vertex_info &, float sum) {
        const double restart = 0.15;
        double base = g_normalised ? restart / ctx.nvertices : restart;
        return base + (1 - restart) * sum;
    }
    float value_to_neighbor(graphchi_context &, vertex_info &v, vid_t, float r) { return r / v.outdegree; }
};

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    metrics m("pagerank");
    std::string file = get_option_string("file");
    int niters = get_option_int("niters", 10);
    g_normalised = get_option_string("form", "normalised") != "unnormalised";
    run_functional_unweighted_synchronous<pr_kernel>(file, niters, m);
    metrics_report(m);
    int ntop = get_option_int("top", 20);
    std::vector<vertex_value<float>> top = get_top_vertices<float>(file, ntop);
    printf("Top %d vertices:\n", ntop);
    for (size_t i = 0; i < top.size(); i++) printf("%zu. %u\t%g\n", i + 1, top[i].vertex, top[i].value);
    return 0;
}
