// Triangle counting on the device.
//   trianglecounting file <edgelist> [nshards 2]
// The graph is ordered by degree (<file>.vertexmap, shards <file>_degord), the
// per-vertex counts are written to <file>_degord.4B.vout and the total printed.
#include <string>

#include "engine/dynamic_graphs/graphchi_dynamicgraph_engine.hpp"
#include "graphchi_basic_includes.hpp"
#include "preprocessing/util/orderbydegree.hpp"

using namespace graphchi;

// Counts triangles through the registered device form
// (graphchi::device::triangle_counting) on the degree-ordered graph.
struct TriangleCountingProgram : public GraphChiProgram<uint32_t, uint32_t> {
    void update(graphchi_vertex<uint32_t, uint32_t> &, graphchi_context &) {}
};

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    metrics m("triangle-counting");
    std::string file = get_option_string("file");
    int nshards = convert_if_notexists_novalues<uint32_t>(file, get_option_string("nshards", "2"));
    nshards = order_by_degree<uint32_t>(file, nshards, m);
    TriangleCountingProgram program;
    graphchi_dynamicgraph_engine<uint32_t, uint32_t> engine(file + "_degord", nshards, true, m);
    engine.run(program, 1);
    size_t sum3 = sum_vertices<vid_t, size_t>(file + "_degord", 0, (vid_t)engine.num_vertices());
    printf("Number of triangles: %zu(%zu)\n", sum3 / 3, sum3);
    metrics_report(m);
    return 0;
}
