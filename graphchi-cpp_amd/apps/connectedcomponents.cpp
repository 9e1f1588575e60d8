// Weakly connected components by minimum label on the device.
//   connectedcomponents file <edgelist> [nshards auto]
// Writes <file>.4B.vout (label = smallest vertex id of the component) and
// <file>.components.
#include <algorithm>
#include <string>

#include "graphchi_basic_includes.hpp"
#include "util/labelanalysis.hpp"

using namespace graphchi;

// Min-label update: the engine runs its registered device form
// (graphchi::device::min_label_components); update() states the rule.
struct ConnectedComponentsProgram : public GraphChiProgram<vid_t, vid_t> {
    void update(graphchi_vertex<vid_t, vid_t> &v, graphchi_context &ctx) {
        vid_t label = ctx.iteration == 0 ? v.id() : v.get_data();
        for (int i = 0; i < v.num_edges(); i++)
            label = std::min(label, ctx.iteration == 0 ? v.edge(i)->vertex_id() : v.edge(i)->get_data());
        v.set_data(label);
        for (int i = 0; i < v.num_edges(); i++)
            if (label < v.edge(i)->get_data()) v.edge(i)->set_data(label);
    }
};

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    metrics m("connected-components");
    std::string file = get_option_string("file");
    int nshards = convert_if_notexists<vid_t>(file, get_option_string("nshards", "auto"));
    ConnectedComponentsProgram program;
    graphchi_engine<vid_t, vid_t> engine(file, nshards, false, m);
    engine.run(program, get_option_int("niters", 1000));
    analyze_labels<vid_t>(file);
    metrics_report(m);
    return 0;
}
