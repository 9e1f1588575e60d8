// Shards an edge list into the reference's on-disk layout with the device sharder.
//   sharder file <edgelist> nshards <n|auto> [edgedatatype float|uint|float-float|none]
#include <string>

#include "graphchi_basic_includes.hpp"

using namespace graphchi;

int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    std::string file = get_option_string("file");
    std::string nshards = get_option_string("nshards", "auto");
    std::string type = get_option_string("edgedatatype", "float");
    int n = 0;
    if (type == "float") n = convert<float, float>(file, nshards);
    else if (type == "uint") n = convert<unsigned int, unsigned int>(file, nshards);
    else if (type == "float-float") n = convert<PairContainer<float>, PairContainer<float>>(file, nshards);
    else if (type == "none") n = convert<dummy, dummy>(file, nshards, true);
    else logstream(LOG_FATAL) << "edgedatatype must be float, uint, float-float or none" << std::endl;
    printf("Created %d shards\n", n);
    return 0;
}
