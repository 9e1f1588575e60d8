// Per-vertex degrees as stored in <base>_degs.bin (src/engine/auxdata/degree_data.hpp:45-48).
#pragma once
#include <string>
#include <vector>

#include "api/chifilenames.hpp"
#include "shards/shard_reader.hpp"

namespace graphchi {

struct degree {
    int indegree;
    int outdegree;
};

inline std::vector<degree> load_degrees(std::string base) {
    std::vector<uint8_t> b = read_file_bytes(filename_degree_data(base));
    std::vector<degree> d(b.size() / sizeof(degree));
    if (!d.empty()) memcpy(d.data(), b.data(), d.size() * sizeof(degree));
    return d;
}

}  // namespace graphchi
