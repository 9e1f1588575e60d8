// Top-k vertices by value from the result file (src/util/toplist.hpp:48-152).
#pragma once
#include <algorithm>
#include <string>
#include <vector>

#include "api/vertex_aggregator.hpp"

namespace graphchi {

template <typename VertexDataType>
struct vertex_value {
    vid_t vertex;
    VertexDataType value;
    vertex_value() {}
    vertex_value(vid_t v, VertexDataType x) : vertex(v), value(x) {}
};

template <typename VertexDataType>
std::vector<vertex_value<VertexDataType>> get_top_vertices(std::string basefilename, int ntop, vid_t from = 0,
                                                           vid_t to = 0) {
    std::vector<VertexDataType> v = read_vertex_values<VertexDataType>(basefilename, from, to);
    std::vector<vertex_value<VertexDataType>> all(v.size());
    for (size_t i = 0; i < v.size(); i++) all[i] = vertex_value<VertexDataType>((vid_t)(from + i), v[i]);
    size_t k = std::min<size_t>((size_t)std::max(ntop, 0), all.size());
    auto greater = [](const vertex_value<VertexDataType> &a, const vertex_value<VertexDataType> &b) {
        return a.value > b.value || (a.value == b.value && a.vertex < b.vertex);
    };
    std::partial_sort(all.begin(), all.begin() + k, all.end(), greater);
    all.resize(k);
    return all;
}

}  // namespace graphchi
