// Readers of the per-vertex result file <base>.<sizeof VT>B.vout
// (src/api/vertex_aggregator.hpp:67-129).
#pragma once
#include <cstdio>
#include <string>
#include <vector>

#include "api/chifilenames.hpp"
#include "api/graph_objects.hpp"
#include "logger/logger.hpp"

namespace graphchi {

template <typename VertexDataType>
class VCallback {
  public:
    // host-and-device with hipcc, so that callbacks defined in an application
    // compiled with it still override these
    GCHI_HD virtual ~VCallback() {}
    GCHI_HD virtual void callback(vid_t vertex_id, VertexDataType &value) = 0;
};

template <typename VertexDataType>
std::vector<VertexDataType> read_vertex_values(std::string base, vid_t from = 0, vid_t to = 0) {
    std::string fname = filename_vertex_data<VertexDataType>(base);
    FILE *f = fopen(fname.c_str(), "rb");
    if (!f) logstream(LOG_FATAL) << "Could not open " << fname << std::endl;
    fseek(f, 0, SEEK_END);
    size_t n = (size_t)ftell(f) / sizeof(VertexDataType);
    if (to == 0 || to > n) to = (vid_t)n;
    std::vector<VertexDataType> v(to > from ? to - from : 0);
    fseek(f, (long)(from * sizeof(VertexDataType)), SEEK_SET);
    if (!v.empty() && fread(v.data(), sizeof(VertexDataType), v.size(), f) != v.size())
        logstream(LOG_FATAL) << "short read from " << fname << std::endl;
    fclose(f);
    return v;
}

template <typename VertexDataType>
void write_vertex_values(std::string base, const std::vector<VertexDataType> &v) {
    std::string fname = filename_vertex_data<VertexDataType>(base);
    FILE *f = fopen(fname.c_str(), "wb");
    if (!f) logstream(LOG_FATAL) << "Could not write " << fname << std::endl;
    if (!v.empty() && fwrite(v.data(), sizeof(VertexDataType), v.size(), f) != v.size())
        logstream(LOG_FATAL) << "short write to " << fname << std::endl;
    fclose(f);
}

template <typename VertexDataType>
void foreach_vertices(std::string base, vid_t from, vid_t to, VCallback<VertexDataType> &cb) {
    std::vector<VertexDataType> v = read_vertex_values<VertexDataType>(base, from, to);
    for (size_t i = 0; i < v.size(); i++) cb.callback((vid_t)(from + i), v[i]);
}

template <typename VertexDataType, typename SumType>
SumType sum_vertices(std::string base, vid_t from, vid_t to) {
    std::vector<VertexDataType> v = read_vertex_values<VertexDataType>(base, from, to);
    SumType s = SumType();
    for (auto &x : v) s += (SumType)x;
    return s;
}

}  // namespace graphchi
