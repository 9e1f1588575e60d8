// Vertex and edge objects of the program API (src/api/graph_objects.hpp:55-376).
// They exist so that update() functions compile unchanged; the device engines
// run registered programs without constructing them per vertex.
#pragma once
#include <algorithm>
#include <vector>

#include "graphchi_types.hpp"

namespace graphchi {

// Deleted edges carry a marker value in their edge data (the convention of
// src/api/graph_objects.hpp:96-140 under SUPPORT_DELETIONS): true for bool,
// all ones for 32-bit integers, zero for float.
inline bool is_deleted_edge_value(bool v) { return v; }
inline bool is_deleted_edge_value(int v) { return (unsigned int)v == 0xffffffffu; }
inline bool is_deleted_edge_value(unsigned int v) { return v == 0xffffffffu; }
inline bool is_deleted_edge_value(float v) { return !(v < 0 || v > 0); }

template <typename T>
inline T deleted_edge_value();
template <>
inline bool deleted_edge_value<bool>() { return true; }
template <>
inline int deleted_edge_value<int>() { return -1; }
template <>
inline unsigned int deleted_edge_value<unsigned int>() { return 0xffffffffu; }
template <>
inline float deleted_edge_value<float>() { return 0.0f; }

template <typename EdgeDataType>
class graphchi_edge {
  public:
    vid_t vertexid = 0;
    EdgeDataType *data_ptr = nullptr;
    graphchi_edge() {}
    graphchi_edge(vid_t v, EdgeDataType *p) : vertexid(v), data_ptr(p) {}
    EdgeDataType get_data() { return *data_ptr; }
    void set_data(EdgeDataType x) { *data_ptr = x; }
    vid_t vertex_id() const { return vertexid; }
};

template <typename VertexDataType, typename EdgeDataType>
class graphchi_vertex {
  public:
    typedef EdgeDataType edge_data_type;
    vid_t vertexid = 0;
    VertexDataType *dataptr = nullptr;
    bool scheduled = true;
    bool parallel_safe = true;
    std::vector<graphchi_edge<EdgeDataType>> inc, outc;
    VertexDataType local = VertexDataType();

    graphchi_vertex() {}
    graphchi_vertex(vid_t id, graphchi_edge<EdgeDataType> *, graphchi_edge<EdgeDataType> *, int, int)
        : vertexid(id) {}
    vid_t id() const { return vertexid; }
    int num_inedges() const { return (int)inc.size(); }
    int num_outedges() const { return (int)outc.size(); }
    int num_edges() const { return (int)(inc.size() + outc.size()); }
    graphchi_edge<EdgeDataType> *inedge(int i) { return &inc[i]; }
    graphchi_edge<EdgeDataType> *outedge(int i) { return &outc[i]; }
    graphchi_edge<EdgeDataType> *edge(int i) {
        return i < (int)inc.size() ? &inc[i] : &outc[i - inc.size()];
    }
    graphchi_edge<EdgeDataType> *random_outedge() { return outc.empty() ? nullptr : &outc[0]; }
    VertexDataType get_data() { return dataptr ? *dataptr : local; }
    void set_data(VertexDataType d) {
        if (dataptr) *dataptr = d;
        else local = d;
    }
    void add_inedge(vid_t src, EdgeDataType *ptr, bool) { inc.push_back(graphchi_edge<EdgeDataType>(src, ptr)); }
    void add_outedge(vid_t dst, EdgeDataType *ptr, bool) { outc.push_back(graphchi_edge<EdgeDataType>(dst, ptr)); }
    void sort_edges_indirect() {
        auto by_id = [](const graphchi_edge<EdgeDataType> &a, const graphchi_edge<EdgeDataType> &b) {
            return a.vertexid < b.vertexid;
        };
        std::sort(inc.begin(), inc.end(), by_id);
        std::sort(outc.begin(), outc.end(), by_id);
    }
    // Marks edge i deleted (its data takes the deletion marker).
    void remove_edge(int i) { edge(i)->set_data(deleted_edge_value<EdgeDataType>()); }
    bool computational_edges() { return false; }
    static bool read_outedges() { return true; }
};

}  // namespace graphchi
