// Vertex and edge objects of the program API (src/api/graph_objects.hpp:55-376).
//
// A vertex is a view: its id, a pointer to its value, and its in- and
// out-edges as arrays of edge objects (each an adjacent vertex id, a copy of
// the edge's value and a pointer to the other endpoint's copy) owned by
// whoever built the view.  The reference's edge object holds a pointer to the
// value instead (data_ptr); programs reach it through get_data / set_data,
// which behave the same.  The generic device
// engine (engine/psw_engine.hpp) builds these views in HBM, so every member a
// program's update() calls is callable on the device when the program is
// compiled with hipcc; with a host compiler they are ordinary inline members.
#pragma once
#include <stddef.h>

#include "graphchi_types.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GCHI_HD __host__ __device__
#else
#define GCHI_HD
#endif

namespace graphchi {

// Deleted edges carry a marker value in their edge data (the convention of
// src/api/graph_objects.hpp:96-140 under SUPPORT_DELETIONS): true for bool,
// all ones for 32-bit integers, zero for float.
GCHI_HD inline bool is_deleted_edge_value(bool v) { return v; }
GCHI_HD inline bool is_deleted_edge_value(int v) { return (unsigned int)v == 0xffffffffu; }
GCHI_HD inline bool is_deleted_edge_value(unsigned int v) { return v == 0xffffffffu; }
GCHI_HD inline bool is_deleted_edge_value(float v) { return !(v < 0 || v > 0); }

template <typename T>
GCHI_HD inline T deleted_edge_value();
template <>
GCHI_HD inline bool deleted_edge_value<bool>() { return true; }
template <>
GCHI_HD inline int deleted_edge_value<int>() { return -1; }
template <>
GCHI_HD inline unsigned int deleted_edge_value<unsigned int>() { return 0xffffffffu; }
template <>
GCHI_HD inline float deleted_edge_value<float>() { return 0.0f; }

// An edge as one of its endpoints sees it: the other endpoint's id and the
// edge's value.  Every edge has two views, the source's out-edge and the
// target's in-edge; each holds a copy of the value and a pointer to the other
// copy (twin), and set_data() writes both.  A read is one load from the view
// itself, next to the vertex's other views, rather than a load through a
// pointer into a shared array.  A view built over outside storage,
// graphchi_edge(v, p), starts from *p and writes through to it.
template <typename EdgeDataType>
class graphchi_edge {
  public:
    vid_t vertexid = 0;
    EdgeDataType value = EdgeDataType();
    EdgeDataType *twin = nullptr;
    GCHI_HD graphchi_edge() {}
    GCHI_HD graphchi_edge(vid_t v, EdgeDataType *p) : vertexid(v), value(p ? *p : EdgeDataType()), twin(p) {}
    GCHI_HD graphchi_edge(vid_t v, EdgeDataType x, EdgeDataType *t) : vertexid(v), value(x), twin(t) {}
    GCHI_HD EdgeDataType get_data() { return value; }
    GCHI_HD void set_data(EdgeDataType x) {
        value = x;
        if (twin) *twin = x;
    }
    GCHI_HD vid_t vertex_id() const { return vertexid; }
    // the view whose value `twin` points at, when both views are views
    GCHI_HD static graphchi_edge *view_of(EdgeDataType *v) {
        return reinterpret_cast<graphchi_edge *>(reinterpret_cast<char *>(v) - offsetof(graphchi_edge, value));
    }
};

template <typename VertexDataType, typename EdgeDataType>
class graphchi_vertex {
  public:
    typedef EdgeDataType edge_data_type;
    vid_t vertexid = 0;
    VertexDataType *dataptr = nullptr;
    bool scheduled = true;
    bool parallel_safe = true;
    bool twins_are_views = false;  // the engine's views: every twin is the value of another view
    int inc = 0, outc = 0;  // in- and out-edge counts (the reference's public members)
    graphchi_edge<EdgeDataType> *inedges_ptr = nullptr, *outedges_ptr = nullptr;  // the view's edge arrays
    VertexDataType local = VertexDataType();  // the value when the view has no storage for it

    GCHI_HD graphchi_vertex() {}
    // The reference's constructor: id, in-edge and out-edge storage, counts.
    GCHI_HD graphchi_vertex(vid_t id, graphchi_edge<EdgeDataType> *iptr, graphchi_edge<EdgeDataType> *optr, int indeg,
                            int outdeg)
        : vertexid(id), inc(indeg), outc(outdeg), inedges_ptr(iptr), outedges_ptr(optr) {}
    GCHI_HD vid_t id() const { return vertexid; }
    GCHI_HD int num_inedges() const { return inc; }
    GCHI_HD int num_outedges() const { return outc; }
    GCHI_HD int num_edges() const { return inc + outc; }
    GCHI_HD graphchi_edge<EdgeDataType> *inedge(int i) { return &inedges_ptr[i]; }
    GCHI_HD graphchi_edge<EdgeDataType> *outedge(int i) { return &outedges_ptr[i]; }
    // In-edges first, then out-edges (graph_objects.hpp:283-286).
    GCHI_HD graphchi_edge<EdgeDataType> *edge(int i) { return i < inc ? &inedges_ptr[i] : &outedges_ptr[i - inc]; }
    GCHI_HD graphchi_edge<EdgeDataType> *random_outedge() { return outc ? &outedges_ptr[0] : nullptr; }
    GCHI_HD VertexDataType get_data() { return dataptr ? *dataptr : local; }
    GCHI_HD void set_data(VertexDataType d) {
        if (dataptr) *dataptr = d;
        else local = d;
    }
    // Appends to the view's edge storage, which the caller sized.
    GCHI_HD void add_inedge(vid_t src, EdgeDataType *ptr, bool) {
        inedges_ptr[inc++] = graphchi_edge<EdgeDataType>(src, ptr);
    }
    GCHI_HD void add_outedge(vid_t dst, EdgeDataType *ptr, bool) {
        outedges_ptr[outc++] = graphchi_edge<EdgeDataType>(dst, ptr);
    }
    // Orders the in- and out-edges by adjacent vertex id (stable).  A moved
    // view's twin is pointed at its new place, one array at a time, so that a
    // self-loop's two views (one in each array) stay linked.
    GCHI_HD void sort_edges_indirect() {
        sort_by_id(inedges_ptr, inc);
        relink(inedges_ptr, inc);
        sort_by_id(outedges_ptr, outc);
        relink(outedges_ptr, outc);
    }
    // Marks edges deleted (their data takes the deletion marker).
    GCHI_HD void remove_edge(int i) { edge(i)->set_data(deleted_edge_value<EdgeDataType>()); }
    GCHI_HD void remove_inedge(int i) { inedge(i)->set_data(deleted_edge_value<EdgeDataType>()); }
    GCHI_HD void remove_outedge(int i) { outedge(i)->set_data(deleted_edge_value<EdgeDataType>()); }
    GCHI_HD void remove_alledges() {
        for (int i = 0; i < num_edges(); i++) remove_edge(i);
    }
    GCHI_HD void remove_alloutedges() {
        for (int i = 0; i < outc; i++) remove_outedge(i);
    }
    GCHI_HD void remove_allinedges() {
        for (int i = 0; i < inc; i++) remove_inedge(i);
    }
    GCHI_HD bool computational_edges() { return false; }
    GCHI_HD static bool read_outedges() { return true; }

  private:
    GCHI_HD void relink(graphchi_edge<EdgeDataType> *e, int n) {
        if (!twins_are_views) return;
        for (int i = 0; i < n; i++)
            if (e[i].twin) graphchi_edge<EdgeDataType>::view_of(e[i].twin)->twin = &e[i].value;
    }
    GCHI_HD static void sort_by_id(graphchi_edge<EdgeDataType> *e, int n) {
        for (int i = 1; i < n; i++) {
            graphchi_edge<EdgeDataType> x = e[i];
            int j = i - 1;
            while (j >= 0 && e[j].vertexid > x.vertexid) {
                e[j + 1] = e[j];
                j--;
            }
            e[j + 1] = x;
        }
    }
};

}  // namespace graphchi
