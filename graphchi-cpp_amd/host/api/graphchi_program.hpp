// The vertex-program interface (src/api/graphchi_program.hpp:42-87).  With
// hipcc the members are host-and-device, so a program's own update() can
// override them from device-compilable code (engine/psw_engine.hpp).
#pragma once
#include "api/graph_objects.hpp"
#include "api/graphchi_context.hpp"

namespace graphchi {

template <typename VertexDataType_, typename EdgeDataType_,
          typename vertex_t = graphchi_vertex<VertexDataType_, EdgeDataType_>>
class GraphChiProgram {
  public:
    typedef VertexDataType_ VertexDataType;
    typedef EdgeDataType_ EdgeDataType;
    GCHI_HD virtual ~GraphChiProgram() {}
    GCHI_HD virtual void before_iteration(int iteration, graphchi_context &gcontext) {}
    GCHI_HD virtual void after_iteration(int iteration, graphchi_context &gcontext) {}
    GCHI_HD virtual void before_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
    GCHI_HD virtual void after_exec_interval(vid_t window_st, vid_t window_en, graphchi_context &gcontext) {}
    GCHI_HD virtual void update(vertex_t &v, graphchi_context &gcontext) = 0;
    GCHI_HD virtual bool repeat_updates(graphchi_context &gcontext) { return false; }
};

}  // namespace graphchi
