// Per-run information handed to programs (src/api/graphchi_context.hpp:45-105)
// and the scheduler interface (src/api/ischeduler.hpp).
#pragma once
#include <stddef.h>

#include <vector>

#include "api/graph_objects.hpp"
#include "graphchi_types.hpp"

namespace graphchi {

// The scheduler interface (src/api/ischeduler.hpp:38-46).  Host-and-device
// with hipcc: the generic device engine gives update() a scheduler object
// built on the device (engine/psw_engine.hpp).  This base class schedules
// nothing and reports every vertex scheduled (the reference's non_scheduler).
class ischeduler {
  public:
    GCHI_HD virtual ~ischeduler() {}
    GCHI_HD virtual void add_task(vid_t, bool = false) {}
    GCHI_HD virtual void remove_tasks(vid_t, vid_t) {}
    GCHI_HD virtual void add_task_to_all() {}
    GCHI_HD virtual bool is_scheduled(vid_t) { return true; }
    GCHI_HD virtual size_t num_tasks() { return 0; }
    GCHI_HD virtual void new_iteration(int) {}
};

// The device engine copies the context into HBM for every pass and reads only
// its plain fields there; log_change() adds to device_delta when it is called
// on the device.
struct graphchi_context {
    size_t nvertices = 0;
    size_t nedges = 0;
    int iteration = 0;
    int num_iterations = 0;
    int last_iteration = -1;
    int execthreads = 1;
    ischeduler *scheduler = nullptr;
    double *device_delta = nullptr;
    std::vector<double> deltas = std::vector<double>(1, 0.0);

    GCHI_HD void set_last_iteration(int it) { last_iteration = it; }
    GCHI_HD bool isfinal() const { return last_iteration >= 0 && iteration >= last_iteration; }
    GCHI_HD void log_change(double delta) {
#if defined(__HIP_DEVICE_COMPILE__)
        if (device_delta) atomicAdd(device_delta, delta);
#else
        deltas[0] += delta;
#endif
    }
    double get_delta() const { return deltas[0]; }
    void reset_deltas(int) { deltas.assign(1, 0.0); }
};

}  // namespace graphchi
