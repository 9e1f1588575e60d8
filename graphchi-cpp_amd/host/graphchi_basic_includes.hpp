// Everything a GraphChi application includes (src/graphchi_basic_includes.hpp),
// backed by the HIP library libgchi.so.
#pragma once
#include <assert.h>
#include <cmath>
#include <cstdio>
#include <string>

#include "api/chifilenames.hpp"
#include "api/graph_objects.hpp"
#include "api/graphchi_context.hpp"
#include "api/graphchi_program.hpp"
#include "api/vertex_aggregator.hpp"
#include "engine/graphchi_engine.hpp"
#include "graphchi_types.hpp"
#include "logger/logger.hpp"
#include "metrics/metrics.hpp"
#include "preprocessing/conversions.hpp"
#include "util/cmdopts.hpp"
#include "util/pthread_tools.hpp"

#if defined(__HIPCC__)
// An application compiled with hipcc runs its programs' own update() on the
// device (engine/psw_engine.hpp).  What follows this header in the
// application's source is compiled for the host and the device alike, so the
// program's members (defined in its class, hence inline) can be called from
// the engine's kernel; on the device pass main() becomes an unused inline
// function, so its host-only body is never compiled for the device.  An
// assert() in update() is reported by the engine after the pass instead of
// trapping the device.
#undef assert
#if defined(NDEBUG)
#define assert(e) ((void)0)
#else
#define assert(e) ::graphchi::psw::check_assert((e), #e, __FILE__, __LINE__)
#endif
#if defined(__HIP_DEVICE_COMPILE__)
#define main inline gchi_device_pass_main_
#endif
#pragma clang force_cuda_host_device begin
#endif
