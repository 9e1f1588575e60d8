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
