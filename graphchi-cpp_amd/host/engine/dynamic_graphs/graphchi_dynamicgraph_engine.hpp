// The dynamic-graph engine name used by triangle counting
// (src/engine/dynamic_graphs/graphchi_dynamicgraph_engine.hpp); it is the
// device engine of engine/graphchi_engine.hpp.
#pragma once
#include "engine/graphchi_engine.hpp"
