// Basic GraphChi types (same names as src/graphchi_types.hpp).
#pragma once
#include <stdint.h>

namespace graphchi {

typedef uint32_t vid_t;

// Two values per edge for bulk-synchronous computation: the value read in
// iteration t and the one written in t (src/graphchi_types.hpp:34-59).
template <typename ET>
struct PairContainer {
    ET left;
    ET right;
    PairContainer() : left(ET()), right(ET()) {}
    PairContainer(ET a, ET b) : left(a), right(b) {}
    ET &oldval(int iter) { return (iter % 2 == 0) ? left : right; }
    void set_newval(int iter, ET x) {
        if (iter % 2 == 0) right = x;
        else left = x;
    }
};

struct dummy {};

}  // namespace graphchi
