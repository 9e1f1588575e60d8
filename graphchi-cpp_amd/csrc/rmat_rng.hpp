// Counter-based R-MAT edge generator shared by the host and device builds.
//
// Edge i of a graph with seed S is a pure function of (S, i): for every pair of
// recursion levels one 64-bit word  h = mix64(key(S) + 64*i + j)  is drawn,
// where mix64 is the SplitMix64 output function (Steele, Lea & Flood, OOPSLA
// 2014) and key(S) = mix64(S).  Level l (l = 0 picks the most significant bit)
// uses the low 32 bits of word l/2 when l is even and the high 32 bits when l
// is odd.  The uniform u in [0, 2^32) picks a quadrant of the adjacency matrix:
//   u <  A            -> (src bit 0, dst bit 0)
//   A <= u < A+B      -> (src bit 0, dst bit 1)
//   A+B <= u < A+B+C  -> (src bit 1, dst bit 0)
//   otherwise         -> (src bit 1, dst bit 1)
// with A, A+B, A+B+C the cumulative probabilities scaled by 2^32 and floored.
// There is no noise and no vertex permutation, so the generator is exactly
// reproducible on any machine.
#pragma once

#include <stdint.h>

#ifndef GCHI_RNG_HD
#if defined(__HIPCC__)
#define GCHI_RNG_HD __host__ __device__ inline
#else
#define GCHI_RNG_HD inline
#endif
#endif

namespace gchi {

GCHI_RNG_HD uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct RmatThresholds {
    uint64_t t0, t1, t2;  // cumulative thresholds on a 2^32 scale
};

GCHI_RNG_HD RmatThresholds rmat_thresholds(const double abcd[4]) {
    const double scale = 4294967296.0;
    double c0 = abcd[0], c1 = abcd[0] + abcd[1], c2 = abcd[0] + abcd[1] + abcd[2];
    RmatThresholds t;
    t.t0 = (uint64_t)(c0 * scale);
    t.t1 = (uint64_t)(c1 * scale);
    t.t2 = (uint64_t)(c2 * scale);
    return t;
}

GCHI_RNG_HD void rmat_edge(uint64_t key, uint64_t i, int scale, RmatThresholds t, uint32_t *src,
                           uint32_t *dst) {
    uint32_t s = 0, d = 0;
    uint64_t h = 0;
    for (int l = 0; l < scale; l++) {
        if ((l & 1) == 0) h = mix64(key + 64ull * i + (uint64_t)(l >> 1));
        uint64_t u = (l & 1) ? (h >> 32) : (h & 0xffffffffull);
        uint32_t bit = 1u << (scale - 1 - l);
        if (u < t.t0) {
        } else if (u < t.t1) {
            d |= bit;
        } else if (u < t.t2) {
            s |= bit;
        } else {
            s |= bit;
            d |= bit;
        }
    }
    *src = s;
    *dst = d;
}

}  // namespace gchi
