// Internal (non-exported) entry points shared between translation units.
#pragma once
#include "common.hpp"

namespace gchi {

size_t sort_workspace_bytes(uint64_t n, int key_bits, bool vals);
int radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, void *work, size_t work_bytes,
                   hipStream_t s);

unsigned grid_for(uint64_t n, unsigned threads, unsigned cap);
int degrees(const uint32_t *src, const uint32_t *dst, uint64_t n, uint32_t nv, uint32_t *outdeg, uint32_t *indeg,
            uint32_t *err, hipStream_t s);
size_t csc_workspace(uint64_t E, uint32_t N);
int csc_build(const uint32_t *src, const uint32_t *dst, uint64_t E, uint32_t N, uint32_t *in_off, uint32_t *in_src,
              uint32_t *outdeg, uint32_t *indeg_out, void *work, size_t work_bytes, hipStream_t s);

}  // namespace gchi
