// Internal (non-exported) entry points shared between translation units.
#pragma once
#include "common.hpp"

namespace gchi {

size_t sort_workspace_bytes(uint64_t n, int key_bits, bool vals);
int radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, void *work, size_t work_bytes,
                   hipStream_t s);
int radix_sort_u64_bits(uint64_t *keys, uint64_t *alt, uint32_t *vals, uint64_t n, int lo, int hi, void *work,
                        size_t work_bytes, hipStream_t s, uint64_t **out);

unsigned grid_for(uint64_t n, unsigned threads, unsigned cap);
// Over keys sorted on a field f = (key >> shift) & mask: cnt[f] = (or +=, with
// accumulate) the number of keys holding f, for f < nv (other entries of cnt
// untouched; fields >= nv skipped).  run_start is scratch of nv entries.
// low_out, when given, receives key & low_mask per key.
int count_field_runs(const uint64_t *keys, uint64_t n, int shift, uint32_t mask, uint32_t nv, uint32_t *run_start,
                     uint32_t *cnt, uint32_t *low_out, uint32_t low_mask, hipStream_t s, bool accumulate = false);
size_t csc_workspace(uint64_t E, uint32_t N);
int csc_build(const uint32_t *src, const uint32_t *dst, uint64_t E, uint32_t N, uint32_t *in_off, uint32_t *in_src,
              uint32_t *outdeg, uint32_t *indeg_out, void *work, size_t work_bytes, hipStream_t s);

}  // namespace gchi
