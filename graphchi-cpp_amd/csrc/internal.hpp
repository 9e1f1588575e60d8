// Internal (non-exported) entry points shared between translation units.
#pragma once
#include "common.hpp"

namespace gchi {

size_t sort_workspace_bytes(uint64_t n, int key_bits, bool vals);
// err: the caller's status word (GCHI_WORK_* bits, cleared by the caller and
// read back with read_work_status once the stream has run); null = the sort
// workspace's own word (its first 4 bytes), cleared by the sort.
int radix_sort_u64(uint64_t *keys, uint32_t *vals, uint64_t n, int key_bits, void *work, size_t work_bytes,
                   hipStream_t s, uint32_t *err);
int radix_sort_u64_bits(uint64_t *keys, uint64_t *alt, uint32_t *vals, uint64_t n, int lo, int hi, void *work,
                        size_t work_bytes, hipStream_t s, uint64_t **out, uint32_t *err);
// The rank-space tiled sweep (launch form 4, pr_seg.hip).
}  // namespace gchi
struct gchi_seg_plan;
namespace gchi {
int seg_plan_create(const gchi_dgraph *g, int hot, gchi_seg_plan **out, hipStream_t s);
void seg_plan_destroy(gchi_seg_plan *p);
void seg_plan_info(const gchi_seg_plan *p, uint32_t *tiles, uint32_t *nsplit, uint64_t *slots, uint32_t *hot,
                   uint32_t *nwg);
int seg_sweep(const gchi_seg_plan *p, const gchi_pr_op *op, const float *x_in, const float *r_in, float *x_out,
              float *r_out, double *d_stats, hipStream_t s);
int seg_sweep_rank(const gchi_seg_plan *p, const gchi_pr_op *op, int src, int dst, double *d_stats, hipStream_t s);
int seg_enter(const gchi_seg_plan *p, const float *x, const float *r, int dst, hipStream_t s);
int seg_leave(const gchi_seg_plan *p, int src, float *x, float *r, hipStream_t s);
int seg_init(const gchi_seg_plan *p, float init, const uint32_t *seed_bits, float seed_value, int dst, hipStream_t s);

// Takes the device address of tc.hip's allocation-free scratch on device dev
// (once per device; gchi_device_open calls it).
int tc_scratch_setup(int dev);
// Waits for s, then maps a status word to GCHI_ERR_TIMEOUT / GCHI_ERR_INVALID.
int read_work_status(const uint32_t *err, hipStream_t s);

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
