// Connected components by minimum label.
//
// Reference: ConnectedComponentsProgram::update (example_apps/connectedcomponents.cpp:79-121)
// propagates min labels over in- and out-edges until no edge label changes
// (:125-138); its fixed point is, for every vertex, the smallest vertex id of
// its weakly connected component.  Two device forms:
//   gchi_cc_minlabel  : lock-free union-find over the edge list.  A root is
//       always hooked under a smaller root with CAS, so every root is the
//       smallest id of its set; a final pass writes each vertex's root.  Result
//       identical to the reference's fixed point, independent of scheduling.
//   gchi_cc_labelprop_sweep : one synchronous min-propagation sweep
//       (label_out[v] = min(label_in[v], label_in[u] for every edge u-v)) with a
//       count of changed labels -- the reference's per-iteration semantics.
//   gchi_cc_scheduled : the reference's update under its selective scheduler
//       (scheduler on: connectedcomponents.cpp:81,85,112 with the two bitsets
//       of bitset_scheduler.hpp:38-104).  Round 0 schedules every vertex; a
//       vertex whose label drops schedules all its neighbours for the next
//       round; the rounds end when no label drops.  Device form: byte flags for
//       the current and next schedule (plain stores of 1 instead of atomic bit
//       sets), the scheduled vertices compacted per round into a list of short
//       rows (one lane each) and 4096-edge chunks of long rows (one wavefront
//       each, atomicMin on the row's label), labels updated in place -- so a
//       drop is seen by rows evaluated later in the same round, as add_task's
//       also_this_iteration does.  Same fixed point as the other forms.
#include <algorithm>
#include <vector>

#include "common.hpp"
#include "internal.hpp"

namespace gchi {

// Follows parent pointers to a root.  Loads may be stale across workgroups:
// every value parent[x] ever held is an ancestor of x in the same set, so a
// stale walk still ends at a member's ancestor; progress is guaranteed by the
// CAS below, which returns the current value.
GCHI_DEV uint32_t uf_find(uint32_t *parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        uint32_t gp = parent[p];
        if (gp != p) parent[x] = gp;  // path halving; gp is an ancestor of x
        x = p;
        p = gp;
    }
    return x;
}

__global__ void cc_init(uint32_t *parent, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) parent[v] = v;
}

// One union per edge.  With `half` (a symmetric list: every u-v present as both
// (u,v) and (v,u)) only the entries with u < v are used: the other half would
// union the same sets again.
__global__ void cc_hook(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint64_t e,
                        uint32_t n, uint32_t *parent, int half) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t u = __builtin_nontemporal_load(&src[i]);
        uint32_t v = __builtin_nontemporal_load(&dst[i]);
        if (u >= n || v >= n) continue;  // ids beyond N are rejected on the host side
        if (half && u >= v) continue;
        u = uf_find(parent, u);
        v = uf_find(parent, v);
        while (u != v) {
            if (u < v) {
                uint32_t t = u;
                u = v;
                v = t;
            }
            // u > v: hook root u under v
            uint32_t old = atomicCAS(&parent[u], u, v);
            if (old == u) break;
            u = uf_find(parent, old);
            v = uf_find(parent, v);
        }
    }
}

// The union of cc_hook for one pair: roots hooked under smaller roots by CAS.
GCHI_DEV void uf_link(uint32_t *parent, uint32_t u, uint32_t v) {
    u = uf_find(parent, u);
    v = uf_find(parent, v);
    while (u != v) {
        if (u < v) {
            uint32_t t = u;
            u = v;
            v = t;
        }
        uint32_t old = atomicCAS(&parent[u], u, v);
        if (old == u) break;
        u = uf_find(parent, old);
        v = uf_find(parent, v);
    }
}

// ---- subgraph sampling (Afforest: Sutton, Ben-Nun, Barak, "Optimizing
// Parallel Graph Connectivity Computation via Subgraph Sampling", IPDPS 2018)
// over a symmetric CSR: link every vertex to its neighbour r for r < kAffRounds,
// compress, find the component that most sampled vertices belong to (nearly
// always the giant one), then link the remaining neighbours of the vertices
// outside it only.  An edge between that component and another vertex is
// linked from the other vertex's row (the CSR holds it both ways), so the sets
// are those of the full union-find: every root is its set's smallest id.
constexpr uint32_t kAffRounds = 2;
constexpr uint32_t kAffSamples = 1024;
constexpr uint32_t kAffLong = 256;  // remaining neighbours above this: the whole wavefront links them

__global__ void cca_sample(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj, uint32_t n,
                           uint32_t r, uint32_t *parent, unsigned long long *bad) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        const uint32_t b = off[v], e = off[v + 1];
        if (e - b <= r) continue;
        const uint32_t u = adj[b + r];
        if (u >= n) {
            atomicAdd(bad, 1ull);
            continue;
        }
        uf_link(parent, v, u);
    }
}

// parent[v] = its root, comp[v] = the same (a snapshot for the skip test)
__global__ void cca_flatten(uint32_t *parent, uint32_t *comp, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        uint32_t r = v, p = parent[v];
        while (p != r) {
            r = p;
            p = parent[r];
        }
        parent[v] = r;
        comp[v] = r;
    }
}

// roots of kAffSamples vertices spread over [0, n) (multiplicative hash)
__global__ void cca_samples(const uint32_t *comp, uint32_t n, uint32_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kAffSamples) out[i] = comp[(uint32_t)(((uint64_t)i * 2654435761ull + 12345u) % n)];
}

// The neighbours from kAffRounds on of every vertex outside component c; rows
// with more than kAffLong of them are linked by the whole wavefront.
__global__ void cca_rest(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj, uint32_t n,
                         const uint32_t *__restrict__ comp, uint32_t c, uint32_t *parent, unsigned long long *bad) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(uint32_t)(kWave - 1)); base < n;
         base += stride) {
        const uint64_t v = base + lane;
        bool is_long = false;
        if (v < n && comp[v] != c) {
            const uint32_t b = off[v] + kAffRounds, e = off[v + 1];
            if (e > b && e - b > kAffLong) {
                is_long = true;
            } else {
                for (uint32_t k = b; k < e; k++) {
                    const uint32_t u = adj[k];
                    if (u < n) uf_link(parent, (uint32_t)v, u);
                    else atomicAdd(bad, 1ull);
                }
            }
        }
        uint64_t m = __ballot(is_long);
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t w = (uint32_t)(base + j);
            const uint32_t b = off[w] + kAffRounds, e = off[w + 1];
            for (uint32_t k = b + lane; k < e; k += kWave) {
                const uint32_t u = adj[k];
                if (u < n) uf_link(parent, w, u);
                else atomicAdd(bad, 1ull);
            }
        }
    }
}

__global__ void cc_compress(uint32_t *parent, uint32_t *labels, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        uint32_t r = v, p = parent[v];
        while (p != r) {
            r = p;
            p = parent[r];
        }
        labels[v] = r;
    }
}

__global__ void cc_copy(const uint32_t *a, uint32_t *b, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) b[v] = a[v];
}

__global__ void cc_prop_edges(const uint32_t *__restrict__ src, const uint32_t *__restrict__ dst, uint64_t e,
                              uint32_t n, const uint32_t *__restrict__ lin, uint32_t *lout) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t u = __builtin_nontemporal_load(&src[i]);
        uint32_t v = __builtin_nontemporal_load(&dst[i]);
        if (u >= n || v >= n) continue;
        uint32_t lu = lin[u], lv = lin[v];
        if (lu < lv) atomicMin(&lout[v], lu);
        else if (lv < lu) atomicMin(&lout[u], lv);
    }
}

__global__ void cc_count_changed(const uint32_t *lin, const uint32_t *lout, uint32_t n, uint64_t *changed) {
    uint32_t c = 0;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) c += lin[v] != lout[v];
    c = wave_sum(c);
    if ((threadIdx.x & (kWave - 1)) == 0 && c) atomicAdd((unsigned long long *)changed, (unsigned long long)c);
}

// One synchronous sweep pulled over a symmetric CSR: l_out[v] = min(l_in[v],
// l_in[u] for u in adj(v)).  Rows up to kPullShort entries: one lane each,
// reading its row in order; rows up to kPullHuge: the whole wavefront,
// lane-strided, then a min across lanes; longer rows (R-MAT's hubs hold
// millions of entries, which one wavefront would walk for milliseconds): l_out
// gets l_in here, and cc_pull_huge splits their entries over many wavefronts.
// *changed += labels that dropped (one atomic per wavefront).
constexpr uint32_t kPullShort = 32;
constexpr uint32_t kPullHuge = 4096;  // also the entries per wavefront of cc_pull_huge

__global__ void cc_pull_sweep(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj, uint32_t n,
                              const uint32_t *__restrict__ lin, uint32_t *__restrict__ lout,
                              unsigned long long *changed) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    uint32_t dropped = 0;
    for (uint64_t base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(uint32_t)(kWave - 1)); base < n;
         base += stride) {
        const uint64_t v = base + lane;
        bool is_long = false;
        if (v < n) {
            const uint32_t b = off[v], e = off[v + 1];
            uint32_t m = lin[v];
            if (e - b > kPullHuge) {
                lout[v] = m;  // cc_pull_huge lowers it
            } else if (e - b > kPullShort) {
                is_long = true;
            } else {
                for (uint32_t k = b; k < e; k++) {
                    const uint32_t u = adj[k];
                    m = min(m, lin[u < n ? u : (uint32_t)v]);
                }
                lout[v] = m;
                dropped += m != lin[v];
            }
        }
        uint64_t mask = __ballot(is_long);
        while (mask) {
            const int j = __builtin_ctzll(mask);
            mask &= mask - 1;
            const uint32_t w = (uint32_t)(base + j);
            const uint32_t b = off[w], e = off[w + 1];
            const uint32_t own = lin[w];
            uint32_t m = own;
            for (uint32_t k = b + lane; k < e; k += kWave) {
                const uint32_t u = adj[k];
                m = min(m, lin[u < n ? u : w]);
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o, kWave));
            if (lane == j) {
                lout[w] = m;
                dropped += m != own;
            }
        }
    }
    dropped = wave_sum(dropped);
    if (lane == 0 && dropped) atomicAdd(changed, (unsigned long long)dropped);
}

// Row holding entry k (the last r with off[r] <= k, off[0] = 0 <= k < off[n]):
// a 64-ary search, one load per lane per step.  Call with all lanes.
GCHI_DEV uint32_t pull_row_of(const uint32_t *__restrict__ off, uint32_t n, uint32_t k, int lane) {
    uint32_t lo = 0, hi = n;  // the row is in [lo, hi)
    while (hi - lo > 1) {
        const uint32_t span = hi - lo;
        const uint32_t step = (span + kWave - 1) / kWave;
        const uint32_t p = lo + (uint32_t)lane * step;
        const bool le = p < hi && off[p] <= k;
        const uint64_t m = __ballot(le);  // lanes 0..j hold probes <= k (lane 0 always)
        const int j = 63 - __builtin_clzll(m);
        lo = lo + (uint32_t)j * step;
        hi = min(hi, lo + step);
    }
    return lo;
}

// Rows of more than kPullHuge entries, split by entry range: wavefront c takes
// entries [c * kPullHuge, (c + 1) * kPullHuge).  Such a row is longer than the
// range, so the only rows of it the range can meet are the row of its first
// entry and the row of its last one.  Each takes the min of its part and lowers
// l_out with atomicMin (exact for integers, so the result does not depend on
// the order); the atomic that first takes a row below l_in counts its drop.
__global__ void cc_pull_huge(const uint32_t *__restrict__ off, const uint32_t *__restrict__ adj, uint32_t n,
                             const uint32_t *__restrict__ lin, uint32_t *lout, unsigned long long *changed) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t nwv = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    const uint32_t E = off[n];
    const uint64_t nchunks = ((uint64_t)E + kPullHuge - 1) / kPullHuge;
    uint32_t dropped = 0;
    for (uint64_t c = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave; c < nchunks; c += nwv) {
        const uint32_t c0 = (uint32_t)(c * kPullHuge);
        const uint32_t c1 = (uint32_t)min<uint64_t>((uint64_t)E, c * kPullHuge + kPullHuge);
        const uint32_t ra = pull_row_of(off, n, c0, lane);
        const uint32_t rb = pull_row_of(off, n, c1 - 1, lane);
        for (int t = 0; t < 2; t++) {
            const uint32_t r = t ? rb : ra;
            if (t && rb == ra) break;
            const uint32_t b = off[r], e = off[r + 1];
            if (e - b <= kPullHuge) continue;
            const uint32_t k0 = max(b, c0), k1 = min(e, c1);
            uint32_t m = 0xffffffffu;
            for (uint32_t k = k0 + lane; k < k1; k += kWave) {
                const uint32_t u = adj[k];
                m = min(m, lin[u < n ? u : r]);
            }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor((int)m, o, kWave));
            if (lane == 0) {
                const uint32_t own = lin[r];
                if (m < own) {
                    const uint32_t old = atomicMin(&lout[r], m);
                    dropped += old == own && m < old;
                }
            }
        }
    }
    dropped = wave_sum(dropped);
    if (lane == 0 && dropped) atomicAdd(changed, (unsigned long long)dropped);
}

__global__ void cc_iota(uint32_t *l, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) l[v] = v;
}

// ---- selective scheduling ----
constexpr uint32_t kCcsShort = 64;    // rows up to this length: one lane each
constexpr uint32_t kCcsChunk = 4096;  // longer rows: one wavefront per chunk

// ctr: [0] short rows, [1] long-row chunks.  stat: [0] labels dropped,
// [1] rows evaluated, [2] neighbour ids >= n seen.  Each wavefront compacts a
// contiguous vertex range in two passes (count, reserve once, write), so the
// shared counters see one atomic per wavefront and round, not one per 64
// vertices: atomics on one address complete one after another.
__global__ void ccs_compact(const uint32_t *off, uint32_t n, const uint8_t *cur, uint32_t *short_list,
                            uint64_t *items, uint32_t *ctr, unsigned long long *stat) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    const uint64_t per = ((n + nw - 1) / nw + kWave - 1) / kWave * kWave;
    const uint64_t lo = wid * per, hi = lo + per < n ? lo + per : n;
    uint32_t ns = 0, ni = 0, na = 0;
    for (uint64_t v = lo + lane; v < hi; v += kWave) {
        if (!cur[v]) continue;
        const uint32_t deg = off[v + 1] - off[v];
        na++;
        if (deg <= kCcsShort) ns++;
        else ni += (deg + kCcsChunk - 1) / kCcsChunk;
    }
    ns = wave_sum(ns);
    ni = wave_sum(ni);
    na = wave_sum(na);
    uint32_t sbase = 0, ibase = 0;
    if (lane == 0) {
        if (ns) sbase = atomicAdd(&ctr[0], ns);
        if (ni) ibase = atomicAdd(&ctr[1], ni);
        if (na) atomicAdd(&stat[1], (unsigned long long)na);
    }
    sbase = __shfl(sbase, 0, kWave);
    ibase = __shfl(ibase, 0, kWave);
    if (ns == 0 && ni == 0) return;
    for (uint64_t base = lo; base < hi; base += kWave) {  // whole wavefronts: ballots and shuffles
        const uint64_t v = base + lane;
        const bool act = v < hi && cur[v];
        const uint32_t deg = act ? off[v + 1] - off[v] : 0u;
        const bool sh = act && deg <= kCcsShort;
        const uint32_t nit = (act && deg > kCcsShort) ? (deg + kCcsChunk - 1) / kCcsChunk : 0u;
        const uint64_t bs = __ballot(sh);
        if (sh) short_list[sbase + __popcll(bs & lanemask_lt())] = (uint32_t)v;
        sbase += (uint32_t)__popcll(bs);
        uint32_t incl = nit;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += t;
        }
        for (uint32_t c = 0; c < nit; c++) items[ibase + incl - nit + c] = (v << 32) | c;
        ibase += __shfl(incl, kWave - 1, kWave);
    }
}

// stat[2] += rows whose offsets decrease or pass off[n].
__global__ void ccs_check_off(const uint32_t *off, uint32_t n, unsigned long long *stat) {
    const uint32_t E = off[n];
    uint32_t bad = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        bad += off[v] > off[v + 1] || off[v + 1] > E;
    bad = wave_sum(bad);
    if ((threadIdx.x & (kWave - 1)) == 0 && bad) atomicAdd(&stat[2], (unsigned long long)bad);
}

// One lane per short row: label = min over the row, neighbours scheduled if it dropped.
__global__ void ccs_short(const uint32_t *off, const uint32_t *adj, uint32_t n, uint32_t *label,
                          const uint32_t *short_list, const uint32_t *ctr, uint8_t *next, unsigned long long *stat) {
    const uint32_t cnt = ctr[0];
    uint32_t dropped = 0, bad = 0;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += stride) {
        const uint32_t v = short_list[i], b = off[v], e = off[v + 1];
        const uint32_t lv = label[v];
        uint32_t m = lv;
#pragma unroll 4
        for (uint32_t k = b; k < e; k++) {
            const uint32_t u = adj[k];
            if (u < n) m = min(m, label[u]);
            else bad = 1;
        }
        if (m < lv) {
            label[v] = m;  // the only writer of a short row's label
            dropped++;
            for (uint32_t k = b; k < e; k++) {
                const uint32_t u = adj[k];
                if (u < n) next[u] = 1;
            }
        }
    }
    dropped = wave_sum(dropped);
    bad = wave_sum(bad);
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (dropped) atomicAdd(&stat[0], (unsigned long long)dropped);
        if (bad) atomicAdd(&stat[2], (unsigned long long)bad);
    }
}

// One wavefront per chunk of a long row: min over the chunk, atomicMin into the row.
__global__ void ccs_long_min(const uint32_t *off, const uint32_t *adj, uint32_t n, uint32_t *label,
                             const uint64_t *items, const uint32_t *ctr, uint8_t *dropped_flag,
                             unsigned long long *stat) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t cnt = ctr[1];
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nw = (gridDim.x * blockDim.x) / kWave;
    uint32_t bad = 0;
    for (uint32_t j = wid; j < cnt; j += nw) {
        const uint64_t it = items[j];
        const uint32_t v = (uint32_t)(it >> 32), c = (uint32_t)it;
        const uint32_t b = off[v] + c * kCcsChunk, e = min(off[v + 1], b + kCcsChunk);
        const uint32_t lv = label[v];
        uint32_t m = 0xffffffffu;
        for (uint32_t k0 = b + lane; k0 < e; k0 += 4 * kWave) {  // four gathers in flight per lane
            uint32_t u[4];
#pragma unroll
            for (int q = 0; q < 4; q++) u[q] = k0 + q * kWave < e ? adj[k0 + q * kWave] : 0xffffffffu;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (u[q] < n) m = min(m, label[u[q]]);
                else if (k0 + q * kWave < e) bad = 1;
            }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m = min(m, (uint32_t)__shfl_xor(m, o, kWave));
        if (lane == 0 && m < lv) {
            const uint32_t old = atomicMin(&label[v], m);
            if (m < old) dropped_flag[v] = 1;
        }
    }
    bad = wave_sum(bad);
    if (lane == 0 && bad) atomicAdd(&stat[2], (unsigned long long)bad);
}

// Schedules the neighbours of every long row whose label dropped (per chunk).
__global__ void ccs_long_mark(const uint32_t *off, const uint32_t *adj, uint32_t n, const uint64_t *items,
                              const uint32_t *ctr, const uint8_t *dropped_flag, uint8_t *next,
                              unsigned long long *stat) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t cnt = ctr[1];
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave, nw = (gridDim.x * blockDim.x) / kWave;
    uint32_t dropped = 0;
    for (uint32_t j = wid; j < cnt; j += nw) {
        const uint64_t it = items[j];
        const uint32_t v = (uint32_t)(it >> 32), c = (uint32_t)it;
        if (!dropped_flag[v]) continue;
        if (c == 0) dropped++;
        const uint32_t b = off[v] + c * kCcsChunk, e = min(off[v + 1], b + kCcsChunk);
        for (uint32_t k0 = b + lane; k0 < e; k0 += 4 * kWave) {
            uint32_t u[4];
#pragma unroll
            for (int q = 0; q < 4; q++) u[q] = k0 + q * kWave < e ? adj[k0 + q * kWave] : 0xffffffffu;
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (u[q] < n) next[u[q]] = 1;
        }
    }
    if (lane == 0 && dropped) atomicAdd(&stat[0], (unsigned long long)dropped);
}

struct CcsWork {
    uint8_t *flags[2];
    uint8_t *dropped;
    uint32_t *short_list;
    uint64_t *items;
    uint32_t *ctr;
    unsigned long long *stat;
};

template <typename C>
static void ccs_layout(C &c, uint32_t n, uint64_t e, CcsWork *w) {
    const size_t nb = align_up((size_t)n + 1, 4);
    auto *f0 = c.template take<uint8_t>(nb);
    auto *f1 = c.template take<uint8_t>(nb);
    auto *dr = c.template take<uint8_t>(nb);
    auto *sl = c.template take<uint32_t>((size_t)n + 1);
    auto *it = c.template take<uint64_t>((size_t)n + e / kCcsChunk + 1);
    auto *ct = c.template take<uint32_t>(2);
    auto *st = c.template take<unsigned long long>(3);
    if (w) *w = CcsWork{{f0, f1}, dr, sl, it, ct, st};
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_cc_workspace(uint32_t nvertices, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    Sizer z;
    z.take<uint32_t>(nvertices ? nvertices : 1);
    *bytes = z.off;
    return GCHI_OK;
}

extern "C" int gchi_cc_minlabel_ex(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                                   int flags, uint32_t *d_labels, void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK((flags & ~GCHI_CC_SYMMETRIC) == 0, GCHI_ERR_INVALID, "cc_minlabel: unknown flags");
    GCHI_CHECK(d_labels && nvertices > 0, GCHI_ERR_INVALID, "cc_minlabel: null labels or no vertices");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "cc_minlabel: null edges");
    size_t need = 0;
    gchi_cc_workspace(nvertices, &need);
    GCHI_CHECK(d_work && work_bytes >= need, GCHI_ERR_WORKSPACE, "cc_minlabel: workspace too small");
    hipStream_t s = as_stream(stream);
    uint32_t *parent = static_cast<uint32_t *>(d_work);
    hipLaunchKernelGGL(cc_init, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, s, parent, nvertices);
    if (nedges)
        hipLaunchKernelGGL(cc_hook, dim3(grid_for(nedges, 256, 1u << 16)), dim3(256), 0, s, d_src, d_dst, nedges,
                           nvertices, parent, (flags & GCHI_CC_SYMMETRIC) ? 1 : 0);
    hipLaunchKernelGGL(cc_compress, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, s, parent, d_labels,
                       nvertices);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_cc_afforest_workspace(uint32_t nvertices, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    Sizer z;
    z.take<uint32_t>(nvertices ? nvertices : 1);  // parent
    z.take<uint32_t>(kAffSamples);
    z.take<unsigned long long>(1);
    *bytes = z.off;
    return GCHI_OK;
}

extern "C" int gchi_cc_afforest(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t *d_labels,
                                void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK(d_off && d_labels && nvertices > 0, GCHI_ERR_INVALID, "cc_afforest: null argument");
    size_t need = 0;
    gchi_cc_afforest_workspace(nvertices, &need);
    GCHI_CHECK(d_work && work_bytes >= need, GCHI_ERR_WORKSPACE, "cc_afforest: workspace too small");
    hipStream_t s = as_stream(stream);
    const uint32_t n = nvertices;
    Carver c(d_work, work_bytes);
    uint32_t *parent = c.take<uint32_t>(n);
    uint32_t *samp = c.take<uint32_t>(kAffSamples);
    unsigned long long *bad = c.take<unsigned long long>(1);
    uint32_t *comp = d_labels;  // the snapshot lives in the output until the last pass
    const unsigned gv = grid_for(n, 256, 65536);
    launch_fill_u64(reinterpret_cast<uint64_t *>(bad), 0ull, 1, s);
    hipLaunchKernelGGL(cc_init, dim3(gv), dim3(256), 0, s, parent, n);
    for (uint32_t r = 0; r < kAffRounds; r++)
        hipLaunchKernelGGL(cca_sample, dim3(gv), dim3(256), 0, s, d_off, d_adj, n, r, parent, bad);
    hipLaunchKernelGGL(cca_flatten, dim3(gv), dim3(256), 0, s, parent, comp, n);
    hipLaunchKernelGGL(cca_samples, dim3(kAffSamples / 256), dim3(256), 0, s, comp, n, samp);
    GCHI_LAUNCH_CHECK();
    std::vector<uint32_t> h(kAffSamples);
    GCHI_HIP(hipMemcpyAsync(h.data(), samp, 4 * (size_t)kAffSamples, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    std::sort(h.begin(), h.end());
    uint32_t best = h[0], best_n = 0;
    for (size_t i = 0; i < h.size();) {
        size_t j = i;
        while (j < h.size() && h[j] == h[i]) j++;
        if (j - i > best_n) {
            best_n = (uint32_t)(j - i);
            best = h[i];
        }
        i = j;
    }
    hipLaunchKernelGGL(cca_rest, dim3(grid_for(n, 256, 65536)), dim3(256), 0, s, d_off, d_adj, n, comp, best, parent,
                       bad);
    hipLaunchKernelGGL(cc_compress, dim3(gv), dim3(256), 0, s, parent, d_labels, n);
    GCHI_LAUNCH_CHECK();
    unsigned long long h_bad = 0;
    GCHI_HIP(hipMemcpyAsync(&h_bad, bad, 8, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    GCHI_CHECK(h_bad == 0, GCHI_ERR_INVALID, "cc_afforest: neighbour id >= nvertices");
    return GCHI_OK;
}

extern "C" int gchi_cc_minlabel(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                                uint32_t *d_labels, void *d_work, size_t work_bytes, void *stream) {
    return gchi_cc_minlabel_ex(d_src, d_dst, nedges, nvertices, 0, d_labels, d_work, work_bytes, stream);
}

extern "C" int gchi_cc_labelprop_init(uint32_t *d_labels, uint32_t nvertices, void *stream) {
    GCHI_CHECK(d_labels, GCHI_ERR_INVALID, "cc_labelprop_init: null labels");
    hipLaunchKernelGGL(cc_iota, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, as_stream(stream), d_labels,
                       nvertices);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_cc_labelprop_sweep(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges,
                                       uint32_t nvertices, const uint32_t *d_labels_in, uint32_t *d_labels_out,
                                       uint64_t *d_changed, void *stream) {
    GCHI_CHECK(d_labels_in && d_labels_out && d_changed && d_labels_in != d_labels_out, GCHI_ERR_INVALID,
               "cc_labelprop_sweep: bad label buffers");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "cc_labelprop_sweep: null edges");
    hipStream_t s = as_stream(stream);
    hipLaunchKernelGGL(cc_copy, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, s, d_labels_in, d_labels_out,
                       nvertices);
    if (nedges)
        hipLaunchKernelGGL(cc_prop_edges, dim3(grid_for(nedges, 256, 1u << 16)), dim3(256), 0, s, d_src, d_dst,
                           nedges, nvertices, d_labels_in, d_labels_out);
    launch_fill_u64(d_changed, 0ull, 1, s);
    hipLaunchKernelGGL(cc_count_changed, dim3(grid_for(nvertices, 256, 4096)), dim3(256), 0, s, d_labels_in,
                       d_labels_out, nvertices, d_changed);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_cc_pull_sweep(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices,
                                  const uint32_t *d_labels_in, uint32_t *d_labels_out, uint64_t *d_changed,
                                  void *stream) {
    GCHI_CHECK(d_off && d_labels_in && d_labels_out && d_changed && nvertices > 0 && d_labels_in != d_labels_out,
               GCHI_ERR_INVALID, "cc_pull_sweep: bad argument");
    hipStream_t s = as_stream(stream);
    launch_fill_u64(d_changed, 0ull, 1, s);
    hipLaunchKernelGGL(cc_pull_sweep, dim3(grid_for(nvertices, 256, 16384)), dim3(256), 0, s, d_off, d_adj, nvertices,
                       d_labels_in, d_labels_out, reinterpret_cast<unsigned long long *>(d_changed));
    hipLaunchKernelGGL(cc_pull_huge, dim3(4096), dim3(256), 0, s, d_off, d_adj, nvertices, d_labels_in, d_labels_out,
                       reinterpret_cast<unsigned long long *>(d_changed));
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_cc_scheduled_workspace(uint32_t nvertices, uint64_t nedges, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    SizerAdapter a;
    ccs_layout(a, nvertices, nedges, nullptr);
    *bytes = a.z.off;
    return GCHI_OK;
}

extern "C" int gchi_cc_scheduled_ex(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices,
                                    uint32_t max_rounds, uint32_t *d_labels, uint32_t *h_rounds, uint64_t *h_evaluated,
                                    gchi_cc_round_fn on_round, void *user, void *d_work, size_t work_bytes,
                                    void *stream) {
    GCHI_CHECK(d_off && d_labels && h_rounds && nvertices > 0, GCHI_ERR_INVALID, "cc_scheduled: null argument");
    hipStream_t s = as_stream(stream);
    uint32_t e32 = 0;
    GCHI_HIP(hipMemcpyAsync(&e32, d_off + nvertices, 4, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    const uint64_t E = e32;
    GCHI_CHECK(E == 0 || d_adj, GCHI_ERR_INVALID, "cc_scheduled: null adjacency");
    size_t need = 0;
    gchi_cc_scheduled_workspace(nvertices, E, &need);
    GCHI_CHECK(d_work && work_bytes >= need, GCHI_ERR_WORKSPACE, "cc_scheduled: workspace too small");
    Carver c(d_work, work_bytes);
    CcsWork w;
    ccs_layout(c, nvertices, E, &w);
    const uint32_t n = nvertices;
    const size_t nwords = align_up((size_t)n + 1, 4) / 4;
    const unsigned gv = grid_for(n, 256, 65536);
    const unsigned gl = grid_for(n, 256, 2048);  // list kernels loop over their lists
    const unsigned gw = 4096;  // wavefronts loop over the chunk list
    launch_fill_u64(reinterpret_cast<uint64_t *>(w.stat), 0ull, 3, s);
    hipLaunchKernelGGL(ccs_check_off, dim3(gv), dim3(256), 0, s, d_off, n, w.stat);
    unsigned long long h_bad = 0;
    GCHI_HIP(hipMemcpyAsync(&h_bad, w.stat + 2, 8, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    GCHI_CHECK(h_bad == 0, GCHI_ERR_INVALID, "cc_scheduled: offsets are not a CSR");
    hipLaunchKernelGGL(cc_iota, dim3(gv), dim3(256), 0, s, d_labels, n);
    launch_fill_u32(reinterpret_cast<uint32_t *>(w.flags[0]), 0x01010101u, nwords, s);  // round 0: every vertex
    // labels only drop and stay ids of the component: at most n rounds
    const uint64_t limit = max_rounds ? max_rounds : (uint64_t)n + 1;
    uint32_t rounds = 0;
    uint64_t evaluated = 0;
    bool fixed = false, stopped = false;
    int cur = 0;
    while (rounds < limit) {
        uint8_t *fc = w.flags[cur], *fn = w.flags[cur ^ 1];
        launch_fill_u32(reinterpret_cast<uint32_t *>(fn), 0u, nwords, s);
        launch_fill_u32(reinterpret_cast<uint32_t *>(w.dropped), 0u, nwords, s);
        launch_fill_u32(w.ctr, 0u, 2, s);
        launch_fill_u64(reinterpret_cast<uint64_t *>(w.stat), 0ull, 3, s);
        hipLaunchKernelGGL(ccs_compact, dim3(grid_for(n, 256 * 16, 1024)), dim3(256), 0, s, d_off, n, fc,
                           w.short_list, w.items, w.ctr, w.stat);
        hipLaunchKernelGGL(ccs_short, dim3(gl), dim3(256), 0, s, d_off, d_adj, n, d_labels, w.short_list, w.ctr, fn,
                           w.stat);
        hipLaunchKernelGGL(ccs_long_min, dim3(gw), dim3(256), 0, s, d_off, d_adj, n, d_labels, w.items, w.ctr,
                           w.dropped, w.stat);
        hipLaunchKernelGGL(ccs_long_mark, dim3(gw), dim3(256), 0, s, d_off, d_adj, n, w.items, w.ctr, w.dropped, fn,
                           w.stat);
        GCHI_LAUNCH_CHECK();
        unsigned long long st[3] = {0, 0, 0};
        GCHI_HIP(hipMemcpyAsync(st, w.stat, sizeof(st), hipMemcpyDeviceToHost, s));
        GCHI_HIP(hipStreamSynchronize(s));
        GCHI_CHECK(st[2] == 0, GCHI_ERR_INVALID, "cc_scheduled: neighbour id >= nvertices");
        rounds++;
        evaluated += st[1];
        if (st[0] == 0) fixed = true;
        if (on_round && on_round(user, rounds - 1, st[0], st[1]) != 0) {
            stopped = true;
            break;
        }
        if (fixed) break;
        cur ^= 1;
    }
    *h_rounds = rounds;
    if (h_evaluated) *h_evaluated = evaluated;
    GCHI_CHECK(fixed || stopped, GCHI_ERR_INVALID, "cc_scheduled: no fixed point within max_rounds");
    return GCHI_OK;
}

extern "C" int gchi_cc_scheduled(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t max_rounds,
                                 uint32_t *d_labels, uint32_t *h_rounds, uint64_t *h_evaluated, void *d_work,
                                 size_t work_bytes, void *stream) {
    return gchi_cc_scheduled_ex(d_off, d_adj, nvertices, max_rounds, d_labels, h_rounds, h_evaluated, nullptr, nullptr,
                                d_work, work_bytes, stream);
}
