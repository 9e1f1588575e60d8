// Triangle counting on the degree-ordered graph.
//
// Reference: example_apps/trianglecounting.cpp.  The input is relabelled by
// order_by_degree (src/preprocessing/util/orderbydegree.hpp:58-100): vertices
// ranked by (in+out degree, id), vertexmap[old] = rank; per-vertex triangle
// counts are stored in the new id space as uint32 (<G>_degord.4B.vout,
// trianglecounting.cpp:72) and the total is their sum / 3 (:476-477).
//
// Device form:
//   gchi_tc_prepare : degrees, (degree,id) radix sort -> vertexmap, edges mapped
//       and oriented from lower to higher rank, sorted and deduplicated into an
//       oriented CSR (N+(v): the higher-ranked neighbours of v, ascending).
//   gchi_tc_count   : every triangle a<b<w is found once, at vertex a, as
//       w in N+(a) n N+(b) for b in N+(a).  One wavefront per vertex a.
//       |N+(a)| <= 1024 (the hash form, tc_count_hash_kernel): N+(a) goes
//       into an LDS hash table and every w of N+(b) is looked up lane-wide.
//       Longer N+(a) (tc_count_kernel): the (b, w in N+(b)) pairs of 64
//       consecutive b are flattened over the 64 lanes and each lane
//       binary-searches its w in N+(a) in global memory.  Counts: a, b and w
//       each get +1 per triangle, gathered per wavefront and flushed with
//       uint64 atomics, converted to uint32 with an overflow flag.
#include <mutex>

#include "common.hpp"
#include "internal.hpp"

namespace gchi {

constexpr int kTcWaves = 4;
constexpr int kTcLds = 1024;

// Degree keys: (dst << sb) | src with sb = the id field width rounded up to a
// whole radix digit; the id field is wide enough to hold n itself, which marks
// self-loops (and out-of-range ids, reported through err) so that run counting
// skips them.
__global__ void tc_degree_pack(const uint32_t *src, const uint32_t *dst, uint64_t e, uint32_t n, int sb,
                               uint64_t *keys, uint32_t *err) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t s = src[i], d = dst[i];
        if (s >= n || d >= n) atomicOr(err, 1u);
        if (s >= n || d >= n || s == d) s = d = n;
        keys[i] = ((uint64_t)d << sb) | s;
    }
}

__global__ void tc_degree_keys(const uint32_t *deg, uint32_t n, int idb, uint64_t *keys) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        keys[v] = ((uint64_t)deg[v] << idb) | v;
}

// Ids >= n appear only in the output of a sort whose look-back gave up (the
// caller reports that and discards the result): skipped so that no write leaves
// the arrays.  The same guard is in the two run kernels below.
__global__ void tc_vertexmap(const uint64_t *keys, uint32_t n, int idb, uint32_t *vmap) {
    uint32_t stride = gridDim.x * blockDim.x;
    const uint64_t m = (1ull << idb) - 1ull;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t v = keys[i] & m;
        if (v < n) vmap[v] = i;
    }
}

__global__ void tc_orient_keys(const uint32_t *src, const uint32_t *dst, uint64_t e, uint32_t n, const uint32_t *vmap, int idb,
                               uint64_t sentinel, uint64_t *keys) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t s = src[i], d = dst[i];
        if (s == d || s >= n || d >= n) {  // out-of-range ids are reported by tc_degree_pack
            keys[i] = sentinel;
            continue;
        }
        uint32_t a = vmap[s], b = vmap[d];
        uint32_t lo = min(a, b), hi = max(a, b);
        keys[i] = ((uint64_t)lo << idb) | hi;
    }
}

__global__ void tc_unique_flags(const uint64_t *keys, uint64_t e, uint64_t sentinel, uint32_t *flags) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint64_t k = keys[i];
        flags[i] = (k != sentinel && (i == 0 || keys[i - 1] != k)) ? 1u : 0u;
    }
}

// Runs of equal lo = key >> idb in the sorted oriented keys (sentinel keys,
// which sort last, excluded).  lo_start[lo] = pos at the run's first key.
__global__ void tc_lo_run_starts(const uint64_t *keys, const uint32_t *pos, uint64_t e, int idb, uint64_t sentinel,
                                 uint32_t n, uint32_t *lo_start) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        const uint64_t k = keys[i];
        if (k == sentinel) continue;
        const uint64_t lo = k >> idb;
        if (lo >= n) continue;
        if (i == 0 || (keys[i - 1] >> idb) != lo) lo_start[lo] = pos[i];
    }
}

// adj[pos] = hi of every unique key; at a run's last key, lo_count[lo] = the
// unique keys of the run (no atomics: one writer per run).
__global__ void tc_scatter_unique(const uint64_t *keys, const uint32_t *flags, const uint32_t *pos, uint64_t e,
                                  int idb, uint64_t sentinel, uint32_t n, const uint32_t *lo_start, uint32_t *adj,
                                  uint32_t *lo_count) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t m = (1ull << idb) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        const uint64_t k = keys[i];
        if (k == sentinel) continue;
        const uint64_t lo = k >> idb;
        if (lo >= n) continue;
        const uint32_t f = flags[i], p = pos[i];
        if (f) adj[p] = (uint32_t)(k & m);
        const bool last = i + 1 == e || keys[i + 1] == sentinel || (keys[i + 1] >> idb) != lo;
        if (last) lo_count[lo] = p + f - lo_start[lo];
    }
}

// Position of x in the ascending a[lo, hi), or -1.
GCHI_DEV int bsearch_u32(const uint32_t *a, uint32_t lo, uint32_t hi, uint32_t x) {
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        uint32_t v = a[mid];
        if (v == x) return (int)mid;
        if (v < x)
            lo = mid + 1;
        else
            hi = mid;
    }
    return -1;
}

__global__ __launch_bounds__(kTcWaves *kWave) void tc_count_kernel(const uint32_t *__restrict__ off,
                                                                   const uint32_t *__restrict__ adj, uint32_t n,
                                                                   uint32_t min_da, unsigned long long *cnt,
                                                                   unsigned long long *total) {
    __shared__ uint32_t sA[kTcWaves][kTcLds];
    __shared__ uint32_t sW[kTcWaves][kTcLds];  // triangles found per position of N+(a) (the w credit)
    __shared__ uint32_t sPre[kTcWaves][kWave];
    __shared__ uint32_t sBeg[kTcWaves][kWave];
    __shared__ uint32_t sCnt[kTcWaves][kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    // A static scan, 64 vertices per step, picks the vertices above min_da (no
    // work counter: the few long rows would not pay for a ticket per vertex).
    const uint64_t step = (uint64_t)gridDim.x * kTcWaves * kWave;
    for (uint64_t base = ((uint64_t)blockIdx.x * kTcWaves + wave) * kWave; base < n; base += step) {
        const uint64_t v = base + lane;
        const uint32_t dv = v < n ? off[v + 1] - off[v] : 0u;
        uint64_t todo = __ballot(dv >= 2 && dv > min_da);
        while (todo) {
        const uint32_t a = (uint32_t)base + (uint32_t)__builtin_ctzll(todo);
        todo &= todo - 1;
        const uint32_t ab = off[a], da = off[a + 1] - ab;
        const uint32_t *A = adj + ab;
        const bool lds = da <= (uint32_t)kTcLds;
        if (lds)
            for (uint32_t k = lane; k < da; k += kWave) {
                sA[wave][k] = A[k];
                sW[wave][k] = 0u;
            }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint32_t *As = lds ? sA[wave] : A;
        unsigned long long acc_a = 0;
        for (uint32_t bc = 0; bc + 1 < da; bc += kWave) {
            const uint32_t k = bc + lane;
            const bool valid = k < da;
            uint32_t b = valid ? As[k] : 0u;
            uint32_t bb = valid ? off[b] : 0u;
            uint32_t lb = valid ? off[b + 1] - bb : 0u;
            uint32_t pre = lb;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                uint32_t t = __shfl_up(pre, o, kWave);
                if (lane >= o) pre += t;
            }
            const uint32_t tot = __shfl(pre, kWave - 1, kWave);
            sPre[wave][lane] = pre;
            sBeg[wave][lane] = bb;
            sCnt[wave][lane] = 0;
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (uint32_t p = lane; p < tot; p += kWave) {
                uint32_t lo = 0, hi = kWave - 1;
                while (lo < hi) {
                    uint32_t mid = (lo + hi) >> 1;
                    if (sPre[wave][mid] > p)
                        hi = mid;
                    else
                        lo = mid + 1;
                }
                const uint32_t j = lo;
                const uint32_t excl = j ? sPre[wave][j - 1] : 0u;
                const uint32_t w = adj[sBeg[wave][j] + (p - excl)];
                // w > b_j > a; N+(a) is ascending, b_j sits at index bc + j
                const int kw = bsearch_u32(As, bc + j + 1, da, w);
                if (kw >= 0) {
                    if (lds)
                        atomicAdd(&sW[wave][kw], 1u);
                    else
                        atomicAdd(&cnt[w], 1ull);
                    atomicAdd(&sCnt[wave][j], 1u);
                }
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const uint32_t c = valid ? sCnt[wave][lane] : 0u;
            if (c) atomicAdd(&cnt[b], (unsigned long long)c);
            acc_a += c;
            __builtin_amdgcn_wave_barrier();
        }
        if (lds) {  // one global atomic per w of N+(a) that closed a triangle
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            for (uint32_t k = lane; k < da; k += kWave) {
                const uint32_t cw = sW[wave][k];
                if (cw) atomicAdd(&cnt[sA[wave][k]], (unsigned long long)cw);
            }
        }
        acc_a = wave_sum(acc_a);
        if (lane == 0 && acc_a) {
            atomicAdd(&cnt[a], acc_a);
            atomicAdd(total, acc_a);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        }
    }
}

// Hash form, for |N+(a)| <= kTcHashMax (every vertex of a degree-ordered
// graph with up to ~500K edges, and nearly all of larger ones).  Vertices are
// taken from the highest rank down: the work per vertex grows with its rank,
// so the longest ones start first.  One wavefront per vertex a puts N+(a) into
// an LDS hash table (open addressing, linear probing, at most half full, a
// quarter when |N+(a)| <= kTcHashSlots / 4), then walks b in N+(a) one at a
// time and looks up every w in N+(b) lane-wide, kTcInflight w per lane in
// flight.  A hit is the triangle a < b < w.  Counts: w's and b's credits go to
// their slots (16-bit halves of LDS words: at most 2|N+(a)| per slot) and are
// added to the global counts once per vertex a.
constexpr int kTcHashWaves = 4;
constexpr uint32_t kTcHashSlots = 2048;
constexpr uint32_t kTcHashMax = kTcHashSlots / 2;
constexpr uint32_t kTcEmpty = 0xffffffffu;
// Vertex tickets come from kTcStripes counters, each in its own 128-byte line:
// stripe k hands out the ranks n-1-k, n-1-k-kTcStripes, ...  Atomics on one
// address complete one after another (about 10 ns each here), so a single
// counter would cap the kernel at ~10^8 vertices per second.
constexpr uint32_t kTcStripes = 32;
constexpr int kTcInflight = 8;  // w loads per lane in flight (one memory round trip per 512 w)
constexpr uint32_t kTcStripeGap = 32;

GCHI_DEV uint32_t tc_slot(uint32_t w, int sh) { return (w * 2654435761u) >> sh; }

GCHI_DEV int tc_find(const uint32_t *keys, uint32_t mask, int sh, uint32_t w) {
    uint32_t h = tc_slot(w, sh);
    for (;;) {
        const uint32_t k = keys[h];
        if (k == w) return (int)h;
        if (k == kTcEmpty) return -1;
        h = (h + 1) & mask;
    }
}

GCHI_DEV void tc_credit(uint32_t *cw, uint32_t slot, uint32_t c) { atomicAdd(&cw[slot >> 1], c << ((slot & 1) * 16)); }

// Entry `lane` of N+(b_j) (b_j, its start and length in lane j of b/bb/lb), or
// kTcEmpty past the list; the load is issued at an in-bounds address either way.
GCHI_DEV uint32_t tc_first_w(const uint32_t *adj, uint32_t bb, uint32_t lb, uint32_t j, int lane) {
    const uint32_t bbj = (uint32_t)__builtin_amdgcn_readlane((int)bb, (int)j);
    const uint32_t lbj = (uint32_t)__builtin_amdgcn_readlane((int)lb, (int)j);
    const uint32_t v = adj[lbj ? bbj + min((uint32_t)lane, lbj - 1) : 0u];
    return (uint32_t)lane < lbj ? v : kTcEmpty;
}

// Triangles (a, b_j, w) for w in N+(b_j): w0 holds the list's first 64
// entries, the rest are loaded here; each w found in the table is credited,
// and b_j once with the wavefront's total.  Returns that total.
GCHI_DEV uint32_t tc_count_b(const uint32_t *adj, const uint32_t *keys, uint32_t *cw, uint32_t mask, int sh,
                             uint32_t b, uint32_t bb, uint32_t lb, uint32_t j, uint32_t w0, int lane) {
    const uint32_t bj = (uint32_t)__builtin_amdgcn_readlane((int)b, (int)j);
    const uint32_t bbj = (uint32_t)__builtin_amdgcn_readlane((int)bb, (int)j);
    const uint32_t lbj = (uint32_t)__builtin_amdgcn_readlane((int)lb, (int)j);
    const uint32_t *B = adj + bbj;
    uint32_t hits = 0;
    if (w0 != kTcEmpty) {
        const int h = tc_find(keys, mask, sh, w0);
        if (h >= 0) {
            tc_credit(cw, (uint32_t)h, 1u);
            hits++;
        }
    }
    for (uint32_t t = kWave + lane; t < lbj; t += kTcInflight * kWave) {
        uint32_t w[kTcInflight];
#pragma unroll
        for (int q = 0; q < kTcInflight; q++) {
            const uint32_t tq = t + q * kWave;
            w[q] = tq < lbj ? B[tq] : kTcEmpty;
        }
#pragma unroll
        for (int q = 0; q < kTcInflight; q++) {
            if (w[q] == kTcEmpty) continue;
            const int h = tc_find(keys, mask, sh, w[q]);
            if (h >= 0) {
                tc_credit(cw, (uint32_t)h, 1u);
                hits++;
            }
        }
    }
    hits = wave_sum(hits);
    if (hits && lane == 0) tc_credit(cw, (uint32_t)tc_find(keys, mask, sh, bj), hits);
    return hits;
}

__global__ __launch_bounds__(kTcHashWaves *kWave) void tc_count_hash_kernel(const uint32_t *__restrict__ off,
                                                                            const uint32_t *__restrict__ adj,
                                                                            uint32_t n, unsigned long long *cnt,
                                                                            unsigned long long *total,
                                                                            uint32_t *work) {
    __shared__ uint32_t sKey[kTcHashWaves][kTcHashSlots];
    __shared__ uint32_t sCw[kTcHashWaves][kTcHashSlots / 2];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    uint32_t *keys = sKey[wave];
    uint32_t *cw = sCw[wave];
    unsigned long long acc_total = 0;
    uint32_t stripe = (blockIdx.x * kTcHashWaves + wave) % kTcStripes, spent = 0;
    for (;;) {
        uint32_t t = 0;
        if (lane == 0) t = atomicAdd(&work[stripe * kTcStripeGap], 1u);
        t = __shfl(t, 0, kWave);
        const uint64_t idx = stripe + (uint64_t)kTcStripes * t;
        if (idx >= n) {  // this stripe is drained: move on, stop after all of them
            if (++spent == kTcStripes) break;
            stripe = (stripe + 1) % kTcStripes;
            continue;
        }
        const uint32_t a = n - 1 - (uint32_t)idx;  // highest rank (most work) first
        const uint32_t ab = off[a], da = off[a + 1] - ab;
        if (da < 2 || da > kTcHashMax) continue;
        const uint32_t want = 4 * da;  // slots: pow2 >= 4 da, at most kTcHashSlots
        const int lg = min(32 - __clz((int)(want - 1)), 11);
        const uint32_t S = 1u << lg, mask = S - 1;
        const int sh = 32 - lg;
        for (uint32_t k = lane; k < S; k += kWave) keys[k] = kTcEmpty;
        for (uint32_t k = lane; k < S / 2; k += kWave) cw[k] = 0u;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        const uint32_t *A = adj + ab;
        for (uint32_t k = lane; k < da; k += kWave) {
            const uint32_t x = A[k];
            uint32_t h = tc_slot(x, sh);
            while (atomicCAS(&keys[h], kTcEmpty, x) != kTcEmpty) h = (h + 1) & mask;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        uint32_t acc_a = 0;
        // the last element of N+(a) closes no triangle here (no w above it)
        for (uint32_t bc = 0; bc + 1 < da; bc += kWave) {
            const uint32_t k = bc + lane;
            const bool valid = k + 1 < da;
            const uint32_t b = valid ? A[k] : 0u;
            const uint32_t bb = valid ? off[b] : 0u;
            const uint32_t lb = valid ? off[b + 1] - bb : 0u;
            const uint32_t nb = min((uint32_t)kWave, da - 1 - bc);
            // The first 64 entries of N+(b_j) are loaded one b ahead (slots w2[0],
            // w2[1] in turn), so that b_j's probes overlap b_{j+1}'s loads.
            uint32_t w2[2];
            w2[0] = tc_first_w(adj, bb, lb, 0, lane);
            for (uint32_t j0 = 0; j0 < nb; j0 += 2) {  // loads unconditional (clamped j), uses not
                w2[1] = tc_first_w(adj, bb, lb, min(j0 + 1, nb - 1), lane);
                acc_a += tc_count_b(adj, keys, cw, mask, sh, b, bb, lb, j0, w2[0], lane);
                w2[0] = tc_first_w(adj, bb, lb, min(j0 + 2, nb - 1), lane);
                if (j0 + 1 < nb) acc_a += tc_count_b(adj, keys, cw, mask, sh, b, bb, lb, j0 + 1, w2[1], lane);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        for (uint32_t k = lane; k < S; k += kWave) {
            const uint32_t key = keys[k];
            const uint32_t c = (cw[k >> 1] >> ((k & 1) * 16)) & 0xffffu;
            if (key != kTcEmpty && c) atomicAdd(&cnt[key], (unsigned long long)c);
        }
        if (lane == 0 && acc_a) atomicAdd(&cnt[a], (unsigned long long)acc_a);
        acc_total += acc_a;
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    if (lane == 0 && acc_total) atomicAdd(total, acc_total);
}

__global__ void tc_to_u32(const unsigned long long *cnt, uint32_t n, uint32_t *out, uint32_t *overflow) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        unsigned long long c = cnt[v];
        if (c > 0xffffffffull) atomicOr(overflow, 1u);
        out[v] = (uint32_t)c;
    }
}

template <typename C>
static void tc_prep_layout(C &c, uint64_t E, uint32_t N, uint32_t **deg, uint64_t **vkeys, uint64_t **ekeys,
                           uint32_t **flags, uint32_t **pos, uint32_t **err, void **sw, size_t *swb) {
    int idb = bits_for(N ? N - 1 : 0);
    if (idb == 0) idb = 1;
    *deg = c.template take<uint32_t>(N + 1);
    *vkeys = c.template take<uint64_t>(N + 1);
    *ekeys = c.template take<uint64_t>(E + 1);
    *flags = c.template take<uint32_t>(E + 1);
    *pos = c.template take<uint32_t>(E + 2);
    *err = c.template take<uint32_t>(1);
    size_t s = sort_workspace_bytes(std::max<uint64_t>(E, N), 64, false);
    s = std::max(s, scan_workspace(std::max<uint64_t>(E, N) + 1));
    *sw = c.template take<char>(s);
    *swb = s;
}

// deg[v] += in + out degree of v without self-loops (deg zeroed by the caller),
// from runs of equal ids: the src passes order the keys by src, the dst passes
// by dst.  vkeys (N+1 entries) serves as run-start scratch, ekeys and sw as the
// sort buffers.
static int tc_degree_counts(const uint32_t *src, const uint32_t *dst, uint64_t E, uint32_t N, uint32_t *deg,
                            uint32_t *err, uint64_t *vkeys, uint64_t *ekeys, void *sw, size_t swb, hipStream_t s) {
    if (!E) return GCHI_OK;
    const int fb = bits_for(N);  // wide enough for n, the skip mark
    const int sb = (fb + 7) / 8 * 8;
    const uint32_t fmask = fb >= 32 ? 0xffffffffu : ((1u << fb) - 1u);
    uint32_t *run_start = reinterpret_cast<uint32_t *>(vkeys);
    hipLaunchKernelGGL(tc_degree_pack, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, src, dst, E, N, sb, ekeys, err);
    GCHI_LAUNCH_CHECK();
    uint64_t *by_src = nullptr, *by_dst = nullptr;
    GCHI_TRY(radix_sort_u64_bits(ekeys, nullptr, nullptr, E, 0, sb, sw, swb, s, &by_src, err));
    GCHI_TRY(count_field_runs(by_src, E, 0, fmask, N, run_start, deg, nullptr, 0u, s, true));
    GCHI_TRY(radix_sort_u64_bits(by_src, by_src == ekeys ? nullptr : ekeys, nullptr, E, sb, sb + fb, sw, swb, s,
                                 &by_dst, err));
    return count_field_runs(by_dst, E, sb, fmask, N, run_start, deg, nullptr, 0u, s, true);
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_tc_prepare_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    SizerAdapter a;
    uint32_t *d, *f, *p, *e;
    uint64_t *vk, *ek;
    void *sw;
    size_t swb;
    tc_prep_layout(a, nedges, nvertices, &d, &vk, &ek, &f, &p, &e, &sw, &swb);
    *bytes = a.z.off;
    return GCHI_OK;
}

extern "C" int gchi_tc_prepare(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                               uint32_t *d_vertexmap, uint32_t *d_off, uint32_t *d_adj, uint64_t *h_noriented,
                               void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK(d_vertexmap && d_off && d_adj && h_noriented && nvertices > 0, GCHI_ERR_INVALID,
               "tc_prepare: null argument");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "tc_prepare: null edges");
    GCHI_CHECK(nedges < (1ull << 31), GCHI_ERR_UNSUPPORTED, "tc_prepare: E must be < 2^31");
    hipStream_t s = as_stream(stream);
    Carver c(d_work, work_bytes);
    uint32_t *deg, *flags, *pos, *err;
    uint64_t *vkeys, *ekeys;
    void *sw;
    size_t swb;
    tc_prep_layout(c, nedges, nvertices, &deg, &vkeys, &ekeys, &flags, &pos, &err, &sw, &swb);
    GCHI_CHECK(c.ok(), GCHI_ERR_WORKSPACE, "tc_prepare: workspace too small");
    const uint32_t N = nvertices;
    const uint64_t E = nedges;
    int idb = bits_for(N - 1);
    if (idb == 0) idb = 1;
    const int degb = bits_for(2 * E);
    GCHI_CHECK(idb + degb <= 64, GCHI_ERR_UNSUPPORTED, "tc_prepare: degree/id keys exceed 64 bits");
    launch_fill_u32(deg, 0u, N, s);
    launch_fill_u32(err, 0u, 1, s);
    GCHI_TRY(tc_degree_counts(d_src, d_dst, E, N, deg, err, vkeys, ekeys, sw, swb, s));
    hipLaunchKernelGGL(tc_degree_keys, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, deg, N, idb, vkeys);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(radix_sort_u64(vkeys, nullptr, N, idb + degb, sw, swb, s, err));
    hipLaunchKernelGGL(tc_vertexmap, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, vkeys, N, idb, d_vertexmap);
    const uint64_t sentinel = (2 * idb >= 64) ? ~0ull : ((1ull << (2 * idb)) - 1ull);
    uint32_t *lo_count = deg;  // reuse: degrees are no longer needed
    launch_fill_u32(lo_count, 0u, N + 1, s);
    if (E) {
        hipLaunchKernelGGL(tc_orient_keys, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, d_src, d_dst, E, N,
                           d_vertexmap, idb, sentinel, ekeys);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(radix_sort_u64(ekeys, nullptr, E, 2 * idb, sw, swb, s, err));
        hipLaunchKernelGGL(tc_unique_flags, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, ekeys, E, sentinel,
                           flags);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(exclusive_scan_u32(flags, pos, E, sw, swb, s));
        uint32_t *lo_start = reinterpret_cast<uint32_t *>(vkeys);  // vertexmap is built: vkeys is free
        hipLaunchKernelGGL(tc_lo_run_starts, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, ekeys, pos, E, idb,
                           sentinel, N, lo_start);
        hipLaunchKernelGGL(tc_scatter_unique, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, ekeys, flags, pos, E,
                           idb, sentinel, N, lo_start, d_adj, lo_count);
        GCHI_LAUNCH_CHECK();
    }
    GCHI_TRY(exclusive_scan_u32(lo_count, d_off, N, sw, swb, s));
    uint32_t h_err = 0, h_m = 0;
    GCHI_HIP(hipMemcpyAsync(&h_err, err, 4, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipMemcpyAsync(&h_m, d_off + N, 4, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    GCHI_CHECK(!(h_err & GCHI_WORK_TIMEOUT), GCHI_ERR_TIMEOUT, "tc_prepare: a sort look-back wait gave up");
    GCHI_CHECK(h_err == 0, GCHI_ERR_INVALID, "tc_prepare: vertex id >= nvertices");
    *h_noriented = h_m;
    return GCHI_OK;
}

namespace gchi {

// The counting launches, over counters cnt[n+1] / tot[1] / ticket stripes wk
// (all cleared here first).
static int tc_count_impl(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t *d_per_vertex,
                         uint64_t *d_total, uint32_t *d_overflow, unsigned long long *cnt, unsigned long long *tot,
                         uint32_t *wk, hipStream_t s) {
    launch_fill_u64(reinterpret_cast<uint64_t *>(cnt), 0ull, nvertices, s);
    launch_fill_u64(reinterpret_cast<uint64_t *>(tot), 0ull, 1, s);
    launch_fill_u32(wk, 0u, kTcStripes * kTcStripeGap, s);
    launch_fill_u32(d_overflow, 0u, 1, s);
    int ncu = 256;
    int dev = 0;
    if (hipGetDevice(&dev) == hipSuccess) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0) ncu = v;
    }
    hipLaunchKernelGGL(tc_count_hash_kernel, dim3(ncu * 6), dim3(kTcHashWaves * kWave), 0, s, d_off, d_adj,
                       nvertices, cnt, tot, wk);
    hipLaunchKernelGGL(tc_count_kernel, dim3(ncu * 8), dim3(kTcWaves * kWave), 0, s, d_off, d_adj, nvertices,
                       kTcHashMax, cnt, tot);
    hipLaunchKernelGGL(tc_to_u32, dim3(grid_for(nvertices, 256, 65536)), dim3(256), 0, s, cnt, nvertices,
                       d_per_vertex, d_overflow);
    GCHI_HIP(hipMemcpyAsync(d_total, tot, 8, hipMemcpyDeviceToDevice, s));
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

template <typename C>
static void tc_count_layout(C &c, uint32_t nvertices, unsigned long long **cnt, unsigned long long **tot,
                            uint32_t **wk) {
    *cnt = c.template take<unsigned long long>((size_t)nvertices + 1);
    *tot = c.template take<unsigned long long>(1);
    *wk = c.template take<uint32_t>(kTcStripes * kTcStripeGap);
}

// ---- allocation-free counting: scratch in a file-scope device array ----
// Counters for up to GCHI_TC_NOALLOC_MAX_VERTICES vertices, laid out as
// tc_count_layout carves them (256-byte aligned pieces).
constexpr size_t kTcScratchCounts = (((size_t)GCHI_TC_NOALLOC_MAX_VERTICES + 1) * 8 + 255) / 256 * 256;
constexpr size_t kTcScratchBytes = kTcScratchCounts + 256 + (size_t)kTcStripes * kTcStripeGap * 4;
static __device__ __align__(256) unsigned char scratch[kTcScratchBytes];

// Device address of each device's copy of scratch (the array's name in host
// code is not a device address), taken once per device with that device
// current.
constexpr int kTcMaxDevices = 64;
static void *g_scratch_ptr[kTcMaxDevices];
static std::mutex g_scratch_mu;

int tc_scratch_setup(int dev) {
    GCHI_CHECK(dev >= 0 && dev < kTcMaxDevices, GCHI_ERR_UNSUPPORTED, "tc scratch: device index >= 64");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return GCHI_OK;
    void *ptr = nullptr;
    GCHI_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    GCHI_CHECK(ptr, GCHI_ERR_HIP, "tc scratch: hipGetSymbolAddress returned null");
    g_scratch_ptr[dev] = ptr;
    return GCHI_OK;
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_tc_count_workspace(uint32_t nvertices, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    SizerAdapter a;
    unsigned long long *cnt, *tot;
    uint32_t *wk;
    tc_count_layout(a, nvertices, &cnt, &tot, &wk);
    *bytes = a.z.off;
    return GCHI_OK;
}

extern "C" int gchi_tc_count(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices,
                             uint32_t *d_per_vertex, uint64_t *d_total, uint32_t *d_overflow, void *d_work,
                             size_t work_bytes, void *stream) {
    GCHI_CHECK(d_off && d_adj && d_per_vertex && d_total && d_overflow && nvertices > 0, GCHI_ERR_INVALID,
               "tc_count: null argument");
    size_t need = 0;
    gchi_tc_count_workspace(nvertices, &need);
    GCHI_CHECK(d_work && work_bytes >= need, GCHI_ERR_WORKSPACE, "tc_count: workspace too small");
    Carver c(d_work, work_bytes);
    unsigned long long *cnt, *tot;
    uint32_t *wk;
    tc_count_layout(c, nvertices, &cnt, &tot, &wk);
    return tc_count_impl(d_off, d_adj, nvertices, d_per_vertex, d_total, d_overflow, cnt, tot, wk,
                         as_stream(stream));
}

extern "C" int gchi_tc_count_noalloc(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices,
                                     uint32_t *d_per_vertex, uint64_t *d_total, uint32_t *d_overflow, void *stream) {
    GCHI_CHECK(d_off && d_adj && d_per_vertex && d_total && d_overflow && nvertices > 0, GCHI_ERR_INVALID,
               "tc_count_noalloc: null argument");
    GCHI_CHECK(nvertices <= GCHI_TC_NOALLOC_MAX_VERTICES, GCHI_ERR_INVALID,
               "tc_count_noalloc: nvertices exceeds GCHI_TC_NOALLOC_MAX_VERTICES (4,194,304); use gchi_tc_count "
               "with a workspace");
    int dev = 0;
    GCHI_HIP(hipGetDevice(&dev));
    GCHI_TRY(tc_scratch_setup(dev));
    Carver c(g_scratch_ptr[dev], kTcScratchBytes);
    unsigned long long *cnt, *tot;
    uint32_t *wk;
    tc_count_layout(c, nvertices, &cnt, &tot, &wk);
    GCHI_CHECK(c.ok(), GCHI_ERR_INVALID, "tc_count_noalloc: scratch layout exceeds its array");
    return tc_count_impl(d_off, d_adj, nvertices, d_per_vertex, d_total, d_overflow, cnt, tot, wk,
                         as_stream(stream));
}

__global__ void relabel_kernel(const uint32_t *src, const uint32_t *dst, uint64_t e, uint32_t n, const uint32_t *map,
                               uint32_t *so, uint32_t *dout, uint32_t *err) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t s = src[i], d = dst[i];
        if (s >= n || d >= n) {
            atomicOr(err, 1u);
            so[i] = dout[i] = 0;
            continue;
        }
        so[i] = map[s];
        dout[i] = map[d];
    }
}

extern "C" int gchi_degree_order_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes) {
    return gchi_tc_prepare_workspace(nedges, nvertices, bytes);
}

// order_by_degree's translation table only (orderbydegree.hpp:64-100)
extern "C" int gchi_degree_order(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                                 uint32_t *d_vertexmap, void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK(d_vertexmap && nvertices > 0, GCHI_ERR_INVALID, "degree_order: null argument");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "degree_order: null edges");
    GCHI_CHECK(nedges < (1ull << 31), GCHI_ERR_UNSUPPORTED, "degree_order: E must be < 2^31");
    hipStream_t s = as_stream(stream);
    Carver c(d_work, work_bytes);
    uint32_t *deg, *flags, *pos, *err;
    uint64_t *vkeys, *ekeys;
    void *sw;
    size_t swb;
    tc_prep_layout(c, nedges, nvertices, &deg, &vkeys, &ekeys, &flags, &pos, &err, &sw, &swb);
    GCHI_CHECK(c.ok(), GCHI_ERR_WORKSPACE, "degree_order: workspace too small");
    const uint32_t N = nvertices;
    int idb = bits_for(N - 1);
    if (idb == 0) idb = 1;
    const int degb = bits_for(2 * nedges);
    launch_fill_u32(deg, 0u, N, s);
    launch_fill_u32(err, 0u, 1, s);
    GCHI_TRY(tc_degree_counts(d_src, d_dst, nedges, N, deg, err, vkeys, ekeys, sw, swb, s));
    hipLaunchKernelGGL(tc_degree_keys, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, deg, N, idb, vkeys);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(radix_sort_u64(vkeys, nullptr, N, idb + degb, sw, swb, s, err));
    hipLaunchKernelGGL(tc_vertexmap, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, vkeys, N, idb, d_vertexmap);
    uint32_t h_err = 0;
    GCHI_HIP(hipMemcpyAsync(&h_err, err, 4, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    GCHI_CHECK(!(h_err & GCHI_WORK_TIMEOUT), GCHI_ERR_TIMEOUT, "degree_order: a sort look-back wait gave up");
    GCHI_CHECK(h_err == 0, GCHI_ERR_INVALID, "degree_order: vertex id >= nvertices");
    return GCHI_OK;
}

extern "C" int gchi_relabel_edges(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                                  const uint32_t *d_map, uint32_t *d_src_out, uint32_t *d_dst_out, void *stream) {
    GCHI_CHECK(d_map && (nedges == 0 || (d_src && d_dst && d_src_out && d_dst_out)), GCHI_ERR_INVALID,
               "relabel_edges: null argument");
    if (!nedges) return GCHI_OK;
    hipStream_t s = as_stream(stream);
    uint32_t *err = nullptr;
    GCHI_HIP(hipMalloc(&err, 4));
    launch_fill_u32(err, 0u, 1, s);
    hipLaunchKernelGGL(relabel_kernel, dim3(grid_for(nedges, 256, 65536)), dim3(256), 0, s, d_src, d_dst, nedges,
                       nvertices, d_map, d_src_out, d_dst_out, err);
    uint32_t h = 0;
    hipError_t e = hipMemcpyAsync(&h, err, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(err);
    if (e != hipSuccess) return hip_fail(e, "relabel_edges", __FILE__, __LINE__);
    GCHI_CHECK(h == 0, GCHI_ERR_INVALID, "relabel_edges: vertex id >= nvertices");
    return GCHI_OK;
}
