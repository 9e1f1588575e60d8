// Batched personalized PageRank: many restart sets swept together.
//
// Kernel per query (the functional kernel with a seed-set restart, see
// example_apps/pagerank_functional.cpp:72-75 for the apply form):
//   r_v = (v in S ? (1 - d)/|S| : 0) + d * sum_{u -> v} r_u / outdeg_u
// iterated from r = teleport vector.  Every term is non-negative, so the
// iterates increase monotonically to the fixed point r*, and with
// res_k = ||r_k - r_{k-1}||_1 the remaining error is at most
// res_k * d / (1 - d); a query is converged when that bound is <= tol.
//
// Layout: 64 queries form a block; ranks are stored vertex-major, 64 values
// per vertex (128 B in BF16, 256 B in FP32), so a wavefront owning one
// destination row (lane = query) turns every in-edge into one coalesced row
// gather shared by the 64 queries; the in-edge index stream is read once per
// query block instead of once per query.  Rows come from the PageRank plan
// (regular blocks of < 2W edges, hub rows cut into 512-edge chunks whose
// 64-wide partials are combined in chunk order).  A persistent kernel pulls
// (query block, tile) items, query-block-major, from kPpStripes device work
// counters (stripe s hands out items s, s + kPpStripes, ...; each counter in
// its own 128-byte line, since atomics on one address complete one after
// another).
//
// Precision: phase 1 stores ranks in BF16 with FP32 accumulation; phase 2
// converts to FP32 and refines with FP32 sweeps until every query meets the
// tolerance.  The per-query L1 residuals are FP64 sums.
#include <type_traits>

#include "common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace gchi {
// plan internals shared with pagerank.hip
struct PlanView {
    uint32_t N, nblk, nhub, nchunk, npersist;
    const uint32_t *reg_off, *reg_src, *blk_row, *hub_bits, *hub_v, *hub_chunk_off;
    const uint64_t *chunk_beg;
    const uint32_t *chunk_len, *hub_src, *outdeg;
};
int pr_plan_view(const gchi_pr_plan *p, PlanView *v);
}  // namespace gchi

struct gchi_ppr_batch {
    gchi::PlanView pv;
    uint32_t nq = 0, nqb = 0;
    float *inv_outdeg = nullptr;   // [N]
    void *r16[2] = {nullptr, nullptr};  // [nqb][N][64] bf16
    float *r32[2] = {nullptr, nullptr};  // [nqb][N][64] fp32
    float *hub_part = nullptr;     // [nqb][nchunk][64]
    double *res = nullptr;         // [nqb*64]
    uint32_t *seed_bits = nullptr; // [nqb][words]
    uint32_t *seed_v = nullptr;    // [nqb][64*64] sorted (vertex, query)
    uint32_t *seed_q = nullptr;
    uint32_t *seed_n = nullptr;    // [nqb]
    float *tele = nullptr;         // [nqb*64] restart mass per seed
    uint32_t *counter = nullptr;
    uint32_t *active = nullptr;    // [nqb] query blocks still swept, then their count at [nqb]
    size_t words = 0;
    int cur = 0;                   // buffer holding the current ranks
    int ncu = 256;
    bool fp32 = false;
    void *mem = nullptr;
    // chunks of sweeps captured once per phase and replayed (the residual is
    // read back only after each chunk)
    hipGraphExec_t chunk_exec[2] = {nullptr, nullptr};  // [0] BF16, [1] FP32
    hipGraph_t chunk_graph[2] = {nullptr, nullptr};
    int chunk_len[2] = {0, 0};
    float chunk_damping[2] = {0.0f, 0.0f};
};

namespace gchi {

constexpr int kPpThreads = 256;
constexpr int kPpWaves = kPpThreads / kWave;
constexpr int kQ = 64;                 // queries per block (one per lane)
constexpr uint32_t kPpStripes = 16;    // work counters of the persistent sweep
constexpr uint32_t kPpStripeGap = 32;  // uint32 words between counters
constexpr int kMaxSeeds = 64;          // restart-set size limit per query
constexpr int kSeedTable = kQ * kMaxSeeds;

GCHI_DEV float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
GCHI_DEV uint16_t f32_to_bf16(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename T>
struct Store;
template <>
struct Store<uint16_t> {
    static GCHI_DEV float ld(const uint16_t *p) { return bf16_to_f32(*p); }
    static GCHI_DEV float round(float f) { return bf16_to_f32(f32_to_bf16(f)); }
    static GCHI_DEV void st(uint16_t *p, float f) { *p = f32_to_bf16(f); }
};
template <>
struct Store<float> {
    static GCHI_DEV float ld(const float *p) { return *p; }
    static GCHI_DEV float round(float f) { return f; }
    static GCHI_DEV void st(float *p, float f) { *p = f; }
};

struct PpArgs {
    PlanView pv;
    const float *inv;
    const uint32_t *seed_bits, *seed_v, *seed_q, *seed_n;
    const float *tele;
    float *hub_part;
    double *res;
    uint32_t *counter;
    const uint32_t *active;  // query blocks still swept; active[nqb] = their count
    size_t words;
    uint32_t nqb;
    float damping;
    int residual;
};

// restart mass of (v, lane) in query block qb
GCHI_DEV float seed_mass(const PpArgs &a, uint32_t qb, uint32_t v, int lane) {
    const uint32_t *bits = a.seed_bits + qb * a.words;
    if (!((bits[v >> 5] >> (v & 31)) & 1u)) return 0.0f;
    const uint32_t *sv = a.seed_v + (size_t)qb * kSeedTable;
    const uint32_t *sq = a.seed_q + (size_t)qb * kSeedTable;
    uint32_t lo = 0, hi = a.seed_n[qb];
    while (lo < hi) {
        uint32_t mid = (lo + hi) >> 1;
        if (sv[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    float m = 0.0f;
    for (uint32_t k = lo; k < a.seed_n[qb] && sv[k] == v; k++)
        if (sq[k] == (uint32_t)lane) m = a.tele[qb * kQ + lane];
    return m;
}

// sum over in-edges [beg, end) of r_u * inv_u for this lane's query
template <typename T>
GCHI_DEV float gather_rows_1(const uint32_t *src_idx, uint32_t beg, uint32_t end, const float *inv, const T *R,
                           int lane) {
    float acc = 0.0f;
    for (uint32_t e = beg; e < end; e += kWave) {
        const uint32_t n = min((uint32_t)kWave, end - e);
        const uint32_t my = lane < (int)n ? __builtin_nontemporal_load(&src_idx[e + lane]) : 0u;
        const float myi = lane < (int)n ? inv[my] : 0.0f;
        uint32_t j = 0;
        // 8 row gathers in flight per wavefront; the additions stay in edge order
        for (; j + 8 <= n; j += 8) {
            float v[8], w[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t sq = __builtin_amdgcn_readlane(my, j + q);
                w[q] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j + q));
                v[q] = Store<T>::ld(R + (size_t)sq * kQ + lane);
            }
#pragma unroll
            for (int q = 0; q < 8; q++) acc += v[q] * w[q];
        }
        for (; j + 4 <= n; j += 4) {
            uint32_t s0 = __builtin_amdgcn_readlane(my, j), s1 = __builtin_amdgcn_readlane(my, j + 1);
            uint32_t s2 = __builtin_amdgcn_readlane(my, j + 2), s3 = __builtin_amdgcn_readlane(my, j + 3);
            float i0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j));
            float i1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j + 1));
            float i2 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j + 2));
            float i3 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j + 3));
            float v0 = Store<T>::ld(R + (size_t)s0 * kQ + lane);
            float v1 = Store<T>::ld(R + (size_t)s1 * kQ + lane);
            float v2 = Store<T>::ld(R + (size_t)s2 * kQ + lane);
            float v3 = Store<T>::ld(R + (size_t)s3 * kQ + lane);
            acc += v0 * i0;
            acc += v1 * i1;
            acc += v2 * i2;
            acc += v3 * i3;
        }
        for (; j < n; j++) {
            uint32_t s0 = __builtin_amdgcn_readlane(my, j);
            float i0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j));
            acc += Store<T>::ld(R + (size_t)s0 * kQ + lane) * i0;
        }
    }
    return acc;
}

// BF16 rows: a 64-query row is 128 bytes, so one load instruction can carry
// two edges.  Lanes 0-31 take the even edges of the row (counted from beg),
// lanes 32-63 the odd ones; lane 32h + p loads queries 2p and 2p+1 of its
// edge's source row as one 32-bit word.  Each query's sum is its even-edge
// sum plus its odd-edge sum (both in edge order), combined at the end and
// moved to lane q for query q.
GCHI_DEV float bf16_lo(uint32_t v) { return __uint_as_float(v << 16); }
GCHI_DEV float bf16_hi(uint32_t v) { return __uint_as_float(v & 0xffff0000u); }

GCHI_DEV float gather_rows_bf16x2(const uint32_t *src_idx, uint32_t beg, uint32_t end, const float *inv,
                                  const uint16_t *R, int lane) {
    const bool odd = lane >= 32;
    const int p = lane & 31;
    const uint32_t *R32 = reinterpret_cast<const uint32_t *>(R);  // row s: words s*32 .. s*32+31
    float a0 = 0.0f, a1 = 0.0f;
    for (uint32_t e = beg; e < end; e += kWave) {
        const uint32_t n = min((uint32_t)kWave, end - e);
        const uint32_t my = lane < (int)n ? __builtin_nontemporal_load(&src_idx[e + lane]) : 0u;
        const float myi = lane < (int)n ? inv[my] : 0.0f;
        uint32_t j = 0;
        for (; j + 16 <= n; j += 16) {  // 8 edge pairs in flight
            uint32_t v[8];
            float w[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t s0 = __builtin_amdgcn_readlane(my, j + 2 * q);
                const uint32_t s1 = __builtin_amdgcn_readlane(my, j + 2 * q + 1);
                const float w0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j + 2 * q));
                const float w1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j + 2 * q + 1));
                w[q] = odd ? w1 : w0;
                v[q] = R32[(size_t)(odd ? s1 : s0) * (kQ / 2) + p];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                a0 += bf16_lo(v[q]) * w[q];
                a1 += bf16_hi(v[q]) * w[q];
            }
        }
        for (; j + 2 <= n; j += 2) {
            const uint32_t s0 = __builtin_amdgcn_readlane(my, j), s1 = __builtin_amdgcn_readlane(my, j + 1);
            const float w0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j));
            const float w1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j + 1));
            const float w = odd ? w1 : w0;
            const uint32_t v = R32[(size_t)(odd ? s1 : s0) * (kQ / 2) + p];
            a0 += bf16_lo(v) * w;
            a1 += bf16_hi(v) * w;
        }
        if (j < n) {  // the chunk's last edge is even: the odd half loads it and discards it
            const uint32_t s0 = __builtin_amdgcn_readlane(my, j);
            const float w0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(myi), j));
            const uint32_t v = R32[(size_t)s0 * (kQ / 2) + p];
            if (!odd) {
                a0 += bf16_lo(v) * w0;
                a1 += bf16_hi(v) * w0;
            }
        }
    }
    // query 2p: even part (lane p) + odd part (lane 32 + p), in that order
    const float o0 = __shfl_xor(a0, 32, kWave), o1 = __shfl_xor(a1, 32, kWave);
    const float t0 = odd ? o0 + a0 : a0 + o0;
    const float t1 = odd ? o1 + a1 : a1 + o1;
    const float u0 = __shfl(t0, lane >> 1, kWave), u1 = __shfl(t1, lane >> 1, kWave);
    return (lane & 1) ? u1 : u0;
}

template <typename T>
GCHI_DEV float gather_rows(const uint32_t *src_idx, uint32_t beg, uint32_t end, const float *inv, const T *R,
                           int lane) {
    if constexpr (std::is_same<T, uint16_t>::value)
        return gather_rows_bf16x2(src_idx, beg, end, inv, R, lane);
    else
        return gather_rows_1<T>(src_idx, beg, end, inv, R, lane);
}

template <typename T>
__global__ __launch_bounds__(kPpThreads) void ppr_sweep_kernel(PpArgs a, const T *__restrict__ Rin_all,
                                                              T *__restrict__ Rout_all) {
    __shared__ float s_part[kPpWaves][kQ];
    __shared__ double s_res[kQ];
    __shared__ uint32_t s_tile;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t T_per_qb = a.pv.nchunk + a.pv.nblk;
    const uint32_t nact = a.active[a.nqb];
    const uint32_t ntiles = T_per_qb * nact;
    const size_t stride_q = (size_t)a.pv.N * kQ;
    uint32_t cur_qb = 0xffffffffu;
    double res = 0.0;  // this lane's query, accumulated by this wave for cur_qb
    if (threadIdx.x < kQ) s_res[threadIdx.x] = 0.0;
    uint32_t stripe = blockIdx.x % kPpStripes, spent = 0;
    for (;;) {
        uint32_t t = ntiles;  // the next tile from the stripes; ntiles once all are drained
        while (spent < kPpStripes) {  // block-uniform
            if (threadIdx.x == 0) s_tile = atomicAdd(&a.counter[stripe * kPpStripeGap], 1u);
            __syncthreads();
            const uint64_t t64 = stripe + (uint64_t)kPpStripes * s_tile;
            __syncthreads();  // s_tile is rewritten by the next ticket
            if (t64 < ntiles) {
                t = (uint32_t)t64;
                break;
            }
            spent++;
            stripe = (stripe + 1) % kPpStripes;
        }
        const uint32_t qi = t < ntiles ? t / T_per_qb : 0xffffffffu;  // index into the active list
        const uint32_t qb = t < ntiles ? a.active[qi] : 0xffffffffu;
        if (qb != cur_qb) {  // block-uniform: every thread read the same tile
            if (a.residual && cur_qb != 0xffffffffu) {
                // add the waves' residuals in wave order, then one atomic per query
                for (int w = 0; w < kPpWaves; w++) {
                    if (wave == w) s_res[lane] += res;
                    __syncthreads();
                }
                if (wave == 0) {
                    const double v = s_res[lane];
                    if (v != 0.0) atomicAdd(&a.res[cur_qb * kQ + lane], v);
                    s_res[lane] = 0.0;
                }
                __syncthreads();
            }
            res = 0.0;
            cur_qb = qb;
        }
        if (t >= ntiles) break;
        const uint32_t lt = t - qi * T_per_qb;
        const T *Rin = Rin_all + qb * stride_q;
        T *Rout = Rout_all + qb * stride_q;
        if (lt < a.pv.nchunk) {
            // hub chunk: 4 waves split <= 512 edges, partials combined in wave order
            const uint64_t cb = a.pv.chunk_beg[lt];
            const uint32_t cl = a.pv.chunk_len[lt];
            const uint32_t per = (cl + kPpWaves - 1) / kPpWaves;
            const uint32_t b0 = min(cl, per * wave), b1 = min(cl, per * (wave + 1));
            float acc = gather_rows<T>(a.pv.hub_src + cb, b0, b1, a.inv, Rin, lane);
            s_part[wave][lane] = acc;
            __syncthreads();
            if (wave == 0) {
                float t4 = 0.0f;
                for (int w = 0; w < kPpWaves; w++) t4 += s_part[w][lane];
                a.hub_part[((size_t)qb * a.pv.nchunk + lt) * kQ + lane] = t4;
            }
            __syncthreads();
            continue;
        }
        const uint32_t b = lt - a.pv.nchunk;
        const uint32_t rb = a.pv.blk_row[b], re = a.pv.blk_row[b + 1];
        for (uint32_t v = rb + wave; v < re; v += kPpWaves) {
            if ((a.pv.hub_bits[v >> 5] >> (v & 31)) & 1u) continue;
            float acc = gather_rows<T>(a.pv.reg_src, a.pv.reg_off[v], a.pv.reg_off[v + 1], a.inv, Rin, lane);
            float rn = Store<T>::round(seed_mass(a, qb, v, lane) + a.damping * acc);
            if (a.residual) res += fabs((double)rn - (double)Store<T>::ld(Rin + (size_t)v * kQ + lane));
            Store<T>::st(Rout + (size_t)v * kQ + lane, rn);
        }
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kPpThreads) void ppr_hub_kernel(PpArgs a, const T *__restrict__ Rin_all,
                                                            T *__restrict__ Rout_all) {
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const uint32_t nw = gridDim.x * kPpWaves;
    const size_t stride_q = (size_t)a.pv.N * kQ;
    const uint32_t items = a.pv.nhub * a.active[a.nqb];
    // residuals summed per wavefront while its items stay in one query block,
    // then one atomic per query (not one per hub row and query)
    double racc = 0.0;
    uint32_t rqb = 0xffffffffu;
    for (uint32_t it = blockIdx.x * kPpWaves + wave; it < items; it += nw) {
        const uint32_t qb = a.active[it / a.pv.nhub], h = it % a.pv.nhub;
        if (a.residual && qb != rqb) {  // wavefront-uniform
            if (rqb != 0xffffffffu && racc != 0.0) atomicAdd(&a.res[rqb * kQ + lane], racc);
            racc = 0.0;
            rqb = qb;
        }
        const uint32_t c0 = a.pv.hub_chunk_off[h], c1 = a.pv.hub_chunk_off[h + 1];
        // the chunk partials added in chunk order, 16 loads in flight
        const float *part = a.hub_part + (size_t)qb * a.pv.nchunk * kQ + lane;
        float acc = 0.0f;
        for (uint32_t c = c0; c < c1; c += 16) {
            float pv[16];
#pragma unroll
            for (uint32_t q = 0; q < 16; q++) pv[q] = part[(size_t)min(c + q, c1 - 1) * kQ];
#pragma unroll
            for (uint32_t q = 0; q < 16; q++)
                if (c + q < c1) acc += pv[q];
        }
        const uint32_t v = a.pv.hub_v[h];
        const T *Rin = Rin_all + qb * stride_q;
        T *Rout = Rout_all + qb * stride_q;
        float rn = Store<T>::round(seed_mass(a, qb, v, lane) + a.damping * acc);
        if (a.residual) racc += fabs((double)rn - (double)Store<T>::ld(Rin + (size_t)v * kQ + lane));
        Store<T>::st(Rout + (size_t)v * kQ + lane, rn);
    }
    if (a.residual && rqb != 0xffffffffu && racc != 0.0) atomicAdd(&a.res[rqb * kQ + lane], racc);
    if (blockIdx.x == 0)  // work queue ready for the next sweep
        for (uint32_t i = threadIdx.x; i < kPpStripes; i += blockDim.x) a.counter[i * kPpStripeGap] = 0u;
}

template <typename T>
__global__ void ppr_init_kernel(PpArgs a, T *R) {
    // r = teleport vector: zero everything, then the seeds
    const size_t total = (size_t)a.nqb * a.pv.N * kQ;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        const size_t qb = i / ((size_t)a.pv.N * kQ);
        const size_t rem = i - qb * ((size_t)a.pv.N * kQ);
        const uint32_t v = (uint32_t)(rem / kQ), lane = (uint32_t)(rem % kQ);
        float m = 0.0f;
        const uint32_t *bits = a.seed_bits + qb * a.words;
        if ((bits[v >> 5] >> (v & 31)) & 1u) {
            const uint32_t *sv = a.seed_v + qb * kSeedTable, *sq = a.seed_q + qb * kSeedTable;
            for (uint32_t k = 0; k < a.seed_n[qb]; k++)
                if (sv[k] == v && sq[k] == lane) m = a.tele[qb * kQ + lane];
        }
        Store<T>::st(R + i, m);
    }
}

__global__ void ppr_widen_kernel(const uint16_t *a, float *b, size_t n) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) b[i] = bf16_to_f32(a[i]);
}

__global__ void inv_outdeg_kernel(const uint32_t *outdeg, uint32_t n, float *inv) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        inv[v] = outdeg[v] ? 1.0f / (float)outdeg[v] : 0.0f;
}

__global__ void ppr_extract_kernel(const void *R, int fp32, uint32_t n, uint32_t q, float *out) {
    uint32_t stride = gridDim.x * blockDim.x;
    const size_t base = (size_t)(q / kQ) * n * kQ;
    const uint32_t lane = q % kQ;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        size_t i = base + (size_t)v * kQ + lane;
        out[v] = fp32 ? ((const float *)R)[i] : bf16_to_f32(((const uint16_t *)R)[i]);
    }
}

static PpArgs make_args(gchi_ppr_batch *b, float damping, int residual) {
    PpArgs a;
    a.pv = b->pv;
    a.inv = b->inv_outdeg;
    a.seed_bits = b->seed_bits;
    a.seed_v = b->seed_v;
    a.seed_q = b->seed_q;
    a.seed_n = b->seed_n;
    a.tele = b->tele;
    a.hub_part = b->hub_part;
    a.res = b->res;
    a.counter = b->counter;
    a.active = b->active;
    a.words = b->words;
    a.nqb = b->nqb;
    a.damping = damping;
    a.residual = residual;
    return a;
}

template <typename T>
static int ppr_sweep(gchi_ppr_batch *b, const T *rin, T *rout, float damping, int residual, hipStream_t s) {
    PpArgs a = make_args(b, damping, residual);
    if (residual) launch_fill_u64(reinterpret_cast<uint64_t *>(b->res), 0ull, (size_t)b->nqb * kQ, s);
    const uint64_t tiles = (uint64_t)b->nqb * (b->pv.nblk + b->pv.nchunk);
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(tiles, (uint64_t)b->ncu * 8));
    hipLaunchKernelGGL(ppr_sweep_kernel<T>, dim3(grid), dim3(kPpThreads), 0, s, a, rin, rout);
    hipLaunchKernelGGL(ppr_hub_kernel<T>, dim3(grid_for((uint64_t)b->pv.nhub * b->nqb * kWave, kPpThreads, 2048) + 0),
                       dim3(kPpThreads), 0, s, a, rin, rout);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

// Captures `len` sweeps of one phase (0: BF16, 1: FP32) from buffer 0, the
// last one with the residual, into a graph kept by the batch; recaptured only
// when the chunk length or the damping changes.  The kernels read the active
// list from device memory, so retiring blocks needs no recapture.
static int ppr_capture_chunk(gchi_ppr_batch *b, int phase, int len, float d) {
    if (b->chunk_exec[phase] && b->chunk_len[phase] == len && b->chunk_damping[phase] == d) return GCHI_OK;
    if (b->chunk_exec[phase]) (void)hipGraphExecDestroy(b->chunk_exec[phase]);
    if (b->chunk_graph[phase]) (void)hipGraphDestroy(b->chunk_graph[phase]);
    b->chunk_exec[phase] = nullptr;
    b->chunk_graph[phase] = nullptr;
    hipStream_t cs;
    GCHI_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int st = GCHI_OK;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) st = hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    if (st == GCHI_OK) {
        int s1 = GCHI_OK;
        for (int k = 0; k < len && s1 == GCHI_OK; k++) {
            const int res = k == len - 1;
            if (phase == 0)
                s1 = ppr_sweep<uint16_t>(b, (const uint16_t *)b->r16[k & 1], (uint16_t *)b->r16[(k + 1) & 1], d, res,
                                         cs);
            else
                s1 = ppr_sweep<float>(b, b->r32[k & 1], b->r32[(k + 1) & 1], d, res, cs);
        }
        e = hipStreamEndCapture(cs, &b->chunk_graph[phase]);
        if (s1 != GCHI_OK) st = s1;
        else if (e != hipSuccess) st = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    }
    if (st == GCHI_OK) {
        e = hipGraphInstantiate(&b->chunk_exec[phase], b->chunk_graph[phase], nullptr, nullptr, 0);
        if (e != hipSuccess) st = hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    (void)hipStreamDestroy(cs);
    if (st != GCHI_OK) {
        if (b->chunk_graph[phase]) (void)hipGraphDestroy(b->chunk_graph[phase]);
        b->chunk_graph[phase] = nullptr;
        b->chunk_exec[phase] = nullptr;
        return st;
    }
    b->chunk_len[phase] = len;
    b->chunk_damping[phase] = d;
    return GCHI_OK;
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_ppr_batch_create(const gchi_pr_plan *plan, uint32_t nqueries, gchi_ppr_batch **out) {
    GCHI_CHECK(plan && out && nqueries > 0, GCHI_ERR_INVALID, "ppr_batch_create: bad argument");
    PlanView pv;
    GCHI_TRY(pr_plan_view(plan, &pv));  // before the batch exists: a refused plan leaks nothing
    gchi_ppr_batch *b = new gchi_ppr_batch();
    b->pv = pv;
    b->nq = nqueries;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
            b->ncu = prop.multiProcessorCount;
    }
    b->nqb = (nqueries + kQ - 1) / kQ;
    b->words = (b->pv.N + 31) / 32;
    const size_t state = (size_t)b->nqb * b->pv.N * kQ;
    Sizer z;
    z.take<float>(b->pv.N);
    z.take<uint16_t>(state);
    z.take<uint16_t>(state);
    z.take<float>(state);
    z.take<float>(state);
    z.take<float>((size_t)b->nqb * std::max<uint32_t>(b->pv.nchunk, 1) * kQ);
    z.take<double>((size_t)b->nqb * kQ);
    z.take<uint32_t>((size_t)b->nqb * b->words);
    z.take<uint32_t>((size_t)b->nqb * kSeedTable);
    z.take<uint32_t>((size_t)b->nqb * kSeedTable);
    z.take<uint32_t>(b->nqb);
    z.take<float>((size_t)b->nqb * kQ);
    z.take<uint32_t>(kPpStripes * kPpStripeGap);
    z.take<uint32_t>((size_t)b->nqb + 1);
    if (hipMalloc(&b->mem, z.off) != hipSuccess) {
        delete b;
        return fail(GCHI_ERR_HIP, "ppr_batch_create: hipMalloc failed (state " + std::to_string(z.off) + " bytes)");
    }
    Carver c(b->mem, z.off);
    b->inv_outdeg = c.take<float>(b->pv.N);
    b->r16[0] = c.take<uint16_t>(state);
    b->r16[1] = c.take<uint16_t>(state);
    b->r32[0] = c.take<float>(state);
    b->r32[1] = c.take<float>(state);
    b->hub_part = c.take<float>((size_t)b->nqb * std::max<uint32_t>(b->pv.nchunk, 1) * kQ);
    b->res = c.take<double>((size_t)b->nqb * kQ);
    b->seed_bits = c.take<uint32_t>((size_t)b->nqb * b->words);
    b->seed_v = c.take<uint32_t>((size_t)b->nqb * kSeedTable);
    b->seed_q = c.take<uint32_t>((size_t)b->nqb * kSeedTable);
    b->seed_n = c.take<uint32_t>(b->nqb);
    b->tele = c.take<float>((size_t)b->nqb * kQ);
    b->counter = c.take<uint32_t>(kPpStripes * kPpStripeGap);
    b->active = c.take<uint32_t>((size_t)b->nqb + 1);
    hipLaunchKernelGGL(inv_outdeg_kernel, dim3(grid_for(b->pv.N, 256, 65536)), dim3(256), 0, 0, b->pv.outdeg, b->pv.N,
                       b->inv_outdeg);
    launch_fill_u32(b->counter, 0u, kPpStripes * kPpStripeGap, 0);
    GCHI_LAUNCH_CHECK();
    GCHI_HIP(hipDeviceSynchronize());
    *out = b;
    return GCHI_OK;
}

extern "C" int gchi_ppr_batch_destroy(gchi_ppr_batch *b) {
    if (!b) return GCHI_OK;
    for (int k = 0; k < 2; k++) {
        if (b->chunk_exec[k]) (void)hipGraphExecDestroy(b->chunk_exec[k]);
        if (b->chunk_graph[k]) (void)hipGraphDestroy(b->chunk_graph[k]);
    }
    if (b->mem) (void)hipFree(b->mem);
    delete b;
    return GCHI_OK;
}

// Runs every query to convergence.  h_seed_off[nq+1] / h_seeds: restart sets.
extern "C" int gchi_ppr_batch_run(gchi_ppr_batch *b, const uint32_t *h_seed_off, const uint32_t *h_seeds,
                                  const gchi_ppr_config *cfg, gchi_ppr_result *result, double *h_residuals,
                                  void *stream) {
    GCHI_CHECK(b && h_seed_off && h_seeds && cfg && result, GCHI_ERR_INVALID, "ppr_batch_run: null argument");
    GCHI_CHECK(cfg->damping > 0 && cfg->damping < 1 && cfg->tol > 0, GCHI_ERR_INVALID, "ppr_batch_run: bad config");
    hipStream_t s = as_stream(stream);
    const uint32_t nqb = b->nqb, N = b->pv.N;
    // seed tables (host-built, uploaded once per run)
    std::vector<uint32_t> bits((size_t)nqb * b->words, 0), sv((size_t)nqb * kSeedTable, 0),
        sq((size_t)nqb * kSeedTable, 0), sn(nqb, 0);
    std::vector<float> tele((size_t)nqb * kQ, 0.0f);
    for (uint32_t q = 0; q < b->nq; q++) {
        uint32_t k0 = h_seed_off[q], k1 = h_seed_off[q + 1];
        GCHI_CHECK(k1 > k0 && k1 - k0 <= (uint32_t)kMaxSeeds, GCHI_ERR_UNSUPPORTED,
                   "ppr_batch_run: each query needs 1..64 seeds");
        std::vector<uint32_t> seeds(h_seeds + k0, h_seeds + k1);
        std::sort(seeds.begin(), seeds.end());
        GCHI_CHECK(std::adjacent_find(seeds.begin(), seeds.end()) == seeds.end(), GCHI_ERR_INVALID,
                   "ppr_batch_run: repeated seed in a restart set");
        const uint32_t qb = q / kQ, lane = q % kQ;
        tele[q] = (float)((1.0 - cfg->damping) / (double)seeds.size());
        for (uint32_t v : seeds) {
            GCHI_CHECK(v < N, GCHI_ERR_INVALID, "ppr_batch_run: seed out of range");
            bits[(size_t)qb * b->words + (v >> 5)] |= 1u << (v & 31);
            size_t slot = (size_t)qb * kSeedTable + sn[qb]++;
            sv[slot] = v;
            sq[slot] = lane;
        }
    }
    for (uint32_t qb = 0; qb < nqb; qb++) {  // sort each table by (vertex, query)
        std::vector<std::pair<uint32_t, uint32_t>> t;
        for (uint32_t k = 0; k < sn[qb]; k++) t.push_back({sv[(size_t)qb * kSeedTable + k], sq[(size_t)qb * kSeedTable + k]});
        std::sort(t.begin(), t.end());
        for (uint32_t k = 0; k < sn[qb]; k++) {
            sv[(size_t)qb * kSeedTable + k] = t[k].first;
            sq[(size_t)qb * kSeedTable + k] = t[k].second;
        }
    }
    GCHI_HIP(hipMemcpyAsync(b->seed_bits, bits.data(), bits.size() * 4, hipMemcpyHostToDevice, s));
    GCHI_HIP(hipMemcpyAsync(b->seed_v, sv.data(), sv.size() * 4, hipMemcpyHostToDevice, s));
    GCHI_HIP(hipMemcpyAsync(b->seed_q, sq.data(), sq.size() * 4, hipMemcpyHostToDevice, s));
    GCHI_HIP(hipMemcpyAsync(b->seed_n, sn.data(), sn.size() * 4, hipMemcpyHostToDevice, s));
    GCHI_HIP(hipMemcpyAsync(b->tele, tele.data(), tele.size() * 4, hipMemcpyHostToDevice, s));

    const float d = (float)cfg->damping;
    const double stop = cfg->tol * (1.0 - cfg->damping) / cfg->damping;  // residual bound
    // sweeps run in chunks of `check` (even, so a chunk ends in the buffer it
    // started from), each chunk one HIP-graph launch whose last sweep computes
    // the residuals; converged query blocks leave the active list
    const int check = std::max(2, (cfg->check_every > 0 ? cfg->check_every : 4) & ~1);
    const int max_sweeps = cfg->max_sweeps > 0 ? cfg->max_sweeps : 1000;
    std::vector<double> res((size_t)nqb * kQ, 0.0), last((size_t)nqb * kQ, 0.0);
    std::vector<uint32_t> active(nqb + 1);
    for (uint32_t qb = 0; qb < nqb; qb++) active[qb] = qb;
    active[nqb] = nqb;
    GCHI_HIP(hipMemcpyAsync(b->active, active.data(), 4 * active.size(), hipMemcpyHostToDevice, s));
    uint32_t nact = nqb;
    auto read_res = [&]() -> int {
        GCHI_HIP(hipMemcpyAsync(res.data(), b->res, res.size() * 8, hipMemcpyDeviceToHost, s));
        GCHI_HIP(hipStreamSynchronize(s));
        for (uint32_t i = 0; i < nact; i++)
            for (int l = 0; l < kQ; l++) last[(size_t)active[i] * kQ + l] = res[(size_t)active[i] * kQ + l];
        return GCHI_OK;
    };
    auto max_res = [&]() {
        double m = 0.0;
        for (uint32_t q = 0; q < b->nq; q++) m = std::max(m, last[q]);
        return m;
    };
    result->sweeps_bf16 = result->sweeps_fp32 = result->block_sweeps_fp32 = 0;
    PpArgs a0 = make_args(b, d, 0);
    int sweeps = 0;
    b->cur = 0;
    b->fp32 = cfg->bf16_sweeps <= 0;
    if (!b->fp32) {
        hipLaunchKernelGGL(ppr_init_kernel<uint16_t>, dim3(65536), dim3(256), 0, s, a0, (uint16_t *)b->r16[0]);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(ppr_capture_chunk(b, 0, check, d));
        double prev = 1e300;
        while (sweeps + check <= max_sweeps && result->sweeps_bf16 + check <= cfg->bf16_sweeps) {
            GCHI_HIP(hipGraphLaunch(b->chunk_exec[0], s));
            sweeps += check;
            result->sweeps_bf16 += check;
            GCHI_TRY(read_res());
            const double m = max_res();
            // leave BF16 once the residual is small or no longer falls (rounding floor)
            if (m <= cfg->bf16_switch_residual || m > 0.7 * prev) break;
            prev = m;
        }
        const size_t state = (size_t)nqb * N * kQ;
        hipLaunchKernelGGL(ppr_widen_kernel, dim3(65536), dim3(256), 0, s, (const uint16_t *)b->r16[0], b->r32[0],
                           state);
        GCHI_LAUNCH_CHECK();
        b->fp32 = true;
    } else {
        hipLaunchKernelGGL(ppr_init_kernel<float>, dim3(65536), dim3(256), 0, s, a0, b->r32[0]);
        GCHI_LAUNCH_CHECK();
    }
    GCHI_TRY(ppr_capture_chunk(b, 1, check, d));
    bool converged = false;
    while (sweeps + check <= max_sweeps) {
        GCHI_HIP(hipGraphLaunch(b->chunk_exec[1], s));
        sweeps += check;
        result->sweeps_fp32 += check;
        GCHI_TRY(read_res());
        // retire the query blocks whose every query meets the bound
        uint32_t keep = 0;
        for (uint32_t i = 0; i < nact; i++) {
            const uint32_t qb = active[i];
            double m = 0.0;
            for (int l = 0; l < kQ; l++) m = std::max(m, last[(size_t)qb * kQ + l]);
            if (m > stop) active[keep++] = qb;
        }
        result->block_sweeps_fp32 += (int)nact * check;
        if (keep == 0) {
            converged = true;
            break;
        }
        if (keep != nact) {
            nact = keep;
            active[nqb] = nact;
            GCHI_HIP(hipMemcpyAsync(b->active, active.data(), 4 * nact, hipMemcpyHostToDevice, s));
            GCHI_HIP(hipMemcpyAsync(b->active + nqb, &active[nqb], 4, hipMemcpyHostToDevice, s));
        }
    }
    GCHI_HIP(hipStreamSynchronize(s));
    result->converged = converged ? 1 : 0;
    result->max_residual = max_res();
    if (h_residuals)
        for (uint32_t q = 0; q < b->nq; q++) h_residuals[q] = last[q];
    return GCHI_OK;
}

extern "C" int gchi_ppr_batch_extract(const gchi_ppr_batch *b, uint32_t query, float *d_out, void *stream) {
    GCHI_CHECK(b && d_out && query < b->nq, GCHI_ERR_INVALID, "ppr_batch_extract: bad argument");
    const void *R = b->fp32 ? (const void *)b->r32[b->cur] : b->r16[b->cur];
    hipLaunchKernelGGL(ppr_extract_kernel, dim3(grid_for(b->pv.N, 256, 65536)), dim3(256), 0, as_stream(stream), R,
                       b->fp32 ? 1 : 0, b->pv.N, query, d_out);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}
