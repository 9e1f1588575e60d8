// Synchronous PageRank pull sweep (FP32 ranks, FP64 apply and residuals).
//
// Reference semantics (bulk-synchronous functional engine):
//   gather : s_v = plus(s_v, op_neighborval(x_u)) over in-edges, FP32 adds
//            (src/api/functional/functional_bulksync.hpp:91-101)
//   apply  : r_v = compute_vertexvalue(s_v) = teleport + damping*s_v evaluated
//            in double, rounded to float (example_apps/pagerank_functional.cpp:72-75)
//   scatter: x_v = value_to_neighbor(r_v) = r_v / outdeg_v in FP32 (:78-81),
//            written to the out-edges' double buffer (functional_bulksync.hpp:107-113)
// The per-edge PairContainer (src/graphchi_types.hpp:34-59) becomes one x value
// per vertex: every out-edge of u carries the same x_u, so a pull over in-edges
// reads exactly what the reference's gather reads.
//
// Layout (built once by gchi_pr_plan_create from the CSC):
//   regular rows (in-degree < W): reg_off[N+G] (graph g's rows at slots
//     v+g, relative to the graph's first regular edge ebase[g], plus one end
//     slot per graph), reg_src[E_reg]; a row's "work" is deg+1, and block b
//     owns rows with reg_off[v+g]+(v-vbase[g]) in [b*W, (b+1)*W) of its graph,
//     so a block sees < 2W edges and < 2W rows.  Offsets are 32-bit per graph
//     and 64-bit across graphs, so a batch of graphs may hold more than 2^32
//     edges in total.  Its in-edges are streamed coalesced
//     (non-temporal), x[src] gathered into LDS, then rows are reduced: one lane
//     per row up to short_row_max edges (sequential adds in src order, as the
//     reference's gather), one wavefront per longer row.
//   hub rows (in-degree >= W): their in-edges are cut into chunks of 512; one
//     wavefront reduces a chunk; a finalize kernel sums each hub's chunk
//     partials in chunk order and applies.
// Many graphs (a disjoint union with vertex bases) share one plan: blocks are
// aligned to graph boundaries, so every block and hub belongs to one graph,
// the apply can use a per-graph constant and FP64 statistics are per graph.
// Tiles (hub chunks first, then regular blocks) are either one workgroup each
// or pulled by persistent workgroups from a device work counter.
// Gather layout (optional): sources are ranked by out-degree, descending, and
// the plan's edges store that rank; each sweep first copies x into rank order
// (xg), so the most referenced values share cache lines, and persistent
// workgroups of the LDS-cached kernel keep xg[0, H) in LDS for the whole sweep
// (on R-MAT graphs the top 32K sources cover about half of all edges).  The
// gathered values, and therefore every sum, are the same as without it.
// Every sum is in a fixed order, so a sweep is bit-reproducible run to run.
#include "common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <vector>

struct gchi_pr_plan {
    uint32_t N = 0;
    uint64_t E = 0;
    uint32_t G = 1;        // graphs (disjoint union)
    int W = 2048;
    int short_max = 32;
    int persistent = 0;
    uint32_t nblk = 0;     // regular blocks
    uint32_t nhub = 0;     // hub rows
    uint32_t nchunk = 0;   // hub chunks
    uint32_t nfin = 0;     // finalize workgroups (single-graph plans)
    uint32_t npersist = 0; // workgroups of the persistent sweep
    uint64_t e_reg = 0, e_hub = 0;
    const uint32_t *outdeg = nullptr;
    uint32_t *reg_off = nullptr;
    uint32_t *reg_src = nullptr;
    uint32_t *hub_bits = nullptr;
    uint32_t *blk_row = nullptr;
    uint32_t *blk_graph = nullptr;
    uint32_t *tile_e0 = nullptr;   // [nblk+1] first regular edge of each block (relative to its graph)
    uint32_t *tile_e1 = nullptr;   // [nblk+1] end regular edge of each block (relative to its graph)
    uint64_t *ebase = nullptr;     // [G] first regular edge of each graph in reg_src
    uint32_t *own_outdeg = nullptr;  // out-degrees copied from per-graph CSCs (batch plans)
    uint32_t *hub_v = nullptr;
    uint32_t *hub_graph = nullptr;
    uint32_t *hub_chunk_off = nullptr;
    uint64_t *chunk_beg = nullptr;
    uint32_t *chunk_len = nullptr;
    uint32_t *hub_src = nullptr;
    uint32_t *gblk = nullptr;  // [G+1] block range per graph
    uint32_t *ghub = nullptr;  // [G+1] hub range per graph
    uint32_t *vbase = nullptr; // [G+1] vertex range per graph
    float *chunk_part = nullptr;
    double *blk_stat = nullptr;
    double *hub_stat = nullptr;
    double *fin_stat = nullptr;
    uint32_t *counters = nullptr;  // [0] finalize ticket, [kPrStripeGap...] tile counter stripes
    int hot = 0;           // tune value (0: vertex ids, -1: gather layout, > 0: layout + LDS cache)
    uint32_t A = 0;        // sources in the gather layout (out-degree > 0)
    uint32_t H = 0;        // layout values cached in LDS by the wavefront-tiled kernel
    uint32_t nwave_wg = 0; // workgroups of the wavefront-tiled kernel (persistent == 2)
    size_t wave_lds = 0;   // its dynamic LDS bytes
    uint32_t nhot_wg = 0;  // workgroups of the cached workgroup-tiled kernel (persistent == 3)
    size_t hotwg_lds = 0;  // its dynamic LDS bytes
    uint32_t *gorder = nullptr;  // [A] vertex at each layout position
    float *xg = nullptr;         // [A] x in layout order, refreshed by every sweep
    void *mem = nullptr;
    gchi_seg_plan *seg = nullptr;  // launch form 4: the rank-space tiled plan (pr_seg.hip) instead of the above
};

namespace gchi {

constexpr int kPrThreads = 256;
constexpr int kPrWaves = kPrThreads / kWave;
constexpr int kHubChunk = 512;
constexpr int kMaxW = 4096;
constexpr int kMaxLong = 512;  // long rows per regular block
constexpr int kFinThreads = 256;
constexpr size_t kStaticLds = 4096;  // slack for static LDS beside the dynamic allocation
constexpr int kWaveThreads = 1024;   // workgroup of the wavefront-tiled sweep (one per CU)
constexpr int kWaveWaves = kWaveThreads / kWave;

struct PrConsts {
    double teleport;
    double damping;
    const double *g_tele;      // per-graph constant term, or null
    const uint32_t *seed_bits; // seed bitmap, or null
    double seed_tele;
};

constexpr uint32_t kPrStripes = 16;    // tile counters of the persistent sweep
constexpr uint32_t kPrStripeGap = 32;  // uint32 words between counters (128 bytes)
constexpr uint32_t kPrCounterWords = kPrStripeGap * (kPrStripes + 1);

GCHI_DEV void pr_reset_stripes(uint32_t *counters) {
    for (uint32_t i = threadIdx.x; i < kPrStripes; i += blockDim.x) counters[kPrStripeGap * (i + 1)] = 0u;
}

struct PrArgs {
    const uint32_t *reg_off, *reg_src, *blk_row, *blk_graph, *tile_e0, *tile_e1, *hub_bits;
    const uint64_t *ebase, *chunk_beg;
    const uint32_t *chunk_len, *hub_src, *outdeg;
    const float *xg;  // gather source: x by vertex id, or the gather layout
    const float *r_in;
    float *x_out, *r_out, *chunk_part;
    double *blk_stat;
    uint32_t *tile_counter;
    uint32_t nchunk, nhubt, nblk, ntiles;  // hub chunks, hub tiles (NW chunks each), regular blocks, all tiles
    int short_max;
};

GCHI_DEV double row_teleport(const PrConsts &c, uint32_t g, uint32_t v) {
    double t = c.g_tele ? c.g_tele[g] : c.teleport;
    if (c.seed_bits && ((c.seed_bits[v >> 5] >> (v & 31)) & 1u)) t += c.seed_tele;
    return t;
}

// r = (float)(teleport + damping * (double)s), without contraction into an FMA:
// the reference evaluates a double multiply and a double add, then rounds.
GCHI_DEV float pr_apply(double teleport, double damping, float s) {
#pragma clang fp contract(off)
    double t = damping * (double)s;
    double r = teleport + t;
    return (float)r;
}

GCHI_DEV float pr_scatter(float r, uint32_t outdeg) { return outdeg ? __fdiv_rn(r, (float)outdeg) : 0.0f; }

GCHI_DEV void pr_finish_row(uint32_t v, uint32_t g, float s, const PrConsts &c, const float *r_in,
                            const uint32_t *outdeg, float *r_out, float *x_out, double &res, double &sum) {
    float r = pr_apply(row_teleport(c, g, v), c.damping, s);
    float ro = r_in[v];
    r_out[v] = r;
    x_out[v] = pr_scatter(r, outdeg[v]);
    res += fabs((double)r - (double)ro);
    sum += (double)r;
}

// Makes a wavefront's LDS writes visible to its own later LDS reads (and
// orders its reads before later overwrites); no other wavefront takes part.
GCHI_DEV void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Partial sum of one hub chunk (<= kHubChunk in-edges) by one wavefront, in
// the one order every kernel uses: lane l adds edges l, l+64, l+128, ... in
// turn, then the 64 lane sums are combined by the butterfly of wave_sum.
// Returns the sum in every lane.  With HOT, layout positions below H are read
// from the LDS cache.
template <bool HOT>
GCHI_DEV float hub_chunk_sum(const PrArgs &a, uint32_t c, const float *hot, uint32_t H, int lane) {
    constexpr int PER = kHubChunk / kWave;
    const uint64_t cb = a.chunk_beg[c];
    const uint32_t cl = a.chunk_len[c];
    // Every load is issued unconditionally at an in-bounds address (lanes past
    // the chunk read its first edge and discard it), so the compiler emits no
    // branch around loads and its wait counts stay exact.
    uint32_t idx[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        uint32_t k = (uint32_t)j * kWave + lane;
        idx[j] = __builtin_nontemporal_load(&a.hub_src[cb + (k < cl ? k : 0u)]);
    }
    float val[PER];
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t u = idx[j];
        const bool in_lds = HOT && u < H;
        float g = a.xg[in_lds ? 0u : u];
        if (HOT) {
            float h = hot[in_lds ? u : 0u];
            g = in_lds ? h : g;
        }
        val[j] = (uint32_t)j * kWave + lane < cl ? g : 0.0f;
    }
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < PER; j++) acc += val[j];
    return wave_sum(acc);
}

// One tile of a T-thread workgroup: NW = T/64 hub chunks (one per wavefront,
// tile < nhubt) or one regular block.
template <int T, bool HOT = false>
GCHI_DEV void pr_tile(const PrArgs &a, const PrConsts &c, uint32_t tile, float *s_val, uint32_t *s_long,
                      uint32_t *s_nlong, double *s_red, const float *hot = nullptr, uint32_t H = 0) {
    constexpr int NW = T / kWave;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    if (tile < a.nhubt) {
        const uint32_t ch = tile * NW + wave;
        if (ch < a.nchunk) {
            float t = hub_chunk_sum<HOT>(a, ch, hot, H, lane);
            if (lane == 0) a.chunk_part[ch] = t;
        }
        return;
    }

    // ---- regular block ----
    const uint32_t b = tile - a.nhubt;
    const uint32_t rb = a.blk_row[b], re = a.blk_row[b + 1];
    const uint32_t g = a.blk_graph[b];
    const uint32_t *ro = a.reg_off + g;  // the graph's offset slots: row v at ro[v]
    const uint32_t e0 = a.tile_e0[b], e1 = a.tile_e1[b];
    const uint32_t span = e1 - e0;
    const uint32_t *rs = a.reg_src + a.ebase[g] + e0;
    if (tid == 0) *s_nlong = 0;
    for (uint32_t base = 0; base < span; base += 8 * T) {
        uint32_t idx[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t k = base + (uint32_t)j * T + tid;
            idx[j] = k < span ? __builtin_nontemporal_load(&rs[k]) : 0u;
        }
        float val[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t k = base + (uint32_t)j * T + tid;
            val[j] = 0.0f;
            if (k < span) val[j] = (HOT && idx[j] < H) ? hot[idx[j]] : a.xg[idx[j]];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t k = base + (uint32_t)j * T + tid;
            if (k < span) s_val[k] = val[j];
        }
    }
    __syncthreads();

    double res = 0.0, sum = 0.0;
    for (uint32_t v = rb + tid; v < re; v += T) {
        if ((a.hub_bits[v >> 5] >> (v & 31)) & 1u) continue;
        uint32_t beg = ro[v] - e0, end = ro[v + 1] - e0;
        if ((int)(end - beg) <= a.short_max) {
            float s = 0.0f;
            for (uint32_t k = beg; k < end; k++) s += s_val[k];
            pr_finish_row(v, g, s, c, a.r_in, a.outdeg, a.r_out, a.x_out, res, sum);
        } else {
            uint32_t slot = atomicAdd(s_nlong, 1u);
            s_long[slot] = v;
        }
    }
    __syncthreads();
    const uint32_t nlong = *s_nlong;
    for (uint32_t li = wave; li < nlong; li += NW) {
        uint32_t v = s_long[li];
        uint32_t beg = ro[v] - e0, end = ro[v + 1] - e0;
        float s = 0.0f;
        for (uint32_t k = beg + lane; k < end; k += kWave) s += s_val[k];
        s = wave_sum(s);
        if (lane == 0) pr_finish_row(v, g, s, c, a.r_in, a.outdeg, a.r_out, a.x_out, res, sum);
    }
    res = wave_sum(res);
    sum = wave_sum(sum);
    if (lane == 0) {
        s_red[2 * wave] = res;
        s_red[2 * wave + 1] = sum;
    }
    __syncthreads();
    if (tid == 0) {
        double r0 = 0.0, s0 = 0.0;
        for (int w = 0; w < NW; w++) {
            r0 += s_red[2 * w];
            s0 += s_red[2 * w + 1];
        }
        a.blk_stat[2 * b] = r0;
        a.blk_stat[2 * b + 1] = s0;
    }
}

__global__ __launch_bounds__(kPrThreads) void pr_sweep_kernel(PrArgs a, PrConsts c) {
    extern __shared__ float s_val[];  // 2W floats
    __shared__ uint32_t s_long[kMaxLong];
    __shared__ uint32_t s_nlong;
    __shared__ double s_red[2 * kPrWaves];
    pr_tile<kPrThreads>(a, c, blockIdx.x, s_val, s_long, &s_nlong, s_red);
}

// Persistent form: workgroups pull tiles from the work counters until the queue
// is drained (hub chunks first, then regular blocks).  Atomics on one address
// complete one after another (about 10 ns each here), so one counter would cap
// a sweep at ~10^8 tiles per second; kPrStripes counters, each in its own
// 128-byte line, hand out the tiles s, s + kPrStripes, ... of stripe s.  A
// workgroup starts on stripe blockIdx % kPrStripes and moves to the next one
// when it is drained, so the tiles in flight stay a narrow window of the tile
// order (the sources the gathers touch stay in cache).  Every workgroup exits
// once all stripes are drained; the finalize kernels reset the counters.
__global__ __launch_bounds__(kPrThreads) void pr_sweep_persistent_kernel(PrArgs a, PrConsts c) {
    extern __shared__ float s_val[];
    __shared__ uint32_t s_long[kMaxLong];
    __shared__ uint32_t s_nlong;
    __shared__ double s_red[2 * kPrWaves];
    __shared__ uint32_t s_tile;
    uint32_t stripe = blockIdx.x % kPrStripes, spent = 0;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&a.tile_counter[stripe * kPrStripeGap], 1u);
        __syncthreads();
        const uint64_t tile = stripe + (uint64_t)kPrStripes * s_tile;
        __syncthreads();  // s_tile is rewritten by the next ticket
        if (tile >= a.ntiles) {  // block-uniform
            if (++spent == kPrStripes) break;
            stripe = (stripe + 1) % kPrStripes;
            continue;
        }
        pr_tile<kPrThreads>(a, c, (uint32_t)tile, s_val, s_long, &s_nlong, s_red);
        __syncthreads();
    }
}

// Copies layout positions [0, H) of xg into the LDS cache, kFillAhead loads per
// thread in flight (one at a time would take a memory round trip per 1024
// values).
constexpr uint32_t kFillAhead = 8;

GCHI_DEV void fill_hot(float *hot, const float *__restrict__ xg, uint32_t H, uint32_t tid, uint32_t nthreads) {
    for (uint32_t i0 = tid; i0 < H; i0 += nthreads * kFillAhead) {
        float v[kFillAhead];
#pragma unroll
        for (uint32_t u = 0; u < kFillAhead; u++) {
            const uint32_t i = i0 + u * nthreads;
            v[u] = xg[i < H ? i : 0u];
        }
#pragma unroll
        for (uint32_t u = 0; u < kFillAhead; u++) {
            const uint32_t i = i0 + u * nthreads;
            if (i < H) hot[i] = v[u];
        }
    }
}

// Persistent workgroups with a static tile list (tile = blockIdx.x, then every
// gridDim.x-th) and an LDS cache of layout positions [0, H), loaded once per
// workgroup; several such workgroups share a CU, each with its own copy.
__global__ __launch_bounds__(kPrThreads) void pr_sweep_hotwg_kernel(PrArgs a, PrConsts c, uint32_t H) {
    extern __shared__ float lds[];  // H + 2W floats
    __shared__ uint32_t s_long[kMaxLong];
    __shared__ uint32_t s_nlong;
    __shared__ double s_red[2 * kPrWaves];
    float *hot = lds;
    float *s_val = lds + H;
    fill_hot(hot, a.xg, H, threadIdx.x, kPrThreads);
    __syncthreads();
    for (uint32_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
        pr_tile<kPrThreads, true>(a, c, tile, s_val, s_long, &s_nlong, s_red, hot, H);
        __syncthreads();
    }
}

// ---- wavefront-tiled persistent sweep ----
// Every wavefront works alone through a fixed list of tiles: hub chunks
// round-robin, then a contiguous range of regular blocks, whose gathered values
// it stages in its own LDS slice (2W floats), so no workgroup barrier separates
// one wavefront's gathers from another's reductions.  The workgroup first loads
// layout positions [0, H) into LDS (the most referenced sources of the gather
// layout) and serves those gathers from there for the whole sweep.
// Blocks are software-pipelined: while block b is reduced, the gathers of
// block b+1 and the edge ids of block b+2 are in flight.  Memory operations
// complete in issue order as far as the wait counters are concerned, so the
// ids of b+2 and the row metadata of b+1 are issued before the gathers of b+1,
// and everything the reduction of b reads was issued before the gathers of b.
// Per row, the summation order is the one of pr_tile (one lane per short row
// in src order, one wavefront per long row, strided), so ranks are
// bit-identical to the workgroup-tiled kernels.

struct TileDesc {
    uint32_t rb, re, e0, e1, g;  // e0/e1 relative to the graph's first regular edge
    uint64_t eb;                 // that edge's position in reg_src
    double tele;                 // the graph's constant term
};

// Row metadata of row v = rb + lane of a block (zero when v >= re).
struct RowMeta {
    uint32_t beg, end, hubw, seedw, od;
    float ro;
};

// Descriptors of 64 consecutive blocks, lane j holding block base + j.
struct DescBatch {
    uint32_t base;
    uint32_t rb, re, e0, e1, g;
    uint32_t tlo, thi, eblo, ebhi;
};

GCHI_DEV void load_desc_batch(const PrArgs &a, const PrConsts &c, uint32_t base, uint32_t bend, int lane,
                              DescBatch &d) {
    d.base = base;
    const uint32_t b = min(base + (uint32_t)lane, bend - 1);  // bend > base
    d.rb = a.blk_row[b];
    d.re = a.blk_row[b + 1];
    d.e0 = a.tile_e0[b];
    d.e1 = a.tile_e1[b];
    d.g = a.blk_graph[b];
    double t = c.teleport;
    if (c.g_tele) t = c.g_tele[d.g];
    d.tlo = (uint32_t)__double_as_longlong(t);
    d.thi = (uint32_t)(__double_as_longlong(t) >> 32);
    const uint64_t eb = a.ebase[d.g];
    d.eblo = (uint32_t)eb;
    d.ebhi = (uint32_t)(eb >> 32);
}

GCHI_DEV TileDesc get_desc(const DescBatch &d, uint32_t b) {
    const int j = (int)(b - d.base);
    TileDesc t;
    t.rb = __builtin_amdgcn_readlane(d.rb, j);
    t.re = __builtin_amdgcn_readlane(d.re, j);
    t.e0 = __builtin_amdgcn_readlane(d.e0, j);
    t.e1 = __builtin_amdgcn_readlane(d.e1, j);
    t.g = __builtin_amdgcn_readlane(d.g, j);
    // readlane returns int: widen through uint32_t so the low word is not sign-extended
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readlane(d.tlo, j);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readlane(d.thi, j);
    t.tele = __longlong_as_double((long long)(lo | (hi << 32)));
    t.eb = (uint64_t)(uint32_t)__builtin_amdgcn_readlane(d.eblo, j) |
           ((uint64_t)(uint32_t)__builtin_amdgcn_readlane(d.ebhi, j) << 32);
    return t;
}

// The loads below are issued unconditionally at in-bounds addresses (a lane
// with nothing to load reads a valid element), and nothing is computed from a
// loaded value until it is consumed one pipeline step later: a load under a
// branch, or an instruction using a loaded register, makes the wavefront wait
// for every earlier load, which would serialise the pipeline.

// Edge ids of a block, raw (lanes past the span hold a discarded value).
template <int PER>
GCHI_DEV void load_ids(const PrArgs &a, const TileDesc &t, int lane, uint32_t (&idx)[PER]) {
    const uint32_t span = t.e1 - t.e0;
    const uint32_t *rs = a.reg_src + t.eb + t.e0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t k = (uint32_t)j * kWave + lane;
        idx[j] = __builtin_nontemporal_load(k < span ? &rs[k] : a.reg_src);
    }
}

// Gathered values of a block, raw: gv from global memory, hv from the LDS
// cache; cls bit j = value j comes from LDS, bit 16+j = no edge (value 0).
template <int PER, bool HOT>
struct Gathered {
    float gv[PER];
    float hv[HOT ? PER : 1];
    uint32_t cls;
};

template <int PER, bool HOT>
GCHI_DEV void gather_ids(const PrArgs &a, const TileDesc &t, const float *hot, uint32_t H, int lane,
                         const uint32_t (&idx)[PER], Gathered<PER, HOT> &gd) {
    const uint32_t span = t.e1 - t.e0;
    uint32_t cls = 0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        const uint32_t u = idx[j];
        const bool none = (uint32_t)j * kWave + lane >= span;
        const bool in_lds = HOT && !none && u < H;
        gd.gv[j] = a.xg[(none || in_lds) ? 0u : u];
        if (HOT) gd.hv[j] = hot[in_lds ? u : 0u];
        cls |= (in_lds ? 1u : 0u) << j;
        cls |= (none ? 1u : 0u) << (16 + j);
    }
    gd.cls = cls;
}

// Row metadata of row v = rb + lane, raw (meaningful only when v < re).
GCHI_DEV void load_meta(const PrArgs &a, const PrConsts &c, const TileDesc &t, int lane, RowMeta &m) {
    const uint32_t v = t.rb + lane;
    const uint32_t w = v < t.re ? v : t.rb;  // rb < re: every block has a row
    m.beg = a.reg_off[w + t.g];
    m.end = a.reg_off[w + t.g + 1];
    m.hubw = a.hub_bits[w >> 5];
    m.seedw = (c.seed_bits ? c.seed_bits : a.hub_bits)[w >> 5];
    m.od = a.outdeg[w];
    m.ro = a.r_in[w];
}

// pr_finish_row with the row's inputs already in registers
GCHI_DEV void finish_row_pf(uint32_t v, double tele, uint32_t seedw, const PrConsts &c, float s, float ro,
                            uint32_t od, float *r_out, float *x_out, double &res, double &sum) {
    double t = tele;
    if (c.seed_bits && ((seedw >> (v & 31)) & 1u)) t += c.seed_tele;
    float r = pr_apply(t, c.damping, s);
    r_out[v] = r;
    x_out[v] = pr_scatter(r, od);
    res += fabs((double)r - (double)ro);
    sum += (double)r;
}

// ---- hub chunks of the wavefront-tiled sweep, software-pipelined ----
// Same order of additions as hub_chunk_sum (lane j*64+l values in turn, then
// the butterfly), with the ids of chunk i+2 and the gathers of chunk i+1 in
// flight while chunk i is summed.
constexpr int kChunkPer = kHubChunk / kWave;

GCHI_DEV void chunk_ids(const PrArgs &a, uint32_t ch, int lane, uint32_t (&idx)[kChunkPer]) {
    const uint64_t cb = a.chunk_beg[ch];
    const uint32_t cl = a.chunk_len[ch];
#pragma unroll
    for (int j = 0; j < kChunkPer; j++) {
        const uint32_t k = (uint32_t)j * kWave + lane;
        idx[j] = __builtin_nontemporal_load(&a.hub_src[cb + (k < cl ? k : 0u)]);
    }
}

template <bool HOT>
struct ChunkVals {
    float gv[kChunkPer];
    float hv[HOT ? kChunkPer : 1];
    uint32_t lds_bits, cl;
};

template <bool HOT>
GCHI_DEV void chunk_gather(const PrArgs &a, uint32_t ch, const float *hot, uint32_t H,
                           const uint32_t (&idx)[kChunkPer], ChunkVals<HOT> &cv) {
    uint32_t bits = 0;
#pragma unroll
    for (int j = 0; j < kChunkPer; j++) {
        const uint32_t u = idx[j];
        const bool in_lds = HOT && u < H;
        cv.gv[j] = a.xg[in_lds ? 0u : u];
        if (HOT) cv.hv[j] = hot[in_lds ? u : 0u];
        bits |= (in_lds ? 1u : 0u) << j;
    }
    cv.lds_bits = bits;
    cv.cl = a.chunk_len[ch];
}

template <bool HOT>
GCHI_DEV float chunk_sum(const ChunkVals<HOT> &cv, int lane) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < kChunkPer; j++) {
        float v = cv.gv[j];
        asm volatile("" : "+v"(v));  // no "select of loads" folding (see wave_step)
        if (HOT) {
            float h = cv.hv[j];
            asm volatile("" : "+v"(h));
            v = ((cv.lds_bits >> j) & 1u) ? h : v;
        }
        acc += (uint32_t)j * kWave + lane < cv.cl ? v : 0.0f;
    }
    return wave_sum(acc);
}

template <bool HOT, int S>
GCHI_DEV void chunk_step(const PrArgs &a, const float *hot, uint32_t H, int lane, uint32_t ch, uint32_t step,
                         uint32_t (&idx)[2][kChunkPer], ChunkVals<HOT> (&cv)[2]) {
    if (ch + step < a.nchunk) chunk_gather<HOT>(a, ch + step, hot, H, idx[1 - S], cv[1 - S]);
    if (ch + 2 * step < a.nchunk) chunk_ids(a, ch + 2 * step, lane, idx[S]);
    const float t = chunk_sum<HOT>(cv[S], lane);
    if (lane == 0) a.chunk_part[ch] = t;
}

// Issue priority from the wavefront's remaining pipeline steps.  Every
// wavefront of the sweep gets about the same number of steps, and a SIMD issues
// the oldest ready wavefront first, so without this the oldest of a
// workgroup's 16 finish well before the youngest, which then run on at low
// occupancy.  A wavefront that falls behind keeps a higher priority than one
// ahead of it (four levels: more than 3/4, 1/2, 1/4 of its steps left).
GCHI_DEV void wave_prio(uint32_t left, uint32_t total) {
    if (4 * (uint64_t)left > 3 * (uint64_t)total) __builtin_amdgcn_s_setprio(3);
    else if (2 * (uint64_t)left > total) __builtin_amdgcn_s_setprio(2);
    else if (4 * (uint64_t)left > total) __builtin_amdgcn_s_setprio(1);
    else __builtin_amdgcn_s_setprio(0);
}

template <bool HOT>
GCHI_DEV void wave_chunks(const PrArgs &a, const float *hot, uint32_t H, int lane, uint32_t ch0, uint32_t step,
                          uint32_t after, uint32_t total) {
    if (ch0 >= a.nchunk) return;
    uint32_t idx[2][kChunkPer];
    ChunkVals<HOT> cv[2];
    chunk_ids(a, ch0, lane, idx[0]);
    if (ch0 + step < a.nchunk) chunk_ids(a, ch0 + step, lane, idx[1]);
    chunk_gather<HOT>(a, ch0, hot, H, idx[0], cv[0]);
    uint32_t ch = ch0;
    uint32_t left = (a.nchunk - ch0 + step - 1) / step + after;
    for (; ch + step < a.nchunk; ch += 2 * step, left -= 2) {
        wave_prio(left, total);
        chunk_step<HOT, 0>(a, hot, H, lane, ch, step, idx, cv);
        chunk_step<HOT, 1>(a, hot, H, lane, ch + step, step, idx, cv);
    }
    if (ch < a.nchunk) chunk_step<HOT, 0>(a, hot, H, lane, ch, step, idx, cv);
}

// Reduces block b (values staged in st, metadata of its first 64 rows in m).
GCHI_DEV void wave_rows(const PrArgs &a, const PrConsts &c, uint32_t b, const TileDesc &t, const RowMeta &m,
                        const float *st, int lane) {
    double res = 0.0, sum = 0.0;
    {
        const uint32_t v = t.rb + lane;
        bool is_long = false;
        if (v < t.re && !((m.hubw >> (v & 31)) & 1u)) {
            const uint32_t beg = m.beg - t.e0, end = m.end - t.e0;
            if ((int)(end - beg) <= a.short_max) {
                float s = 0.0f;
                for (uint32_t k = beg; k < end; k++) s += st[k];
                finish_row_pf(v, t.tele, m.seedw, c, s, m.ro, m.od, a.r_out, a.x_out, res, sum);
            } else {
                is_long = true;
            }
        }
        uint64_t mask = __ballot(is_long);
        while (mask) {
            const int i = __builtin_ctzll(mask);
            mask &= mask - 1;
            const uint32_t u = t.rb + (uint32_t)i;
            const uint32_t beg = (uint32_t)__builtin_amdgcn_readlane(m.beg, i) - t.e0;
            const uint32_t end = (uint32_t)__builtin_amdgcn_readlane(m.end, i) - t.e0;
            const uint32_t seedw = __builtin_amdgcn_readlane(m.seedw, i);
            const uint32_t od = __builtin_amdgcn_readlane(m.od, i);
            const float ro = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(m.ro), i));
            float s = 0.0f;
            for (uint32_t k = beg + lane; k < end; k += kWave) s += st[k];
            s = wave_sum(s);
            if (lane == 0) finish_row_pf(u, t.tele, seedw, c, s, ro, od, a.r_out, a.x_out, res, sum);
        }
    }
    // rows past the first 64 (blocks of many short rows): metadata read here
    for (uint32_t v0 = t.rb + kWave; v0 < t.re; v0 += kWave) {
        const uint32_t v = v0 + lane;
        bool is_long = false;
        if (v < t.re && !((a.hub_bits[v >> 5] >> (v & 31)) & 1u)) {
            const uint32_t beg = a.reg_off[v + t.g] - t.e0, end = a.reg_off[v + t.g + 1] - t.e0;
            if ((int)(end - beg) <= a.short_max) {
                float s = 0.0f;
                for (uint32_t k = beg; k < end; k++) s += st[k];
                pr_finish_row(v, t.g, s, c, a.r_in, a.outdeg, a.r_out, a.x_out, res, sum);
            } else {
                is_long = true;
            }
        }
        uint64_t mask = __ballot(is_long);
        while (mask) {
            const uint32_t u = v0 + (uint32_t)__builtin_ctzll(mask);
            mask &= mask - 1;
            const uint32_t beg = a.reg_off[u + t.g] - t.e0, end = a.reg_off[u + t.g + 1] - t.e0;
            float s = 0.0f;
            for (uint32_t k = beg + lane; k < end; k += kWave) s += st[k];
            s = wave_sum(s);
            if (lane == 0) pr_finish_row(u, t.g, s, c, a.r_in, a.outdeg, a.r_out, a.x_out, res, sum);
        }
    }
    res = wave_sum(res);
    sum = wave_sum(sum);
    if (lane == 0) {
        a.blk_stat[2 * b] = res;
        a.blk_stat[2 * b + 1] = sum;
    }
}

// One pipeline step for block b (slot S holds b's gathered values and
// metadata, slot 1-S receives b+1's; idx[S] receives b+2's edge ids).
template <int PER, bool HOT, int S>
GCHI_DEV void wave_step(const PrArgs &a, const PrConsts &c, const float *hot, uint32_t H, float *st, int lane,
                        uint32_t b, uint32_t b1, const DescBatch &db, uint32_t (&idx)[2][PER],
                        Gathered<PER, HOT> (&gd)[2], RowMeta (&meta)[2]) {
    const TileDesc t = get_desc(db, b);
    if (b + 2 < b1) load_ids<PER>(a, get_desc(db, b + 2), lane, idx[S]);
    if (b + 1 < b1) {
        const TileDesc t1 = get_desc(db, b + 1);
        load_meta(a, c, t1, lane, meta[1 - S]);
        gather_ids<PER, HOT>(a, t1, hot, H, lane, idx[1 - S], gd[1 - S]);
    }
    // stage block b: the whole 2W slice, zeros past the span
    const uint32_t cls = gd[S].cls;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        // the empty asm keeps the compiler from folding "select of two loaded
        // values" into a load through a selected address (a stack array)
        float v = gd[S].gv[j];
        asm volatile("" : "+v"(v));
        if (HOT) {
            float h = gd[S].hv[j];
            asm volatile("" : "+v"(h));
            v = ((cls >> j) & 1u) ? h : v;
        }
        v = ((cls >> (16 + j)) & 1u) ? 0.0f : v;
        st[(uint32_t)j * kWave + lane] = v;
    }
    wave_lds_sync();
    wave_rows(a, c, b, t, meta[S], st, lane);
    wave_lds_sync();
}

// Blocks [b0, b1) of one wavefront, b1 - b0 <= 64: descriptors in one batch,
// then the pipeline (prologue, steps in slot pairs).
template <int PER, bool HOT>
GCHI_DEV void wave_blocks(const PrArgs &a, const PrConsts &c, const float *hot, uint32_t H, float *st, int lane,
                          uint32_t b0, uint32_t b1, uint32_t bend, uint32_t total) {
    DescBatch db;
    load_desc_batch(a, c, b0, b1, lane, db);
    uint32_t idx[2][PER];
    Gathered<PER, HOT> gd[2];
    RowMeta meta[2];
    const TileDesc t0 = get_desc(db, b0);
    load_ids<PER>(a, t0, lane, idx[0]);
    load_meta(a, c, t0, lane, meta[0]);
    if (b0 + 1 < b1) load_ids<PER>(a, get_desc(db, b0 + 1), lane, idx[1]);
    gather_ids<PER, HOT>(a, t0, hot, H, lane, idx[0], gd[0]);
    uint32_t b = b0;
    for (; b + 1 < b1; b += 2) {
        wave_prio(bend - b, total);
        wave_step<PER, HOT, 0>(a, c, hot, H, st, lane, b, b1, db, idx, gd, meta);
        wave_step<PER, HOT, 1>(a, c, hot, H, st, lane, b + 1, b1, db, idx, gd, meta);
    }
    if (b < b1) wave_step<PER, HOT, 0>(a, c, hot, H, st, lane, b, b1, db, idx, gd, meta);
}

template <int T, int PER, bool HOT>
__global__ __launch_bounds__(T) void pr_sweep_wave_kernel(PrArgs a, PrConsts c, uint32_t H) {
    constexpr int NW = T / kWave;
    extern __shared__ float lds[];  // H + NW * PER*64 floats
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    float *hot = lds;
    float *st = lds + H + (size_t)wave * PER * kWave;
    if (HOT) {
        fill_hot(hot, a.xg, H, tid, T);
    }
    __syncthreads();
    const uint32_t nwv = gridDim.x * NW;
    const uint32_t gw = blockIdx.x * NW + wave;
    const uint64_t per = ((uint64_t)a.nblk + nwv - 1) / nwv;
    const uint32_t b0 = (uint32_t)min<uint64_t>(a.nblk, per * gw);
    const uint32_t b1 = (uint32_t)min<uint64_t>(a.nblk, per * gw + per);
    const uint32_t nch = gw < a.nchunk ? (a.nchunk - gw + nwv - 1) / nwv : 0u;
    wave_chunks<HOT>(a, hot, H, lane, gw, nwv, b1 - b0, nch + (b1 - b0));
    for (uint32_t p = b0; p < b1; p += kWave)
        wave_blocks<PER, HOT>(a, c, hot, H, st, lane, p, min(b1, p + (uint32_t)kWave), b1, nch + (b1 - b0));
}

// Gather-layout refresh: xg[i] = x[order[i]] for the A layout positions.
__global__ void pr_gather_layout_kernel(const uint32_t *__restrict__ order, uint32_t A, const float *__restrict__ x,
                                        float *__restrict__ xg) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < A; i += stride) xg[i] = x[order[i]];
}

// Hub rows' in-edge sums: each row's 512-edge chunk partials added one after
// another in chunk order (every launch form and both finalize forms use this
// order).  Lane slot i of the finalize handles hub hub_of_slot(i): a bijection
// of [0, nhub) (multiplication by a prime above nhub, mod nhub) that scatters
// the long rows -- R-MAT's hubs have the lowest ids, so consecutive hubs are
// the longest -- over many wavefronts.  Rows of up to kHubLaneChunks chunks
// are summed by their own lane, 16 partial loads in flight; longer ones (a
// row of 10^5 in-edges has hundreds of chunks) by the whole wavefront, whose
// lanes load 64 partials at a time while every lane adds them in order: the
// same float additions in the same order.  Call hub_sums_wave with all lanes
// of the wavefront.
constexpr uint32_t kHubLaneChunks = 32;
constexpr uint32_t kFinStatAhead = 4;  // block statistics per finalize thread loaded up front
constexpr uint32_t kHubLaneBatch = 16;
constexpr uint64_t kHubSlotPrime = 4294967291ull;  // largest prime below 2^32

GCHI_DEV uint32_t hub_of_slot(uint32_t i, uint32_t nhub) { return (uint32_t)(((uint64_t)i * kHubSlotPrime) % nhub); }

GCHI_DEV float hub_sums_wave(const float *__restrict__ chunk_part, const uint32_t *__restrict__ hub_chunk_off,
                             uint32_t h, bool valid, int lane) {
    uint32_t c0 = 0, c1 = 0;
    if (valid) {
        c0 = hub_chunk_off[h];
        c1 = hub_chunk_off[h + 1];
    }
    const bool big = c1 - c0 > kHubLaneChunks;
    float mine = 0.0f;
    if (!big)
        for (uint32_t k0 = c0; k0 < c1; k0 += kHubLaneBatch) {  // loads in flight, then the adds in order
            float v[kHubLaneBatch];
#pragma unroll
            for (uint32_t q = 0; q < kHubLaneBatch; q++) v[q] = k0 + q < c1 ? chunk_part[k0 + q] : 0.0f;
#pragma unroll
            for (uint32_t q = 0; q < kHubLaneBatch; q++)
                if (k0 + q < c1) mine += v[q];
        }
    uint64_t m = __ballot(big);
    while (m) {
        const int j = __builtin_ctzll(m);
        m &= m - 1;
        const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)c0, j);
        const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)c1, j);
        float s = 0.0f;
        for (uint32_t k0 = b0; k0 < b1; k0 += kWave) {
            const uint32_t k = k0 + lane;
            const float v = k < b1 ? chunk_part[k] : 0.0f;
            const uint32_t cnt = min((uint32_t)kWave, b1 - k0);
            for (uint32_t q = 0; q < cnt; q++) s += __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)q));
        }
        if (lane == j) mine = s;
    }
    return mine;
}

// Single-graph finalize: hub rows (sum of chunk partials, apply) + the
// deterministic reduction of per-block statistics; the last workgroup writes
// d_stats and resets the counters for the next sweep / graph replay.
__global__ __launch_bounds__(kFinThreads) void pr_finalize_kernel(
    const uint32_t *__restrict__ hub_v, const uint32_t *__restrict__ hub_chunk_off,
    const float *__restrict__ chunk_part, const uint32_t *__restrict__ outdeg, const float *__restrict__ r_in,
    float *__restrict__ r_out, float *__restrict__ x_out, const double *__restrict__ blk_stat, uint32_t nblk,
    uint32_t nhub, double *fin_stat, uint32_t *counters, double *d_stats, PrConsts c) {
    __shared__ double s_red[2 * (kFinThreads / kWave)];
    __shared__ bool s_last;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    double res = 0.0, sum = 0.0;
    const uint32_t gt = blockIdx.x * kFinThreads + tid;
    const uint32_t nthreads_total = gridDim.x * kFinThreads;
    // a lane per hub row: slot k*T + lane*(T/64) + w for wavefront w, so that
    // even a few hundred hubs spread over all wavefronts
    // The first kFinStatAhead block statistics of this thread are loaded before
    // the hub rows and added after them, as the loop below adds the rest: the
    // order of additions is the same either way.
    double bs[2 * kFinStatAhead];
#pragma unroll
    for (uint32_t q = 0; q < kFinStatAhead; q++) {
        const uint32_t i = min(gt + q * nthreads_total, nblk ? nblk - 1 : 0u);
        bs[2 * q] = nblk ? blk_stat[2 * i] : 0.0;
        bs[2 * q + 1] = nblk ? blk_stat[2 * i + 1] : 0.0;
    }
    const uint32_t nwv = nthreads_total / kWave, wv = gt / kWave;
    for (uint32_t base = wv; base < nhub; base += nthreads_total) {
        const uint32_t slot = base + lane * nwv;
        const bool valid = slot < nhub;
        const uint32_t h = valid ? hub_of_slot(slot, nhub) : 0u;
        // the row's own inputs are loaded beside its chunk partials
        const uint32_t v = hub_v[h];
        const float ro = r_in[v];
        const uint32_t od = outdeg[v];
        const float hs = hub_sums_wave(chunk_part, hub_chunk_off, h, valid, lane);
        if (valid) {
            const float r = pr_apply(row_teleport(c, 0, v), c.damping, hs);
            r_out[v] = r;
            x_out[v] = pr_scatter(r, od);
            res += fabs((double)r - (double)ro);
            sum += (double)r;
        }
    }
#pragma unroll
    for (uint32_t q = 0; q < kFinStatAhead; q++) {
        if (gt + q * nthreads_total < nblk) {
            res += bs[2 * q];
            sum += bs[2 * q + 1];
        }
    }
    for (uint32_t i = gt + kFinStatAhead * nthreads_total; i < nblk; i += nthreads_total) {
        res += blk_stat[2 * i];
        sum += blk_stat[2 * i + 1];
    }
    res = wave_sum(res);
    sum = wave_sum(sum);
    if (lane == 0) {
        s_red[2 * wave] = res;
        s_red[2 * wave + 1] = sum;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, s2 = 0.0;
        for (int w = 0; w < kFinThreads / kWave; w++) {
            a += s_red[2 * w];
            s2 += s_red[2 * w + 1];
        }
        st_relaxed_agent(&fin_stat[2 * blockIdx.x], a);
        st_relaxed_agent(&fin_stat[2 * blockIdx.x + 1], s2);
        __threadfence();
        uint32_t ticket = atomicAdd(&counters[0], 1u);
        s_last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (s_last) {  // block-uniform
        __threadfence();
        double a = 0.0, s2 = 0.0;
        for (uint32_t j = tid; j < gridDim.x; j += kFinThreads) {
            a += ld_relaxed_agent(&fin_stat[2 * j]);
            s2 += ld_relaxed_agent(&fin_stat[2 * j + 1]);
        }
        a = wave_sum(a);
        s2 = wave_sum(s2);
        if (lane == 0) {
            s_red[2 * wave] = a;
            s_red[2 * wave + 1] = s2;
        }
        __syncthreads();
        if (tid == 0) {
            double ta = 0.0, ts = 0.0;
            for (int w = 0; w < kFinThreads / kWave; w++) {
                ta += s_red[2 * w];
                ts += s_red[2 * w + 1];
            }
            d_stats[0] = ta;
            d_stats[1] = ts;
            counters[0] = 0u;
        }
        pr_reset_stripes(counters);
    }
}

// Multi-graph finalize, step 1: hub rows (lane per hub), per-hub statistics.
__global__ __launch_bounds__(kFinThreads) void pr_hub_apply_kernel(
    const uint32_t *__restrict__ hub_v, const uint32_t *__restrict__ hub_graph,
    const uint32_t *__restrict__ hub_chunk_off, const float *__restrict__ chunk_part,
    const uint32_t *__restrict__ outdeg, const float *__restrict__ r_in, float *__restrict__ r_out,
    float *__restrict__ x_out, uint32_t nhub, double *hub_stat, PrConsts c) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t stride = gridDim.x * kFinThreads;
    // a lane per hub row, slots spread over all wavefronts as in the single-graph finalize
    const uint32_t nwv = stride / kWave, wv = (blockIdx.x * kFinThreads + threadIdx.x) / kWave;
    for (uint32_t base = wv; base < nhub; base += stride) {
        const uint32_t slot = base + lane * nwv;
        const bool valid = slot < nhub;
        const uint32_t h = valid ? hub_of_slot(slot, nhub) : 0u;
        const float hs = hub_sums_wave(chunk_part, hub_chunk_off, h, valid, lane);
        if (valid) {
            double res = 0.0, sum = 0.0;
            pr_finish_row(hub_v[h], hub_graph[h], hs, c, r_in, outdeg, r_out, x_out, res, sum);
            hub_stat[2 * h] = res;
            hub_stat[2 * h + 1] = sum;
        }
    }
}

// Multi-graph finalize, step 2: one workgroup per graph reduces its blocks'
// and hubs' statistics in a fixed order into d_stats[2g], d_stats[2g+1].
__global__ __launch_bounds__(kFinThreads) void pr_graph_reduce_kernel(const uint32_t *__restrict__ gblk,
                                                                      const uint32_t *__restrict__ ghub,
                                                                      const double *__restrict__ blk_stat,
                                                                      const double *__restrict__ hub_stat,
                                                                      uint32_t G, uint32_t *counters,
                                                                      double *d_stats) {
    __shared__ double s_red[2 * (kFinThreads / kWave)];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    for (uint32_t g = blockIdx.x; g < G; g += gridDim.x) {
        double res = 0.0, sum = 0.0;
        for (uint32_t i = gblk[g] + tid; i < gblk[g + 1]; i += kFinThreads) {
            res += blk_stat[2 * i];
            sum += blk_stat[2 * i + 1];
        }
        for (uint32_t h = ghub[g] + tid; h < ghub[g + 1]; h += kFinThreads) {
            res += hub_stat[2 * h];
            sum += hub_stat[2 * h + 1];
        }
        res = wave_sum(res);
        sum = wave_sum(sum);
        if (lane == 0) {
            s_red[2 * wave] = res;
            s_red[2 * wave + 1] = sum;
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0.0, s2 = 0.0;
            for (int w = 0; w < kFinThreads / kWave; w++) {
                a += s_red[2 * w];
                s2 += s_red[2 * w + 1];
            }
            d_stats[2 * g] = a;
            d_stats[2 * g + 1] = s2;
        }
        __syncthreads();
    }
    if (blockIdx.x == 0) pr_reset_stripes(counters);
}

__global__ void pr_init_kernel(const uint32_t *outdeg, uint32_t n, float init, const uint32_t *seed_bits,
                               float seed_value, float *x, float *r) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        float r0 = init;
        if (seed_bits) r0 = ((seed_bits[v >> 5] >> (v & 31)) & 1u) ? seed_value : 0.0f;
        r[v] = r0;
        x[v] = pr_scatter(r0, outdeg[v]);
    }
}

// ---------------------------------------------------------------- plan build
GCHI_DEV uint32_t graph_of(const uint32_t *vbase, uint32_t G, uint32_t v) {
    uint32_t lo = 0, hi = G - 1;  // largest g with vbase[g] <= v
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo + 1) / 2;
        if (vbase[mid] <= v) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The plan build reads every graph's in-edges through one view: graph g owns
// vertices [vbase[g], vbase[g+1]) and offset slots [vbase[g]+g, vbase[g+1]+g]
// of off_ext, which hold its in-edge offsets relative to its first in-edge at
// address g_src[g]; a stored source id plus g_idoff[g] is the plan vertex id.
struct CscView {
    const uint32_t *off_ext;  // [N+G]
    const uint64_t *g_src;    // [G] device addresses of each graph's in_src
    const uint32_t *g_idoff;  // [G] added to stored ids (vbase for per-graph ids, 0 for a union)
    const uint32_t *vbase;    // [G+1]
    uint32_t G, N;
    int local_ids;            // ids are per graph (checked against its vertex count)
    uint32_t src_lim;         // bound of stored ids for a union CSC (N, or more for a row slice)
};

GCHI_DEV uint32_t view_deg(const CscView &cv, uint32_t v, uint32_t g) {
    return cv.off_ext[v + g + 1] - cv.off_ext[v + g];
}

// off_ext of a union CSC (global offsets, global ids): slot v+g = in_off[v] -
// in_off[vbase[g]] for v in [vbase[g], vbase[g+1]]; g_src[g] = in_src + in_off[vbase[g]].
__global__ void plan_ext_union(const uint32_t *in_off, const uint32_t *in_src, const uint32_t *vbase, uint32_t G,
                               uint32_t N, uint32_t *off_ext, uint64_t *g_src, uint32_t *g_idoff) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) {
        const uint32_t g = graph_of(vbase, G, v);
        off_ext[v + g] = in_off[v] - in_off[vbase[g]];
    }
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += stride) {
        off_ext[vbase[g + 1] + g] = in_off[vbase[g + 1]] - in_off[vbase[g]];
        g_src[g] = (uint64_t)(uintptr_t)(in_src + in_off[vbase[g]]);
        g_idoff[g] = 0u;
    }
}

__global__ void plan_rows(CscView cv, int W, uint32_t *reg_deg, uint32_t *hub_flag, uint32_t *hub_bits) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < cv.N; v += stride) {
        const uint32_t d = view_deg(cv, v, graph_of(cv.vbase, cv.G, v));
        bool hub = d >= (uint32_t)W;
        reg_deg[v] = hub ? 0u : d;
        hub_flag[v] = hub ? 1u : 0u;
        if (hub) atomicOr(&hub_bits[v >> 5], 1u << (v & 31));
    }
}

__global__ void plan_hubs(CscView cv, const uint32_t *hub_flag, const uint32_t *hub_idx, uint32_t *hub_v,
                          uint32_t *hub_deg, uint32_t *hub_nchunk) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < cv.N; v += stride) {
        if (hub_flag[v]) {
            uint32_t h = hub_idx[v];
            uint32_t d = view_deg(cv, v, graph_of(cv.vbase, cv.G, v));
            hub_v[h] = v;
            hub_deg[h] = d;
            hub_nchunk[h] = (d + kHubChunk - 1) / kHubChunk;
        }
    }
}

// regular offsets relative to each graph (slot v+g, plus the end slot of each
// graph) and each graph's first regular edge, from the global u64 scan reg64[N+1]
__global__ void plan_reg_rel(const uint64_t *reg64, const uint32_t *vbase, uint32_t G, uint32_t N, uint32_t *reg_rel,
                             uint64_t *ebase) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) {
        const uint32_t g = graph_of(vbase, G, v);
        reg_rel[v + g] = (uint32_t)(reg64[v] - reg64[vbase[g]]);
    }
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += stride) {
        reg_rel[vbase[g + 1] + g] = (uint32_t)(reg64[vbase[g + 1]] - reg64[vbase[g]]);
        ebase[g] = reg64[vbase[g]];
    }
}

// Copies the rows' in-edges into the regular and hub arrays, mapped to
// gather-layout positions when grank is given.  A source id outside its graph,
// or one without a layout position (its out-degree says 0), sets *bad and is
// stored as 0, so no sweep ever gathers out of bounds; the plan is refused.
// Regular rows (fewer than W <= 4096 in-edges) take one wavefront each; hub
// rows are copied per 512-edge chunk, one wavefront per chunk, so that a row
// of 10^5 or more in-edges does not hold up the whole copy.
GCHI_DEV uint32_t plan_map_src(uint32_t u, uint32_t idoff, uint32_t lim, const uint32_t *grank, uint32_t *bad) {
    bool ok = u < lim;
    u += idoff;
    if (ok && grank) {
        u = grank[u];
        ok = u != 0xffffffffu;
    }
    if (!ok) {
        atomicOr(bad, 1u);
        u = 0;
    }
    return u;
}

__global__ void plan_copy(CscView cv, const uint32_t *hub_flag, const uint32_t *reg_rel, const uint64_t *ebase,
                          const uint32_t *grank, uint32_t *reg_src, uint32_t *bad) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    for (uint64_t v = wid; v < cv.N; v += nw) {
        if (hub_flag[v]) continue;
        const uint32_t g = graph_of(cv.vbase, cv.G, (uint32_t)v);
        const uint32_t b = cv.off_ext[v + g], e = cv.off_ext[v + g + 1];
        const uint32_t *src = reinterpret_cast<const uint32_t *>((uintptr_t)cv.g_src[g]);
        const uint32_t idoff = cv.g_idoff[g];
        const uint32_t lim = cv.local_ids ? cv.vbase[g + 1] - cv.vbase[g] : cv.src_lim;
        uint32_t *dst = reg_src + ebase[g] + reg_rel[v + g];
        for (uint32_t k = b + lane; k < e; k += kWave) dst[k - b] = plan_map_src(src[k], idoff, lim, grank, bad);
    }
}

__global__ void plan_copy_hubs(CscView cv, const uint32_t *hub_v, const uint32_t *hub_chunk_off, uint32_t nhub,
                               uint32_t nchunk, const uint64_t *chunk_beg, const uint32_t *chunk_len,
                               const uint32_t *grank, uint32_t *hub_src, uint32_t *bad) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint32_t wid = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint32_t nw = (gridDim.x * blockDim.x) / kWave;
    for (uint32_t c = wid; c < nchunk; c += nw) {
        uint32_t lo = 0, hi = nhub - 1;  // the hub of chunk c: largest h with hub_chunk_off[h] <= c
        while (lo < hi) {
            const uint32_t mid = lo + (hi - lo + 1) / 2;
            if (hub_chunk_off[mid] <= c) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t v = hub_v[lo], j = c - hub_chunk_off[lo];
        const uint32_t g = graph_of(cv.vbase, cv.G, v);
        const uint32_t b = cv.off_ext[v + g] + j * kHubChunk;
        const uint32_t *src = reinterpret_cast<const uint32_t *>((uintptr_t)cv.g_src[g]) + b;
        const uint32_t idoff = cv.g_idoff[g];
        const uint32_t lim = cv.local_ids ? cv.vbase[g + 1] - cv.vbase[g] : cv.src_lim;
        uint32_t *dst = hub_src + chunk_beg[c];
        const uint32_t len = chunk_len[c];
        for (uint32_t k = lane; k < len; k += kWave) dst[k] = plan_map_src(src[k], idoff, lim, grank, bad);
    }
}

// gather-layout keys: ascending order = out-degree descending, then id
__global__ void plan_order_keys(const uint32_t *outdeg, uint32_t N, uint64_t *keys) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride)
        keys[v] = ((uint64_t)(~outdeg[v]) << 32) | v;
}

// layout position of every vertex with out-degree > 0; *A = their count
__global__ void plan_gather_rank(const uint64_t *keys, uint32_t N, uint32_t *order, uint32_t *grank, uint32_t *A) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += stride) {
        const uint32_t v = (uint32_t)keys[i];
        const uint32_t od = ~(uint32_t)(keys[i] >> 32);
        if (od == 0) continue;
        order[i] = v;
        grank[v] = i;
        if (i + 1 == N || ~(uint32_t)(keys[i + 1] >> 32) == 0u) *A = i + 1;
    }
}

__global__ void plan_chunks(const uint64_t *hub_off, const uint32_t *hub_chunk_off, uint32_t nhub, uint64_t *chunk_beg,
                            uint32_t *chunk_len) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nhub; h += stride) {
        const uint64_t b = hub_off[h];
        const uint32_t d = (uint32_t)(hub_off[h + 1] - b);
        uint32_t c0 = hub_chunk_off[h], c1 = hub_chunk_off[h + 1];
        for (uint32_t c = c0; c < c1; c++) {
            uint32_t j = c - c0;
            chunk_beg[c] = b + (uint64_t)j * kHubChunk;
            chunk_len[c] = min((uint32_t)kHubChunk, d - j * kHubChunk);
        }
    }
}

// blocks per graph: the last row's work position within its graph / W + 1
__global__ void plan_graph_blocks(const uint32_t *reg_rel, const uint32_t *vbase, uint32_t G, int W,
                                  uint32_t *nblk_g) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < G; g += stride) {
        uint32_t v0 = vbase[g], v1 = vbase[g + 1];
        uint64_t p_last = (uint64_t)reg_rel[v1 - 1 + g] + (v1 - 1 - v0);
        nblk_g[g] = (uint32_t)(p_last / (uint64_t)W + 1);
    }
}

__global__ void plan_blocks(const uint32_t *reg_rel, const uint32_t *vbase, uint32_t G, const uint32_t *gblk,
                            uint32_t N, int W, uint32_t nblk, uint32_t *blk_row, uint32_t *blk_graph) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) {
        const uint32_t g = graph_of(vbase, G, v);
        const uint32_t v0 = vbase[g];
        const uint64_t base = gblk[g];
        const uint64_t bv = base + ((uint64_t)reg_rel[v + g] + (v - v0)) / (uint64_t)W;
        int64_t bp = (int64_t)base - 1;
        if (v > v0) bp = (int64_t)(base + ((uint64_t)reg_rel[v - 1 + g] + (v - 1 - v0)) / (uint64_t)W);
        for (int64_t bb = bp + 1; bb <= (int64_t)bv; bb++) {
            blk_row[bb] = v;
            blk_graph[bb] = g;
        }
        if (v == N - 1) {
            blk_row[nblk] = N;
            blk_graph[nblk] = G - 1;
        }
    }
}

// first and end regular edge of every block, relative to its graph
__global__ void plan_tile_edges(const uint32_t *reg_rel, const uint32_t *blk_row, const uint32_t *blk_graph,
                                uint32_t nblk, uint32_t *tile_e0, uint32_t *tile_e1) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nblk; b += stride) {
        const uint32_t g = blk_graph[b];
        tile_e0[b] = reg_rel[blk_row[b] + g];
        tile_e1[b] = reg_rel[blk_row[b + 1] + g];
    }
}

__global__ void plan_hub_graphs(const uint32_t *hub_v, uint32_t nhub, const uint32_t *vbase, uint32_t G,
                                uint32_t *hub_graph) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nhub; h += stride)
        hub_graph[h] = graph_of(vbase, G, hub_v[h]);
}

// ghub[g] = first hub with vertex >= vbase[g]
__global__ void plan_graph_hubs(const uint32_t *hub_v, uint32_t nhub, const uint32_t *vbase, uint32_t G,
                                uint32_t *ghub) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g <= G; g += stride) {
        uint32_t target = vbase[g], lo = 0, hi = nhub;
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (hub_v[mid] >= target) hi = mid;
            else lo = mid + 1;
        }
        ghub[g] = g == G ? nhub : lo;
    }
}

// Temporary device allocations of the plan build, freed on every exit path.
struct TmpPool {
    std::vector<void *> ptrs;
    ~TmpPool() {
        for (void *q : ptrs) (void)hipFree(q);
    }
    template <typename T>
    T *alloc(size_t count) {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        ptrs.push_back(q);
        return static_cast<T *>(q);
    }
};

static int read_u32(const uint32_t *d, uint32_t *h, hipStream_t s) {
    GCHI_HIP(hipMemcpyAsync(h, d, 4, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    return GCHI_OK;
}

static int read_u64(const uint64_t *d, uint64_t *h, hipStream_t s) {
    GCHI_HIP(hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    return GCHI_OK;
}

// The graphs a plan is built from: one CSC whose vertex ranges are the graphs
// (global ids and offsets), or one CSC per graph (local ids and offsets).
struct PlanInput {
    uint32_t N = 0, G = 1;
    uint64_t E = 0;
    const uint32_t *h_vbase = nullptr;  // [G+1]
    const gchi_dgraph *unionc = nullptr;
    const gchi_dgraph *graphs = nullptr;  // [G]
    uint32_t src_limit = 0;  // single-graph plans over a slice of rows: sources may be any id < src_limit
};

static int plan_build(const PlanInput &in, int W, gchi_pr_plan *p, hipStream_t s) {
    const uint32_t N = in.N;
    const uint64_t E = in.E;
    const uint32_t G = in.G;
    TmpPool tp;
    const size_t nwords = (N + 31) / 32;
    uint32_t *off_ext = tp.alloc<uint32_t>((size_t)N + G);
    uint64_t *g_src = tp.alloc<uint64_t>(G);
    uint32_t *g_idoff = tp.alloc<uint32_t>(G);
    uint32_t *reg_deg = tp.alloc<uint32_t>(N);
    uint32_t *hub_flag = tp.alloc<uint32_t>(N);
    uint32_t *hub_idx = tp.alloc<uint32_t>(N + 1);
    uint64_t *reg64 = tp.alloc<uint64_t>(N + 1);
    uint32_t *reg_rel = tp.alloc<uint32_t>((size_t)N + G);
    uint64_t *ebase = tp.alloc<uint64_t>(G);
    uint32_t *bits = tp.alloc<uint32_t>(nwords);
    uint32_t *vbase = tp.alloc<uint32_t>(G + 1);
    uint32_t *nblk_g = tp.alloc<uint32_t>(G + 1);
    uint32_t *gblk = tp.alloc<uint32_t>(G + 1);
    const size_t swb = scan_workspace((size_t)std::max(N, G) + 1);
    void *sw = tp.alloc<char>(swb);
    GCHI_CHECK(off_ext && g_src && g_idoff && reg_deg && hub_flag && hub_idx && reg64 && reg_rel && ebase && bits &&
                   vbase && nblk_g && gblk && sw,
               GCHI_ERR_HIP, "pr_plan: hipMalloc failed");
    GCHI_HIP(hipMemcpyAsync(vbase, in.h_vbase, 4 * (size_t)(G + 1), hipMemcpyHostToDevice, s));
    if (in.unionc) {
        p->outdeg = in.unionc->outdeg;
        hipLaunchKernelGGL(plan_ext_union, dim3(grid_for(std::max(N, G), 256, 65536)), dim3(256), 0, s,
                           in.unionc->in_off, in.unionc->in_src, vbase, G, N, off_ext, g_src, g_idoff);
        GCHI_LAUNCH_CHECK();
    } else {
        // per-graph CSCs: offsets and out-degrees copied side by side, ids offset by vbase
        GCHI_HIP(hipMalloc(&p->own_outdeg, (size_t)std::max<uint32_t>(N, 1) * 4));
        p->outdeg = p->own_outdeg;
        std::vector<uint64_t> h_src(G);
        for (uint32_t k = 0; k < G; k++) {
            const gchi_dgraph &gk = in.graphs[k];
            const uint32_t nk = gk.nvertices;
            GCHI_HIP(hipMemcpyAsync(off_ext + in.h_vbase[k] + k, gk.in_off, (size_t)(nk + 1) * 4,
                                    hipMemcpyDeviceToDevice, s));
            GCHI_HIP(hipMemcpyAsync(p->own_outdeg + in.h_vbase[k], gk.outdeg, (size_t)nk * 4,
                                    hipMemcpyDeviceToDevice, s));
            h_src[k] = (uint64_t)(uintptr_t)gk.in_src;
        }
        GCHI_HIP(hipMemcpyAsync(g_src, h_src.data(), 8 * (size_t)G, hipMemcpyHostToDevice, s));
        GCHI_HIP(hipMemcpyAsync(g_idoff, in.h_vbase, 4 * (size_t)G, hipMemcpyHostToDevice, s));
        GCHI_HIP(hipStreamSynchronize(s));  // h_src is a host temporary
    }
    const CscView cv{off_ext, g_src, g_idoff, vbase, G, N, in.unionc ? 0 : 1, in.src_limit ? in.src_limit : N};
    launch_fill_u32(bits, 0u, nwords, s);
    hipLaunchKernelGGL(plan_rows, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, cv, W, reg_deg, hub_flag, bits);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(exclusive_scan_u32_to_u64(reg_deg, reg64, N, sw, swb, s));
    GCHI_TRY(exclusive_scan_u32(hub_flag, hub_idx, N, sw, swb, s));
    hipLaunchKernelGGL(plan_reg_rel, dim3(grid_for(std::max(N, G), 256, 65536)), dim3(256), 0, s, reg64, vbase, G, N,
                       reg_rel, ebase);
    GCHI_LAUNCH_CHECK();
    hipLaunchKernelGGL(plan_graph_blocks, dim3(grid_for(G, 256, 4096)), dim3(256), 0, s, reg_rel, vbase, G, W, nblk_g);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(exclusive_scan_u32(nblk_g, gblk, G, sw, swb, s));
    uint64_t e_reg = 0;
    uint32_t nhub = 0, nblk = 0;
    GCHI_TRY(read_u64(reg64 + N, &e_reg, s));
    GCHI_TRY(read_u32(hub_idx + N, &nhub, s));
    GCHI_TRY(read_u32(gblk + G, &nblk, s));
    p->e_reg = e_reg;
    p->nhub = nhub;
    p->nblk = nblk;

    uint32_t *hub_v = tp.alloc<uint32_t>(nhub + 1);
    uint32_t *hub_deg = tp.alloc<uint32_t>(nhub + 1);
    uint32_t *hub_nch = tp.alloc<uint32_t>(nhub + 1);
    uint64_t *hub_off = tp.alloc<uint64_t>(nhub + 1);
    uint32_t *hub_choff = tp.alloc<uint32_t>(nhub + 1);
    const size_t swb2 = scan_workspace((size_t)nhub + 1);
    void *sw2 = tp.alloc<char>(swb2);
    GCHI_CHECK(hub_v && hub_deg && hub_nch && hub_off && hub_choff && sw2, GCHI_ERR_HIP, "pr_plan: hipMalloc failed");
    if (nhub) {
        hipLaunchKernelGGL(plan_hubs, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, cv, hub_flag, hub_idx, hub_v,
                           hub_deg, hub_nch);
        GCHI_LAUNCH_CHECK();
    }
    GCHI_TRY(exclusive_scan_u32_to_u64(hub_deg, hub_off, nhub, sw2, swb2, s));
    GCHI_TRY(exclusive_scan_u32(hub_nch, hub_choff, nhub, sw2, swb2, s));
    uint32_t nchunk = 0;
    uint64_t e_hub = 0;
    GCHI_TRY(read_u32(hub_choff + nhub, &nchunk, s));
    GCHI_TRY(read_u64(hub_off + nhub, &e_hub, s));
    p->e_hub = e_hub;
    p->nchunk = nchunk;
    GCHI_CHECK(e_reg + e_hub == E, GCHI_ERR_INVALID, "pr_plan_create: in-edge offsets do not add up to nedges");
    // one lane per hub row; at least 64 workgroups for the statistics.  Each
    // workgroup takes a ticket on one counter at its end, so fewer is better.
    p->nfin = std::min<uint32_t>(1024, std::max<uint32_t>(64, (nhub + kFinThreads - 1) / kFinThreads));

    // gather layout: sources ranked by out-degree (descending, ties by id)
    uint32_t *grank = nullptr, *gorder_tmp = nullptr, *bad = tp.alloc<uint32_t>(1);
    GCHI_CHECK(bad, GCHI_ERR_HIP, "pr_plan: hipMalloc failed");
    launch_fill_u32(bad, 0u, 1, s);
    uint32_t A = 0;
    if (p->hot) {
        uint64_t *keys = tp.alloc<uint64_t>(N);
        grank = tp.alloc<uint32_t>(N);
        gorder_tmp = tp.alloc<uint32_t>(N);
        uint32_t *dA = tp.alloc<uint32_t>(1);
        size_t sb = 0;
        GCHI_TRY(gchi_sort_workspace(N, 64, 0, &sb));
        void *sws = tp.alloc<char>(sb);
        GCHI_CHECK(keys && grank && gorder_tmp && dA && sws, GCHI_ERR_HIP, "pr_plan: hipMalloc failed");
        launch_fill_u32(grank, 0xffffffffu, N, s);
        launch_fill_u32(dA, 0u, 1, s);
        hipLaunchKernelGGL(plan_order_keys, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, p->outdeg, N, keys);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(gchi_sort_u64(keys, nullptr, N, 64, sws, sb, s));
        GCHI_TRY(read_work_status(static_cast<const uint32_t *>(sws), s));
        hipLaunchKernelGGL(plan_gather_rank, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, keys, N, gorder_tmp,
                           grank, dA);
        GCHI_LAUNCH_CHECK();
        GCHI_TRY(read_u32(dA, &A, s));
    }
    p->A = A;

    Sizer zz;
    zz.take<uint32_t>((size_t)N + G);
    zz.take<uint32_t>(std::max<uint64_t>(p->e_reg, 1));
    zz.take<uint32_t>(nwords);
    zz.take<uint32_t>(p->nblk + 1);
    zz.take<uint32_t>(p->nblk + 1);
    zz.take<uint32_t>(nhub + 1);
    zz.take<uint32_t>(nhub + 1);
    zz.take<uint32_t>(nhub + 1);
    zz.take<uint64_t>(nchunk + 1);
    zz.take<uint32_t>(nchunk + 1);
    zz.take<uint32_t>(std::max<uint64_t>(p->e_hub, 1));
    zz.take<uint32_t>(G + 1);
    zz.take<uint32_t>(G + 1);
    zz.take<uint32_t>(G + 1);
    zz.take<float>(nchunk + 1);
    zz.take<double>(2 * (size_t)p->nblk);
    zz.take<double>(2 * (size_t)nhub + 2);
    zz.take<double>(2 * (size_t)p->nfin);
    zz.take<uint32_t>(kPrCounterWords);
    zz.take<uint32_t>(std::max<uint32_t>(A, 1));
    zz.take<float>(std::max<uint32_t>(A, 1));
    zz.take<uint32_t>(p->nblk + 1);
    zz.take<uint32_t>(p->nblk + 1);
    zz.take<uint64_t>(G);
    GCHI_HIP(hipMalloc(&p->mem, zz.off));
    Carver c(p->mem, zz.off);
    p->reg_off = c.take<uint32_t>((size_t)N + G);
    p->reg_src = c.take<uint32_t>(std::max<uint64_t>(p->e_reg, 1));
    p->hub_bits = c.take<uint32_t>(nwords);
    p->blk_row = c.take<uint32_t>(p->nblk + 1);
    p->blk_graph = c.take<uint32_t>(p->nblk + 1);
    p->hub_v = c.take<uint32_t>(nhub + 1);
    p->hub_graph = c.take<uint32_t>(nhub + 1);
    p->hub_chunk_off = c.take<uint32_t>(nhub + 1);
    p->chunk_beg = c.take<uint64_t>(nchunk + 1);
    p->chunk_len = c.take<uint32_t>(nchunk + 1);
    p->hub_src = c.take<uint32_t>(std::max<uint64_t>(p->e_hub, 1));
    p->gblk = c.take<uint32_t>(G + 1);
    p->ghub = c.take<uint32_t>(G + 1);
    p->vbase = c.take<uint32_t>(G + 1);
    p->chunk_part = c.take<float>(nchunk + 1);
    p->blk_stat = c.take<double>(2 * (size_t)p->nblk);
    p->hub_stat = c.take<double>(2 * (size_t)nhub + 2);
    p->fin_stat = c.take<double>(2 * (size_t)p->nfin);
    p->counters = c.take<uint32_t>(kPrCounterWords);
    p->gorder = c.take<uint32_t>(std::max<uint32_t>(A, 1));
    p->xg = c.take<float>(std::max<uint32_t>(A, 1));
    p->tile_e0 = c.take<uint32_t>(p->nblk + 1);
    p->tile_e1 = c.take<uint32_t>(p->nblk + 1);
    p->ebase = c.take<uint64_t>(G);
    if (A) GCHI_HIP(hipMemcpyAsync(p->gorder, gorder_tmp, (size_t)A * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->reg_off, reg_rel, ((size_t)N + G) * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->ebase, ebase, (size_t)G * 8, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->hub_bits, bits, nwords * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->hub_v, hub_v, (size_t)(nhub + 1) * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->hub_chunk_off, hub_choff, (size_t)(nhub + 1) * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->gblk, gblk, (size_t)(G + 1) * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->vbase, vbase, (size_t)(G + 1) * 4, hipMemcpyDeviceToDevice, s));
    launch_fill_u32(p->counters, 0u, kPrCounterWords, s);
    if (E)
        hipLaunchKernelGGL(plan_copy, dim3(grid_for((uint64_t)N * kWave, 256, 1u << 16)), dim3(256), 0, s, cv,
                           hub_flag, p->reg_off, p->ebase, grank, p->reg_src, bad);
    if (nhub) {
        hipLaunchKernelGGL(plan_chunks, dim3(grid_for(nhub, 256, 4096)), dim3(256), 0, s, hub_off, hub_choff, nhub,
                           p->chunk_beg, p->chunk_len);
        hipLaunchKernelGGL(plan_copy_hubs, dim3(grid_for((uint64_t)nchunk * kWave, 256, 1u << 16)), dim3(256), 0, s,
                           cv, hub_v, hub_choff, nhub, nchunk, p->chunk_beg, p->chunk_len, grank, p->hub_src, bad);
        hipLaunchKernelGGL(plan_hub_graphs, dim3(grid_for(nhub, 256, 4096)), dim3(256), 0, s, p->hub_v, nhub,
                           p->vbase, G, p->hub_graph);
    }
    hipLaunchKernelGGL(plan_graph_hubs, dim3(grid_for(G + 1, 256, 4096)), dim3(256), 0, s, p->hub_v, nhub, p->vbase, G,
                       p->ghub);
    hipLaunchKernelGGL(plan_blocks, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, p->reg_off, p->vbase, G, p->gblk, N,
                       W, p->nblk, p->blk_row, p->blk_graph);
    GCHI_LAUNCH_CHECK();
    hipLaunchKernelGGL(plan_tile_edges, dim3(grid_for(p->nblk, 256, 65536)), dim3(256), 0, s, p->reg_off, p->blk_row,
                       p->blk_graph, p->nblk, p->tile_e0, p->tile_e1);
    GCHI_LAUNCH_CHECK();
    uint32_t h_bad = 0;
    GCHI_TRY(read_u32(bad, &h_bad, s));
    GCHI_CHECK(!h_bad, GCHI_ERR_INVALID,
               "pr_plan_create: an in-edge source is outside its graph or has out-degree 0 (in_src and outdeg "
               "disagree)");
    return GCHI_OK;
}

static int plan_create(const PlanInput &in, const gchi_pr_tune *tune, gchi_pr_plan **out, hipStream_t s) {
    const uint32_t G = in.G;
    const uint32_t *h_vbase = in.h_vbase;
    GCHI_CHECK(out && h_vbase, GCHI_ERR_INVALID, "pr_plan_create: null argument");
    GCHI_CHECK(in.N > 0 && G >= 1, GCHI_ERR_INVALID, "pr_plan_create: graph has no vertices");
    GCHI_CHECK(h_vbase[0] == 0 && h_vbase[G] == in.N, GCHI_ERR_INVALID,
               "pr_plan_create: vertex bases must start at 0 and end at N");
    for (uint32_t i = 0; i < G; i++)
        GCHI_CHECK(h_vbase[i + 1] > h_vbase[i], GCHI_ERR_INVALID, "pr_plan_create: every graph needs a vertex");
    const int W = (tune && tune->edges_per_block > 0) ? tune->edges_per_block : 2048;
    GCHI_CHECK(W >= 256 && W <= kMaxW && W % 256 == 0, GCHI_ERR_INVALID,
               "pr_plan_create: edges_per_block must be a multiple of 256 in [256, 4096]");
    const int short_max = (tune && tune->short_row_max > 0) ? tune->short_row_max : 32;
    GCHI_CHECK(short_max >= 1 && short_max <= 1024, GCHI_ERR_INVALID, "pr_plan_create: short_row_max out of range");
    GCHI_CHECK(2 * W / (short_max + 1) <= kMaxLong, GCHI_ERR_INVALID,
               "pr_plan_create: short_row_max too small for edges_per_block (more than 512 long rows per block)");
    gchi_pr_plan *p = new gchi_pr_plan();
    p->N = in.N;
    p->E = in.E;
    p->G = G;
    p->W = W;
    p->short_max = short_max;
    p->persistent = tune ? tune->persistent : 0;
    p->hot = tune ? tune->hot_vertices : 0;
    if (p->persistent < 0 || p->persistent > 4) {
        delete p;
        return fail(GCHI_ERR_INVALID, "pr_plan_create: persistent must be 0, 1, 2, 3 or 4");
    }
    if (p->persistent == 4) {
        // the rank-space tiled sweep: one graph given as one CSC
        if (G != 1 || !in.unionc || in.src_limit) {
            delete p;
            return fail(GCHI_ERR_UNSUPPORTED, "pr_plan_create: launch form 4 takes one graph (gchi_pr_plan_create)");
        }
        p->outdeg = in.unionc->outdeg;
        const int st4 = seg_plan_create(in.unionc, p->hot, &p->seg, s);
        if (st4 != GCHI_OK) {
            delete p;
            return st4;
        }
        *out = p;
        return GCHI_OK;
    }
    // the LDS cache lives in the persistent kernels 2 and 3 (default 2)
    if (p->hot > 0 && p->persistent != 3) p->persistent = 2;
    if (p->hot) {
        if (p->hot < -1) {
            delete p;
            return fail(GCHI_ERR_INVALID, "pr_plan_create: hot_vertices must be -1, 0 or positive");
        }
        if (in.N >= (1u << 30)) {
            delete p;
            return fail(GCHI_ERR_UNSUPPORTED, "pr_plan_create: the gather layout needs N < 2^30");
        }
    }
    int st = plan_build(in, W, p, s);
    if (st == GCHI_OK) {
        int dev = 0, ncu = 256;
        size_t lds_cu = 160 * 1024, lds_wg = 160 * 1024;
        hipDeviceProp_t prop;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
            ncu = prop.multiProcessorCount;
            lds_cu = prop.maxSharedMemoryPerMultiProcessor;
            lds_wg = prop.sharedMemPerBlock;
        }
        const uint32_t nhubt = (p->nchunk + kPrWaves - 1) / kPrWaves;
        p->npersist = (uint32_t)std::min<uint64_t>((uint64_t)ncu * 8, (uint64_t)p->nblk + nhubt);
        if (p->persistent == 3) {
            // 256-thread workgroups, four per CU: each holds H cached values and
            // its 2W staging floats within a quarter of the CU's LDS
            const size_t cap = std::min(lds_cu / 4, lds_wg);
            const size_t stage = (size_t)8 * W;
            uint32_t H = p->hot > 0 ? std::min<uint32_t>((uint32_t)p->hot, p->A) / 1024 * 1024 : 0;
            const size_t hmax = cap > stage + kStaticLds ? (cap - stage - kStaticLds) / 4 / 1024 * 1024 : 0;
            H = std::min<uint32_t>(H, (uint32_t)hmax);
            p->H = H;
            p->hotwg_lds = stage + 4 * (size_t)H;
            p->nhot_wg = (uint32_t)std::min<uint64_t>((uint64_t)ncu * 4, (uint64_t)p->nblk + nhubt);
            hipError_t e = hipFuncSetAttribute((const void *)pr_sweep_hotwg_kernel,
                                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->hotwg_lds);
            if (e != hipSuccess) st = hip_fail(e, "hipFuncSetAttribute", __FILE__, __LINE__);
        }
        if (p->persistent == 2) {
            // one 1024-thread workgroup per CU: H cached layout values + a 2W-float
            // staging slice per wavefront + static LDS (bounded by kStaticLds)
            const size_t cap = std::min(lds_cu, lds_wg);
            const size_t stage = (size_t)kWaveWaves * 8 * W;
            uint32_t H = p->hot > 0 ? std::min<uint32_t>((uint32_t)p->hot, p->A) / 1024 * 1024 : 0;
            if (W != 256) {
                st = fail(GCHI_ERR_INVALID, "pr_plan_create: the wavefront-tiled sweep needs edges_per_block 256");
            } else if (stage + kStaticLds > cap) {
                st = fail(GCHI_ERR_INVALID, "pr_plan_create: edges_per_block too large for the wavefront-tiled sweep");
            } else {
                const size_t hmax = (cap - stage - kStaticLds) / 4 / 1024 * 1024;
                H = std::min<uint32_t>(H, (uint32_t)hmax);
                p->H = H;
                p->wave_lds = stage + 4 * (size_t)H;
                p->nwave_wg = (uint32_t)ncu;
                const void *fn = H ? (const void *)pr_sweep_wave_kernel<kWaveThreads, 8, true>
                                   : (const void *)pr_sweep_wave_kernel<kWaveThreads, 8, false>;
                hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->wave_lds);
                if (e != hipSuccess) st = hip_fail(e, "hipFuncSetAttribute", __FILE__, __LINE__);
            }
        }
    }
    if (st != GCHI_OK) {
        if (p->mem) (void)hipFree(p->mem);
        if (p->own_outdeg) (void)hipFree(p->own_outdeg);
        delete p;
        return st;
    }
    *out = p;
    return GCHI_OK;
}

int pr_init(const gchi_pr_plan *p, float init, const uint32_t *seed_bits, float seed_value, float *x, float *r,
            hipStream_t s) {
    hipLaunchKernelGGL(pr_init_kernel, dim3(grid_for(p->N, 256, 65536)), dim3(256), 0, s, p->outdeg, p->N, init,
                       seed_bits, seed_value, x, r);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

int pr_sweep(const gchi_pr_plan *p, const gchi_pr_op *op, const float *x_in, const float *r_in, float *x_out,
             float *r_out, double *d_stats, hipStream_t s) {
    if (p->seg) return seg_sweep(p->seg, op, x_in, r_in, x_out, r_out, d_stats, s);
    PrConsts c{op->teleport, op->damping, op->d_graph_teleport, op->d_seed_bits, op->seed_teleport};
    const float *xg = x_in;
    if (p->hot) {
        if (p->A)
            hipLaunchKernelGGL(pr_gather_layout_kernel, dim3(grid_for(p->A, 256, 65536)), dim3(256), 0, s, p->gorder,
                               p->A, x_in, p->xg);
        xg = p->xg;
    }
    const uint32_t nhubt = (p->nchunk + kPrWaves - 1) / kPrWaves;
    PrArgs a{p->reg_off, p->reg_src, p->blk_row, p->blk_graph, p->tile_e0, p->tile_e1, p->hub_bits, p->ebase,
             p->chunk_beg, p->chunk_len, p->hub_src,
             p->outdeg, xg, r_in, x_out, r_out, p->chunk_part, p->blk_stat, p->counters + kPrStripeGap,
             p->nchunk, nhubt, p->nblk, nhubt + p->nblk, p->short_max};
    const size_t shmem = (size_t)2 * p->W * sizeof(float);
    if (p->persistent == 3) {
        hipLaunchKernelGGL(pr_sweep_hotwg_kernel, dim3(p->nhot_wg), dim3(kPrThreads), p->hotwg_lds, s, a, c, p->H);
    } else if (p->persistent == 2) {
        const dim3 grid(p->nwave_wg), blk(kWaveThreads);
        if (p->H)
            hipLaunchKernelGGL((pr_sweep_wave_kernel<kWaveThreads, 8, true>), grid, blk, p->wave_lds, s, a, c, p->H);
        else
            hipLaunchKernelGGL((pr_sweep_wave_kernel<kWaveThreads, 8, false>), grid, blk, p->wave_lds, s, a, c, 0u);
    } else if (p->persistent == 1) {
        hipLaunchKernelGGL(pr_sweep_persistent_kernel, dim3(p->npersist), dim3(kPrThreads), shmem, s, a, c);
    } else {
        hipLaunchKernelGGL(pr_sweep_kernel, dim3(nhubt + p->nblk), dim3(kPrThreads), shmem, s, a, c);
    }
    if (p->G == 1) {
        hipLaunchKernelGGL(pr_finalize_kernel, dim3(p->nfin), dim3(kFinThreads), 0, s, p->hub_v, p->hub_chunk_off,
                           p->chunk_part, p->outdeg, r_in, r_out, x_out, p->blk_stat, p->nblk, p->nhub, p->fin_stat,
                           p->counters, d_stats, c);
    } else {
        if (p->nhub)
            hipLaunchKernelGGL(pr_hub_apply_kernel, dim3(grid_for((uint64_t)p->nhub, kFinThreads, 1024)),
                               dim3(kFinThreads), 0, s, p->hub_v, p->hub_graph, p->hub_chunk_off, p->chunk_part,
                               p->outdeg, r_in, r_out, x_out, p->nhub, p->hub_stat, c);
        hipLaunchKernelGGL(pr_graph_reduce_kernel, dim3(std::min<uint32_t>(p->G, 4096)), dim3(kFinThreads), 0, s,
                           p->gblk, p->ghub, p->blk_stat, p->hub_stat, p->G, p->counters, d_stats);
    }
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

// Form-4 plans timed in the state a HIP-graph exec keeps them in (rank space).
int pr_init_rank_space(const gchi_pr_plan *p, float init, hipStream_t s) {
    GCHI_CHECK(p->seg, GCHI_ERR_INVALID, "pr_init_rank_space: not a form-4 plan");
    return seg_init(p->seg, init, nullptr, 0.0f, 0, s);
}

int pr_sweep_rank_space(const gchi_pr_plan *p, const gchi_pr_op *op, int src, int dst, double *d_stats,
                        hipStream_t s) {
    GCHI_CHECK(p->seg, GCHI_ERR_INVALID, "pr_sweep_rank_space: not a form-4 plan");
    return seg_sweep_rank(p->seg, op, src, dst, d_stats, s);
}

struct PlanView {
    uint32_t N, nblk, nhub, nchunk, npersist;
    const uint32_t *reg_off, *reg_src, *blk_row, *hub_bits, *hub_v, *hub_chunk_off;
    const uint64_t *chunk_beg;
    const uint32_t *chunk_len, *hub_src, *outdeg;
};

// Row structure of a single-graph plan whose edges hold vertex ids (no gather
// layout): reg_off[v] is row v's first regular edge in reg_src.
int pr_plan_view(const gchi_pr_plan *p, PlanView *v) {
    GCHI_CHECK(p && v, GCHI_ERR_INVALID, "plan view: null argument");
    GCHI_CHECK(p->G == 1, GCHI_ERR_UNSUPPORTED, "plan view: needs a single-graph plan");
    GCHI_CHECK(!p->seg, GCHI_ERR_UNSUPPORTED, "plan view: launch form 4 keeps the graph in rank order; build the "
                                              "plan with persistent 0-3");
    GCHI_CHECK(!p->hot, GCHI_ERR_UNSUPPORTED, "plan view: the plan's edges hold gather-layout positions "
                                              "(hot_vertices != 0); build it with hot_vertices = 0");
    *v = PlanView{p->N, p->nblk, p->nhub, p->nchunk, p->npersist, p->reg_off, p->reg_src, p->blk_row, p->hub_bits,
                  p->hub_v, p->hub_chunk_off, p->chunk_beg, p->chunk_len, p->hub_src, p->outdeg};
    return GCHI_OK;
}

}  // namespace gchi

using namespace gchi;

static int check_csc(const gchi_dgraph *g) {
    GCHI_CHECK(g, GCHI_ERR_INVALID, "pr_plan_create: null graph");
    GCHI_CHECK(g->nvertices > 0, GCHI_ERR_INVALID, "pr_plan_create: graph has no vertices");
    GCHI_CHECK(g->in_off && g->outdeg && (g->nedges == 0 || g->in_src), GCHI_ERR_INVALID,
               "pr_plan_create: graph arrays missing");
    GCHI_CHECK(g->nedges <= kMaxEdges, GCHI_ERR_UNSUPPORTED, "pr_plan_create: a graph's E must be <= 2^32 - 2^25");
    return GCHI_OK;
}

extern "C" int gchi_pr_plan_create(const gchi_dgraph *g, const gchi_pr_tune *tune, gchi_pr_plan **out, void *stream) {
    GCHI_TRY(check_csc(g));
    uint32_t vb[2] = {0, g->nvertices};
    PlanInput in;
    in.N = g->nvertices;
    in.E = g->nedges;
    in.h_vbase = vb;
    in.unionc = g;
    return plan_create(in, tune, out, as_stream(stream));
}

extern "C" int gchi_pr_plan_create_multi(const gchi_dgraph *g, const uint32_t *h_vbase, uint32_t ngraphs,
                                         const gchi_pr_tune *tune, gchi_pr_plan **out, void *stream) {
    GCHI_TRY(check_csc(g));
    GCHI_CHECK(ngraphs >= 1, GCHI_ERR_INVALID, "pr_plan_create_multi: no graphs");
    PlanInput in;
    in.N = g->nvertices;
    in.E = g->nedges;
    in.G = ngraphs;
    in.h_vbase = h_vbase;
    in.unionc = g;
    return plan_create(in, tune, out, as_stream(stream));
}

extern "C" int gchi_pr_plan_create_batch(const gchi_dgraph *graphs, uint32_t ngraphs, const gchi_pr_tune *tune,
                                         gchi_pr_plan **out, void *stream) {
    GCHI_CHECK(graphs && ngraphs >= 1, GCHI_ERR_INVALID, "pr_plan_create_batch: no graphs");
    std::vector<uint32_t> vb(ngraphs + 1, 0);
    uint64_t n = 0, e = 0;
    for (uint32_t k = 0; k < ngraphs; k++) {
        GCHI_TRY(check_csc(&graphs[k]));
        n += graphs[k].nvertices;
        e += graphs[k].nedges;
        GCHI_CHECK(n < (1ull << 32), GCHI_ERR_UNSUPPORTED, "pr_plan_create_batch: total vertices must be < 2^32");
        vb[k + 1] = (uint32_t)n;
    }
    PlanInput in;
    in.N = (uint32_t)n;
    in.E = e;
    in.G = ngraphs;
    in.h_vbase = vb.data();
    in.graphs = graphs;
    return plan_create(in, tune, out, as_stream(stream));
}

extern "C" int gchi_pr_plan_destroy(gchi_pr_plan *p) {
    if (!p) return GCHI_OK;
    seg_plan_destroy(p->seg);
    if (p->mem) (void)hipFree(p->mem);
    if (p->own_outdeg) (void)hipFree(p->own_outdeg);
    delete p;
    return GCHI_OK;
}

extern "C" int gchi_pr_plan_info_get(const gchi_pr_plan *p, gchi_pr_plan_info *info) {
    GCHI_CHECK(p && info, GCHI_ERR_INVALID, "pr_plan_info: null argument");
    info->nvertices = p->N;
    info->nedges = p->E;
    info->edges_per_block = p->W;
    info->short_row_max = p->short_max;
    info->regular_blocks = p->nblk;
    info->hub_rows = p->nhub;
    info->hub_chunks = p->nchunk;
    info->regular_edges = p->e_reg;
    info->hub_edges = p->e_hub;
    info->ngraphs = p->G;
    info->persistent = p->persistent;
    info->persistent_workgroups = p->persistent == 3   ? p->nhot_wg
                                  : p->persistent == 2 ? p->nwave_wg
                                  : p->persistent      ? p->npersist
                                                       : 0u;
    info->hot_vertices = p->H;
    info->gather_sources = p->A;
    info->hot_threads = p->persistent == 2 ? (uint32_t)kWaveThreads : p->persistent == 3 ? (uint32_t)kPrThreads : 0u;
    if (p->seg) {  // tiles as blocks, split rows as hub rows, every slot as a regular edge
        uint32_t tiles = 0, nsplit = 0, hot = 0, nwg = 0;
        uint64_t slots = 0;
        seg_plan_info(p->seg, &tiles, &nsplit, &slots, &hot, &nwg);
        info->edges_per_block = 512;
        info->short_row_max = 0;
        info->regular_blocks = tiles;
        info->hub_rows = nsplit;
        info->hub_chunks = 0;
        info->regular_edges = slots;
        info->hub_edges = 0;
        info->persistent_workgroups = nwg;
        info->hot_vertices = hot;
        info->gather_sources = p->N;
        info->hot_threads = 1024;
    }
    return GCHI_OK;
}

extern "C" int gchi_pr_init(const gchi_pr_plan *p, float init, float *d_x, float *d_r, void *stream) {
    GCHI_CHECK(p && d_x && d_r, GCHI_ERR_INVALID, "pr_init: null argument");
    return pr_init(p, init, nullptr, 0.0f, d_x, d_r, as_stream(stream));
}

extern "C" int gchi_pr_init_seeds(const gchi_pr_plan *p, const uint32_t *d_seed_bits, float seed_value, float *d_x,
                                  float *d_r, void *stream) {
    GCHI_CHECK(p && d_seed_bits && d_x && d_r, GCHI_ERR_INVALID, "pr_init_seeds: null argument");
    return pr_init(p, 0.0f, d_seed_bits, seed_value, d_x, d_r, as_stream(stream));
}

extern "C" int gchi_pr_sweep(const gchi_pr_plan *p, const gchi_pr_op *op, const float *d_x_in, const float *d_r_in,
                             float *d_x_out, float *d_r_out, double *d_stats, void *stream) {
    GCHI_CHECK(p && op && d_x_in && d_r_in && d_x_out && d_r_out && d_stats, GCHI_ERR_INVALID,
               "pr_sweep: null argument");
    GCHI_CHECK(d_x_in != d_x_out && d_r_in != d_r_out, GCHI_ERR_INVALID, "pr_sweep: in and out buffers must differ");
    GCHI_CHECK(p->G == 1 || op->d_graph_teleport || op->teleport >= 0, GCHI_ERR_INVALID, "pr_sweep: bad operator");
    return pr_sweep(p, op, d_x_in, d_r_in, d_x_out, d_r_out, d_stats, as_stream(stream));
}

// ------------------------------------------------------ semi-synchronous sweeps
// The functional engine's semisync mode (functional_semisync.hpp:76-110 under
// the interval loop of graphchi_engine.hpp:825-921): the vertex blocks (the
// engine's sub-intervals) run in order within a sweep, and block j gathers the
// x written this sweep by earlier blocks and the previous sweep's x of the
// rest.  Each block has its own plan over its rows (sources keep their global
// ids), and x lives in one buffer: block j's sweep reads it whole, writes its
// new x to a scratch copy, and a copy kernel then moves that slice into x, so
// rows of block j never see each other's new values.
namespace gchi {

__global__ void blk_slice_offsets(const uint32_t *__restrict__ in_off, uint32_t a, uint32_t n, uint32_t *out) {
    const uint32_t base = in_off[a];
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k <= n; k += stride) out[k] = in_off[a + k] - base;
}

__global__ void blk_copy_f32(const float *__restrict__ src, float *__restrict__ dst, uint32_t n) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += stride) dst[k] = src[k];
}

// out bit k = bit a+k of bits, k < n (one word per thread)
__global__ void blk_shift_bits(const uint32_t *__restrict__ bits, uint32_t a, uint32_t n, uint32_t *out) {
    const uint32_t words = (n + 31) / 32;
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
        const uint64_t first = (uint64_t)a + 32ull * w;  // global index of bit 0 of this word
        const uint32_t q = (uint32_t)(first >> 5), r = (uint32_t)(first & 31);
        const uint32_t left = n - 32 * w, take = left < 32 ? left : 32;
        uint32_t word = bits[q] >> r;
        // the next global word holds bits of this one only when they are inside
        // [a, a + n), so it exists
        if (r && r + take > 32) word |= bits[q + 1] << (32 - r);
        if (take < 32) word &= (1u << take) - 1u;
        out[w] = word;
    }
}

// d_stats[0..1] = sums of the blocks' statistics, in block order
__global__ void blk_stats_sum(const double *__restrict__ bst, uint32_t nb, double *out) {
    if (blockIdx.x || threadIdx.x) return;
    double r = 0.0, t = 0.0;
    for (uint32_t j = 0; j < nb; j++) {
        r += bst[2 * j];
        t += bst[2 * j + 1];
    }
    out[0] = r;
    out[1] = t;
}

}  // namespace gchi

struct gchi_pr_blocked {
    uint32_t N = 0, nb = 0;
    std::vector<uint32_t> bstart;        // [nb+1]
    std::vector<gchi_pr_plan *> plans;   // [nb]
    std::vector<uint32_t *> seed_local;  // [nb] block-local seed bitmaps
    const uint32_t *outdeg = nullptr;
    float *xnew = nullptr;
    double *bstat = nullptr;
    void *mem = nullptr;
};

namespace gchi {

static void blocked_free(gchi_pr_blocked *b) {
    for (gchi_pr_plan *p : b->plans) gchi_pr_plan_destroy(p);
    if (b->mem) (void)hipFree(b->mem);
    delete b;
}

}  // namespace gchi

extern "C" int gchi_pr_blocked_create(const gchi_dgraph *g, const uint32_t *h_bstart, uint32_t nblocks,
                                      const gchi_pr_tune *tune, gchi_pr_blocked **out, void *stream) {
    GCHI_TRY(check_csc(g));
    GCHI_CHECK(h_bstart && out && nblocks >= 1, GCHI_ERR_INVALID, "pr_blocked_create: null argument or no blocks");
    const uint32_t N = g->nvertices;
    GCHI_CHECK(h_bstart[0] == 0 && h_bstart[nblocks] == N, GCHI_ERR_INVALID,
               "pr_blocked_create: blocks must start at 0 and end at N");
    for (uint32_t j = 0; j < nblocks; j++)
        GCHI_CHECK(h_bstart[j + 1] > h_bstart[j], GCHI_ERR_INVALID, "pr_blocked_create: empty or unordered block");
    hipStream_t s = as_stream(stream);
    gchi_pr_tune t = tune ? *tune : gchi_pr_tune{0, 0, 0, 0};
    t.hot_vertices = 0;  // block plans gather by vertex id (a gather layout would rank each slice's rows only)
    if (t.persistent == 3) t.persistent = 2;
    if (t.persistent == 2 && t.edges_per_block == 0) t.edges_per_block = 256;
    gchi_pr_blocked *b = new gchi_pr_blocked();
    b->N = N;
    b->nb = nblocks;
    b->bstart.assign(h_bstart, h_bstart + nblocks + 1);
    b->outdeg = g->outdeg;
    std::vector<uint32_t> eoff(nblocks + 1);
    Sizer z;
    z.take<float>(N);
    z.take<double>(2 * (size_t)nblocks);
    for (uint32_t j = 0; j < nblocks; j++) {
        z.take<uint32_t>((size_t)(h_bstart[j + 1] - h_bstart[j]) + 1);
        z.take<uint32_t>((size_t)(h_bstart[j + 1] - h_bstart[j] + 31) / 32);
    }
    if (hipMalloc(&b->mem, z.off) != hipSuccess) {
        b->mem = nullptr;
        blocked_free(b);
        return fail(GCHI_ERR_HIP, "pr_blocked_create: hipMalloc failed");
    }
    Carver c(b->mem, z.off);
    b->xnew = c.take<float>(N);
    b->bstat = c.take<double>(2 * (size_t)nblocks);
    std::vector<uint32_t *> offs(nblocks);
    for (uint32_t j = 0; j < nblocks; j++) {
        offs[j] = c.take<uint32_t>((size_t)(h_bstart[j + 1] - h_bstart[j]) + 1);
        b->seed_local.push_back(c.take<uint32_t>((size_t)(h_bstart[j + 1] - h_bstart[j] + 31) / 32));
    }
    int st = GCHI_OK;
    for (uint32_t j = 0; j <= nblocks && st == GCHI_OK; j++)
        st = read_u32(g->in_off + h_bstart[j], &eoff[j], s);
    for (uint32_t j = 0; j < nblocks && st == GCHI_OK; j++) {
        const uint32_t a = h_bstart[j], nj = h_bstart[j + 1] - a;
        hipLaunchKernelGGL(blk_slice_offsets, dim3(grid_for((uint64_t)nj + 1, 256, 65536)), dim3(256), 0, s,
                           g->in_off, a, nj, offs[j]);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            st = hip_fail(e, "pr_blocked_create", __FILE__, __LINE__);
            break;
        }
        const gchi_dgraph gj{nj, 0, (uint64_t)(eoff[j + 1] - eoff[j]), offs[j], g->in_src ? g->in_src + eoff[j] : nullptr,
                             g->outdeg + a, g->indeg ? g->indeg + a : nullptr};
        uint32_t vb[2] = {0, nj};
        PlanInput in;
        in.N = nj;
        in.E = gj.nedges;
        in.h_vbase = vb;
        in.unionc = &gj;
        in.src_limit = N;
        gchi_pr_plan *pj = nullptr;
        st = plan_create(in, &t, &pj, s);
        if (st == GCHI_OK) b->plans.push_back(pj);
    }
    if (st == GCHI_OK) {
        hipError_t e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = hip_fail(e, "pr_blocked_create", __FILE__, __LINE__);
    }
    if (st != GCHI_OK) {
        blocked_free(b);
        return st;
    }
    *out = b;
    return GCHI_OK;
}

extern "C" int gchi_pr_blocked_destroy(gchi_pr_blocked *b) {
    if (b) blocked_free(b);
    return GCHI_OK;
}

extern "C" int gchi_pr_blocked_init(const gchi_pr_blocked *b, float init, const uint32_t *d_seed_bits,
                                    float seed_value, float *d_x, float *d_r, void *stream) {
    GCHI_CHECK(b && d_x && d_r, GCHI_ERR_INVALID, "pr_blocked_init: null argument");
    hipLaunchKernelGGL(pr_init_kernel, dim3(grid_for(b->N, 256, 65536)), dim3(256), 0, as_stream(stream), b->outdeg,
                       b->N, init, d_seed_bits, seed_value, d_x, d_r);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_pr_blocked_sweep(const gchi_pr_blocked *b, const gchi_pr_op *op, float *d_x, const float *d_r_in,
                                     float *d_r_out, double *d_stats, void *stream) {
    GCHI_CHECK(b && op && d_x && d_r_in && d_r_out && d_stats, GCHI_ERR_INVALID, "pr_blocked_sweep: null argument");
    GCHI_CHECK(d_r_in != d_r_out, GCHI_ERR_INVALID, "pr_blocked_sweep: rank buffers must differ");
    GCHI_CHECK(!op->d_graph_teleport, GCHI_ERR_UNSUPPORTED, "pr_blocked_sweep: one graph only (no graph_teleport)");
    hipStream_t s = as_stream(stream);
    for (uint32_t j = 0; j < b->nb; j++) {
        const uint32_t a = b->bstart[j], nj = b->bstart[j + 1] - a;
        gchi_pr_op oj = *op;
        if (op->d_seed_bits) {
            hipLaunchKernelGGL(blk_shift_bits, dim3(grid_for((nj + 31) / 32, 256, 4096)), dim3(256), 0, s,
                               op->d_seed_bits, a, nj, b->seed_local[j]);
            oj.d_seed_bits = b->seed_local[j];
        }
        GCHI_TRY(pr_sweep(b->plans[j], &oj, d_x, d_r_in + a, b->xnew + a, d_r_out + a, b->bstat + 2 * (size_t)j, s));
        hipLaunchKernelGGL(blk_copy_f32, dim3(grid_for(nj, 256, 65536)), dim3(256), 0, s, b->xnew + a, d_x + a, nj);
        GCHI_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(blk_stats_sum, dim3(1), dim3(64), 0, s, b->bstat, b->nb, d_stats);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

// ---------------------------------------------------------------- graph exec
struct gchi_pr_exec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nsweeps = 0;
};

extern "C" int gchi_pr_exec_create_ex(const gchi_pr_plan *p, const gchi_pr_op *op, float init, int do_init,
                                      int nsweeps, float *d_x0, float *d_x1, float *d_r0, float *d_r1,
                                      double *d_stats_all, gchi_pr_exec **out) {
    GCHI_CHECK(p && op && d_x0 && d_x1 && d_r0 && d_r1 && d_stats_all && out, GCHI_ERR_INVALID,
               "pr_exec_create: null argument");
    GCHI_CHECK(nsweeps >= 0, GCHI_ERR_INVALID, "pr_exec_create: nsweeps < 0");
    hipStream_t cs;
    GCHI_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    gchi_pr_exec *ex = new gchi_pr_exec();
    ex->nsweeps = nsweeps;
    int st = GCHI_OK;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) st = hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
    if (st == GCHI_OK) {
        float *xs[2] = {d_x0, d_x1};
        float *rs[2] = {d_r0, d_r1};
        int s1 = GCHI_OK;
        if (p->seg) {
            // every sweep in rank space; the caller's buffers only on entry and exit
            if (do_init)
                s1 = op->d_seed_bits ? seg_init(p->seg, 0.0f, op->d_seed_bits, (float)op->seed_teleport, 0, cs)
                                     : seg_init(p->seg, init, nullptr, 0.0f, 0, cs);
            else
                s1 = seg_enter(p->seg, xs[0], rs[0], 0, cs);
            for (int k = 0; k < nsweeps && s1 == GCHI_OK; k++)
                s1 = seg_sweep_rank(p->seg, op, k & 1, (k + 1) & 1, d_stats_all + 2 * (size_t)k, cs);
            if (s1 == GCHI_OK) s1 = seg_leave(p->seg, nsweeps & 1, xs[nsweeps & 1], rs[nsweeps & 1], cs);
            do_init = 0;
            nsweeps = 0;  // captured above
        }
        if (do_init) {
            if (op->d_seed_bits)
                s1 = pr_init(p, 0.0f, op->d_seed_bits, (float)op->seed_teleport, xs[0], rs[0], cs);
            else
                s1 = pr_init(p, init, nullptr, 0.0f, xs[0], rs[0], cs);
        }
        const size_t stride = 2 * (size_t)p->G;
        for (int k = 0; k < nsweeps && s1 == GCHI_OK; k++)
            s1 = pr_sweep(p, op, xs[k & 1], rs[k & 1], xs[(k + 1) & 1], rs[(k + 1) & 1], d_stats_all + stride * k, cs);
        e = hipStreamEndCapture(cs, &ex->graph);
        if (s1 != GCHI_OK) st = s1;
        else if (e != hipSuccess) st = hip_fail(e, "hipStreamEndCapture", __FILE__, __LINE__);
    }
    if (st == GCHI_OK) {
        e = hipGraphInstantiate(&ex->exec, ex->graph, nullptr, nullptr, 0);
        if (e != hipSuccess) st = hip_fail(e, "hipGraphInstantiate", __FILE__, __LINE__);
    }
    (void)hipStreamDestroy(cs);
    if (st != GCHI_OK) {
        if (ex->graph) (void)hipGraphDestroy(ex->graph);
        delete ex;
        return st;
    }
    *out = ex;
    return GCHI_OK;
}

extern "C" int gchi_pr_exec_create(const gchi_pr_plan *p, const gchi_pr_op *op, float init, int nsweeps, float *d_x0,
                                   float *d_x1, float *d_r0, float *d_r1, double *d_stats_all, gchi_pr_exec **out) {
    return gchi_pr_exec_create_ex(p, op, init, 1, nsweeps, d_x0, d_x1, d_r0, d_r1, d_stats_all, out);
}

extern "C" int gchi_pr_exec_launch(gchi_pr_exec *ex, void *stream) {
    GCHI_CHECK(ex, GCHI_ERR_INVALID, "pr_exec_launch: null exec");
    GCHI_HIP(hipGraphLaunch(ex->exec, as_stream(stream)));
    return GCHI_OK;
}

extern "C" int gchi_pr_exec_destroy(gchi_pr_exec *ex) {
    if (!ex) return GCHI_OK;
    if (ex->exec) (void)hipGraphExecDestroy(ex->exec);
    if (ex->graph) (void)hipGraphDestroy(ex->graph);
    delete ex;
    return GCHI_OK;
}
