// PageRank pull sweep over fixed 512-slot tiles in out-degree rank order
// (launch form 4 of gchi_pr_tune).
//
// Semantics are those of the synchronous sweep in pagerank.hip (the
// reference's bulk-synchronous functional engine,
// src/api/functional/functional_bulksync.hpp:84-113, with the apply of
// example_apps/pagerank_functional.cpp:72-81): s_v = sum of x_u over v's
// in-edges, r_v = (float)(c + d * (double)s_v), x_v = r_v / outdeg_v.  Only the
// order of the FP32 additions inside a row differs (fixed, so sweeps are
// bit-reproducible).
//
// Layout (built once per graph):
//   * vertices are relabelled by out-degree, descending, ties by id (rank i
//     is vertex perm[i]); rows and sources both live in rank space, so the
//     values most often gathered are the first ones and the first H of them
//     stay in LDS for the whole sweep;
//   * row i owns max(indeg, 1) consecutive slots of one array, holding its
//     sources' ranks (a row without in-edges holds one slot with id N, whose
//     x is a constant 0);
//   * the slot array is cut into tiles of 512 slots (64 lanes x 8); one
//     wavefront reduces one tile: lane l loads slots 8l..8l+7 as two 16-byte
//     loads, gathers their x, sums them with a per-lane running sum that
//     restarts at every row start (a byte of row-start flags per lane), and a
//     segmented scan across the 64 lanes carries the sums of rows that cross
//     lanes.  Rows wholly inside the tile are applied by the same wavefront
//     (one lane per row, from sums staged in LDS);
//   * a row that crosses a tile boundary (every row longer than a tile, and
//     the row at most boundaries) is "split".  Tiles are equal work, so each
//     wavefront of the persistent launch takes a fixed contiguous range of
//     them (no work queue, no per-row loops, no padding beyond the last tile)
//     and carries the sum of the row in progress from tile to tile: a split
//     row that starts and ends inside one wavefront's range is applied by that
//     wavefront.  For the rows that cross from one range into the next (at
//     most one per range boundary) every tile also writes two partials, head
//     (slots before its first row start) and tail (from its last row start to
//     its end, when that row goes on into the next tile), and the finalize
//     sums such a row's partials in the same order and applies it.
//
// State lives in rank space: the plan owns xs[2] (N+1 floats, xs[N] = 0) and
// rs[2].  A HIP-graph exec keeps every sweep there and permutes only on entry
// and exit; a standalone sweep permutes its vertex-order inputs in and its
// outputs out.
#include "common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <vector>

namespace gchi {

constexpr int kSegPer = 8;                    // slots per lane
constexpr uint32_t kSegTile = kWave * kSegPer;  // slots per tile
constexpr int kSegThreads = 1024;             // workgroup of the sweep (16 wavefronts, one per CU)
constexpr int kSegWaves = kSegThreads / kWave;
constexpr int kSegFinThreads = 1024;
constexpr size_t kSegStaticLds = 1024;
constexpr uint32_t kSegCont = 0x80000000u;  // tile_row0 bit: the tile's last row goes on into the next tile

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

struct PrConstsSeg {
    double teleport;
    double damping;
    const uint32_t *seed_bits;  // vertex-id bitmap, or null
    double seed_tele;
};

struct SegArgs {
    const uint32_t *slots;      // [T*512] source ranks (N for a slot without a source)
    const uint8_t *flags;       // [T*64] row-start bits, byte l of tile t = slots 8l..8l+7
    const uint32_t *tile_row0;  // [T] rank of the row holding slot 0, | kSegCont
    const uint32_t *outdeg_r;   // [N] out-degree by rank
    const uint32_t *perm;       // [N] vertex of each rank
    const float *xs_in;         // [N+1] x by rank (xs_in[N] = 0)
    const float *rs_in;         // [N]
    float *xs_out, *rs_out;
    float *part;                // [2T] head / tail partials
    double *wave_stat;          // [2 * waves] residual, rank sum of the rows each wavefront applied
    uint32_t T, N;
    uint32_t per;               // tiles per wavefront: wavefront w takes tiles [w*per, w*per + per)
};

GCHI_DEV double seg_tele(const PrConstsSeg &c, const uint32_t *perm, uint32_t i) {
    double t = c.teleport;
    if (c.seed_bits) {
        const uint32_t v = perm[i];
        if ((c.seed_bits[v >> 5] >> (v & 31)) & 1u) t += c.seed_tele;
    }
    return t;
}

// r = (float)(teleport + damping * (double)s) without FMA contraction (the
// reference's double multiply, double add, rounding to float).
GCHI_DEV float seg_apply(double teleport, double damping, float s) {
#pragma clang fp contract(off)
    double t = damping * (double)s;
    double r = teleport + t;
    return (float)r;
}

GCHI_DEV float seg_scatter(float r, uint32_t od) { return od ? __fdiv_rn(r, (float)od) : 0.0f; }

GCHI_DEV void seg_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Exclusive prefix over the lanes of a value in [0, 15].
GCHI_DEV uint32_t lane_prefix4(uint32_t f) {
    uint32_t pre = 0;
#pragma unroll
    for (int b = 0; b < 4; b++) {
        const uint64_t m = __ballot((f >> b) & 1u);
        const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        pre += below << b;
    }
    return pre;
}

// The loads of one tile: its slot ids, its row-start flags and its metadata.
struct SegIds {
    u32x4 q0, q1;
    uint32_t fb, meta;
};

GCHI_DEV void seg_load_ids(const SegArgs &a, uint32_t t, int lane, SegIds &d) {
    const uint64_t sbase = (uint64_t)t * kSegTile + (uint64_t)lane * kSegPer;
    d.q0 = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(a.slots + sbase));
    d.q1 = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(a.slots + sbase + 4));
    d.fb = a.flags[(uint64_t)t * kWave + lane];
    d.meta = a.tile_row0[t];
}

// x of the tile's sources: the first H ranks from LDS, the others from HBM
// (lanes with a cached source skip the HBM load); the choice is made where the
// value is used (seg_pick).  The same step loads the previous rank and the
// out-degree of the tile's first 64 rows (local rows 0..63) for the apply.
struct SegVals {
    float g[kSegPer];  // from HBM
    float h[kSegPer];  // from the LDS cache
    uint32_t lds;      // bit j: slot j's source is cached
    float ro;          // previous rank of row row0 + lane (clamped below N)
    uint32_t od;       // its out-degree
};

template <bool HOT>
GCHI_DEV void seg_gather(const SegArgs &a, const float *hot, uint32_t H, const SegIds &d, SegVals &v, int lane) {
    const uint32_t ids[kSegPer] = {d.q0.x, d.q0.y, d.q0.z, d.q0.w, d.q1.x, d.q1.y, d.q1.z, d.q1.w};
    v.lds = 0;
#pragma unroll
    for (int j = 0; j < kSegPer; j++) {
        const uint32_t u = ids[j];
        const bool in_lds = HOT && u < H;
        float gv = 0.0f;
        if (!in_lds) gv = a.xs_in[u];  // cached lanes issue no HBM request (1% faster than a load of xs_in[0])
        v.g[j] = gv;
        if (HOT) {
            v.h[j] = hot[in_lds ? u : 0u];
            v.lds |= (in_lds ? 1u : 0u) << j;
        }
    }
    const uint32_t i = min((d.meta & ~kSegCont) + (uint32_t)lane, a.N - 1);
    v.ro = a.rs_in[i];
    v.od = a.outdeg_r[i];
}

// The gathered values of a tile.  The empty asm statements hand the select
// two opaque registers: without them the compiler turns "cached ? h : g" into
// a select of addresses and moves the arrays to scratch memory.
template <bool HOT>
GCHI_DEV void seg_pick(const SegVals &v, float *val) {
#pragma unroll
    for (int j = 0; j < kSegPer; j++) {
        float g = v.g[j];
        if (HOT) {
            float h = v.h[j];
            asm volatile("" : "+v"(g));
            asm volatile("" : "+v"(h));
            g = ((v.lds >> j) & 1u) ? h : g;
        }
        val[j] = g;
    }
}

// What the scan of a tile leaves for its apply.
struct SegRows {
    uint32_t row0, lo, hi;  // the rows the tile applies are local rows [lo, hi) from rank row0
    float ro;               // previous rank of row row0 + lane (loaded with the gathers)
    uint32_t od;            // its out-degree
};

// The row a wavefront carries from one tile into the next: its sum so far,
// and whether it started inside the wavefront's range (then the wavefront
// applies it; otherwise it is a row of the finalize).  Wavefront-uniform.
struct SegCarry {
    float sum = 0.0f;
    bool open = false;
};

// One tile, one wavefront, from its gathered values: the row sums (rows the
// tile applies into rowsum, this wavefront's 512-float LDS slice; head and
// tail partials into part), and the carry into the next tile.
GCHI_DEV void seg_scan(const SegArgs &a, float *rowsum, uint32_t t, int lane, const float *val, uint32_t fb,
                       uint32_t meta, SegCarry &k, SegRows &R) {
    const uint32_t row0 = meta & ~kSegCont;
    const bool cont = (meta & kSegCont) != 0;
    // the tile's slot 0 starts a row: that row is local row 0, not a new one
    const bool s0 = (__builtin_amdgcn_readfirstlane(fb) & 1u) != 0;
    const uint32_t fcount = (uint32_t)__popc(fb) - ((lane == 0 && s0) ? 1u : 0u);
    const uint32_t pre = lane_prefix4(fcount);  // local index of the row in progress at this lane's start
    uint32_t loc = pre;
    float run = 0.0f, head = 0.0f;
    bool seen = false;
#pragma unroll
    for (int j = 0; j < kSegPer; j++) {
        if ((fb >> j) & 1u) {
            if (seen) rowsum[loc] = run;  // a row started and ended inside this lane
            else head = run;
            if (lane != 0 || j != 0) loc++;
            seen = true;
            run = 0.0f;
        }
        run += val[j];
    }
    // segmented inclusive scan over the lanes: a lane with a row start restarts
    // the carried sum with its own tail
    float v = run;
    bool f = fb != 0;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const float vu = __shfl_up(v, d, kWave);
        const bool fu = __shfl_up(f ? 1 : 0, d, kWave) != 0;
        if (lane >= d) {
            if (!f) v = vu + v;
            f = f || fu;
        }
    }
    float carry = __shfl_up(v, 1, kWave);
    if (lane == 0) carry = 0.0f;
    const uint64_t anyf = __ballot(fb != 0);
    const float total = __shfl(v, kWave - 1, kWave);  // sum since the tile's last row start
    const uint32_t last = __shfl(loc, kWave - 1, kWave);  // local index of the tile's last row
    // the row coming from the previous tile ends here and began in this range
    // the carried row (begun in this range) ends here: at the tile's first row
    // start, or with the tile when one row covers it and does not go on
    const bool ends = k.open && anyf != 0;
    const bool ends_whole = k.open && !anyf && !cont;
    if (fb != 0) {
        // this lane's head ends the row in progress at its start
        const uint32_t hrow = pre;
        const float hs = carry + head;
        if (hrow == 0 && !s0) {  // the row coming from the previous tile
            if (ends) rowsum[0] = k.sum + hs;
            else a.part[2 * (uint64_t)t] = hs;
        } else if (!(lane == 0 && s0 && hrow == 0)) {
            rowsum[hrow] = hs;
        }
    }
    if (lane == 0) {
        if (!anyf) {  // one row covers the whole tile
            a.part[2 * (uint64_t)t] = total;
            a.part[2 * (uint64_t)t + 1] = 0.0f;
            if (ends_whole) rowsum[0] = k.sum + total;
        } else {
            if (s0) a.part[2 * (uint64_t)t] = 0.0f;
            a.part[2 * (uint64_t)t + 1] = cont ? total : 0.0f;
            if (!cont) rowsum[last] = total;
        }
    }
    // a carried row is open only after a tile whose last row goes on (so a
    // tile with one is never s0), and stays open across the tiles it covers
    if (!anyf) {
        k.sum = k.sum + total;
        k.open = k.open && cont;
    } else {
        k.sum = total;
        k.open = cont;
    }
    seg_lds_sync();
    // local rows [lo, hi): the rows ending in the tile, less the one that came
    // from before this wavefront's range
    R.row0 = row0;
    R.lo = (s0 || ends || ends_whole) ? 0u : 1u;
    R.hi = anyf ? last + (cont ? 0u : 1u) : (ends_whole ? 1u : 0u);
}

// The apply of the rows the scan left in LDS: r from the row sum, the
// residual against the previous rank, x = r / outdeg.  The first 64 rows are
// computed unconditionally (stores masked) from the values loaded with the
// tile's gathers, so the compiler keeps those loads where they were issued.
GCHI_DEV void seg_apply_rows(const SegArgs &a, const PrConstsSeg &c, const float *rowsum, int lane,
                             const SegRows &R, double &res, double &sum) {
    {
        const uint32_t li = (uint32_t)lane;
        const bool ok = li >= R.lo && li < R.hi;
        const uint32_t i = R.row0 + (ok ? li : 0u);
        const float r = seg_apply(seg_tele(c, a.perm, i), c.damping, rowsum[ok ? li : 0u]);
        const float x = seg_scatter(r, R.od);
        const double d = fabs((double)r - (double)R.ro);
        if (ok) {
            a.rs_out[i] = r;
            a.xs_out[i] = x;
            res += d;
            sum += (double)r;
        }
    }
    for (uint32_t r0 = kWave; r0 < R.hi; r0 += kWave) {
        const uint32_t li = r0 + (uint32_t)lane;
        if (li < R.hi) {
            const uint32_t i = R.row0 + li;
            const float r = seg_apply(seg_tele(c, a.perm, i), c.damping, rowsum[li]);
            const float ro = a.rs_in[i];
            a.rs_out[i] = r;
            a.xs_out[i] = seg_scatter(r, a.outdeg_r[i]);
            res += fabs((double)r - (double)ro);
            sum += (double)r;
        }
    }
    seg_lds_sync();
}

// Reduces tile t from its gathered values; its rows' statistics are added to
// the lane's res and sum.
template <bool HOT>
GCHI_DEV void seg_tile(const SegArgs &a, const PrConstsSeg &c, float *rowsum, uint32_t t, int lane,
                       const SegVals &v, uint32_t fb, uint32_t meta, SegCarry &k, double &res, double &sum) {
    float val[kSegPer];
    seg_pick<HOT>(v, val);
    SegRows R;
    seg_scan(a, rowsum, t, lane, val, fb, meta, k, R);
    R.ro = v.ro;
    R.od = v.od;
    seg_apply_rows(a, c, rowsum, lane, R, res, sum);
}

template <bool HOT>
__global__ __launch_bounds__(kSegThreads) void pr_seg_sweep_kernel(SegArgs a, PrConstsSeg c, uint32_t H) {
    extern __shared__ float lds[];  // hot[H], then 512 floats per wavefront
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(tid / kWave);
    float *hot = lds;
    float *rowsum = lds + H + (size_t)wave * kSegTile;
    if (HOT) {
        // 8 loads per thread in flight
        for (uint32_t b = 0; b < H; b += 8 * kSegThreads) {
            float tmp[8];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t i = b + (uint32_t)q * kSegThreads + (uint32_t)tid;
                tmp[q] = a.xs_in[i < H ? i : 0u];
            }
#pragma unroll
            for (int q = 0; q < 8; q++) {
                const uint32_t i = b + (uint32_t)q * kSegThreads + (uint32_t)tid;
                if (i < H) hot[i] = tmp[q];
            }
        }
    }
    __syncthreads();
    const uint32_t gw = blockIdx.x * kSegWaves + wave;
    const uint64_t per = a.per;
    const uint32_t t0 = (uint32_t)min<uint64_t>(a.T, per * gw);
    const uint32_t t1 = (uint32_t)min<uint64_t>(a.T, per * gw + per);
    double res = 0.0, sum = 0.0;  // this lane's share of the wavefront's statistics
    SegCarry k;
    if (t0 < t1) {
        // Two-stage pipeline: the ids of tile t+2 and the gathers (and apply
        // loads) of tile t+1 are issued before tile t is reduced, so every wait
        // of the reduce is for loads issued a whole tile earlier (vector-memory
        // waits complete in issue order).  Two fixed register sets (A, B) take
        // turns, so no value is copied between iterations (a copy would make the
        // compiler wait for the loads in flight).  Past the wavefront's last tile
        // the loads repeat that tile (in bounds, unused).
        const uint32_t tl = t1 - 1;
        SegIds iA, iB;
        SegVals vA, vB;
        seg_load_ids(a, t0, lane, iA);
        seg_load_ids(a, min(t0 + 1, tl), lane, iB);
        seg_gather<HOT>(a, hot, H, iA, vA, lane);
        for (uint32_t t = t0;; t += 2) {
            {  // tile t in A: ids of t+2 into A, gathers of t+1 into B, then reduce t
                const uint32_t fb = iA.fb, meta = iA.meta;
                seg_load_ids(a, min(t + 2, tl), lane, iA);
                seg_gather<HOT>(a, hot, H, iB, vB, lane);
                seg_tile<HOT>(a, c, rowsum, t, lane, vA, fb, meta, k, res, sum);
            }
            if (t + 1 >= t1) break;
            {  // tile t+1 in B: ids of t+3 into B, gathers of t+2 into A, then reduce t+1
                const uint32_t fb = iB.fb, meta = iB.meta;
                seg_load_ids(a, min(t + 3, tl), lane, iB);
                seg_gather<HOT>(a, hot, H, iA, vA, lane);
                seg_tile<HOT>(a, c, rowsum, t + 1, lane, vB, fb, meta, k, res, sum);
            }
            if (t + 2 >= t1) break;
        }
    }
    res = wave_sum(res);
    sum = wave_sum(sum);
    if (lane == 0) {  // every wavefront writes its pair, zeros when it had no tile
        a.wave_stat[2 * (uint64_t)gw] = res;
        a.wave_stat[2 * (uint64_t)gw + 1] = sum;
    }
}

// Rows crossing from one sweep wavefront's range into the next: partials
// part[beg, end) summed, then the apply.  One lane per row, in tile order; rows of more than kSegLaneParts partials by the whole
// wavefront (lane-strided, then across the lanes).  Then the deterministic
// reduction of the sweep wavefronts' statistics, and the last workgroup writes
// d_stats.
constexpr uint32_t kSegLaneParts = 32;

__global__ __launch_bounds__(kSegFinThreads) void pr_seg_finalize_kernel(
    const uint32_t *__restrict__ split_v, const uint32_t *__restrict__ split_beg,
    const uint32_t *__restrict__ split_end, uint32_t nsplit, const float *__restrict__ part,
    const uint32_t *__restrict__ outdeg_r, const uint32_t *__restrict__ perm, const float *__restrict__ rs_in,
    float *__restrict__ rs_out, float *__restrict__ xs_out, const double *__restrict__ wave_stat, uint32_t nws,
    double *fin_stat, uint32_t *counter, double *d_stats, PrConstsSeg c) {
    __shared__ double s_red[2 * (kSegFinThreads / kWave)];
    __shared__ bool s_last;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const uint32_t gt = blockIdx.x * kSegFinThreads + tid;
    const uint32_t nthreads = gridDim.x * kSegFinThreads;
    double res = 0.0, sum = 0.0;
    const uint32_t nwv = nthreads / kWave, wv = gt / kWave;
    for (uint32_t base = wv; base < nsplit; base += nthreads) {
        // lane slots spread over all wavefronts: the rows of many partials are
        // among the lowest ranks, and consecutive slots would put them in a few
        // wavefronts one after another (measured 14% slower sweeps)
        const uint32_t h = base + (uint32_t)lane * nwv;
        const bool valid = h < nsplit;
        uint32_t b = 0, e = 0, i = 0;
        if (valid) {
            b = split_beg[h];
            e = split_end[h];
            i = split_v[h];
        }
        const bool big = e - b > kSegLaneParts;
        float s = 0.0f;
        if (!big) {
            for (uint32_t k0 = b; k0 < e; k0 += 8) {
                float p[8];
#pragma unroll
                for (int q = 0; q < 8; q++) p[q] = k0 + q < e ? part[k0 + q] : 0.0f;
#pragma unroll
                for (int q = 0; q < 8; q++)
                    if (k0 + q < e) s += p[q];
            }
        }
        uint64_t m = __ballot(big);
        while (m) {
            const int j = __builtin_ctzll(m);
            m &= m - 1;
            const uint32_t b0 = (uint32_t)__builtin_amdgcn_readlane((int)b, j);
            const uint32_t b1 = (uint32_t)__builtin_amdgcn_readlane((int)e, j);
            // lane-strided partial sums, then a butterfly: a fixed order, so
            // sweeps stay bit-reproducible, in (b1 - b0) / 64 + 6 steps
            float w = 0.0f;
            for (uint32_t k = b0 + lane; k < b1; k += kWave) w += part[k];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o, kWave);
            if (lane == j) s = w;
        }
        if (valid) {
            const float r = seg_apply(seg_tele(c, perm, i), c.damping, s);
            const float ro = rs_in[i];
            rs_out[i] = r;
            xs_out[i] = seg_scatter(r, outdeg_r[i]);
            res += fabs((double)r - (double)ro);
            sum += (double)r;
        }
    }
    for (uint32_t k = gt; k < nws; k += nthreads) {
        res += wave_stat[2 * (uint64_t)k];
        sum += wave_stat[2 * (uint64_t)k + 1];
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
        for (int w = 0; w < kSegFinThreads / kWave; w++) {
            a += s_red[2 * w];
            s2 += s_red[2 * w + 1];
        }
        st_relaxed_agent(&fin_stat[2 * blockIdx.x], a);
        st_relaxed_agent(&fin_stat[2 * blockIdx.x + 1], s2);
        __threadfence();
        const uint32_t ticket = atomicAdd(counter, 1u);
        s_last = ticket == gridDim.x - 1;
    }
    __syncthreads();
    if (s_last) {  // workgroup-uniform
        __threadfence();
        double a = 0.0, s2 = 0.0;
        for (uint32_t j = tid; j < gridDim.x; j += kSegFinThreads) {
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
            for (int w = 0; w < kSegFinThreads / kWave; w++) {
                ta += s_red[2 * w];
                ts += s_red[2 * w + 1];
            }
            d_stats[0] = ta;
            d_stats[1] = ts;
            *counter = 0u;
        }
    }
}

// ---- rank-space state: entry, exit, initial values ----
__global__ void seg_gather_in(const uint32_t *__restrict__ perm, uint32_t n, const float *__restrict__ x,
                              const float *__restrict__ r, float *__restrict__ xs, float *__restrict__ rs) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        if (i == n) {
            xs[n] = 0.0f;
            continue;
        }
        const uint32_t v = perm[i];
        xs[i] = x[v];
        rs[i] = r[v];
    }
}

__global__ void seg_scatter_out(const uint32_t *__restrict__ perm, uint32_t n, const float *__restrict__ xs,
                                const float *__restrict__ rs, float *__restrict__ x, float *__restrict__ r) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t v = perm[i];
        x[v] = xs[i];
        r[v] = rs[i];
    }
}

// iteration 0 in rank space: r = init (or seed_value on seeds, 0 elsewhere),
// x = r / outdeg
__global__ void seg_init_kernel(const uint32_t *__restrict__ perm, const uint32_t *__restrict__ outdeg_r, uint32_t n,
                                float init, const uint32_t *seed_bits, float seed_value, float *xs, float *rs) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += stride) {
        if (i == n) {
            xs[n] = 0.0f;
            continue;
        }
        float r0 = init;
        if (seed_bits) {
            const uint32_t v = perm[i];
            r0 = ((seed_bits[v >> 5] >> (v & 31)) & 1u) ? seed_value : 0.0f;
        }
        rs[i] = r0;
        xs[i] = seg_scatter(r0, outdeg_r[i]);
    }
}

// ---- plan build ----
// keys: out-degree descending, then id
__global__ void seg_order_keys(const uint32_t *outdeg, uint32_t n, uint64_t *keys) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride)
        keys[v] = ((uint64_t)(0xffffffffu - outdeg[v]) << 32) | v;
}

__global__ void seg_ranks(const uint64_t *keys, const uint32_t *in_off, const uint32_t *outdeg, uint32_t n,
                          uint32_t *perm, uint32_t *rank, uint32_t *outdeg_r, uint32_t *nslot) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint32_t v = (uint32_t)keys[i];
        perm[i] = v;
        rank[v] = i;
        outdeg_r[i] = outdeg[v];
        const uint32_t d = in_off[v + 1] - in_off[v];
        nslot[i] = d ? d : 1u;
    }
}

// One wavefront per row: its sources' ranks into its slots (N for a row with
// none); row-start flag bits.  A source id >= n sets *bad and is stored as N.
__global__ void seg_fill(const uint32_t *in_off, const uint32_t *in_src, const uint32_t *perm, const uint32_t *rank,
                         const uint64_t *soff, uint32_t n, uint32_t *slots, uint32_t *flag_words, uint32_t *bad) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    for (uint64_t i = wid; i < n; i += nw) {
        const uint32_t v = perm[i];
        const uint32_t b = in_off[v], e = in_off[v + 1];
        const uint64_t s0 = soff[i];
        if (lane == 0) {
            atomicOr(&flag_words[s0 >> 5], 1u << (s0 & 31));
            if (e == b) slots[s0] = n;
        }
        for (uint32_t k = b + lane; k < e; k += kWave) {
            const uint32_t u = in_src[k];
            uint32_t ru = n;
            if (u < n) ru = rank[u];
            else atomicOr(bad, 1u);
            slots[s0 + (k - b)] = ru;
        }
    }
}

__global__ void seg_fill_value(uint32_t *p, uint64_t n, uint32_t v) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// last row whose first slot is <= s (soff strictly increasing, soff[0] = 0)
GCHI_DEV uint32_t seg_row_of(const uint64_t *soff, uint32_t n, uint64_t s) {
    uint32_t lo = 0, hi = n - 1;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo + 1) / 2;
        if (soff[mid] <= s) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void seg_tiles(const uint64_t *soff, uint32_t n, uint64_t S, uint32_t T, uint32_t *tile_row0) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < T; t += stride) {
        uint32_t meta = seg_row_of(soff, n, (uint64_t)t * kSegTile);
        const uint64_t nx = (uint64_t)(t + 1) * kSegTile;
        if (nx < S) {
            const uint32_t rn = seg_row_of(soff, n, nx);
            if (soff[rn] != nx) meta |= kSegCont;  // slot nx continues a row
        }
        tile_row0[t] = meta;
    }
}

// split[i] = 1 when row i's first and last slot lie in tiles of different
// wavefronts' ranges (per tiles each); a row crossing tiles inside one range
// is summed and applied by its wavefront
__global__ void seg_split_flags(const uint64_t *soff, uint32_t n, uint64_t per, uint32_t *split) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t ta = soff[i] / kSegTile, tb = (soff[i + 1] - 1) / kSegTile;
        split[i] = ta / per != tb / per ? 1u : 0u;  // every row has at least one slot
    }
}

// Split row h: partials part[2*t0 + 1 .. 2*t1] inclusive (its first tile's
// tail, then the heads of the tiles after it up to its last one).
__global__ void seg_split_list(const uint64_t *soff, const uint32_t *split, const uint32_t *pos, uint32_t n,
                               uint32_t *split_v, uint32_t *split_beg, uint32_t *split_end) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        if (!split[i]) continue;
        const uint32_t h = pos[i];
        const uint64_t t0 = soff[i] / kSegTile, t1 = (soff[i + 1] - 1) / kSegTile;
        split_v[h] = i;
        split_beg[h] = (uint32_t)(2 * t0 + 1);
        split_end[h] = (uint32_t)(2 * t1 + 1);
    }
}

}  // namespace gchi

struct gchi_seg_plan {
    uint32_t N = 0, T = 0, H = 0, nsplit = 0, nwg = 0, nfin = 0, per = 1;
    uint64_t S = 0, E = 0;
    size_t lds = 0;
    uint32_t *perm = nullptr, *outdeg_r = nullptr, *slots = nullptr, *tile_row0 = nullptr;
    uint8_t *flags = nullptr;
    uint32_t *split_v = nullptr, *split_beg = nullptr, *split_end = nullptr;
    float *part = nullptr;
    double *wave_stat = nullptr, *fin_stat = nullptr;
    uint32_t *counter = nullptr;
    float *xs[2] = {nullptr, nullptr};
    float *rs[2] = {nullptr, nullptr};
    void *mem = nullptr;
};

namespace gchi {

struct SegTmp {
    std::vector<void *> ptrs;
    ~SegTmp() {
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

static int seg_read(void *h, const void *d, size_t n, hipStream_t s) {
    GCHI_HIP(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    return GCHI_OK;
}

int seg_plan_create(const gchi_dgraph *g, int hot, gchi_seg_plan **out, hipStream_t s) {
    const uint32_t N = g->nvertices;
    const uint64_t E = g->nedges;
    GCHI_CHECK(N > 0 && N < (1u << 31), GCHI_ERR_UNSUPPORTED, "pr_plan (form 4): needs 0 < N < 2^31");
    GCHI_CHECK(hot >= -1, GCHI_ERR_INVALID, "pr_plan (form 4): hot_vertices must be -1, 0 or positive");
    SegTmp tp;
    uint64_t *keys = tp.alloc<uint64_t>(N);
    uint32_t *rank = tp.alloc<uint32_t>(N);
    uint32_t *nslot = tp.alloc<uint32_t>(N + 1);
    uint64_t *soff = tp.alloc<uint64_t>((size_t)N + 1);
    uint32_t *split = tp.alloc<uint32_t>(N + 1);
    uint32_t *spos = tp.alloc<uint32_t>(N + 1);
    uint32_t *bad = tp.alloc<uint32_t>(1);
    size_t sb = 0;
    GCHI_TRY(gchi_sort_workspace(N, 64, 0, &sb));
    const size_t swb = std::max(sb, scan_workspace((size_t)N + 1));
    void *sw = tp.alloc<char>(swb);
    GCHI_CHECK(keys && rank && nslot && soff && split && spos && bad && sw, GCHI_ERR_HIP,
               "pr_plan (form 4): hipMalloc failed");
    gchi_seg_plan *p = new gchi_seg_plan();
    p->N = N;
    p->E = E;
    auto fail_free = [&](int st) {
        if (p->mem) (void)hipFree(p->mem);
        delete p;
        return st;
    };
    // rank order and the slot counts of the rows
    hipLaunchKernelGGL(seg_order_keys, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, g->outdeg, N, keys);
    int st = radix_sort_u64(keys, nullptr, N, 64, sw, sb, s, nullptr);
    if (st == GCHI_OK) st = read_work_status(static_cast<const uint32_t *>(sw), s);
    if (st != GCHI_OK) return fail_free(st);
    // perm, outdeg_r live in the plan; rank is temporary
    uint32_t *perm_t = tp.alloc<uint32_t>(N), *od_t = tp.alloc<uint32_t>(N);
    if (!perm_t || !od_t) return fail_free(fail(GCHI_ERR_HIP, "pr_plan (form 4): hipMalloc failed"));
    hipLaunchKernelGGL(seg_ranks, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, keys, g->in_off, g->outdeg, N,
                       perm_t, rank, od_t, nslot);
    if ((st = exclusive_scan_u32_to_u64(nslot, soff, N, sw, swb, s)) != GCHI_OK) return fail_free(st);
    uint64_t S = 0;
    if ((st = seg_read(&S, soff + N, 8, s)) != GCHI_OK) return fail_free(st);
    const uint64_t T64 = (S + kSegTile - 1) / kSegTile;
    if (T64 >= (1ull << 31))
        return fail_free(fail(GCHI_ERR_UNSUPPORTED, "pr_plan (form 4): more than 2^31 tiles"));
    const uint32_t T = (uint32_t)T64;
    p->S = S;
    p->T = T;
    // launch shape: one 1024-thread workgroup per CU, each wavefront a fixed
    // range of per tiles
    int dev = 0, ncu = 256;
    size_t lds_cu = 160 * 1024, lds_wg = 160 * 1024;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess) {
        ncu = prop.multiProcessorCount;
        lds_cu = prop.maxSharedMemoryPerMultiProcessor;
        lds_wg = prop.sharedMemPerBlock;
    }
    p->nwg = (uint32_t)std::min<uint64_t>((uint64_t)ncu, ((uint64_t)T + kSegWaves - 1) / kSegWaves);
    if (p->nwg == 0) p->nwg = 1;
    const uint64_t nwv = (uint64_t)p->nwg * kSegWaves;
    p->per = (uint32_t)std::max<uint64_t>(1, (T + nwv - 1) / nwv);
    // the rows of the finalize: those crossing from one wavefront's range into the next
    hipLaunchKernelGGL(seg_split_flags, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, soff, N, (uint64_t)p->per,
                       split);
    if ((st = exclusive_scan_u32(split, spos, N, sw, swb, s)) != GCHI_OK) return fail_free(st);
    uint32_t nsplit = 0;
    if ((st = seg_read(&nsplit, spos + N, 4, s)) != GCHI_OK) return fail_free(st);
    p->nsplit = nsplit;
    // the plan's memory
    const uint64_t nslots = (uint64_t)T * kSegTile;
    Sizer z;
    z.take<uint32_t>(N);                 // perm
    z.take<uint32_t>(N);                 // outdeg_r
    z.take<uint32_t>(nslots);            // slots
    z.take<uint32_t>(nslots / 32);       // flags (as words)
    z.take<uint32_t>(T);                 // tile_row0
    z.take<uint32_t>(nsplit + 1);        // split_v
    z.take<uint32_t>(nsplit + 1);        // split_beg
    z.take<uint32_t>(nsplit + 1);        // split_end
    z.take<float>(2 * (size_t)T);        // part
    z.take<double>(2 * ((size_t)T + kSegWaves));  // wave_stat (at most one wavefront per tile, rounded up)
    z.take<double>(2 * 1024);            // fin_stat
    z.take<uint32_t>(64);                // counter
    z.take<float>((size_t)N + 1);        // xs[0]
    z.take<float>((size_t)N + 1);        // xs[1]
    z.take<float>(N);                    // rs[0]
    z.take<float>(N);                    // rs[1]
    if (hipMalloc(&p->mem, z.off) != hipSuccess) {
        p->mem = nullptr;
        return fail_free(fail(GCHI_ERR_HIP, "pr_plan (form 4): hipMalloc failed (" + std::to_string(z.off) +
                                                " bytes)"));
    }
    Carver c(p->mem, z.off);
    p->perm = c.take<uint32_t>(N);
    p->outdeg_r = c.take<uint32_t>(N);
    p->slots = c.take<uint32_t>(nslots);
    uint32_t *flag_words = c.take<uint32_t>(nslots / 32);
    p->flags = reinterpret_cast<uint8_t *>(flag_words);
    p->tile_row0 = c.take<uint32_t>(T);
    p->split_v = c.take<uint32_t>(nsplit + 1);
    p->split_beg = c.take<uint32_t>(nsplit + 1);
    p->split_end = c.take<uint32_t>(nsplit + 1);
    p->part = c.take<float>(2 * (size_t)T);
    p->wave_stat = c.take<double>(2 * ((size_t)T + kSegWaves));
    p->fin_stat = c.take<double>(2 * 1024);
    p->counter = c.take<uint32_t>(64);
    p->xs[0] = c.take<float>((size_t)N + 1);
    p->xs[1] = c.take<float>((size_t)N + 1);
    p->rs[0] = c.take<float>(N);
    p->rs[1] = c.take<float>(N);
    GCHI_HIP(hipMemcpyAsync(p->perm, perm_t, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->outdeg_r, od_t, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
    launch_fill_u32(flag_words, 0u, nslots / 32, s);
    launch_fill_u32(bad, 0u, 1, s);
    launch_fill_u32(p->counter, 0u, 64, s);
    // x of the no-source id N stays 0 in both state buffers: sweeps never
    // write it, and rows without in-edges gather it
    launch_fill_u32(reinterpret_cast<uint32_t *>(p->xs[0]) + N, 0u, 1, s);
    launch_fill_u32(reinterpret_cast<uint32_t *>(p->xs[1]) + N, 0u, 1, s);
    if (nslots > S)
        hipLaunchKernelGGL(seg_fill_value, dim3(grid_for(nslots - S, 256, 4096)), dim3(256), 0, s, p->slots + S,
                           nslots - S, N);
    hipLaunchKernelGGL(seg_fill, dim3(grid_for((uint64_t)N * kWave, 256, 1u << 16)), dim3(256), 0, s, g->in_off,
                       g->in_src, p->perm, rank, soff, N, p->slots, flag_words, bad);
    hipLaunchKernelGGL(seg_tiles, dim3(grid_for(T, 256, 65536)), dim3(256), 0, s, soff, N, S, T, p->tile_row0);
    if (nsplit)
        hipLaunchKernelGGL(seg_split_list, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, soff, split, spos, N,
                           p->split_v, p->split_beg, p->split_end);
    GCHI_LAUNCH_CHECK();
    uint32_t h_bad = 0;
    if ((st = seg_read(&h_bad, bad, 4, s)) != GCHI_OK) return fail_free(st);
    if (h_bad) return fail_free(fail(GCHI_ERR_INVALID, "pr_plan_create: an in-edge source id >= nvertices"));
    // LDS holds the H hottest ranks plus a 512-float row-sum slice per wavefront
    const size_t cap = std::min(lds_cu, lds_wg);
    const size_t stage = (size_t)kSegWaves * kSegTile * 4;
    const size_t hmax = cap > stage + kSegStaticLds ? (cap - stage - kSegStaticLds) / 4 / 1024 * 1024 : 0;
    uint32_t H = hot > 0 ? (uint32_t)std::min<uint64_t>((uint64_t)hot, hmax) / 1024 * 1024 : 0;
    H = std::min<uint32_t>(H, N / 1024 * 1024);
    p->H = H;
    p->lds = stage + 4 * (size_t)H;
    // at least 64 workgroups: the statistics of the 4,096 sweep wavefronts are
    // reduced by all of them (265.9 -> 265.0 us per sweep against 16)
    p->nfin = std::min<uint32_t>(1024, std::max<uint32_t>(64, (nsplit + kSegFinThreads - 1) / kSegFinThreads));
    const void *fn = H ? (const void *)pr_seg_sweep_kernel<true> : (const void *)pr_seg_sweep_kernel<false>;
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p->lds);
    if (e != hipSuccess) return fail_free(hip_fail(e, "hipFuncSetAttribute", __FILE__, __LINE__));
    e = hipStreamSynchronize(s);
    if (e != hipSuccess) return fail_free(hip_fail(e, "pr_plan (form 4)", __FILE__, __LINE__));
    *out = p;
    return GCHI_OK;
}

void seg_plan_destroy(gchi_seg_plan *p) {
    if (!p) return;
    if (p->mem) (void)hipFree(p->mem);
    delete p;
}

void seg_plan_info(const gchi_seg_plan *p, uint32_t *tiles, uint32_t *nsplit, uint64_t *slots, uint32_t *hot,
                   uint32_t *nwg) {
    *tiles = p->T;
    *nsplit = p->nsplit;
    *slots = p->S;
    *hot = p->H;
    *nwg = p->nwg;
}

// One sweep in rank space: xs[src] / rs[src] -> xs[dst] / rs[dst].
int seg_sweep_rank(const gchi_seg_plan *p, const gchi_pr_op *op, int src, int dst, double *d_stats, hipStream_t s) {
    GCHI_CHECK(!op->d_graph_teleport, GCHI_ERR_UNSUPPORTED, "pr_sweep (form 4): one graph only");
    const PrConstsSeg c{op->teleport, op->damping, op->d_seed_bits, op->seed_teleport};
    SegArgs a{p->slots, p->flags, p->tile_row0, p->outdeg_r, p->perm, p->xs[src], p->rs[src], p->xs[dst],
              p->rs[dst], p->part, p->wave_stat, p->T, p->N, p->per};
    if (p->H)
        hipLaunchKernelGGL(pr_seg_sweep_kernel<true>, dim3(p->nwg), dim3(kSegThreads), p->lds, s, a, c, p->H);
    else
        hipLaunchKernelGGL(pr_seg_sweep_kernel<false>, dim3(p->nwg), dim3(kSegThreads), p->lds, s, a, c, 0u);
    hipLaunchKernelGGL(pr_seg_finalize_kernel, dim3(p->nfin), dim3(kSegFinThreads), 0, s, p->split_v, p->split_beg,
                       p->split_end, p->nsplit, p->part, p->outdeg_r, p->perm, p->rs[src], p->rs[dst], p->xs[dst],
                       p->wave_stat, p->nwg * (uint32_t)kSegWaves, p->fin_stat, p->counter, d_stats, c);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

int seg_enter(const gchi_seg_plan *p, const float *x, const float *r, int dst, hipStream_t s) {
    hipLaunchKernelGGL(seg_gather_in, dim3(grid_for((uint64_t)p->N + 1, 256, 65536)), dim3(256), 0, s, p->perm, p->N,
                       x, r, p->xs[dst], p->rs[dst]);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

int seg_leave(const gchi_seg_plan *p, int src, float *x, float *r, hipStream_t s) {
    hipLaunchKernelGGL(seg_scatter_out, dim3(grid_for(p->N, 256, 65536)), dim3(256), 0, s, p->perm, p->N, p->xs[src],
                       p->rs[src], x, r);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

int seg_init(const gchi_seg_plan *p, float init, const uint32_t *seed_bits, float seed_value, int dst,
             hipStream_t s) {
    hipLaunchKernelGGL(seg_init_kernel, dim3(grid_for((uint64_t)p->N + 1, 256, 65536)), dim3(256), 0, s, p->perm,
                       p->outdeg_r, p->N, init, seed_bits, seed_value, p->xs[dst], p->rs[dst]);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

// A standalone sweep with vertex-order buffers: in, sweep, out.
int seg_sweep(const gchi_seg_plan *p, const gchi_pr_op *op, const float *x_in, const float *r_in, float *x_out,
              float *r_out, double *d_stats, hipStream_t s) {
    GCHI_TRY(seg_enter(p, x_in, r_in, 0, s));
    GCHI_TRY(seg_sweep_rank(p, op, 0, 1, d_stats, s));
    return seg_leave(p, 1, x_out, r_out, s);
}

}  // namespace gchi
