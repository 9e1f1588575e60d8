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
//   regular rows (in-degree < W): reg_off[N+1], reg_src[E_reg]; a row's "work"
//     is deg+1, and block b owns rows with reg_off[v]+v in [b*W, (b+1)*W), so a
//     block sees < 2W edges and < 2W rows.  Its in-edges are streamed coalesced
//     (non-temporal), x[src] gathered into LDS, then rows are reduced: one lane
//     per row up to short_row_max edges (sequential adds in src order, as the
//     reference's gather), one wavefront per longer row.
//   hub rows (in-degree >= W): their in-edges are cut into chunks of 4096; one
//     workgroup reduces a chunk; a finalize kernel sums each hub's chunk
//     partials in chunk order and applies.
// Every sum is in a fixed order, so a sweep is bit-reproducible run to run.
#include "common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <vector>

struct gchi_pr_plan {
    uint32_t N = 0;
    uint64_t E = 0;
    int W = 2048;
    int short_max = 32;
    uint32_t nblk = 0;     // regular blocks
    uint32_t nhub = 0;     // hub rows
    uint32_t nchunk = 0;   // hub chunks
    uint32_t nfin = 0;     // finalize workgroups
    uint64_t e_reg = 0, e_hub = 0;
    const uint32_t *outdeg = nullptr;
    uint32_t *reg_off = nullptr;
    uint32_t *reg_src = nullptr;
    uint32_t *hub_bits = nullptr;
    uint32_t *blk_row = nullptr;
    uint32_t *hub_v = nullptr;
    uint32_t *hub_chunk_off = nullptr;
    uint32_t *chunk_beg = nullptr;
    uint32_t *chunk_len = nullptr;
    uint32_t *hub_src = nullptr;
    float *chunk_part = nullptr;
    double *blk_stat = nullptr;
    double *fin_stat = nullptr;
    uint32_t *fin_counter = nullptr;
    void *mem = nullptr;
};

namespace gchi {

constexpr int kPrThreads = 256;
constexpr int kPrWaves = kPrThreads / kWave;
constexpr int kHubChunk = 4096;
constexpr int kMaxW = 4096;
constexpr int kMaxLong = 512;  // long rows per regular block
constexpr int kFinThreads = 256;

struct PrConsts {
    double teleport;
    double damping;
};

// r = (float)(teleport + damping * (double)s), without contraction into an FMA:
// the reference evaluates a double multiply and a double add, then rounds.
GCHI_DEV float pr_apply(PrConsts c, float s) {
#pragma clang fp contract(off)
    double t = c.damping * (double)s;
    double r = c.teleport + t;
    return (float)r;
}

GCHI_DEV float pr_scatter(float r, uint32_t outdeg) { return outdeg ? __fdiv_rn(r, (float)outdeg) : 0.0f; }

GCHI_DEV void pr_finish_row(uint32_t v, float s, PrConsts c, const float *r_in, const uint32_t *outdeg, float *r_out,
                            float *x_out, double &res, double &sum) {
    float r = pr_apply(c, s);
    float ro = r_in[v];
    r_out[v] = r;
    x_out[v] = pr_scatter(r, outdeg[v]);
    res += fabs((double)r - (double)ro);
    sum += (double)r;
}

template <typename T>
GCHI_DEV T block_sum_ordered(T v, T *s_red) {
    // wave butterfly, then waves in index order
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    if (lane == 0) s_red[wave] = v;
    __syncthreads();
    T t = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x / kWave); w++) t += s_red[w];
    __syncthreads();
    return t;  // valid in thread 0
}

__global__ __launch_bounds__(kPrThreads) void pr_sweep_kernel(
    const uint32_t *__restrict__ reg_off, const uint32_t *__restrict__ reg_src, const uint32_t *__restrict__ blk_row,
    const uint32_t *__restrict__ hub_bits, const uint32_t *__restrict__ chunk_beg,
    const uint32_t *__restrict__ chunk_len, const uint32_t *__restrict__ hub_src, const uint32_t *__restrict__ outdeg,
    const float *__restrict__ x_in, const float *__restrict__ r_in, float *__restrict__ x_out,
    float *__restrict__ r_out, float *__restrict__ chunk_part, double *__restrict__ blk_stat, uint32_t nchunk,
    int short_max, PrConsts c) {
    extern __shared__ float s_val[];  // 2W floats
    __shared__ uint32_t s_long[kMaxLong];
    __shared__ uint32_t s_nlong;
    __shared__ double s_red[2 * kPrWaves];
    __shared__ float s_redf[kPrWaves];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;

    if (blockIdx.x < nchunk) {
        // ---- hub chunk: partial sum of <= 4096 in-edges of one hub row ----
        const uint32_t cb = chunk_beg[blockIdx.x], cl = chunk_len[blockIdx.x];
        uint32_t idx[kHubChunk / kPrThreads];
#pragma unroll
        for (int j = 0; j < kHubChunk / kPrThreads; j++) {
            uint32_t k = (uint32_t)j * kPrThreads + tid;
            idx[j] = k < cl ? __builtin_nontemporal_load(&hub_src[cb + k]) : 0u;
        }
        float acc = 0.0f;
#pragma unroll
        for (int j = 0; j < kHubChunk / kPrThreads; j++) {
            uint32_t k = (uint32_t)j * kPrThreads + tid;
            float v = k < cl ? x_in[idx[j]] : 0.0f;
            acc += v;
        }
        acc = wave_sum(acc);
        if (lane == 0) s_redf[wave] = acc;
        __syncthreads();
        if (tid == 0) {
            float t = 0.0f;
            for (int w = 0; w < kPrWaves; w++) t += s_redf[w];
            chunk_part[blockIdx.x] = t;
        }
        return;
    }

    // ---- regular block ----
    const uint32_t b = blockIdx.x - nchunk;
    const uint32_t rb = blk_row[b], re = blk_row[b + 1];
    const uint32_t e0 = reg_off[rb], e1 = reg_off[re];
    const uint32_t span = e1 - e0;
    if (tid == 0) s_nlong = 0;
    for (uint32_t base = 0; base < span; base += 8 * kPrThreads) {
        uint32_t idx[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t k = base + (uint32_t)j * kPrThreads + tid;
            idx[j] = k < span ? __builtin_nontemporal_load(&reg_src[e0 + k]) : 0u;
        }
        float val[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t k = base + (uint32_t)j * kPrThreads + tid;
            val[j] = k < span ? x_in[idx[j]] : 0.0f;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            uint32_t k = base + (uint32_t)j * kPrThreads + tid;
            if (k < span) s_val[k] = val[j];
        }
    }
    __syncthreads();

    double res = 0.0, sum = 0.0;
    for (uint32_t v = rb + tid; v < re; v += kPrThreads) {
        if ((hub_bits[v >> 5] >> (v & 31)) & 1u) continue;
        uint32_t beg = reg_off[v] - e0, end = reg_off[v + 1] - e0;
        if ((int)(end - beg) <= short_max) {
            float s = 0.0f;
            for (uint32_t k = beg; k < end; k++) s += s_val[k];
            pr_finish_row(v, s, c, r_in, outdeg, r_out, x_out, res, sum);
        } else {
            uint32_t slot = atomicAdd(&s_nlong, 1u);
            s_long[slot] = v;
        }
    }
    __syncthreads();
    const uint32_t nlong = s_nlong;
    for (uint32_t li = wave; li < nlong; li += kPrWaves) {
        uint32_t v = s_long[li];
        uint32_t beg = reg_off[v] - e0, end = reg_off[v + 1] - e0;
        float s = 0.0f;
        for (uint32_t k = beg + lane; k < end; k += kWave) s += s_val[k];
        s = wave_sum(s);
        if (lane == 0) pr_finish_row(v, s, c, r_in, outdeg, r_out, x_out, res, sum);
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
        for (int w = 0; w < kPrWaves; w++) {
            a += s_red[2 * w];
            s2 += s_red[2 * w + 1];
        }
        blk_stat[2 * b] = a;
        blk_stat[2 * b + 1] = s2;
    }
}

// Hub rows (sum of chunk partials, apply) + deterministic reduction of the
// per-block statistics; the last finalize workgroup writes d_stats.
__global__ __launch_bounds__(kFinThreads) void pr_finalize_kernel(
    const uint32_t *__restrict__ hub_v, const uint32_t *__restrict__ hub_chunk_off,
    const float *__restrict__ chunk_part, const uint32_t *__restrict__ outdeg, const float *__restrict__ r_in,
    float *__restrict__ r_out, float *__restrict__ x_out, const double *__restrict__ blk_stat, uint32_t nblk,
    uint32_t nhub, double *fin_stat, uint32_t *fin_counter, double *d_stats, PrConsts c) {
    __shared__ double s_red[2 * (kFinThreads / kWave)];
    __shared__ bool s_last;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wave = tid / kWave;
    const uint32_t nwaves_total = gridDim.x * (kFinThreads / kWave);
    const uint32_t gw = blockIdx.x * (kFinThreads / kWave) + wave;
    double res = 0.0, sum = 0.0;
    for (uint32_t h = gw; h < nhub; h += nwaves_total) {
        uint32_t c0 = hub_chunk_off[h], c1 = hub_chunk_off[h + 1];
        float s = 0.0f;
        for (uint32_t k = c0 + lane; k < c1; k += kWave) s += chunk_part[k];
        s = wave_sum(s);
        if (lane == 0) pr_finish_row(hub_v[h], s, c, r_in, outdeg, r_out, x_out, res, sum);
    }
    const uint32_t gt = blockIdx.x * kFinThreads + tid;
    const uint32_t nthreads_total = gridDim.x * kFinThreads;
    for (uint32_t i = gt; i < nblk; i += nthreads_total) {
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
        uint32_t ticket = atomicAdd(fin_counter, 1u);
        s_last = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (s_last && tid == 0) {
        __threadfence();
        double a = 0.0, s2 = 0.0;
        for (uint32_t j = 0; j < gridDim.x; j++) {
            a += ld_relaxed_agent(&fin_stat[2 * j]);
            s2 += ld_relaxed_agent(&fin_stat[2 * j + 1]);
        }
        d_stats[0] = a;
        d_stats[1] = s2;
        *fin_counter = 0u;  // ready for the next sweep / graph replay
    }
}

__global__ void pr_init_kernel(const uint32_t *outdeg, uint32_t n, float init, float *x, float *r) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        r[v] = init;
        x[v] = pr_scatter(init, outdeg[v]);
    }
}

// ---------------------------------------------------------------- plan build
__global__ void plan_rows(const uint32_t *in_off, uint32_t N, int W, uint32_t *reg_deg, uint32_t *hub_flag,
                          uint32_t *hub_bits) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) {
        uint32_t d = in_off[v + 1] - in_off[v];
        bool hub = d >= (uint32_t)W;
        reg_deg[v] = hub ? 0u : d;
        hub_flag[v] = hub ? 1u : 0u;
        if (hub) atomicOr(&hub_bits[v >> 5], 1u << (v & 31));
    }
}

__global__ void plan_hubs(const uint32_t *in_off, uint32_t N, const uint32_t *hub_flag, const uint32_t *hub_idx,
                          uint32_t *hub_v, uint32_t *hub_deg, uint32_t *hub_nchunk) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) {
        if (hub_flag[v]) {
            uint32_t h = hub_idx[v];
            uint32_t d = in_off[v + 1] - in_off[v];
            hub_v[h] = v;
            hub_deg[h] = d;
            hub_nchunk[h] = (d + kHubChunk - 1) / kHubChunk;
        }
    }
}

// one wavefront per row: copy the row's in-edges into the regular or hub array
__global__ void plan_copy(const uint32_t *in_off, const uint32_t *in_src, uint32_t N, const uint32_t *hub_flag,
                          const uint32_t *hub_idx, const uint32_t *reg_off, const uint32_t *hub_off,
                          uint32_t *reg_src, uint32_t *hub_src) {
    const int lane = threadIdx.x & (kWave - 1);
    uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    for (uint64_t v = wid; v < N; v += nw) {
        uint32_t b = in_off[v], e = in_off[v + 1];
        uint32_t *dst = hub_flag[v] ? hub_src + hub_off[hub_idx[v]] : reg_src + reg_off[v];
        for (uint32_t k = b + lane; k < e; k += kWave) dst[k - b] = in_src[k];
    }
}

__global__ void plan_chunks(const uint32_t *hub_off, const uint32_t *hub_chunk_off, uint32_t nhub, uint32_t *chunk_beg,
                            uint32_t *chunk_len) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t h = blockIdx.x * blockDim.x + threadIdx.x; h < nhub; h += stride) {
        uint32_t b = hub_off[h], d = hub_off[h + 1] - b;
        uint32_t c0 = hub_chunk_off[h], c1 = hub_chunk_off[h + 1];
        for (uint32_t c = c0; c < c1; c++) {
            uint32_t j = c - c0;
            chunk_beg[c] = b + j * kHubChunk;
            chunk_len[c] = min((uint32_t)kHubChunk, d - j * kHubChunk);
        }
    }
}

__global__ void plan_blocks(const uint32_t *reg_off, uint32_t N, int W, uint32_t nblk, uint32_t *blk_row) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < N; v += stride) {
        uint64_t bv = ((uint64_t)reg_off[v] + v) / (uint64_t)W;
        int64_t bp = v == 0 ? -1 : (int64_t)(((uint64_t)reg_off[v - 1] + v - 1) / (uint64_t)W);
        for (int64_t bb = bp + 1; bb <= (int64_t)bv; bb++) blk_row[bb] = v;
        if (v == N - 1) blk_row[nblk] = N;
    }
}

static int pr_default_W(const gchi_pr_tune *t) {
    int W = (t && t->edges_per_block > 0) ? t->edges_per_block : 2048;
    return W;
}

}  // namespace gchi

using namespace gchi;

namespace gchi {

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

static int plan_build(const gchi_dgraph *g, int W, int short_max, gchi_pr_plan *p, hipStream_t s) {
    const uint32_t N = g->nvertices;
    const uint64_t E = g->nedges;
    TmpPool tp;
    const size_t nwords = (N + 31) / 32;
    uint32_t *reg_deg = tp.alloc<uint32_t>(N);
    uint32_t *hub_flag = tp.alloc<uint32_t>(N);
    uint32_t *hub_idx = tp.alloc<uint32_t>(N + 1);
    uint32_t *reg_off = tp.alloc<uint32_t>(N + 1);
    uint32_t *bits = tp.alloc<uint32_t>(nwords);
    const size_t swb = scan_workspace((size_t)N + 1);
    void *sw = tp.alloc<char>(swb);
    GCHI_CHECK(reg_deg && hub_flag && hub_idx && reg_off && bits && sw, GCHI_ERR_HIP, "pr_plan: hipMalloc failed");
    launch_fill_u32(bits, 0u, nwords, s);
    hipLaunchKernelGGL(plan_rows, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, g->in_off, N, W, reg_deg, hub_flag,
                       bits);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(exclusive_scan_u32(reg_deg, reg_off, N, sw, swb, s));
    GCHI_TRY(exclusive_scan_u32(hub_flag, hub_idx, N, sw, swb, s));
    uint32_t e_reg = 0, nhub = 0, reg_last = 0;
    GCHI_TRY(read_u32(reg_off + N, &e_reg, s));
    GCHI_TRY(read_u32(hub_idx + N, &nhub, s));
    GCHI_TRY(read_u32(reg_off + (N - 1), &reg_last, s));
    p->e_reg = e_reg;
    p->e_hub = E - e_reg;
    p->nhub = nhub;
    p->nblk = (uint32_t)(((uint64_t)reg_last + (N - 1)) / (uint64_t)W + 1);

    uint32_t *hub_v = tp.alloc<uint32_t>(nhub + 1);
    uint32_t *hub_deg = tp.alloc<uint32_t>(nhub + 1);
    uint32_t *hub_nch = tp.alloc<uint32_t>(nhub + 1);
    uint32_t *hub_off = tp.alloc<uint32_t>(nhub + 1);
    uint32_t *hub_choff = tp.alloc<uint32_t>(nhub + 1);
    const size_t swb2 = scan_workspace((size_t)nhub + 1);
    void *sw2 = tp.alloc<char>(swb2);
    GCHI_CHECK(hub_v && hub_deg && hub_nch && hub_off && hub_choff && sw2, GCHI_ERR_HIP, "pr_plan: hipMalloc failed");
    if (nhub) {
        hipLaunchKernelGGL(plan_hubs, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, g->in_off, N, hub_flag, hub_idx,
                           hub_v, hub_deg, hub_nch);
        GCHI_LAUNCH_CHECK();
    }
    GCHI_TRY(exclusive_scan_u32(hub_deg, hub_off, nhub, sw2, swb2, s));
    GCHI_TRY(exclusive_scan_u32(hub_nch, hub_choff, nhub, sw2, swb2, s));
    uint32_t nchunk = 0;
    GCHI_TRY(read_u32(hub_choff + nhub, &nchunk, s));
    p->nchunk = nchunk;
    p->nfin = 64;

    Sizer zz;
    zz.take<uint32_t>(N + 1);
    zz.take<uint32_t>(std::max<uint64_t>(p->e_reg, 1));
    zz.take<uint32_t>(nwords);
    zz.take<uint32_t>(p->nblk + 1);
    zz.take<uint32_t>(nhub + 1);
    zz.take<uint32_t>(nhub + 1);
    zz.take<uint32_t>(nchunk + 1);
    zz.take<uint32_t>(nchunk + 1);
    zz.take<uint32_t>(std::max<uint64_t>(p->e_hub, 1));
    zz.take<float>(nchunk + 1);
    zz.take<double>(2 * (size_t)p->nblk);
    zz.take<double>(2 * (size_t)p->nfin);
    zz.take<uint32_t>(1);
    GCHI_HIP(hipMalloc(&p->mem, zz.off));
    Carver c(p->mem, zz.off);
    p->reg_off = c.take<uint32_t>(N + 1);
    p->reg_src = c.take<uint32_t>(std::max<uint64_t>(p->e_reg, 1));
    p->hub_bits = c.take<uint32_t>(nwords);
    p->blk_row = c.take<uint32_t>(p->nblk + 1);
    p->hub_v = c.take<uint32_t>(nhub + 1);
    p->hub_chunk_off = c.take<uint32_t>(nhub + 1);
    p->chunk_beg = c.take<uint32_t>(nchunk + 1);
    p->chunk_len = c.take<uint32_t>(nchunk + 1);
    p->hub_src = c.take<uint32_t>(std::max<uint64_t>(p->e_hub, 1));
    p->chunk_part = c.take<float>(nchunk + 1);
    p->blk_stat = c.take<double>(2 * (size_t)p->nblk);
    p->fin_stat = c.take<double>(2 * (size_t)p->nfin);
    p->fin_counter = c.take<uint32_t>(1);
    GCHI_HIP(hipMemcpyAsync(p->reg_off, reg_off, (size_t)(N + 1) * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->hub_bits, bits, nwords * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->hub_v, hub_v, (size_t)(nhub + 1) * 4, hipMemcpyDeviceToDevice, s));
    GCHI_HIP(hipMemcpyAsync(p->hub_chunk_off, hub_choff, (size_t)(nhub + 1) * 4, hipMemcpyDeviceToDevice, s));
    launch_fill_u32(p->fin_counter, 0u, 1, s);
    if (E)
        hipLaunchKernelGGL(plan_copy, dim3(grid_for((uint64_t)N * kWave, 256, 1u << 16)), dim3(256), 0, s, g->in_off,
                           g->in_src, N, hub_flag, hub_idx, reg_off, hub_off, p->reg_src, p->hub_src);
    if (nhub)
        hipLaunchKernelGGL(plan_chunks, dim3(grid_for(nhub, 256, 4096)), dim3(256), 0, s, hub_off, hub_choff, nhub,
                           p->chunk_beg, p->chunk_len);
    hipLaunchKernelGGL(plan_blocks, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, p->reg_off, N, W, p->nblk,
                       p->blk_row);
    GCHI_LAUNCH_CHECK();
    GCHI_HIP(hipStreamSynchronize(s));
    return GCHI_OK;
}

}  // namespace gchi

extern "C" int gchi_pr_plan_create(const gchi_dgraph *g, const gchi_pr_tune *tune, gchi_pr_plan **out, void *stream) {
    GCHI_CHECK(g && out, GCHI_ERR_INVALID, "pr_plan_create: null argument");
    GCHI_CHECK(g->nvertices > 0, GCHI_ERR_INVALID, "pr_plan_create: graph has no vertices");
    GCHI_CHECK(g->in_off && g->outdeg && (g->nedges == 0 || g->in_src), GCHI_ERR_INVALID,
               "pr_plan_create: graph arrays missing");
    GCHI_CHECK(g->nedges < (1ull << 32), GCHI_ERR_UNSUPPORTED, "pr_plan_create: E must be < 2^32");
    const int W = pr_default_W(tune);
    GCHI_CHECK(W >= 256 && W <= kMaxW && W % 256 == 0, GCHI_ERR_INVALID,
               "pr_plan_create: edges_per_block must be a multiple of 256 in [256, 4096]");
    const int short_max = (tune && tune->short_row_max > 0) ? tune->short_row_max : 32;
    GCHI_CHECK(short_max >= 1 && short_max <= 1024, GCHI_ERR_INVALID, "pr_plan_create: short_row_max out of range");
    GCHI_CHECK(2 * W / (short_max + 1) <= kMaxLong, GCHI_ERR_INVALID,
               "pr_plan_create: short_row_max too small for edges_per_block (more than 512 long rows per block)");
    gchi_pr_plan *p = new gchi_pr_plan();
    p->N = g->nvertices;
    p->E = g->nedges;
    p->W = W;
    p->short_max = short_max;
    p->outdeg = g->outdeg;
    int st = plan_build(g, W, short_max, p, as_stream(stream));
    if (st != GCHI_OK) {
        if (p->mem) (void)hipFree(p->mem);
        delete p;
        return st;
    }
    *out = p;
    return GCHI_OK;
}

extern "C" int gchi_pr_plan_destroy(gchi_pr_plan *p) {
    if (!p) return GCHI_OK;
    if (p->mem) (void)hipFree(p->mem);
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
    return GCHI_OK;
}

namespace gchi {

int pr_init(const gchi_pr_plan *p, float init, float *x, float *r, hipStream_t s) {
    hipLaunchKernelGGL(pr_init_kernel, dim3(grid_for(p->N, 256, 65536)), dim3(256), 0, s, p->outdeg, p->N, init, x, r);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

int pr_sweep(const gchi_pr_plan *p, const gchi_pr_op *op, const float *x_in, const float *r_in, float *x_out,
             float *r_out, double *d_stats, hipStream_t s) {
    PrConsts c{op->teleport, op->damping};
    const unsigned grid = p->nchunk + p->nblk;
    const size_t shmem = (size_t)2 * p->W * sizeof(float);
    hipLaunchKernelGGL(pr_sweep_kernel, dim3(grid), dim3(kPrThreads), shmem, s, p->reg_off, p->reg_src, p->blk_row,
                       p->hub_bits, p->chunk_beg, p->chunk_len, p->hub_src, p->outdeg, x_in, r_in, x_out, r_out,
                       p->chunk_part, p->blk_stat, p->nchunk, p->short_max, c);
    hipLaunchKernelGGL(pr_finalize_kernel, dim3(p->nfin), dim3(kFinThreads), 0, s, p->hub_v, p->hub_chunk_off,
                       p->chunk_part, p->outdeg, r_in, r_out, x_out, p->blk_stat, p->nblk, p->nhub, p->fin_stat,
                       p->fin_counter, d_stats, c);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

}  // namespace gchi

extern "C" int gchi_pr_init(const gchi_pr_plan *p, float init, float *d_x, float *d_r, void *stream) {
    GCHI_CHECK(p && d_x && d_r, GCHI_ERR_INVALID, "pr_init: null argument");
    return pr_init(p, init, d_x, d_r, as_stream(stream));
}

extern "C" int gchi_pr_sweep(const gchi_pr_plan *p, const gchi_pr_op *op, const float *d_x_in, const float *d_r_in,
                             float *d_x_out, float *d_r_out, double *d_stats, void *stream) {
    GCHI_CHECK(p && op && d_x_in && d_r_in && d_x_out && d_r_out && d_stats, GCHI_ERR_INVALID,
               "pr_sweep: null argument");
    GCHI_CHECK(d_x_in != d_x_out && d_r_in != d_r_out, GCHI_ERR_INVALID, "pr_sweep: in and out buffers must differ");
    return pr_sweep(p, op, d_x_in, d_r_in, d_x_out, d_r_out, d_stats, as_stream(stream));
}

// ---------------------------------------------------------------- graph exec
struct gchi_pr_exec {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nsweeps = 0;
};

extern "C" int gchi_pr_exec_create(const gchi_pr_plan *p, const gchi_pr_op *op, float init, int nsweeps, float *d_x0,
                                   float *d_x1, float *d_r0, float *d_r1, double *d_stats_all, gchi_pr_exec **out) {
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
        int s1 = pr_init(p, init, xs[0], rs[0], cs);
        for (int k = 0; k < nsweeps && s1 == GCHI_OK; k++)
            s1 = pr_sweep(p, op, xs[k & 1], rs[k & 1], xs[(k + 1) & 1], rs[(k + 1) & 1], d_stats_all + 2 * k, cs);
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
