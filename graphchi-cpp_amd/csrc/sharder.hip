// Device sharder: reference-format shards from an edge list resident in HBM.
//
// Reference pipeline (src/preprocessing/sharder.hpp): shovels sorted by dst
// (:161-202), k-way merged into one dst-ordered stream (:938-981), cut into
// intervals by a streaming rule (add/createnextshard/done, :861-927), each
// shard stably sorted by src (:640) and encoded (finish_shard, :620-848).
//
// Device form (same bytes, different route):
//   1. in/out-degree histograms (the _degs.bin content, :748-751);
//   2. the streaming cut replayed exactly on the in-degree prefix P (P[v] =
//      edges with dst < v): with S edges already sharded and target eps, the
//      shard closes at the first dst u with P[u+1] >= S + eps provided a later
//      edge exists; eps is then rebalanced to (E - S) / (remaining shards).
//      One thread, one binary search per shard;
//   3. one onesweep radix sort of the composite key (shard, src, dst) -- the
//      stable src sort of each dst-ordered shard gives exactly (src, dst)
//      order, since records with equal (src, dst) are identical;
//   4. encoding: runs of equal src are found by a flag scan; each run's byte
//      size (zero-run bytes for the src gap before it, a 1- or 5-byte count,
//      4 bytes per dst) is scanned into offsets and written in parallel.
//      .adjidx entries (every >= 1M edges at a run boundary, :765-771) are
//      found greedily by binary search over run starts.
#include <zlib.h>

#include "common.hpp"
#include "internal.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <sys/stat.h>
#include <vector>

namespace gchi {

constexpr uint32_t kIndexInterval = 1024 * 1024;  // sharder.hpp:674

__global__ void sh_degrees_interleave(const uint32_t *indeg, const uint32_t *outdeg, uint32_t n, int32_t *degs) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
        degs[2 * v] = (int32_t)indeg[v];
        degs[2 * v + 1] = (int32_t)outdeg[v];
    }
}

// Replays createnextshard/done on the in-degree prefix P[0..n] (P[n] = E).
// ends[k], counts[k] per created shard; *nshards_out = shards created;
// status: 0 ok, 1 the reference would assert (more shards than requested).
__global__ void sh_interval_cut(const uint64_t *P, uint32_t n, uint32_t max_id, int nshards, uint32_t *ends,
                                uint64_t *counts, int *nshards_out, int *status) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const uint64_t E = P[n];
    uint64_t eps = E / (uint64_t)nshards + 1;
    uint64_t S = 0;
    int shardnum = 0;
    // last dst with edges (prevvid at done()); 0xffffffff when there are none
    uint32_t last_dst = 0xffffffffu;
    if (E > 0) {
        uint32_t lo = 0, hi = n;  // largest v with P[v] < E
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo + 1) / 2;
            if (P[mid] < E) lo = mid;
            else hi = mid - 1;
        }
        last_dst = lo;
    }
    *status = 0;
    for (;;) {
        // smallest t in [1, n] with P[t] >= S + eps
        // a shard always holds at least the dst group that opened it
        uint64_t target = S + (eps > 0 ? eps : 1);
        bool close_mid = false;
        uint32_t t = 0;
        if (target <= E) {
            uint32_t lo = 1, hi = n;
            while (lo < hi) {
                uint32_t mid = lo + (hi - lo) / 2;
                if (P[mid] >= target) hi = mid;
                else lo = mid + 1;
            }
            t = lo;
            close_mid = P[t] < E;  // a later edge exists to trigger the close
        }
        if (!close_mid) break;
        if (shardnum >= nshards) {
            *status = 1;
            return;
        }
        ends[shardnum] = (shardnum == nshards - 1) ? max_id : (t - 1);
        counts[shardnum] = P[t] - S;
        shardnum++;
        S = P[t];
        if (shardnum < nshards) eps = (E - S) / (uint64_t)(nshards - shardnum);
    }
    if (shardnum >= nshards) {
        *status = 1;
        return;
    }
    ends[shardnum] = (shardnum == nshards - 1) ? max_id : last_dst;
    counts[shardnum] = E - S;
    shardnum++;
    *nshards_out = shardnum;
}

// (src << db) | dst, db = the id bits rounded up to a whole radix digit, so that
// the first db/8 sort passes order the edges by dst alone.
__global__ void sh_pack(const uint32_t *src, const uint32_t *dst, uint64_t e, uint32_t nv, int db, uint64_t *keys,
                        uint32_t *err) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        const uint32_t sv = src[i], dv = dst[i];
        if (sv >= nv || dv >= nv) atomicOr(err, 1u);
        keys[i] = ((uint64_t)sv << db) | dv;
    }
}

// (src << db | dst) -> (shard << 2*idb | src << idb | dst), shard = the first
// interval whose end is >= dst (the last one for larger ids).
__global__ void sh_rekey(uint64_t *keys, uint64_t e, int db, int idb, const uint32_t *ends, int ns) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t m = (1ull << idb) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        const uint64_t k = keys[i];
        const uint32_t d = (uint32_t)(k & m), sv = (uint32_t)((k >> db) & m);
        int lo = 0, hi = ns - 1;
        while (lo < hi) {
            const int mid = (lo + hi) / 2;
            if (d <= ends[mid]) hi = mid;
            else lo = mid + 1;
        }
        keys[i] = ((uint64_t)lo << (2 * idb)) | ((uint64_t)sv << idb) | d;
    }
}

__global__ void sh_unpack(const uint64_t *keys, uint64_t e, int idb, uint32_t *src, uint32_t *dst) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t m = (1ull << idb) - 1ull;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint64_t k = keys[i];
        dst[i] = (uint32_t)(k & m);
        src[i] = (uint32_t)((k >> idb) & m);
    }
}

// run-start flags over one shard [0, e) of (src,dst)-sorted records
__global__ void sh_run_flags(const uint32_t *src, uint64_t e, uint32_t *flags) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride)
        flags[i] = (i == 0 || src[i] != src[i - 1]) ? 1u : 0u;
}

__global__ void sh_run_starts(const uint32_t *flags, const uint32_t *ridx, uint64_t e, uint32_t *run_start) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride)
        if (flags[i]) run_start[ridx[i]] = (uint32_t)i;
}

GCHI_HD uint64_t zero_bytes(uint32_t gap) { return gap ? 2ull * ((gap + 254u) / 255u) : 0ull; }

// byte size of run k (zeros before it + count + dsts); run_start[nruns] = e
__global__ void sh_run_sizes(const uint32_t *src, const uint32_t *run_start, uint32_t nruns, uint32_t *sizes) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nruns; k += stride) {
        uint32_t b = run_start[k], c = run_start[k + 1] - b;
        uint32_t s = src[b];
        uint32_t gap = k == 0 ? s : (s - src[run_start[k - 1]] - 1);
        uint64_t sz = zero_bytes(gap) + (c < 255 ? 1 : 5) + 4ull * c;
        sizes[k] = (uint32_t)sz;
    }
}

// headers and zero runs (one thread per run)
__global__ void sh_write_heads(const uint32_t *src, const uint32_t *run_start, uint32_t nruns, const uint64_t *off,
                               uint8_t *adj) {
    uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nruns; k += stride) {
        uint32_t b = run_start[k], c = run_start[k + 1] - b;
        uint32_t s = src[b];
        uint32_t gap = k == 0 ? s : (s - src[run_start[k - 1]] - 1);
        uint64_t p = off[k];
        int64_t nz = gap;
        while (nz > 0) {  // sharder.hpp:808-818 with nz >= 1
            adj[p++] = 0;
            nz--;
            int64_t t = nz < 254 ? nz : 254;
            adj[p++] = (uint8_t)t;
            nz -= t;
        }
        if (c < 255) {
            adj[p++] = (uint8_t)c;
        } else {
            adj[p++] = 0xff;
            adj[p++] = (uint8_t)(c & 0xff);
            adj[p++] = (uint8_t)((c >> 8) & 0xff);
            adj[p++] = (uint8_t)((c >> 16) & 0xff);
            adj[p++] = (uint8_t)(c >> 24);
        }
    }
}

// dst words (one thread per edge; unaligned little-endian bytes)
__global__ void sh_write_dsts(const uint32_t *src, const uint32_t *dst, const uint32_t *flags, const uint32_t *ridx,
                              const uint32_t *run_start, uint64_t e, const uint64_t *off, uint8_t *adj) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride) {
        uint32_t k = ridx[i] + flags[i] - 1;  // run of record i
        uint32_t b = run_start[k], c = run_start[k + 1] - b;
        uint32_t s = src[b];
        uint32_t gap = k == 0 ? s : (s - src[run_start[k - 1]] - 1);
        uint64_t p = off[k] + zero_bytes(gap) + (c < 255 ? 1 : 5) + 4ull * (i - b);
        uint32_t d = dst[i];
        adj[p] = (uint8_t)(d & 0xff);
        adj[p + 1] = (uint8_t)((d >> 8) & 0xff);
        adj[p + 2] = (uint8_t)((d >> 16) & 0xff);
        adj[p + 3] = (uint8_t)(d >> 24);
    }
}

// .adjidx: greedy over run starts (first run start >= last + 1M), one thread
__global__ void sh_index(const uint32_t *src, const uint32_t *run_start, uint32_t nruns, const uint64_t *off,
                         uint64_t *triples, uint32_t max_entries, uint32_t *count) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    uint64_t last = 0;
    uint32_t ni = 0;
    uint32_t lo_k = 0;
    for (;;) {
        uint64_t target = last + kIndexInterval;
        uint32_t lo = lo_k, hi = nruns;  // smallest k with run_start[k] >= target
        while (lo < hi) {
            uint32_t mid = lo + (hi - lo) / 2;
            if (run_start[mid] >= target) hi = mid;
            else lo = mid + 1;
        }
        if (lo >= nruns || ni >= max_entries) break;
        uint32_t k = lo, b = run_start[k], s = src[b];
        uint32_t gap = k == 0 ? s : (s - src[run_start[k - 1]] - 1);
        triples[3 * ni] = s;
        triples[3 * ni + 1] = off[k] + zero_bytes(gap);
        triples[3 * ni + 2] = b;
        ni++;
        last = b;
        lo_k = k + 1;
    }
    *count = ni;
}

static int sh_read(void *h, const void *d, size_t n, hipStream_t s) {
    GCHI_HIP(hipMemcpyAsync(h, d, n, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    return GCHI_OK;
}

}  // namespace gchi

using namespace gchi;

struct gchi_shards {
    int nshards_req = 0;       // nshards requested (names the files)
    int nshards = 0;           // shards created
    uint32_t max_id = 0;
    uint64_t nedges = 0;
    int idb = 1;
    std::vector<uint32_t> ends;
    std::vector<uint64_t> counts, first;
    uint32_t *src = nullptr;   // [E] shard-major, (src,dst) inside a shard
    uint32_t *dst = nullptr;
    int32_t *degs = nullptr;   // [2N] {indegree, outdegree}
    void *mem = nullptr;
};

extern "C" int gchi_shard_auto_count(uint64_t nedges, uint32_t edge_value_bytes, int membudget_mb, int *nshards) {
    GCHI_CHECK(nshards && membudget_mb > 0, GCHI_ERR_INVALID, "shard_auto_count: bad argument");
    double max_shardsize = membudget_mb * 1024. * 1024. / 8;  // sharder.hpp:572-575
    *nshards = (int)(1 + (nedges * (double)edge_value_bytes / max_shardsize) + 0.5);
    return GCHI_OK;
}

extern "C" int gchi_shard_build(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t max_id,
                                int nshards, gchi_shards **out, void *stream) {
    GCHI_CHECK(out && nshards >= 1 && nshards <= 65536, GCHI_ERR_INVALID, "shard_build: bad argument");
    GCHI_CHECK(nedges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "shard_build: null edges");
    GCHI_CHECK(nedges <= kMaxEdges, GCHI_ERR_UNSUPPORTED, "shard_build: E must be <= 2^32 - 2^25");
    hipStream_t s = as_stream(stream);
    const uint32_t N = max_id + 1;
    GCHI_CHECK(N > max_id, GCHI_ERR_UNSUPPORTED, "shard_build: max_id too large");
    int idb = bits_for(max_id);
    if (idb == 0) idb = 1;
    const int pb = bits_for((uint64_t)nshards - 1);
    GCHI_CHECK(pb + 2 * idb <= 64, GCHI_ERR_UNSUPPORTED, "shard_build: key exceeds 64 bits");
    const uint64_t E = nedges;

    gchi_shards *sh = new gchi_shards();
    sh->nshards_req = nshards;
    sh->max_id = max_id;
    sh->nedges = E;
    sh->idb = idb;
    Sizer zz;
    zz.take<uint32_t>(E + 1);
    zz.take<uint32_t>(E + 1);
    zz.take<int32_t>(2 * (size_t)N);
    if (hipMalloc(&sh->mem, zz.off) != hipSuccess) {
        delete sh;
        return fail(GCHI_ERR_HIP, "shard_build: hipMalloc failed");
    }
    Carver c(sh->mem, zz.off);
    sh->src = c.take<uint32_t>(E + 1);
    sh->dst = c.take<uint32_t>(E + 1);
    sh->degs = c.take<int32_t>(2 * (size_t)N);

    // scratch
    Sizer z;
    z.take<uint32_t>(N);            // indeg
    z.take<uint32_t>(N);            // outdeg
    z.take<uint32_t>(N);            // run starts
    z.take<uint64_t>((size_t)N + 1);  // P
    z.take<uint32_t>(nshards);      // ends
    z.take<uint64_t>(nshards);      // counts
    z.take<int>(4);                 // nshards_out, status
    z.take<uint32_t>(1);            // err
    z.take<uint64_t>(E + 1);        // keys
    const int db = (idb + 7) / 8 * 8;
    size_t swb = std::max(sort_workspace_bytes(E, std::max(db, pb), false), scan_workspace((size_t)N + 1));
    z.take<char>(swb);
    void *scratch = nullptr;
    if (hipMalloc(&scratch, z.off) != hipSuccess) {
        (void)hipFree(sh->mem);
        delete sh;
        return fail(GCHI_ERR_HIP, "shard_build: hipMalloc failed");
    }
    Carver sc(scratch, z.off);
    uint32_t *indeg = sc.take<uint32_t>(N);
    uint32_t *outdeg = sc.take<uint32_t>(N);
    uint32_t *run_start = sc.take<uint32_t>(N);
    uint64_t *P = sc.take<uint64_t>((size_t)N + 1);
    uint32_t *d_ends = sc.take<uint32_t>(nshards);
    uint64_t *d_counts = sc.take<uint64_t>(nshards);
    int *d_ns = sc.take<int>(4);
    uint32_t *err = sc.take<uint32_t>(1);
    uint64_t *keys = sc.take<uint64_t>(E + 1);
    void *sw = sc.take<char>(swb);
    auto done = [&](int st) {
        (void)hipFree(scratch);
        if (st != GCHI_OK) {
            (void)hipFree(sh->mem);
            delete sh;
        } else {
            *out = sh;
        }
        return st;
    };
    // Degrees from runs of equal ids in sorted keys (device-wide atomics on
    // random counters would be far slower): the dst passes give in-degrees,
    // the src passes the (src, dst) order and out-degrees.
    launch_fill_u32(err, 0u, 1, s);
    launch_fill_u32(indeg, 0u, N, s);
    launch_fill_u32(outdeg, 0u, N, s);
    const uint32_t idmask = idb >= 32 ? 0xffffffffu : ((1u << idb) - 1u);
    uint64_t *by_src_dst = keys;
    int st = GCHI_OK;
    if (E) {
        hipLaunchKernelGGL(sh_pack, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, d_src, d_dst, E, N, db, keys, err);
        uint64_t *by_dst = nullptr;
        st = radix_sort_u64_bits(keys, nullptr, nullptr, E, 0, db, sw, swb, s, &by_dst, err);
        if (st == GCHI_OK) st = count_field_runs(by_dst, E, 0, idmask, N, run_start, indeg, nullptr, 0u, s);
        if (st == GCHI_OK)
            st = radix_sort_u64_bits(by_dst, by_dst == keys ? nullptr : keys, nullptr, E, db, db + idb, sw, swb, s,
                                     &by_src_dst, err);
        if (st == GCHI_OK) st = count_field_runs(by_src_dst, E, db, idmask, N, run_start, outdeg, nullptr, 0u, s);
    }
    if (st == GCHI_OK) {
        hipLaunchKernelGGL(sh_degrees_interleave, dim3(grid_for(N, 256, 65536)), dim3(256), 0, s, indeg, outdeg, N,
                           sh->degs);
        st = exclusive_scan_u32_to_u64(indeg, P, N, sw, swb, s);
    }
    if (st == GCHI_OK) {
        hipLaunchKernelGGL(sh_interval_cut, dim3(1), dim3(64), 0, s, P, N, max_id, nshards, d_ends, d_counts, d_ns,
                           d_ns + 1);
        int h_ns[2] = {0, 0};
        uint32_t h_err = 0;
        st = sh_read(h_ns, d_ns, sizeof(h_ns), s);
        if (st == GCHI_OK) st = sh_read(&h_err, err, 4, s);
        if (st == GCHI_OK && (h_err & GCHI_WORK_TIMEOUT))
            st = fail(GCHI_ERR_TIMEOUT, "shard_build: a sort look-back wait gave up");
        if (st == GCHI_OK && h_err) st = fail(GCHI_ERR_INVALID, "shard_build: vertex id > max_id");
        if (st == GCHI_OK && h_ns[1]) st = fail(GCHI_ERR_FORMAT, "shard_build: the streaming cut needs more shards "
                                                                 "than requested (the reference asserts here)");
        if (st == GCHI_OK) {
            sh->nshards = h_ns[0];
            sh->ends.resize(sh->nshards);
            sh->counts.resize(sh->nshards);
            st = sh_read(sh->ends.data(), d_ends, 4 * sh->nshards, s);
            if (st == GCHI_OK) st = sh_read(sh->counts.data(), d_counts, 8 * sh->nshards, s);
        }
    }
    if (st == GCHI_OK && E) {
        // stable pass on the shard id: (shard, src, dst) order
        hipLaunchKernelGGL(sh_rekey, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, by_src_dst, E, db, idb, d_ends,
                           sh->nshards);
        uint64_t *sorted = nullptr;
        st = radix_sort_u64_bits(by_src_dst, by_src_dst == keys ? nullptr : keys, nullptr, E, 2 * idb,
                                 2 * idb + pb, sw, swb, s, &sorted, err);
        if (st == GCHI_OK)
            hipLaunchKernelGGL(sh_unpack, dim3(grid_for(E, 256, 65536)), dim3(256), 0, s, sorted, E, idb, sh->src,
                               sh->dst);
        if (st == GCHI_OK) st = read_work_status(err, s);  // waits for the stream
        hipError_t e = hipStreamSynchronize(s);
        if (st == GCHI_OK && e != hipSuccess) st = hip_fail(e, "shard_build", __FILE__, __LINE__);
    }
    sh->first.resize(sh->nshards + 1);
    sh->first[0] = 0;
    for (int p = 0; p < sh->nshards; p++) sh->first[p + 1] = sh->first[p] + sh->counts[p];
    return done(st);
}

extern "C" int gchi_shard_destroy(gchi_shards *sh) {
    if (!sh) return GCHI_OK;
    if (sh->mem) (void)hipFree(sh->mem);
    delete sh;
    return GCHI_OK;
}

extern "C" int gchi_shard_info(const gchi_shards *sh, int *nshards_created, uint32_t *h_ends, uint64_t *h_counts) {
    GCHI_CHECK(sh && nshards_created, GCHI_ERR_INVALID, "shard_info: null argument");
    *nshards_created = sh->nshards;
    if (h_ends) memcpy(h_ends, sh->ends.data(), 4 * sh->ends.size());
    if (h_counts) memcpy(h_counts, sh->counts.data(), 8 * sh->counts.size());
    return GCHI_OK;
}

extern "C" int gchi_shard_edges(const gchi_shards *sh, int shard, const uint32_t **d_src, const uint32_t **d_dst,
                                uint64_t *count) {
    GCHI_CHECK(sh && d_src && d_dst && count, GCHI_ERR_INVALID, "shard_edges: null argument");
    GCHI_CHECK(shard >= 0 && shard < sh->nshards, GCHI_ERR_INVALID, "shard_edges: shard out of range");
    *d_src = sh->src + sh->first[shard];
    *d_dst = sh->dst + sh->first[shard];
    *count = sh->counts[shard];
    return GCHI_OK;
}

extern "C" int gchi_shard_degrees(const gchi_shards *sh, const int32_t **d_degs) {
    GCHI_CHECK(sh && d_degs, GCHI_ERR_INVALID, "shard_degrees: null argument");
    *d_degs = sh->degs;
    return GCHI_OK;
}

// Encodes one shard's adjacency.  h_adj == NULL: only *adj_bytes is returned.
extern "C" int gchi_shard_encode(const gchi_shards *sh, int shard, uint8_t *h_adj, uint64_t *adj_bytes,
                                 uint64_t *h_index, uint32_t max_index, uint32_t *nindex, void *stream) {
    GCHI_CHECK(sh && adj_bytes, GCHI_ERR_INVALID, "shard_encode: null argument");
    GCHI_CHECK(shard >= 0 && shard < sh->nshards, GCHI_ERR_INVALID, "shard_encode: shard out of range");
    hipStream_t s = as_stream(stream);
    const uint64_t e = sh->counts[shard];
    const uint32_t *src = sh->src + sh->first[shard];
    const uint32_t *dst = sh->dst + sh->first[shard];
    if (nindex) *nindex = 0;
    if (e == 0) {
        *adj_bytes = 0;
        return GCHI_OK;
    }
    Sizer z;
    z.take<uint32_t>(e + 1);  // flags
    z.take<uint32_t>(e + 1);  // ridx (exclusive scan of flags)
    z.take<uint32_t>(e + 1);  // run_start
    z.take<uint32_t>(e + 1);  // sizes
    z.take<uint64_t>(e + 2);  // offsets
    z.take<uint64_t>(3 * (size_t)(max_index + 1));
    z.take<uint32_t>(1);
    size_t swb = scan_workspace(e + 1);
    z.take<char>(swb);
    z.take<uint8_t>(1);
    void *scratch = nullptr;
    GCHI_HIP(hipMalloc(&scratch, z.off));
    Carver c(scratch, z.off);
    uint32_t *flags = c.take<uint32_t>(e + 1);
    uint32_t *ridx = c.take<uint32_t>(e + 1);
    uint32_t *rstart = c.take<uint32_t>(e + 1);
    uint32_t *sizes = c.take<uint32_t>(e + 1);
    uint64_t *off = c.take<uint64_t>(e + 2);
    uint64_t *d_idx = c.take<uint64_t>(3 * (size_t)(max_index + 1));
    uint32_t *d_nidx = c.take<uint32_t>(1);
    void *sw = c.take<char>(swb);
    uint8_t *adj = nullptr;
    int st = GCHI_OK;
    uint32_t nruns = 0;
    uint64_t total = 0;
    hipLaunchKernelGGL(sh_run_flags, dim3(grid_for(e, 256, 65536)), dim3(256), 0, s, src, e, flags);
    st = exclusive_scan_u32(flags, ridx, e, sw, swb, s);
    if (st == GCHI_OK) st = sh_read(&nruns, ridx + e, 4, s);
    if (st == GCHI_OK) {
        hipLaunchKernelGGL(sh_run_starts, dim3(grid_for(e, 256, 65536)), dim3(256), 0, s, flags, ridx, e, rstart);
        uint32_t ee = (uint32_t)e;
        if (hipMemcpyAsync(rstart + nruns, &ee, 4, hipMemcpyHostToDevice, s) != hipSuccess)
            st = fail(GCHI_ERR_HIP, "shard_encode: copy failed");
    }
    if (st == GCHI_OK) {
        hipLaunchKernelGGL(sh_run_sizes, dim3(grid_for(nruns, 256, 65536)), dim3(256), 0, s, src, rstart, nruns,
                           sizes);
        st = exclusive_scan_u32_to_u64(sizes, off, nruns, sw, swb, s);
    }
    if (st == GCHI_OK) st = sh_read(&total, off + nruns, 8, s);
    *adj_bytes = total;
    if (st == GCHI_OK && h_adj) {
        if (hipMalloc(&adj, total) != hipSuccess) st = fail(GCHI_ERR_HIP, "shard_encode: hipMalloc failed");
        if (st == GCHI_OK) {
            hipLaunchKernelGGL(sh_write_heads, dim3(grid_for(nruns, 256, 65536)), dim3(256), 0, s, src, rstart, nruns,
                               off, adj);
            hipLaunchKernelGGL(sh_write_dsts, dim3(grid_for(e, 256, 65536)), dim3(256), 0, s, src, dst, flags, ridx,
                               rstart, e, off, adj);
            if (h_index && nindex)
                hipLaunchKernelGGL(sh_index, dim3(1), dim3(64), 0, s, src, rstart, nruns, off, d_idx, max_index,
                                   d_nidx);
            hipError_t le = hipGetLastError();
            if (le != hipSuccess) st = hip_fail(le, "shard_encode kernels", __FILE__, __LINE__);
        }
        if (st == GCHI_OK) st = sh_read(h_adj, adj, total, s);
        if (st == GCHI_OK && h_index && nindex) {
            st = sh_read(nindex, d_nidx, 4, s);
            if (st == GCHI_OK && *nindex) st = sh_read(h_index, d_idx, 24 * (size_t)*nindex, s);
        }
    }
    if (adj) (void)hipFree(adj);
    (void)hipFree(scratch);
    return st;
}

// ------------------------------------------------------------------ files
namespace {

std::string part(int p, int n) { return "." + std::to_string(p) + "_" + std::to_string(n); }

int write_file(const std::string &path, const void *data, size_t n) {
    FILE *f = fopen(path.c_str(), "wb");
    if (!f) return fail(GCHI_ERR_IO, "cannot open " + path);
    size_t w = n ? fwrite(data, 1, n, f) : 0;
    int rc = fclose(f);
    if (w != n || rc != 0) return fail(GCHI_ERR_IO, "short write to " + path);
    return GCHI_OK;
}

}  // namespace

// Writes <base>.<P>.intervals, <base>.numvertices, <base>_degs.bin and, per
// shard, <base>.edata_azv.<p>_<P>.adj/.adjidx plus the edge-data block
// directory (<edata>_blockdir_<B>/<id>, default-valued records of
// edge_value_bytes, zlib-compressed with GCHI_SHARD_ZLIB) and <edata>.size --
// the layout of src/api/chifilenames.hpp:65-153 that find_shards (:200-275)
// accepts.
//
// zlib deflate of one edge-data block at level 1 (Z_BEST_SPEED), one
// Z_FINISH stream: write_compressed of src/util/ioutil.hpp:144-199.
static int deflate_block(const uint8_t *in, size_t n, std::vector<uint8_t> &out) {
    z_stream z;
    memset(&z, 0, sizeof(z));
    if (deflateInit(&z, Z_BEST_SPEED) != Z_OK) return fail(GCHI_ERR_IO, "shard_write_files: deflateInit failed");
    out.resize(deflateBound(&z, (uLong)n) + 64);
    z.next_in = const_cast<Bytef *>(in);
    z.avail_in = (uInt)n;
    z.next_out = out.data();
    z.avail_out = (uInt)out.size();
    const int r = deflate(&z, Z_FINISH);
    const size_t have = out.size() - z.avail_out;
    deflateEnd(&z);
    if (r != Z_STREAM_END) return fail(GCHI_ERR_IO, "shard_write_files: deflate did not finish");
    out.resize(have);
    return GCHI_OK;
}

extern "C" int gchi_shard_write_files(const gchi_shards *sh, const char *basefilename, uint32_t edge_value_bytes,
                                      void *stream) {
    return gchi_shard_write_files_ex(sh, basefilename, edge_value_bytes, 0, stream);
}

extern "C" int gchi_shard_write_files_ex(const gchi_shards *sh, const char *basefilename, uint32_t edge_value_bytes,
                                         int flags, void *stream) {
    GCHI_CHECK(sh && basefilename, GCHI_ERR_INVALID, "shard_write_files: null argument");
    GCHI_CHECK((flags & ~GCHI_SHARD_ZLIB) == 0, GCHI_ERR_INVALID, "shard_write_files: unknown flags");
    const bool zlib = (flags & GCHI_SHARD_ZLIB) != 0;
    const std::string base(basefilename);
    const int P = sh->nshards_req;
    hipStream_t s = as_stream(stream);
    {  // intervals: one line per created shard (sharder.hpp:917-919)
        std::string txt;
        for (int p = 0; p < sh->nshards; p++) txt += std::to_string(sh->ends[p]) + "\n";
        GCHI_TRY(write_file(base + "." + std::to_string(P) + ".intervals", txt.data(), txt.size()));
        std::string nv = std::to_string((uint64_t)sh->max_id + 1) + "\n";
        GCHI_TRY(write_file(base + ".numvertices", nv.data(), nv.size()));
    }
    {
        std::vector<int32_t> degs(2 * ((size_t)sh->max_id + 1));
        GCHI_TRY(sh_read(degs.data(), sh->degs, degs.size() * 4, s));
        GCHI_TRY(write_file(base + "_degs.bin", degs.data(), degs.size() * 4));
    }
    size_t block = 1024 * 1024;
    if (edge_value_bytes)
        while (block % edge_value_bytes) block++;
    for (int p = 0; p < sh->nshards; p++) {
        uint64_t bytes = 0;
        GCHI_TRY(gchi_shard_encode(sh, p, nullptr, &bytes, nullptr, 0, nullptr, stream));
        std::vector<uint8_t> adj(bytes ? bytes : 1);
        std::vector<uint64_t> idx(3 * 4096);
        uint32_t nidx = 0;
        GCHI_TRY(gchi_shard_encode(sh, p, adj.data(), &bytes, idx.data(), 4096, &nidx, stream));
        std::string adjname = base + ".edata_azv" + part(p, P) + ".adj";
        GCHI_TRY(write_file(adjname, adj.data(), bytes));
        GCHI_TRY(write_file(adjname + "idx", idx.data(), 24 * (size_t)nidx));
        if (edge_value_bytes) {
            // filename_shard_edata (src/api/chifilenames.hpp:91-105): ".edata." then
            // ".Z." when compressed, hence the double dot
            std::string ed = base + ".edata." + (zlib ? ".Z." : "") + "e" + std::to_string(edge_value_bytes) + "B" +
                             part(p, P);
            uint64_t total = sh->counts[p] * edge_value_bytes;
            std::string sz = std::to_string(total);
            GCHI_TRY(write_file(ed + ".size", sz.data(), sz.size()));
            std::string dir = ed + "_blockdir_" + std::to_string(block);
            mkdir(dir.c_str(), 0777);
            std::vector<uint8_t> zeros(block, 0), packed;
            uint64_t nblocks = total ? (total + block - 1) / block : 1;
            for (uint64_t b = 0; b < nblocks; b++) {
                uint64_t len = std::min<uint64_t>(block, total - std::min<uint64_t>(total, b * block));
                if (zlib) {
                    GCHI_TRY(deflate_block(zeros.data(), len, packed));
                    GCHI_TRY(write_file(dir + "/" + std::to_string(b), packed.data(), packed.size()));
                } else {
                    GCHI_TRY(write_file(dir + "/" + std::to_string(b), zeros.data(), len));
                }
            }
        }
    }
    return GCHI_OK;
}

// ------------------------------------------------------------------ edge values
// A shard's edge values in its edge-data block directory (the reference's
// memory_shard load and commit, src/shards/memoryshard.hpp:160-210, 268-300,
// with read_compressed / write_compressed, src/util/ioutil.hpp:144-265): blocks
// of B bytes (1 MiB rounded up to a multiple of the value size; the last one
// shorter), each one zlib stream when the file is compressed, and
// <edata>.size holding the total.
namespace {

size_t edata_block_bytes(uint32_t vb) {
    size_t b = 1024 * 1024;
    while (b % vb) b++;
    return b;
}

int read_whole(const std::string &path, std::vector<uint8_t> &out) {
    FILE *f = fopen(path.c_str(), "rb");
    if (!f) return fail(GCHI_ERR_IO, "cannot open " + path);
    out.clear();
    uint8_t buf[1 << 16];
    size_t r;
    while ((r = fread(buf, 1, sizeof(buf), f)) > 0) out.insert(out.end(), buf, buf + r);
    const bool bad = ferror(f) != 0;
    fclose(f);
    if (bad) return fail(GCHI_ERR_IO, "read error on " + path);
    return GCHI_OK;
}

}  // namespace

extern "C" int gchi_edata_read(const char *edata_name, int compressed, uint32_t value_bytes, uint64_t nvalues,
                               void *host_out) {
    GCHI_CHECK(edata_name && value_bytes && (host_out || nvalues == 0), GCHI_ERR_INVALID, "edata_read: bad argument");
    const std::string ed(edata_name);
    const uint64_t total = nvalues * value_bytes;
    std::vector<uint8_t> buf;
    GCHI_TRY(read_whole(ed + ".size", buf));
    buf.push_back(0);
    const unsigned long long stated = strtoull(reinterpret_cast<const char *>(buf.data()), nullptr, 10);
    GCHI_CHECK(stated == total, GCHI_ERR_FORMAT,
               ed + ".size states " + std::to_string(stated) + " bytes, the shard has " + std::to_string(total));
    const size_t B = edata_block_bytes(value_bytes);
    const std::string dir = ed + "_blockdir_" + std::to_string(B);
    uint8_t *out = static_cast<uint8_t *>(host_out);
    for (uint64_t b = 0; b * B < total; b++) {
        const size_t len = (size_t)std::min<uint64_t>(B, total - b * B);
        const std::string path = dir + "/" + std::to_string(b);
        GCHI_TRY(read_whole(path, buf));
        if (!compressed) {
            GCHI_CHECK(buf.size() == len, GCHI_ERR_FORMAT, path + ": block of the wrong size");
            memcpy(out + b * B, buf.data(), len);
            continue;
        }
        z_stream z;
        memset(&z, 0, sizeof(z));
        GCHI_CHECK(inflateInit(&z) == Z_OK, GCHI_ERR_IO, "edata_read: inflateInit failed");
        z.next_in = buf.data();
        z.avail_in = (uInt)buf.size();
        z.next_out = out + b * B;
        z.avail_out = (uInt)len;
        const int r = inflate(&z, Z_FINISH);
        const size_t have = len - z.avail_out;
        inflateEnd(&z);
        GCHI_CHECK(r == Z_STREAM_END && have == len, GCHI_ERR_FORMAT, path + ": not a zlib block of the shard's size");
    }
    return GCHI_OK;
}

extern "C" int gchi_edata_write(const char *edata_name, int compressed, uint32_t value_bytes, uint64_t nvalues,
                                const void *host_in) {
    GCHI_CHECK(edata_name && value_bytes && (host_in || nvalues == 0), GCHI_ERR_INVALID, "edata_write: bad argument");
    const std::string ed(edata_name);
    const uint64_t total = nvalues * value_bytes;
    const size_t B = edata_block_bytes(value_bytes);
    const std::string dir = ed + "_blockdir_" + std::to_string(B);
    mkdir(dir.c_str(), 0777);
    const std::string sz = std::to_string(total);
    GCHI_TRY(write_file(ed + ".size", sz.data(), sz.size()));
    const uint8_t *in = static_cast<const uint8_t *>(host_in);
    std::vector<uint8_t> packed;
    for (uint64_t b = 0; b * B < total; b++) {
        const size_t len = (size_t)std::min<uint64_t>(B, total - b * B);
        const std::string path = dir + "/" + std::to_string(b);
        if (compressed) {
            GCHI_TRY(deflate_block(in + b * B, len, packed));
            GCHI_TRY(write_file(path, packed.data(), packed.size()));
        } else {
            GCHI_TRY(write_file(path, in + b * B, len));
        }
    }
    return GCHI_OK;
}
