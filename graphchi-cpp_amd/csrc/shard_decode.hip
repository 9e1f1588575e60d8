// Device decode of a shard's adjacency stream (the bytes of a .adj file).
//
// Reference: memory_shard::load_vertices (src/shards/memoryshard.hpp:448-497)
// parses the stream record by record: a 0 byte and the byte nz after it skip
// 1 + nz vertices; a count byte c (0xff: a uint32 count follows) is followed by
// c uint32 dst ids of the current vertex, which then advances by one.  The
// records have no markers, so where one starts depends on every record before
// it.  Device form, for a stream of n bytes cut into 32 KB chunks of four 8 KB
// pieces:
//   1. sd_exits: every byte position p is taken as if a record started there
//      (next(p) = p + its length); one workgroup per chunk resolves, by pointer
//      jumping in LDS, the first position at or beyond the end of p's piece
//      (X8[p]), then of p's chunk (X[p]), that the chain from p reaches.
//   2. sd_chain: one thread walks the chunks from position 0 -- entry[c] is the
//      first record start in chunk c, X[entry[c]] the first in the next chunk
//      the chain reaches -- one dependent load per chunk instead of per record,
//      and checks that the chain ends exactly at n.  sd_pieces then finds each
//      piece's first record start from its chunk's (up to four X8 loads).
//   3. sd_count / sd_emit: each piece parses its own records from its entry,
//      from the piece's bytes in LDS: counts (edges and vertex advance, scanned
//      into bases), then (src, dst) pairs, written through an LDS queue so the
//      workgroup copies many short records' dst words at once.
// The output is the same edge sequence as the reference's parse: the shard's
// edges in stream order, src ids counted from 0.  Several streams (a graph's
// shards) go through one set of launches: a stream table maps every global
// chunk and piece to its stream, one thread per stream walks its chain, and
// the edges land back to back in stream order.
#include "common.hpp"
#include "internal.hpp"

namespace gchi {

constexpr uint32_t kSdChunk = 32768;  // bytes per chunk (the chain walk's step)
constexpr uint32_t kSdPiece = 8192;   // bytes per piece (one record walk each)
constexpr uint32_t kSdPieces = kSdChunk / kSdPiece;
constexpr int kSdThreads = 1024;      // exit-table workgroup
constexpr int kSdCopyThreads = 256;   // count / emit workgroups
constexpr uint32_t kSdNone = 0xffffffffu;
constexpr uint32_t kSdQueue = 1024;   // records per emit batch

// The streams of one call: stream s has n[s] bytes at adj[s]; its chunks are
// global chunks [cbase[s], cbase[s+1]), its pieces [4 cbase[s], 4 cbase[s+1]),
// its positions [pbase[s], pbase[s] + n[s]) of the exit tables.
struct SdStreams {
    const uint64_t *adj, *n, *pbase;
    const uint32_t *cbase;
    uint32_t count;
};

GCHI_DEV uint32_t sd_stream_of(const SdStreams &t, uint32_t g) {  // largest s with cbase[s] <= g
    uint32_t lo = 0, hi = t.count - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (t.cbase[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Position after the record starting at p, or ~0 when it runs past n.
GCHI_DEV uint64_t sd_next(const uint8_t *__restrict__ adj, uint64_t n, uint64_t p) {
    const uint32_t b = adj[p];
    uint64_t q;
    if (b == 0) {
        q = p + 2;
    } else if (b != 0xff) {
        q = p + 1 + 4ull * b;
    } else {
        if (p + 5 > n) return ~0ull;
        const uint32_t c = (uint32_t)adj[p + 1] | ((uint32_t)adj[p + 2] << 8) | ((uint32_t)adj[p + 3] << 16) |
                           ((uint32_t)adj[p + 4] << 24);
        q = p + 5 + 4ull * c;
    }
    return q <= n ? q : ~0ull;
}

// One workgroup per global chunk.
__global__ __launch_bounds__(kSdThreads) void sd_exits(SdStreams t, uint32_t *__restrict__ X8,
                                                       uint32_t *__restrict__ X) {
    __shared__ uint32_t L[kSdChunk];  // chain position relative to the chunk
    __shared__ int s_more[2];
    const uint32_t sid = sd_stream_of(t, blockIdx.x);
    const uint8_t *adj = reinterpret_cast<const uint8_t *>((uintptr_t)t.adj[sid]);
    const uint64_t n = t.n[sid];
    const uint64_t start = (uint64_t)(blockIdx.x - t.cbase[sid]) * kSdChunk;
    uint32_t *x8 = X8 + t.pbase[sid], *x = X + t.pbase[sid];
    const uint32_t len = n - start < kSdChunk ? (uint32_t)(n - start) : kSdChunk;
    for (uint32_t i = threadIdx.x; i < len; i += kSdThreads) {
        const uint64_t q = sd_next(adj, n, start + i);
        L[i] = q == ~0ull ? kSdNone : (uint32_t)(q - start);
    }
    // Phase 0 stops at the end of each position's piece, phase 1 at the end of
    // the chunk.  Each round replaces an inside position by the one it leads
    // to; values only move forward along the chain, so reading an entry that
    // another thread has already advanced this round is harmless.  Records are
    // at least 2 bytes long: at most log2(kSdChunk / 2) + 1 rounds per phase.
    for (int phase = 0; phase < 2; phase++) {
        if (threadIdx.x < 2) s_more[threadIdx.x] = 1;
        __syncthreads();
        for (int round = 0; round < 32 && s_more[round & 1]; round++) {
            if (threadIdx.x == 0) s_more[(round + 1) & 1] = 0;
            __syncthreads();
            int more = 0;
            for (uint32_t i = threadIdx.x; i < len; i += kSdThreads) {
                const uint32_t lim = phase ? len : min(len, (i / kSdPiece + 1) * kSdPiece);
                const uint32_t v = L[i];
                if (v < lim) {
                    const uint32_t y = L[v];
                    L[i] = y;
                    more |= y < lim;
                }
            }
            if (more) s_more[(round + 1) & 1] = 1;
            __syncthreads();
        }
        uint32_t *out = phase ? x : x8;
        for (uint32_t i = threadIdx.x; i < len; i += kSdThreads) {
            const uint32_t v = L[i];
            out[start + i] = v == kSdNone ? kSdNone : (uint32_t)(start + v);
        }
        __syncthreads();
    }
}

// One thread per stream.  status: [0] a malformed stream.
__global__ void sd_chain(SdStreams t, const uint32_t *__restrict__ X, uint32_t *__restrict__ entry,
                         uint32_t *__restrict__ status) {
    const uint32_t sid = blockIdx.x * blockDim.x + threadIdx.x;
    if (sid >= t.count) return;
    const uint64_t n = t.n[sid];
    const uint32_t nch = t.cbase[sid + 1] - t.cbase[sid];
    const uint32_t *x = X + t.pbase[sid];
    uint32_t *en = entry + t.cbase[sid];
    uint64_t s = 0;
    for (uint32_t c = 0; c < nch; c++) {
        const uint64_t end = n < (uint64_t)(c + 1) * kSdChunk ? n : (uint64_t)(c + 1) * kSdChunk;
        if (s >= end) continue;  // chunk c lies inside one record: no record starts there
        en[c] = (uint32_t)s;
        const uint32_t v = x[s];
        if (v == kSdNone) {
            status[0] = 1;
            return;
        }
        s = v;
    }
    if (s != n) status[0] = 1;
}

// First record start of every piece, from its chunk's entry (thread per global chunk).
__global__ void sd_pieces(SdStreams t, uint32_t nch_total, const uint32_t *__restrict__ X8,
                          const uint32_t *__restrict__ entry, uint32_t *__restrict__ pentry) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nch_total) return;
    const uint32_t sid = sd_stream_of(t, g);
    const uint64_t n = t.n[sid];
    const uint32_t c = g - t.cbase[sid];
    const uint32_t *x8 = X8 + t.pbase[sid];
    uint32_t s = entry[g];
    for (uint32_t j = 0; j < kSdPieces; j++) {
        const uint64_t pend = (uint64_t)c * kSdChunk + (uint64_t)(j + 1) * kSdPiece;
        const bool here = s != kSdNone && (uint64_t)s < pend && (uint64_t)s < n;
        pentry[g * kSdPieces + j] = here ? s : kSdNone;
        if (here) s = x8[s];  // a valid chain: X8 is never kSdNone on it
    }
}

// Loads piece c of a stream (its bytes and the 5 after it, when inside the stream) into LDS.
GCHI_DEV uint32_t sd_load_piece(const uint8_t *__restrict__ adj, uint64_t n, uint32_t c, uint8_t *B) {
    const uint64_t start = (uint64_t)c * kSdPiece;
    const uint32_t len = n - start < kSdPiece ? (uint32_t)(n - start) : kSdPiece;
    const uint32_t ext = n - start < kSdPiece + 5 ? (uint32_t)(n - start) : kSdPiece + 5;
    for (uint32_t i = threadIdx.x; i < ext; i += blockDim.x) B[i] = adj[start + i];
    return len;
}

GCHI_DEV uint32_t sd_u32(const uint8_t *B, uint32_t i) {
    return (uint32_t)B[i] | ((uint32_t)B[i + 1] << 8) | ((uint32_t)B[i + 2] << 16) | ((uint32_t)B[i + 3] << 24);
}

// Per global piece: edges and vertex advance of the records starting in it.
__global__ __launch_bounds__(kSdCopyThreads) void sd_count(SdStreams t, const uint32_t *__restrict__ pentry,
                                                           uint32_t *__restrict__ nedge, uint32_t *__restrict__ vadv,
                                                           uint32_t *__restrict__ status) {
    __shared__ uint8_t B[kSdPiece + 8];
    const uint32_t gp = blockIdx.x;
    const uint32_t e0 = pentry[gp];
    if (e0 == kSdNone) {
        if (threadIdx.x == 0) nedge[gp] = vadv[gp] = 0u;
        return;
    }
    const uint32_t sid = sd_stream_of(t, gp / kSdPieces);
    const uint8_t *adj = reinterpret_cast<const uint8_t *>((uintptr_t)t.adj[sid]);
    const uint32_t c = gp - t.cbase[sid] * kSdPieces;  // the stream's own piece index
    const uint32_t len = sd_load_piece(adj, t.n[sid], c, B);
    __syncthreads();
    if (threadIdx.x != 0) return;
    // 64-bit position: a record that runs past the stream (the host then
    // reports the stream as truncated) must not wrap the walk around
    uint64_t p = e0 - (uint64_t)c * kSdPiece;
    uint64_t ne = 0, va = 0;
    while (p < len) {
        const uint32_t b = B[p];
        if (b == 0) {
            va += 1 + (uint32_t)B[p + 1];
            p += 2;
            continue;
        }
        const uint32_t cnt = b != 0xff ? b : sd_u32(B, (uint32_t)p + 1);
        ne += cnt;
        va += 1;
        p += (b != 0xff ? 1u : 5u) + 4ull * cnt;
    }
    if (ne > 0xffffffffull || va > 0xffffffffull) status[1] = 1;
    nedge[gp] = (uint32_t)ne;
    vadv[gp] = (uint32_t)va;
}

// Per global piece: (src, dst) pairs of its records at their global edge
// positions; src = the vertex advance before the record within its stream.
__global__ __launch_bounds__(kSdCopyThreads) void sd_emit(SdStreams t, const uint32_t *__restrict__ pentry,
                                                          const uint64_t *__restrict__ ebase,
                                                          const uint64_t *__restrict__ vbase,
                                                          uint32_t *__restrict__ out_src,
                                                          uint32_t *__restrict__ out_dst) {
    __shared__ uint8_t B[kSdPiece + 8];
    __shared__ uint32_t q_src[kSdQueue];
    __shared__ uint64_t q_pos[kSdQueue];       // byte position of the record's first dst word
    __shared__ uint64_t q_eoff[kSdQueue + 1];  // edge offset of each record (exclusive), + the batch end
    __shared__ uint32_t s_nq;
    __shared__ uint64_t s_p;
    __shared__ uint64_t s_e, s_v;
    __shared__ int s_done;
    const uint32_t gp = blockIdx.x;
    const uint32_t e0 = pentry[gp];
    if (e0 == kSdNone) return;
    const uint32_t sid = sd_stream_of(t, gp / kSdPieces);
    const uint8_t *adj = reinterpret_cast<const uint8_t *>((uintptr_t)t.adj[sid]);
    const uint32_t c = gp - t.cbase[sid] * kSdPieces;
    const uint32_t len = sd_load_piece(adj, t.n[sid], c, B);
    const uint64_t start = (uint64_t)c * kSdPiece;
    if (threadIdx.x == 0) {
        s_p = e0 - start;
        s_e = ebase[gp];
        s_v = vbase[gp] - vbase[t.cbase[sid] * kSdPieces];  // ids restart at 0 in every stream
        s_done = 0;
    }
    __syncthreads();
    for (;;) {
        if (threadIdx.x == 0) {  // walk: queue up to kSdQueue records with edges
            uint64_t p = s_p;
            uint32_t nq = 0;
            uint64_t e = s_e, v = s_v;
            while (p < len && nq < kSdQueue) {
                const uint32_t b = B[p];
                if (b == 0) {
                    v += 1 + (uint32_t)B[p + 1];
                    p += 2;
                    continue;
                }
                const uint32_t cnt = b != 0xff ? b : sd_u32(B, (uint32_t)p + 1);
                const uint32_t hdr = b != 0xff ? 1u : 5u;
                if (cnt) {
                    q_src[nq] = (uint32_t)v;
                    q_pos[nq] = start + p + hdr;
                    q_eoff[nq] = e;
                    nq++;
                }
                e += cnt;
                v += 1;
                p += hdr + 4ull * cnt;
            }
            q_eoff[nq] = e;
            s_nq = nq;
            s_p = p;
            s_e = e;
            s_v = v;
            s_done = p >= len;
        }
        __syncthreads();
        const uint32_t nq = s_nq;
        if (nq) {
            const uint64_t b0 = q_eoff[0], total = q_eoff[nq] - b0;
            for (uint64_t k = threadIdx.x; k < total; k += kSdCopyThreads) {
                const uint64_t eo = b0 + k;
                uint32_t lo = 0, hi = nq - 1;  // the record holding edge eo: largest r with q_eoff[r] <= eo
                while (lo < hi) {
                    const uint32_t mid = (lo + hi + 1) >> 1;
                    if (q_eoff[mid] <= eo) lo = mid;
                    else hi = mid - 1;
                }
                const uint64_t bp = q_pos[lo] + 4ull * (eo - q_eoff[lo]);
                const uint32_t d = (uint32_t)adj[bp] | ((uint32_t)adj[bp + 1] << 8) | ((uint32_t)adj[bp + 2] << 16) |
                                   ((uint32_t)adj[bp + 3] << 24);
                out_src[eo] = q_src[lo];
                out_dst[eo] = d;
            }
        }
        const int done = s_done;
        __syncthreads();  // the queue is rewritten by the next walk
        if (done) break;
    }
}

struct SdWork {
    uint64_t *t_adj, *t_n, *t_pbase;
    uint32_t *t_cbase;
    uint32_t *X8, *X, *entry, *pentry, *nedge, *vadv, *status;
    uint64_t *ebase, *vbase;
    void *scan;
    size_t scan_bytes;
};

struct SdShape {
    uint64_t bytes = 0, chunks = 0;
    uint32_t count = 0;
};

static SdShape sd_shape(const uint64_t *h_nbytes, uint32_t count) {
    SdShape z;
    z.count = count;
    for (uint32_t i = 0; i < count; i++) {
        z.bytes += h_nbytes[i];
        z.chunks += (h_nbytes[i] + kSdChunk - 1) / kSdChunk;
    }
    return z;
}

template <typename C>
static void sd_layout(C &c, const SdShape &z, SdWork *w) {
    const uint64_t nch = z.chunks, npc = nch * kSdPieces;
    auto *ta = c.template take<uint64_t>(z.count + 1);
    auto *tn = c.template take<uint64_t>(z.count + 1);
    auto *tp = c.template take<uint64_t>(z.count + 1);
    auto *tc = c.template take<uint32_t>(z.count + 1);
    auto *X8 = c.template take<uint32_t>(z.bytes ? z.bytes : 1);
    auto *X = c.template take<uint32_t>(z.bytes ? z.bytes : 1);
    auto *en = c.template take<uint32_t>(nch + 1);
    auto *pe = c.template take<uint32_t>(npc + 1);
    auto *ne = c.template take<uint32_t>(npc + 1);
    auto *va = c.template take<uint32_t>(npc + 1);
    auto *st = c.template take<uint32_t>(2);
    auto *eb = c.template take<uint64_t>(npc + 1);
    auto *vb = c.template take<uint64_t>(npc + 1);
    const size_t sb = scan_workspace((size_t)npc + 1);
    void *sw = c.template take<char>(sb);
    if (w) *w = SdWork{ta, tn, tp, tc, X8, X, en, pe, ne, va, st, eb, vb, sw, sb};
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_shard_decode_many_workspace(const uint64_t *h_nbytes, uint32_t nstreams, size_t *bytes) {
    GCHI_CHECK(bytes && (nstreams == 0 || h_nbytes), GCHI_ERR_INVALID, "shard_decode: null argument");
    SizerAdapter a;
    sd_layout(a, sd_shape(h_nbytes, nstreams), nullptr);
    *bytes = a.z.off;
    return GCHI_OK;
}

extern "C" int gchi_shard_decode_workspace(uint64_t nbytes, size_t *bytes) {
    return gchi_shard_decode_many_workspace(&nbytes, 1, bytes);
}

extern "C" int gchi_shard_decode_many(const uint8_t *const *h_adj, const uint64_t *h_nbytes, uint32_t nstreams,
                                      uint32_t *d_src, uint32_t *d_dst, uint64_t max_edges, uint64_t *h_nedges,
                                      uint64_t *h_vend, void *d_work, size_t work_bytes, void *stream) {
    GCHI_CHECK(nstreams == 0 || (h_adj && h_nbytes && h_nedges), GCHI_ERR_INVALID, "shard_decode: null argument");
    GCHI_CHECK((d_src == nullptr) == (d_dst == nullptr), GCHI_ERR_INVALID, "shard_decode: give both outputs or none");
    std::vector<uint64_t> t_adj(nstreams + 1, 0), t_n(nstreams + 1, 0), t_p(nstreams + 1, 0);
    std::vector<uint32_t> t_c(nstreams + 1, 0);
    for (uint32_t i = 0; i < nstreams; i++) {
        GCHI_CHECK(h_nbytes[i] == 0 || h_adj[i], GCHI_ERR_INVALID, "shard_decode: null stream");
        GCHI_CHECK(h_nbytes[i] < 0xffffffffull, GCHI_ERR_UNSUPPORTED, "shard_decode: streams of 4 GiB or more");
        h_nedges[i] = 0;
        if (h_vend) h_vend[i] = 0;
        t_adj[i] = (uint64_t)(uintptr_t)h_adj[i];
        t_n[i] = h_nbytes[i];
        t_p[i + 1] = t_p[i] + h_nbytes[i];
        t_c[i + 1] = t_c[i] + (uint32_t)((h_nbytes[i] + kSdChunk - 1) / kSdChunk);
    }
    const SdShape z = sd_shape(h_nbytes, nstreams);
    if (z.chunks == 0) return GCHI_OK;  // no stream holds a byte
    GCHI_CHECK(z.chunks * kSdPieces < 0xffffffffull, GCHI_ERR_UNSUPPORTED, "shard_decode: too many bytes in one call");
    size_t need = 0;
    gchi_shard_decode_many_workspace(h_nbytes, nstreams, &need);
    GCHI_CHECK(d_work && work_bytes >= need, GCHI_ERR_WORKSPACE, "shard_decode: workspace too small");
    hipStream_t s = as_stream(stream);
    Carver c(d_work, work_bytes);
    SdWork w;
    sd_layout(c, z, &w);
    GCHI_HIP(hipMemcpyAsync(w.t_adj, t_adj.data(), 8 * (size_t)nstreams, hipMemcpyHostToDevice, s));
    GCHI_HIP(hipMemcpyAsync(w.t_n, t_n.data(), 8 * (size_t)nstreams, hipMemcpyHostToDevice, s));
    GCHI_HIP(hipMemcpyAsync(w.t_pbase, t_p.data(), 8 * (size_t)(nstreams + 1), hipMemcpyHostToDevice, s));
    GCHI_HIP(hipMemcpyAsync(w.t_cbase, t_c.data(), 4 * (size_t)(nstreams + 1), hipMemcpyHostToDevice, s));
    const SdStreams T{w.t_adj, w.t_n, w.t_pbase, w.t_cbase, nstreams};
    const uint32_t nch = (uint32_t)z.chunks, npc = nch * kSdPieces;
    launch_fill_u32(w.entry, kSdNone, nch, s);
    launch_fill_u32(w.status, 0u, 2, s);
    hipLaunchKernelGGL(sd_exits, dim3(nch), dim3(kSdThreads), 0, s, T, w.X8, w.X);
    hipLaunchKernelGGL(sd_chain, dim3((nstreams + 63) / 64), dim3(64), 0, s, T, w.X, w.entry, w.status);
    hipLaunchKernelGGL(sd_pieces, dim3((nch + 255) / 256), dim3(256), 0, s, T, nch, w.X8, w.entry, w.pentry);
    hipLaunchKernelGGL(sd_count, dim3(npc), dim3(kSdCopyThreads), 0, s, T, w.pentry, w.nedge, w.vadv, w.status);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(exclusive_scan_u32_to_u64(w.nedge, w.ebase, npc, w.scan, w.scan_bytes, s));
    GCHI_TRY(exclusive_scan_u32_to_u64(w.vadv, w.vbase, npc, w.scan, w.scan_bytes, s));
    uint32_t st[2] = {0, 0};
    std::vector<uint64_t> eb(npc + 1), vb(npc + 1);
    GCHI_HIP(hipMemcpyAsync(st, w.status, sizeof(st), hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipMemcpyAsync(eb.data(), w.ebase, 8 * (size_t)(npc + 1), hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipMemcpyAsync(vb.data(), w.vbase, 8 * (size_t)(npc + 1), hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    GCHI_CHECK(st[0] == 0, GCHI_ERR_FORMAT, "shard_decode: truncated stream (a record runs past its end)");
    GCHI_CHECK(st[1] == 0, GCHI_ERR_OVERFLOW, "shard_decode: a piece's edge count or vertex advance exceeds 32 bits");
    for (uint32_t i = 0; i < nstreams; i++) {
        const uint64_t p0 = (uint64_t)t_c[i] * kSdPieces, p1 = (uint64_t)t_c[i + 1] * kSdPieces;
        h_nedges[i] = eb[p1] - eb[p0];
        const uint64_t vend = vb[p1] - vb[p0];
        GCHI_CHECK(vend <= 0x100000000ull, GCHI_ERR_OVERFLOW, "shard_decode: vertex ids exceed 32 bits");
        if (h_vend) h_vend[i] = vend;
    }
    if (!d_src) return GCHI_OK;
    GCHI_CHECK(eb[npc] <= max_edges, GCHI_ERR_INVALID, "shard_decode: the streams hold more than max_edges edges");
    hipLaunchKernelGGL(sd_emit, dim3(npc), dim3(kSdCopyThreads), 0, s, T, w.pentry, w.ebase, w.vbase, d_src, d_dst);
    GCHI_LAUNCH_CHECK();
    return GCHI_OK;
}

extern "C" int gchi_shard_decode(const uint8_t *d_adj, uint64_t nbytes, uint32_t *d_src, uint32_t *d_dst,
                                 uint64_t max_edges, uint64_t *h_nedges, uint64_t *h_vend, void *d_work,
                                 size_t work_bytes, void *stream) {
    GCHI_CHECK(h_nedges, GCHI_ERR_INVALID, "shard_decode: null h_nedges");
    return gchi_shard_decode_many(&d_adj, &nbytes, 1, d_src, d_dst, max_edges, h_nedges, h_vend, d_work, work_bytes,
                                  stream);
}
