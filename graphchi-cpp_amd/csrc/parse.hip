// Edge-list text parsed on the device: the input path of
// convert_edgelist (src/preprocessing/conversions.hpp:210-288) feeding
// sharder::preprocessing_add_edge (src/preprocessing/sharder.hpp:427-446).
//
// The reference reads the file line by line with fgets (lines of at most 1023
// bytes), drops the trailing newline, skips lines whose first byte is '#' or
// '%', splits the rest with strtok on tab, comma and space, takes from =
// atoi(token 1) and to = atoi(token 2) (a third token is the edge value), stops
// with an assertion when a line has fewer than two tokens, and drops
// self-loops.  Here the text is in HBM and parsed in parallel:
//   * one wavefront per 4 KiB chunk, one lane per 64-byte slice; a line
//     belongs to the slice its first byte lies in (the file start, or the byte
//     after a newline), and that lane parses it to its end (past the slice and
//     the chunk if need be);
//   * a count pass gives every chunk its number of kept edges, a scan gives
//     its output offset, and a write pass parses again and writes the pairs in
//     file order (lanes in order, lines in order within a lane).
// A line with fewer than two tokens, or of 1023 bytes or more (fgets would
// split it), is an error reported with the byte offset where it starts.
#include "common.hpp"
#include "internal.hpp"

namespace gchi {

constexpr uint32_t kParseChunk = 4096;
constexpr uint32_t kParseSlice = kParseChunk / kWave;  // 64 bytes per lane
constexpr uint32_t kParseMaxLine = 1023;               // fgets(s, 1024, f) reads at most 1023 bytes

GCHI_DEV bool is_delim(uint8_t c) { return c == '\t' || c == ',' || c == ' '; }

// atoi on the token at p (its bytes up to the next delimiter or the line end):
// leading white space, an optional sign, then decimal digits up to the first
// other byte; the value wraps to 32 bits as the reference's int-to-vid_t cast.
GCHI_DEV uint32_t parse_atoi(const uint8_t *t, uint64_t p, uint64_t e) {
    while (p < e && (t[p] == '\r' || t[p] == '\v' || t[p] == '\f')) p++;
    bool neg = false;
    if (p < e && (t[p] == '+' || t[p] == '-')) {
        neg = t[p] == '-';
        p++;
    }
    uint32_t v = 0;
    while (p < e && t[p] >= '0' && t[p] <= '9') {
        v = v * 10u + (uint32_t)(t[p] - '0');
        p++;
    }
    return neg ? 0u - v : v;
}

// Parses the line starting at s (its end: the next newline or the file end).
// Returns 1 and the pair for an edge line, 0 for a comment, -1 for a line
// with fewer than two tokens or one too long.  *next = the byte after its
// newline.
GCHI_DEV int parse_line(const uint8_t *t, uint64_t n, uint64_t s, uint32_t &from, uint32_t &to, uint64_t &next) {
    uint64_t e = s;
    while (e < n && t[e] != '\n') e++;
    next = e + 1;
    if (e - s >= kParseMaxLine) return -1;
    if (e > s && (t[s] == '#' || t[s] == '%')) return 0;
    uint64_t p = s;
    while (p < e && is_delim(t[p])) p++;
    if (p == e) return -1;
    const uint64_t a = p;
    while (p < e && !is_delim(t[p])) p++;
    const uint64_t a_end = p;
    while (p < e && is_delim(t[p])) p++;
    if (p == e) return -1;
    const uint64_t b = p;
    while (p < e && !is_delim(t[p])) p++;
    from = parse_atoi(t, a, a_end);
    to = parse_atoi(t, b, p);
    return 1;
}

// First byte of the line holding position i's slice start: lines start at 0 and
// after each newline.
GCHI_DEV bool line_starts_at(const uint8_t *t, uint64_t i) { return i == 0 || t[i - 1] == '\n'; }

// WRITE = 0: count kept edges per chunk, max id, errors.  WRITE = 1: write the
// pairs at the chunk's scanned offset.
template <int WRITE>
__global__ void parse_pass(const uint8_t *__restrict__ t, uint64_t n, uint64_t nchunks, uint32_t *chunk_cnt,
                           const uint64_t *__restrict__ chunk_off, uint32_t *src, uint32_t *dst, uint64_t max_edges,
                           unsigned long long *bad_pos) {
    const int lane = threadIdx.x & (kWave - 1);
    const uint64_t wid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
    const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) / kWave;
    for (uint64_t c = wid; c < nchunks; c += nw) {
        const uint64_t lo = c * kParseChunk + (uint64_t)lane * kParseSlice;
        const uint64_t hi = lo + kParseSlice < n ? lo + kParseSlice : n;
        uint32_t kept = 0;
        // pass over this lane's line starts: count first (both passes), so the
        // write pass knows each lane's offset inside the chunk
        for (uint64_t i = lo; i < hi;) {
            if (!line_starts_at(t, i)) {
                i++;
                continue;
            }
            uint32_t f = 0, to = 0;
            uint64_t next = 0;
            const int r = parse_line(t, n, i, f, to, next);
            if (r < 0 && !WRITE) atomicMin(bad_pos, (unsigned long long)i);
            if (r > 0 && f != to) kept++;
            i = next;
        }
        if (!WRITE) {
            const uint32_t tot = wave_sum(kept);
            if (lane == 0) chunk_cnt[c] = tot;
            continue;
        }
        // lane offset = kept edges of the lanes before it
        uint32_t pre = 0;
        {
            uint32_t v = kept;
#pragma unroll
            for (int o = 1; o < kWave; o <<= 1) {
                const uint32_t u = __shfl_up(v, o, kWave);
                if (lane >= o) v += u;
            }
            pre = v - kept;
        }
        uint64_t out = chunk_off[c] + pre;
        for (uint64_t i = lo; i < hi;) {
            if (!line_starts_at(t, i)) {
                i++;
                continue;
            }
            uint32_t f = 0, to = 0;
            uint64_t next = 0;
            const int r = parse_line(t, n, i, f, to, next);
            if (r > 0 && f != to) {
                if (out < max_edges) {
                    src[out] = f;
                    dst[out] = to;
                }
                out++;
            }
            i = next;
        }
    }
}

// max vertex id over kept edges, from the written pairs
__global__ void parse_maxid(const uint32_t *src, const uint32_t *dst, uint64_t e, uint32_t *maxid) {
    uint32_t m = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e; i += stride)
        m = max(m, max(src[i], dst[i]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, kWave));
    if ((threadIdx.x & (kWave - 1)) == 0 && m) atomicMax(maxid, m);
}

template <typename C>
static void parse_layout(C &c, uint64_t nbytes, uint32_t **cnt, uint64_t **off, uint32_t **maxid,
                         unsigned long long **bad, void **sw, size_t *swb) {
    const uint64_t nchunks = (nbytes + kParseChunk - 1) / kParseChunk;
    *cnt = c.template take<uint32_t>(nchunks + 1);
    *off = c.template take<uint64_t>(nchunks + 1);
    *maxid = c.template take<uint32_t>(1);
    *bad = c.template take<unsigned long long>(1);
    *swb = scan_workspace(nchunks + 1);
    *sw = c.template take<char>(*swb);
}

}  // namespace gchi

using namespace gchi;

extern "C" int gchi_parse_edgelist_workspace(uint64_t nbytes, size_t *bytes) {
    GCHI_CHECK(bytes, GCHI_ERR_INVALID, "null bytes");
    SizerAdapter a;
    uint32_t *cnt, *mx;
    uint64_t *off;
    unsigned long long *bad;
    void *sw;
    size_t swb;
    parse_layout(a, nbytes, &cnt, &off, &mx, &bad, &sw, &swb);
    *bytes = a.z.off;
    return GCHI_OK;
}

extern "C" int gchi_parse_edgelist(const uint8_t *d_text, uint64_t nbytes, uint32_t *d_src, uint32_t *d_dst,
                                   uint64_t max_edges, uint64_t *h_nedges, uint32_t *h_maxid, void *d_work,
                                   size_t work_bytes, void *stream) {
    GCHI_CHECK(h_nedges && h_maxid && (nbytes == 0 || d_text), GCHI_ERR_INVALID, "parse_edgelist: null argument");
    GCHI_CHECK(max_edges == 0 || (d_src && d_dst), GCHI_ERR_INVALID, "parse_edgelist: null output");
    Carver c(d_work, work_bytes);
    uint32_t *cnt, *mx;
    uint64_t *off;
    unsigned long long *bad;
    void *sw;
    size_t swb;
    parse_layout(c, nbytes, &cnt, &off, &mx, &bad, &sw, &swb);
    GCHI_CHECK(d_work && c.ok(), GCHI_ERR_WORKSPACE, "parse_edgelist: workspace too small");
    *h_nedges = 0;
    *h_maxid = 0;
    if (nbytes == 0) return GCHI_OK;
    hipStream_t s = as_stream(stream);
    const uint64_t nchunks = (nbytes + kParseChunk - 1) / kParseChunk;
    launch_fill_u32(mx, 0u, 1, s);
    launch_fill_u64(reinterpret_cast<uint64_t *>(bad), ~0ull, 1, s);
    const unsigned grid = grid_for(nchunks * kWave, 256, 1u << 15);
    hipLaunchKernelGGL(parse_pass<0>, dim3(grid), dim3(256), 0, s, d_text, nbytes, nchunks, cnt, off, d_src, d_dst,
                       max_edges, bad);
    GCHI_LAUNCH_CHECK();
    GCHI_TRY(exclusive_scan_u32_to_u64(cnt, off, nchunks, sw, swb, s));
    unsigned long long h_bad = 0;
    uint64_t total = 0;
    GCHI_HIP(hipMemcpyAsync(&h_bad, bad, 8, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipMemcpyAsync(&total, off + nchunks, 8, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    if (h_bad != ~0ull)
        return fail(GCHI_ERR_FORMAT, "parse_edgelist: the line at byte " + std::to_string(h_bad) +
                                         " has fewer than two fields or is 1023 bytes or longer");
    *h_nedges = total;
    GCHI_CHECK(total <= max_edges || max_edges == 0, GCHI_ERR_INVALID, "parse_edgelist: more edges than max_edges");
    if (max_edges == 0 || total == 0) return GCHI_OK;
    hipLaunchKernelGGL(parse_pass<1>, dim3(grid), dim3(256), 0, s, d_text, nbytes, nchunks, cnt, off, d_src, d_dst,
                       max_edges, bad);
    hipLaunchKernelGGL(parse_maxid, dim3(grid_for(total, 256, 4096)), dim3(256), 0, s, d_src, d_dst, total, mx);
    GCHI_LAUNCH_CHECK();
    GCHI_HIP(hipMemcpyAsync(h_maxid, mx, 4, hipMemcpyDeviceToHost, s));
    GCHI_HIP(hipStreamSynchronize(s));
    return GCHI_OK;
}
