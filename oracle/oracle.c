/*
 * oracle.c -- CPU restatement of the GraphChi hot path.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * always as what results are compared against (or the timed CPU baseline),
 * never as the product.
 *
 * Pinning (what checks each function, beyond agreeing with the GPU):
 *   orc_csc_build          - closed forms on hand graphs; offsets == in-degree scan
 *   orc_pagerank_sync_f32  - golden .vout of the reference pagerank_functional
 *                            (sync mode) and of the 0.15+0.85*s harness kernel,
 *                            built from the reference sources (oracle/Makefile);
 *                            closed forms (no in-edges -> teleport; chain/star)
 *   orc_pagerank_semisync_f32 - golden .vout of the reference pagerank_functional
 *                            in its default (semisync) mode, one and several
 *                            sub-intervals per interval; one block == sync
 *   orc_semisync_blocks    - the same goldens (a wrong cut changes the ranks);
 *                            hand cases of the window rule
 *   orc_pagerank_sync_f64  - equals the f32 variant to FP32 rounding; linear-
 *                            system residual of its fixed point
 *   orc_ppr_f64            - fixed-point residual ||r - (t + aP'r)||_1; sum rule
 *                            on dangling-free graphs (sum r = 1)
 *   orc_cc_*               - golden labels of the reference connectedcomponents;
 *                            the two restatements agree; union-find by scipy
 *   orc_stable_sort_u64    - equals numpy stable argsort on random keys
 *   orc_shard_*            - golden .intervals / .adj / .adjidx / _degs.bin of the
 *                            reference sharder (several nshards, both edge sizes)
 *   orc_shard_decode       - the golden .adj bytes of the reference sharder decode
 *                            to the golden cases' edges; encode -> decode round
 *                            trips (gaps, 0xff counts); truncations are refused
 *   orc_degree_order       - golden .vertexmap of the reference order_by_degree
 *   orc_triangles          - golden totals/per-vertex .vout of the reference
 *                            trianglecounting; brute force on tiny graphs; K_n
 *   orc_pr_psw_omp         - equals orc_pagerank_sync_f32 within FP32 tolerance
 *   orc_psw_pagerank       - the reference engine's ranks, iteration count and
 *                            change log for the same program (golden fixtures)
 *   orc_psw_run            - the reference engine's own vertex values for the
 *                            probe program (oracle/Makefile builds it against the
 *                            reference headers; golden fixtures, 1-3 shards);
 *                            zero violations of basic_smoketest.cpp:65's
 *                            invariant, and violations when run in reverse
 *   orc_psw_sched          - the reference engine's vertex values, deltalog
 *                            and printed task counts for the same program
 *                            (golden fixtures, 1-4 shards, up to 26 iterations)
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ helpers */
static int cmp_u32(const void *a, const void *b) {
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return x < y ? -1 : (x > y);
}

int orc_csc_build(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, uint32_t *in_off,
                  uint32_t *in_src, uint32_t *outdeg, uint32_t *indeg) {
    /* counting sort by dst, then each row sorted by src (rows small: qsort) */
    uint64_t *cnt = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
    if (!cnt) return -1;
    if (outdeg) memset(outdeg, 0, (size_t)n * 4);
    for (uint64_t i = 0; i < e; i++) {
        if (src[i] >= n || dst[i] >= n) {
            free(cnt);
            return -1;
        }
        cnt[dst[i] + 1]++;
        if (outdeg) outdeg[src[i]]++;
    }
    for (uint32_t v = 0; v < n; v++) cnt[v + 1] += cnt[v];
    for (uint32_t v = 0; v <= n; v++) in_off[v] = (uint32_t)cnt[v];
    if (indeg)
        for (uint32_t v = 0; v < n; v++) indeg[v] = (uint32_t)(cnt[v + 1] - cnt[v]);
    for (uint64_t i = 0; i < e; i++) in_src[cnt[dst[i]]++] = src[i];
    for (uint32_t v = 0; v < n; v++) {
        uint32_t b = in_off[v], en = in_off[v + 1];
        if (en - b > 1) qsort(in_src + b, en - b, 4, cmp_u32);
    }
    free(cnt);
    return 0;
}

/* ------------------------------------------------------------------ PageRank */
static float pr_apply_f32(double teleport, double damping, float s) {
    /* example_apps/pagerank_functional.cpp:74 -- double arithmetic, float result */
    volatile double t = damping * (double)s;
    double r = teleport + t;
    return (float)r;
}

int orc_pagerank_sync_f32(uint32_t n, const uint32_t *in_off, const uint32_t *in_src, const uint32_t *outdeg,
                          double teleport, double damping, float init, int nsweeps, float *r_out, double *stats) {
    float *x = (float *)malloc((size_t)n * 4), *xn = (float *)malloc((size_t)n * 4);
    float *r = (float *)malloc((size_t)n * 4);
    if (!x || !xn || !r) return -1;
    for (uint32_t v = 0; v < n; v++) {
        r[v] = init;
        x[v] = outdeg[v] ? init / (float)outdeg[v] : 0.0f; /* value_to_neighbor, :78-81 */
    }
    for (int it = 0; it < nsweeps; it++) {
        double res = 0, sum = 0;
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : res, sum)
        for (uint32_t v = 0; v < n; v++) {
            float s = 0.0f; /* reset() */
            for (uint32_t k = in_off[v]; k < in_off[v + 1]; k++) s = s + x[in_src[k]]; /* plus */
            float rn = pr_apply_f32(teleport, damping, s);
            res += fabs((double)rn - (double)r[v]);
            sum += (double)rn;
            r[v] = rn;
            xn[v] = outdeg[v] ? rn / (float)outdeg[v] : 0.0f;
        }
        float *t = x;
        x = xn;
        xn = t;
        if (stats) {
            stats[2 * it] = res;
            stats[2 * it + 1] = sum;
        }
    }
    memcpy(r_out, r, (size_t)n * 4);
    free(x);
    free(xn);
    free(r);
    return 0;
}

/* Semi-synchronous form: run_functional_unweighted_semisynchronous
 * (src/api/functional/functional_api.hpp:119-137) with the vertex program of
 * src/api/functional/functional_semisync.hpp:76-110 under the interval loop of
 * src/engine/graphchi_engine.hpp:825-921 and the phase order of
 * src/engine/functional/functional_engine.hpp:50-121.  Every edge holds one
 * value, overwritten when its source is updated; blocks (the engine's
 * sub-intervals) run in order, each gathering all its in-edges before any of
 * its vertices is updated and scattered.  So block j sees this sweep's x of
 * sources in earlier blocks and the previous sweep's x of the others.
 * bstart[0..nblocks] are the block boundaries (bstart[nblocks] = n). */
int orc_pagerank_semisync_f32(uint32_t n, const uint32_t *in_off, const uint32_t *in_src, const uint32_t *outdeg,
                              uint32_t nblocks, const uint32_t *bstart, double teleport, double damping, float init,
                              int nsweeps, float *r_out, double *stats) {
    if (nblocks == 0 || bstart[0] != 0 || bstart[nblocks] != n) return -1;
    float *x = (float *)malloc((size_t)n * 4), *r = (float *)malloc((size_t)n * 4);
    float *rn = (float *)malloc((size_t)n * 4);
    if (!x || !r || !rn) return -1;
    for (uint32_t v = 0; v < n; v++) {
        r[v] = init; /* iteration 0: first_iteration() then the scatter, :86-88, :104-106 */
        x[v] = outdeg[v] ? init / (float)outdeg[v] : 0.0f;
    }
    for (int it = 0; it < nsweeps; it++) {
        double res = 0, sum = 0;
        for (uint32_t j = 0; j < nblocks; j++) {
            const uint32_t a = bstart[j], b = bstart[j + 1];
            /* gather + apply of the whole block, then its scatter */
#pragma omp parallel for schedule(dynamic, 4096)
            for (uint32_t v = a; v < b; v++) {
                float s = 0.0f;
                for (uint32_t k = in_off[v]; k < in_off[v + 1]; k++) s = s + x[in_src[k]];
                rn[v] = pr_apply_f32(teleport, damping, s);
            }
            for (uint32_t v = a; v < b; v++) {
                res += fabs((double)rn[v] - (double)r[v]);
                sum += (double)rn[v];
                r[v] = rn[v];
                x[v] = outdeg[v] ? rn[v] / (float)outdeg[v] : 0.0f;
            }
        }
        if (stats) {
            stats[2 * it] = res;
            stats[2 * it + 1] = sum;
        }
    }
    memcpy(r_out, r, (size_t)n * 4);
    free(x);
    free(r);
    free(rn);
    return 0;
}

/* The engine's sub-intervals for a functional kernel: each interval
 * [intervals[p-1]+1, intervals[p]] (the first from 0) is cut into windows of
 * at most maxwindow + 1 vertices (src/engine/graphchi_engine.hpp:862-869 with
 * determine_next_window returning its upper bound for computational edges,
 * :311-313, and maxwindow = membudget_mb * 2^20 / 3 / 100, :729-732).  Writes
 * the block starts and n to bstart (capacity cap) and returns the block count,
 * or -1. */
int orc_semisync_blocks(uint32_t n, uint32_t nintervals, const uint32_t *interval_ends, int membudget_mb,
                        uint32_t *bstart, uint32_t cap) {
    const uint64_t maxwindow = (uint64_t)membudget_mb * 1024 * 1024 / 3 / 100;
    uint32_t nb = 0;
    uint64_t st = 0;
    for (uint32_t p = 0; p < nintervals; p++) {
        const uint64_t en = interval_ends[p];
        if (st > en) continue; /* empty interval (:841) */
        uint64_t sub = st;
        while (sub <= en) {
            uint64_t sub_en = sub + maxwindow < en ? sub + maxwindow : en;
            if (nb >= cap) return -1;
            bstart[nb++] = (uint32_t)sub;
            sub = sub_en + 1;
        }
        st = en + 1;
    }
    if (nb >= cap) return -1;
    bstart[nb] = n;
    return (int)nb;
}

int orc_pagerank_sync_f64(uint32_t n, const uint32_t *in_off, const uint32_t *in_src, const uint32_t *outdeg,
                          double teleport, double damping, double init, int nsweeps, double *r_out) {
    double *x = (double *)malloc((size_t)n * 8), *xn = (double *)malloc((size_t)n * 8);
    if (!x || !xn) return -1;
    for (uint32_t v = 0; v < n; v++) {
        r_out[v] = init;
        x[v] = outdeg[v] ? init / outdeg[v] : 0.0;
    }
    for (int it = 0; it < nsweeps; it++) {
#pragma omp parallel for schedule(dynamic, 4096)
        for (uint32_t v = 0; v < n; v++) {
            double s = 0.0;
            for (uint32_t k = in_off[v]; k < in_off[v + 1]; k++) s += x[in_src[k]];
            double rn = teleport + damping * s;
            r_out[v] = rn;
            xn[v] = outdeg[v] ? rn / outdeg[v] : 0.0;
        }
        double *t = x;
        x = xn;
        xn = t;
    }
    free(x);
    free(xn);
    return 0;
}

int orc_ppr_f64(uint32_t n, const uint32_t *in_off, const uint32_t *in_src, const uint32_t *outdeg,
                const uint32_t *seeds, uint32_t nseeds, double damping, double tol, int max_sweeps, double *r_out) {
    double *x = (double *)malloc((size_t)n * 8), *t = (double *)calloc(n, 8);
    if (!x || !t || nseeds == 0) return -1;
    for (uint32_t i = 0; i < nseeds; i++) t[seeds[i]] += (1.0 - damping) / (double)nseeds;
    for (uint32_t v = 0; v < n; v++) {
        r_out[v] = t[v];
        x[v] = outdeg[v] ? t[v] / outdeg[v] : 0.0;
    }
    int it = 0;
    for (; it < max_sweeps; it++) {
        double res = 0;
        double *rn = (double *)malloc((size_t)n * 8);
#pragma omp parallel for schedule(dynamic, 4096) reduction(+ : res)
        for (uint32_t v = 0; v < n; v++) {
            double s = 0.0;
            for (uint32_t k = in_off[v]; k < in_off[v + 1]; k++) s += x[in_src[k]];
            rn[v] = t[v] + damping * s;
            res += fabs(rn[v] - r_out[v]);
        }
        for (uint32_t v = 0; v < n; v++) {
            r_out[v] = rn[v];
            x[v] = outdeg[v] ? rn[v] / outdeg[v] : 0.0;
        }
        free(rn);
        if (res <= tol) {
            it++;
            break;
        }
    }
    free(x);
    free(t);
    return it;
}

/* ------------------------------------------------------------------ CC */
static uint32_t uf_find(uint32_t *p, uint32_t x) {
    while (p[x] != x) {
        p[x] = p[p[x]];
        x = p[x];
    }
    return x;
}

int orc_cc_unionfind(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, uint32_t *labels) {
    for (uint32_t v = 0; v < n; v++) labels[v] = v;
    for (uint64_t i = 0; i < e; i++) {
        if (src[i] >= n || dst[i] >= n) return -1;
        uint32_t a = uf_find(labels, src[i]), b = uf_find(labels, dst[i]);
        if (a == b) continue;
        if (a < b)
            labels[b] = a; /* the root is always the smallest id of its set */
        else
            labels[a] = b;
    }
    for (uint32_t v = 0; v < n; v++) labels[v] = uf_find(labels, v);
    return 0;
}

int orc_cc_labelprop(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, uint32_t *labels) {
    /* iteration 0: label = min(id, neighbour ids); then min over neighbour
     * labels until nothing changes (connectedcomponents.cpp:83-120). */
    uint32_t *nl = (uint32_t *)malloc((size_t)n * 4);
    if (!nl) return -1;
    for (uint32_t v = 0; v < n; v++) labels[v] = v;
    int sweeps = 0;
    for (;;) {
        memcpy(nl, labels, (size_t)n * 4);
        for (uint64_t i = 0; i < e; i++) {
            uint32_t a = src[i], b = dst[i];
            if (labels[b] < nl[a]) nl[a] = labels[b];
            if (labels[a] < nl[b]) nl[b] = labels[a];
        }
        sweeps++;
        int changed = memcmp(nl, labels, (size_t)n * 4) != 0;
        memcpy(labels, nl, (size_t)n * 4);
        if (!changed) break;
    }
    free(nl);
    return sweeps;
}

/* ------------------------------------------------------------------ sort */
int orc_stable_sort_u64(uint64_t n, uint64_t *keys, uint32_t *vals, int key_bits) {
    if (n <= 1 || key_bits <= 0) return 0;
    uint64_t *kb = (uint64_t *)malloc(n * 8);
    uint32_t *vb = vals ? (uint32_t *)malloc(n * 4) : NULL;
    if (!kb || (vals && !vb)) return -1;
    uint64_t *ka = keys;
    uint32_t *va = vals;
    for (int sh = 0; sh < key_bits; sh += 8) {
        int bits = key_bits - sh < 8 ? key_bits - sh : 8;
        uint64_t mask = (1ull << bits) - 1, cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (uint64_t i = 0; i < n; i++) cnt[((ka[i] >> sh) & mask) + 1]++;
        for (int d = 0; d < 256; d++) cnt[d + 1] += cnt[d];
        for (uint64_t i = 0; i < n; i++) {
            uint64_t d = (ka[i] >> sh) & mask;
            uint64_t p = cnt[d]++;
            kb[p] = ka[i];
            if (vals) vb[p] = va[i];
        }
        uint64_t *tk = ka;
        ka = kb;
        kb = tk;
        uint32_t *tv = va;
        va = vb;
        vb = tv;
    }
    if (ka != keys) {
        memcpy(keys, ka, n * 8);
        if (vals) memcpy(vals, va, n * 4);
        free(ka);
        if (vals) free(va);
    } else {
        free(kb);
        if (vals) free(vb);
    }
    return 0;
}

/* ------------------------------------------------------------------ sharder */
int orc_auto_nshards(uint64_t nedges, size_t edge_value_bytes, int membudget_mb) {
    /* sharder.hpp:572-575 */
    double max_shardsize = membudget_mb * 1024. * 1024. / 8;
    return (int)(1 + (nedges * edge_value_bytes / max_shardsize) + 0.5);
}

int orc_shard_intervals(uint64_t e, const uint32_t *dst_sorted, uint32_t max_id, int nshards, uint32_t *ends,
                        uint64_t *counts) {
    /* sharder.hpp:861-927, write_shards :961-968 */
    uint64_t eps = e / nshards + 1, cur = 0, sharded = 0;
    int shardnum = 0;
    uint32_t prevvid = 0xffffffffu;
    for (uint64_t i = 0; i <= e; i++) {
        int is_done = (i == e);
        if (is_done || (cur >= eps && dst_sorted[i] != prevvid)) {
            if (shardnum >= nshards) return -1; /* the reference asserts here */
            ends[shardnum] = (shardnum == nshards - 1) ? max_id : prevvid;
            counts[shardnum] = cur;
            shardnum++;
            cur = 0;
            if (shardnum < nshards) eps = (e - sharded) / (uint64_t)(nshards - shardnum);
            if (is_done) break;
        }
        cur++;
        prevvid = dst_sorted[i];
        sharded++;
    }
    return shardnum;
}

int64_t orc_shard_encode(uint64_t e, const uint32_t *src, const uint32_t *dst, uint8_t *adj, uint64_t *idx_triples,
                         uint64_t *nidx) {
    /* sharder.hpp:701-822 (non-dynamic edge data) */
    uint64_t pos = 0, istart = 0, last_index_output = 0, ni = 0;
    uint32_t curvid = 0;
#define PUT8(b)                                  \
    do {                                         \
        if (adj) adj[pos] = (uint8_t)(b);        \
        pos++;                                   \
    } while (0)
#define PUT32(w)                                 \
    do {                                         \
        uint32_t _w = (w);                       \
        if (adj) memcpy(adj + pos, &_w, 4);      \
        pos += 4;                                \
    } while (0)
    for (uint64_t i = 0; i <= e; i++) {
        int stopper = (i == e);
        uint32_t es = stopper ? 0 : src[i];
        if (es != curvid || stopper) {
            uint64_t count = i - istart;
            if (istart - last_index_output >= 1024 * 1024) {
                if (idx_triples) {
                    idx_triples[3 * ni] = curvid;
                    idx_triples[3 * ni + 1] = pos;
                    idx_triples[3 * ni + 2] = istart;
                }
                ni++;
                last_index_output = istart;
            }
            if (count > 0) {
                if (count < 255) {
                    PUT8(count);
                } else {
                    PUT8(0xff);
                    PUT32((uint32_t)count);
                }
            }
            for (uint64_t j = istart; j < i; j++) PUT32(dst[j]);
            istart = i;
            if (!stopper) {
                if (es - curvid > 1 || (i == 0 && es > 0)) {
                    int64_t nz = (int64_t)es - (int64_t)curvid - 1;
                    if (i == 0 && es > 0) nz = es;
                    do {
                        PUT8(0);
                        nz--;
                        int64_t tnz = nz < 254 ? nz : 254;
                        PUT8(tnz);
                        nz -= tnz;
                    } while (nz > 0);
                }
            }
            curvid = es;
        }
    }
#undef PUT8
#undef PUT32
    if (nidx) *nidx = ni;
    return (int64_t)pos;
}

/* ------------------------------------------------------------------ TC */
static const uint32_t *g_deg;
static int cmp_deg_id(const void *a, const void *b) {
    /* vertex_degree_less, orderbydegree.hpp:23-25 */
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    if (g_deg[x] != g_deg[y]) return g_deg[x] < g_deg[y] ? -1 : 1;
    return x < y ? -1 : (x > y);
}

int orc_degree_order(uint32_t n, const uint32_t *deg, uint32_t *vertexmap) {
    uint32_t *ids = (uint32_t *)malloc((size_t)n * 4);
    if (!ids) return -1;
    for (uint32_t v = 0; v < n; v++) ids[v] = v;
    g_deg = deg;
    qsort(ids, n, 4, cmp_deg_id);
    for (uint32_t i = 0; i < n; i++) vertexmap[ids[i]] = i;
    free(ids);
    return 0;
}

int orc_triangles(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, uint32_t *vertexmap,
                  uint32_t *per_vertex, uint64_t *total) {
    uint32_t *deg = (uint32_t *)calloc(n, 4);
    if (!deg) return -1;
    for (uint64_t i = 0; i < e; i++) {
        if (src[i] == dst[i]) continue;
        deg[src[i]]++;
        deg[dst[i]]++;
    }
    orc_degree_order(n, deg, vertexmap);
    /* oriented adjacency (low rank -> high rank), sorted and deduplicated */
    uint32_t *a = (uint32_t *)malloc((e ? e : 1) * 4), *b = (uint32_t *)malloc((e ? e : 1) * 4);
    uint64_t m = 0;
    for (uint64_t i = 0; i < e; i++) {
        if (src[i] == dst[i]) continue;
        uint32_t x = vertexmap[src[i]], y = vertexmap[dst[i]];
        a[m] = x < y ? x : y;
        b[m] = x < y ? y : x;
        m++;
    }
    uint32_t *off = (uint32_t *)calloc((size_t)n + 1, 4), *adj = (uint32_t *)malloc((m ? m : 1) * 4);
    for (uint64_t i = 0; i < m; i++) off[a[i] + 1]++;
    for (uint32_t v = 0; v < n; v++) off[v + 1] += off[v];
    uint32_t *fill = (uint32_t *)malloc((size_t)n * 4);
    memcpy(fill, off, (size_t)n * 4);
    for (uint64_t i = 0; i < m; i++) adj[fill[a[i]]++] = b[i];
    /* sort + dedup each list, compacting in place */
    uint32_t *noff = (uint32_t *)calloc((size_t)n + 1, 4);
    uint64_t w = 0;
    for (uint32_t v = 0; v < n; v++) {
        uint32_t s = off[v], en = off[v + 1];
        qsort(adj + s, en - s, 4, cmp_u32);
        noff[v] = (uint32_t)w;
        for (uint32_t k = s; k < en; k++)
            if (k == s || adj[k] != adj[k - 1]) adj[w++] = adj[k];
    }
    noff[n] = (uint32_t)w;
    uint64_t *cnt = (uint64_t *)calloc(n, 8);
    uint64_t tot = 0;
    for (uint32_t u = 0; u < n; u++) {
        for (uint32_t k = noff[u]; k < noff[u + 1]; k++) {
            uint32_t v = adj[k];
            uint32_t i = noff[u], j = noff[v];
            uint64_t c = 0;
            while (i < noff[u + 1] && j < noff[v + 1]) {
                if (adj[i] < adj[j])
                    i++;
                else if (adj[i] > adj[j])
                    j++;
                else {
                    cnt[adj[i]]++;
                    c++;
                    i++;
                    j++;
                }
            }
            cnt[u] += c;
            cnt[v] += c;
            tot += c;
        }
    }
    int overflow = 0;
    for (uint32_t v = 0; v < n; v++) {
        if (cnt[v] > 0xffffffffull) overflow = 1;
        per_vertex[v] = (uint32_t)cnt[v];
    }
    *total = tot;
    free(deg);
    free(a);
    free(b);
    free(off);
    free(adj);
    free(fill);
    free(noff);
    free(cnt);
    return overflow ? 1 : 0;
}

/* ------------------------------------------------------------------ CPU baseline */
static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : (x > y);
}

double orc_pr_psw_omp(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, int nshards,
                      double teleport, double damping, float init, int nsweeps, int nthreads, float *r_out) {
    if (nthreads > 0) omp_set_num_threads(nthreads);
    /* preprocessing (untimed): shards = dst intervals, each sorted by (src,dst) */
    uint64_t *keys = (uint64_t *)malloc((e ? e : 1) * 8);
    for (uint64_t i = 0; i < e; i++) keys[i] = ((uint64_t)dst[i] << 32) | src[i];
    qsort(keys, e, 8, cmp_u64);
    uint32_t *dsorted = (uint32_t *)malloc((e ? e : 1) * 4);
    for (uint64_t i = 0; i < e; i++) dsorted[i] = (uint32_t)(keys[i] >> 32);
    uint32_t *ends = (uint32_t *)malloc(nshards * 4);
    uint64_t *counts = (uint64_t *)malloc(nshards * 8);
    int ns = orc_shard_intervals(e, dsorted, n - 1, nshards, ends, counts);
    if (ns <= 0) return -1.0;
    /* shard-major arrays, (src,dst) inside each shard */
    uint32_t *s_src = (uint32_t *)malloc((e ? e : 1) * 4), *s_dst = (uint32_t *)malloc((e ? e : 1) * 4);
    uint64_t *sh_off = (uint64_t *)calloc(ns + 1, 8);
    for (int p = 0; p < ns; p++) sh_off[p + 1] = sh_off[p] + counts[p];
    for (int p = 0; p < ns; p++) {
        for (uint64_t i = sh_off[p]; i < sh_off[p + 1]; i++)
            keys[i] = ((uint64_t)(uint32_t)keys[i] << 32) | (keys[i] >> 32); /* (src,dst) */
        qsort(keys + sh_off[p], counts[p], 8, cmp_u64);
        for (uint64_t i = sh_off[p]; i < sh_off[p + 1]; i++) {
            s_src[i] = (uint32_t)(keys[i] >> 32);
            s_dst[i] = (uint32_t)keys[i];
        }
    }
    /* out-edge slots of every vertex across all shards (sliding-shard walk) */
    uint32_t *outdeg = (uint32_t *)calloc(n, 4);
    for (uint64_t i = 0; i < e; i++) outdeg[s_src[i]]++;
    uint64_t *out_off = (uint64_t *)calloc((size_t)n + 1, 8);
    for (uint32_t v = 0; v < n; v++) out_off[v + 1] = out_off[v] + outdeg[v];
    uint64_t *fill = (uint64_t *)malloc((size_t)n * 8);
    memcpy(fill, out_off, (size_t)n * 8);
    uint64_t *out_pos = (uint64_t *)malloc((e ? e : 1) * 8);
    for (uint64_t i = 0; i < e; i++) out_pos[fill[s_src[i]]++] = i;
    float *val = (float *)calloc(2 * (e ? e : 1), 4); /* PairContainer<float> per edge */
    float *r = (float *)malloc((size_t)n * 4), *cum = (float *)malloc((size_t)n * 4);
    uint32_t *istart = (uint32_t *)malloc((ns + 1) * 4);
    istart[0] = 0;
    for (int p = 0; p < ns; p++) istart[p + 1] = ends[p] + 1;
    if (istart[ns] > n) istart[ns] = n;
    /* iteration 0: initial value + scatter into right slot */
#pragma omp parallel for schedule(dynamic, 1024)
    for (uint32_t v = 0; v < n; v++) {
        r[v] = init;
        float x = outdeg[v] ? init / (float)outdeg[v] : 0.0f;
        for (uint64_t k = out_off[v]; k < out_off[v + 1]; k++) val[2 * out_pos[k] + 1] = x;
    }
    double t0 = omp_get_wtime();
    for (int it = 1; it <= nsweeps; it++) {
        int rd = (it % 2 == 0) ? 0 : 1, wr = 1 - rd; /* PairContainer oldval/set_newval */
        for (int p = 0; p < ns; p++) {
            uint32_t a = istart[p], b = istart[p + 1];
#pragma omp parallel for schedule(static)
            for (uint32_t v = a; v < b; v++) cum[v] = 0.0f;
#pragma omp parallel for schedule(dynamic, 65536)
            for (uint64_t i = sh_off[p]; i < sh_off[p + 1]; i++) {
                float xv = val[2 * i + rd];
#pragma omp atomic
                cum[s_dst[i]] += xv;
            }
#pragma omp parallel for schedule(dynamic, 1024)
            for (uint32_t v = a; v < b; v++) {
                float rn = pr_apply_f32(teleport, damping, cum[v]);
                r[v] = rn;
                float x = outdeg[v] ? rn / (float)outdeg[v] : 0.0f;
                for (uint64_t k = out_off[v]; k < out_off[v + 1]; k++) val[2 * out_pos[k] + wr] = x;
            }
        }
    }
    double secs = omp_get_wtime() - t0;
    if (r_out) memcpy(r_out, r, (size_t)n * 4);
    free(keys);
    free(dsorted);
    free(ends);
    free(counts);
    free(s_src);
    free(s_dst);
    free(sh_off);
    free(outdeg);
    free(out_off);
    free(fill);
    free(out_pos);
    free(val);
    free(r);
    free(cum);
    free(istart);
    return secs;
}

int64_t orc_shard_decode(uint64_t nbytes, const uint8_t *adj, uint32_t *src, uint32_t *dst, uint64_t max_edges) {
    /* memory_shard::load_vertices' parse (src/shards/memoryshard.hpp:448-497):
     * a 0 byte and the byte nz after it skip 1 + nz vertices; a count byte
     * (0xff: a uint32 count follows) and that many uint32 dst ids belong to the
     * current vertex, which then advances by one. */
    uint64_t p = 0, ne = 0, vid = 0;
    while (p < nbytes) {
        uint8_t ns = adj[p++];
        if (ns == 0) {
            if (p >= nbytes) return -1;
            vid += 1 + (uint64_t)adj[p++];
            continue;
        }
        uint64_t cnt = ns;
        if (ns == 0xff) {
            uint32_t c32;
            if (p + 4 > nbytes) return -1;
            memcpy(&c32, adj + p, 4);
            p += 4;
            cnt = c32;
        }
        if (p + 4 * cnt > nbytes) return -1;
        for (uint64_t k = 0; k < cnt; k++, p += 4) {
            if (src && ne < max_edges) {
                uint32_t d;
                memcpy(&d, adj + p, 4);
                src[ne] = (uint32_t)vid;
                dst[ne] = d;
            }
            ne++;
        }
        vid++;
    }
    return (int64_t)ne;
}

/* ---- orc_psw_run: updates one after another in id order ---- */
static uint32_t psw_mix(uint32_t a, uint32_t b) {
    uint32_t h = (a * 0x9E3779B1u) ^ (b + 0x7F4A7C15u + (a << 6) + (a >> 2));
    h ^= h >> 15;
    h *= 0x2C1B3C6Du;
    h ^= h >> 12;
    return h;
}

long orc_psw_run(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, int program, int niters,
                 int reverse, uint32_t *vdata_out) {
    if (program != 0 && program != 1) return -1;
    /* incidence lists by edge index (the programs here fold them order-free) */
    uint64_t *ioff = calloc((size_t)n + 1, 8), *ooff = calloc((size_t)n + 1, 8);
    uint64_t *iedge = malloc((e ? e : 1) * 8), *oedge = malloc((e ? e : 1) * 8);
    uint32_t *ev = calloc(e ? e : 1, 4), *vv = calloc(n ? n : 1, 4);
    if (!ioff || !ooff || !iedge || !oedge || !ev || !vv) return -2;
    for (uint64_t k = 0; k < e; k++)
        if (src[k] != dst[k]) {
            ioff[dst[k] + 1]++;
            ooff[src[k] + 1]++;
        }
    for (uint32_t v = 0; v < n; v++) {
        ioff[v + 1] += ioff[v];
        ooff[v + 1] += ooff[v];
    }
    uint64_t *ic = malloc(((size_t)n + 1) * 8), *oc = malloc(((size_t)n + 1) * 8);
    memcpy(ic, ioff, ((size_t)n + 1) * 8);
    memcpy(oc, ooff, ((size_t)n + 1) * 8);
    for (uint64_t k = 0; k < e; k++)
        if (src[k] != dst[k]) {
            iedge[ic[dst[k]]++] = k;
            oedge[oc[src[k]]++] = k;
        }
    long violations = 0;
    for (int it = 0; it < niters; it++) {
        for (uint32_t j = 0; j < n; j++) {
            const uint32_t v = reverse ? n - 1 - j : j;
            if (program == 0) { /* basic_smoketest.cpp:47-72 */
                if (it == 0) {
                    for (uint64_t q = ooff[v]; q < ooff[v + 1]; q++) ev[oedge[q]] = v;
                } else {
                    for (uint64_t q = ioff[v]; q < ioff[v + 1]; q++) {
                        const uint32_t u = src[iedge[q]];
                        const uint32_t expected = u + (uint32_t)it - (u > v ? 1u : 0u);
                        if (ev[iedge[q]] != expected) violations++;
                    }
                    for (uint64_t q = ooff[v]; q < ooff[v + 1]; q++) ev[oedge[q]] = v + (uint32_t)it;
                }
                vv[v] = (uint32_t)it + 1u;
                continue;
            }
            /* psw_probe.cpp */
            if (it == 0) {
                for (uint64_t q = ooff[v]; q < ooff[v + 1]; q++) ev[oedge[q]] = psw_mix(v, dst[oedge[q]]);
                vv[v] = v;
                continue;
            }
            uint32_t s_in = 0, x_in = 0, s_out = 0, x_out = 0;
            for (uint64_t q = ioff[v]; q < ioff[v + 1]; q++) {
                s_in += ev[iedge[q]];
                x_in ^= psw_mix(ev[iedge[q]], 1u);
            }
            for (uint64_t q = ooff[v]; q < ooff[v + 1]; q++) {
                s_out += ev[oedge[q]];
                x_out ^= psw_mix(ev[oedge[q]], 2u);
            }
            const uint32_t acc = psw_mix(psw_mix(vv[v], s_in ^ x_out), x_in + s_out);
            for (uint64_t q = ooff[v]; q < ooff[v + 1]; q++) ev[oedge[q]] = psw_mix(acc, dst[oedge[q]]);
            for (uint64_t q = ioff[v]; q < ioff[v + 1]; q++)
                if (src[iedge[q]] > v) ev[iedge[q]] = psw_mix(acc ^ 0x5bd1e995u, src[iedge[q]]);
            vv[v] = acc;
        }
    }
    if (vdata_out) memcpy(vdata_out, vv, (size_t)n * 4);
    free(ioff);
    free(ooff);
    free(iedge);
    free(oedge);
    free(ic);
    free(oc);
    free(ev);
    free(vv);
    return violations;
}

/* ---- orc_psw_pagerank ---- */
static const uint32_t *g_psw_key_src, *g_psw_key_dst;
static int psw_in_cmp(const void *a, const void *b) {
    const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    const uint64_t kx = ((uint64_t)g_psw_key_dst[x] << 32) | g_psw_key_src[x];
    const uint64_t ky = ((uint64_t)g_psw_key_dst[y] << 32) | g_psw_key_src[y];
    if (kx != ky) return kx < ky ? -1 : 1;
    return x < y ? -1 : (x > y ? 1 : 0);
}

int orc_psw_pagerank(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, int niters, double tol,
                     int inmem, float *vdata_out, int *iters_run, double *deltas_out) {
    uint64_t m = 0;
    uint64_t *in_sorted = malloc((e ? e : 1) * 8);
    uint64_t *ioff = calloc((size_t)n + 1, 8), *ooff = calloc((size_t)n + 1, 8);
    uint64_t *oedge = malloc((e ? e : 1) * 8);
    float *ev = calloc(e ? e : 1, 4), *vv = calloc(n ? n : 1, 4);
    if (!in_sorted || !ioff || !ooff || !oedge || !ev || !vv) return -2;
    for (uint64_t k = 0; k < e; k++)
        if (src[k] != dst[k]) {
            in_sorted[m++] = k;
            ioff[dst[k] + 1]++;
            ooff[src[k] + 1]++;
        }
    g_psw_key_src = src;
    g_psw_key_dst = dst;
    qsort(in_sorted, m, 8, psw_in_cmp); /* (dst, src, input index) */
    for (uint32_t v = 0; v < n; v++) {
        ioff[v + 1] += ioff[v];
        ooff[v + 1] += ooff[v];
    }
    uint64_t *oc = malloc(((size_t)n + 1) * 8);
    memcpy(oc, ooff, ((size_t)n + 1) * 8);
    for (uint64_t k = 0; k < e; k++)
        if (src[k] != dst[k]) oedge[oc[src[k]]++] = k;
    int it = 0;
    double delta = 0.0;
    for (; it < niters; it++) {
        if (!inmem) delta = 0.0;
        for (uint32_t v = 0; v < n; v++) {
            const uint64_t nout = ooff[v + 1] - ooff[v];
            if (it == 0) {
                for (uint64_t q = ooff[v]; q < ooff[v + 1]; q++) ev[oedge[q]] = (float)(1.0 / (double)(int)nout);
                vv[v] = 0.15f;
                continue;
            }
            float sum = 0.0f;
            for (uint64_t q = ioff[v]; q < ioff[v + 1]; q++) sum += ev[in_sorted[q]];
            const float rank = (float)(0.15 + 0.85 * (double)sum);
            if (nout > 0) {
                const float share = rank / (float)(int)nout;
                for (uint64_t q = ooff[v]; q < ooff[v + 1]; q++) ev[oedge[q]] = share;
            }
            delta += (double)fabsf(rank - vv[v]);
            vv[v] = rank;
        }
        if (deltas_out) deltas_out[it] = delta;
        if (it > 0 && delta < tol) {
            it++;
            break;
        }
    }
    if (iters_run) *iters_run = it;
    if (vdata_out) memcpy(vdata_out, vv, (size_t)n * 4);
    free(in_sorted);
    free(ioff);
    free(ooff);
    free(oedge);
    free(oc);
    free(ev);
    free(vv);
    return 0;
}

/* ---- orc_psw_sched: apps/psw_sched.cpp under the bitset scheduler ---- */
static uint32_t sched_mix(uint32_t a, uint32_t b) {
    uint32_t h = (a ^ 0x85EBCA6Bu) * 0x9E3779B1u + b;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    return h;
}

typedef struct {
    uint8_t *cur, *next;
    int has_new;
    uint32_t n;
} orc_sched;

static void sched_add(orc_sched *s, uint32_t v, int also) { /* bitset_scheduler.hpp:66-73 */
    if (v >= s->n) return;
    s->next[v] = 1;
    if (also) s->cur[v] = 1;
    s->has_new = 1;
}

static void sched_remove(orc_sched *s, uint32_t from, uint32_t to) { /* :86-88, [from, to] */
    for (uint64_t v = from; v <= to && v < s->n; v++) s->next[v] = 0;
}

static void sched_new_iteration(orc_sched *s, int it) { /* :50-58 */
    if (it == 0) return;
    uint8_t *t = s->cur;
    s->cur = s->next;
    s->next = t;
    memset(s->next, 0, s->n);
}

int orc_psw_sched(uint32_t n, uint64_t e, const uint32_t *src, const uint32_t *dst, int nivl,
                  const uint32_t *ivl_end, int inmem, int niters, uint64_t *vdata_out, uint64_t *log_out,
                  int *nlog, uint64_t *tasks_out, int *ntasks) {
    /* incidence lists by edge index; the program folds them order-free */
    uint64_t *ioff = calloc((size_t)n + 1, 8), *ooff = calloc((size_t)n + 1, 8);
    uint64_t *iedge = malloc((e ? e : 1) * 8), *oedge = malloc((e ? e : 1) * 8);
    uint32_t *ev = calloc(e ? e : 1, 4);
    uint64_t *vv = calloc(n ? n : 1, 8);
    uint8_t *ba = malloc(n ? n : 1), *bb = calloc(n ? n : 1, 1), *snap = malloc(n ? n : 1);
    if (!ioff || !ooff || !iedge || !oedge || !ev || !vv || !ba || !bb || !snap) return -2;
    for (uint64_t k = 0; k < e; k++)
        if (src[k] != dst[k]) {
            ioff[dst[k] + 1]++;
            ooff[src[k] + 1]++;
        }
    for (uint32_t v = 0; v < n; v++) {
        ioff[v + 1] += ioff[v];
        ooff[v + 1] += ooff[v];
    }
    uint64_t *ic = malloc(((size_t)n + 1) * 8), *oc = malloc(((size_t)n + 1) * 8);
    memcpy(ic, ioff, ((size_t)n + 1) * 8);
    memcpy(oc, ooff, ((size_t)n + 1) * 8);
    for (uint64_t k = 0; k < e; k++)
        if (src[k] != dst[k]) {
            iedge[ic[dst[k]]++] = k;
            oedge[oc[src[k]]++] = k;
        }
    free(ic);
    free(oc);
    memset(ba, 1, n); /* initialize_scheduler: add_task_to_all */
    orc_sched S = {ba, bb, 1, n};
    uint64_t updates = 0, work = 0;
    int it_last = 0, started = 0;
    *nlog = 0;
    *ntasks = 0;
    for (int it = 0; it < niters; it++) {
        it_last = it;
        started = 1;
        if (it == 2 && n > 0) sched_add(&S, 0, 0); /* before_iteration */
        /* passes: the intervals in order, or the whole graph in memory */
        int np = inmem ? 1 : nivl;
        if (inmem) {
            if (it > 0 && !S.has_new) break;
            sched_new_iteration(&S, it);
        } else {
            if (!S.has_new) break;
            S.has_new = 0;
            sched_new_iteration(&S, it);
        }
        uint32_t lo = 0;
        int stop = 0;
        for (int p = 0; p < np && !stop; p++) {
            uint64_t hi64 = inmem ? n : (uint64_t)ivl_end[p] + 1;
            uint32_t hi = hi64 > n ? n : (uint32_t)hi64;
            uint64_t cnt = 0;
            for (uint32_t v = lo; v < hi; v++) {
                snap[v] = (inmem && it == 0) ? 1 : S.cur[v];
                if (snap[v]) {
                    cnt++;
                    work += (ioff[v + 1] - ioff[v]) + (ooff[v + 1] - ooff[v]);
                }
            }
            updates += cnt;
            if (inmem) {
                if (cnt == 0) {
                    stop = 1;
                    break;
                }
                S.has_new = 0;
            }
            for (uint32_t v = lo; v < hi; v++) {
                if (!snap[v]) continue;
                const uint32_t label = it == 0 ? v : (uint32_t)vv[v];
                uint32_t trace = it == 0 ? 0u : (uint32_t)(vv[v] >> 32);
                uint32_t m = label;
                for (int side = 0; side < 2; side++) {
                    const uint64_t *off = side ? ooff : ioff, *lst = side ? oedge : iedge;
                    for (uint64_t q = off[v]; q < off[v + 1]; q++) {
                        const uint32_t x = ev[lst[q]];
                        if (x != 0 && x - 1 < m) m = x - 1;
                    }
                }
                trace = sched_mix(trace, (uint32_t)it * 0x10001u + m);
                for (int side = 0; side < 2; side++) {
                    const uint64_t *off = side ? ooff : ioff, *lst = side ? oedge : iedge;
                    for (uint64_t q = off[v]; q < off[v + 1]; q++) {
                        const uint64_t k = lst[q];
                        if (ev[k] != m + 1) {
                            ev[k] = m + 1;
                            const uint32_t nb = side ? dst[k] : src[k];
                            if (nb != v) sched_add(&S, nb, ((nb ^ v) & 1u) != 0);
                        }
                    }
                }
                if (v % 7 == 3 && (it & 1)) sched_remove(&S, v, v);
                vv[v] = ((uint64_t)trace << 32) | m;
            }
            lo = hi;
        }
        if (stop) break;
        if (it == 1) sched_remove(&S, 5, 9); /* after_iteration */
        uint64_t t = 0;
        for (uint32_t v = 0; v < n; v++) t += S.cur[v];
        tasks_out[(*ntasks)++] = t;
        if (!inmem) {
            log_out[3 * *nlog] = (uint64_t)it;
            log_out[3 * *nlog + 1] = updates;
            log_out[3 * *nlog + 2] = work;
            (*nlog)++;
        }
    }
    if (inmem && started) {
        log_out[0] = (uint64_t)it_last;
        log_out[1] = updates;
        log_out[2] = work;
        *nlog = 1;
    }
    memcpy(vdata_out, vv, (size_t)n * 8);
    free(ioff);
    free(ooff);
    free(iedge);
    free(oedge);
    free(ev);
    free(vv);
    free(S.cur);
    free(S.next);
    free(snap);
    return 0;
}
