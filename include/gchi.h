/*
 * gchi.h -- C ABI of the MI355X (gfx950) GraphChi hot-path library (libgchi.so).
 *
 * Every entry point is extern "C", takes plain pointers and sizes, returns an
 * int status (GCHI_OK == 0, negative gchi_status on error) and never throws or
 * aborts across the boundary.  Device pointers ("d_" prefix) must point to HBM
 * on the current HIP device; host pointers ("h_" prefix) to host memory.
 * Streams are hipStream_t passed as void* (NULL == default stream).
 *
 * Workspace convention: entry points that need scratch device memory take
 * (d_work, work_bytes); the matching *_workspace() call reports the size.
 * Asynchronous compute calls (sweeps, sorts, CC, TC counting, CSC builds)
 * never allocate device memory, so they can be captured into a HIP graph.
 * Calls documented as host-synchronous read results back to the host.
 *
 * Each declaration names the reference interface it replaces (paths relative
 * to the GraphChi source tree).
 */
#ifndef GCHI_H
#define GCHI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ status */
typedef enum gchi_status {
    GCHI_OK = 0,
    GCHI_ERR_INVALID = -1,      /* bad argument (NULL, size, range)            */
    GCHI_ERR_HIP = -2,          /* HIP runtime error (message in last_error)   */
    GCHI_ERR_WORKSPACE = -3,    /* workspace too small                          */
    GCHI_ERR_UNSUPPORTED = -4,  /* shape/operator not supported on the device   */
    GCHI_ERR_FORMAT = -5,       /* malformed shard / file content               */
    GCHI_ERR_OVERFLOW = -6,     /* result does not fit the output type          */
    GCHI_ERR_NO_DEVICE = -7,    /* no HIP device / library not initialised      */
    GCHI_ERR_IO = -8,           /* file system error                            */
    GCHI_ERR_TIMEOUT = -9       /* bounded device wait exceeded its spin limit  */
} gchi_status;

/* Library version (major*10000 + minor*100 + patch). */
int gchi_version(void);
/* Short text for a status code. */
const char *gchi_strerror(int status);
/* Copies the last error message of the calling thread into buf. */
int gchi_last_error(char *buf, size_t len);

/* Selects the HIP device used by the calling thread and warms the runtime.
 * Replaces the engine/iomgr construction of src/engine/graphchi_engine.hpp:148-200. */
int gchi_device_open(int device);
/* Number of compute units and GPU arch name of the current device. */
int gchi_device_info(int *num_cus, char *arch, size_t arch_len);

/* Device memory for host-side callers (the engine's RAII buffers). */
int gchi_malloc(void **d_ptr, size_t bytes);
int gchi_free(void *d_ptr);
int gchi_synchronize(void *stream);
/* Pinned (page-locked) host memory for the one-time shard upload: the engines
 * read shard files straight into it, so the upload runs at DMA speed.  Replaces
 * the pread buffers of src/io/stripedio.hpp:480-640. */
int gchi_host_alloc(void **h_ptr, size_t bytes);
int gchi_host_free(void *h_ptr);
/* Copies bytes between any host/device pointers on the stream and waits for it
 * (the pinned one-time upload/readback of the engine). */
int gchi_memcpy(void *dst, const void *src, size_t bytes, void *stream);

/* ----------------------------------------------------- synthetic R-MAT input */
/* Deterministic R-MAT edge generator (counter-based RNG: edge i of a graph is a
 * pure function of (seed, i)); self-loops are kept here and removed by
 * gchi_drop_self_loops, which restates sharder::preprocessing_add_edge
 * (src/preprocessing/sharder.hpp:427-446).  abcd = quadrant probabilities. */
int gchi_rmat_generate(uint32_t *d_src, uint32_t *d_dst, uint64_t first_edge, uint64_t nedges,
                       int scale, uint64_t seed, const double abcd[4], void *stream);
int gchi_rmat_generate_host(uint32_t *h_src, uint32_t *h_dst, uint64_t first_edge, uint64_t nedges,
                            int scale, uint64_t seed, const double abcd[4], int nthreads);

/* Stable compaction dropping src==dst pairs (sharder.hpp:428-431).  Writes the
 * kept count to *d_count (uint64) and the maximum vertex id over kept edges to
 * *d_maxid (uint32; 0 when empty) -- sharder.hpp:445. */
int gchi_drop_self_loops_workspace(uint64_t nedges, size_t *bytes);
int gchi_drop_self_loops(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges,
                         uint32_t *d_src_out, uint32_t *d_dst_out, uint64_t *d_count,
                         uint32_t *d_maxid, void *d_work, size_t work_bytes, void *stream);

/* Edge-list text parsed on the device: convert_edgelist
 * (src/preprocessing/conversions.hpp:210-288) feeding preprocessing_add_edge
 * (src/preprocessing/sharder.hpp:427-446).  d_text holds the file's bytes.
 * Lines whose first byte is '#' or '%' are comments; other lines split on tab,
 * comma and space, from = atoi(field 1), to = atoi(field 2) (further fields are
 * edge values and ignored here); self-loops are dropped.  The kept pairs are
 * written in file order to d_src/d_dst (up to max_edges; 0 = count only),
 * *h_nedges = their number, *h_maxid = the largest id among them.  A line with
 * fewer than two fields, or of 1023 bytes or more, is GCHI_ERR_FORMAT (the
 * reference asserts on the first and splits the second).  Host-synchronous. */
int gchi_parse_edgelist_workspace(uint64_t nbytes, size_t *bytes);
int gchi_parse_edgelist(const uint8_t *d_text, uint64_t nbytes, uint32_t *d_src, uint32_t *d_dst, uint64_t max_edges,
                        uint64_t *h_nedges, uint32_t *h_maxid, void *d_work, size_t work_bytes, void *stream);

/* ---------------------------------------------------------------- radix sort */
/* Stable LSD onesweep radix sort of 64-bit keys on their low key_bits bits
 * (8-bit digits; optional 32-bit payload), n < 2^32.  Replaces
 * iSort/radixStep/eBits of src/util/radixSort.hpp:49-116.  d_keys/d_vals are
 * sorted in place (the library ping-pongs through the workspace). */
int gchi_sort_workspace(uint64_t n, int key_bits, int with_values, size_t *bytes);
int gchi_sort_u64(uint64_t *d_keys, uint32_t *d_vals, uint64_t n, int key_bits,
                  void *d_work, size_t work_bytes, void *stream);

/* Device-side status of an asynchronous call.  gchi_sort_u64,
 * gchi_sort_pairs_dst_src and gchi_csc_build keep a status word in the first 4
 * bytes of their workspace: GCHI_WORK_BAD_ID when an input vertex id was out of
 * range, GCHI_WORK_TIMEOUT when a bounded device wait (the sort's decoupled
 * look-back) gave up -- the output is then invalid.  gchi_work_status waits for
 * the stream and returns GCHI_OK, GCHI_ERR_INVALID or GCHI_ERR_TIMEOUT.  Calls
 * that are host-synchronous anyway (gchi_shard_build, gchi_tc_prepare,
 * gchi_degree_order, gchi_pr_plan_create*) check their word themselves. */
#define GCHI_WORK_BAD_ID 1u
#define GCHI_WORK_TIMEOUT 2u
int gchi_work_status(const void *d_work, void *stream);
/* Test hook: bound of every sort look-back wait in polls (0 = the default,
 * 2^26) and, with force_stall, waits that treat every predecessor as not yet
 * published, so that the bound is reached and the status reports it while
 * every kernel still exits.  Process-wide; not for production use. */
int gchi_debug_sort_lookback(uint32_t spin_limit, int force_stall);

/* Sorts uint32 (src,dst) pairs by (dst,src) -- the order the sharder's shovel
 * sort (sharder.hpp:192, dst key) followed by a full tie order produces.
 * Pairs are read from d_src/d_dst and written sorted to d_src_out/d_dst_out,
 * which may be d_src/d_dst themselves (an in-place sort).
 * id_bits = bits needed for the largest vertex id. */
int gchi_sort_pairs_dst_src_workspace(uint64_t n, int id_bits, size_t *bytes);
int gchi_sort_pairs_dst_src(const uint32_t *d_src, const uint32_t *d_dst, uint64_t n, int id_bits,
                            uint32_t *d_src_out, uint32_t *d_dst_out,
                            void *d_work, size_t work_bytes, void *stream);

/* ------------------------------------------------------------- device graph */
/* Device-resident in-edge (CSC) view of a graph: in_off[N+1], in_src[E] sorted
 * by (dst, src) -- the memory-shard adjacency of every interval at once
 * (src/shards/memoryshard.hpp:384-519) -- plus out-degrees
 * (src/engine/auxdata/degree_data.hpp:45-181).  Plain struct of device
 * pointers owned by the caller. */
typedef struct gchi_dgraph {
    uint32_t nvertices;        /* N = 1 + max vertex id */
    uint32_t reserved;
    uint64_t nedges;           /* E (<= 2^32 - 2^25) */
    const uint32_t *in_off;    /* [N+1] */
    const uint32_t *in_src;    /* [E]   */
    const uint32_t *outdeg;    /* [N]   */
    const uint32_t *indeg;     /* [N]   (optional, may be NULL) */
} gchi_dgraph;

/* Builds the CSC arrays from an edge list already resident in HBM;
 * nedges <= 2^32 - 2^25. */
int gchi_csc_build_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes);
int gchi_csc_build(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                   uint32_t *d_in_off, uint32_t *d_in_src, uint32_t *d_outdeg, uint32_t *d_indeg,
                   void *d_work, size_t work_bytes, void *stream);

/* ------------------------------------------------- generic vertex programs */
/* Adjacency for running a program's own update() on the device in the
 * reference's PSW order (graphchi_engine::exec_updates,
 * src/engine/graphchi_engine.hpp:396-454; the order basic_smoketest.cpp:65
 * checks: a vertex sees what lower ids wrote this iteration).  From an edge
 * list in HBM: out-edges in stable (src, dst) order (d_out_off[N+1],
 * d_out_dst[E]; edge values are stored in this order), in-edges in stable
 * (dst, src) order (d_in_off[N+1], d_in_src[E]) with d_in_eidx[E] = the
 * index of each in-edge's value in out-edge order, and d_lower[N] = the number
 * of incident edges whose other end has a smaller id.  nedges <= 2^32 - 2^25.
 * Status word in the workspace (gchi_work_status): a bad id or a sort wait. */
int gchi_psw_graph_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes);
int gchi_psw_graph_build(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                         uint32_t *d_out_off, uint32_t *d_out_dst, uint32_t *d_in_off, uint32_t *d_in_src,
                         uint32_t *d_in_eidx, uint32_t *d_lower, void *d_work, size_t work_bytes, void *stream);
/* As gchi_psw_graph_build; d_out_eidx[E] (may be NULL) gets the input index of
 * each out-edge, which places edge values read in input (shard) order. */
int gchi_psw_graph_build_ex(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                            uint32_t *d_out_off, uint32_t *d_out_dst, uint32_t *d_in_off, uint32_t *d_in_src,
                            uint32_t *d_in_eidx, uint32_t *d_lower, uint32_t *d_out_eidx, void *d_work,
                            size_t work_bytes, void *stream);

/* ------------------------------------------------------------------ PageRank */
/* Affine functional kernel of example_apps/pagerank_functional.cpp:48-83:
 *   value_to_neighbor : x_u = r_u / outdeg_u            (FP32, :78-81)
 *   op_neighborval/plus: s_v = sum of x_u over in-edges  (FP32, :62-69)
 *   compute_vertexvalue: r_v = (float)(c_v + damping*(double)s_v)  (:72-75)
 * with c_v = teleport (0.15/N reproduces pagerank_functional, 0.15 the
 * pagerank.cpp form), or d_graph_teleport[g] for the graph g of v in a
 * multi-graph plan, plus seed_teleport when v is set in d_seed_bits
 * (personalized PageRank: restart mass on a seed set). */
typedef struct gchi_pr_op {
    double teleport;
    double damping;
    const double *d_graph_teleport;  /* optional, [ngraphs] (device) */
    const uint32_t *d_seed_bits;     /* optional, bitmap over N (device) */
    double seed_teleport;
} gchi_pr_op;

/* Launch shape of the pull sweep (filled by gchi_autotune_pr or defaults). */
typedef struct gchi_pr_tune {
    int edges_per_block;   /* W: multiple of 256 in [256,4096]; 0 = 2048.  Rows with
                              in-degree >= W are hub rows, cut into 512-edge chunks */
    int short_row_max;     /* rows up to this length use one lane per row; 0 = 32 */
    int persistent;        /* 0: one workgroup per tile; 1: persistent workgroups pull
                              tiles from a device work counter; 2: one 1024-thread
                              workgroup per CU whose wavefronts each work through a
                              fixed list of tiles, software-pipelined (W = 256),
                              optionally with the LDS cache of hot_vertices; 3: four
                              persistent 256-thread workgroups per CU with a static
                              tile list, each with its own (smaller) LDS cache;
                              4: the rank-space tiled sweep -- vertices relabelled by
                              out-degree, every row's in-edges packed into 512-slot
                              tiles reduced by one wavefront each (segmented sums),
                              rows crossing tiles summed from per-tile partials; the
                              first hot_vertices ranks stay in LDS.  edges_per_block
                              and short_row_max do not apply; one graph only
                              (gchi_pr_plan_create); a HIP-graph exec keeps the state
                              in rank order between sweeps, a standalone sweep maps its
                              vertex-order buffers in and out.  N < 2^31 */
    int hot_vertices;      /* 0: gathers read x by vertex id.  Otherwise the plan keeps a
                              gather layout: sources ranked by out-degree (descending,
                              ties by id), edges store the rank, and every sweep first
                              copies x into that order.  > 0: persistent workgroups also
                              keep the first hot_vertices values of the layout in LDS
                              (rounded down to a multiple of 1024, capped by LDS;
                              implies persistent = 2 unless 3 is given); -1: the
                              layout without the LDS cache.  N < 2^30. */
} gchi_pr_tune;

/* Sweep plan: the graph's in-edges regrouped into regular rows (LDS-staged
 * blocks of < 2W edges) and hub rows (chunked), plus the sweep's scratch.
 * Built once per graph (the one-time upload that replaces the per-iteration
 * shard loads of src/engine/graphchi_engine.hpp:825-936).  A plan's scratch is
 * shared by its sweeps: sweeps of one plan must not overlap in time. */
typedef struct gchi_pr_plan gchi_pr_plan;
typedef struct gchi_pr_plan_info {
    uint32_t nvertices;
    uint32_t edges_per_block;
    uint64_t nedges;
    uint32_t short_row_max;
    uint32_t regular_blocks;
    uint32_t hub_rows;
    uint32_t hub_chunks;
    uint64_t regular_edges;
    uint64_t hub_edges;
    uint32_t ngraphs;
    uint32_t persistent;
    uint32_t persistent_workgroups;
    uint32_t hot_vertices;     /* values cached in LDS per workgroup (0: none) */
    uint32_t gather_sources;   /* vertices in the gather layout (0: no layout) */
    uint32_t hot_threads;      /* workgroup size of the LDS-cached sweep */
} gchi_pr_plan_info;

int gchi_pr_plan_create(const gchi_dgraph *g, const gchi_pr_tune *tune, gchi_pr_plan **out, void *stream);
/* Plan over a disjoint union of ngraphs graphs: graph k owns vertices
 * [h_vbase[k], h_vbase[k+1]) of g (h_vbase[0] = 0, h_vbase[ngraphs] = N, no edge
 * crosses graphs).  Blocks never straddle graphs; sweeps write per-graph
 * statistics (d_stats[2k], d_stats[2k+1]).  Many independent PageRank problems
 * run in one launch per sweep. */
int gchi_pr_plan_create_multi(const gchi_dgraph *g, const uint32_t *h_vbase, uint32_t ngraphs,
                              const gchi_pr_tune *tune, gchi_pr_plan **out, void *stream);
/* Plan over ngraphs separate graphs, each its own CSC (local vertex ids 0..N_k-1,
 * its own in_off/in_src/outdeg, any device allocations): graph k owns plan
 * vertices [V_k, V_k + N_k), V_k = N_0 + ... + N_{k-1}, in the x/r vectors;
 * d_graph_teleport and d_stats are per graph as for gchi_pr_plan_create_multi.
 * Each graph needs E_k <= 2^32 - 2^25; the total may exceed 2^32 (offsets are 32-bit
 * per graph, 64-bit across graphs), total vertices < 2^32.  The plan copies
 * what it needs: the graphs may be freed once it is created.  All graphs of a
 * batch are swept by one launch of the sweep kernel. */
int gchi_pr_plan_create_batch(const gchi_dgraph *graphs, uint32_t ngraphs, const gchi_pr_tune *tune,
                              gchi_pr_plan **out, void *stream);
int gchi_pr_plan_destroy(gchi_pr_plan *plan);
int gchi_pr_plan_info_get(const gchi_pr_plan *plan, gchi_pr_plan_info *info);

/* Iteration 0 of the synchronous functional engine (functional_bulksync.hpp:84-87,
 * 107-113): r = init, x = r/outdeg (0 where outdeg == 0). */
int gchi_pr_init(const gchi_pr_plan *plan, float init, float *d_x, float *d_r, void *stream);
/* Personalized start: r = seed_value on seed vertices, 0 elsewhere. */
int gchi_pr_init_seeds(const gchi_pr_plan *plan, const uint32_t *d_seed_bits, float seed_value, float *d_x,
                       float *d_r, void *stream);

/* One synchronous gather->apply->scatter sweep (functional_bulksync.hpp:91-113,
 * functional_engine.hpp:50-121): two kernels (sweep, finalize), three with a
 * gather layout (its refresh from d_x_in comes first).
 * d_stats[0] = sum |r_out - r_in| (FP64), d_stats[1] = sum r_out (FP64), both
 * reduced in a fixed order (the log_change deltas of
 * src/api/graphchi_context.hpp:78-102). */
int gchi_pr_sweep(const gchi_pr_plan *plan, const gchi_pr_op *op, const float *d_x_in, const float *d_r_in,
                  float *d_x_out, float *d_r_out, double *d_stats, void *stream);

/* init + nsweeps sweeps captured into one HIP graph (instantiated once,
 * replayed by gchi_pr_exec_launch).  Sweep k reads buffers [k&1] and writes
 * [(k+1)&1]; the final ranks are in d_r{nsweeps&1}.  d_stats_all receives
 * 2*ngraphs*nsweeps doubles.  Replaces the iteration loop of graphchi_engine::run
 * (src/engine/graphchi_engine.hpp:779-956) for functional kernels. */
typedef struct gchi_pr_exec gchi_pr_exec;
int gchi_pr_exec_create(const gchi_pr_plan *plan, const gchi_pr_op *op, float init, int nsweeps, float *d_x0,
                        float *d_x1, float *d_r0, float *d_r1, double *d_stats_all, gchi_pr_exec **out);
/* As gchi_pr_exec_create; do_init = 0 captures sweeps only (the caller's
 * buffers carry the state), so replays continue an iteration. */
int gchi_pr_exec_create_ex(const gchi_pr_plan *plan, const gchi_pr_op *op, float init, int do_init, int nsweeps,
                           float *d_x0, float *d_x1, float *d_r0, float *d_r1, double *d_stats_all,
                           gchi_pr_exec **out);
int gchi_pr_exec_launch(gchi_pr_exec *exec, void *stream);
int gchi_pr_exec_destroy(gchi_pr_exec *exec);

/* Semi-synchronous sweeps: the default mode of the reference's
 * pagerank_functional (run_functional_unweighted_semisynchronous,
 * src/api/functional/functional_api.hpp:119-137; vertex program
 * src/api/functional/functional_semisync.hpp:76-110; phase order
 * src/engine/functional/functional_engine.hpp:50-121; interval and
 * sub-interval loop src/engine/graphchi_engine.hpp:825-921).  Every edge holds
 * one value, overwritten when its source is updated, and the vertex blocks
 * [h_bstart[j], h_bstart[j+1]) -- the engine's sub-intervals, in execution
 * order -- run one after another: block j gathers the x written this sweep by
 * the blocks before it and the previous sweep's x of the rest, then applies and
 * scatters.  One block is the synchronous sweep.  Each block gets a plan over
 * its rows built with the tune's shape (vertex-id gathers: hot_vertices is
 * ignored); the graph's arrays must outlive the object. */
typedef struct gchi_pr_blocked gchi_pr_blocked;
int gchi_pr_blocked_create(const gchi_dgraph *g, const uint32_t *h_bstart, uint32_t nblocks, const gchi_pr_tune *tune,
                           gchi_pr_blocked **out, void *stream);
int gchi_pr_blocked_destroy(gchi_pr_blocked *b);
/* Iteration 0: r = init and x = r/outdeg, or with d_seed_bits r = seed_value
 * on the seeds and 0 elsewhere. */
int gchi_pr_blocked_init(const gchi_pr_blocked *b, float init, const uint32_t *d_seed_bits, float seed_value,
                         float *d_x, float *d_r, void *stream);
/* One sweep: d_x is updated in place, block by block; ranks go from d_r_in to
 * d_r_out (distinct); d_stats[0..1] = sum |r_out - r_in| and sum r_out (FP64,
 * block by block in order).  op->d_graph_teleport must be NULL. */
int gchi_pr_blocked_sweep(const gchi_pr_blocked *b, const gchi_pr_op *op, float *d_x, const float *d_r_in,
                          float *d_r_out, double *d_stats, void *stream);

/* Autotuner: times every candidate launch shape (edges_per_block in
 * {512..4096}, short_row_max in {8..64}, persistent 0/1, the
 * wavefront-tiled sweep, W = 256, with hot_vertices in {0, -1, 16384, max},
 * and the cached persistent workgroups, W in {1024, 2048}) on the caller's
 * graph with HIP events (sweeps_per_candidate sweeps after a warm-up) and
 * returns the fastest; cached per (N, E) for the process unless h_tried is
 * given (then every candidate is re-timed and reported). */
int gchi_autotune_pr(const gchi_dgraph *g, int sweeps_per_candidate, gchi_pr_tune *best, double *best_ms,
                     gchi_pr_tune *h_tried, double *h_ms, int max_tried, int *ntried, void *stream);

/* ------------------------------------------------- batched personalized PageRank */
/* Many personalized-PageRank queries (restart sets of 1..64 seeds; kernel
 * r_v = (v in S ? (1-d)/|S| : 0) + d * sum r_u/outdeg_u, from r = restart vector)
 * swept together: 64 queries per block, ranks stored vertex-major (64 per
 * vertex), one wavefront per destination row with one lane per query, tiles
 * (query block, row block / hub chunk) pulled from a device work queue by a
 * persistent kernel.  Phase 1 stores ranks in BF16 (FP32 accumulation) for up to
 * bf16_sweeps sweeps or until every query's L1 residual <= bf16_switch_residual
 * or stops falling; phase 2 refines in FP32 until every query satisfies
 * residual * d/(1-d) <= tol (a bound on the L1 distance to the fixed point,
 * iterates being monotone).  Residuals are FP64.  Sweeps run in chunks of
 * check_every (rounded down to even, at least 2), each chunk one HIP-graph
 * launch (captured once per phase and batch) whose last sweep computes the
 * residuals, the only host readback; a query block whose 64 queries all meet
 * the bound leaves the work queue, so later chunks sweep only the rest.
 * max_sweeps is honoured in whole chunks. */
typedef struct gchi_ppr_config {
    double damping;
    double tol;
    double bf16_switch_residual;
    int bf16_sweeps;     /* 0: FP32 only */
    int max_sweeps;
    int check_every;     /* residual checked every k sweeps (default 4) */
    int reserved;
} gchi_ppr_config;
typedef struct gchi_ppr_result {
    int sweeps_bf16;
    int sweeps_fp32;
    int converged;
    int block_sweeps_fp32;  /* FP32 sweeps summed over the query blocks (a converged block stops) */
    double max_residual;
} gchi_ppr_result;
typedef struct gchi_ppr_batch gchi_ppr_batch;
int gchi_ppr_batch_create(const gchi_pr_plan *plan, uint32_t nqueries, gchi_ppr_batch **out);
int gchi_ppr_batch_destroy(gchi_ppr_batch *batch);
/* h_seed_off[nqueries+1], h_seeds: restart set of query q is
 * h_seeds[h_seed_off[q] .. h_seed_off[q+1]) (distinct ids).  h_residuals
 * (optional) receives the last L1 residual of every query. */
int gchi_ppr_batch_run(gchi_ppr_batch *batch, const uint32_t *h_seed_off, const uint32_t *h_seeds,
                       const gchi_ppr_config *cfg, gchi_ppr_result *result, double *h_residuals, void *stream);
/* Copies query q's rank vector (N floats) to d_out. */
int gchi_ppr_batch_extract(const gchi_ppr_batch *batch, uint32_t query, float *d_out, void *stream);

/* --------------------------------------------------------- connected components */
/* Labels = smallest vertex id of each weakly connected component: the fixed
 * point of ConnectedComponentsProgram (example_apps/connectedcomponents.cpp:79-138),
 * reached here by a lock-free union-find (roots always hooked under smaller
 * roots).  Edges are read in either direction. */
int gchi_cc_workspace(uint32_t nvertices, size_t *bytes);
int gchi_cc_minlabel(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                     uint32_t *d_labels, void *d_work, size_t work_bytes, void *stream);
/* As gchi_cc_minlabel.  flags: GCHI_CC_SYMMETRIC promises a symmetric list (every
 * u-v present as both (u,v) and (v,u), the form the reference's undirected
 * inputs take in its edge files), so only the entries with u < v are unioned;
 * the labels are the same. */
#define GCHI_CC_SYMMETRIC 1
int gchi_cc_minlabel_ex(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices, int flags,
                        uint32_t *d_labels, void *d_work, size_t work_bytes, void *stream);
/* The reference's per-iteration form: labels = ids (iteration 0), then one
 * synchronous min-propagation sweep over every edge in both directions;
 * *d_changed = number of labels that changed (0 = converged,
 * connectedcomponents.cpp:125-138). */
int gchi_cc_labelprop_init(uint32_t *d_labels, uint32_t nvertices, void *stream);
int gchi_cc_labelprop_sweep(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                            const uint32_t *d_labels_in, uint32_t *d_labels_out, uint64_t *d_changed, void *stream);

/* The same synchronous sweep pulled over a symmetric CSR (d_off[N+1], d_adj;
 * every edge in both rows): d_labels_out[v] = min(d_labels_in[v], d_labels_in[u]
 * for every neighbour u), one write per vertex, no atomics on labels;
 * *d_changed = labels that dropped. */
int gchi_cc_pull_sweep(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, const uint32_t *d_labels_in,
                       uint32_t *d_labels_out, uint64_t *d_changed, void *stream);

/* The reference's CC under its selective scheduler (ConnectedComponentsProgram
 * with scheduler on, connectedcomponents.cpp:79-121 and :81,85,112; the two
 * bitsets of src/engine/bitset_scheduler.hpp:38-104): round 0 evaluates every
 * vertex (label = min of its id and its neighbours' labels); a vertex whose label
 * drops schedules all its neighbours for the next round; the rounds end when no
 * label drops.  Input: a symmetric adjacency in CSR form (d_off[N+1] with
 * d_off[N] = E, d_adj[E]; every edge present in both rows, as gchi_csc_build
 * gives for a list holding each edge both ways).  d_labels[N] ends equal to
 * gchi_cc_minlabel's.  max_rounds 0 = no limit; a limit reached before the fixed
 * point is an error.  *h_rounds = rounds run, *h_evaluated (optional) = vertex
 * evaluations summed over the rounds.  Host-synchronous (one 24-byte read per
 * round). */
int gchi_cc_scheduled_workspace(uint32_t nvertices, uint64_t nedges, size_t *bytes);
int gchi_cc_scheduled(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t max_rounds,
                      uint32_t *d_labels, uint32_t *h_rounds, uint64_t *h_evaluated, void *d_work, size_t work_bytes,
                      void *stream);

/* The same labels from a symmetric CSR (d_off[N+1], d_adj; every edge in both
 * rows) by union-find with subgraph sampling (Afforest, Sutton et al., IPDPS
 * 2018): each vertex is linked to its first two neighbours, the component most
 * sampled vertices fall in is found, and only the vertices outside it link
 * their remaining neighbours -- the giant component's edges are never read.
 * Host-synchronous (reads 1,024 sampled roots, then a status word). */
int gchi_cc_afforest_workspace(uint32_t nvertices, size_t *bytes);
int gchi_cc_afforest(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t *d_labels,
                     void *d_work, size_t work_bytes, void *stream);

/* As gchi_cc_scheduled, reporting each round to on_round (host, between
 * rounds): round index from 0, labels that dropped in it, vertices evaluated in
 * it.  A nonzero return stops after that round (not an error; the labels are
 * those of the rounds run).  This is the engine's per-iteration hook
 * (before/after_iteration and set_last_iteration of
 * src/engine/graphchi_engine.hpp:779-956). */
typedef int (*gchi_cc_round_fn)(void *user, uint32_t round, uint64_t dropped, uint64_t evaluated);
int gchi_cc_scheduled_ex(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t max_rounds,
                         uint32_t *d_labels, uint32_t *h_rounds, uint64_t *h_evaluated, gchi_cc_round_fn on_round,
                         void *user, void *d_work, size_t work_bytes, void *stream);

/* ------------------------------------------------------------ triangle counting */
/* Degree ordering (order_by_degree, src/preprocessing/util/orderbydegree.hpp:58-100:
 * rank by (in+out degree, id), d_vertexmap[old] = new id) and the oriented,
 * deduplicated adjacency N+(v) = neighbours of higher rank, ascending
 * (d_off[N+1], d_adj[<= E]).  *h_noriented = number of oriented edges. */
int gchi_tc_prepare_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes);
int gchi_tc_prepare(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                    uint32_t *d_vertexmap, uint32_t *d_off, uint32_t *d_adj, uint64_t *h_noriented,
                    void *d_work, size_t work_bytes, void *stream);
/* The translation table of order_by_degree alone, and an edge relabel
 * (src' = map[src], dst' = map[dst]) -- the re-sharding input of
 * orderbydegree.hpp:112-142.  Both read back an id-range check, so they are
 * host-synchronous; gchi_relabel_edges allocates its 4-byte check flag. */
int gchi_degree_order_workspace(uint64_t nedges, uint32_t nvertices, size_t *bytes);
int gchi_degree_order(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                      uint32_t *d_vertexmap, void *d_work, size_t work_bytes, void *stream);
int gchi_relabel_edges(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t nvertices,
                       const uint32_t *d_map, uint32_t *d_src_out, uint32_t *d_dst_out, void *stream);
/* Per-vertex incident-triangle counts in the new id space (uint32, as
 * <G>_degord.4B.vout of example_apps/trianglecounting.cpp:72) and the uint64
 * total (:476-477).  *d_overflow != 0 when a count does not fit 32 bits. */
int gchi_tc_count_workspace(uint32_t nvertices, size_t *bytes);
int gchi_tc_count(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t *d_per_vertex,
                  uint64_t *d_total, uint32_t *d_overflow, void *d_work, size_t work_bytes, void *stream);

/* gchi_tc_count without a workspace and without any allocation.  No
 * counterpart in the reference: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Its counters live in a library-owned device array of about
 * 32 MiB per GPU (held until the process ends), so nvertices may not exceed
 * GCHI_TC_NOALLOC_MAX_VERTICES (GCHI_ERR_INVALID otherwise).  Every call on one
 * GPU shares that array: calls that overlap in time on one GPU (other streams,
 * other threads) race and must be serialised by the caller.  Results equal
 * gchi_tc_count's. */
#define GCHI_TC_NOALLOC_MAX_VERTICES (1u << 22)
int gchi_tc_count_noalloc(const uint32_t *d_off, const uint32_t *d_adj, uint32_t nvertices, uint32_t *d_per_vertex,
                          uint64_t *d_total, uint32_t *d_overflow, void *stream);

/* ------------------------------------------------------------------- sharder */
/* determine_number_of_shards "auto" (src/preprocessing/sharder.hpp:555-587). */
int gchi_shard_auto_count(uint64_t nedges, uint32_t edge_value_bytes, int membudget_mb, int *nshards);
/* Shards an edge list resident in HBM (self-loops already dropped): interval
 * cut of sharder::add/createnextshard/done (:861-927), per-shard (src,dst)
 * order of finish_shard (:640), degrees (:748-751).  The handle owns the
 * shard-major edge arrays in HBM.  nedges <= 2^32 - 2^25. */
typedef struct gchi_shards gchi_shards;
int gchi_shard_build(const uint32_t *d_src, const uint32_t *d_dst, uint64_t nedges, uint32_t max_id, int nshards,
                     gchi_shards **out, void *stream);
int gchi_shard_destroy(gchi_shards *shards);
int gchi_shard_info(const gchi_shards *shards, int *nshards_created, uint32_t *h_ends, uint64_t *h_counts);
int gchi_shard_edges(const gchi_shards *shards, int shard, const uint32_t **d_src, const uint32_t **d_dst,
                     uint64_t *count);
int gchi_shard_degrees(const gchi_shards *shards, const int32_t **d_degs);
/* The .adj byte stream of one shard (finish_shard, sharder.hpp:701-822) and its
 * .adjidx entries as (vertexid, filepos, edgecounter) uint64 triples. */
int gchi_shard_encode(const gchi_shards *shards, int shard, uint8_t *h_adj, uint64_t *adj_bytes,
                      uint64_t *h_index, uint32_t max_index, uint32_t *nindex, void *stream);
/* Writes the reference's on-disk layout (src/api/chifilenames.hpp:65-153):
 * .intervals, .numvertices, _degs.bin, .adj/.adjidx per shard and
 * uncompressed default-valued edge-data blocks + .size. */
/* The edge values of one shard, nvalues records of value_bytes, in the shard's
 * edge order: its edge-data file edata_name (filename_shard_edata,
 * src/api/chifilenames.hpp:91-105), blocks <edata_name>_blockdir_<B>/<id> of
 * B = 1 MiB rounded up to a multiple of value_bytes, each one zlib stream when
 * `compressed`, and <edata_name>.size.  gchi_edata_read is the load of
 * src/shards/memoryshard.hpp:268-300 with read_compressed
 * (src/util/ioutil.hpp:201-265); a .size or block that does not match nvalues
 * is GCHI_ERR_FORMAT.  gchi_edata_write rewrites the blocks and .size, as the
 * reference's memory_shard::commit (memoryshard.hpp:160-210, write_compressed
 * :144-199).  Host memory only; no device work. */
int gchi_edata_read(const char *edata_name, int compressed, uint32_t value_bytes, uint64_t nvalues, void *host_out);
int gchi_edata_write(const char *edata_name, int compressed, uint32_t value_bytes, uint64_t nvalues,
                     const void *host_in);
int gchi_shard_write_files(const gchi_shards *shards, const char *basefilename, uint32_t edge_value_bytes,
                           void *stream);
/* As gchi_shard_write_files; flags GCHI_SHARD_ZLIB writes the edge-data blocks
 * zlib-compressed (deflate level 1, one stream per block) into the ".edata..Z."
 * directories -- write_compressed of src/util/ioutil.hpp:144-199, what the
 * reference's sharder does unless the application defines
 * GRAPHCHI_DISABLE_COMPRESSION (connectedcomponents does not). */
#define GCHI_SHARD_ZLIB 1
int gchi_shard_write_files_ex(const gchi_shards *shards, const char *basefilename, uint32_t edge_value_bytes,
                              int flags, void *stream);
/* Decodes one shard's .adj byte stream, already in HBM, into (src, dst) pairs
 * in stream order -- the parse of memory_shard::load_vertices
 * (src/shards/memoryshard.hpp:448-497): a 0 byte and the byte nz after it skip
 * 1 + nz vertices; a count byte (0xff: a uint32 count follows) and that many
 * uint32 dst ids belong to the current vertex, which then advances by one.
 * src ids count from 0 as in the file.  Streams of less than 4 GiB.
 * *h_nedges = the stream's edge count, *h_vend (optional) = the vertex id after
 * the last record.  d_src == d_dst == NULL: count only.  A record that runs
 * past the end of the stream is GCHI_ERR_FORMAT; more than max_edges edges is
 * GCHI_ERR_INVALID.  Host-synchronous (reads the counts). */
int gchi_shard_decode_workspace(uint64_t nbytes, size_t *bytes);
int gchi_shard_decode(const uint8_t *d_adj, uint64_t nbytes, uint32_t *d_src, uint32_t *d_dst, uint64_t max_edges,
                      uint64_t *h_nedges, uint64_t *h_vend, void *d_work, size_t work_bytes, void *stream);
/* Several streams (a graph's shards) in one call: h_adj[i] (device pointers, in
 * a host array) of h_nbytes[i] bytes.  The edges land back to back in stream
 * order in d_src/d_dst, each stream's src ids counting from 0 as in its file;
 * h_nedges[i] / h_vend[i] per stream. */
int gchi_shard_decode_many_workspace(const uint64_t *h_nbytes, uint32_t nstreams, size_t *bytes);
int gchi_shard_decode_many(const uint8_t *const *h_adj, const uint64_t *h_nbytes, uint32_t nstreams, uint32_t *d_src,
                           uint32_t *d_dst, uint64_t max_edges, uint64_t *h_nedges, uint64_t *h_vend, void *d_work,
                           size_t work_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GCHI_H */
