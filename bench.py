#!/usr/bin/env python3
"""Benchmark of the GraphChi hot path on one MI355X.

Primary workload (the JSON line's `value`): synchronous FP32 PageRank,
r = 0.15 + 0.85 * sum(r_u / outdeg_u), on a directed R-MAT graph of scale 22
(edge factor 16, a/b/c/d = 0.57/0.19/0.19/0.05, seed 2, self-loops dropped,
duplicates kept) resident in HBM.  A step is one sweep (gather -> apply ->
scatter over every edge); value = edges per second per sweep.

The same line carries the other rows of the hot path under "extras":
batched personalized PageRank (1,024 queries, BF16 -> FP32) against plain FP32
one-query-at-a-time PPR, both checked against the FP64 oracle; the (dst,src)
radix sort of the 67M pairs; the device sharder; connected components on an
undirected R-MAT scale 24; triangle counting on R-MAT scale 21; many graphs of
uneven size in one launch per sweep.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--only pagerank]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "graphchi-cpp_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "PageRank edges/s/iter & % HBM peak; batched PPR queries/s at L1<=1e-6 vs FP64"
ABCD = (0.57, 0.19, 0.19, 0.05)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def shape_key(info):
    """Launch shape of a PageRank plan as used to key profiles."""
    return (f"W{info['edges_per_block']}_s{info['short_row_max']}_p{info['persistent']}"
            f"_h{info['hot_vertices']}_l{1 if info['gather_sources'] else 0}")


def load_traffic(name, key):
    """Bytes per sweep between L2 and the fabric, measured with rocprofv3 --pmc
    (FETCH_SIZE, WRITE_SIZE) for this launch shape and kept in
    profiles/<name>.json; None when that shape was not profiled."""
    p = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        shapes = json.load(f).get("shapes", {})
    entry = shapes.get(key)
    return entry.get("bytes_per_sweep") if entry else None


def timed(fn, torch):
    torch.cuda.synchronize()
    t = time.perf_counter()
    out = fn()
    torch.cuda.synchronize()
    return out, time.perf_counter() - t


def oracle_mod():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    return oracle


# ------------------------------------------------------------------ PageRank
def bench_pagerank(gchi, torch, args, src, dst, E, N):
    g = gchi.DeviceGraph(src, dst, N)
    tune = {"edges_per_block": args.edges_per_block, "short_row_max": args.short_row_max,
            "persistent": args.persistent, "hot_vertices": args.hot_vertices}
    tuned = None
    if args.autotune and not args.edges_per_block:
        best, best_ms, tried = gchi.autotune_pr(g, sweeps=4, report=True)
        tune = best
        tuned = {"best": best, "best_ms": round(best_ms, 4), "candidates": len(tried),
                 "worst_ms": round(max(t[1] for t in tried), 4)}
        log(f"autotune: best {best} {best_ms:.4f} ms/sweep over {len(tried)} candidates")
    plan = gchi.PageRankPlan(g, **tune)
    info = plan.info()
    op = gchi.pr_op(0.15, 0.85)
    log(f"graph N={N} E={E} plan={info}")
    if args.warmup > 0:
        w = gchi.PageRankExec(plan, op, 1.0, args.warmup)
        w.launch()
        torch.cuda.synchronize()
        del w
    ex = gchi.PageRankExec(plan, op, 1.0, args.steps)
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(s)
    ex.launch()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = e0.elapsed_time(e1) / args.steps
    stats = ex.stats.cpu().numpy()[:2 * args.steps].reshape(-1, 2)
    log(f"timed {args.steps} sweeps: {ms:.4f} ms/sweep (wall {wall * 1e3:.2f} ms); last residual "
        f"{stats[-1, 0]:.3e}, rank sum {stats[-1, 1]:.6e}")
    return g, plan, info, ms, tuned, stats


def bench_load(gchi, torch, src, dst, E, N, tune):
    """What a caller holding the edges in host memory pays once per graph: the
    upload of the (src, dst) arrays from pinned host memory, the CSC build and
    the sweep plan; reported beside the HBM-resident sweep rate, never as it."""
    hs = torch.empty(E, dtype=torch.int32, pin_memory=True)
    hd = torch.empty(E, dtype=torch.int32, pin_memory=True)
    hs.copy_(src[:E])
    hd.copy_(dst[:E])
    torch.cuda.synchronize()
    (ds, dd), t_up = timed(lambda: (hs.to("cuda", non_blocking=True), hd.to("cuda", non_blocking=True)), torch)
    g, t_csc = timed(lambda: gchi.DeviceGraph(ds, dd, N), torch)
    plan, t_plan = timed(lambda: gchi.PageRankPlan(g, **tune), torch)
    del plan, g, ds, dd
    # the reference's on-disk form: the graph's shards (.adj byte streams, from
    # the device sharder) decoded in HBM back into edge lists
    sh = gchi.Shards(src[:E], dst[:E], N - 1, gchi.shard_auto_count(E, 8))
    streams = [sh.encode(p)[0] for p in range(sh.nshards)]
    sh.close()
    d_streams = [torch.from_numpy(np.frombuffer(b, np.uint8).copy()).to("cuda") for b in streams]
    pairs = [(d, len(b)) for d, b in zip(d_streams, streams)]
    gchi.shard_decode_many(pairs)
    out, t_dec = timed(lambda: gchi.shard_decode_many(pairs), torch)
    nbytes = sum(len(b) for b in streams)
    assert int(out[0].numel()) == E
    del out, d_streams, pairs
    return {"upload_ms": round(t_up * 1e3, 2), "upload_GBps": round(8 * E / t_up / 1e9, 1),
            "csc_build_ms": round(t_csc * 1e3, 2), "plan_build_ms": round(t_plan * 1e3, 2),
            "shard_decode_ms": round(t_dec * 1e3, 2), "shard_decode_streams": len(streams),
            "shard_decode_bytes": nbytes, "shard_decode_GBps": round(nbytes / t_dec / 1e9, 1)}


def sweep_kernels(info):
    """Kernel names of one sweep of the plan (what rocprofv3 lists)."""
    if info["persistent"] == 4:  # rank space: no per-sweep layout refresh
        return "pr_seg_sweep_kernel + pr_seg_finalize_kernel"
    main = {0: "pr_sweep_kernel", 1: "pr_sweep_persistent_kernel", 2: "pr_sweep_wave_kernel",
            3: "pr_sweep_hotwg_kernel"}[info["persistent"]]
    ks = ([] if not info["gather_sources"] else ["pr_gather_layout_kernel"]) + [main]
    ks += ["pr_finalize_kernel"] if info["ngraphs"] == 1 else ["pr_hub_apply_kernel", "pr_graph_reduce_kernel"]
    return " + ".join(ks)


# ------------------------------------------------------------------ PPR
def bench_ppr(gchi, torch, args, g, plan, N):
    oracle = oracle_mod()
    queries = gchi.ppr_workload_queries(N, args.ppr_queries)
    info = plan.info()
    # the batched kernel reads vertex ids from the plan's edges: a vertex-layout
    # plan (the tuned block shape where it has one)
    if info["persistent"] == 4:
        vplan = gchi.PageRankPlan(g, edges_per_block=2048, persistent=1)
    elif info["gather_sources"]:
        vplan = gchi.PageRankPlan(g, edges_per_block=info["edges_per_block"], short_row_max=info["short_row_max"],
                                  persistent=1)
    else:
        vplan = plan
    b = gchi.PprBatch(vplan, len(queries))
    b.run(queries[:1] * len(queries), max_sweeps=8)  # warm-up (seed tables, caches)
    (info, resid), t_batch = timed(lambda: b.run(queries), torch)
    log(f"ppr batched: {len(queries)} queries in {t_batch:.3f}s, sweeps bf16={info['sweeps_bf16']} "
        f"fp32={info['sweeps_fp32']} max residual {info['max_residual']:.2e}")
    plain = gchi.PlainPpr(vplan)
    plain.run(queries[0], max_sweeps=8)
    nplain = min(args.ppr_plain_queries, len(queries))
    sweeps_plain = []

    def run_plain():
        for q in queries[:nplain]:
            _, k = plain.run(q)
            sweeps_plain.append(k)

    _, t_plain = timed(run_plain, torch)
    log(f"ppr plain: {nplain} queries in {t_plain:.3f}s (sweeps {min(sweeps_plain)}..{max(sweeps_plain)})")
    # accuracy vs the FP64 oracle on a sample of queries (CPU)
    hs, hd = gchi.to_host_u32(g.in_src, g.E), None
    in_off = gchi.to_host_u32(g.in_off, N + 1)
    outdeg = gchi.to_host_u32(g.outdeg, N)
    errs_b, errs_p = [], []
    for qi in range(min(args.ppr_check, len(queries))):
        q = len(queries) - 1 - qi if qi % 2 else qi
        r64, _ = oracle.ppr_f64(N, in_off, hs, outdeg, queries[q], 0.85, tol=1e-11)
        rb = b.ranks(q).cpu().numpy().astype(np.float64)
        rp, _ = plain.run(queries[q])
        rp = rp.cpu().numpy().astype(np.float64)
        errs_b.append((float(np.abs(rb - r64).sum()), float(np.abs(rb - r64).max())))
        errs_p.append((float(np.abs(rp - r64).sum()), float(np.abs(rp - r64).max())))
    b.close()
    qps_b, qps_p = len(queries) / t_batch, nplain / t_plain
    return {"workload": f"{len(queries)} PPR queries, restart sets of 1-64 uniform seeds, damping 0.85, "
                        "scale-22 graph; tol L1<=1e-6 (residual bound)",
            "batched_queries_per_s": round(qps_b, 2), "plain_fp32_queries_per_s": round(qps_p, 2),
            "speedup": round(qps_b / qps_p, 2), "batched_sweeps_bf16": info["sweeps_bf16"],
            "batched_sweeps_fp32": info["sweeps_fp32"], "batched_block_sweeps_fp32": info["block_sweeps_fp32"],
            "query_blocks": (len(queries) + 63) // 64, "plain_sweeps_mean": float(np.mean(sweeps_plain)),
            "plain_queries_timed": nplain,
            "l1_vs_fp64_batched_max": max(e[0] for e in errs_b) if errs_b else None,
            "maxabs_vs_fp64_batched_max": max(e[1] for e in errs_b) if errs_b else None,
            "l1_vs_fp64_plain_max": max(e[0] for e in errs_p) if errs_p else None,
            "maxabs_vs_fp64_plain_max": max(e[1] for e in errs_p) if errs_p else None,
            "queries_checked_vs_fp64": len(errs_b)}


# ------------------------------------------------------------------ sort / sharder
def bench_sort_and_shards(gchi, torch, src, dst, E, maxid):
    idb = max(1, int(maxid).bit_length())
    gchi.sort_pairs_dst_src(src, dst, idb)  # warm-up
    reps = 5

    def sort_reps():  # outputs dropped each time: their memory is reused, not allocated anew
        for _ in range(reps):
            gchi.sort_pairs_dst_src(src, dst, idb)

    _, t = timed(sort_reps, torch)
    t /= reps
    passes = -(-2 * idb // 8)
    algo = 16 * E * passes
    nsh = gchi.shard_auto_count(E, 8)
    sh, t_sh = timed(lambda: gchi.Shards(src, dst, maxid, nsh), torch)
    out = {"sort_workload": f"{E} (src,dst) uint32 pairs by (dst,src), {2 * idb}-bit keys, {passes} onesweep passes",
           "sort_pairs_per_s": round(E / t, 1), "sort_ms": round(t * 1e3, 3),
           "sort_algorithmic_GBps": round(algo / t / 1e9, 1), "sort_hbm_frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4),
           "shard_build_ms": round(t_sh * 1e3, 2), "shards": int(sh.nshards)}
    sh.close()
    return out


# ------------------------------------------------------------------ CC
def bench_cc(gchi, torch, scale):
    s, d, cnt, mx = gchi.rmat_device(scale, 8 << scale, 3)
    s2 = torch.cat([s, d])
    d2 = torch.cat([d, s])
    del s, d
    n = mx + 1
    e = int(s2.numel())
    gchi.cc_minlabel(s2, d2, n, symmetric=True)
    lab, t = timed(lambda: gchi.cc_minlabel(s2, d2, n, symmetric=True), torch)
    ncomp = int(torch.unique(lab[:n]).numel())
    lab_all, t_all = timed(lambda: gchi.cc_minlabel(s2, d2, n), torch)
    assert torch.equal(lab_all[:n], lab[:n])
    _, t_lp = timed(lambda: gchi.cc_labelprop(s2, d2, n), torch)
    _, sweeps = gchi.cc_labelprop(s2, d2, n)
    # forms over the symmetric adjacency (CSR built once, outside the timing):
    # selective scheduling, and union-find with subgraph sampling
    g = gchi.DeviceGraph(s2, d2, n)
    del s2, d2
    gchi.cc_scheduled(g.in_off, g.in_src, n)
    (lab_s, rounds, evals), t_s = timed(lambda: gchi.cc_scheduled(g.in_off, g.in_src, n), torch)
    assert torch.equal(lab_s[:n], lab[:n])
    gchi.cc_afforest(g.in_off, g.in_src, n)
    lab_a, t_a = timed(lambda: gchi.cc_afforest(g.in_off, g.in_src, n), torch)
    assert torch.equal(lab_a[:n], lab[:n])
    (lab_p, psweeps), t_p = timed(lambda: gchi.cc_pull(g.in_off, g.in_src, n), torch)
    assert torch.equal(lab_p[:n], lab[:n])
    del g
    algo = 8 * e + 12 * n + 4  # one synchronous sweep's bytes: the workload's algorithmic unit
    return {"workload": f"undirected R-MAT scale {scale}, {e} directed edges (both directions), seed 3",
            "time_to_fixed_point_ms": round(t * 1e3, 3), "edges_per_s": round(e / t, 1), "components": ncomp,
            "algorithmic_bytes_per_sweep": algo,
            "unionfind_hbm_frac": round(algo / t / 1e9 / HBM_PEAK_GBS, 4),
            "afforest_ms": round(t_a * 1e3, 3), "afforest_hbm_frac": round(algo / t_a / 1e9 / HBM_PEAK_GBS, 4),
            "pull_sweeps": psweeps, "pull_ms_per_sweep": round(t_p * 1e3 / psweeps, 3),
            "pull_sweep_hbm_frac": round(algo * psweeps / t_p / 1e9 / HBM_PEAK_GBS, 4),
            "symmetric_half_list": True, "time_all_entries_ms": round(t_all * 1e3, 3),
            "labelprop_sweeps": sweeps, "labelprop_time_ms": round(t_lp * 1e3, 2),
            "labelprop_edges_sweeps_per_s": round(e * sweeps / t_lp, 1),
            "scheduled_rounds": rounds, "scheduled_evaluations": evals, "scheduled_time_ms": round(t_s * 1e3, 2)}


# ------------------------------------------------------------------ TC
def bench_tc(gchi, torch, scale):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    hs, hd = gchi.rmat_host(scale, 16 << scale, 4)
    keep = hs != hd
    a, b = np.minimum(hs[keep], hd[keep]), np.maximum(hs[keep], hd[keep])
    key = np.unique((a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64))
    s = gchi.to_dev_u32((key >> np.uint64(32)).astype(np.uint32))
    d = gchi.to_dev_u32((key & np.uint64(0xffffffff)).astype(np.uint32))
    n = int(max(int(key.max() >> np.uint64(32)), int((key & np.uint64(0xffffffff)).max()))) + 1
    prep, t_prep = timed(lambda: gchi.TrianglePrep(s, d, n), torch)
    prep.count()
    (pv, tot, ovf), t = timed(prep.count, torch)
    off = gchi.to_host_u32(prep.off, n + 1).astype(np.int64)
    adj = gchi.to_host_u32(prep.adj, prep.noriented).astype(np.int64)
    deg = np.diff(off)
    src_of = np.repeat(np.arange(n), deg)
    algo = int((4 * (deg[src_of] + deg[adj])).sum() + 4 * prep.noriented)
    total = int(tot.item())
    return {"workload": f"simple undirected R-MAT scale {scale}, {len(key)} edges, seed 4",
            "triangles": total, "count_ms": round(t * 1e3, 3), "prepare_ms": round(t_prep * 1e3, 2),
            "triangles_per_s": round(total / t, 1), "oriented_edges_per_s": round(prep.noriented / t, 1),
            "algorithmic_GBps": round(algo / t / 1e9, 1), "overflow": int(ovf.item())}


# ------------------------------------------------------------------ generic engine
def bench_generic(gchi, scale=17, ef=16, seed=11, niters=3):
    """A program's own update() on the device (apps/bin/psw_probe, built with
    hipcc): the app's execute-updates timer over niters iterations in the
    reference engine's update order, on an R-MAT edge list it shards itself."""
    import re
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "graphchi-cpp_amd", "apps", "bin", "psw_probe")
    if not os.path.exists(exe):
        return {"skipped": "apps/bin/psw_probe not built"}
    s, d = gchi.rmat_host(scale, ef << scale, seed)
    keep = s != d
    s, d = s[keep], d[keep]
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "g")
        with open(path, "w") as f:
            f.write("".join(f"{a}\t{b}\n" for a, b in zip(s.tolist(), d.tolist())))
        p = subprocess.run([exe, "file", path, "niters", str(niters)], cwd=tmp, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, stdin=subprocess.DEVNULL, timeout=300)
        out = p.stdout.decode(errors="replace")
    m = re.search(r"execute-updates:\s*([0-9.]+) s", out)
    if p.returncode != 0 or not m:
        return {"failed": out[-500:]}
    return {"workload": f"apps/psw_probe on R-MAT scale {scale}, edge factor {ef}, seed {seed}, {int(keep.sum())} "
                        f"edges, {niters} iterations", "execute_updates_s": float(m.group(1))}


# ------------------------------------------------------------------ many graphs
def bench_multigraph(gchi, torch, ngraphs, sweeps, seed=9):
    """Many graphs of uneven size in one plan: each graph's CSC is built on its
    own (local ids), the batch plan copies them side by side (64-bit offsets
    across graphs, so the total may exceed 2^32 edges), and every sweep of all
    graphs is one launch of the persistent sweep kernel plus the finalize pair."""
    graphs, ns = [], []
    for scale, ne, gseed in gchi.multigraph_workload(ngraphs, seed):
        s, d, cnt, mx = gchi.rmat_device(scale, ne, gseed)
        graphs.append(gchi.DeviceGraph(s, d, mx + 1))
        ns.append(mx + 1)
        del s, d
    E_total = sum(g.E for g in graphs)
    plan, t_plan = timed(lambda: gchi.PageRankPlan.batch(graphs, persistent=1), torch)
    del graphs
    torch.cuda.empty_cache()
    tele = torch.tensor([0.15 / n for n in ns], dtype=torch.float64, device="cuda")
    op = gchi.pr_op(0.15, 0.85, graph_teleport=tele)
    ex = gchi.PageRankExec(plan, op, 1.0, sweeps)
    ex.launch()
    _, t = timed(ex.launch, torch)
    info = plan.info()
    return {"workload": f"{ngraphs} directed R-MAT graphs, scales uniform 10-20, edge factor uniform 8-32, "
                        f"{E_total} edges, {info['nvertices']} vertices in one batch plan; kernel 0.15/N_g, "
                        f"{sweeps} sweeps (init + sweeps replayed as one HIP graph)",
            "launches_per_sweep": len(sweep_kernels(info).split(" + ")),
            "sweep_kernels": sweep_kernels(info), "plan_build_s": round(t_plan, 3),
            "ms_per_sweep_all_graphs": round(t * 1e3 / sweeps, 3), "edges_per_s": round(E_total * sweeps / t, 1),
            "graphs_per_s": round(ngraphs * sweeps / t, 1)}


def host_cpus():
    """Host CPUs this process may use: the affinity mask, capped by a cgroup
    CPU quota when one is set; with the machine's count and CPU model."""
    avail = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, period = f.read().split()[:2]
            if q != "max":
                quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = avail if quota is None else max(1, min(avail, int(quota)))
    return {"usable": usable, "affinity": avail, "nproc": os.cpu_count(), "cgroup_quota_cpus": quota,
            "model": model}


def cpu_baseline(hs, hd, n, sweeps, nshards):
    oracle = oracle_mod()
    cpus = host_cpus()
    threads = cpus["usable"]
    t = time.time()
    _, secs = oracle.pr_psw_omp(n, hs, hd, nshards, 0.15, 0.85, 1.0, sweeps, threads)
    log(f"cpu baseline: {sweeps} sweeps in {secs:.2f}s on {threads} threads "
        f"(+{time.time() - t - secs:.1f}s preprocessing); host {cpus}")
    return {"value": len(hs) * sweeps / secs, "unit": "edges/s/iter", "cores": threads, "kind": "port",
            "sample": f"same scale-22 graph, {sweeps} sweeps of the oracle's OpenMP restatement of the "
                      f"reference PSW sweep (shards preloaded, {nshards} dst-interval shards), one thread per "
                      f"usable host CPU",
            "host": cpus}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--cpu-sweeps", type=int, default=40)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--edges-per-block", type=int, default=0)
    ap.add_argument("--short-row-max", type=int, default=0)
    ap.add_argument("--persistent", type=int, default=0)
    ap.add_argument("--hot-vertices", type=int, default=0)
    ap.add_argument("--no-autotune", dest="autotune", action="store_false")
    ap.add_argument("--only", default="",
                    help="comma list of sections: pagerank,load,ppr,sort,cc,tc,multigraph,generic")
    ap.add_argument("--ppr-queries", type=int, default=1024)
    ap.add_argument("--ppr-plain-queries", type=int, default=64)
    ap.add_argument("--ppr-check", type=int, default=8)
    ap.add_argument("--cc-scale", type=int, default=24)
    ap.add_argument("--tc-scale", type=int, default=21)
    ap.add_argument("--graphs", type=int, default=2048)
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run a device copy of this many MiB (counter calibration for HBM bytes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the package has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)
    only = set(x for x in args.only.split(",") if x)
    want = (lambda k: not only or k in only)

    import torch
    import gchi
    torch.cuda.set_device(0)
    gchi.open_device(0)
    if args.calibrate_copy:
        a = torch.empty(args.calibrate_copy << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + written per launch, 3 launches")
        del a, b
    scale, ef, seed = args.scale, 16, 2
    log(f"generating R-MAT scale {scale} on device")
    src, dst, E, maxid = gchi.rmat_device(scale, ef << scale, seed)
    N = maxid + 1
    g, plan, info, ms, tuned, stats = bench_pagerank(gchi, torch, args, src, dst, E, N)

    algo_bytes = 8 * E + 20 * N + 4
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": load_traffic("pr_sweep_rmat22_traffic", shape_key(info)), "shape": shape_key(info),
                "kernel": sweep_kernels(info) + ", one sweep (all launches)", "algorithmic_bytes": algo_bytes}
    out = {
        "metric": METRIC, "value": E / (ms * 1e-3), "unit": "edges/s/iter", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": f"synthetic R-MAT scale {scale}, edge factor {ef}, a/b/c/d 0.57/0.19/0.19/0.05, seed {seed}, "
                "self-loops dropped, duplicates kept (device generator)",
        "config": {"workload": f"pagerank_sync_fp32_rmat{scale}", "scale": scale, "edge_factor": ef, "seed": seed,
                   "nvertices": N, "nedges": E, "formula": "0.15 + 0.85*sum(r/outdeg)", "parallelism": "none (1 GPU)",
                   "edges_per_block": info["edges_per_block"], "short_row_max": info["short_row_max"],
                   "persistent": info["persistent"], "autotune": tuned},
        "roofline": roofline,
    }
    extras = {}
    tune_used = {k: info[k] for k in ("edges_per_block", "short_row_max", "persistent")}
    tune_used["hot_vertices"] = (tuned["best"]["hot_vertices"] if tuned else args.hot_vertices)
    sections = [("load", lambda: bench_load(gchi, torch, src, dst, E, N, tune_used)),
                ("ppr", lambda: bench_ppr(gchi, torch, args, g, plan, N)),
                ("sort", lambda: bench_sort_and_shards(gchi, torch, src, dst, E, maxid)),
                ("cc", lambda: bench_cc(gchi, torch, args.cc_scale)),
                ("tc", lambda: bench_tc(gchi, torch, args.tc_scale)),
                ("multigraph", lambda: bench_multigraph(gchi, torch, args.graphs, 10)),
                ("generic", lambda: bench_generic(gchi))]
    for name, fn in sections:
        if not want(name):
            continue
        log(f"== {name}")
        t = time.time()
        extras[name] = fn()
        log(f"   {extras[name]} ({time.time() - t:.1f}s)")
        torch.cuda.empty_cache()
    out["extras"] = extras
    if not args.no_cpu_baseline and want("pagerank"):
        hs, hd = gchi.to_host_u32(src, E), gchi.to_host_u32(dst, E)
        nshards = gchi.shard_auto_count(E, 8)  # the reference's auto rule, PairContainer<float> edges
        out["cpu_baseline"] = cpu_baseline(hs, hd, N, args.cpu_sweeps, nshards)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
