#!/usr/bin/env python3
"""Benchmark of the GraphChi hot path on one MI355X.

Workload (one JSON line): synchronous FP32 PageRank, r = 0.15 + 0.85 * sum(x),
x = r / outdeg, on a directed R-MAT graph of scale 22 (edge factor 16,
a/b/c/d = 0.57/0.19/0.19/0.05, seed 2, self-loops dropped, duplicates kept),
resident in HBM.  A step is one sweep (gather -> apply -> scatter over every
edge); `value` = edges per second per sweep over the timed sweeps.

  python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--cpu-sweeps 2]
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE),
    kept in profiles/<name>.json by the profiling runs; None when absent."""
    p = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def cpu_baseline(hs, hd, n, sweeps, nshards):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = min(16, os.cpu_count() or 1)
    t = time.time()
    _, secs = oracle.pr_psw_omp(n, hs, hd, nshards, 0.15, 0.85, 1.0, sweeps, threads)
    log(f"cpu baseline: {sweeps} sweeps in {secs:.2f}s (+{time.time() - t - secs:.1f}s preprocessing)")
    return {"value": len(hs) * sweeps / secs, "unit": "edges/s/iter", "cores": threads, "kind": "port",
            "sample": f"same scale-22 graph, {sweeps} sweeps of the oracle's OpenMP restatement of the "
                      f"reference PSW sweep (shards preloaded, {nshards} dst-interval shards)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--cpu-sweeps", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--edges-per-block", type=int, default=0)
    ap.add_argument("--short-row-max", type=int, default=0)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import gchi
    torch.cuda.set_device(0)
    gchi.open_device(0)
    scale, ef, seed = args.scale, 16, 2
    log(f"generating R-MAT scale {scale} on device")
    src, dst, E, maxid = gchi.rmat_device(scale, ef << scale, seed)
    N = maxid + 1
    g = gchi.DeviceGraph(src, dst, N)
    plan = gchi.PageRankPlan(g, args.edges_per_block, args.short_row_max)
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

    algo_bytes = 8 * E + 20 * N + 4
    achieved = algo_bytes / (ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": load_traffic("pr_sweep_rmat22_traffic"),
                "kernel": "pr_sweep_kernel + pr_finalize_kernel (one sweep)", "algorithmic_bytes": algo_bytes}
    out = {
        "metric": METRIC, "value": E / (ms * 1e-3), "unit": "edges/s/iter", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 5), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "data": f"synthetic R-MAT scale {scale}, edge factor {ef}, a/b/c/d 0.57/0.19/0.19/0.05, seed {seed}, "
                "self-loops dropped, duplicates kept (device generator)",
        "config": {"workload": f"pagerank_sync_fp32_rmat{scale}", "scale": scale, "edge_factor": ef, "seed": seed,
                   "nvertices": N, "nedges": E, "formula": "0.15 + 0.85*sum(r/outdeg)", "parallelism": "none (1 GPU)",
                   "edges_per_block": info["edges_per_block"], "short_row_max": info["short_row_max"]},
        "roofline": roofline,
    }
    if not args.no_cpu_baseline:
        hs, hd = gchi.to_host_u32(src, E), gchi.to_host_u32(dst, E)
        nshards = 6  # auto rule of the reference sharder for E*8 bytes at membudget 800 MB
        out["cpu_baseline"] = cpu_baseline(hs, hd, N, args.cpu_sweeps, nshards)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
