"""The generic device engine (graphchi-cpp_amd/host/engine/psw_engine.hpp): a
program's own update() run on the device in the reference engine's order.

* gchi_psw_graph_build against numpy: stable (src, dst) out-edges, stable
  (dst, src) in-edges with the index of each in-edge's value among the
  out-edges, and the per-vertex count of lower-id incident edges;
* apps/psw_probe (hipcc; its update() folds iteration-sensitive hashes of the
  edge values) against the vertex values the reference's own engine computed
  for the same program (golden fixtures, 1-3 shards, duplicates both ways, a
  two-way path) and against the oracle on a 2M-edge R-MAT graph, bit for bit;
* apps/psw_pagerank (FP32 values on the edges, log_change, a stop rule set in
  after_iteration) against the reference engine's ranks, iteration count and
  deltalog bookkeeping, including its in-memory mode for one shard;
* apps/psw_sched (selective scheduling: add_task for this and the next
  iteration, remove_tasks, host calls from the hooks, num_tasks) against the
  reference engine's vertex values, deltalog (updates and work of every
  iteration) and printed task counts, at 1-4 shards; and a second run on the
  same shards, which reads the edge values the first wrote back;
* the reference's src/tests/basic_smoketest.cpp, compiled unchanged with hipcc
  (apps/_dropin/basic_smoketest): its update() asserts the order's invariant
  on every in-edge read, on the device, and its vertex-data check passes."""
import os
import re
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "graphchi-cpp_amd", "apps", "bin")
DROPIN = os.path.join(ROOT, "graphchi-cpp_amd", "apps", "_dropin")


@pytest.fixture
def workdir():
    d = tempfile.mkdtemp(prefix="gchi_psw_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def write_edgelist(d, name, src, dst):
    path = os.path.join(d, name)
    with open(path, "w") as f:
        f.write("".join(f"{a}\t{b}\n" for a, b in zip(src.tolist(), dst.tolist())))
    return path


def run(exe, args, cwd):
    print(f"\nrun {os.path.basename(exe)} {' '.join(str(a) for a in args)}", flush=True)
    p = subprocess.run([exe] + [str(a) for a in args], cwd=cwd, stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=300)
    out = p.stdout.decode(errors="replace")
    assert p.returncode == 0, out[-3000:]
    return out


def test_psw_graph_build_vs_numpy(dev):
    rng = np.random.default_rng(5)
    n = 5000
    src = rng.integers(0, n, 60000).astype(np.uint32)
    dst = rng.integers(0, n, 60000).astype(np.uint32)
    src[:500], dst[:500] = dst[500:1000], src[500:1000]  # reversed copies
    src[1000:1100] = dst[1000:1100]  # self-loops: in both lists, never awaited
    e = len(src)
    out_off, out_dst, in_off, in_src, in_eidx, lower = dev.psw_graph_build(dev.to_dev_u32(src), dev.to_dev_u32(dst),
                                                                           n)
    idx = np.arange(e)
    perm = np.lexsort((idx, dst, src))
    q = np.lexsort((idx, src, dst))
    pos_of = np.empty(e, np.int64)
    pos_of[perm] = np.arange(e)
    np.testing.assert_array_equal(dev.to_host_u32(out_dst), dst[perm])
    np.testing.assert_array_equal(dev.to_host_u32(in_src), src[q])
    np.testing.assert_array_equal(dev.to_host_u32(in_eidx), pos_of[q])
    np.testing.assert_array_equal(dev.to_host_u32(out_off), np.concatenate([[0], np.cumsum(np.bincount(src, minlength=n))]))
    np.testing.assert_array_equal(dev.to_host_u32(in_off), np.concatenate([[0], np.cumsum(np.bincount(dst, minlength=n))]))
    keep = src != dst
    np.testing.assert_array_equal(dev.to_host_u32(lower),
                                  np.bincount(np.maximum(src, dst)[keep], minlength=n).astype(np.uint32))


@pytest.mark.parametrize("case", graphs.PSW_CASES, ids=[c["name"] for c in graphs.PSW_CASES])
def test_psw_probe_equals_reference_engine(workdir, golden, case):
    src, dst, n = graphs.make(case)
    path = write_edgelist(workdir, case["name"], src, dst)
    run(os.path.join(BIN, "psw_probe"), ["file", path, "niters", case["niters"], "nshards", case["nshards"]], workdir)
    np.testing.assert_array_equal(np.fromfile(path + ".4B.vout", np.uint32), golden[f"{case['name']}__psw_probe"])


@pytest.mark.parametrize("case", graphs.PSW_CASES, ids=[c["name"] for c in graphs.PSW_CASES])
def test_psw_pagerank_equals_reference_engine(workdir, golden, manifest, case):
    """apps/psw_pagerank (values on the edges, log_change, a stop rule in
    after_iteration): ranks bit-exact with the reference engine's, the same
    iterations, the reference's deltalog lines (per iteration, or one line in
    its in-memory mode for one shard) with the change equal to the oracle's
    exact sum (device additions in any order: 1e-9 relative)."""
    src, dst, n = graphs.make(case)
    path = write_edgelist(workdir, case["name"], src, dst)
    out = run(os.path.join(BIN, "psw_pagerank"), ["file", path, "niters", graphs.PSW_PR_MAX_ITERS, "tol",
                                                   case["pr_tol"], "nshards", case["nshards"]], workdir)
    m = manifest[case["name"]]
    np.testing.assert_array_equal(np.fromfile(path + ".4B.vout", np.float32), golden[f"{case['name']}__psw_pagerank"])
    assert f"iterations run: {m['psw_pagerank_iterations']}" in out
    inmem = case["nshards"] == 1
    _, _, deltas = oracle.psw_pagerank(n, src, dst, graphs.PSW_PR_MAX_ITERS, case["pr_tol"], inmem)
    with open(path + ".deltalog") as f:
        got = [line.split(",") for line in f.read().split()]
    want = [m["psw_pagerank_deltalog"][-1]] if inmem else m["psw_pagerank_deltalog"]
    assert len(got) == len(want)
    for g, w, d in zip(got, want, deltas[-1:] if inmem else deltas):
        assert [int(x) for x in g[:3]] == w[:3]
        assert abs(float(g[3]) - d) <= 1e-9 * abs(d) + 1e-6


@pytest.mark.parametrize("case", graphs.PSW_CASES, ids=[c["name"] for c in graphs.PSW_CASES])
def test_psw_sched_equals_reference_engine(workdir, golden, manifest, case):
    src, dst, n = graphs.make(case)
    path = write_edgelist(workdir, case["name"], src, dst)
    out = run(os.path.join(BIN, "psw_sched"), ["file", path, "niters", graphs.PSW_SCHED_MAX_ITERS, "nshards",
                                                case["nshards"]], workdir)
    m = manifest[case["name"]]
    np.testing.assert_array_equal(np.fromfile(path + ".8B.vout", np.uint64), golden[f"{case['name']}__psw_sched"])
    tasks = [[int(a), int(b)] for a, b in re.findall(r"psw_sched: iteration (\d+) tasks (\d+)", out)]
    assert tasks == m["psw_sched_tasks"]
    with open(path + ".deltalog") as f:
        got = [[int(a), int(b), int(c), float(d)] for a, b, c, d in (line.split(",") for line in f.read().split())]
    assert got == m["psw_sched_deltalog"]
    # a second run on the same shards: with several shards it starts from the
    # edge values the first wrote back into the shard files (zlib blocks), as
    # the reference's does; in the in-memory mode of one shard nothing is
    # written back
    run(os.path.join(BIN, "psw_sched"), ["file", path, "niters", graphs.PSW_SCHED_MAX_ITERS, "nshards",
                                         case["nshards"]], workdir)
    np.testing.assert_array_equal(np.fromfile(path + ".8B.vout", np.uint64),
                                  golden[f"{case['name']}__psw_sched_rerun"])


def test_psw_probe_rmat17_vs_oracle(workdir):
    """2.1M edges, automatic shard count, 3 iterations; run twice (the second
    run reuses the shards): identical, and equal to the oracle."""
    case = {"name": "psw_rmat17", "rmat": (17, 16, 11)}
    src, dst, n = graphs.make(case)
    path = write_edgelist(workdir, case["name"], src, dst)
    out = run(os.path.join(BIN, "psw_probe"), ["file", path, "niters", 3], workdir)
    got = np.fromfile(path + ".4B.vout", np.uint32)
    print("\n".join(l for l in out.splitlines() if "execute-updates" in l or "load_shards" in l))
    run(os.path.join(BIN, "psw_probe"), ["file", path, "niters", 3], workdir)
    np.testing.assert_array_equal(np.fromfile(path + ".4B.vout", np.uint32), got)
    want, _ = oracle.psw_run(n, src, dst, oracle.PSW_PROBE, 3)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("nshards", [1, 3])
def test_psw_sched_rmat17_vs_oracle(workdir, nshards):
    """2.1M edges under the scheduler: vertex values, deltalog and task counts
    equal the oracle's, with the intervals the sharder wrote."""
    case = {"name": "psw_rmat17", "rmat": (17, 16, 11)}
    src, dst, n = graphs.make(case)
    path = write_edgelist(workdir, case["name"], src, dst)
    out = run(os.path.join(BIN, "psw_sched"), ["file", path, "niters", graphs.PSW_SCHED_MAX_ITERS, "nshards", nshards],
              workdir)
    print("\n".join(l for l in out.splitlines() if "execute-updates" in l))
    with open(f"{path}.{nshards}.intervals") as f:
        ends = [int(x) for x in f.read().split()]
    want, wlog, wtasks = oracle.psw_sched(n, src, dst, ends, nshards == 1, graphs.PSW_SCHED_MAX_ITERS)
    np.testing.assert_array_equal(np.fromfile(path + ".8B.vout", np.uint64), want)
    tasks = [int(b) for _, b in re.findall(r"psw_sched: iteration (\d+) tasks (\d+)", out)]
    assert tasks == wtasks
    with open(path + ".deltalog") as f:
        got = [[int(x) for x in line.split(",")[:3]] for line in f.read().split()]
    assert got == wlog


@pytest.mark.skipif(not os.path.exists(os.path.join(DROPIN, "basic_smoketest")), reason="drop-in not built")
@pytest.mark.parametrize("name", ["psw_misc", "psw_path", "psw_rmat12_p3", "psw_rmat14_p2"])
def test_reference_smoketest_source_on_device(workdir, name):
    case = next(c for c in graphs.PSW_CASES if c["name"] == name)
    src, dst, n = graphs.make(case)
    path = write_edgelist(workdir, name, src, dst)
    out = run(os.path.join(DROPIN, "basic_smoketest"), ["file", path, "niters", 5, "nshards", case["nshards"]],
              workdir)
    assert "Smoketest passed successfully" in out
    np.testing.assert_array_equal(np.fromfile(path + ".4B.vout", np.uint32), np.full(n, 5, np.uint32))
