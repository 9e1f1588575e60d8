"""GPU parity of the synchronous PageRank sweep against the reference's own
outputs (golden .vout), the FP32 oracle and the FP64 oracle.

Tolerance (FP32 ranks; the reference's gather order is itself nondeterministic
across index chunks): per vertex |r - ref| <= 2e-5 * |ref| + 1e-30, and
L1(r - ref) <= 2e-6 * L1(ref).  Rows short enough for one lane per row are
summed in the reference's order, so hand graphs must match bit for bit.
"""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu

DAMPING = 1 - 0.15
RTOL, L1TOL = 2e-5, 2e-6
TUNES = [(0, 0), (256, 4), (1024, 8)]


def run_pr(dev, src, dst, n, teleport, sweeps, tune=(0, 0), init=1.0):
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    plan = dev.PageRankPlan(g, *tune)
    ex = dev.PageRankExec(plan, dev.pr_op(teleport, DAMPING), init, sweeps)
    ex.launch()
    r = ex.ranks().cpu().numpy().copy()
    stats = ex.stats.cpu().numpy()[:2 * sweeps].reshape(-1, 2).copy()
    return r, stats, (g, plan, ex)


def assert_close(r, ref, rtol=RTOL, l1tol=L1TOL):
    r = r.astype(np.float64)
    ref = ref.astype(np.float64)
    err = np.abs(r - ref)
    assert (err <= rtol * np.abs(ref) + 1e-30).all(), f"max rel {np.max(err / np.maximum(np.abs(ref), 1e-30)):.3e}"
    assert err.sum() <= l1tol * np.abs(ref).sum(), f"L1 rel {err.sum() / np.abs(ref).sum():.3e}"


@pytest.mark.parametrize("tune", TUNES, ids=[f"W{t[0]}_s{t[1]}" for t in TUNES])
@pytest.mark.parametrize("case", graphs.PR_CASES, ids=[c["name"] for c in graphs.PR_CASES])
def test_pagerank_vs_reference_golden(dev, golden, case, tune):
    src, dst, n = graphs.make(case)
    for kind, teleport in (("pr_norm", 0.15 / n), ("pr_affine", 0.15)):
        r, stats, _ = run_pr(dev, src, dst, n, teleport, case["sweeps"], tune)
        ref = golden[f"{case['name']}__{kind}"]
        max_indeg = int(np.bincount(dst, minlength=n).max())
        if "hand" in case and max_indeg <= (tune[1] or 32):
            np.testing.assert_array_equal(r, ref)  # every row summed sequentially in src order
        else:
            assert_close(r, ref)
        # FP64 statistics of the last sweep vs the oracle's (same definition)
        in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
        r32, ostats = oracle.pagerank_f32(n, in_off, in_src, outdeg, teleport, DAMPING, 1.0, case["sweeps"])
        np.testing.assert_allclose(stats[:, 1], ostats[:, 1], rtol=1e-6)
        np.testing.assert_allclose(stats[:, 0], ostats[:, 0], rtol=1e-3, atol=1e-6 * ostats[0, 0])


def test_pagerank_hub_path_and_reproducible(dev):
    # W=256 turns every vertex with in-degree >= 256 into a chunked hub row
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat16")
    src, dst, n = graphs.make(case)
    r1, _, (g, plan, ex) = run_pr(dev, src, dst, n, 0.15, 9, (256, 4))
    info = plan.info()
    assert info["hub_rows"] > 0 and info["hub_chunks"] >= info["hub_rows"]
    assert info["regular_edges"] + info["hub_edges"] == len(src)
    r2 = None
    for _ in range(3):  # graph replays, with eager work between them
        ex.launch()
        r2 = ex.ranks().cpu().numpy()
        np.testing.assert_array_equal(r1, r2)
    # eager sweeps equal the captured graph bit for bit
    import torch
    x = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
    r = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
    st = torch.zeros(2, dtype=torch.float64, device="cuda")
    plan.init(1.0, x[0], r[0])
    for k in range(9):
        plan.sweep(dev.pr_op(0.15, DAMPING), x[k & 1], r[k & 1], x[(k + 1) & 1], r[(k + 1) & 1], st)
    np.testing.assert_array_equal(r[1].cpu().numpy(), r1)


def test_pagerank_isolated_and_dangling(dev):
    # vertex 3 isolated (no edges), vertex 4 dangling (in-edges only), id gap
    src = np.array([0, 1, 2, 2, 6], np.uint32)
    dst = np.array([1, 2, 0, 4, 4], np.uint32)
    n = 7
    r, _, _ = run_pr(dev, src, dst, n, 0.15 / n, 4)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    want, _ = oracle.pagerank_f32(n, in_off, in_src, outdeg, 0.15 / n, DAMPING, 1.0, 4)
    np.testing.assert_array_equal(r, want)
    assert r[3] == np.float32(0.15 / n) and r[5] == np.float32(0.15 / n)


def test_pagerank_scale22_vs_fp64(dev):
    """The bench workload: R-MAT scale 22, edge factor 16, seed 2, 20 sweeps of
    0.15 + 0.85*s.  Integer CSC bit-exact vs the oracle at full size; ranks
    within tolerance of the FP64 oracle (errors printed)."""
    import time
    s, d, cnt, mx = dev.rmat_device(22, 16 << 22, 2)
    n = mx + 1
    hs, hd = dev.to_host_u32(s, cnt), dev.to_host_u32(d, cnt)
    g = dev.DeviceGraph(s, d, n)
    plan = dev.PageRankPlan(g)
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, DAMPING), 1.0, 20)
    ex.launch()
    r = ex.ranks().cpu().numpy().astype(np.float64)
    t = time.time()
    in_off, in_src, outdeg, _ = oracle.csc(n, hs, hd)
    np.testing.assert_array_equal(dev.to_host_u32(g.in_off, n + 1), in_off)
    np.testing.assert_array_equal(dev.to_host_u32(g.outdeg, n), outdeg)
    r64 = oracle.pagerank_f64(n, in_off, in_src, outdeg, 0.15, DAMPING, 1.0, 20)
    err = np.abs(r - r64)
    print(f"\nscale22: N={n} E={cnt} max_abs={err.max():.3e} max_rel={(err / r64).max():.3e} "
          f"L1={err.sum():.3e} L1_rel={err.sum() / r64.sum():.3e} (oracle {time.time() - t:.1f}s)")
    assert (err <= 5e-5 * r64).all()
    assert err.sum() <= 5e-6 * r64.sum()


HOT_TUNES = [dict(hot_vertices=-1), dict(hot_vertices=-1, persistent=1),
             dict(hot_vertices=1024, edges_per_block=256), dict(hot_vertices=1 << 20, edges_per_block=256),
             dict(hot_vertices=4096, edges_per_block=256, short_row_max=4),
             dict(hot_vertices=2048, edges_per_block=256, short_row_max=64),
             dict(persistent=2, edges_per_block=256), dict(persistent=2, hot_vertices=-1, edges_per_block=256),
             dict(persistent=3, hot_vertices=1 << 20), dict(persistent=3, hot_vertices=2048, edges_per_block=256,
                                                              short_row_max=4)]


@pytest.mark.parametrize("hot", HOT_TUNES, ids=["layout", "layout_pers", "hot1k", "hot_max", "hot_s4",
                                                "hot_s64", "wave", "wave_layout", "hotwg", "hotwg_W256"])
def test_gather_layout_bitwise_equal(dev, golden, hot):
    """The gather layout, its LDS cache and the wavefront-tiled launch change
    where values are read from and which wavefront sums a row, never which
    values are summed or in what order: ranks are bit-identical to one
    workgroup per tile in vertex-id layout with the same W and short-row
    limit (FP64 statistics equal to 1e-12: their grouping follows the launch)."""
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat16")
    src, dst, n = graphs.make(case)
    base = {k: v for k, v in hot.items() if k in ("edges_per_block", "short_row_max")}
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    outs = []
    for tune in (base, hot):
        plan = dev.PageRankPlan(g, **tune)
        ex = dev.PageRankExec(plan, dev.pr_op(0.15 / n, DAMPING), 1.0, 9)
        ex.launch()
        outs.append((ex.ranks().cpu().numpy().copy(), ex.stats.cpu().numpy().copy(), plan.info()))
    info = outs[1][2]
    nsrc = int(np.count_nonzero(np.bincount(src, minlength=n)))
    hv = hot.get("hot_vertices", 0)
    assert info["gather_sources"] == (nsrc if hv else 0)
    if hv > 0:
        h = info["hot_vertices"]
        assert 0 < h <= min(hv, nsrc) and h % 1024 == 0
        assert (info["persistent"], info["hot_threads"]) in ((2, 1024), (3, 256))
    else:
        assert info["hot_vertices"] == 0
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_allclose(outs[0][1], outs[1][1], rtol=1e-12)
    # and the reference's own output still holds
    assert_close(outs[1][0], golden[f"{case['name']}__pr_norm"])


def test_gather_layout_rejects_inconsistent_outdeg(dev):
    """A CSC whose out-degrees do not cover its in-edge sources cannot be put
    in the gather layout; plan creation fails instead of gathering out of bounds."""
    import torch
    src = np.array([0, 1, 2, 3], np.uint32)
    dst = np.array([1, 2, 3, 0], np.uint32)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), 4)
    bad = dev.graph_from_csc(g.in_off, g.in_src, torch.zeros_like(g.outdeg), g.indeg, 4, 4)
    with pytest.raises(dev.GchiError):
        dev.PageRankPlan(bad, hot_vertices=-1)
    dev.PageRankPlan(g, hot_vertices=-1)  # the consistent graph plans fine


def test_gather_layout_scale22_matches_vertex_layout(dev):
    """At the bench size (about 68 regular blocks per wavefront, so several
    descriptor batches each) the LDS-cached sweep gives the same bits as the
    vertex-id layout with the same block shape."""
    s, d, cnt, mx = dev.rmat_device(22, 16 << 22, 2)
    n = mx + 1
    g = dev.DeviceGraph(s, d, n)
    outs = []
    for tune in (dict(edges_per_block=256, short_row_max=32), dict(edges_per_block=256, short_row_max=32,
                                                                    hot_vertices=1 << 20)):
        plan = dev.PageRankPlan(g, **tune)
        ex = dev.PageRankExec(plan, dev.pr_op(0.15, DAMPING), 1.0, 20)
        ex.launch()
        outs.append(ex.ranks().cpu().numpy().copy())
        print("\n", tune, plan.info())
    np.testing.assert_array_equal(outs[0], outs[1])


def test_wave_sweep_long_blocks(dev):
    """Blocks of more than 64 rows (runs of rows without in-edges) and skewed
    sources, in the wavefront-tiled sweep with and without the LDS cache:
    bit-identical to one workgroup per tile.  (More blocks per wavefront than
    one descriptor batch of 64 is covered at scale 22 below.)"""
    rng = np.random.default_rng(11)
    n = 1 << 21
    e = 6 << 20
    src = (rng.zipf(1.6, e) % n).astype(np.uint32)
    dst = rng.integers(0, n // 2, e).astype(np.uint32)  # upper half: no in-edges
    keep = src != dst
    src, dst = src[keep], dst[keep]
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    outs = []
    for tune in (dict(edges_per_block=256, short_row_max=16), dict(edges_per_block=256, short_row_max=16,
                                                                    persistent=2),
                 dict(edges_per_block=256, short_row_max=16, hot_vertices=1 << 20)):
        plan = dev.PageRankPlan(g, **tune)
        ex = dev.PageRankExec(plan, dev.pr_op(0.15 / n, DAMPING), 1.0, 6)
        ex.launch()
        outs.append(ex.ranks().cpu().numpy().copy())
    np.testing.assert_array_equal(outs[0], outs[1])
    np.testing.assert_array_equal(outs[0], outs[2])


@pytest.mark.parametrize("tune", [dict(), dict(persistent=1), dict(edges_per_block=256, persistent=2),
                                  dict(edges_per_block=256, hot_vertices=1024), dict(persistent=3, hot_vertices=1024),
                                  dict(hot_vertices=-1)],
                         ids=["wg", "persistent", "wave", "wave_hot", "hotwg", "layout"])
def test_pagerank_graph_without_edges(dev, tune):
    """No edges at all: every rank is the constant term after each sweep, in
    every launch form (empty plans, no gather layout sources)."""
    import torch
    n = 5
    empty = torch.empty(0, dtype=torch.int32, device="cuda")
    g = dev.DeviceGraph(empty, empty, n)
    plan = dev.PageRankPlan(g, **tune)
    ex = dev.PageRankExec(plan, dev.pr_op(0.15 / n, DAMPING), 1.0, 3)
    ex.launch()
    r = ex.ranks().cpu().numpy()
    np.testing.assert_array_equal(r, np.full(n, np.float32(0.15 / n)))


def test_pagerank_single_vertex_and_self_loop_only(dev):
    """A lone vertex, and a graph whose only edges were self-loops (dropped by
    the generator's preprocessing; here absent from the edge list)."""
    import torch
    empty = torch.empty(0, dtype=torch.int32, device="cuda")
    g = dev.DeviceGraph(empty, empty, 1)
    ex = dev.PageRankExec(dev.PageRankPlan(g), dev.pr_op(0.15, DAMPING), 1.0, 2)
    ex.launch()
    assert ex.ranks().cpu().numpy()[0] == np.float32(0.15)
