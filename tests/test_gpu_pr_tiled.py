"""GPU parity of the rank-space tiled PageRank sweep (launch form 4,
pr_seg.hip): vertices relabelled by out-degree, every row's in-edges packed
into 512-slot tiles, rows crossing tiles summed from per-tile partials.

Against the reference's own outputs (golden .vout), the FP32 and FP64
oracles, and the other launch forms.  Tolerance (FP32 ranks, a different but
fixed order of additions inside long rows): per vertex |r - ref| <=
2e-5 * |ref| + 1e-30 and L1(r - ref) <= 2e-6 * L1(ref); FP64 statistics of
every sweep to 1e-6 relative."""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu
DAMPING = 1 - 0.15
FORMS = [dict(persistent=4), dict(persistent=4, hot_vertices=1024), dict(persistent=4, hot_vertices=1 << 20)]
FORM_IDS = ["tiled", "tiled_hot1k", "tiled_hotmax"]


def assert_close(r, ref, rtol=2e-5, l1tol=2e-6):
    r, ref = r.astype(np.float64), ref.astype(np.float64)
    err = np.abs(r - ref)
    assert (err <= rtol * np.abs(ref) + 1e-30).all(), f"max rel {np.max(err / np.maximum(np.abs(ref), 1e-30)):.3e}"
    assert err.sum() <= l1tol * np.abs(ref).sum(), f"L1 rel {err.sum() / np.abs(ref).sum():.3e}"


def run(dev, g, tune, teleport, sweeps, init=1.0):
    plan = dev.PageRankPlan(g, **tune)
    ex = dev.PageRankExec(plan, dev.pr_op(teleport, DAMPING), init, sweeps)
    ex.launch()
    return ex.ranks().cpu().numpy().copy(), ex.stats.cpu().numpy()[:2 * sweeps].reshape(-1, 2).copy(), plan, ex


@pytest.mark.parametrize("tune", FORMS, ids=FORM_IDS)
@pytest.mark.parametrize("case", graphs.PR_CASES, ids=[c["name"] for c in graphs.PR_CASES])
def test_tiled_vs_reference_golden(dev, golden, case, tune):
    src, dst, n = graphs.make(case)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    for kind, teleport in (("pr_norm", 0.15 / n), ("pr_affine", 0.15)):
        r, stats, plan, _ = run(dev, g, tune, teleport, case["sweeps"])
        assert_close(r, golden[f"{case['name']}__{kind}"])
        _, ostats = oracle.pagerank_f32(n, in_off, in_src, outdeg, teleport, DAMPING, 1.0, case["sweeps"])
        # rank sums: |sum r - sum ref| <= L1(r - ref), bounded by the rank tolerance
        np.testing.assert_allclose(stats[:, 1], ostats[:, 1], rtol=2e-6)
        np.testing.assert_allclose(stats[:, 0], ostats[:, 0], rtol=1e-3, atol=1e-6 * ostats[0, 0])
    info = plan.info()
    assert info["persistent"] == 4 and info["regular_edges"] >= len(src)


def test_tiled_hub_rows_exec_equals_eager_and_replays(dev):
    """pr_rmat16 has rows of thousands of in-edges (tens of tiles each): the
    captured exec (rank space throughout) equals eager standalone sweeps
    (vertex order in and out every sweep) bit for bit, on every replay."""
    import torch
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat16")
    src, dst, n = graphs.make(case)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    r1, _, plan, ex = run(dev, g, dict(persistent=4, hot_vertices=4096), 0.15, 9)
    assert int(np.bincount(dst).max()) > 20 * 512
    for _ in range(3):
        _ = torch.ones(1 << 20, device="cuda").sum()  # eager work between replays
        ex.launch()
        np.testing.assert_array_equal(ex.ranks().cpu().numpy(), r1)
    x = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
    r = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
    st = torch.zeros(2, dtype=torch.float64, device="cuda")
    plan.init(1.0, x[0], r[0])
    for k in range(9):
        plan.sweep(dev.pr_op(0.15, DAMPING), x[k & 1], r[k & 1], x[(k + 1) & 1], r[(k + 1) & 1], st)
    np.testing.assert_array_equal(r[1].cpu().numpy(), r1)


def test_tiled_scale22_vs_fp64(dev):
    """The bench workload at full size (R-MAT 22, 20 sweeps of 0.15 + 0.85 s)
    with the largest LDS cache: within tolerance of the FP64 oracle."""
    s, d, cnt, mx = dev.rmat_device(22, 16 << 22, 2)
    n = mx + 1
    hs, hd = dev.to_host_u32(s, cnt), dev.to_host_u32(d, cnt)
    g = dev.DeviceGraph(s, d, n)
    r, stats, plan, _ = run(dev, g, dict(persistent=4, hot_vertices=1 << 20), 0.15, 20)
    in_off, in_src, outdeg, _ = oracle.csc(n, hs, hd)
    r64 = oracle.pagerank_f64(n, in_off, in_src, outdeg, 0.15, DAMPING, 1.0, 20)
    err = np.abs(r.astype(np.float64) - r64)
    print(f"\ntiled scale22: {plan.info()} max_abs={err.max():.3e} max_rel={(err / r64).max():.3e} "
          f"L1_rel={err.sum() / r64.sum():.3e}")
    assert (err <= 5e-5 * r64).all()
    assert err.sum() <= 5e-6 * r64.sum()


@pytest.mark.parametrize("tune", FORMS[:2], ids=FORM_IDS[:2])
def test_tiled_small_and_empty_graphs(dev, tune):
    import torch
    empty = torch.empty(0, dtype=torch.int32, device="cuda")
    for n in (1, 5, 700):
        g = dev.DeviceGraph(empty, empty, n)
        r, _, _, _ = run(dev, g, tune, 0.15 / n, 3)
        np.testing.assert_array_equal(r, np.full(n, np.float32(0.15 / n)))
    # isolated and dangling vertices, an id gap
    src = np.array([0, 1, 2, 2, 6], np.uint32)
    dst = np.array([1, 2, 0, 4, 4], np.uint32)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), 7)
    r, _, _, _ = run(dev, g, tune, 0.15 / 7, 4)
    in_off, in_src, outdeg, _ = oracle.csc(7, src, dst)
    want, _ = oracle.pagerank_f32(7, in_off, in_src, outdeg, 0.15 / 7, DAMPING, 1.0, 4)
    np.testing.assert_array_equal(r, want)  # rows of at most two in-edges: sums in the reference's order


def test_tiled_rows_across_lanes_and_tiles(dev):
    """Rows of every length from 0 to 3000 in-edges, in an order that puts row
    starts at every lane and tile offset: vs the FP32 oracle."""
    rng = np.random.default_rng(7)
    n = 4000
    deg = rng.integers(0, 40, n)
    deg[rng.choice(n, 30, replace=False)] = rng.integers(400, 3000, 30)
    dst = np.repeat(np.arange(n), deg).astype(np.uint32)
    src = rng.integers(0, n, len(dst)).astype(np.uint32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    want, _ = oracle.pagerank_f32(n, in_off, in_src, outdeg, 0.15, DAMPING, 1.0, 5)
    for tune in FORMS:
        r, _, plan, _ = run(dev, g, tune, 0.15, 5)
        assert_close(r, want)
        assert plan.info()["hub_rows"] > 30  # split rows: the long ones and rows at tile boundaries


def test_tiled_split_rows_inside_and_across_wavefront_ranges(dev):
    """More than 8,192 tiles, so each of the 4,096 sweep wavefronts takes a
    range of at least 3 tiles: rows crossing tiles inside a range are summed
    and applied by their wavefront, rows crossing a range boundary by the
    finalize (at most one per boundary).  vs the FP32 oracle and the FP64
    statistics, and the captured exec bit for bit on replays."""
    rng = np.random.default_rng(11)
    n = 250_000
    deg = rng.integers(0, 30, n)
    deg[rng.choice(n, 300, replace=False)] = rng.integers(600, 12_000, 300)
    dst = np.repeat(rng.permutation(n), deg).astype(np.uint32)
    src = rng.integers(0, n, len(dst)).astype(np.uint32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    want, ostats = oracle.pagerank_f32(n, in_off, in_src, outdeg, 0.15, DAMPING, 1.0, 6)
    for tune in FORMS:
        r, stats, plan, ex = run(dev, g, tune, 0.15, 6)
        info = plan.info()
        assert info["regular_blocks"] > 2 * 4096 and info["persistent_workgroups"] == 256
        assert 0 < info["hub_rows"] < 4096  # only rows across a range boundary go to the finalize
        assert_close(r, want)
        np.testing.assert_allclose(stats[:, 1], ostats[:, 1], rtol=2e-6)
        np.testing.assert_allclose(stats[:, 0], ostats[:, 0], rtol=1e-3, atol=1e-6 * ostats[0, 0])
        for _ in range(2):
            ex.launch()
            np.testing.assert_array_equal(ex.ranks().cpu().numpy(), r)


def test_tiled_personalized_matches_vertex_layout(dev):
    """A seed-set operator (personalized form) through the tiled sweep equals
    the workgroup-tiled form within tolerance: the seed test maps ranks back
    to vertex ids."""
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat14_p3")
    src, dst, n = graphs.make(case)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    seeds = np.array([3, 77, 1000, n - 1], np.uint32)
    bits = dev.seed_bitmap(seeds, n)
    op = dev.pr_op(0.0, DAMPING, seed_bits=bits, seed_teleport=0.15 / len(seeds))
    outs = []
    for tune in (dict(), dict(persistent=4, hot_vertices=2048)):
        plan = dev.PageRankPlan(g, **tune)
        ex = dev.PageRankExec(plan, op, 0.0, 30)
        ex.launch()
        outs.append(ex.ranks().cpu().numpy().copy())
    assert_close(outs[1], outs[0])
