"""GPU tests of many independent graphs in one launch per sweep (disjoint-union
plan): every graph's ranks are bit-identical to sweeping it alone, with a
per-graph teleport (0.15/N_g) and per-graph FP64 statistics; the persistent
work-queue launch gives the same bits as one workgroup per tile."""
import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu
D = 1 - 0.15


def make_graphs(count, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        scale = int(rng.integers(4, 13))
        ef = int(rng.integers(8, 33))
        s, d = graphs.rmat_pairs(scale, ef, 1000 + i)
        s, d = graphs.preprocess(s, d)
        n = int(max(s.max(), d.max())) + 1
        out.append((s, d, n))
    return out


def run_single(dev, s, d, n, sweeps, tune):
    g = dev.DeviceGraph(dev.to_dev_u32(s), dev.to_dev_u32(d), n)
    plan = dev.PageRankPlan(g, **tune)
    ex = dev.PageRankExec(plan, dev.pr_op(0.15 / n, D), 1.0, sweeps)
    ex.launch()
    return ex.ranks().cpu().numpy().copy(), ex.stats.cpu().numpy().copy()


@pytest.mark.parametrize("persistent", [0, 1])
@pytest.mark.parametrize("tune", [{}, {"edges_per_block": 256, "short_row_max": 4}], ids=["W2048", "W256"])
def test_union_equals_individual_runs(dev, tune, persistent):
    import torch
    gl = make_graphs(24, 5)
    src, dst, vbase = dev.union_of(gl)
    N = int(vbase[-1])
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), N)
    plan = dev.PageRankPlan(g, persistent=persistent, vbase=vbase, **tune)
    assert plan.info()["ngraphs"] == len(gl)
    tele = torch.tensor([0.15 / n for (_, _, n) in gl], dtype=torch.float64, device="cuda")
    sweeps = 6
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, D, graph_teleport=tele), 1.0, sweeps)
    ex.launch()
    r = ex.ranks().cpu().numpy()
    st = ex.stats.cpu().numpy().reshape(sweeps, len(gl), 2)
    for i, (s, d, n) in enumerate(gl):
        ri, sti = run_single(dev, s, d, n, sweeps, tune)
        np.testing.assert_array_equal(r[vbase[i]:vbase[i + 1]], ri)
        np.testing.assert_allclose(st[:, i, :], sti.reshape(sweeps, 2), rtol=1e-12)


def test_persistent_single_graph_bitwise(dev):
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat16")
    s, d, n = graphs.make(case)
    a, _ = run_single(dev, s, d, n, 9, {"edges_per_block": 512, "short_row_max": 8})
    b, _ = run_single(dev, s, d, n, 9, {"edges_per_block": 512, "short_row_max": 8, "persistent": 1})
    np.testing.assert_array_equal(a, b)


def test_union_gather_layout_equals_individual_runs(dev):
    """A multi-graph plan in the gather layout with the LDS cache: every
    graph's ranks are still bit-identical to its solo vertex-id run."""
    import torch
    gl = make_graphs(16, 9)
    src, dst, vbase = dev.union_of(gl)
    N = int(vbase[-1])
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), N)
    plan = dev.PageRankPlan(g, vbase=vbase, hot_vertices=2048, edges_per_block=256)
    tele = torch.tensor([0.15 / n for (_, _, n) in gl], dtype=torch.float64, device="cuda")
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, D, graph_teleport=tele), 1.0, 5)
    ex.launch()
    r = ex.ranks().cpu().numpy()
    for i, (s, d, n) in enumerate(gl):
        ri, _ = run_single(dev, s, d, n, 5, {"edges_per_block": 256})
        np.testing.assert_array_equal(r[vbase[i]:vbase[i + 1]], ri)


@pytest.mark.parametrize("tune", [{}, {"edges_per_block": 256, "short_row_max": 4}, {"persistent": 1},
                                  {"edges_per_block": 256, "hot_vertices": 2048}],
                         ids=["W2048", "W256", "persistent", "wave_hot"])
def test_batch_plan_equals_individual_runs(dev, tune):
    """A batch plan built from separate per-graph CSCs (local ids, freed after
    planning) sweeps every graph to the same bits as its solo run, with
    per-graph teleport and statistics."""
    import torch
    gl = make_graphs(20, 13)
    graphs = [dev.DeviceGraph(dev.to_dev_u32(s), dev.to_dev_u32(d), n) for (s, d, n) in gl]
    plan = dev.PageRankPlan.batch(graphs, **tune)
    del graphs
    torch.cuda.synchronize()
    info = plan.info()
    assert info["ngraphs"] == len(gl) and info["nedges"] == sum(len(s) for (s, _, _) in gl)
    vb = plan.vbase.astype(np.int64)
    tele = torch.tensor([0.15 / n for (_, _, n) in gl], dtype=torch.float64, device="cuda")
    sweeps = 5
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, D, graph_teleport=tele), 1.0, sweeps)
    ex.launch()
    r = ex.ranks().cpu().numpy()
    st = ex.stats.cpu().numpy().reshape(sweeps, len(gl), 2)
    solo = {k: v for k, v in tune.items() if k != "persistent"}
    for i, (s, d, n) in enumerate(gl):
        ri, sti = run_single(dev, s, d, n, sweeps, solo)
        np.testing.assert_array_equal(r[vb[i]:vb[i + 1]], ri)
        np.testing.assert_allclose(st[:, i, :], sti.reshape(sweeps, 2), rtol=1e-12)


def test_batch_plan_rejects_out_of_graph_ids(dev):
    """A per-graph CSC whose source ids reach past its own vertex count is
    refused at plan creation (it would read another graph's ranks)."""
    import torch
    g0 = dev.DeviceGraph(dev.to_dev_u32(np.array([0, 1], np.uint32)), dev.to_dev_u32(np.array([1, 0], np.uint32)), 2)
    g1 = dev.DeviceGraph(dev.to_dev_u32(np.array([0, 1], np.uint32)), dev.to_dev_u32(np.array([1, 2], np.uint32)), 3)
    bad_src = torch.tensor([0, 5], dtype=torch.int32, device="cuda")  # id 5 >= 3 vertices of this graph
    gbad = dev.graph_from_csc(g1.in_off, bad_src, g1.outdeg, g1.indeg, 3, 2)
    with pytest.raises(dev.GchiError):
        dev.PageRankPlan.batch([g0, gbad])
    dev.PageRankPlan.batch([g0, g1])


def test_batch_plan_beyond_2p32_edges(dev):
    """More than 2^32 edges in one plan (64-bit offsets across graphs): 160
    copies of one scale-20 graph (27M edges each, 4.3G in total) in one batch;
    the first, a middle and the last graph match the solo sweep bit for bit."""
    import torch
    s, d, cnt, mx = dev.rmat_device(20, 26 << 20, 77)
    n = mx + 1
    g = dev.DeviceGraph(s, d, n)
    del s, d
    copies = 160
    assert copies * cnt > (1 << 32)
    plan = dev.PageRankPlan.batch([g] * copies)
    info = plan.info()
    assert info["nedges"] == copies * cnt and info["ngraphs"] == copies
    sweeps = 3
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, D), 1.0, sweeps)
    ex.launch()
    solo = dev.PageRankPlan(g)
    ex1 = dev.PageRankExec(solo, dev.pr_op(0.15, D), 1.0, sweeps)
    ex1.launch()
    r1 = ex1.ranks().cpu().numpy()
    r = ex.ranks()
    for k in (0, copies // 2, copies - 1):
        np.testing.assert_array_equal(r[k * n:(k + 1) * n].cpu().numpy(), r1)


def test_batch_and_union_with_edgeless_graphs(dev):
    """Graphs without edges inside a batch or union plan: their ranks are the
    per-graph constant, and their neighbours are swept as if alone."""
    import torch
    gl = make_graphs(6, 21)
    n_empty = 7
    empty = torch.empty(0, dtype=torch.int32, device="cuda")
    devs = []
    for i, (s, d, n) in enumerate(gl):
        devs.append(dev.DeviceGraph(dev.to_dev_u32(s), dev.to_dev_u32(d), n))
        if i % 2 == 0:
            devs.append(dev.DeviceGraph(empty, empty, n_empty))
    sizes = [g.N for g in devs]
    plan = dev.PageRankPlan.batch(devs)
    tele = torch.tensor([0.15 / n for n in sizes], dtype=torch.float64, device="cuda")
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, D, graph_teleport=tele), 1.0, 4)
    ex.launch()
    r = ex.ranks().cpu().numpy()
    vb = plan.vbase.astype(np.int64)
    k = 0
    for i, g in enumerate(devs):
        part = r[vb[i]:vb[i + 1]]
        if g.E == 0:
            np.testing.assert_array_equal(part, np.full(g.N, np.float32(0.15 / g.N)))
        else:
            s, d, n = gl[k]
            ri, _ = run_single(dev, s, d, n, 4, {})
            np.testing.assert_array_equal(part, ri)
            k += 1
