"""GPU parity of personalized PageRank: plain FP32 per-query sweeps and the
batched BF16 -> FP32 form, both against the FP64 oracle (L1 <= 1e-6 per query,
stated in the workload) and against the reference's own FP32 PPR run."""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu
D = 0.85
TOL = 1e-6


def graph_and_plan(dev, case, **tune):
    src, dst, n = graphs.make(case)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    return src, dst, n, g, dev.PageRankPlan(g, **tune)


def random_queries(n, nq, seed, max_seeds=64):
    rng = np.random.default_rng(seed)
    return [rng.choice(n, size=int(rng.integers(1, max_seeds + 1)), replace=False).astype(np.uint32)
            for _ in range(nq)]


def test_plain_ppr_vs_reference_and_fp64(dev, golden):
    case = next(c for c in graphs.PR_CASES if c.get("ppr_seeds"))
    src, dst, n, g, plan = graph_and_plan(dev, case)
    r, sweeps = dev.PlainPpr(plan, D).run(case["ppr_seeds"], TOL)
    r = r.cpu().numpy().astype(np.float64)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    r64, _ = oracle.ppr_f64(n, in_off, in_src, outdeg, case["ppr_seeds"], D)
    assert np.abs(r - r64).sum() <= TOL
    ref = golden[f"{case['name']}__ppr"].astype(np.float64)  # reference, 199 sweeps from r = 1
    assert np.abs(r - ref).sum() <= 2 * TOL


@pytest.mark.parametrize("mode", ["bf16_fp32", "fp32"])
@pytest.mark.parametrize("tune", [{}, {"edges_per_block": 256, "short_row_max": 4}], ids=["W2048", "W256"])
def test_batched_ppr_vs_fp64(dev, mode, tune):
    case = {"name": "ppr14", "rmat": (14, 16, 7)}
    src, dst, n, g, plan = graph_and_plan(dev, case, **tune)
    queries = random_queries(n, 70, 11)  # two query blocks, the second partial
    b = dev.PprBatch(plan, len(queries))
    info, resid = b.run(queries, D, TOL, bf16_sweeps=60 if mode == "bf16_fp32" else 0)
    assert info["converged"] == 1
    if mode == "bf16_fp32":
        assert info["sweeps_bf16"] > 0
    # chunks of 4 sweeps; converged query blocks stop being swept
    assert info["sweeps_fp32"] % 4 == 0 and info["sweeps_bf16"] % 4 == 0
    assert 0 < info["block_sweeps_fp32"] <= 2 * info["sweeps_fp32"]
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    worst = 0.0
    for q in (0, 1, 63, 64, 69):
        r = b.ranks(q).cpu().numpy().astype(np.float64)
        r64, _ = oracle.ppr_f64(n, in_off, in_src, outdeg, queries[q], D)
        l1 = np.abs(r - r64).sum()
        worst = max(worst, l1)
        assert l1 <= TOL, (q, l1)
    print(f"\nbatched {mode} {tune}: sweeps bf16={info['sweeps_bf16']} fp32={info['sweeps_fp32']} worst L1={worst:.2e}")


def test_batched_ppr_matches_plain(dev):
    case = {"name": "ppr12", "rmat": (12, 16, 9)}
    src, dst, n, g, plan = graph_and_plan(dev, case)
    queries = random_queries(n, 5, 3)
    b = dev.PprBatch(plan, len(queries))
    b.run(queries, D, TOL)
    plain = dev.PlainPpr(plan, D)
    for q, seeds in enumerate(queries):
        rp, _ = plain.run(seeds, TOL)
        assert np.abs(b.ranks(q).cpu().numpy().astype(np.float64) - rp.cpu().numpy()).sum() <= 2 * TOL


def test_ppr_with_gather_layout_plan(dev):
    """Plain PPR over a gather-layout plan gives the vertex-layout bits; the
    batched form needs vertex ids in the plan's edges and refuses such a plan."""
    case = next(c for c in graphs.PR_CASES if c.get("ppr_seeds"))
    src, dst, n, g, plan = graph_and_plan(dev, case, edges_per_block=256)
    hot = dev.PageRankPlan(g, hot_vertices=1024, edges_per_block=256)
    a, sa = dev.PlainPpr(plan, D).run(case["ppr_seeds"], TOL)
    b, sb = dev.PlainPpr(hot, D).run(case["ppr_seeds"], TOL)
    assert sa == sb
    np.testing.assert_array_equal(a.cpu().numpy(), b.cpu().numpy())
    with pytest.raises(dev.GchiError):
        dev.PprBatch(hot, 64)


def test_batched_ppr_retires_converged_blocks(dev):
    """Query blocks converge at different sweeps: one block of 1-seed queries on
    high-out-degree vertices, one of 64-seed queries.  Every query still meets
    the tolerance against the FP64 oracle, and the FP32 work counted per block
    is below (blocks x sweeps) when one block retires first."""
    case = {"name": "ppr14r", "rmat": (14, 16, 7)}
    src, dst, n, g, plan = graph_and_plan(dev, case)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    hubs = np.argsort(-np.diff(in_off.astype(np.int64)))[:64].astype(np.uint32)
    rng = np.random.default_rng(2)
    queries = [np.array([h], np.uint32) for h in hubs] + [rng.choice(n, 64, replace=False).astype(np.uint32)
                                                          for _ in range(64)]
    b = dev.PprBatch(plan, len(queries))
    info, resid = b.run(queries, D, TOL, bf16_sweeps=0, check_every=2)
    assert info["converged"] == 1
    assert info["block_sweeps_fp32"] <= 2 * info["sweeps_fp32"]
    for q in (0, 63, 64, 127):
        r = b.ranks(q).cpu().numpy().astype(np.float64)
        r64, _ = oracle.ppr_f64(n, in_off, in_src, outdeg, queries[q], D)
        assert np.abs(r - r64).sum() <= TOL, q
    assert (resid * D / (1 - D) <= TOL).all()
