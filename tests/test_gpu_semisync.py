"""GPU parity of the semi-synchronous sweeps (gchi_pr_blocked_*), the default
mode of the reference's pagerank_functional: against the reference's own
outputs for that mode (golden .vout at two memory budgets, hence 1-19 blocks),
against the oracle's restatement on larger graphs with many blocks, and
against the synchronous sweep when there is one block.

Tolerance as for the synchronous sweep (FP32 ranks): per vertex
|r - ref| <= 2e-5 * |ref| + 1e-30 and L1(r - ref) <= 2e-6 * L1(ref)."""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu
DAMPING = 1 - 0.15
RTOL, L1TOL = 2e-5, 2e-6
TUNES = [dict(), dict(edges_per_block=256, short_row_max=4), dict(edges_per_block=256, persistent=2),
         dict(persistent=1)]
TUNE_IDS = ["W2048", "W256_s4", "wave", "persistent"]


def assert_close(r, ref, rtol=RTOL, l1tol=L1TOL):
    r, ref = r.astype(np.float64), ref.astype(np.float64)
    err = np.abs(r - ref)
    assert (err <= rtol * np.abs(ref) + 1e-30).all(), f"max rel {np.max(err / np.maximum(np.abs(ref), 1e-30)):.3e}"
    assert err.sum() <= l1tol * np.abs(ref).sum(), f"L1 rel {err.sum() / np.abs(ref).sum():.3e}"


@pytest.mark.parametrize("tune", TUNES, ids=TUNE_IDS)
@pytest.mark.parametrize("name", ["pr_rmat14_p3", "pr_rmat16"])
@pytest.mark.parametrize("membudget_mb,key", [(800, "pr_semisync"), (1, "pr_semisync_mb1")], ids=["mb800", "mb1"])
def test_semisync_vs_reference_golden(dev, golden, manifest, name, membudget_mb, key, tune):
    case = next(c for c in graphs.PR_CASES if c["name"] == name)
    src, dst, n = graphs.make(case)
    bstart = oracle.semisync_blocks(n, manifest[name]["shards"]["intervals"], membudget_mb)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    b = dev.PageRankBlocked(g, bstart, **tune)
    r, stats = b.run(dev.pr_op(0.15 / n, DAMPING), 1.0, case["sweeps"])
    assert_close(r.cpu().numpy(), golden[f"{name}__{key}"])
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    _, ostats = oracle.pagerank_semisync_f32(n, in_off, in_src, outdeg, bstart, 0.15 / n, DAMPING, 1.0,
                                             case["sweeps"])
    # rank sums: |sum r - sum ref| <= L1(r - ref), bounded by the rank tolerance
    np.testing.assert_allclose(stats.cpu().numpy()[:, 1], ostats[:, 1], rtol=L1TOL)


@pytest.mark.parametrize("tune", TUNES, ids=TUNE_IDS)
def test_semisync_many_blocks_vs_oracle(dev, tune):
    """R-MAT scale 18 cut into 37 uneven blocks (hub rows of the highest
    in-degree fall in several of them), 8 sweeps of the 0.15 + 0.85 s form."""
    case = {"name": "ss18", "rmat": (18, 16, 3)}
    src, dst, n = graphs.make(case)
    rng = np.random.default_rng(1)
    cuts = np.unique(rng.integers(1, n, 36))
    bstart = np.concatenate([[0], cuts, [n]]).astype(np.uint32)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    b = dev.PageRankBlocked(g, bstart, **tune)
    r, _ = b.run(dev.pr_op(0.15, DAMPING), 1.0, 8)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    want, _ = oracle.pagerank_semisync_f32(n, in_off, in_src, outdeg, bstart, 0.15, DAMPING, 1.0, 8)
    assert_close(r.cpu().numpy(), want)


def test_semisync_one_block_equals_sync_sweep(dev):
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat16")
    src, dst, n = graphs.make(case)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    b = dev.PageRankBlocked(g, [0, n])
    r1, s1 = b.run(dev.pr_op(0.15, DAMPING), 1.0, 6)
    plan = dev.PageRankPlan(g)
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, DAMPING), 1.0, 6)
    ex.launch()
    np.testing.assert_array_equal(r1.cpu().numpy(), ex.ranks().cpu().numpy())
    np.testing.assert_allclose(s1.cpu().numpy(), ex.stats.cpu().numpy()[:12].reshape(-1, 2), rtol=1e-12)


def test_semisync_personalized_seeds_across_blocks(dev):
    """Seed bits are re-based per block (blocks start at ids that are not
    multiples of 32): a PPR operator over 5 blocks equals the oracle's
    semisync restatement with the seed teleport applied by global id."""
    case = {"name": "ss14", "rmat": (14, 16, 5)}
    src, dst, n = graphs.make(case)
    bstart = np.array([0, 1001, 1002, 5555, 9999, n], np.uint32)
    seeds = np.array([0, 31, 32, 1000, 1001, 1002, 5554, 5555, n - 1], np.uint32)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    b = dev.PageRankBlocked(g, bstart)
    bits = dev.seed_bitmap(seeds, n)
    tele = 0.15 / len(seeds)
    r, _ = b.run(dev.pr_op(0.0, DAMPING, seed_bits=bits, seed_teleport=tele), 0.0, 12, seed_bits=bits,
                 seed_value=tele)
    # restatement: c_v = tele on seeds, else 0; start r = tele on seeds (the
    # restart vector); blocks in order, each gathering before it scatters
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    od = np.maximum(outdeg, 1).astype(np.float64)
    rr = np.zeros(n)
    rr[seeds] = tele
    x = np.where(outdeg > 0, rr / od, 0.0)
    c = np.zeros(n)
    c[seeds] = tele
    rows = np.repeat(np.arange(n), np.diff(in_off.astype(np.int64)))
    for _ in range(12):
        for j in range(len(bstart) - 1):
            lo, hi = int(bstart[j]), int(bstart[j + 1])
            e0, e1 = int(in_off[lo]), int(in_off[hi])
            s = np.bincount(rows[e0:e1] - lo, weights=x[in_src[e0:e1]], minlength=hi - lo)
            rr[lo:hi] = c[lo:hi] + DAMPING * s
            x[lo:hi] = np.where(outdeg[lo:hi] > 0, rr[lo:hi] / od[lo:hi], 0.0)
    assert_close(r.cpu().numpy(), rr)
