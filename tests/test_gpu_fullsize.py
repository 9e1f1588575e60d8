"""GPU parity at the sizes the benchmark runs: the (dst,src) sort of the 67.1M
scale-22 pairs, the 1,024-query batched PPR on the scale-22 graph, and the
2,048-graph batch (scales 10-20, edge factors 8-32), each against the oracle;
then sizes beyond 2^30: a sort of 2^30 + 12,345 keys and a 1.5G-edge graph in
one PageRank plan, through size-independent checks.

Sort: bit-exact against numpy.  PPR: L1 <= 1e-6 per query against the FP64
oracle (run to a 1e-11 fixed-point residual, far below the tolerance).  Many
graphs: every sampled graph within 5e-5 relative per vertex of the FP64 oracle
and 2e-5 relative of the FP32 oracle after the same 10 sweeps."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu
SCALE, EF, SEED = 22, 16, 2
D = 0.85


@pytest.fixture(scope="module")
def rmat22(dev):
    import torch
    torch.cuda.empty_cache()
    s, d, e, maxid = dev.rmat_device(SCALE, EF << SCALE, SEED)
    return s, d, e, maxid


def test_sort_pairs_full_scale22(dev, rmat22):
    """gchi_sort_pairs_dst_src on the benchmark's 67.1M pairs (44-bit keys, six
    passes): the sorted (dst, src) keys equal numpy's sort of the same keys."""
    s, d, e, maxid = rmat22
    idb = int(maxid).bit_length()
    so, do = dev.sort_pairs_dst_src(s, d, idb)
    hs, hd = dev.to_host_u32(s, e).astype(np.uint64), dev.to_host_u32(d, e).astype(np.uint64)
    want = np.sort((hd << np.uint64(idb)) | hs)
    got = (dev.to_host_u32(do, e).astype(np.uint64) << np.uint64(idb)) | dev.to_host_u32(so, e).astype(np.uint64)
    assert e > 67_000_000
    np.testing.assert_array_equal(got, want)


@pytest.fixture(scope="module")
def ppr1024(dev, rmat22):
    """The 1,024-query batch of the benchmark (restart sets of 1-64 seeds, seed 5)
    run once; the oracle's CSC of the same graph."""
    s, d, e, maxid = rmat22
    n = maxid + 1
    g = dev.DeviceGraph(s, d, n)
    plan = dev.PageRankPlan(g)
    queries = dev.ppr_workload_queries(n, 1024)
    b = dev.PprBatch(plan, len(queries))
    info, resid = b.run(queries, D, 1e-6)
    in_off, in_src, outdeg, _ = oracle.csc(n, dev.to_host_u32(s, e), dev.to_host_u32(d, e))
    return n, g, plan, queries, b, info, resid, (in_off, in_src, outdeg)


def _ppr_cases():
    # first and last blocks, both ends of each, plus one 1-seed and one 64-seed
    # query (found in the fixture's query list)
    return [0, 1, 63, 64, 511, 960, 1022, 1023, "one_seed", "64_seeds"]


@pytest.mark.parametrize("which", _ppr_cases())
def test_batched_ppr_1024_queries_vs_fp64(dev, ppr1024, which):
    n, g, plan, queries, b, info, resid, csc = ppr1024
    assert info["converged"] == 1
    if which == "one_seed":
        q = next(i for i, x in enumerate(queries) if len(x) == 1)
    elif which == "64_seeds":
        q = next(i for i, x in enumerate(queries) if len(x) == 64)
    else:
        q = which
    r64, _ = oracle.ppr_f64(n, *csc, queries[q], D, tol=1e-11)
    r = b.ranks(q).cpu().numpy().astype(np.float64)
    l1 = float(np.abs(r - r64).sum())
    mx = float(np.abs(r - r64).max())
    print(f"\nquery {q} ({len(queries[q])} seeds): L1 {l1:.3e} max-abs {mx:.3e}; "
          f"sweeps bf16={info['sweeps_bf16']} fp32={info['sweeps_fp32']}")
    assert l1 <= 1e-6, (q, l1)


@pytest.fixture(scope="module")
def multigraph2048(dev):
    import torch
    torch.cuda.empty_cache()
    spec = dev.multigraph_workload(2048)
    graphs, ns = [], []
    for scale, ne, gseed in spec:
        s, d, cnt, mx = dev.rmat_device(scale, ne, gseed)
        graphs.append(dev.DeviceGraph(s, d, mx + 1))
        ns.append(mx + 1)
        del s, d
    e_total = sum(g.E for g in graphs)
    plan = dev.PageRankPlan.batch(graphs, persistent=1)
    del graphs
    torch.cuda.empty_cache()
    sweeps = 10
    tele = torch.tensor([0.15 / n for n in ns], dtype=torch.float64, device="cuda")
    ex = dev.PageRankExec(plan, dev.pr_op(0.15, D, graph_teleport=tele), 1.0, sweeps)
    ex.launch()
    r = ex.ranks().cpu().numpy()
    yield spec, ns, plan.vbase, r, sweeps, e_total
    ex.close()
    plan.close()
    torch.cuda.empty_cache()


def _multigraph_samples():
    """Eight graphs: both ends of the batch, three inside, the largest and the
    smallest by edge count, and the last of the largest scale."""
    import gchi
    spec = gchi.multigraph_workload(2048)
    big = max(range(len(spec)), key=lambda i: spec[i][1])
    small = min(range(len(spec)), key=lambda i: spec[i][1])
    top = max(i for i in range(len(spec)) if spec[i][0] == 20)
    return sorted({0, 1, 777, 1024, 1500, 2047, big, small, top})


@pytest.mark.parametrize("gi", _multigraph_samples())
def test_multigraph_2048_vs_oracle(dev, multigraph2048, gi):
    spec, ns, vbase, r, sweeps, e_total = multigraph2048
    assert len(spec) == 2048 and e_total > 7_000_000_000
    scale, ne, gseed = spec[gi]
    hs, hd = dev.drop_self_loops_host(*dev.rmat_host(scale, ne, gseed))
    n = ns[gi]
    assert n == int(max(hs.max(), hd.max())) + 1
    in_off, in_src, outdeg, _ = oracle.csc(n, hs, hd)
    got = r[int(vbase[gi]):int(vbase[gi + 1])].astype(np.float64)
    r64 = oracle.pagerank_f64(n, in_off, in_src, outdeg, 0.15 / n, D, 1.0, sweeps)
    r32, _ = oracle.pagerank_f32(n, in_off, in_src, outdeg, 0.15 / n, D, 1.0, sweeps)
    rel64 = float((np.abs(got - r64) / np.abs(r64)).max())
    rel32 = float((np.abs(got - r32) / np.abs(r32.astype(np.float64))).max())
    print(f"\ngraph {gi}: scale {scale}, N {n}, E {len(hs)}: max rel vs FP64 {rel64:.2e}, vs FP32 oracle {rel32:.2e}, "
          f"L1 vs FP64 {np.abs(got - r64).sum():.3e}")
    assert rel64 <= 5e-5, rel64
    assert rel32 <= 2e-5, rel32


def test_sort_beyond_2p30_keys(dev):
    """gchi_sort_u64 on 2^30 + 12,345 keys (more than the 30-bit look-back
    counts of earlier versions allowed), 40-bit keys in five passes: sorted,
    and a permutation of the input (the count of every top-byte value and the
    wrapping sum of the keys unchanged).  Size-independent checks: numpy would
    take minutes."""
    import torch
    torch.cuda.empty_cache()
    n = (1 << 30) + 12345
    g = torch.Generator(device="cuda").manual_seed(7)
    keys = torch.randint(0, 1 << 40, (n,), dtype=torch.int64, device="cuda", generator=g)
    hist0 = torch.bincount(keys >> 32, minlength=256)
    sum0 = int(keys.sum().item())
    ws = dev.sort_u64(keys, key_bits=40)
    del ws
    assert bool((keys[1:] >= keys[:-1]).all())
    assert torch.equal(torch.bincount(keys >> 32, minlength=256), hist0)
    assert int(keys.sum().item()) == sum0
    del keys
    torch.cuda.empty_cache()


def test_pagerank_one_plan_beyond_2p30_edges(dev):
    """R-MAT 27 with 1.5G edges (more than 2^30) in one CSC and one form-4
    plan, one sweep from r = 1, x = 1/outdeg.  That sweep has a closed form:
    r_v = 0.15 + 0.85 * sum over in-edges u->v of 1/outdeg(u), so the rank
    sum is 0.15 N + 0.85 * (vertices with out-edges).  Checked: that sum
    against the sweep's FP64 statistics and the ranks, and 40 rows (the 8
    largest in-degrees, rows without in-edges, random rows) against an FP64
    host sum over the raw edge list."""
    import torch
    torch.cuda.empty_cache()
    s, d, e, maxid = dev.rmat_device(27, 1_500_000_000, 5)
    n = maxid + 1
    assert e > (1 << 30)
    g = dev.DeviceGraph(s, d, n)
    outdeg = dev.to_host_u32(g.outdeg).astype(np.int64)
    indeg = dev.to_host_u32(g.indeg).astype(np.int64)
    assert int(outdeg.sum()) == e and int(indeg.sum()) == e
    plan = dev.PageRankPlan(g, persistent=4, hot_vertices=1 << 20)
    x = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
    r = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
    st = torch.zeros(2, dtype=torch.float64, device="cuda")
    plan.init(1.0, x[0], r[0])
    plan.sweep(dev.pr_op(0.15, D), x[0], r[0], x[1], r[1], st)
    got = r[1].cpu().numpy().astype(np.float64)
    stats = st.cpu().numpy()
    want_sum = 0.15 * n + D * float(np.count_nonzero(outdeg))
    assert abs(stats[1] - want_sum) <= 1e-6 * want_sum, (stats[1], want_sum)
    assert abs(got.sum() - want_sum) <= 1e-6 * want_sum
    rng = np.random.default_rng(3)
    rows = np.concatenate([np.argsort(indeg)[-8:], np.flatnonzero(indeg == 0)[:8],
                           rng.integers(0, n, 24)]).astype(np.int64)
    mark = torch.zeros(n, dtype=torch.bool, device="cuda")
    mark[torch.from_numpy(rows).cuda()] = True
    sel = mark[d.long()]
    hs = s[sel].cpu().numpy().view(np.uint32).astype(np.int64)
    hd = d[sel].cpu().numpy().view(np.uint32).astype(np.int64)
    del sel, mark
    acc = np.zeros(n, np.float64)
    np.add.at(acc, hd, 1.0 / outdeg[hs])
    want = 0.15 + D * acc[rows]
    err = np.abs(got[rows] - want)
    print(f"\nR-MAT 27: N {n}, E {e}, max in-degree {indeg.max()}, {plan.info()['regular_blocks']} tiles; "
          f"sampled rows max rel {float((err / want).max()):.2e}")
    assert (err <= 2e-5 * want).all()
    del plan, g, s, d, x, r
    torch.cuda.empty_cache()
