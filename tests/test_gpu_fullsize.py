"""GPU parity at the sizes the benchmark runs: the (dst,src) sort of the 67.1M
scale-22 pairs, the 1,024-query batched PPR on the scale-22 graph, and the
2,048-graph batch (scales 10-20, edge factors 8-32), each against the oracle.

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
