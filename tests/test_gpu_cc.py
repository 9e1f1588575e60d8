"""GPU parity of connected components: labels bit-exact vs the reference's
connectedcomponents output (golden) and vs the union-find oracle at full size."""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", graphs.CC_CASES, ids=[c["name"] for c in graphs.CC_CASES])
def test_cc_vs_reference_golden(dev, golden, case):
    src, dst, n = graphs.make(case)
    ds, dd = dev.to_dev_u32(src), dev.to_dev_u32(dst)
    ref = golden[f"{case['name']}__labels"]
    np.testing.assert_array_equal(dev.to_host_u32(dev.cc_minlabel(ds, dd, n), n), ref)
    lab, sweeps = dev.cc_labelprop(ds, dd, n)
    np.testing.assert_array_equal(dev.to_host_u32(lab, n), ref)
    _, osweeps = oracle.cc_labelprop(n, src, dst)
    assert sweeps == osweeps  # synchronous propagation: same sweep count as the restatement


def test_cc_directed_input_and_isolated(dev):
    # weak connectivity over one-directional edges; vertex 6 isolated; 9 -> 0
    src = np.array([3, 5, 9, 2, 8], np.uint32)
    dst = np.array([1, 3, 0, 7, 2], np.uint32)
    n = 10
    lab = dev.to_host_u32(dev.cc_minlabel(dev.to_dev_u32(src), dev.to_dev_u32(dst), n), n)
    np.testing.assert_array_equal(lab, oracle.cc_unionfind(n, src, dst))
    assert lab[6] == 6 and lab[9] == 0 and lab[5] == 1 and lab[8] == 2


def test_cc_scale24_vs_unionfind(dev):
    """The CC workload: undirected R-MAT scale 24 (8 x 2^24 raw pairs stored in
    both directions, seed 3): labels bit-exact vs the union-find oracle."""
    src, dst = graphs.rmat_pairs(24, 16, 3, undirected=True)
    src, dst = graphs.preprocess(src, dst, undirected=True)
    n = int(max(src.max(), dst.max())) + 1
    ds, dd = dev.to_dev_u32(src), dev.to_dev_u32(dst)
    lab = dev.to_host_u32(dev.cc_minlabel(ds, dd, n), n)
    want = oracle.cc_unionfind(n, src, dst)
    np.testing.assert_array_equal(lab, want)
    # the list holds every edge both ways: the half-list union gives the same labels
    np.testing.assert_array_equal(dev.to_host_u32(dev.cc_minlabel(ds, dd, n, symmetric=True), n), want)
    print(f"\nscale24 CC: N={n} E={len(src)} components={len(np.unique(want))}")


@pytest.mark.parametrize("case", graphs.CC_CASES, ids=[c["name"] for c in graphs.CC_CASES])
def test_cc_symmetric_flag_on_symmetrised_golden(dev, golden, case):
    """The reference's CC treats edges as undirected: symmetrising a golden
    input and unioning only u < v reproduces the reference's labels."""
    src, dst, n = graphs.make(case)
    s2 = np.concatenate([src, dst]).astype(np.uint32)
    d2 = np.concatenate([dst, src]).astype(np.uint32)
    lab = dev.to_host_u32(dev.cc_minlabel(dev.to_dev_u32(s2), dev.to_dev_u32(d2), n, symmetric=True), n)
    np.testing.assert_array_equal(lab, golden[f"{case['name']}__labels"])


def _sym_csr(dev, src, dst, n):
    """CSR of the symmetrised list (each edge in both rows) on the device."""
    s2 = np.concatenate([src, dst]).astype(np.uint32)
    d2 = np.concatenate([dst, src]).astype(np.uint32)
    g = dev.DeviceGraph(dev.to_dev_u32(s2), dev.to_dev_u32(d2), n)
    return g


@pytest.mark.parametrize("case", graphs.CC_CASES, ids=[c["name"] for c in graphs.CC_CASES])
def test_cc_scheduled_vs_reference_golden(dev, golden, case):
    """Selective scheduling reaches the reference's labels (its own output)."""
    src, dst, n = graphs.make(case)
    g = _sym_csr(dev, src, dst, n)
    lab, rounds, evals = dev.cc_scheduled(g.in_off, g.in_src, n)
    np.testing.assert_array_equal(dev.to_host_u32(lab, n), golden[f"{case['name']}__labels"])
    assert rounds >= 1 and evals >= n  # round 0 evaluates every vertex


def test_cc_scheduled_long_rows_and_many_rounds(dev):
    """Rows longer than one 4096-edge chunk (a 20000-leaf star whose centre has
    the largest id), a 3000-vertex path with ids falling along it (the minimum
    starts at the far end), and isolated vertices."""
    star_c, leaves = 30000, np.arange(1, 20001, dtype=np.uint32)
    path = np.arange(23000, 20000, -1, dtype=np.uint32)  # 23000 - 22999 - ... - 20001
    src = np.concatenate([np.full(len(leaves), star_c, np.uint32), path[:-1]])
    dst = np.concatenate([leaves, path[1:]])
    n = 30005  # 0, 23001..29999 and 30001..30004 have no edges
    g = _sym_csr(dev, src, dst, n)
    lab, rounds, _ = dev.cc_scheduled(g.in_off, g.in_src, n)
    lab = dev.to_host_u32(lab, n)
    np.testing.assert_array_equal(lab, oracle.cc_unionfind(n, src, dst))
    assert lab[star_c] == 1 and lab[23000] == 20001 and lab[0] == 0 and lab[30004] == 30004
    assert rounds >= 2


def test_cc_scheduled_scale24_vs_unionfind(dev):
    """The CC workload through the scheduled form: labels bit-exact vs union-find."""
    src, dst = graphs.rmat_pairs(24, 16, 3, undirected=True)
    src, dst = graphs.preprocess(src, dst, undirected=True)
    n = int(max(src.max(), dst.max())) + 1
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)  # already both ways
    lab, rounds, evals = dev.cc_scheduled(g.in_off, g.in_src, n)
    np.testing.assert_array_equal(dev.to_host_u32(lab, n), oracle.cc_unionfind(n, src, dst))
    print(f"\nscale24 scheduled CC: rounds={rounds} evaluations={evals} (N={n})")


def test_cc_scheduled_rejects_bad_input(dev):
    off = dev.to_dev_u32(np.array([0, 1, 2], np.uint32))
    adj = dev.to_dev_u32(np.array([1, 5], np.uint32))  # 5 >= n
    with pytest.raises(dev.GchiError):
        dev.cc_scheduled(off, adj, 2)
    bad_off = dev.to_dev_u32(np.array([0, 2, 1], np.uint32))  # decreasing
    with pytest.raises(dev.GchiError):
        dev.cc_scheduled(bad_off, dev.to_dev_u32(np.array([1, 0], np.uint32)), 2)
    # a path needs more than one round: a limit of 1 is an error, not a wrong answer
    path = np.arange(100, 0, -1, dtype=np.uint32)
    g = _sym_csr(dev, path[:-1], path[1:], 101)
    with pytest.raises(dev.GchiError):
        dev.cc_scheduled(g.in_off, g.in_src, 101, max_rounds=1)


@pytest.mark.parametrize("scale,ef,seed", [(12, 16, 3), (16, 8, 5), (24, 8, 3)])
def test_cc_afforest_vs_unionfind(dev, scale, ef, seed):
    """Sampled union-find over the symmetric CSR (the giant component's edges
    skipped) gives the same labels as the oracle's union-find, up to the bench
    size (R-MAT 24, both directions, 268M entries: checked against the full
    device union-find, itself pinned to the oracle at smaller sizes)."""
    s, d, cnt, mx = dev.rmat_device(scale, ef << scale, seed)
    import torch
    s2, d2 = torch.cat([s, d]), torch.cat([d, s])
    n = mx + 1
    g = dev.DeviceGraph(s2, d2, n)
    lab = dev.to_host_u32(dev.cc_afforest(g.in_off, g.in_src, n), n)
    if scale <= 16:
        want = oracle.cc_unionfind(n, dev.to_host_u32(s, cnt), dev.to_host_u32(d, cnt))
    else:
        want = dev.to_host_u32(dev.cc_minlabel(s2, d2, n, symmetric=True), n)
    np.testing.assert_array_equal(lab, want)


def test_cc_afforest_no_giant_component(dev):
    """Many small components and isolated vertices (no component holds most
    samples): every vertex's remaining neighbours are linked."""
    rng = np.random.default_rng(3)
    n = 50000
    a = rng.integers(0, n, 30000).astype(np.uint32)
    b = ((a // 100) * 100 + rng.integers(0, 100, len(a))).astype(np.uint32)  # edges inside blocks of 100 ids
    b = np.minimum(b, n - 1)
    keep = a != b
    a, b = a[keep], b[keep]
    src, dst = np.concatenate([a, b]), np.concatenate([b, a])
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    lab = dev.to_host_u32(dev.cc_afforest(g.in_off, g.in_src, n), n)
    np.testing.assert_array_equal(lab, oracle.cc_unionfind(n, a, b))


@pytest.mark.parametrize("case", graphs.CC_CASES, ids=[c["name"] for c in graphs.CC_CASES])
def test_cc_pull_sweeps_vs_reference_golden(dev, golden, case):
    """Synchronous sweeps pulled over the symmetric CSR: the reference's labels,
    in as many sweeps as the oracle's synchronous label propagation."""
    import torch
    src, dst, n = graphs.make(case)
    a = np.concatenate([src, dst]).astype(np.uint32)
    b = np.concatenate([dst, src]).astype(np.uint32)
    g = dev.DeviceGraph(dev.to_dev_u32(a), dev.to_dev_u32(b), n)
    lab, sweeps = dev.cc_pull(g.in_off, g.in_src, n)
    np.testing.assert_array_equal(dev.to_host_u32(lab, n), golden[f"{case['name']}__labels"])
    _, osweeps = oracle.cc_labelprop(n, src, dst)
    assert sweeps == osweeps


def test_cc_pull_rows_split_across_wavefronts(dev):
    """Rows longer than one wavefront's 4096-entry range (split over many
    wavefronts, combined with atomicMin), rows that start and end inside such
    ranges, and a path that needs many sweeps: the oracle's labels in the
    oracle's number of sweeps."""
    rng = np.random.default_rng(11)
    n = 40000
    src, dst = [], []
    for hub, deg in ((5, 4097), (17, 9000), (30001, 25000), (39999, 4096), (123, 12345)):
        nb = rng.choice(n, size=deg, replace=False)
        nb = nb[nb != hub]
        src.append(np.full(len(nb), hub)), dst.append(nb)
    path = np.arange(20000, 20300)  # a path of 300 vertices: labels travel one hop per sweep
    src.append(path[1:]), dst.append(path[:-1])
    rnd = rng.integers(0, n, (2, 30000))
    src.append(rnd[0]), dst.append(rnd[1])
    src = np.concatenate(src).astype(np.uint32)
    dst = np.concatenate(dst).astype(np.uint32)
    keep = src != dst
    src, dst = src[keep], dst[keep]
    a = np.concatenate([src, dst]).astype(np.uint32)
    b = np.concatenate([dst, src]).astype(np.uint32)
    g = dev.DeviceGraph(dev.to_dev_u32(a), dev.to_dev_u32(b), n)
    assert int(np.bincount(b, minlength=n).max()) > 3 * 4096
    lab, sweeps = dev.cc_pull(g.in_off, g.in_src, n)
    want, osweeps = oracle.cc_labelprop(n, src, dst)
    np.testing.assert_array_equal(dev.to_host_u32(lab, n), want)
    assert sweeps == osweeps


def test_cc_pull_scale24_equals_unionfind(dev):
    s, d, cnt, mx = dev.rmat_device(24, 8 << 24, 3)
    import torch
    s2, d2 = torch.cat([s, d]), torch.cat([d, s])
    n = mx + 1
    want = dev.to_host_u32(dev.cc_minlabel(s2, d2, n, symmetric=True), n)
    g = dev.DeviceGraph(s2, d2, n)
    del s2, d2
    lab, sweeps = dev.cc_pull(g.in_off, g.in_src, n)
    np.testing.assert_array_equal(dev.to_host_u32(lab, n), want)
    print(f"\npull sweeps at scale 24: {sweeps}")
