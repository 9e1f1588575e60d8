"""GPU parity of the building blocks: device R-MAT generator, self-loop
compaction, onesweep radix sort, CSC build.  All integer work: bit-exact."""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu


def test_device_rmat_equals_host(dev):
    for scale, n, seed in ((10, 100000, 1), (22, 300000, 2), (1, 1000, 5)):
        s, d, _, _ = dev.rmat_device(scale, n, seed, drop_self_loops=False)
        hs, hd = dev.rmat_host(scale, n, seed)
        np.testing.assert_array_equal(dev.to_host_u32(s, n), hs)
        np.testing.assert_array_equal(dev.to_host_u32(d, n), hd)


def test_drop_self_loops_and_maxid(dev):
    for scale, n, seed in ((4, 5000, 3), (16, 1 << 20, 1), (12, 4097, 8)):
        s, d, cnt, mx = dev.rmat_device(scale, n, seed)
        hs, hd = dev.drop_self_loops_host(*dev.rmat_host(scale, n, seed))
        assert cnt == len(hs)
        np.testing.assert_array_equal(dev.to_host_u32(s, cnt), hs)
        np.testing.assert_array_equal(dev.to_host_u32(d, cnt), hd)
        assert mx == max(hs.max(), hd.max())


@pytest.mark.parametrize("n,bits", [(0, 8), (1, 64), (4095, 13), (4096, 8), (4097, 1), (100003, 37),
                                    (1 << 20, 44), (3 << 20, 64)])
def test_radix_sort_stable_with_payload(dev, n, bits):
    import torch
    rng = np.random.default_rng(n + bits)
    keys = rng.integers(0, np.iinfo(np.uint64).max, max(n, 1), dtype=np.uint64)[:n]
    if n > 10:
        keys[::4] = keys[::4] & np.uint64(0xff)  # duplicates across digits
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xffffffffffffffff)
    vals = np.arange(n, dtype=np.uint32)
    kt = torch.from_numpy(keys.view(np.int64).copy()).to("cuda")
    vt = torch.from_numpy(vals.view(np.int32).copy()).to("cuda")
    if n:
        dev.sort_u64(kt, vt, bits)
    order = np.argsort(keys & mask, kind="stable")
    got_k = kt.cpu().numpy().view(np.uint64)
    got_v = vt.cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got_v[:n], vals[order])
    np.testing.assert_array_equal(got_k[:n], keys[order])


def test_sort_pairs_dst_src(dev):
    s, d, cnt, mx = dev.rmat_device(16, 1 << 20, 4)
    so, do = dev.sort_pairs_dst_src(s, d, 16)
    hs, hd = dev.to_host_u32(s, cnt), dev.to_host_u32(d, cnt)
    o = np.lexsort((hs, hd))
    np.testing.assert_array_equal(dev.to_host_u32(so, cnt), hs[o])
    np.testing.assert_array_equal(dev.to_host_u32(do, cnt), hd[o])


@pytest.mark.parametrize("n,id_bits", [(1, 1), (2, 1), (1000, 3), (5000, 4), (70001, 11), (300007, 32), (8192, 13)])
def test_sort_pairs_dst_src_sizes(dev, n, id_bits):
    """Pair sorts whose first pass reads the (src, dst) arrays and whose last
    pass writes them: one- and two-digit keys (an extra all-zero pass), ragged
    tile counts, 32-bit ids; then the same sort in place (outputs = inputs)."""
    import ctypes
    rng = np.random.default_rng(n + id_bits)
    hi = (1 << id_bits) - 1
    hs = rng.integers(0, hi + 1, n, dtype=np.uint64).astype(np.uint32)
    hd = rng.integers(0, hi + 1, n, dtype=np.uint64).astype(np.uint32)
    hd[: n // 3] = hd[0]  # long runs of one destination
    o = np.lexsort((hs, hd))
    s, d = dev.to_dev_u32(hs), dev.to_dev_u32(hd)
    so, do = dev.sort_pairs_dst_src(s, d, id_bits)
    np.testing.assert_array_equal(dev.to_host_u32(so, n), hs[o])
    np.testing.assert_array_equal(dev.to_host_u32(do, n), hd[o])
    L = dev.lib()
    nb = dev.c_sz()
    dev.check(L.gchi_sort_pairs_dst_src_workspace(n, id_bits, ctypes.byref(nb)))
    ws = dev.workspace(nb.value)
    dev.check(L.gchi_sort_pairs_dst_src(dev.dptr(s), dev.dptr(d), n, id_bits, dev.dptr(s), dev.dptr(d), dev.dptr(ws),
                                        nb.value, dev.stream()), "sort_pairs in place")
    np.testing.assert_array_equal(dev.to_host_u32(s, n), hs[o])
    np.testing.assert_array_equal(dev.to_host_u32(d, n), hd[o])


@pytest.mark.parametrize("case", [c for c in graphs.PR_CASES] + graphs.CC_CASES[:1],
                         ids=[c["name"] for c in graphs.PR_CASES] + [graphs.CC_CASES[0]["name"]])
def test_csc_build_bit_exact(dev, case):
    src, dst, n = graphs.make(case)
    g = dev.DeviceGraph(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    in_off, in_src, outdeg, indeg = oracle.csc(n, src, dst)
    np.testing.assert_array_equal(dev.to_host_u32(g.in_off, n + 1), in_off)
    np.testing.assert_array_equal(dev.to_host_u32(g.in_src, len(src)), in_src)
    np.testing.assert_array_equal(dev.to_host_u32(g.outdeg, n), outdeg)
    np.testing.assert_array_equal(dev.to_host_u32(g.indeg, n), indeg)


def test_sort_lookback_timeout_is_reported(dev):
    """A look-back wait that reaches its bound must not pass silently: the
    test hook makes every wait see unpublished predecessors and gives it a
    bound of 8 polls, so each tile gives up (every kernel still exits), and the
    status word reports the failure to the asynchronous sort, the pair sort,
    the CSC build, the sharder and the triangle-counting preparation.  The
    passes after a timed-out one read buffers with unwritten holes (digits the
    histogram never counted): the sort drops every position outside [0, n),
    and the kernels that index by sorted ids skip ids outside the graph, so
    nothing is written out of bounds (checked under a host address sanitizer
    on a CPU build of the kernels)."""
    import torch
    s, d, cnt, mx = dev.rmat_device(16, 1 << 20, 4)
    keys = torch.arange(1 << 20, 0, -1, dtype=torch.int64, device="cuda")
    try:
        dev.debug_sort_lookback(8, True)
        ws = dev.sort_u64(keys, key_bits=24, check_status=False)
        with pytest.raises(dev.GchiError, match="TIMEOUT|timed out|wait"):
            dev.work_status(ws)
        with pytest.raises(dev.GchiError):
            dev.sort_pairs_dst_src(s, d, 16)
        with pytest.raises(dev.GchiError):
            dev.DeviceGraph(s, d, mx + 1)
        with pytest.raises(dev.GchiError):
            dev.Shards(s, d, mx, 2)
        with pytest.raises(dev.GchiError, match="TIMEOUT|timed out|wait"):
            dev.TrianglePrep(s, d, mx + 1)
    finally:
        dev.debug_sort_lookback(0, False)
    # the same calls succeed again once the hook is off (fresh keys: the
    # stalled sort left its output invalid)
    keys = torch.arange(1 << 20, 0, -1, dtype=torch.int64, device="cuda")
    dev.sort_u64(keys, key_bits=24)
    assert bool((keys[1:] >= keys[:-1]).all())
    so, do = dev.sort_pairs_dst_src(s, d, 16)
    hs, hd = dev.to_host_u32(s, cnt), dev.to_host_u32(d, cnt)
    o = np.lexsort((hs, hd))
    np.testing.assert_array_equal(dev.to_host_u32(so, cnt), hs[o])
