"""gchi_tc_count_noalloc: triangle counting with no workspace and no
allocation (its counters live in a library-owned device array).  Called
through the C ABI and checked against gchi_tc_count on the same oriented
adjacency, bit-exact, on two calls in a row with the second input smaller
(the shared array keeps the first call's counts, which must not leak into the
second); sizes beyond the array's limit are refused."""
import ctypes

import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu


def _prep(dev, scale, ef, seed):
    s, d = graphs.rmat_pairs(scale, ef, seed)
    keep = s != d
    a, b = np.minimum(s[keep], d[keep]), np.maximum(s[keep], d[keep])
    key = np.unique((a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64))
    hs = (key >> np.uint64(32)).astype(np.uint32)
    hd = (key & np.uint64(0xffffffff)).astype(np.uint32)
    n = int(max(hs.max(), hd.max())) + 1
    return dev.TrianglePrep(dev.to_dev_u32(hs), dev.to_dev_u32(hd), n), n


def _noalloc(dev, prep, n):
    import torch
    pv = dev.empty_u32(n)
    tot = torch.zeros(1, dtype=torch.int64, device="cuda")
    ovf = torch.zeros(1, dtype=torch.int32, device="cuda")
    dev.check(dev.lib().gchi_tc_count_noalloc(dev.dptr(prep.off), dev.dptr(prep.adj), n, dev.dptr(pv), dev.dptr(tot),
                                              dev.dptr(ovf), dev.stream()), "tc_count_noalloc")
    return pv, tot, ovf


def test_tc_count_noalloc_equals_workspace_form(dev):
    big, nb = _prep(dev, 16, 16, 4)
    small, ns = _prep(dev, 11, 8, 6)
    assert ns < nb
    for prep, n in ((big, nb), (small, ns)):
        pv_w, tot_w, ovf_w = prep.count()
        pv_n, tot_n, ovf_n = _noalloc(dev, prep, n)
        np.testing.assert_array_equal(dev.to_host_u32(pv_n, n), dev.to_host_u32(pv_w, n))
        assert int(tot_n.item()) == int(tot_w.item()) > 0
        assert int(ovf_n.item()) == int(ovf_w.item()) == 0
    # again on the big graph after the small one: the same counts
    pv_n, tot_n, _ = _noalloc(dev, big, nb)
    pv_w, tot_w, _ = big.count()
    np.testing.assert_array_equal(dev.to_host_u32(pv_n, nb), dev.to_host_u32(pv_w, nb))
    assert int(tot_n.item()) == int(tot_w.item())


def test_tc_count_noalloc_refuses_oversize(dev):
    L = dev.lib()
    one = ctypes.c_void_p(1)
    assert L.gchi_tc_count_noalloc(one, one, (1 << 22) + 1, one, one, one, None) == -1
    buf = ctypes.create_string_buffer(256)
    L.gchi_last_error(buf, 256)
    assert b"GCHI_TC_NOALLOC_MAX_VERTICES" in buf.value
