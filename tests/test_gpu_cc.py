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
