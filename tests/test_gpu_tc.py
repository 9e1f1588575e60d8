"""GPU parity of triangle counting: vertexmap, per-vertex counts and total
bit-exact vs the reference's trianglecounting output (golden) and the oracle."""
import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu


def run_tc(dev, src, dst, n):
    prep = dev.TrianglePrep(dev.to_dev_u32(src), dev.to_dev_u32(dst), n)
    pv, tot, ovf = prep.count()
    return prep, dev.to_host_u32(pv, n), int(tot.item()), int(ovf.item())


@pytest.mark.parametrize("case", graphs.TC_CASES, ids=[c["name"] for c in graphs.TC_CASES])
def test_tc_vs_reference_golden(dev, manifest, golden, case):
    src, dst, n = graphs.make(case)
    prep, pv, tot, ovf = run_tc(dev, src, dst, n)
    assert ovf == 0
    g = manifest[case["name"]]
    assert tot == g["total"]
    np.testing.assert_array_equal(dev.to_host_u32(prep.vertexmap, n), golden[f"{case['name']}__vertexmap"])
    np.testing.assert_array_equal(pv, golden[f"{case['name']}__tc_pervertex"])


def test_tc_duplicates_and_reciprocal_edges(dev):
    # the same triangle given with reversed and repeated edges counts once
    src = np.array([0, 1, 2, 1, 0, 2, 3], np.uint32)
    dst = np.array([1, 2, 0, 0, 2, 1, 3], np.uint32)
    prep, pv, tot, ovf = run_tc(dev, src, dst, 4)
    assert tot == 1 and prep.noriented == 3
    vm = dev.to_host_u32(prep.vertexmap, 4)
    for v in range(3):
        assert pv[vm[v]] == 1
    assert pv[vm[3]] == 0


def test_tc_long_lists_global_search(dev):
    # K_1100: oriented lists longer than the 1024-entry LDS stage
    n = 1100
    a, b = np.triu_indices(n, 1)
    prep, pv, tot, ovf = run_tc(dev, a.astype(np.uint32), b.astype(np.uint32), n)
    assert tot == n * (n - 1) * (n - 2) // 6
    assert (pv == (n - 1) * (n - 2) // 2).all()


def test_tc_scale21_vs_oracle(dev):
    """The TC workload: simple undirected R-MAT scale 21 (edge factor 16, seed 4)."""
    case = {"name": "tc_rmat21", "rmat": (21, 16, 4), "simple": True}
    src, dst, n = graphs.make(case)
    prep, pv, tot, ovf = run_tc(dev, src, dst, n)
    vm, opv, otot = oracle.triangles(n, src, dst)
    print(f"\nscale21 TC: N={n} E={len(src)} oriented={prep.noriented} triangles={otot}")
    assert tot == otot
    np.testing.assert_array_equal(dev.to_host_u32(prep.vertexmap, n), vm)
    np.testing.assert_array_equal(pv, opv)
