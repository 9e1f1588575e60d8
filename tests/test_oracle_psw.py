"""The oracle's restatement of the reference engine's update order
(orc_psw_run), pinned without a GPU:

* the vertex values of the order-probe program (graphchi-cpp_amd/apps/
  psw_probe.cpp) equal what the reference's own engine computed for it
  (golden fixtures written by tests/golden/make_golden.py from the program
  built against the reference headers), for 1, 2 and 3 shards;
* the smoke-test program of src/tests/basic_smoketest.cpp breaks none of its
  invariants (in-edge u->v holds u + iteration - (u > v)), and a planted
  mistake in the oracle (vertices in reverse order) breaks them."""
import numpy as np
import pytest

import graphs
import oracle


@pytest.mark.parametrize("case", graphs.PSW_CASES, ids=[c["name"] for c in graphs.PSW_CASES])
def test_psw_probe_oracle_equals_reference_engine(case, golden, manifest):
    src, dst, n = graphs.make(case)
    assert manifest[case["name"]]["input_sha"] == graphs.input_sha(src, dst)
    got, _ = oracle.psw_run(n, src, dst, oracle.PSW_PROBE, case["niters"])
    np.testing.assert_array_equal(got, golden[f"{case['name']}__psw_probe"])


@pytest.mark.parametrize("case", graphs.PSW_CASES, ids=[c["name"] for c in graphs.PSW_CASES])
def test_psw_smoketest_invariant(case):
    src, dst, n = graphs.make(case)
    vd, viol = oracle.psw_run(n, src, dst, oracle.PSW_SMOKETEST, 5)
    assert viol == 0
    np.testing.assert_array_equal(vd, np.full(n, 5, np.uint32))
    _, viol_rev = oracle.psw_run(n, src, dst, oracle.PSW_SMOKETEST, 5, reverse=True)
    assert viol_rev > 0


def test_psw_probe_sees_order():
    """The probe tells the two orders apart, so its goldens pin the order."""
    src, dst, n = graphs.make(graphs.PSW_CASES[2])
    a, _ = oracle.psw_run(n, src, dst, oracle.PSW_PROBE, 3)
    b, _ = oracle.psw_run(n, src, dst, oracle.PSW_PROBE, 3, reverse=True)
    assert (a != b).mean() > 0.5


@pytest.mark.parametrize("case", graphs.PSW_CASES, ids=[c["name"] for c in graphs.PSW_CASES])
def test_psw_pagerank_oracle_equals_reference_engine(case, golden, manifest):
    """The PageRank program with values on the edges (apps/psw_pagerank.cpp):
    ranks bit-exact, the same number of iterations before the stop rule, and the
    reference's deltalog bookkeeping (one line per iteration, or one line for a
    graph of one shard, run in the engine's in-memory mode).  The reference
    logs the change from several threads into per-thread slots that can collide
    (nested OpenMP teams), so its logged change can only be lower than the
    exact sum: checked within 5% from below (up to 3.2% lost here)."""
    src, dst, n = graphs.make(case)
    m = manifest[case["name"]]
    inmem = case["nshards"] == 1
    r, iters, deltas = oracle.psw_pagerank(n, src, dst, graphs.PSW_PR_MAX_ITERS, case["pr_tol"], inmem)
    np.testing.assert_array_equal(r, golden[f"{case['name']}__psw_pagerank"])
    assert iters == m["psw_pagerank_iterations"]
    log = m["psw_pagerank_deltalog"]
    want = [log[-1]] if inmem else log
    got_deltas = deltas[-1:] if inmem else deltas
    for line, d in zip(want, got_deltas):
        assert d * (1 - 0.05) - 1e-6 <= line[3] <= d + 1e-6
