"""CPU tests: pin the oracle (the CPU restatement) against the golden outputs of
the reference's own binaries, closed forms and brute force.  No GPU needed."""
import itertools

import numpy as np
import pytest

import graphs
import oracle

DAMPING = 1 - 0.15  # (1 - RANDOMRESETPROB) of example_apps/pagerank_functional.cpp:74


def ids(cases):
    return [c["name"] for c in cases]


@pytest.mark.parametrize("case", graphs.all_cases(), ids=ids(graphs.all_cases()))
def test_inputs_match_fixtures(case, manifest):
    src, dst, n = graphs.make(case)
    m = manifest[case["name"]]
    assert m["n"] == n and m["e"] == len(src)
    assert m["input_sha"] == graphs.input_sha(src, dst)


@pytest.mark.parametrize("case", graphs.PR_CASES, ids=ids(graphs.PR_CASES))
def test_pagerank_f32_oracle_bit_exact_vs_reference(case, golden):
    # Shards under 1M edges are gathered by one index chunk, in src order, so the
    # reference's FP32 sums have the same order as the oracle's: bit-exact.
    src, dst, n = graphs.make(case)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    for kind, teleport in (("pr_norm", 0.15 / n), ("pr_affine", 0.15)):
        r, _ = oracle.pagerank_f32(n, in_off, in_src, outdeg, teleport, DAMPING, 1.0, case["sweeps"])
        ref = golden[f"{case['name']}__{kind}"]
        assert ref.shape == (n,)
        np.testing.assert_array_equal(r, ref, err_msg=kind)


SEMISYNC = [c for c in graphs.PR_CASES if c["name"] in ("pr_rmat14_p3", "pr_rmat16")]


@pytest.mark.parametrize("case", SEMISYNC, ids=ids(SEMISYNC))
@pytest.mark.parametrize("membudget_mb,key", [(800, "pr_semisync"), (1, "pr_semisync_mb1")])
def test_pagerank_semisync_oracle_bit_exact_vs_reference(case, membudget_mb, key, golden, manifest):
    """The reference pagerank_functional's default mode, with its sub-interval
    cut at two memory budgets (3 and 7 blocks on the 3-shard graph, 1 and 19 on
    the 1-shard one): bit-exact, so the block order, the per-block gather-then-
    scatter and the window rule are all pinned; a wrong cut changes the bits."""
    src, dst, n = graphs.make(case)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    b = oracle.semisync_blocks(n, manifest[case["name"]]["shards"]["intervals"], membudget_mb)
    r, _ = oracle.pagerank_semisync_f32(n, in_off, in_src, outdeg, b, 0.15 / n, DAMPING, 1.0, case["sweeps"])
    np.testing.assert_array_equal(r, golden[f"{case['name']}__{key}"])
    # a cut shifted by one vertex changes the result on the multi-block runs
    if len(b) > 2:
        b2 = b.copy()
        b2[1] += 1
        r2, _ = oracle.pagerank_semisync_f32(n, in_off, in_src, outdeg, b2, 0.15 / n, DAMPING, 1.0, case["sweeps"])
        assert not np.array_equal(r2, r)


def test_semisync_one_block_is_sync(golden):
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat14_p3")
    src, dst, n = graphs.make(case)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    r1, s1 = oracle.pagerank_semisync_f32(n, in_off, in_src, outdeg, [0, n], 0.15, DAMPING, 1.0, 5)
    r2, s2 = oracle.pagerank_f32(n, in_off, in_src, outdeg, 0.15, DAMPING, 1.0, 5)
    np.testing.assert_array_equal(r1, r2)
    np.testing.assert_allclose(s1, s2, rtol=1e-12)


def test_semisync_window_rule():
    # maxwindow = membudget_mb * 2^20 / 300: 3495 at 1 MB; a window holds up to
    # maxwindow + 1 vertices; empty intervals are skipped
    np.testing.assert_array_equal(oracle.semisync_blocks(10000, [9999], 1), [0, 3496, 6992, 10000])
    np.testing.assert_array_equal(oracle.semisync_blocks(10, [3, 3, 9], 800), [0, 4, 10])
    np.testing.assert_array_equal(oracle.semisync_blocks(5000, [100, 4999], 1), [0, 101, 3597, 5000])


def test_pagerank_f64_tracks_f32(golden):
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat16")
    src, dst, n = graphs.make(case)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    r64 = oracle.pagerank_f64(n, in_off, in_src, outdeg, 0.15, DAMPING, 1.0, case["sweeps"])
    ref = golden["pr_rmat16__pr_affine"].astype(np.float64)
    rel = np.abs(r64 - ref) / np.abs(r64)
    assert rel.max() < 2e-5
    assert np.abs(r64 - ref).sum() / np.abs(r64).sum() < 2e-6


def test_pagerank_closed_form_chain():
    # chain 0->1->2->3->4: x_u = r_u (outdeg 1); after k >= i sweeps
    # r_i = c * sum_{j<=i} d^j, and a vertex without in-edges stays at c.
    src = np.array([0, 1, 2, 3], np.uint32)
    dst = np.array([1, 2, 3, 4], np.uint32)
    in_off, in_src, outdeg, _ = oracle.csc(5, src, dst)
    c, d = 0.15, DAMPING
    r = oracle.pagerank_f64(5, in_off, in_src, outdeg, c, d, 1.0, 6)
    want = [c * sum(d ** j for j in range(i + 1)) for i in range(5)]
    np.testing.assert_allclose(r, want, rtol=1e-15)


def test_pagerank_closed_form_star():
    # star: leaves 1..9 -> hub 0, hub -> 1.  Leaves 2..9 have no in-edges: r = c.
    # Sweep 1 from r = 1: hub = c + 9d (every leaf has outdeg 1), vertex 1 = c + d.
    # Sweep 2: hub = c + d*(8c + (c + d)), vertex 1 = c + d*(c + 9d).
    src = np.array(list(range(1, 10)) + [0], np.uint32)
    dst = np.array([0] * 9 + [1], np.uint32)
    in_off, in_src, outdeg, _ = oracle.csc(10, src, dst)
    c, d = 0.15, DAMPING
    r1 = oracle.pagerank_f64(10, in_off, in_src, outdeg, c, d, 1.0, 1)
    assert r1[0] == pytest.approx(c + 9 * d, rel=1e-15) and r1[1] == pytest.approx(c + d, rel=1e-15)
    r2 = oracle.pagerank_f64(10, in_off, in_src, outdeg, c, d, 1.0, 2)
    assert r2[0] == pytest.approx(c + d * (8 * c + c + d), rel=1e-15)
    assert r2[1] == pytest.approx(c + d * (c + 9 * d), rel=1e-15)
    assert (r2[2:] == c).all()


def test_ppr_f64_vs_reference_ppr(golden):
    case = next(c for c in graphs.PR_CASES if c.get("ppr_seeds"))
    src, dst, n = graphs.make(case)
    in_off, in_src, outdeg, _ = oracle.csc(n, src, dst)
    r64, sweeps = oracle.ppr_f64(n, in_off, in_src, outdeg, case["ppr_seeds"], DAMPING)
    ref = golden[f"{case['name']}__ppr"].astype(np.float64)
    # the reference ran 199 sweeps from r=1: converged to FP32 rounding
    assert np.abs(r64 - ref).sum() < 1e-6
    assert np.abs(r64 - ref).max() < 1e-6
    # fixed-point residual of the FP64 solution
    t = np.zeros(n)
    t[case["ppr_seeds"]] = 0.15 / len(case["ppr_seeds"])
    x = np.where(outdeg > 0, r64 / np.maximum(outdeg, 1), 0.0)
    s = np.zeros(n)
    np.add.at(s, dst, x[src])
    assert np.abs(r64 - (t + DAMPING * s)).sum() < 1e-12


@pytest.mark.parametrize("case", graphs.CC_CASES, ids=ids(graphs.CC_CASES))
def test_cc_oracles_bit_exact_vs_reference(case, golden):
    src, dst, n = graphs.make(case)
    ref = golden[f"{case['name']}__labels"]
    np.testing.assert_array_equal(oracle.cc_unionfind(n, src, dst), ref)
    lab, sweeps = oracle.cc_labelprop(n, src, dst)
    np.testing.assert_array_equal(lab, ref)


def test_cc_scipy_agrees():
    scipy_sparse = pytest.importorskip("scipy.sparse")
    from scipy.sparse.csgraph import connected_components
    case = next(c for c in graphs.CC_CASES if c["name"] == "cc_rmat12")
    src, dst, n = graphs.make(case)
    a = scipy_sparse.coo_matrix((np.ones(len(src)), (src, dst)), shape=(n, n))
    k, comp = connected_components(a, directed=True, connection="weak")
    lab = oracle.cc_unionfind(n, src, dst)
    # same partition, and each label is the smallest member id
    assert len(np.unique(lab)) == k
    mins = {}
    for v, c in enumerate(comp):
        mins.setdefault(c, v)
    np.testing.assert_array_equal(lab, np.array([mins[c] for c in comp], np.uint32))


def test_stable_sort_matches_numpy():
    rng = np.random.default_rng(7)
    keys = rng.integers(0, 1 << 40, 200000, dtype=np.uint64) & np.uint64((1 << 37) - 1)
    keys[::3] = keys[::3] & np.uint64(0xfff)  # many duplicates
    vals = np.arange(len(keys), dtype=np.uint32)
    k, v = oracle.stable_sort_u64(keys, vals, 37)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


def reference_shards(src, dst, n, nshards):
    """Shard contents as the reference sharder lays them out (oracle)."""
    order = np.argsort(dst, kind="stable")
    d_sorted = dst[order]
    ends, counts = oracle.shard_intervals(d_sorted, n - 1, nshards)
    out, pos = [], 0
    for c in counts:
        sel = order[pos:pos + int(c)]
        s, d = src[sel], dst[sel]
        o = np.lexsort((d, s))
        out.append((s[o], d[o]))
        pos += int(c)
    return ends, out


def degs_bytes(src, dst, n):
    deg = np.zeros((n, 2), np.int32)
    np.add.at(deg[:, 0], dst, 1)
    np.add.at(deg[:, 1], src, 1)
    return deg.tobytes()


SHARDED = [c for c in graphs.SHARD_CASES] + [c for c in graphs.PR_CASES]


@pytest.mark.parametrize("case", SHARDED, ids=ids(SHARDED))
def test_sharder_oracle_bit_exact_vs_reference(case, manifest):
    import hashlib
    src, dst, n = graphs.make(case)
    m = manifest[case["name"]]["shards"]
    nshards = case.get("nshards", 1)
    ends, shards = reference_shards(src, dst, n, nshards)
    assert ends.tolist() == m["intervals"]
    assert m["numvertices"] == n
    assert hashlib.sha256(degs_bytes(src, dst, n)).hexdigest() == m["degs_sha"]
    for (s, d), g in zip(shards, m["shards"]):
        adj, idx = oracle.shard_encode(s, d)
        assert len(adj) == g["adj_len"]
        assert hashlib.sha256(adj).hexdigest() == g["adj_sha"]
        assert idx.reshape(-1).tolist() == g["adjidx"]
        assert g["edata_size"] == 8 * len(s)
        assert g["edata_blocks"] == max(1, -(-8 * len(s) // 1048576))


@pytest.mark.parametrize("case", SHARDED, ids=ids(SHARDED))
def test_shard_decode_reads_reference_bytes(case, manifest):
    """The decoder restatement reads the reference's .adj bytes back into the
    shard's edges: the literal bytes where the manifest holds them, else the
    oracle encoding, whose SHA-256 equals the reference file's (test above)."""
    import hashlib
    src, dst, n = graphs.make(case)
    m = manifest[case["name"]]["shards"]
    _, shards = reference_shards(src, dst, n, case.get("nshards", 1))
    for (s, d), g in zip(shards, m["shards"]):
        adj = bytes.fromhex(g["adj_hex"]) if g.get("adj_hex") else oracle.shard_encode(s, d)[0]
        assert hashlib.sha256(adj).hexdigest() == g["adj_sha"]
        got = oracle.shard_decode(adj)
        assert got is not None
        np.testing.assert_array_equal(got[0], s)
        np.testing.assert_array_equal(got[1], d)


def test_shard_decode_round_trip_and_truncation():
    rng = np.random.default_rng(7)
    # first source far from 0 (leading zero runs), gaps over 255 vertices, a
    # 70000-edge source (0xff count) and many short rows
    src = np.concatenate([np.full(3, 600), np.full(70000, 601), rng.integers(1000, 5000, 20000),
                          np.full(5, 9000)]).astype(np.uint32)
    dst = rng.integers(0, 1 << 32, len(src), dtype=np.uint64).astype(np.uint32)
    o = np.lexsort((dst, src))
    src, dst = src[o], dst[o]
    adj, _ = oracle.shard_encode(src, dst)
    s, d = oracle.shard_decode(adj)
    np.testing.assert_array_equal(s, src)
    np.testing.assert_array_equal(d, dst)
    assert oracle.shard_decode(b"") is not None and len(oracle.shard_decode(b"")[0]) == 0
    assert oracle.shard_decode(b"\x00") is None                # zero run without its count byte
    assert oracle.shard_decode(b"\xff\x01\x00") is None        # 0xff without its 4-byte count
    assert oracle.shard_decode(b"\x02\x01\x00\x00\x00") is None  # two dst words announced, one present
    assert oracle.shard_decode(adj[:-1]) is None


def test_auto_nshards_rule():
    # sharder.hpp:572-575 with membudget 800 (conf/graphchi.cnf): configs of the workloads
    assert oracle.auto_nshards(67_108_864, 8, 800) == 6
    assert oracle.auto_nshards(67_108_864, 4, 800) == 4  # int(1 + 2.56 + 0.5)
    assert oracle.auto_nshards(1_048_576, 8, 800) == 1


@pytest.mark.parametrize("case", graphs.TC_CASES, ids=ids(graphs.TC_CASES))
def test_triangles_oracle_bit_exact_vs_reference(case, manifest, golden):
    src, dst, n = graphs.make(case)
    vm, pv, tot = oracle.triangles(n, src, dst)
    g = manifest[case["name"]]
    assert tot == g["total"]
    assert int(pv.astype(np.uint64).sum()) == g["sum3"]
    np.testing.assert_array_equal(vm, golden[f"{case['name']}__vertexmap"])
    np.testing.assert_array_equal(pv, golden[f"{case['name']}__tc_pervertex"])


def test_triangles_brute_force_and_closed_form():
    # K6: C(6,3)=20 triangles, C(5,2)=10 per vertex
    e = [(a, b) for a, b in itertools.combinations(range(6), 2)]
    src = np.array([a for a, _ in e], np.uint32)
    dst = np.array([b for _, b in e], np.uint32)
    vm, pv, tot = oracle.triangles(6, src, dst)
    assert tot == 20 and (pv == 10).all()
    # random small graph vs brute force
    rng = np.random.default_rng(3)
    n = 40
    adj = np.triu(rng.random((n, n)) < 0.2, 1)
    a, b = np.nonzero(adj)
    vm, pv, tot = oracle.triangles(n, a.astype(np.uint32), b.astype(np.uint32))
    sym = adj | adj.T
    bf_tot, bf_pv = 0, np.zeros(n, np.uint64)
    for i, j, k in itertools.combinations(range(n), 3):
        if sym[i, j] and sym[j, k] and sym[i, k]:
            bf_tot += 1
            for v in (i, j, k):
                bf_pv[vm[v]] += 1
    assert tot == bf_tot
    np.testing.assert_array_equal(pv, bf_pv.astype(np.uint32))


def test_psw_cpu_baseline_matches_oracle(golden):
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat14_p3")
    src, dst, n = graphs.make(case)
    r, secs = oracle.pr_psw_omp(n, src, dst, 3, 0.15, DAMPING, 1.0, case["sweeps"], nthreads=4)
    ref = golden["pr_rmat14_p3__pr_affine"]
    rel = np.abs(r.astype(np.float64) - ref) / np.abs(ref)
    assert rel.max() < 2e-5


def test_parse_edgelist_restatement():
    """The oracle's edge-list parse (the reference's convert_edgelist) on the
    cases its loop distinguishes: delimiters in runs, comments, CRLF, a value
    column, signs and trailing bytes (the C library's atoi), self-loops, no
    final newline; lines without two fields raise (the reference asserts)."""
    text = (b"# comment\n% another\n0\t1\n2,3\n  4   5  \n6\t7\t0.5\n8 8\n9\t10\r\n"
            b"+11,-1\n12abc 13x\n14 15")
    s, d = oracle.parse_edgelist(text)
    assert s.tolist() == [0, 2, 4, 6, 9, 11, 12, 14]
    assert d.tolist() == [1, 3, 5, 7, 10, 0xFFFFFFFF, 13, 15]
    for bad in (b"1 2\n\n3 4\n", b"7\n", b" , \n"):
        with pytest.raises(ValueError):
            oracle.parse_edgelist(bad)
    # the golden inputs are written as "<src>\t<dst>\n": the parse is the identity
    case = graphs.PR_CASES[0]
    src, dst, n = graphs.make(case)
    txt = "".join(f"{a}\t{b}\n" for a, b in zip(src.tolist(), dst.tolist())).encode()
    s, d = oracle.parse_edgelist(txt)
    np.testing.assert_array_equal(s, src)
    np.testing.assert_array_equal(d, dst)
