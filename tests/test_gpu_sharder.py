"""GPU parity of the device sharder: intervals, shard order, .adj bytes,
.adjidx entries and degrees bit-exact vs the reference sharder (golden), and
the reference's own pagerank_functional consuming the device-written shards."""
import hashlib
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import graphs
import oracle

pytestmark = pytest.mark.gpu

SHARDED = graphs.SHARD_CASES + graphs.PR_CASES


def sha(b):
    return hashlib.sha256(b).hexdigest()


@pytest.mark.parametrize("case", SHARDED, ids=[c["name"] for c in SHARDED])
def test_sharder_vs_reference_golden(dev, manifest, case):
    src, dst, n = graphs.make(case)
    nshards = case.get("nshards", 1)
    m = manifest[case["name"]]["shards"]
    sh = dev.Shards(dev.to_dev_u32(src), dev.to_dev_u32(dst), n - 1, nshards)
    assert sh.ends.tolist() == m["intervals"]
    for p, g in enumerate(m["shards"]):
        adj, idx = sh.encode(p)
        assert len(adj) == g["adj_len"] and sha(adj) == g["adj_sha"], f"shard {p}"
        assert idx.reshape(-1).tolist() == g["adjidx"]
    tmp = tempfile.mkdtemp(prefix="gchi_sh_")
    try:
        base = os.path.join(tmp, "g")
        sh.write_files(base, 8)
        with open(f"{base}_degs.bin", "rb") as f:
            assert sha(f.read()) == m["degs_sha"]
        with open(f"{base}.{nshards}.intervals") as f:
            assert [int(x) for x in f.read().split()] == m["intervals"]
        with open(f"{base}.numvertices") as f:
            assert int(f.read()) == m["numvertices"]
        for p, g in enumerate(m["shards"]):
            ed = f"{base}.edata.e8B.{p}_{nshards}"
            assert int(open(ed + ".size").read()) == g["edata_size"]
            assert len(os.listdir(ed + "_blockdir_1048576")) == g["edata_blocks"]
            with open(f"{base}.edata_azv.{p}_{nshards}.adj", "rb") as f:
                assert sha(f.read()) == g["adj_sha"]
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_sharder_fewer_shards_than_requested(dev):
    # one dst group cannot be split: the stream rule creates fewer shards
    src = np.arange(1, 200, dtype=np.uint32)
    dst = np.zeros(199, np.uint32)
    sh = dev.Shards(dev.to_dev_u32(src), dev.to_dev_u32(dst), 199, 4)
    ends, counts = oracle.shard_intervals(np.sort(dst), 199, 4)
    assert sh.nshards == len(ends) == 1
    assert sh.ends.tolist() == ends.tolist() and sh.counts.tolist() == counts.tolist()


def test_sharder_scale22_vs_oracle(dev):
    """The sharding workload: R-MAT scale 22 (seed 2), 6 shards (the reference's
    auto rule for PairContainer<float> edge data).  Intervals, shard order and
    .adj bytes bit-exact vs the oracle restatement at full size."""
    s, d, cnt, mx = dev.rmat_device(22, 16 << 22, 2)
    assert dev.shard_auto_count(cnt, 8) == 6
    sh = dev.Shards(s, d, mx, 6)
    hs, hd = dev.to_host_u32(s, cnt), dev.to_host_u32(d, cnt)
    ends, counts = oracle.shard_intervals(np.sort(hd), mx, 6)
    assert sh.ends.tolist() == ends.tolist() and sh.counts.tolist() == counts.tolist()
    order = np.lexsort((hs, hd))  # (dst, src) -> shard slices, then (src, dst) inside
    pos = 0
    for p in (0, sh.nshards - 1):
        ss, sd = sh.edges_host(p)
        pos = int(counts[:p].sum())
        sel = order[pos:pos + int(counts[p])]
        o = np.lexsort((hd[sel], hs[sel]))
        np.testing.assert_array_equal(ss, hs[sel][o])
        np.testing.assert_array_equal(sd, hd[sel][o])
        adj, idx = sh.encode(p)
        oadj, oidx = oracle.shard_encode(ss, sd)
        assert adj == oadj
        np.testing.assert_array_equal(idx, oidx)


REF_APP = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                       "pagerank_functional")


@pytest.mark.skipif(not os.path.exists(REF_APP), reason="reference binary not built (needs the reference tree)")
def test_reference_app_consumes_device_shards(dev, golden):
    """Shards written by the device sharder are loaded by the reference's own
    pagerank_functional (no input edge list present): its ranks equal the
    golden ranks it produced from its own sharder."""
    case = next(c for c in graphs.PR_CASES if c["name"] == "pr_rmat14_p3")
    src, dst, n = graphs.make(case)
    tmp = tempfile.mkdtemp(prefix="gchi_ref_")
    try:
        os.makedirs(os.path.join(tmp, "conf"))
        with open(os.path.join(tmp, "conf", "graphchi.cnf"), "w") as f:
            f.write("loadthreads = 4\nexecthreads = 4\nniothreads = 2\nmembudget_mb = 800\ncachesize_mb = 0\n"
                    "io.blocksize = 1048576\nmmap = 0\nmetrics.reporter = console\n")
        base = os.path.join(tmp, "g")
        sh = dev.Shards(dev.to_dev_u32(src), dev.to_dev_u32(dst), n - 1, case["nshards"])
        sh.write_files(base, 8)
        p = subprocess.run([REF_APP, "file", base, "filetype", "edgelist", "mode", "sync", "niters",
                            str(case["sweeps"] + 1), "nshards", str(case["nshards"])], cwd=tmp,
                           env=dict(os.environ, GRAPHCHI_ROOT=tmp, OMP_NUM_THREADS="4"), stdin=subprocess.DEVNULL,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=300)
        out = p.stdout.decode(errors="replace")
        assert p.returncode == 0, out[-2000:]
        assert "Found preprocessed files" in out
        r = np.fromfile(base + ".4B.vout", dtype=np.float32)
        np.testing.assert_array_equal(r, golden["pr_rmat14_p3__pr_norm"])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


REF_CC = os.path.join(os.path.dirname(REF_APP), "connectedcomponents")


@pytest.mark.skipif(not os.path.exists(REF_CC), reason="reference binary not built (needs the reference tree)")
def test_reference_cc_consumes_device_shards_with_zlib_edata(dev, golden):
    """connectedcomponents reads zlib-compressed edge-data blocks (it does not
    define GRAPHCHI_DISABLE_COMPRESSION): shards written by the device sharder
    with GCHI_SHARD_ZLIB are loaded by the reference's own binary without
    re-sharding, and its labels equal the golden labels.  Every block inflates
    to the uncompressed block (default edge values, the .size total)."""
    import zlib
    case = next(c for c in graphs.CC_CASES if c["name"] == "cc_rmat14_p3")
    src, dst, n = graphs.make(case)
    tmp = tempfile.mkdtemp(prefix="gchi_refcc_")
    try:
        os.makedirs(os.path.join(tmp, "conf"))
        with open(os.path.join(tmp, "conf", "graphchi.cnf"), "w") as f:
            f.write("loadthreads = 4\nexecthreads = 4\nniothreads = 2\nmembudget_mb = 800\ncachesize_mb = 0\n"
                    "io.blocksize = 1048576\nmmap = 0\nmetrics.reporter = console\n")
        base = os.path.join(tmp, "g")
        sh = dev.Shards(dev.to_dev_u32(src), dev.to_dev_u32(dst), n - 1, 3)
        sh.write_files(base, 4, zlib=True)
        for p in range(sh.nshards):
            ed = f"{base}.edata..Z.e4B.{p}_3"
            total = int(open(ed + ".size").read())
            d = f"{ed}_blockdir_1048576"
            got = b"".join(zlib.decompress(open(os.path.join(d, str(b)), "rb").read())
                           for b in range(len(os.listdir(d))))
            assert got == bytes(total)
        p = subprocess.run([REF_CC, "file", base, "filetype", "edgelist", "nshards", "3"], cwd=tmp,
                           env=dict(os.environ, GRAPHCHI_ROOT=tmp, OMP_NUM_THREADS="4"), stdin=subprocess.DEVNULL,
                           stdout=subprocess.PIPE, stderr=subprocess.STDOUT, timeout=300)
        out = p.stdout.decode(errors="replace")
        assert p.returncode == 0, out[-2000:]
        assert "Found preprocessed files" in out
        np.testing.assert_array_equal(np.fromfile(base + ".4B.vout", dtype=np.uint32), golden["cc_rmat14_p3__labels"])
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
