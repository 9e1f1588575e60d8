"""End-to-end drop-in tests: applications run from an edge-list text file
through the C++ host layer (graphchi-cpp_amd/host) and the HIP library.

* the reference's own pagerank_functional.cpp and connectedcomponents.cpp,
  compiled unchanged against the host headers (apps/_dropin, built where the
  reference tree is present; both modes of pagerank_functional);
* the reference's trianglecounting.cpp, likewise;
* this repository's apps (apps/bin): pagerank, connectedcomponents, sharder.
Outputs (.vout files, shard files) are compared with what the reference's
binaries wrote for the same input (golden fixtures)."""
import hashlib
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import graphs

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "graphchi-cpp_amd", "apps", "bin")
DROPIN = os.path.join(ROOT, "graphchi-cpp_amd", "apps", "_dropin")


def case(name):
    return next(c for c in graphs.all_cases() if c["name"] == name)


@pytest.fixture
def workdir():
    d = tempfile.mkdtemp(prefix="gchi_app_")
    yield d
    shutil.rmtree(d, ignore_errors=True)


def write_edgelist(d, name):
    src, dst, n = graphs.make(case(name))
    path = os.path.join(d, name)
    with open(path, "w") as f:
        f.write("".join(f"{a}\t{b}\n" for a, b in zip(src.tolist(), dst.tolist())))
    return path, n


def run(exe, args, cwd):
    p = subprocess.run([exe] + [str(a) for a in args], cwd=cwd, stdin=subprocess.DEVNULL, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, timeout=300)
    out = p.stdout.decode(errors="replace")
    assert p.returncode == 0, out[-3000:]
    return out


def assert_close(r, ref, rtol=2e-5, l1=2e-6):
    r, ref = r.astype(np.float64), ref.astype(np.float64)
    err = np.abs(r - ref)
    assert (err <= rtol * np.abs(ref) + 1e-30).all()
    assert err.sum() <= l1 * np.abs(ref).sum()


def shard_shas(base, nshards, m):
    for p, g in enumerate(m["shards"]["shards"]):
        with open(f"{base}.edata_azv.{p}_{nshards}.adj", "rb") as f:
            assert hashlib.sha256(f.read()).hexdigest() == g["adj_sha"]


@pytest.mark.skipif(not os.path.exists(os.path.join(DROPIN, "pagerank_functional")), reason="drop-in not built")
def test_reference_pagerank_functional_source_on_device(workdir, golden, manifest):
    path, n = write_edgelist(workdir, "pr_rmat14_p3")
    c = case("pr_rmat14_p3")
    out = run(os.path.join(DROPIN, "pagerank_functional"),
              ["file", path, "filetype", "edgelist", "mode", "sync", "niters", c["sweeps"] + 1, "nshards", 3], workdir)
    assert "Print top 20 vertices" in out
    assert_close(np.fromfile(path + ".4B.vout", np.float32), golden["pr_rmat14_p3__pr_norm"])
    shard_shas(path, 3, manifest["pr_rmat14_p3"])


@pytest.mark.skipif(not os.path.exists(os.path.join(DROPIN, "pagerank_functional")), reason="drop-in not built")
@pytest.mark.parametrize("name", ["pr_rmat14_p3", "pr_rmat16"])
def test_reference_pagerank_functional_default_mode(workdir, golden, name):
    """No `mode` argument: the reference's default, semi-synchronous mode, at
    the default memory budget and at membudget_mb 1 (several sub-intervals per
    interval), against what the reference binary wrote in the same runs."""
    path, n = write_edgelist(workdir, name)
    c = case(name)
    args = ["file", path, "filetype", "edgelist", "niters", c["sweeps"] + 1, "nshards", c.get("nshards", 1)]
    run(os.path.join(DROPIN, "pagerank_functional"), args, workdir)
    assert_close(np.fromfile(path + ".4B.vout", np.float32), golden[f"{name}__pr_semisync"])
    run(os.path.join(DROPIN, "pagerank_functional"), args + ["membudget_mb", 1], workdir)
    assert_close(np.fromfile(path + ".4B.vout", np.float32), golden[f"{name}__pr_semisync_mb1"])


@pytest.mark.skipif(not os.path.exists(os.path.join(DROPIN, "connectedcomponents")), reason="drop-in not built")
def test_reference_connectedcomponents_source_on_device(workdir, golden):
    path, n = write_edgelist(workdir, "cc_rmat14_p3")
    out = run(os.path.join(DROPIN, "connectedcomponents"), ["file", path, "filetype", "edgelist", "nshards", 3],
              workdir)
    np.testing.assert_array_equal(np.fromfile(path + ".4B.vout", np.uint32), golden["cc_rmat14_p3__labels"])
    assert os.path.exists(path + ".components")
    # the program's before/after_iteration ran every round: its convergence
    # message, and one progress line per iteration
    assert "Converged!" in out
    with open(path + ".deltalog") as f:
        lines = [l.split(",") for l in f.read().split()]
    assert len(lines) >= 2 and [int(l[0]) for l in lines] == list(range(len(lines)))
    # niters bounds the iterations
    run(os.path.join(DROPIN, "connectedcomponents"), ["file", path, "filetype", "edgelist", "nshards", 3, "niters", 1],
        workdir)
    with open(path + ".deltalog") as f:
        assert len(f.read().split()) == 1


@pytest.mark.skipif(not os.path.exists(os.path.join(DROPIN, "trianglecounting")), reason="drop-in not built")
@pytest.mark.parametrize("name", ["tc_k5", "tc_rmat10", "tc_rmat13"])
def test_reference_trianglecounting_source_on_device(workdir, golden, manifest, name):
    """example_apps/trianglecounting.cpp compiled unchanged: its degree
    ordering (.vertexmap), per-vertex counts (<G>_degord.4B.vout) and printed
    total equal what the reference binary wrote for the same edge list."""
    path, n = write_edgelist(workdir, name)
    out = run(os.path.join(DROPIN, "trianglecounting"), ["file", path, "filetype", "edgelist", "nshards", 2], workdir)
    g = manifest[name]
    assert f"Number of triangles: {g['total']}({g['sum3']})" in out
    np.testing.assert_array_equal(np.fromfile(path + ".vertexmap", np.uint32), golden[f"{name}__vertexmap"])
    np.testing.assert_array_equal(np.fromfile(path + "_degord.4B.vout", np.uint32), golden[f"{name}__tc_pervertex"])


def test_app_pagerank_both_forms(workdir, golden):
    path, n = write_edgelist(workdir, "pr_rmat16")
    sweeps = case("pr_rmat16")["sweeps"]
    run(os.path.join(BIN, "pagerank"), ["file", path, "niters", sweeps + 1, "form", "unnormalised", "nshards", 1],
        workdir)
    assert_close(np.fromfile(path + ".4B.vout", np.float32), golden["pr_rmat16__pr_affine"])
    run(os.path.join(BIN, "pagerank"), ["file", path, "niters", sweeps + 1, "form", "normalised"], workdir)
    assert_close(np.fromfile(path + ".4B.vout", np.float32), golden["pr_rmat16__pr_norm"])


def test_app_connectedcomponents(workdir, golden):
    path, n = write_edgelist(workdir, "cc_rmat12")
    run(os.path.join(BIN, "connectedcomponents"), ["file", path], workdir)
    np.testing.assert_array_equal(np.fromfile(path + ".4B.vout", np.uint32), golden["cc_rmat12__labels"])


@pytest.mark.parametrize("name", ["sh_rmat12_p7", "sh_hand_gaps"])
def test_app_sharder(workdir, manifest, name):
    path, n = write_edgelist(workdir, name)
    nshards = case(name)["nshards"]
    run(os.path.join(BIN, "sharder"), ["file", path, "nshards", nshards, "edgedatatype", "float-float"], workdir)
    shard_shas(path, nshards, manifest[name])
    with open(f"{path}_degs.bin", "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == manifest[name]["shards"]["degs_sha"]
