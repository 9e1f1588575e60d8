"""CPU tests of the C++ host layer's option handling (no GPU): without a
configuration file the reference's shipped conf/graphchi.cnf values apply, so
membudget_mb -- which sets the automatic shard count -- is 800, as in the
reference; a conf file under GRAPHCHI_ROOT and the command line override it.
The automatic shard count the sharder derives from it is the oracle's."""
import os
import subprocess
import tempfile

import pytest

import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "graphchi-cpp_amd", "host")

PROG = r'''
#include <cstdio>
#include "util/cmdopts.hpp"
using namespace graphchi;
int main(int argc, const char **argv) {
    graphchi_init(argc, argv);
    std::printf("%d %d\n", get_option_int("membudget_mb", 1024), get_option_int("loadthreads", 1));
    return 0;
}
'''


@pytest.fixture(scope="module")
def prog():
    d = tempfile.mkdtemp(prefix="gchi_opts_")
    src = os.path.join(d, "opts.cpp")
    with open(src, "w") as f:
        f.write(PROG)
    exe = os.path.join(d, "opts")
    subprocess.run(["g++", "-std=c++17", "-O0", "-I" + HOST, src, "-o", exe], check=True)
    return d, exe


def _run(exe, args, cwd, root=None):
    env = dict(os.environ)
    env.pop("GRAPHCHI_ROOT", None)
    if root:
        env["GRAPHCHI_ROOT"] = root
    out = subprocess.run([exe] + args, cwd=cwd, env=env, capture_output=True, text=True, check=True).stdout
    return [int(x) for x in out.split()]


def test_defaults_without_conf_file(prog):
    d, exe = prog
    empty = tempfile.mkdtemp(prefix="gchi_noconf_")
    assert _run(exe, [], empty) == [800, 4]


def test_conf_file_and_command_line_override(prog):
    d, exe = prog
    root = tempfile.mkdtemp(prefix="gchi_root_")
    os.makedirs(os.path.join(root, "conf"))
    with open(os.path.join(root, "conf", "graphchi.cnf"), "w") as f:
        f.write("# test\nmembudget_mb = 3000\nloadthreads = 8  # comment\n")
    assert _run(exe, [], d, root) == [3000, 8]
    assert _run(exe, ["membudget_mb", "123"], d, root) == [123, 8]
    assert _run(exe, ["--membudget_mb=64"], d) == [64, 4]


def test_auto_shard_count_at_default_budget():
    """The scale-22 graph with PairContainer<float> edges (8 bytes) takes 6
    shards at the reference's 800 MB budget (5 at 1024 MB, which the host layer
    used to assume without a conf file)."""
    import gchi
    e = 67_107_143
    assert oracle.auto_nshards(e, 8, 800) == 6
    assert oracle.auto_nshards(e, 8, 1024) == 5
    assert gchi.shard_auto_count(e, 8) == 6
