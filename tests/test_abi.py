"""CPU tests of the C-ABI boundary: libgchi.so loads without a GPU and exports
every entry point include/*.h declares; error paths return status codes."""
import ctypes
import glob
import os
import re

import pytest

import gchi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(gchi_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gchi.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_functions()) >= 20


def test_python_binding_covers_header():
    assert set(gchi.declared_symbols()) <= set(declared_functions())


def test_status_strings_and_invalid_arguments():
    L = gchi.lib()
    assert L.gchi_version() >= 100
    assert L.gchi_strerror(-1) == b"invalid argument"
    # null pointers are rejected with a status, never a crash
    assert L.gchi_pr_plan_create(None, None, None, None) == -1
    buf = ctypes.create_string_buffer(256)
    assert L.gchi_last_error(buf, 256) == 0 and b"null" in buf.value
    assert L.gchi_sort_u64(None, None, 10, 64, None, 0, None) == -1
    nb = ctypes.c_size_t()
    assert L.gchi_sort_workspace(1 << 20, 44, 0, ctypes.byref(nb)) == 0 and nb.value > (1 << 23)
    # sort refuses sizes its 32-bit scatter positions cannot hold
    assert L.gchi_sort_u64(ctypes.c_void_p(1), None, 1 << 32, 64, ctypes.c_void_p(1), 1 << 40, None) == -1
    assert L.gchi_last_error(buf, 256) == 0 and b"2^32" in buf.value


def test_host_generator_deterministic():
    a = gchi.rmat_host(12, 5000, 9)
    b = gchi.rmat_host(12, 5000, 9, nthreads=3)
    assert (a[0] == b[0]).all() and (a[1] == b[1]).all()
    # a suffix generated from an offset equals the tail of the full run
    c = gchi.rmat_host(12, 1000, 9, first=4000)
    assert (c[0] == a[0][4000:]).all() and (c[1] == a[1][4000:]).all()
    assert a[0].max() < 4096 and a[1].max() < 4096


def test_size_limits_refused_before_device_work():
    """Edge counts beyond each entry point's stated limit come back as
    GCHI_ERR_UNSUPPORTED (-4) before any device memory is touched (the
    pointers here are placeholders): TC preparation and degree ordering
    E < 2^31; CSC build, sharder and the generic engine's adjacency
    E <= 2^32 - 2^25 (a 32-bit edge index plus a grid stride never wraps)."""
    L = gchi.lib()
    one = ctypes.c_void_p(1)
    big = (1 << 32) - (1 << 25) + 1
    m = ctypes.c_uint64()
    L.gchi_degree_order.restype = ctypes.c_int
    L.gchi_degree_order.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32,
                                    ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    assert L.gchi_tc_prepare(one, one, 1 << 31, 10, one, one, one, ctypes.byref(m), one, 1 << 40, None) == -4
    assert L.gchi_degree_order(one, one, 1 << 31, 10, one, one, 1 << 40, None) == -4
    assert L.gchi_csc_build(one, one, big, 10, one, one, one, one, one, 1 << 40, None) == -4
    sh = ctypes.c_void_p()
    assert L.gchi_shard_build(one, one, big, 9, 2, ctypes.byref(sh), None) == -4
    assert L.gchi_psw_graph_build(one, one, big, 10, one, one, one, one, one, one, one, 1 << 40, None) == -4
    buf = ctypes.create_string_buffer(256)
    assert L.gchi_last_error(buf, 256) == 0 and b"2^32 - 2^25" in buf.value
