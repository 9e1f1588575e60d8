"""ctypes wrapper of liboracle.so -- the CPU restatement results are compared against.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py; never by the product.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB_PATH)
        vp, u32, u64, i32, dbl, f32 = (ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_float)
        sig = {
            "orc_csc_build": (i32, [u32, u64, vp, vp, vp, vp, vp, vp]),
            "orc_pagerank_sync_f32": (i32, [u32, vp, vp, vp, dbl, dbl, f32, i32, vp, vp]),
            "orc_pagerank_sync_f64": (i32, [u32, vp, vp, vp, dbl, dbl, dbl, i32, vp]),
            "orc_pagerank_semisync_f32": (i32, [u32, vp, vp, vp, u32, vp, dbl, dbl, f32, i32, vp, vp]),
            "orc_semisync_blocks": (i32, [u32, u32, vp, i32, vp, u32]),
            "orc_ppr_f64": (i32, [u32, vp, vp, vp, vp, u32, dbl, dbl, i32, vp]),
            "orc_cc_unionfind": (i32, [u32, u64, vp, vp, vp]),
            "orc_cc_labelprop": (i32, [u32, u64, vp, vp, vp]),
            "orc_stable_sort_u64": (i32, [u64, vp, vp, i32]),
            "orc_auto_nshards": (i32, [u64, ctypes.c_size_t, i32]),
            "orc_shard_intervals": (i32, [u64, vp, u32, i32, vp, vp]),
            "orc_shard_encode": (ctypes.c_int64, [u64, vp, vp, vp, vp, vp]),
            "orc_shard_decode": (ctypes.c_int64, [u64, vp, vp, vp, u64]),
            "orc_degree_order": (i32, [u32, vp, vp]),
            "orc_triangles": (i32, [u32, u64, vp, vp, vp, vp, vp]),
            "orc_pr_psw_omp": (dbl, [u32, u64, vp, vp, i32, dbl, dbl, f32, i32, i32, vp]),
            "orc_psw_run": (ctypes.c_long, [u32, u64, vp, vp, i32, i32, i32, vp]),
            "orc_psw_pagerank": (i32, [u32, u64, vp, vp, i32, dbl, i32, vp, vp, vp]),
            "orc_psw_sched": (i32, [u32, u64, vp, vp, i32, vp, i32, i32, vp, vp, vp, vp, vp]),
        }
        for k, (r, a) in sig.items():
            f = getattr(_lib, k)
            f.restype, f.argtypes = r, a
    return _lib


def _p(a):
    return a.ctypes.data if a is not None else None


def u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def csc(n, src, dst):
    src, dst = u32(src), u32(dst)
    in_off = np.zeros(n + 1, np.uint32)
    in_src = np.zeros(max(len(src), 1), np.uint32)
    outdeg = np.zeros(n, np.uint32)
    indeg = np.zeros(n, np.uint32)
    rc = lib().orc_csc_build(n, len(src), _p(src), _p(dst), _p(in_off), _p(in_src), _p(outdeg), _p(indeg))
    assert rc == 0
    return in_off, in_src[:len(src)], outdeg, indeg


def pagerank_f32(n, in_off, in_src, outdeg, teleport, damping, init, nsweeps):
    r = np.zeros(n, np.float32)
    stats = np.zeros(max(2 * nsweeps, 2), np.float64)
    in_src = in_src if len(in_src) else np.zeros(1, np.uint32)
    rc = lib().orc_pagerank_sync_f32(n, _p(in_off), _p(in_src), _p(outdeg), teleport, damping, init, nsweeps,
                                     _p(r), _p(stats))
    assert rc == 0
    return r, stats[:2 * nsweeps].reshape(-1, 2)


def pagerank_semisync_f32(n, in_off, in_src, outdeg, bstart, teleport, damping, init, nsweeps):
    bstart = u32(bstart)
    r = np.zeros(n, np.float32)
    stats = np.zeros(max(2 * nsweeps, 2), np.float64)
    in_src = in_src if len(in_src) else np.zeros(1, np.uint32)
    rc = lib().orc_pagerank_semisync_f32(n, _p(in_off), _p(in_src), _p(outdeg), len(bstart) - 1, _p(bstart),
                                         teleport, damping, init, nsweeps, _p(r), _p(stats))
    assert rc == 0
    return r, stats[:2 * nsweeps].reshape(-1, 2)


def semisync_blocks(n, interval_ends, membudget_mb=800):
    ends = u32(interval_ends)
    cap = len(ends) + n + 2
    b = np.zeros(cap, np.uint32)
    nb = lib().orc_semisync_blocks(n, len(ends), _p(ends), membudget_mb, _p(b), cap)
    assert nb > 0
    return b[:nb + 1]


def pagerank_f64(n, in_off, in_src, outdeg, teleport, damping, init, nsweeps):
    r = np.zeros(n, np.float64)
    in_src = in_src if len(in_src) else np.zeros(1, np.uint32)
    rc = lib().orc_pagerank_sync_f64(n, _p(in_off), _p(in_src), _p(outdeg), teleport, damping, init, nsweeps, _p(r))
    assert rc == 0
    return r


def ppr_f64(n, in_off, in_src, outdeg, seeds, damping=0.85, tol=1e-14, max_sweeps=2000):
    r = np.zeros(n, np.float64)
    seeds = u32(seeds)
    in_src = in_src if len(in_src) else np.zeros(1, np.uint32)
    it = lib().orc_ppr_f64(n, _p(in_off), _p(in_src), _p(outdeg), _p(seeds), len(seeds), damping, tol, max_sweeps,
                           _p(r))
    assert it > 0
    return r, it


def cc_unionfind(n, src, dst):
    src, dst = u32(src), u32(dst)
    lab = np.zeros(n, np.uint32)
    assert lib().orc_cc_unionfind(n, len(src), _p(src), _p(dst), _p(lab)) == 0
    return lab


def cc_labelprop(n, src, dst):
    src, dst = u32(src), u32(dst)
    lab = np.zeros(n, np.uint32)
    sweeps = lib().orc_cc_labelprop(n, len(src), _p(src), _p(dst), _p(lab))
    assert sweeps > 0
    return lab, sweeps


def stable_sort_u64(keys, vals=None, key_bits=64):
    k = np.ascontiguousarray(keys, dtype=np.uint64).copy()
    v = None if vals is None else u32(vals).copy()
    assert lib().orc_stable_sort_u64(len(k), _p(k), _p(v), key_bits) == 0
    return k, v


def auto_nshards(nedges, edge_value_bytes, membudget_mb=800):
    return lib().orc_auto_nshards(nedges, edge_value_bytes, membudget_mb)


def shard_intervals(dst_sorted, max_id, nshards):
    d = u32(dst_sorted)
    ends = np.zeros(nshards, np.uint32)
    counts = np.zeros(nshards, np.uint64)
    ns = lib().orc_shard_intervals(len(d), _p(d) if len(d) else None, max_id, nshards, _p(ends), _p(counts))
    assert ns > 0
    return ends[:ns], counts[:ns]


def shard_encode(src, dst):
    """Adjacency bytes + .adjidx triples of one shard sorted by (src,dst)."""
    src, dst = u32(src), u32(dst)
    n = len(src)
    size = lib().orc_shard_encode(n, _p(src) if n else None, _p(dst) if n else None, None, None, None)
    adj = np.zeros(max(size, 1), np.uint8)
    nidx = ctypes.c_uint64()
    lib().orc_shard_encode(n, _p(src) if n else None, _p(dst) if n else None, _p(adj), None, ctypes.byref(nidx))
    idx = np.zeros(max(3 * nidx.value, 3), np.uint64)
    lib().orc_shard_encode(n, _p(src) if n else None, _p(dst) if n else None, _p(adj), _p(idx), ctypes.byref(nidx))
    return adj[:size].tobytes(), idx[:3 * nidx.value].reshape(-1, 3)


def shard_decode(adj):
    """(src, dst) pairs of an adjacency stream, or None when it is truncated."""
    a = np.ascontiguousarray(np.frombuffer(bytes(adj), np.uint8))
    n = lib().orc_shard_decode(len(a), _p(a) if len(a) else None, None, None, 0)
    if n < 0:
        return None
    src, dst = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    lib().orc_shard_decode(len(a), _p(a) if len(a) else None, _p(src) if n else None, _p(dst) if n else None, n)
    return src, dst

def degree_order(deg):
    deg = u32(deg)
    vm = np.zeros(len(deg), np.uint32)
    assert lib().orc_degree_order(len(deg), _p(deg), _p(vm)) == 0
    return vm


def triangles(n, src, dst):
    src, dst = u32(src), u32(dst)
    vm = np.zeros(n, np.uint32)
    pv = np.zeros(n, np.uint32)
    tot = ctypes.c_uint64()
    rc = lib().orc_triangles(n, len(src), _p(src), _p(dst), _p(vm), _p(pv), ctypes.byref(tot))
    assert rc in (0, 1)
    return vm, pv, tot.value


def pr_psw_omp(n, src, dst, nshards, teleport, damping, init, nsweeps, nthreads=0):
    src, dst = u32(src), u32(dst)
    r = np.zeros(n, np.float32)
    secs = lib().orc_pr_psw_omp(n, len(src), _p(src), _p(dst), nshards, teleport, damping, init, nsweeps, nthreads,
                                _p(r))
    assert secs >= 0
    return r, secs


def _libc():
    return ctypes.CDLL(None)


def parse_edgelist(text):
    """convert_edgelist of src/preprocessing/conversions.hpp:210-288 restated:
    fgets lines of at most 1023 bytes, the trailing newline dropped, lines
    starting with '#' or '%' skipped, strtok on tab / comma / space, from and
    to by the C library's own atoi, self-loops dropped (sharder.hpp:428).
    Returns (src, dst) as uint32 arrays, or raises ValueError where the
    reference asserts (a line with fewer than two tokens)."""
    atoi = _libc().atoi
    atoi.argtypes = [ctypes.c_char_p]
    atoi.restype = ctypes.c_int
    src, dst = [], []
    pos = 0
    data = bytes(text)
    while pos < len(data):
        # fgets(s, 1024, f): up to 1023 bytes, through the first newline
        nl = data.find(b"\n", pos, pos + 1023)
        end = nl + 1 if nl >= 0 else min(len(data), pos + 1023)
        line = data[pos:end]
        pos = end
        if line.endswith(b"\n"):
            line = line[:-1]  # FIXLINE
        if line[:1] in (b"#", b"%"):
            continue
        toks = [t for t in re.split(rb"[\t, ]", line) if t]
        if len(toks) < 2:
            raise ValueError(f"line {line!r} has fewer than two tokens")
        a = atoi(toks[0]) & 0xFFFFFFFF
        b = atoi(toks[1]) & 0xFFFFFFFF
        if a != b:
            src.append(a)
            dst.append(b)
    return np.array(src, np.uint32), np.array(dst, np.uint32)


PSW_SMOKETEST, PSW_PROBE = 0, 1


def psw_run(n, src, dst, program, niters, reverse=False):
    """The reference engine's update order, one vertex after another in id
    order (orc_psw_run): (vertex values, invariant violations of the smoke
    test's program)."""
    src, dst = u32(src), u32(dst)
    vd = np.zeros(max(n, 1), np.uint32)
    viol = lib().orc_psw_run(n, len(src), _p(src), _p(dst), program, niters, 1 if reverse else 0, _p(vd))
    assert viol >= 0
    return vd[:n], int(viol)


def psw_pagerank(n, src, dst, niters, tol, inmem=False):
    """apps/psw_pagerank.cpp's program in the reference's update order
    (orc_psw_pagerank): (ranks, iterations run, logged change after each)."""
    src, dst = u32(src), u32(dst)
    r = np.zeros(max(n, 1), np.float32)
    deltas = np.zeros(max(niters, 1), np.float64)
    it = ctypes.c_int()
    assert lib().orc_psw_pagerank(n, len(src), _p(src), _p(dst), niters, tol, 1 if inmem else 0, _p(r),
                                  ctypes.byref(it), _p(deltas)) == 0
    return r[:n], it.value, deltas[:it.value]


def psw_sched(n, src, dst, interval_ends, inmem, niters):
    """apps/psw_sched.cpp's program under the reference's bitset scheduler
    (orc_psw_sched): (vertex values, deltalog lines [iteration, updates, work],
    task counts printed after each iteration)."""
    src, dst = u32(src), u32(dst)
    ends = u32(interval_ends)
    vd = np.zeros(max(n, 1), np.uint64)
    log = np.zeros(3 * max(niters, 1), np.uint64)
    tasks = np.zeros(max(niters, 1), np.uint64)
    nlog, ntasks = ctypes.c_int(), ctypes.c_int()
    assert lib().orc_psw_sched(n, len(src), _p(src), _p(dst), len(ends), _p(ends), 1 if inmem else 0, niters, _p(vd),
                               _p(log), ctypes.byref(nlog), _p(tasks), ctypes.byref(ntasks)) == 0
    return vd[:n], log[:3 * nlog.value].reshape(-1, 3).tolist(), tasks[:ntasks.value].tolist()
