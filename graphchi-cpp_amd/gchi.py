"""ctypes binding of libgchi.so for the tests and bench.py.

The library's C ABI (include/gchi.h) takes plain device pointers and a HIP
stream; torch is used here only to own device memory and to provide the
stream (torch.cuda.current_stream()).  Every call goes to the HIP library:
there is no CPU fallback, and a missing or unloadable library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgchi.so")

RMAT_ABCD = (0.57, 0.19, 0.19, 0.05)


class GchiError(RuntimeError):
    pass


c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_i32 = ctypes.c_int
c_vp = ctypes.c_void_p
c_sz = ctypes.c_size_t


class DGraph(ctypes.Structure):
    _fields_ = [("nvertices", c_u32), ("reserved", c_u32), ("nedges", c_u64),
                ("in_off", c_vp), ("in_src", c_vp), ("outdeg", c_vp), ("indeg", c_vp)]


class PrOp(ctypes.Structure):
    _fields_ = [("teleport", ctypes.c_double), ("damping", ctypes.c_double), ("d_graph_teleport", c_vp),
                ("d_seed_bits", c_vp), ("seed_teleport", ctypes.c_double)]


class PrTune(ctypes.Structure):
    _fields_ = [("edges_per_block", c_i32), ("short_row_max", c_i32),
                ("persistent", c_i32), ("hot_vertices", c_i32)]


class PrPlanInfo(ctypes.Structure):
    _fields_ = [("nvertices", c_u32), ("edges_per_block", c_u32), ("nedges", c_u64),
                ("short_row_max", c_u32), ("regular_blocks", c_u32), ("hub_rows", c_u32),
                ("hub_chunks", c_u32), ("regular_edges", c_u64), ("hub_edges", c_u64),
                ("ngraphs", c_u32), ("persistent", c_u32), ("persistent_workgroups", c_u32), ("hot_vertices", c_u32),
                ("gather_sources", c_u32), ("hot_threads", c_u32)]


class PprConfig(ctypes.Structure):
    _fields_ = [("damping", ctypes.c_double), ("tol", ctypes.c_double), ("bf16_switch_residual", ctypes.c_double),
                ("bf16_sweeps", c_i32), ("max_sweeps", c_i32), ("check_every", c_i32), ("reserved", c_i32)]


class PprResult(ctypes.Structure):
    _fields_ = [("sweeps_bf16", c_i32), ("sweeps_fp32", c_i32), ("converged", c_i32), ("block_sweeps_fp32", c_i32),
                ("max_residual", ctypes.c_double)]


# name -> (restype, argtypes)
_SIGS = {
    "gchi_version": (c_i32, []),
    "gchi_strerror": (ctypes.c_char_p, [c_i32]),
    "gchi_last_error": (c_i32, [ctypes.c_char_p, c_sz]),
    "gchi_device_open": (c_i32, [c_i32]),
    "gchi_device_info": (c_i32, [ctypes.POINTER(c_i32), ctypes.c_char_p, c_sz]),
    "gchi_memcpy": (c_i32, [c_vp, c_vp, c_sz, c_vp]),
    "gchi_rmat_generate": (c_i32, [c_vp, c_vp, c_u64, c_u64, c_i32, c_u64, ctypes.POINTER(ctypes.c_double), c_vp]),
    "gchi_rmat_generate_host": (c_i32, [c_vp, c_vp, c_u64, c_u64, c_i32, c_u64, ctypes.POINTER(ctypes.c_double), c_i32]),
    "gchi_drop_self_loops_workspace": (c_i32, [c_u64, ctypes.POINTER(c_sz)]),
    "gchi_drop_self_loops": (c_i32, [c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "gchi_parse_edgelist_workspace": (c_i32, [c_u64, ctypes.POINTER(c_sz)]),
    "gchi_parse_edgelist": (c_i32, [c_vp, c_u64, c_vp, c_vp, c_u64, ctypes.POINTER(c_u64), ctypes.POINTER(c_u32),
                                    c_vp, c_sz, c_vp]),
    "gchi_sort_workspace": (c_i32, [c_u64, c_i32, c_i32, ctypes.POINTER(c_sz)]),
    "gchi_sort_u64": (c_i32, [c_vp, c_vp, c_u64, c_i32, c_vp, c_sz, c_vp]),
    "gchi_sort_pairs_dst_src_workspace": (c_i32, [c_u64, c_i32, ctypes.POINTER(c_sz)]),
    "gchi_sort_pairs_dst_src": (c_i32, [c_vp, c_vp, c_u64, c_i32, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "gchi_work_status": (c_i32, [c_vp, c_vp]),
    "gchi_debug_sort_lookback": (c_i32, [c_u32, c_i32]),
    "gchi_csc_build_workspace": (c_i32, [c_u64, c_u32, ctypes.POINTER(c_sz)]),
    "gchi_csc_build": (c_i32, [c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "gchi_pr_plan_create": (c_i32, [ctypes.POINTER(DGraph), ctypes.POINTER(PrTune), ctypes.POINTER(c_vp), c_vp]),
    "gchi_pr_plan_destroy": (c_i32, [c_vp]),
    "gchi_pr_plan_info_get": (c_i32, [c_vp, ctypes.POINTER(PrPlanInfo)]),
    "gchi_pr_init": (c_i32, [c_vp, ctypes.c_float, c_vp, c_vp, c_vp]),
    "gchi_pr_sweep": (c_i32, [c_vp, ctypes.POINTER(PrOp), c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "gchi_pr_exec_create": (c_i32, [c_vp, ctypes.POINTER(PrOp), ctypes.c_float, c_i32, c_vp, c_vp, c_vp, c_vp,
                                    c_vp, ctypes.POINTER(c_vp)]),
    "gchi_pr_exec_create_ex": (c_i32, [c_vp, ctypes.POINTER(PrOp), ctypes.c_float, c_i32, c_i32, c_vp, c_vp, c_vp,
                                       c_vp, c_vp, ctypes.POINTER(c_vp)]),
    "gchi_pr_plan_create_multi": (c_i32, [ctypes.POINTER(DGraph), c_vp, c_u32, ctypes.POINTER(PrTune),
                                          ctypes.POINTER(c_vp), c_vp]),
    "gchi_pr_init_seeds": (c_i32, [c_vp, c_vp, ctypes.c_float, c_vp, c_vp, c_vp]),
    "gchi_pr_exec_launch": (c_i32, [c_vp, c_vp]),
    "gchi_pr_blocked_create": (c_i32, [ctypes.POINTER(DGraph), c_vp, c_u32, ctypes.POINTER(PrTune),
                                       ctypes.POINTER(c_vp), c_vp]),
    "gchi_pr_blocked_destroy": (c_i32, [c_vp]),
    "gchi_pr_blocked_init": (c_i32, [c_vp, ctypes.c_float, c_vp, ctypes.c_float, c_vp, c_vp, c_vp]),
    "gchi_pr_blocked_sweep": (c_i32, [c_vp, ctypes.POINTER(PrOp), c_vp, c_vp, c_vp, c_vp, c_vp]),
    "gchi_host_alloc": (c_i32, [ctypes.POINTER(c_vp), c_sz]),
    "gchi_host_free": (c_i32, [c_vp]),
    "gchi_pr_exec_destroy": (c_i32, [c_vp]),
    "gchi_autotune_pr": (c_i32, [ctypes.POINTER(DGraph), c_i32, ctypes.POINTER(PrTune), ctypes.POINTER(ctypes.c_double),
                                 c_vp, c_vp, c_i32, ctypes.POINTER(c_i32), c_vp]),
    "gchi_ppr_batch_create": (c_i32, [c_vp, c_u32, ctypes.POINTER(c_vp)]),
    "gchi_ppr_batch_destroy": (c_i32, [c_vp]),
    "gchi_ppr_batch_run": (c_i32, [c_vp, c_vp, c_vp, ctypes.POINTER(PprConfig), ctypes.POINTER(PprResult), c_vp,
                                   c_vp]),
    "gchi_ppr_batch_extract": (c_i32, [c_vp, c_u32, c_vp, c_vp]),
    "gchi_cc_workspace": (c_i32, [c_u32, ctypes.POINTER(c_sz)]),
    "gchi_cc_minlabel": (c_i32, [c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_sz, c_vp]),
    "gchi_cc_labelprop_init": (c_i32, [c_vp, c_u32, c_vp]),
    "gchi_cc_labelprop_sweep": (c_i32, [c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_vp, c_vp]),
    "gchi_cc_scheduled_workspace": (c_i32, [c_u32, c_u64, ctypes.POINTER(c_sz)]),
    "gchi_cc_afforest_workspace": (c_i32, [c_u32, ctypes.POINTER(c_sz)]),
    "gchi_psw_graph_workspace": (c_i32, [c_u64, c_u32, ctypes.POINTER(c_sz)]),
    "gchi_psw_graph_build": (c_i32, [c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "gchi_psw_graph_build_ex": (c_i32, [c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_sz,
                                        c_vp]),
    "gchi_edata_read": (c_i32, [ctypes.c_char_p, c_i32, c_u32, c_u64, c_vp]),
    "gchi_edata_write": (c_i32, [ctypes.c_char_p, c_i32, c_u32, c_u64, c_vp]),
    "gchi_cc_pull_sweep": (c_i32, [c_vp, c_vp, c_u32, c_vp, c_vp, c_vp, c_vp]),
    "gchi_cc_afforest": (c_i32, [c_vp, c_vp, c_u32, c_vp, c_vp, c_sz, c_vp]),
    "gchi_shard_decode_workspace": (c_i32, [c_u64, ctypes.POINTER(c_sz)]),
    "gchi_shard_decode_many_workspace": (c_i32, [c_vp, c_u32, ctypes.POINTER(c_sz)]),
    "gchi_shard_decode_many": (c_i32, [c_vp, c_vp, c_u32, c_vp, c_vp, c_u64, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "gchi_shard_decode": (c_i32, [c_vp, c_u64, c_vp, c_vp, c_u64, ctypes.POINTER(c_u64), ctypes.POINTER(c_u64), c_vp,
                                  c_sz, c_vp]),
    "gchi_cc_scheduled": (c_i32, [c_vp, c_vp, c_u32, c_u32, c_vp, ctypes.POINTER(c_u32), ctypes.POINTER(c_u64), c_vp,
                                  c_sz, c_vp]),
    "gchi_tc_prepare_workspace": (c_i32, [c_u64, c_u32, ctypes.POINTER(c_sz)]),
    "gchi_tc_prepare": (c_i32, [c_vp, c_vp, c_u64, c_u32, c_vp, c_vp, c_vp, ctypes.POINTER(c_u64), c_vp, c_sz,
                                c_vp]),
    "gchi_tc_count_workspace": (c_i32, [c_u32, ctypes.POINTER(c_sz)]),
    "gchi_tc_count": (c_i32, [c_vp, c_vp, c_u32, c_vp, c_vp, c_vp, c_vp, c_sz, c_vp]),
    "gchi_tc_count_noalloc": (c_i32, [c_vp, c_vp, c_u32, c_vp, c_vp, c_vp, c_vp]),
    "gchi_shard_auto_count": (c_i32, [c_u64, c_u32, c_i32, ctypes.POINTER(c_i32)]),
    "gchi_shard_build": (c_i32, [c_vp, c_vp, c_u64, c_u32, c_i32, ctypes.POINTER(c_vp), c_vp]),
    "gchi_shard_destroy": (c_i32, [c_vp]),
    "gchi_shard_info": (c_i32, [c_vp, ctypes.POINTER(c_i32), c_vp, c_vp]),
    "gchi_shard_edges": (c_i32, [c_vp, c_i32, ctypes.POINTER(c_vp), ctypes.POINTER(c_vp), ctypes.POINTER(c_u64)]),
    "gchi_shard_degrees": (c_i32, [c_vp, ctypes.POINTER(c_vp)]),
    "gchi_shard_encode": (c_i32, [c_vp, c_i32, c_vp, ctypes.POINTER(c_u64), c_vp, c_u32, ctypes.POINTER(c_u32),
                                  c_vp]),
    "gchi_shard_write_files": (c_i32, [c_vp, ctypes.c_char_p, c_u32, c_vp]),
    "gchi_shard_write_files_ex": (c_i32, [c_vp, ctypes.c_char_p, c_u32, c_i32, c_vp]),
}

_lib = None


def lib():
    """Loads libgchi.so (raises GchiError when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GchiError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        # torch bundles its own HIP runtime (same SONAME libamdhip64.so.7): load it
        # first so that this process holds exactly one HIP runtime.
        import torch  # noqa: F401
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def declared_symbols():
    return sorted(_SIGS)


def check(status, what=""):
    if status != 0:
        buf = ctypes.create_string_buffer(1024)
        lib().gchi_last_error(buf, 1024)
        msg = lib().gchi_strerror(status).decode()
        raise GchiError(f"{what}: {msg} ({status}): {buf.value.decode()}")


# ---------------------------------------------------------------- torch glue
def _torch():
    import torch
    return torch


def dptr(t):
    return c_vp(t.data_ptr()) if t is not None else c_vp(0)


def stream():
    torch = _torch()
    return c_vp(torch.cuda.current_stream().cuda_stream)


def empty_u32(n):
    torch = _torch()
    return torch.empty(max(int(n), 1), dtype=torch.int32, device="cuda")


def workspace(nbytes):
    torch = _torch()
    return torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device="cuda")


def to_host_u32(t, n=None):
    a = t.cpu().numpy().view(np.uint32)
    return a if n is None else a[:n]


def to_dev_u32(a):
    torch = _torch()
    a = np.ascontiguousarray(a, dtype=np.uint32)
    return torch.from_numpy(a.view(np.int32)).to("cuda")


def open_device(dev=0):
    check(lib().gchi_device_open(dev), "gchi_device_open")


def _abcd(abcd):
    return (ctypes.c_double * 4)(*abcd)


# ---------------------------------------------------------------- R-MAT input
def rmat_host(scale, nedges, seed, abcd=RMAT_ABCD, first=0, nthreads=0):
    """Raw R-MAT pairs on the host (self-loops kept), as numpy uint32 arrays."""
    src = np.empty(nedges, dtype=np.uint32)
    dst = np.empty(nedges, dtype=np.uint32)
    check(lib().gchi_rmat_generate_host(src.ctypes.data, dst.ctypes.data, first, nedges, scale, seed,
                                        _abcd(abcd), nthreads), "rmat_host")
    return src, dst


def drop_self_loops_host(src, dst):
    keep = src != dst
    return src[keep], dst[keep]


def rmat_device(scale, nedges, seed, abcd=RMAT_ABCD, drop_self_loops=True):
    """R-MAT pairs generated in HBM; returns (src, dst, count, max_id)."""
    torch = _torch()
    L = lib()
    s0, d0 = empty_u32(nedges), empty_u32(nedges)
    check(L.gchi_rmat_generate(dptr(s0), dptr(d0), 0, nedges, scale, seed, _abcd(abcd), stream()), "rmat")
    if not drop_self_loops:
        return s0, d0, nedges, None
    s1, d1 = empty_u32(nedges), empty_u32(nedges)
    nb = c_sz()
    check(L.gchi_drop_self_loops_workspace(nedges, ctypes.byref(nb)))
    ws = workspace(nb.value)
    cnt = torch.zeros(1, dtype=torch.int64, device="cuda")
    mx = torch.zeros(1, dtype=torch.int32, device="cuda")
    check(L.gchi_drop_self_loops(dptr(s0), dptr(d0), nedges, dptr(s1), dptr(d1), dptr(cnt), dptr(mx),
                                 dptr(ws), nb.value, stream()), "drop_self_loops")
    del s0, d0, ws
    n = int(cnt.item())
    maxid = int(mx.cpu().numpy().view(np.uint32)[0])
    return s1[:n], d1[:n], n, maxid


# ---------------------------------------------------------------- text input
def parse_edgelist(text):
    """Edge-list bytes (host bytes or a device uint8 tensor) parsed on the
    device: returns (src, dst, count, max_id) with self-loops dropped."""
    torch = _torch()
    L = lib()
    if isinstance(text, (bytes, bytearray)):
        t = torch.from_numpy(np.frombuffer(bytes(text), np.uint8).copy()).to("cuda") if text else None
        n = len(text)
    else:
        t, n = text, int(text.numel())
    nb = c_sz()
    check(L.gchi_parse_edgelist_workspace(n, ctypes.byref(nb)))
    ws = workspace(nb.value)
    ne, mx = c_u64(), c_u32()
    check(L.gchi_parse_edgelist(dptr(t), n, None, None, 0, ctypes.byref(ne), ctypes.byref(mx), dptr(ws), nb.value,
                                stream()), "parse_edgelist (count)")
    src, dst = empty_u32(ne.value), empty_u32(ne.value)
    check(L.gchi_parse_edgelist(dptr(t), n, dptr(src), dptr(dst), max(ne.value, 1), ctypes.byref(ne),
                                ctypes.byref(mx), dptr(ws), nb.value, stream()), "parse_edgelist")
    return src[:ne.value], dst[:ne.value], ne.value, mx.value


# ---------------------------------------------------------------- sort
def work_status(ws, what="work_status"):
    """Waits for the stream and raises if the call that used workspace ws
    reported a device-side failure (status word in its first 4 bytes)."""
    check(lib().gchi_work_status(dptr(ws), stream()), what)


def sort_u64(keys, vals=None, key_bits=64, check_status=True):
    L = lib()
    n = keys.numel()
    nb = c_sz()
    check(L.gchi_sort_workspace(n, key_bits, 1 if vals is not None else 0, ctypes.byref(nb)))
    ws = workspace(nb.value)
    check(L.gchi_sort_u64(dptr(keys), dptr(vals), n, key_bits, dptr(ws), nb.value, stream()), "sort_u64")
    if check_status:
        work_status(ws, "sort_u64")
    return ws


def sort_pairs_dst_src(src, dst, id_bits, check_status=True, out=None):
    """(src, dst) sorted by (dst, src); out = (src_out, dst_out) to write into
    given tensors (they may be src, dst themselves)."""
    L = lib()
    n = src.numel()
    so, do = out if out is not None else (empty_u32(n), empty_u32(n))
    nb = c_sz()
    check(L.gchi_sort_pairs_dst_src_workspace(n, id_bits, ctypes.byref(nb)))
    ws = workspace(nb.value)
    check(L.gchi_sort_pairs_dst_src(dptr(src), dptr(dst), n, id_bits, dptr(so), dptr(do), dptr(ws), nb.value,
                                    stream()), "sort_pairs_dst_src")
    if check_status:
        work_status(ws, "sort_pairs_dst_src")
    return so, do


def debug_sort_lookback(spin_limit=0, force_stall=False):
    """Test hook: bound of the sort's look-back waits (0 = default) and forced stalls."""
    check(lib().gchi_debug_sort_lookback(spin_limit, 1 if force_stall else 0), "debug_sort_lookback")


# ---------------------------------------------------------------- graph
class DeviceGraph:
    """CSC of a graph resident in HBM (owned torch tensors + gchi_dgraph view)."""

    def __init__(self, src, dst, nvertices):
        L = lib()
        E = int(src.numel()) if src is not None else 0
        N = int(nvertices)
        self.N, self.E = N, E
        self.in_off = empty_u32(N + 1)
        self.in_src = empty_u32(E)
        self.outdeg = empty_u32(N)
        self.indeg = empty_u32(N)
        nb = c_sz()
        check(L.gchi_csc_build_workspace(E, N, ctypes.byref(nb)))
        ws = workspace(nb.value)
        check(L.gchi_csc_build(dptr(src), dptr(dst), E, N, dptr(self.in_off), dptr(self.in_src), dptr(self.outdeg),
                               dptr(self.indeg), dptr(ws), nb.value, stream()), "csc_build")
        work_status(ws, "csc_build")
        self.dg = DGraph(N, 0, E, self.in_off.data_ptr(), self.in_src.data_ptr(), self.outdeg.data_ptr(),
                         self.indeg.data_ptr())


def pr_op(teleport, damping=0.85, graph_teleport=None, seed_bits=None, seed_teleport=0.0):
    """gchi_pr_op; graph_teleport / seed_bits are device tensors (kept alive by the caller)."""
    op = PrOp(teleport, damping, graph_teleport.data_ptr() if graph_teleport is not None else None,
              seed_bits.data_ptr() if seed_bits is not None else None, seed_teleport)
    op._keep = (graph_teleport, seed_bits)
    return op


def seed_bitmap(seeds, n):
    """Device bitmap (uint32 words) with the seed vertices set."""
    words = np.zeros((n + 31) // 32, np.uint32)
    seeds = np.asarray(seeds, dtype=np.uint64)
    np.bitwise_or.at(words, (seeds >> np.uint64(5)).astype(np.int64),
                     (np.uint32(1) << (seeds & np.uint64(31)).astype(np.uint32)))
    return to_dev_u32(words)


def graph_from_csc(in_off, in_src, outdeg, indeg, n, e):
    """DeviceGraph over existing device CSC tensors (no build)."""
    g = DeviceGraph.__new__(DeviceGraph)
    g.N, g.E = int(n), int(e)
    g.in_off, g.in_src, g.outdeg, g.indeg = in_off, in_src, outdeg, indeg
    g.dg = DGraph(g.N, 0, g.E, in_off.data_ptr(), in_src.data_ptr(), outdeg.data_ptr(), indeg.data_ptr())
    return g


class PageRankPlan:
    def __init__(self, graph, edges_per_block=0, short_row_max=0, persistent=0, vbase=None, hot_vertices=0):
        self.graph = graph
        self.handle = c_vp()
        tune = PrTune(edges_per_block, short_row_max, persistent, hot_vertices)
        if vbase is None:
            self.ngraphs = 1
            check(lib().gchi_pr_plan_create(ctypes.byref(graph.dg), ctypes.byref(tune), ctypes.byref(self.handle),
                                            stream()), "pr_plan_create")
        else:
            vb = np.ascontiguousarray(vbase, dtype=np.uint32)
            self.ngraphs = len(vb) - 1
            check(lib().gchi_pr_plan_create_multi(ctypes.byref(graph.dg), c_vp(vb.ctypes.data), self.ngraphs,
                                                  ctypes.byref(tune), ctypes.byref(self.handle), stream()),
                  "pr_plan_create_multi")

    @classmethod
    def batch(cls, graphs, edges_per_block=0, short_row_max=0, persistent=0, hot_vertices=0):
        """Plan over separate graphs (each a DeviceGraph with local ids): graph k
        owns plan vertices [vbase[k], vbase[k+1]) (self.vbase).  The total edge
        count may exceed 2^32; the graphs may be freed after this call."""
        self = cls.__new__(cls)
        self.graph = None
        self.handle = c_vp()
        tune = PrTune(edges_per_block, short_row_max, persistent, hot_vertices)
        arr = (DGraph * len(graphs))(*[g.dg for g in graphs])
        check(lib().gchi_pr_plan_create_batch(arr, len(graphs), ctypes.byref(tune), ctypes.byref(self.handle),
                                              stream()), "pr_plan_create_batch")
        self.ngraphs = len(graphs)
        self.vbase = np.concatenate([[0], np.cumsum([g.N for g in graphs])]).astype(np.uint64)
        self.N = int(self.vbase[-1])
        return self

    def info(self):
        inf = PrPlanInfo()
        check(lib().gchi_pr_plan_info_get(self.handle, ctypes.byref(inf)))
        return {k: getattr(inf, k) for k, _ in PrPlanInfo._fields_}

    def close(self):
        if self.handle:
            lib().gchi_pr_plan_destroy(self.handle)
            self.handle = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init(self, init, x, r):
        check(lib().gchi_pr_init(self.handle, init, dptr(x), dptr(r), stream()), "pr_init")

    def init_seeds(self, seed_bits, seed_value, x, r):
        check(lib().gchi_pr_init_seeds(self.handle, dptr(seed_bits), seed_value, dptr(x), dptr(r), stream()),
              "pr_init_seeds")

    def sweep(self, op, x_in, r_in, x_out, r_out, stats):
        check(lib().gchi_pr_sweep(self.handle, ctypes.byref(op), dptr(x_in), dptr(r_in), dptr(x_out), dptr(r_out),
                                  dptr(stats), stream()), "pr_sweep")


class PageRankExec:
    """init + nsweeps sweeps captured once into a HIP graph."""

    def __init__(self, plan, op, init, nsweeps, do_init=True, x=None, r=None):
        torch = _torch()
        N = plan.graph.N if plan.graph is not None else plan.N
        self.plan, self.nsweeps = plan, nsweeps
        self.x = x if x is not None else [torch.empty(N, dtype=torch.float32, device="cuda") for _ in range(2)]
        self.r = r if r is not None else [torch.empty(N, dtype=torch.float32, device="cuda") for _ in range(2)]
        self.stats = torch.zeros(max(2 * plan.ngraphs * nsweeps, 2), dtype=torch.float64, device="cuda")
        self.op = op
        self.handle = c_vp()
        check(lib().gchi_pr_exec_create_ex(plan.handle, ctypes.byref(op), init, 1 if do_init else 0, nsweeps,
                                           dptr(self.x[0]), dptr(self.x[1]), dptr(self.r[0]), dptr(self.r[1]),
                                           dptr(self.stats), ctypes.byref(self.handle)), "pr_exec_create")

    def launch(self):
        check(lib().gchi_pr_exec_launch(self.handle, stream()), "pr_exec_launch")

    def ranks(self):
        return self.r[self.nsweeps & 1]

    def close(self):
        if self.handle:
            lib().gchi_pr_exec_destroy(self.handle)
            self.handle = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PageRankBlocked:
    """Semi-synchronous sweeps over vertex blocks [bstart[j], bstart[j+1])
    (gchi_pr_blocked_*): x is updated in place block by block."""

    def __init__(self, graph, bstart, edges_per_block=0, short_row_max=0, persistent=0):
        self.graph = graph
        self.bstart = np.ascontiguousarray(bstart, dtype=np.uint32)
        self.handle = c_vp()
        tune = PrTune(edges_per_block, short_row_max, persistent, 0)
        check(lib().gchi_pr_blocked_create(ctypes.byref(graph.dg), c_vp(self.bstart.ctypes.data),
                                           len(self.bstart) - 1, ctypes.byref(tune), ctypes.byref(self.handle),
                                           stream()), "pr_blocked_create")

    def run(self, op, init, nsweeps, seed_bits=None, seed_value=0.0):
        """init then nsweeps sweeps; returns (ranks, stats[nsweeps, 2])."""
        torch = _torch()
        n = self.graph.N
        x = torch.empty(n, dtype=torch.float32, device="cuda")
        r = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(2)]
        stats = torch.zeros(max(2 * nsweeps, 2), dtype=torch.float64, device="cuda")
        check(lib().gchi_pr_blocked_init(self.handle, init, dptr(seed_bits), seed_value, dptr(x), dptr(r[0]),
                                         stream()), "pr_blocked_init")
        for k in range(nsweeps):
            check(lib().gchi_pr_blocked_sweep(self.handle, ctypes.byref(op), dptr(x), dptr(r[k & 1]),
                                              dptr(r[(k + 1) & 1]), c_vp(stats.data_ptr() + 16 * k), stream()),
                  "pr_blocked_sweep")
        return r[nsweeps & 1], stats[:2 * nsweeps].reshape(-1, 2)

    def close(self):
        if self.handle:
            lib().gchi_pr_blocked_destroy(self.handle)
            self.handle = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- connected components
CC_SYMMETRIC = 1


def cc_minlabel(src, dst, n, symmetric=False):
    """Minimum-id label of every weakly connected component.  symmetric=True
    promises that every edge is listed in both directions (only u < v is used)."""
    L = lib()
    labels = empty_u32(n)
    nb = c_sz()
    check(L.gchi_cc_workspace(n, ctypes.byref(nb)))
    ws = workspace(nb.value)
    e = src.numel() if src is not None else 0
    check(L.gchi_cc_minlabel_ex(dptr(src), dptr(dst), e, n, CC_SYMMETRIC if symmetric else 0, dptr(labels),
                                dptr(ws), nb.value, stream()), "cc_minlabel")
    return labels


def cc_labelprop(src, dst, n, max_sweeps=100000):
    """Synchronous min-label propagation until no label changes; returns
    (labels, sweeps including the final no-change sweep)."""
    torch = _torch()
    L = lib()
    lab = [empty_u32(n), empty_u32(n)]
    changed = torch.zeros(1, dtype=torch.int64, device="cuda")
    check(L.gchi_cc_labelprop_init(dptr(lab[0]), n, stream()))
    e = src.numel()
    for k in range(max_sweeps):
        check(L.gchi_cc_labelprop_sweep(dptr(src), dptr(dst), e, n, dptr(lab[k & 1]), dptr(lab[(k + 1) & 1]),
                                        dptr(changed), stream()), "cc_labelprop_sweep")
        if int(changed.item()) == 0:
            return lab[(k + 1) & 1], k + 1
    raise GchiError("cc_labelprop: no fixed point within max_sweeps")


def cc_pull(off, adj, n, max_sweeps=100000):
    """Synchronous min-label sweeps pulled over a symmetric CSR until no label
    drops; returns (labels, sweeps)."""
    torch = _torch()
    L = lib()
    a = empty_u32(n)
    b = empty_u32(n)
    check(L.gchi_cc_labelprop_init(dptr(a), n, stream()), "cc_labelprop_init")
    ch = torch.zeros(1, dtype=torch.int64, device="cuda")
    sweeps = 0
    while sweeps < max_sweeps:
        check(L.gchi_cc_pull_sweep(dptr(off), dptr(adj), n, dptr(a), dptr(b), dptr(ch), stream()), "cc_pull_sweep")
        sweeps += 1
        a, b = b, a
        if int(ch.item()) == 0:
            break
    return a, sweeps


def psw_graph_build(src, dst, n):
    """Adjacency of the generic update() engine: (out_off, out_dst, in_off,
    in_src, in_eidx, lower) device tensors (gchi_psw_graph_build)."""
    L = lib()
    e = int(src.numel())
    out_off, in_off, lower = empty_u32(n + 1), empty_u32(n + 1), empty_u32(n)
    out_dst, in_src, in_eidx = empty_u32(max(e, 1)), empty_u32(max(e, 1)), empty_u32(max(e, 1))
    nb = c_sz()
    check(L.gchi_psw_graph_workspace(e, n, ctypes.byref(nb)))
    ws = workspace(nb.value)
    check(L.gchi_psw_graph_build(dptr(src), dptr(dst), e, n, dptr(out_off), dptr(out_dst), dptr(in_off), dptr(in_src),
                                 dptr(in_eidx), dptr(lower), dptr(ws), nb.value, stream()), "psw_graph_build")
    work_status(ws, "psw_graph_build")
    return out_off, out_dst[:e], in_off, in_src[:e], in_eidx[:e], lower


def cc_afforest(off, adj, n):
    """Min-id component labels from a symmetric CSR by sampled union-find."""
    L = lib()
    labels = empty_u32(n)
    nb = c_sz()
    check(L.gchi_cc_afforest_workspace(n, ctypes.byref(nb)))
    ws = workspace(nb.value)
    check(L.gchi_cc_afforest(dptr(off), dptr(adj), n, dptr(labels), dptr(ws), nb.value, stream()), "cc_afforest")
    return labels


def cc_scheduled(off, adj, n, max_rounds=0):
    """Selective-scheduling min-label components over a symmetric CSR
    (off[n+1], adj[E]); returns (labels, rounds, vertex evaluations)."""
    L = lib()
    labels = empty_u32(n)
    e = int(adj.numel()) if adj is not None else 0
    nb = c_sz()
    check(L.gchi_cc_scheduled_workspace(n, e, ctypes.byref(nb)))
    ws = workspace(nb.value)
    rounds, evals = c_u32(), c_u64()
    check(L.gchi_cc_scheduled(dptr(off), dptr(adj), n, max_rounds, dptr(labels), ctypes.byref(rounds),
                              ctypes.byref(evals), dptr(ws), nb.value, stream()), "cc_scheduled")
    return labels, rounds.value, evals.value


def shard_decode(d_adj, nbytes):
    """Device decode of one shard's .adj stream (a uint8 device tensor);
    returns (src, dst, vend) device tensors / int."""
    L = lib()
    nb = c_sz()
    check(L.gchi_shard_decode_workspace(nbytes, ctypes.byref(nb)))
    ws = workspace(nb.value)
    ne, vend = c_u64(), c_u64()
    cap = nbytes // 4  # every edge takes 4 bytes of the stream
    src, dst = empty_u32(cap), empty_u32(cap)
    check(L.gchi_shard_decode(dptr(d_adj), nbytes, dptr(src), dptr(dst), cap, ctypes.byref(ne), ctypes.byref(vend),
                              dptr(ws), nb.value, stream()), "shard_decode")
    return src[:ne.value], dst[:ne.value], vend.value


def shard_decode_many(streams):
    """Device decode of several .adj streams [(uint8 device tensor, nbytes), ...]
    in one call; returns (src, dst, per-stream edge counts, per-stream vend)."""
    L = lib()
    k = len(streams)
    nb = np.array([int(n) for _, n in streams], np.uint64)
    ptrs = np.array([t.data_ptr() if n else 0 for t, n in streams], np.uint64)
    wsz = c_sz()
    check(L.gchi_shard_decode_many_workspace(c_vp(nb.ctypes.data), k, ctypes.byref(wsz)))
    ws = workspace(wsz.value)
    cap = int(sum(int(n) // 4 for n in nb))
    src, dst = empty_u32(cap), empty_u32(cap)
    ne, vend = np.zeros(max(k, 1), np.uint64), np.zeros(max(k, 1), np.uint64)
    check(L.gchi_shard_decode_many(c_vp(ptrs.ctypes.data), c_vp(nb.ctypes.data), k, dptr(src), dptr(dst), cap,
                                   c_vp(ne.ctypes.data), c_vp(vend.ctypes.data), dptr(ws), wsz.value, stream()),
          "shard_decode_many")
    tot = int(ne[:k].sum())
    return src[:tot], dst[:tot], ne[:k].astype(np.int64), vend[:k].astype(np.int64)


# ---------------------------------------------------------------- triangle counting
class TrianglePrep:
    """Degree order + oriented CSR (device)."""

    def __init__(self, src, dst, n):
        L = lib()
        e = src.numel()
        self.n = n
        self.vertexmap = empty_u32(n)
        self.off = empty_u32(n + 1)
        self.adj = empty_u32(e)
        nb = c_sz()
        check(L.gchi_tc_prepare_workspace(e, n, ctypes.byref(nb)))
        ws = workspace(nb.value)
        m = c_u64()
        check(L.gchi_tc_prepare(dptr(src), dptr(dst), e, n, dptr(self.vertexmap), dptr(self.off), dptr(self.adj),
                                ctypes.byref(m), dptr(ws), nb.value, stream()), "tc_prepare")
        self.noriented = m.value

    def count(self):
        torch = _torch()
        L = lib()
        pv = empty_u32(self.n)
        tot = torch.zeros(1, dtype=torch.int64, device="cuda")
        ovf = torch.zeros(1, dtype=torch.int32, device="cuda")
        nb = c_sz()
        check(L.gchi_tc_count_workspace(self.n, ctypes.byref(nb)))
        ws = workspace(nb.value)
        check(L.gchi_tc_count(dptr(self.off), dptr(self.adj), self.n, dptr(pv), dptr(tot), dptr(ovf), dptr(ws),
                              nb.value, stream()), "tc_count")
        return pv, tot, ovf


# ---------------------------------------------------------------- sharder
def shard_auto_count(nedges, edge_value_bytes, membudget_mb=800):
    n = c_i32()
    check(lib().gchi_shard_auto_count(nedges, edge_value_bytes, membudget_mb, ctypes.byref(n)))
    return n.value


class Shards:
    def __init__(self, src, dst, max_id, nshards):
        self.handle = c_vp()
        self.nshards_req = nshards
        check(lib().gchi_shard_build(dptr(src), dptr(dst), src.numel() if src is not None else 0, max_id, nshards,
                                     ctypes.byref(self.handle), stream()), "shard_build")
        ns = c_i32()
        check(lib().gchi_shard_info(self.handle, ctypes.byref(ns), None, None))
        self.ends = np.zeros(ns.value, np.uint32)
        self.counts = np.zeros(ns.value, np.uint64)
        check(lib().gchi_shard_info(self.handle, ctypes.byref(ns), self.ends.ctypes.data, self.counts.ctypes.data))
        self.nshards = ns.value

    def edges(self, p):
        s, d, c = c_vp(), c_vp(), c_u64()
        check(lib().gchi_shard_edges(self.handle, p, ctypes.byref(s), ctypes.byref(d), ctypes.byref(c)))
        return s.value, d.value, c.value

    def edges_host(self, p):
        torch = _torch()
        s, d, c = self.edges(p)
        hs, hd = np.zeros(c, np.uint32), np.zeros(c, np.uint32)
        if c:
            check(lib().gchi_memcpy(c_vp(hs.ctypes.data), c_vp(s), 4 * c, stream()), "memcpy")
            check(lib().gchi_memcpy(c_vp(hd.ctypes.data), c_vp(d), 4 * c, stream()), "memcpy")
        return hs, hd

    def encode(self, p, max_index=4096):
        L = lib()
        nbytes = c_u64()
        check(L.gchi_shard_encode(self.handle, p, None, ctypes.byref(nbytes), None, 0, None, stream()))
        adj = np.zeros(max(nbytes.value, 1), np.uint8)
        idx = np.zeros(3 * max_index, np.uint64)
        ni = c_u32()
        check(L.gchi_shard_encode(self.handle, p, adj.ctypes.data, ctypes.byref(nbytes), idx.ctypes.data, max_index,
                                  ctypes.byref(ni), stream()), "shard_encode")
        return adj[:nbytes.value].tobytes(), idx[:3 * ni.value].reshape(-1, 3)

    def write_files(self, base, edge_value_bytes, zlib=False):
        check(lib().gchi_shard_write_files_ex(self.handle, base.encode(), edge_value_bytes, 1 if zlib else 0,
                                              stream()), "shard_write_files")

    def close(self):
        if self.handle:
            lib().gchi_shard_destroy(self.handle)
            self.handle = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- personalized PageRank
def ppr_workload_queries(n, nqueries=1024, seed=5, max_seeds=64):
    """The batched-PPR workload: nqueries restart sets, each of 1..max_seeds
    distinct vertices drawn uniformly from [0, n)."""
    rng = np.random.default_rng(seed)
    return [rng.choice(n, size=int(rng.integers(1, max_seeds + 1)), replace=False).astype(np.uint32)
            for _ in range(nqueries)]


def multigraph_workload(ngraphs=2048, seed=9):
    """The many-graphs workload: graph i is an R-MAT graph of scale
    scales[i] (uniform 10..20, so vertices log-uniform 2^10..2^20), edge
    factor efs[i] (uniform 8..32), seed 10000 + i, self-loops dropped.
    Returns [(scale, nedges_generated, seed)]."""
    rng = np.random.default_rng(seed)
    scales = rng.integers(10, 21, ngraphs)
    efs = rng.integers(8, 33, ngraphs)
    return [(int(scales[i]), int(efs[i]) << int(scales[i]), 10000 + i) for i in range(ngraphs)]


def seed_csr(queries):
    off = np.zeros(len(queries) + 1, np.uint32)
    for i, q in enumerate(queries):
        off[i + 1] = off[i] + len(q)
    flat = np.concatenate([np.asarray(q, np.uint32) for q in queries]) if queries else np.zeros(0, np.uint32)
    return off, np.ascontiguousarray(flat, dtype=np.uint32)


class PprBatch:
    """Batched personalized PageRank over a PageRankPlan (64 queries per block)."""

    def __init__(self, plan, nqueries):
        self.plan, self.nq = plan, nqueries
        self.handle = c_vp()
        check(lib().gchi_ppr_batch_create(plan.handle, nqueries, ctypes.byref(self.handle)), "ppr_batch_create")

    def run(self, queries, damping=0.85, tol=1e-6, bf16_sweeps=60, bf16_switch_residual=2e-3, max_sweeps=1000,
            check_every=4):
        assert len(queries) == self.nq
        off, flat = seed_csr(queries)
        cfg = PprConfig(damping, tol, bf16_switch_residual, bf16_sweeps, max_sweeps, check_every, 0)
        res = PprResult()
        resid = np.zeros(self.nq, np.float64)
        check(lib().gchi_ppr_batch_run(self.handle, c_vp(off.ctypes.data), c_vp(flat.ctypes.data), ctypes.byref(cfg),
                                       ctypes.byref(res), c_vp(resid.ctypes.data), stream()), "ppr_batch_run")
        return {k: getattr(res, k) for k, _ in PprResult._fields_}, resid

    def ranks(self, q):
        torch = _torch()
        out = torch.empty(self.plan.graph.N, dtype=torch.float32, device="cuda")
        check(lib().gchi_ppr_batch_extract(self.handle, q, dptr(out), stream()), "ppr_batch_extract")
        return out

    def close(self):
        if self.handle:
            lib().gchi_ppr_batch_destroy(self.handle)
            self.handle = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlainPpr:
    """One query at a time with the FP32 PageRank sweep (seed bitmap apply):
    sweeps launched one per launch, residual read every check_every sweeps."""

    def __init__(self, plan, damping=0.85):
        torch = _torch()
        self.plan, self.d = plan, damping
        N = plan.graph.N
        self.x = [torch.empty(N, dtype=torch.float32, device="cuda") for _ in range(2)]
        self.r = [torch.empty(N, dtype=torch.float32, device="cuda") for _ in range(2)]
        self.stats = torch.zeros(2, dtype=torch.float64, device="cuda")

    def run(self, seeds, tol=1e-6, max_sweeps=1000, check_every=4):
        torch = _torch()
        n = self.plan.graph.N
        bits = seed_bitmap(seeds, n)
        mass = (1.0 - self.d) / len(seeds)
        op = pr_op(0.0, self.d, seed_bits=bits, seed_teleport=mass)
        self.plan.init_seeds(bits, mass, self.x[0], self.r[0])
        stop = tol * (1.0 - self.d) / self.d
        k = 0
        while k < max_sweeps:
            self.plan.sweep(op, self.x[k & 1], self.r[k & 1], self.x[(k + 1) & 1], self.r[(k + 1) & 1], self.stats)
            k += 1
            if k % check_every == 0 and float(self.stats[0].item()) <= stop:
                break
        torch.cuda.synchronize()
        return self.r[k & 1], k


# ---------------------------------------------------------------- many graphs
def union_of(graphs_list):
    """Disjoint union of (src, dst, n) host graphs: concatenated edges with
    vertex offsets, and the vertex bases [0, n0, n0+n1, ...]."""
    vbase = np.zeros(len(graphs_list) + 1, np.uint64)
    for i, (_, _, n) in enumerate(graphs_list):
        vbase[i + 1] = vbase[i] + n
    assert vbase[-1] < 2 ** 32
    src = np.concatenate([s.astype(np.uint64) + vbase[i] for i, (s, _, _) in enumerate(graphs_list)])
    dst = np.concatenate([d.astype(np.uint64) + vbase[i] for i, (_, d, _) in enumerate(graphs_list)])
    return src.astype(np.uint32), dst.astype(np.uint32), vbase.astype(np.uint32)


def autotune_pr(graph, sweeps=4, report=False):
    """Best gchi_pr_tune for the graph -> (dict, ms per sweep, [(dict, ms)...])."""
    best, ms, n = PrTune(), ctypes.c_double(), c_i32()
    tried = (PrTune * 128)()
    tms = (ctypes.c_double * 128)()
    check(lib().gchi_autotune_pr(ctypes.byref(graph.dg), sweeps, ctypes.byref(best), ctypes.byref(ms),
                                 ctypes.cast(tried, c_vp) if report else None,
                                 ctypes.cast(tms, c_vp) if report else None, 128, ctypes.byref(n), stream()),
          "autotune_pr")
    asd = lambda t: {"edges_per_block": t.edges_per_block, "short_row_max": t.short_row_max,  # noqa: E731
                     "persistent": t.persistent, "hot_vertices": t.hot_vertices}
    return asd(best), ms.value, [(asd(tried[i]), tms[i]) for i in range(n.value)] if report else []
