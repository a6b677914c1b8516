"""ctypes bindings for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the ``cpu_baseline`` leg of ``bench.py`` -- as the checker, never as
the thing measured or shipped.  The product package ``beedrill_amd`` does not
import this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
REF_HARNESS = os.path.join(_HERE, "_ref", "ref_harness")

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")


class LevelRec(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "level", "dir", "frontier", "frontier_deg", "unvisited", "unvisited_deg",
        "edges_to_check", "result", "next", "next_deg", "edges_examined",
        "max_frontier_deg")]


TD, BU = 0, 1


def build() -> None:
    """Compile liboracle.so (and oracle/_ref when the reference is mounted)."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)
    if os.path.isdir("/root/reference"):
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_sitmo_word.argtypes = [ctypes.c_uint32, ctypes.c_uint64]
        L.oracle_sitmo_word.restype = ctypes.c_uint32
        L.oracle_rmat_raw.argtypes = [ctypes.c_int64, ctypes.c_int64] + [ctypes.c_double] * 4 + [
            ctypes.c_uint32, _i64p, _i64p]
        L.oracle_flip_sort_dedup.argtypes = [ctypes.c_int64, _i64p, _i64p]
        L.oracle_flip_sort_dedup.restype = ctypes.c_int64
        L.oracle_vperm.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint64]
        L.oracle_vperm.restype = ctypes.c_int64
        L.oracle_shuffle_key.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.oracle_shuffle_key.restype = ctypes.c_uint64
        L.oracle_make_rmat.argtypes = [ctypes.c_int64, ctypes.c_int64] + [ctypes.c_double] * 4 + [
            ctypes.c_uint32, _i64p, _i64p, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_make_er.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32, _i64p, _i64p,
                                     ctypes.POINTER(ctypes.c_int64)]
        L.oracle_make_lattice.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_uint32, _i64p,
                                          _i64p, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_write_el64.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int64, _i64p, _i64p]
        L.oracle_build_csr.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, _i64p, _i64p, _i32p]
        L.oracle_check_graph.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, _i64p, _i64p, _i32p]
        L.oracle_beamer.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, _i32p, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_int64, _i64p, _i32p,
                                    ctypes.c_void_p, ctypes.c_int64]
        L.oracle_beamer.restype = ctypes.c_int64
        L.oracle_beamer_omp.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, _i32p,
                                        ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, _i64p,
                                        np.ctypeslib.ndpointer(dtype=np.int8, flags="C_CONTIGUOUS"),
                                        ctypes.c_int64]
        L.oracle_beamer_omp.restype = ctypes.c_int64
        L.oracle_omp_threads.restype = ctypes.c_int
        L.oracle_td_only.argtypes = [ctypes.c_int64, ctypes.c_int64, _i64p, _i32p, ctypes.c_int64,
                                     _i64p, _i32p]
        L.oracle_td_only.restype = ctypes.c_int64
        L.oracle_bfs_depth.argtypes = [ctypes.c_int64, _i64p, _i32p, ctypes.c_int64, _i32p]
        L.oracle_check.argtypes = [ctypes.c_int64, _i64p, _i32p, ctypes.c_int64, _i64p]
        L.oracle_count_traversed.argtypes = [ctypes.c_int64, _i64p, _i64p]
        L.oracle_count_traversed.restype = ctypes.c_int64
        L.oracle_lcg_init.argtypes = [ctypes.c_uint64]
        L.oracle_lcg_init.restype = ctypes.c_uint64
        L.oracle_lcg_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.oracle_lcg_next.restype = ctypes.c_uint64
        L.oracle_pick_roots.argtypes = [ctypes.c_int64, _i64p, ctypes.c_int64, _i64p]
        L.oracle_aggregate.argtypes = [_f64p, ctypes.c_int64, _f64p]
        L.oracle_set_mutation.argtypes = [ctypes.c_int]
        L.oracle_head_index.argtypes = [ctypes.c_int64, _i64p, _i32p, _i32p]
        L.oracle_set_bu_head.argtypes = [ctypes.c_void_p]
        L.oracle_mlp_nparams.argtypes = [ctypes.c_int, ctypes.c_int]
        L.oracle_mlp_nparams.restype = ctypes.c_int64
        L.oracle_mlp_forward.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p, ctypes.c_int64,
                                         _f32p]
        L.oracle_mlp_loss_grad.argtypes = [_f32p, ctypes.c_int, ctypes.c_int, _f32p, _f32p,
                                           ctypes.c_void_p, ctypes.c_int64, _f32p]
        L.oracle_mlp_loss_grad.restype = ctypes.c_float
        L.oracle_mlp_adam.argtypes = [_f32p, _f32p, _f32p, _f32p, ctypes.c_int64, ctypes.c_float,
                                      ctypes.c_int64]
        L.oracle_cc.argtypes = [ctypes.c_int64, _i64p, _i32p, _i64p, ctypes.POINTER(ctypes.c_int64)]
        L.oracle_cc.restype = ctypes.c_int64
        L.oracle_cc_check.argtypes = [ctypes.c_int64, _i64p, _i32p, _i64p]
        _lib = L
    return _lib


# ----------------------------------------------------------------- generator

def sitmo_words(seed: int, n: int, start: int = 0) -> np.ndarray:
    L = lib()
    return np.array([L.oracle_sitmo_word(seed, start + i) for i in range(n)], dtype=np.uint32)


def rmat_raw(nv, ne, a=0.57, b=0.19, c=0.19, d=0.05, seed=0):
    src = np.empty(ne, np.int64)
    dst = np.empty(ne, np.int64)
    if lib().oracle_rmat_raw(nv, ne, a, b, c, d, seed, src, dst):
        raise ValueError("bad rmat args")
    return src, dst


def flip_sort_dedup(src, dst):
    s = np.ascontiguousarray(src, np.int64).copy()
    d = np.ascontiguousarray(dst, np.int64).copy()
    m = lib().oracle_flip_sort_dedup(len(s), s, d)
    return s[:m], d[:m]


def vperm(x: int, nv: int, seed: int) -> int:
    return lib().oracle_vperm(x, nv, seed)


def shuffle_key(j: int, seed: int) -> int:
    return lib().oracle_shuffle_key(j, seed)


def make_rmat(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, d=0.05, seed=0):
    nv = 1 << scale
    ne = edge_factor * nv
    src = np.empty(ne, np.int64)
    dst = np.empty(ne, np.int64)
    m = ctypes.c_int64(0)
    if lib().oracle_make_rmat(nv, ne, a, b, c, d, seed, src, dst, ctypes.byref(m)):
        raise ValueError("bad rmat args")
    return nv, src[:m.value].copy(), dst[:m.value].copy()


def make_er(nv, ne, seed=0):
    src = np.empty(ne, np.int64)
    dst = np.empty(ne, np.int64)
    m = ctypes.c_int64(0)
    if lib().oracle_make_er(nv, ne, seed, src, dst, ctypes.byref(m)):
        raise ValueError("bad er args")
    return nv, src[:m.value].copy(), dst[:m.value].copy()


def make_lattice(rows, cols, seed=0):
    ne = 2 * rows * cols
    src = np.empty(max(ne, 1), np.int64)
    dst = np.empty(max(ne, 1), np.int64)
    m = ctypes.c_int64(0)
    if lib().oracle_make_lattice(rows, cols, seed, src, dst, ctypes.byref(m)):
        raise ValueError("bad lattice args")
    return rows * cols, src[:m.value].copy(), dst[:m.value].copy()


def write_el64(path, nv, src, dst):
    s = np.ascontiguousarray(src, np.int64)
    d = np.ascontiguousarray(dst, np.int64)
    if lib().oracle_write_el64(path.encode(), nv, len(s), s, d):
        raise OSError(path)


# --------------------------------------------------------------------- graph

def build_csr(nv, src, dst):
    s = np.ascontiguousarray(src, np.int64)
    d = np.ascontiguousarray(dst, np.int64)
    off = np.empty(nv + 1, np.int64)
    adj = np.empty(max(2 * len(s), 1), np.int32)
    if lib().oracle_build_csr(nv, len(s), s, d, off, adj):
        raise ValueError("edge list has an out-of-range id")
    return off, adj[:2 * len(s)]


def check_graph(nv, src, dst, off, adj) -> bool:
    s = np.ascontiguousarray(src, np.int64)
    d = np.ascontiguousarray(dst, np.int64)
    return bool(lib().oracle_check_graph(nv, len(s), s, d, off, np.ascontiguousarray(adj, np.int32)))


# ----------------------------------------------------------------------- BFS

def head_index(off, adj):
    """Bottom-up head index: per vertex the neighbour of largest degree (first
    in CSR order among equals), -1 for degree 0."""
    nv = len(off) - 1
    head = np.empty(nv, np.int32)
    lib().oracle_head_index(nv, off, np.ascontiguousarray(adj, np.int32), head)
    return head


def beamer(off, adj, m, source, alpha=15, beta=18, log_cap=1 << 16, head=None):
    """Returns (parent[int64, reference encoding], depth[int32], levels[list]).
    With `head` (head_index), bottom-up edges_examined counts the head-first
    scan of the product's kernel."""
    nv = len(off) - 1
    parent = np.empty(nv, np.int64)
    depth = np.empty(nv, np.int32)
    recs = (LevelRec * log_cap)()
    L = lib()
    h = None if head is None else np.ascontiguousarray(head, np.int32)
    L.oracle_set_bu_head(None if h is None else h.ctypes.data_as(ctypes.c_void_p))
    try:
        n = L.oracle_beamer(nv, m, off, np.ascontiguousarray(adj, np.int32), source, alpha, beta,
                            parent, depth, ctypes.cast(recs, ctypes.c_void_p), log_cap)
    finally:
        L.oracle_set_bu_head(None)
    if n < 0:
        raise ValueError("bad beamer args")
    levels = [{f: getattr(recs[i], f) for f, _ in LevelRec._fields_} for i in range(min(n, log_cap))]
    return parent, depth, levels


def beamer_timed(off, adj, m, source, alpha=15, beta=18):
    """run_beamer alone (no level log), for timing."""
    nv = len(off) - 1
    parent = np.empty(nv, np.int64)
    n = lib().oracle_beamer(nv, m, off, adj, source, alpha, beta, parent,
                            np.empty(nv, np.int32), None, 0)
    if n < 0:
        raise ValueError("bad beamer args")
    return parent


def beamer_omp(off, adj, m, source, alpha=15, beta=18):
    """run_beamer on every host core (oracle_bfs_omp.c): (parent, dirs).
    For timing the all-core CPU baseline; directions equal beamer()'s."""
    nv = len(off) - 1
    parent = np.empty(nv, np.int64)
    dirs = np.empty(nv + 2, np.int8)
    n = lib().oracle_beamer_omp(nv, m, off, np.ascontiguousarray(adj, np.int32), source, alpha,
                                beta, parent, dirs, len(dirs))
    if n < 0:
        raise ValueError("bad beamer args")
    return parent, dirs[:n].copy()


def omp_threads():
    return int(lib().oracle_omp_threads())


def td_only(off, adj, m, source):
    nv = len(off) - 1
    parent = np.empty(nv, np.int64)
    depth = np.empty(nv, np.int32)
    lib().oracle_td_only(nv, m, off, np.ascontiguousarray(adj, np.int32), source, parent, depth)
    return parent, depth


def bfs_depth(off, adj, source):
    nv = len(off) - 1
    depth = np.empty(nv, np.int32)
    lib().oracle_bfs_depth(nv, off, np.ascontiguousarray(adj, np.int32), source, depth)
    return depth


def check(off, adj, source, parent) -> bool:
    nv = len(off) - 1
    return bool(lib().oracle_check(nv, off, np.ascontiguousarray(adj, np.int32), source,
                                   np.ascontiguousarray(parent, np.int64)))


def count_traversed(off, parent) -> int:
    return lib().oracle_count_traversed(len(off) - 1, off, np.ascontiguousarray(parent, np.int64))


def lcg_outputs(step: int, n: int) -> np.ndarray:
    L = lib()
    st = ctypes.c_uint64(L.oracle_lcg_init(step))
    return np.array([L.oracle_lcg_next(ctypes.byref(st)) for _ in range(n)], dtype=np.uint64)


def pick_roots(off, n):
    roots = np.empty(n, np.int64)
    lib().oracle_pick_roots(len(off) - 1, off, n, roots)
    return roots


class mutation:
    """Context manager planting mistake `id` in the oracle (tests only)."""

    def __init__(self, mid):
        self.mid = mid

    def __enter__(self):
        lib().oracle_set_mutation(self.mid)
        return self

    def __exit__(self, *a):
        lib().oracle_set_mutation(0)


def aggregate(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.empty(6, np.float64)
    lib().oracle_aggregate(x, len(x), out)
    return dict(zip(("sum", "mean", "hmean", "stddev", "min", "max"), out.tolist()))


# ----------------------------------------------------------------------- MLP

def mlp_nparams(h1=32, h2=32):
    return lib().oracle_mlp_nparams(h1, h2)


def mlp_forward(params, x, h1=32, h2=32):
    x = np.ascontiguousarray(x, np.float32)
    y = np.empty((x.shape[0], 2), np.float32)
    lib().oracle_mlp_forward(np.ascontiguousarray(params, np.float32), h1, h2, x, x.shape[0], y)
    return y


def mlp_loss_grad(params, x, target, mask=None, h1=32, h2=32):
    x = np.ascontiguousarray(x, np.float32)
    t = np.ascontiguousarray(target, np.float32)
    g = np.empty(mlp_nparams(h1, h2), np.float32)
    mk = None
    if mask is not None:
        mk_arr = np.ascontiguousarray(mask, np.float32)
        mk = mk_arr.ctypes.data_as(ctypes.c_void_p)
    loss = lib().oracle_mlp_loss_grad(np.ascontiguousarray(params, np.float32), h1, h2, x, t, mk,
                                      x.shape[0], g)
    return float(loss), g


def mlp_adam(params, grad, m, v, lr, t):
    lib().oracle_mlp_adam(params, np.ascontiguousarray(grad, np.float32), m, v, len(params), lr, t)


# --------------------------------------------------------- connected components

def components(off, adj):
    """components::run (components.cc:38-104): (labels, num_components,
    num_iters); each component is labelled by its smallest vertex id."""
    nv = len(off) - 1
    lab = np.empty(nv, np.int64)
    n = ctypes.c_int64(0)
    it = lib().oracle_cc(nv, np.ascontiguousarray(off, np.int64),
                         np.ascontiguousarray(adj, np.int32), lab, ctypes.byref(n))
    return lab, n.value, it


def components_check(off, adj, labels) -> bool:
    """components::check (components.cc:152-202)."""
    return bool(lib().oracle_cc_check(len(off) - 1, np.ascontiguousarray(off, np.int64),
                                      np.ascontiguousarray(adj, np.int32),
                                      np.ascontiguousarray(labels, np.int64)))
