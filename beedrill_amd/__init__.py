"""beedrill_amd -- MI355X (gfx950) direction-optimizing BFS.

Python bindings over the C ABI in ``include/beedrill_amd.h`` (ctypes).  The
library ``beedrill_amd/libbeedrill_amd.so`` must be built (``make -C
beedrill_amd``); importing this package without it raises, and every call
runs on the GPU -- there is no CPU fallback.

Reference interface mirrored (ehein6/beedrill):
  * ``Context.bfs``            -> hybrid_bfs::run_beamer / run_with_migrating_threads
                                 (hybrid_bfs.cc:152-213), preceded by clear()
  * ``Context.check``          -> hybrid_bfs::check (hybrid_bfs.cc:281-379)
  * ``Context.count_traversed`` -> count_num_traversed_edges (hybrid_bfs.cc:401-414)
  * ``Context.load_el64``      -> dist_edge_list::load_binary + create_graph_from_edge_list
  * ``Context.pick_roots``     -> pick_random_vertex (hybrid_bfs_main.cc:136-144)
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbeedrill_amd.so")

POLICY_BEAMER, POLICY_LEARNED, POLICY_TD_ONLY, POLICY_BU_ONLY, POLICY_SCRIPT = range(5)
POLICIES = {"beamer": 0, "learned": 1, "td_only": 2, "bu_only": 3, "script": 4,
            "remote_writes": 5, "remote_writes_hybrid": 6}
DIR_TD, DIR_BU = 0, 1
GEN_RMAT, GEN_ER, GEN_LATTICE = 0, 1, 2
STATUS = {0: "OK", -1: "EINVAL", -2: "ENOMEM", -3: "EHIP", -4: "ESTATE", -5: "EIO",
          -6: "ERANGE", -7: "ENOEXT"}

_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


class LevelRec(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "level", "dir", "frontier", "frontier_deg", "unvisited", "unvisited_deg",
        "edges_to_check", "result", "next", "next_deg", "edges_examined",
        "max_frontier_deg", "tasks", "queued")] + [
        ("ms", ctypes.c_double), ("ms_other", ctypes.c_double),
        ("launch", ctypes.c_int64), ("launches", ctypes.c_int64)]


class CCStats(ctypes.Structure):
    _fields_ = [("num_components", ctypes.c_int64), ("num_iters", ctypes.c_int64),
                ("ms", ctypes.c_double), ("pass_ms", ctypes.c_double),
                ("labels_read", ctypes.c_int64), ("edge_bytes", ctypes.c_int64),
                ("pass_launches", ctypes.c_int64), ("prep_ms", ctypes.c_double),
                ("prep_bytes", ctypes.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class RunStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int64) for n in (
        "source", "steps", "td_steps", "bu_steps", "reached", "traversed_edges")] + [
        ("ms", ctypes.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


TUNING_FIELDS = ("block", "grid_per_cu", "wave_threshold", "task_threshold", "task_edges",
                 "bu_words", "bu_cap", "level_driver", "time_steps", "bu_head", "step_kernels",
                 "bu_grid_per_cu", "learned_margin_milli", "ms_alpha", "learned_min_skew",
                 "td_split_edges", "vertex_order", "bu_batch", "learned_near")


class Tuning(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in TUNING_FIELDS]

    def as_dict(self):
        return {f: getattr(self, f) for f in TUNING_FIELDS}


class BeedrillError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libbeedrill_amd.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C beedrill_amd` "
                          "(the HIP library is required; there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64, i32, f32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_float
    sig = {
        "bd_device_count": ([], ctypes.c_int),
        "bd_create": ([ctypes.POINTER(vp), ctypes.c_int], ctypes.c_int),
        "bd_destroy": ([vp], ctypes.c_int),
        "bd_last_error": ([vp], ctypes.c_char_p),
        "bd_graph_upload": ([vp, _i64p, _i32p, i64, i64, i64], ctypes.c_int),
        "bd_graph_from_edges": ([vp, _i64p, _i64p, i64, i64], ctypes.c_int),
        "bd_graph_load_el64": ([vp, ctypes.c_char_p], ctypes.c_int),
        "bd_graph_generate": ([vp, ctypes.c_int, i64, i64, ctypes.c_double, ctypes.c_double,
                               ctypes.c_double, ctypes.c_double, ctypes.c_uint32,
                               ctypes.POINTER(i64)], ctypes.c_int),
        "bd_graph_info": ([vp, ctypes.POINTER(i64), ctypes.POINTER(i64), ctypes.POINTER(i64)],
                          ctypes.c_int),
        "bd_graph_download": ([vp, vp, vp], ctypes.c_int),
        "bd_graph_edges_download": ([vp, vp, vp], ctypes.c_int),
        "bd_graph_save_el64": ([vp, ctypes.c_char_p], ctypes.c_int),
        "bd_pick_roots": ([vp, i64, _i64p], ctypes.c_int),
        "bd_bfs_run": ([vp, i64, i64, i64, ctypes.c_int, ctypes.POINTER(RunStats)], ctypes.c_int),
        "bd_bfs_set_script": ([vp, _i32p, i64], ctypes.c_int),
        "bd_bfs_fetch": ([vp, vp, vp], ctypes.c_int),
        "bd_bfs_levels": ([vp, vp, i64, ctypes.POINTER(i64)], ctypes.c_int),
        "bd_bfs_count_traversed": ([vp, ctypes.POINTER(i64)], ctypes.c_int),
        "bd_bfs_check": ([vp, i64, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "bd_bfs_check_parents": ([vp, i64, _i64p, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "bd_bfs_profile": ([vp, i64, i64, i64, ctypes.c_int, ctypes.POINTER(RunStats)],
                           ctypes.c_int),
        "bd_set_tuning": ([vp, ctypes.POINTER(Tuning)], ctypes.c_int),
        "bd_get_tuning": ([vp, ctypes.POINTER(Tuning)], ctypes.c_int),
        "bd_bfs_multi_root": ([vp, _i64p, ctypes.c_int, vp, ctypes.POINTER(ctypes.c_double),
                               ctypes.POINTER(i64)], ctypes.c_int),
        "bd_bfs_launch_times": ([vp, vp, i64, ctypes.POINTER(i64)], ctypes.c_int),
        "bd_copy_bandwidth": ([vp, i64, ctypes.c_int, ctypes.POINTER(ctypes.c_double)],
                              ctypes.c_int),
        "bd_bfs_multi_graph": ([vp, ctypes.c_int, _i64p, _i64p, vp,
                                ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "bd_bfs_multi_graph_static": ([vp, ctypes.c_int, vp, vp,
                                       ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
        "bd_cc_run": ([vp, ctypes.POINTER(CCStats)], ctypes.c_int),
        "bd_cc_fetch": ([vp, _i64p], ctypes.c_int),
        "bd_cc_check": ([vp, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "bd_mlp_set": ([vp, _f32p, ctypes.c_int, ctypes.c_int], ctypes.c_int),
        "bd_mlp_get": ([vp, vp, i64, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
                       ctypes.c_int),
        "bd_mlp_features": ([ctypes.POINTER(LevelRec), i64, i64, i64, _f32p], ctypes.c_int),
        "bd_mlp_infer": ([vp, _f32p, i64, _f32p, ctypes.c_int], ctypes.c_int),
        "bd_mlp_loss_grad": ([vp, _f32p, _f32p, vp, i64, _f32p, ctypes.POINTER(f32)],
                             ctypes.c_int),
        "bd_mlp_train": ([vp, _f32p, _f32p, vp, i64, ctypes.c_int, f32, ctypes.c_uint32,
                          ctypes.POINTER(f32)], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def union_csr(graphs):
    """Disjoint union of CSR graphs [(offsets, cols, m), ...]: graph g's
    vertices are renumbered from vbase[g].  Returns (offsets, cols, m, vbase)."""
    nvs = [len(o) - 1 for o, _, _ in graphs]
    vbase = np.concatenate([[0], np.cumsum(nvs)]).astype(np.int64)
    ebase = np.concatenate([[0], np.cumsum([len(c) for _, c, _ in graphs])]).astype(np.int64)
    off = np.empty(vbase[-1] + 1, np.int64)
    cols = np.empty(ebase[-1], np.int32)
    for g, (o, c, _) in enumerate(graphs):
        off[vbase[g]:vbase[g + 1]] = np.asarray(o[:-1], np.int64) + ebase[g]
        cols[ebase[g]:ebase[g + 1]] = np.asarray(c, np.int64) + vbase[g]
    off[-1] = ebase[-1]
    return off, cols, int(sum(m for _, _, m in graphs)), vbase


def exported_symbols():
    """Names declared in include/beedrill_amd.h."""
    import re
    hdr = os.path.join(os.path.dirname(_HERE), "include", "beedrill_amd.h")
    with open(hdr) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(bd_[a-z0-9_]+)\s*\(", text)))


def device_count() -> int:
    return lib().bd_device_count()


def _level_dicts(recs, n):
    return [{f: getattr(recs[i], f) for f, _ in LevelRec._fields_} for i in range(n)]


class Context:
    """One GPU, one stream, one resident graph (a ``hybrid_bfs`` instance)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.bd_create(ctypes.byref(h), device)
        if rc != 0:
            raise BeedrillError(rc, "bd_create failed (no usable GPU?)")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            self._L.bd_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _ok(self, rc):
        if rc != 0:
            raise BeedrillError(rc, self._L.bd_last_error(self._h).decode(errors="replace"))

    # ---------------------------------------------------------------- graphs
    def upload_csr(self, off, cols, m_undirected):
        off = np.ascontiguousarray(off, np.int64)
        cols = np.ascontiguousarray(cols, np.int32)
        if cols.size == 0:
            cols = np.zeros(1, np.int32)
            nnz = 0
        else:
            nnz = cols.size
        self._ok(self._L.bd_graph_upload(self._h, off, cols, len(off) - 1, nnz, m_undirected))

    def from_edges(self, nv, src, dst):
        s = np.ascontiguousarray(src, np.int64)
        d = np.ascontiguousarray(dst, np.int64)
        if s.size == 0:
            s = np.zeros(1, np.int64)
            d = np.zeros(1, np.int64)
            m = 0
        else:
            m = s.size
        self._ok(self._L.bd_graph_from_edges(self._h, s, d, nv, m))

    def load_el64(self, path):
        self._ok(self._L.bd_graph_load_el64(self._h, os.fsencode(path)))

    def save_el64(self, path):
        self._ok(self._L.bd_graph_save_el64(self._h, os.fsencode(path)))

    def generate(self, kind, nv, num_edges, a=0.57, b=0.19, c=0.19, d=0.05, seed=0):
        m = ctypes.c_int64(0)
        self._ok(self._L.bd_graph_generate(self._h, kind, nv, num_edges, a, b, c, d, seed,
                                           ctypes.byref(m)))
        return m.value

    def generate_rmat(self, scale, edge_factor=16, a=0.57, b=0.19, c=0.19, d=0.05, seed=0):
        return self.generate(GEN_RMAT, 1 << scale, edge_factor << scale, a, b, c, d, seed)

    def generate_er(self, nv, num_edges, seed=0):
        return self.generate(GEN_ER, nv, num_edges, seed=seed)

    def generate_lattice(self, rows, cols, seed=0):
        return self.generate(GEN_LATTICE, rows * cols, cols, seed=seed)

    def info(self):
        nv, nnz, m = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        self._ok(self._L.bd_graph_info(self._h, ctypes.byref(nv), ctypes.byref(nnz),
                                       ctypes.byref(m)))
        return nv.value, nnz.value, m.value

    def csr(self):
        nv, nnz, _ = self.info()
        off = np.empty(nv + 1, np.int64)
        cols = np.empty(max(nnz, 1), np.int32)
        self._ok(self._L.bd_graph_download(self._h, off.ctypes.data, cols.ctypes.data))
        return off, cols[:nnz]

    def edges(self):
        _, _, m = self.info()
        s = np.empty(max(m, 1), np.int64)
        d = np.empty(max(m, 1), np.int64)
        self._ok(self._L.bd_graph_edges_download(self._h, s.ctypes.data, d.ctypes.data))
        return s[:m], d[:m]

    def pick_roots(self, n):
        r = np.empty(max(n, 1), np.int64)
        self._ok(self._L.bd_pick_roots(self._h, n, r))
        return r[:n]

    # ------------------------------------------------------------------- BFS
    def bfs(self, source, alpha=15, beta=18, policy="beamer"):
        pol = POLICIES[policy] if isinstance(policy, str) else int(policy)
        st = RunStats()
        self._ok(self._L.bd_bfs_run(self._h, int(source), alpha, beta, pol, ctypes.byref(st)))
        return st.as_dict()

    def profile(self, source, alpha=15, beta=18, policy="beamer"):
        pol = POLICIES[policy] if isinstance(policy, str) else int(policy)
        st = RunStats()
        self._ok(self._L.bd_bfs_profile(self._h, int(source), alpha, beta, pol, ctypes.byref(st)))
        return st.as_dict()

    def set_script(self, dirs):
        d = np.ascontiguousarray(dirs, np.int32)
        self._ok(self._L.bd_bfs_set_script(self._h, d, d.size))

    def fetch(self, depth=True, parent=True):
        nv, _, _ = self.info()
        dp = np.empty(nv, np.int32) if depth else None
        pa = np.empty(nv, np.int64) if parent else None
        self._ok(self._L.bd_bfs_fetch(self._h, dp.ctypes.data if dp is not None else None,
                                      pa.ctypes.data if pa is not None else None))
        return dp, pa

    def levels(self):
        n = ctypes.c_int64(0)
        self._ok(self._L.bd_bfs_levels(self._h, None, 0, ctypes.byref(n)))
        recs = (LevelRec * max(n.value, 1))()
        self._ok(self._L.bd_bfs_levels(self._h, ctypes.cast(recs, ctypes.c_void_p), n.value,
                                       ctypes.byref(n)))
        return _level_dicts(recs, n.value)

    def count_traversed(self):
        v = ctypes.c_int64(0)
        self._ok(self._L.bd_bfs_count_traversed(self._h, ctypes.byref(v)))
        return v.value

    def check(self, source):
        """hybrid_bfs::check of the last search (hybrid_bfs.cc:281-379)."""
        ok = ctypes.c_int(0)
        self._ok(self._L.bd_bfs_check(self._h, int(source), ctypes.byref(ok)))
        return bool(ok.value)

    def check_parents(self, source, parent):
        """hybrid_bfs::check of a given parent array (negative = unreached)."""
        p = np.ascontiguousarray(parent, np.int64)
        nv, _, _ = self.info()
        if p.size != nv:
            raise ValueError(f"parent array has {p.size} entries, the graph {nv} vertices")
        ok = ctypes.c_int(0)
        self._ok(self._L.bd_bfs_check_parents(self._h, int(source), p, ctypes.byref(ok)))
        return bool(ok.value)

    # ------------------------------------------------- connected components
    def components(self):
        """components::run (components.cc:38-104) on the resident graph:
        {num_components, num_iters, ms}."""
        st = CCStats()
        self._ok(self._L.bd_cc_run(self._h, ctypes.byref(st)))
        return st.as_dict()

    def component_labels(self):
        nv, _, _ = self.info()
        lab = np.empty(nv, np.int64)
        self._ok(self._L.bd_cc_fetch(self._h, lab))
        return lab

    def components_check(self):
        ok = ctypes.c_int(0)
        self._ok(self._L.bd_cc_check(self._h, ctypes.byref(ok)))
        return bool(ok.value)

    def get_tuning(self):
        t = Tuning()
        self._ok(self._L.bd_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = Tuning()
        for k, v in kw.items():
            setattr(t, k, int(v))
        self._ok(self._L.bd_set_tuning(self._h, ctypes.byref(t)))

    def launch_times(self):
        """Kernel time (ms) of every step launch of the last search run with
        time_steps=3."""
        n = ctypes.c_int64(0)
        self._ok(self._L.bd_bfs_launch_times(self._h, None, 0, ctypes.byref(n)))
        out = np.empty(n.value, np.float64)
        self._ok(self._L.bd_bfs_launch_times(self._h, out.ctypes.data, n.value, ctypes.byref(n)))
        return out

    def copy_bandwidth(self, nbytes, reps=3):
        """Best rate of a plain device copy of nbytes (read + write GB/s)."""
        g = ctypes.c_double(0)
        self._ok(self._L.bd_copy_bandwidth(self._h, int(nbytes), int(reps), ctypes.byref(g)))
        return g.value

    # --------------------------------------------------------------- batches
    def multi_root(self, roots, want_depths=True):
        r = np.ascontiguousarray(roots, np.int64)
        nv, _, _ = self.info()
        dp = np.empty((len(r), nv), np.int32) if want_depths else None
        ms = ctypes.c_double(0)
        lv = ctypes.c_int64(0)
        self._ok(self._L.bd_bfs_multi_root(self._h, r, len(r),
                                           dp.ctypes.data if dp is not None else None,
                                           ctypes.byref(ms), ctypes.byref(lv)))
        return dp, ms.value, lv.value

    def multi_graph(self, vbase, roots, want_depths=True):
        """Search every graph of a disjoint union (the resident graph) from its
        root at once; vbase[g] is graph g's first vertex, roots are local."""
        vbase = np.ascontiguousarray(vbase, np.int64)
        roots = np.ascontiguousarray(roots, np.int64)
        dp = np.empty(int(vbase[-1]), np.int32) if want_depths else None
        ms = ctypes.c_double(0)
        self._ok(self._L.bd_bfs_multi_graph(self._h, len(roots), vbase, roots,
                                            dp.ctypes.data if dp is not None else None,
                                            ctypes.byref(ms)))
        return dp, ms.value

    def multi_graph_static(self, vbase, roots):
        """multi_graph through the allocation-free entry point (at most 4096
        graphs; no depths output: read them with fetch()).  Returns the
        search's device ms."""
        vbase = np.ascontiguousarray(vbase, np.int64)
        roots = np.ascontiguousarray(roots, np.int64)
        ms = ctypes.c_double(0)
        self._ok(self._L.bd_bfs_multi_graph_static(self._h, len(roots), vbase.ctypes.data,
                                                   roots.ctypes.data, ctypes.byref(ms)))
        return ms.value

    def upload_union(self, graphs):
        """Upload the disjoint union of CSR graphs [(offsets, cols, m), ...];
        returns vbase (ngraphs+1 entries)."""
        off, cols, m, vbase = union_csr(graphs)
        self.upload_csr(off, cols, m)
        return vbase

    # ------------------------------------------------------------------- MLP
    def mlp_set(self, params, h1=32, h2=32):
        self._ok(self._L.bd_mlp_set(self._h, np.ascontiguousarray(params, np.float32), h1, h2))

    def mlp_get(self):
        h1, h2 = ctypes.c_int(), ctypes.c_int()
        self._ok(self._L.bd_mlp_get(self._h, None, 0, ctypes.byref(h1), ctypes.byref(h2)))
        n = h1.value * 8 + h1.value + h2.value * h1.value + h2.value + 2 * h2.value + 2
        p = np.empty(n, np.float32)
        self._ok(self._L.bd_mlp_get(self._h, p.ctypes.data, n, ctypes.byref(h1),
                                    ctypes.byref(h2)))
        return p, h1.value, h2.value

    def mlp_infer(self, x, mfma=True):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty((x.shape[0], 2), np.float32)
        self._ok(self._L.bd_mlp_infer(self._h, x, x.shape[0], y, 1 if mfma else 0))
        return y

    def mlp_loss_grad(self, x, t, mask=None):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        mk = None if mask is None else np.ascontiguousarray(mask, np.float32)
        p, _, _ = self.mlp_get()
        g = np.empty(p.size, np.float32)
        loss = ctypes.c_float(0)
        self._ok(self._L.bd_mlp_loss_grad(self._h, x, t, None if mk is None else mk.ctypes.data,
                                          x.shape[0], g, ctypes.byref(loss)))
        return loss.value, g

    def mlp_train(self, x, t, mask=None, epochs=500, lr=3e-3, seed=0):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        mk = None if mask is None else np.ascontiguousarray(mask, np.float32)
        loss = ctypes.c_float(0)
        self._ok(self._L.bd_mlp_train(self._h, x, t, None if mk is None else mk.ctypes.data,
                                      x.shape[0], epochs, lr, seed, ctypes.byref(loss)))
        return loss.value


def features(rec, nv, m, prev_dir):
    r = LevelRec(**{k: rec[k] for k, _ in LevelRec._fields_})
    x = np.empty(8, np.float32)
    lib().bd_mlp_features(ctypes.byref(r), nv, m, prev_dir, x)
    return x
