"""Bindings for the reference built from its own sources (oracle/_ref/) -- TEST
INFRASTRUCTURE ONLY.

`make -C oracle ref` compiles the reference's kernel translation units with
their main() renamed, plus small shims (ref_shim_*.cc) that take a graph as
plain CSR arrays.  Present only where the reference sources were available
at build time; `available()` says whether a given shim was built.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_REF = os.path.join(_HERE, "_ref")
_LIBS: dict[str, C.CDLL] = {}
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


def available(name: str) -> bool:
    return os.path.exists(os.path.join(_REF, f"libref_{name}.so"))


def _lib(name: str) -> C.CDLL:
    if name not in _LIBS:
        L = C.CDLL(os.path.join(_REF, f"libref_{name}.so"))
        if name != "build":
            L.ref_graph_new.restype = C.c_void_p
            L.ref_graph_new.argtypes = [C.c_int64, _i64p, C.c_void_p, C.c_int, _i64p, C.c_void_p]
            L.ref_graph_free.argtypes = [C.c_void_p]
            L.ref_graph_free.restype = None
        _LIBS[name] = L
    return _LIBS[name]


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class RefGraph:
    """The reference's CSRGraph built from a host graph's arrays."""

    def __init__(self, name: str, g):
        self.lib = _lib(name)
        self.n = g.num_nodes
        off = g.out_offsets
        dst = np.ascontiguousarray(g.out_dests_raw.reshape(-1))
        ioff = idst = None
        if g.directed:
            ioff = _p(g.in_offsets, C.c_int64)
            idst = np.ascontiguousarray(g.in_dests_raw.reshape(-1)).ctypes.data
        self.h = self.lib.ref_graph_new(g.num_nodes, _p(off, C.c_int64), dst.ctypes.data,
                                        int(g.directed), ioff, idst)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_graph_free(self.h)
            self.h = None


def build(args: list[str], weighted: bool = False):
    """(offsets, dests) of the reference builder's graph for these flags."""
    L = _lib("build")
    L.ref_build.restype = C.c_void_p
    L.ref_build.argtypes = [C.c_int, C.POINTER(C.c_char_p), C.c_int]
    L.ref_build_shape.argtypes = [C.c_void_p, C.c_int, _i64p, _i64p, C.POINTER(C.c_int)]
    L.ref_build_copy.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p]
    L.ref_build_free.argtypes = [C.c_void_p, C.c_int]
    a = [b"ref"] + [s.encode() for s in args]
    h = L.ref_build(len(a), (C.c_char_p * len(a))(*a), int(weighted))
    if not h:
        raise RuntimeError("reference builder rejected the flags")
    n, e, d = C.c_int64(), C.c_int64(), C.c_int()
    L.ref_build_shape(h, int(weighted), C.byref(n), C.byref(e), C.byref(d))
    off = np.empty(n.value + 1, dtype=np.int64)
    dst = np.empty(e.value * (2 if weighted else 1), dtype=np.int32)
    L.ref_build_copy(h, int(weighted), off.ctypes.data, dst.ctypes.data, None, None)
    L.ref_build_free(h, int(weighted))
    return off, dst


def pick_sources(g, k: int, given: int = -1) -> np.ndarray:
    L = _lib("build")
    L.ref_pick_sources.argtypes = [C.c_int64, _i64p, _i32p, C.c_int, _i64p, _i32p, C.c_int64,
                                   C.c_int, _i32p]
    L.ref_pick_sources.restype = None
    out = np.empty(k, dtype=np.int32)
    off = g.out_offsets
    dst = np.ascontiguousarray(g.out_neighbors)
    ioff = idst = None
    if g.directed:
        ioff = _p(g.in_offsets, C.c_int64)
        idst = _p(np.ascontiguousarray(g.in_neighbors), C.c_int32)
    L.ref_pick_sources(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), int(g.directed),
                       ioff, idst, given, k, _p(out, C.c_int32))
    return out


def bfs(rg: RefGraph, roots) -> tuple[np.ndarray, float]:
    """Reference DOBFS from each root: (parents[k, n], summed seconds)."""
    rg.lib.ref_bfs.restype = C.c_double
    rg.lib.ref_bfs.argtypes = [C.c_void_p, _i32p, C.c_int, _i32p]
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    out = np.empty((len(roots), rg.n), dtype=np.int32)
    secs = rg.lib.ref_bfs(rg.h, _p(roots, C.c_int32), len(roots), _p(out, C.c_int32))
    return out, secs


def bfs_time(rg: RefGraph, roots) -> float:
    rg.lib.ref_bfs.restype = C.c_double
    rg.lib.ref_bfs.argtypes = [C.c_void_p, _i32p, C.c_int, _i32p]
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    return rg.lib.ref_bfs(rg.h, _p(roots, C.c_int32), len(roots), None)


def sssp(rg: RefGraph, roots, delta: int) -> tuple[np.ndarray, float]:
    """Reference DeltaStep from each root: (dist[k, n], summed seconds)."""
    rg.lib.ref_sssp.restype = C.c_double
    rg.lib.ref_sssp.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_int32, _i32p]
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    out = np.empty((len(roots), rg.n), dtype=np.int32)
    secs = rg.lib.ref_sssp(rg.h, _p(roots, C.c_int32), len(roots), delta, _p(out, C.c_int32))
    return out, secs


def sssp_time(rg: RefGraph, roots, delta: int) -> float:
    rg.lib.ref_sssp.restype = C.c_double
    rg.lib.ref_sssp.argtypes = [C.c_void_p, _i32p, C.c_int, C.c_int32, _i32p]
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    return rg.lib.ref_sssp(rg.h, _p(roots, C.c_int32), len(roots), delta, None)


def pagerank(name: str, g, max_iters: int = 20, epsilon: float = 1e-4):
    """Reference PageRankPullGS (name="pr") or PageRankPull (name="pr_spmv"): (scores, secs)."""
    rg = RefGraph(name, g)
    rg.lib.ref_pr.restype = C.c_double
    rg.lib.ref_pr.argtypes = [C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_float)]
    out = np.empty(g.num_nodes, dtype=np.float32)
    secs = rg.lib.ref_pr(rg.h, max_iters, epsilon, _p(out, C.c_float))
    return out, secs


def afforest(g):
    """Reference Afforest: (comp, secs)."""
    rg = RefGraph("cc", g)
    rg.lib.ref_cc.restype = C.c_double
    rg.lib.ref_cc.argtypes = [C.c_void_p, _i32p]
    out = np.empty(g.num_nodes, dtype=np.int32)
    secs = rg.lib.ref_cc(rg.h, _p(out, C.c_int32))
    return out, secs


def brandes(g, num_iters: int, given: int = -1):
    """Reference Brandes with its own SourcePicker roots: (scores, secs)."""
    rg = RefGraph("bc", g)
    rg.lib.ref_bc.restype = C.c_double
    rg.lib.ref_bc.argtypes = [C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_float)]
    out = np.empty(g.num_nodes, dtype=np.float32)
    secs = rg.lib.ref_bc(rg.h, num_iters, given, _p(out, C.c_float))
    return out, secs


def tc(g, hybrid: bool = True):
    """Reference Hybrid (relabel when worth it) or OrderedCount: (count, secs)."""
    rg = RefGraph("tc", g)
    rg.lib.ref_tc.restype = C.c_double
    rg.lib.ref_tc.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_uint64)]
    c = C.c_uint64()
    secs = rg.lib.ref_tc(rg.h, 0 if hybrid else 1, C.byref(c))
    return int(c.value), secs


def worth_relabelling(g) -> bool:
    rg = RefGraph("tc", g)
    rg.lib.ref_tc_worth_relabelling.argtypes = [C.c_void_p]
    return bool(rg.lib.ref_tc_worth_relabelling(rg.h))
