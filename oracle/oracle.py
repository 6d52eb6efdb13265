"""ctypes bindings for the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg,
never by the product package.  Each function names the reference code it
restates; see oracle.h.
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class BFSStats(C.Structure):
    _fields_ = [("td_steps", C.c_int64), ("bu_steps", C.c_int64),
                ("edges_examined", C.c_int64), ("vertices_examined", C.c_int64),
                ("reached", C.c_int64), ("reached_edges", C.c_int64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BFSStep(C.Structure):
    _fields_ = [("bottom_up", C.c_int32), ("pad", C.c_int32), ("frontier", C.c_int64),
                ("vertices_examined", C.c_int64), ("edges_examined", C.c_int64),
                ("discovered", C.c_int64)]


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle" + os.environ.get("GAPBS_CPU_LIB_SUFFIX", "") + ".so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        L = C.CDLL(path)
        L.orc_bfs_depth.argtypes = [C.c_int64, _i64p, _i32p, C.c_int32, _i32p]
        L.orc_bfs_depth.restype = None
        L.orc_dobfs.argtypes = [C.c_int64, _i64p, _i32p, _i64p, _i32p, C.c_int64, C.c_int32,
                                C.c_int, C.c_int, _i32p, C.POINTER(BFSStats)]
        L.orc_dobfs.restype = C.c_double
        L.orc_dobfs_trace.argtypes = L.orc_dobfs.argtypes + [C.POINTER(BFSStep), C.c_int,
                                                             C.POINTER(C.c_int)]
        L.orc_dobfs_trace.restype = C.c_double
        L.orc_dijkstra.argtypes = [C.c_int64, _i64p, _i32p, C.c_int32, _i32p]
        L.orc_dijkstra.restype = None
        L.orc_delta_step.argtypes = [C.c_int64, _i64p, _i32p, C.c_int64, C.c_int32, C.c_int32, _i32p]
        L.orc_delta_step.restype = C.c_double
        for nm, ct in (("orc_pr_pull_f32", C.c_float), ("orc_pr_pull_f64", C.c_double)):
            getattr(L, nm).argtypes = [C.c_int64, _i64p, _i32p, _i64p, C.c_int, C.c_double,
                                       C.POINTER(ct), C.POINTER(C.c_double)]
            getattr(L, nm).restype = C.c_int
        L.orc_afforest.argtypes = [C.c_int64, _i64p, _i32p, C.c_int, _i64p, _i32p, C.c_int, _i32p]
        L.orc_afforest.restype = C.c_double
        L.orc_cc_min_labels.argtypes = [C.c_int64, _i64p, _i32p, _i32p]
        L.orc_cc_min_labels.restype = None
        for nm, ct in (("orc_brandes_f32", C.c_float), ("orc_brandes_f64", C.c_double)):
            getattr(L, nm).argtypes = [C.c_int64, _i64p, _i32p, _i32p, C.c_int, C.POINTER(ct)]
            getattr(L, nm).restype = None
        L.orc_ordered_count.argtypes = [C.c_int64, _i64p, _i32p, C.POINTER(C.c_double)]
        L.orc_ordered_count.restype = C.c_uint64
        L.orc_ordered_count_sample.argtypes = [C.c_int64, _i64p, _i32p, C.c_int64, C.c_int64,
                                               C.POINTER(C.c_double)]
        L.orc_ordered_count_sample.restype = C.c_uint64
        L.orc_tc_intersect.argtypes = [C.c_int64, _i64p, _i32p]
        L.orc_tc_intersect.restype = C.c_uint64
        L.orc_distance_certificate.argtypes = [C.c_int64, _i64p, _i32p, _i32p, C.c_int32, _i32p,
                                               C.c_int32]
        L.orc_distance_certificate.restype = C.c_int64
        L.orc_depth_from_parent.argtypes = [C.c_int64, _i32p, C.c_int32, _i32p]
        _LIB = L
    return _LIB


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def _csr(g):
    off = g.out_offsets
    dst = g.out_neighbors if not g.weighted else np.ascontiguousarray(g.out_neighbors)
    return off, np.ascontiguousarray(dst)


def bfs_depth(g, src: int) -> np.ndarray:
    """Serial BFS depths (src/bfs.cc:204-215)."""
    off, dst = _csr(g)
    depth = np.empty(g.num_nodes, dtype=np.int32)
    lib().orc_bfs_depth(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), src,
                        _p(depth, C.c_int32))
    return depth


def dobfs(g, src: int, alpha: int = 15, beta: int = 18):
    """OpenMP direction-optimizing BFS (src/bfs.cc:122-179): (parent, seconds, stats)."""
    off, dst = _csr(g)
    ioff, idst = g.in_offsets, np.ascontiguousarray(g.in_neighbors)
    parent = np.empty(g.num_nodes, dtype=np.int32)
    st = BFSStats()
    secs = lib().orc_dobfs(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32),
                           _p(ioff, C.c_int64), _p(idst, C.c_int32), g.num_edges_directed,
                           src, alpha, beta, _p(parent, C.c_int32), C.byref(st))
    return parent, secs, st.as_dict()


def dobfs_steps(g, src: int, alpha: int = 15, beta: int = 18) -> list[dict]:
    """Per-step direction and work of the reference schedule."""
    off, dst = _csr(g)
    ioff, idst = g.in_offsets, np.ascontiguousarray(g.in_neighbors)
    parent = np.empty(g.num_nodes, dtype=np.int32)
    st = BFSStats()
    steps = (BFSStep * 4096)()
    ns = C.c_int()
    lib().orc_dobfs_trace(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32),
                          _p(ioff, C.c_int64), _p(idst, C.c_int32), g.num_edges_directed, src,
                          alpha, beta, _p(parent, C.c_int32), C.byref(st), steps, 4096,
                          C.byref(ns))
    return [{k: getattr(steps[i], k) for k, _ in BFSStep._fields_ if k != "pad"}
            for i in range(min(ns.value, 4096))]


K_DIST_INF = (2**31 - 1) // 2


def _wcsr(g):
    if not g.weighted:
        raise ValueError("needs a weighted graph")
    return g.out_offsets, np.ascontiguousarray(g.out_dests_raw.reshape(-1))


def dijkstra(g, src: int) -> np.ndarray:
    """Serial Dijkstra distances (src/sssp.cc:164-194); unreached = INT32_MAX/2."""
    off, wd = _wcsr(g)
    dist = np.empty(g.num_nodes, dtype=np.int32)
    lib().orc_dijkstra(g.num_nodes, _p(off, C.c_int64), _p(wd, C.c_int32), src, _p(dist, C.c_int32))
    return dist


def distance_certificate(g, src: int, dist: np.ndarray) -> int:
    """Vertices breaking the shortest-path certificate of dist from src on a
    symmetric graph: 0 iff dist equals the distances (weights of a weighted
    graph, unreached = INT32_MAX/2; BFS depths of an unweighted one,
    unreached = -1).  Linear in the edges and parallel, for graphs where a
    serial Dijkstra or BFS per root is too slow."""
    dist = np.ascontiguousarray(dist, dtype=np.int32)
    if g.weighted:
        off, wd = _wcsr(g)
        return lib().orc_distance_certificate(g.num_nodes, _p(off, C.c_int64), None, _p(wd, C.c_int32),
                                              src, _p(dist, C.c_int32), np.iinfo(np.int32).max // 2)
    off, dst = _csr(g)
    return lib().orc_distance_certificate(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), None,
                                          src, _p(dist, C.c_int32), -1)


def delta_step(g, src: int, delta: int):
    """OpenMP delta-stepping (src/sssp.cc:87-153): (dist, seconds)."""
    off, wd = _wcsr(g)
    dist = np.empty(g.num_nodes, dtype=np.int32)
    secs = lib().orc_delta_step(g.num_nodes, _p(off, C.c_int64), _p(wd, C.c_int32),
                                g.num_edges_directed, src, delta, _p(dist, C.c_int32))
    return dist, secs


def pagerank(g, max_iters: int = 20, epsilon: float = 1e-4, fp64: bool = False):
    """Jacobi pull PageRank (src/pr_spmv.cc:35-61): (scores, iterations, last L1 change).

    fp64=False restates the reference in float; fp64=True is the FP64 oracle."""
    ioff, idst = g.in_offsets, np.ascontiguousarray(g.in_neighbors)
    scores = np.empty(g.num_nodes, dtype=np.float64 if fp64 else np.float32)
    err = C.c_double()
    fn = lib().orc_pr_pull_f64 if fp64 else lib().orc_pr_pull_f32
    it = fn(g.num_nodes, _p(ioff, C.c_int64), _p(idst, C.c_int32), _p(g.out_offsets, C.c_int64),
            max_iters, epsilon, _p(scores, C.c_double if fp64 else C.c_float), C.byref(err))
    return scores, it, err.value


def afforest(g, neighbor_rounds: int = 2):
    """OpenMP Afforest (src/cc.cc:95-151): (comp, seconds)."""
    off, dst = _csr(g)
    comp = np.empty(g.num_nodes, dtype=np.int32)
    ioff = _p(g.in_offsets, C.c_int64) if g.directed else None
    idst = _p(np.ascontiguousarray(g.in_neighbors), C.c_int32) if g.directed else None
    secs = lib().orc_afforest(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), int(g.directed),
                              ioff, idst, neighbor_rounds, _p(comp, C.c_int32))
    return comp, secs


def cc_min_labels(g) -> np.ndarray:
    """Components by serial union-find over the out-CSR; label = smallest id."""
    off, dst = _csr(g)
    comp = np.empty(g.num_nodes, dtype=np.int32)
    lib().orc_cc_min_labels(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), _p(comp, C.c_int32))
    return comp


def brandes(g, roots, fp64: bool = False) -> np.ndarray:
    """Serial Brandes, max-normalised (src/bc.cc:169-220); float or FP64."""
    off, dst = _csr(g)
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    out = np.empty(g.num_nodes, dtype=np.float64 if fp64 else np.float32)
    fn = lib().orc_brandes_f64 if fp64 else lib().orc_brandes_f32
    fn(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), _p(roots, C.c_int32), len(roots),
       _p(out, C.c_double if fp64 else C.c_float))
    return out


def ordered_count(g):
    """OpenMP OrderedCount (src/tc.cc:52-71): (count, seconds)."""
    off, dst = _csr(g)
    secs = C.c_double()
    c = lib().orc_ordered_count(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), C.byref(secs))
    return int(c), secs.value


def ordered_count_sample(g, stride: int, first: int = 0):
    """OrderedCount over u = first (mod stride) only: (partial count, seconds)."""
    off, dst = _csr(g)
    secs = C.c_double()
    c = lib().orc_ordered_count_sample(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32), stride, first,
                                       C.byref(secs))
    return int(c), secs.value


def tc_intersect(g) -> int:
    """Triangles by full sorted-list intersection / 6 (src/tc.cc:109-128)."""
    off, dst = _csr(g)
    return int(lib().orc_tc_intersect(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32)))


def depth_from_parent(parent: np.ndarray, src: int) -> np.ndarray:
    """Depths implied by a parent tree; raises if it is not a tree rooted at src."""
    parent = np.ascontiguousarray(parent, dtype=np.int32)
    depth = np.empty(len(parent), dtype=np.int32)
    rc = lib().orc_depth_from_parent(len(parent), _p(parent, C.c_int32), src, _p(depth, C.c_int32))
    if rc != 0:
        raise ValueError("parent array is not a tree rooted at the source")
    return depth
