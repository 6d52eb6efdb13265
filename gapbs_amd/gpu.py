"""Python face of the gfx950 kernel library (libgapgpu.so, include/gapgpu.h).

The functions named like the reference's kernels take the same arguments and
return the same result arrays as the reference functions they replace:

    DOBFS(g, source, alpha=15, beta=18)      src/bfs.cc:122-179   -> parent[n]
    DeltaStep(g, source, delta)              src/sssp.cc:87-153   -> dist[n]
    PageRankPull(g, max_iters, epsilon)      src/pr_spmv.cc:35-61 -> scores[n]
    Afforest(g, neighbor_rounds=2)           src/cc.cc:95-151     -> comp[n]
    Brandes(g, roots)                        src/bc.cc:95-150     -> scores[n]
    OrderedCount(g) / Hybrid(host_graph)     src/tc.cc:52-100     -> triangles

A `DeviceGraph` is the HBM-resident copy of a host `graph.Graph`.  Every
call goes through the C-ABI; if libgapgpu.so is missing or no GPU is
present the calls raise -- there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

from .graph import Graph

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_INIT_DEVICE = None

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class GpuError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class Stats(C.Structure):
    _fields_ = [
        ("device_ms", C.c_double),
        ("total_ms", C.c_double),
        ("iterations", C.c_int64),
        ("td_steps", C.c_int64),
        ("bu_steps", C.c_int64),
        ("reached", C.c_int64),
        ("reached_edges", C.c_int64),
        ("edges_examined", C.c_int64),
        ("vertices_examined", C.c_int64),
        ("residual", C.c_double),
        ("launches", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [("bfs_alpha", C.c_int32), ("bfs_beta", C.c_int32), ("batch_alpha", C.c_int32),
                ("sssp_delta", C.c_int32), ("pr_small_max", C.c_int32), ("pr_medium_max", C.c_int32),
                ("pr_blocks_per_cu", C.c_int32), ("tuned", C.c_int32), ("best_ms", C.c_double * 4),
                ("pr_segment_log2", C.c_int32 * 2), ("pr_lds_log2", C.c_int32 * 2),
                ("pr_seg_min_density", C.c_int32), ("bfs_blocks_pct", C.c_int32), ("pr_blocks_log2", C.c_int32)]
    _ARRAYS = ("best_ms", "pr_segment_log2", "pr_lds_log2")

    def as_dict(self) -> dict:
        d = {k: getattr(self, k) for k, _ in self._fields_ if k not in self._ARRAYS}
        for k in self._ARRAYS:
            d[k] = list(getattr(self, k))
        return d


TUNE_BFS, TUNE_BATCH, TUNE_SSSP, TUNE_PR = 1, 2, 4, 8


EXPORTED = [
    "gg_init", "gg_graph_upload", "gg_graph_free", "gg_graph_device_bytes",
    "gg_bfs", "gg_bfs_batch", "gg_bfs_packed", "gg_sssp", "gg_sssp_batch", "gg_pagerank", "gg_pagerank_layout",
    "gg_pagerank_blocks",
    "gg_cc", "gg_cc_noalloc", "gg_bc", "gg_bc_batch", "gg_tc",
    "gg_relabel_by_degree", "gg_graph_copy_out", "gg_graph_order_by_degree", "gg_graph_vertex_order",
    "gg_graph_build", "gg_graph_copy_in", "gg_graph_generate", "gg_graph_read", "gg_cc_sv",
    "gg_autotune", "gg_tuning_get", "gg_tuning_set",
    "gg_profile_enable", "gg_profile_reset", "gg_profile_get", "gg_last_error",
]


def lib_path() -> str:
    return os.path.join(_HERE, "libgapgpu.so")


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.gg_init.argtypes = [C.c_int]
        L.gg_graph_upload.argtypes = [C.c_int64, _i64p, _i32p, C.c_int, C.c_int, _i64p, _i32p,
                                      C.POINTER(vp)]
        L.gg_graph_free.argtypes = [vp]
        L.gg_graph_device_bytes.argtypes = [vp]
        L.gg_graph_device_bytes.restype = C.c_int64
        L.gg_bfs.argtypes = [vp, C.c_int32, C.c_int, C.c_int, _i32p, _i32p, C.POINTER(Stats)]
        L.gg_bfs_batch.argtypes = [vp, _i32p, C.c_int, C.c_int, _i32p, C.POINTER(Stats)]
        L.gg_bfs_packed.argtypes = [vp, _i64p, C.c_int, _i32p, C.c_int, C.c_int, _i32p, C.POINTER(Stats)]
        L.gg_sssp.argtypes = [vp, C.c_int32, C.c_int32, _i32p, C.POINTER(Stats)]
        L.gg_sssp_batch.argtypes = [vp, _i32p, C.c_int, C.c_int32, _i32p, C.POINTER(Stats)]
        L.gg_pagerank.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.POINTER(C.c_float),
                                  C.POINTER(Stats)]
        L.gg_pagerank_layout.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.gg_pagerank_blocks.argtypes = [vp, C.c_int, C.POINTER(C.c_int64)]
        L.gg_cc.argtypes = [vp, C.c_int, _i32p, C.POINTER(Stats)]
        L.gg_cc_noalloc.argtypes = [vp, C.c_int, _i32p, C.POINTER(Stats)]
        L.gg_cc_sv.argtypes = [vp, _i32p, C.POINTER(Stats)]
        L.gg_bc.argtypes = [vp, _i32p, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(Stats)]
        L.gg_bc_batch.argtypes = [vp, _i32p, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(Stats)]
        L.gg_tc.argtypes = [vp, C.POINTER(C.c_uint64), C.POINTER(Stats)]
        L.gg_relabel_by_degree.argtypes = [vp, C.POINTER(vp), C.POINTER(Stats)]
        L.gg_graph_copy_out.argtypes = [vp, _i64p, _i32p]
        L.gg_graph_copy_in.argtypes = [vp, _i64p, _i32p]
        L.gg_graph_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp), _i64p, C.POINTER(Stats)]
        L.gg_graph_generate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int, _i32p, C.POINTER(vp),
                                        _i64p, C.POINTER(Stats)]
        L.gg_graph_build.argtypes = [_i32p, _i32p, _i32p, C.c_int64, C.c_int, C.c_int, C.POINTER(vp),
                                     _i64p, C.POINTER(Stats)]
        L.gg_graph_order_by_degree.argtypes = [vp, C.POINTER(Stats)]
        L.gg_graph_vertex_order.argtypes = [vp, _i32p]
        L.gg_autotune.argtypes = [vp, C.c_int, _i32p, C.c_int, C.POINTER(Tuning)]
        L.gg_tuning_get.argtypes = [vp, C.POINTER(Tuning)]
        L.gg_tuning_set.argtypes = [vp, C.POINTER(Tuning)]
        L.gg_profile_enable.argtypes = [vp, C.c_int]
        L.gg_profile_reset.argtypes = [vp]
        L.gg_profile_get.argtypes = [vp, C.c_int, C.c_char_p, C.c_int, _i64p,
                                     C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.gg_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(rc: int) -> int:
    if rc < 0:
        raise GpuError(rc, lib().gg_last_error().decode())
    return rc


def _out(shape, dtype, out):
    """The caller's preallocated result array (reused across calls, so
    timing loops do not fault in fresh host pages), or a new one."""
    if out is None:
        return np.empty(shape, dtype=dtype)
    if out.shape != tuple(shape) or out.dtype != np.dtype(dtype) or not out.flags.c_contiguous:
        raise ValueError(f"out must be a C-contiguous {np.dtype(dtype)} array of shape {tuple(shape)}")
    return out


def init(device: int = 0) -> None:
    global _INIT_DEVICE
    if _INIT_DEVICE != device:
        _check(lib().gg_init(device))
        _INIT_DEVICE = device


def _p(a: np.ndarray, ct):
    return a.ctypes.data_as(C.POINTER(ct))


class DeviceGraph:
    """A graph copied into HBM (int64 offsets, int32 ids, AoS weights)."""

    def __init__(self, g: Graph, device: int = 0, order_by_degree: bool = False):
        init(device)
        self.host = g
        self.ordered = False
        self.order_stats = None
        self.num_nodes = g.num_nodes
        self.weighted = g.weighted
        self.directed = g.directed
        h = C.c_void_p()
        off = g.out_offsets
        dst = g.out_dests_raw.reshape(-1)
        ioff = idst = None
        if g.directed:
            ioff = _p(g.in_offsets, C.c_int64)
            idst = _p(g.in_dests_raw.reshape(-1), C.c_int32)
        _check(lib().gg_graph_upload(g.num_nodes, _p(off, C.c_int64), _p(dst, C.c_int32),
                                     int(g.weighted), int(g.directed), ioff, idst, C.byref(h)))
        self._h = h
        if order_by_degree:
            self.order_by_degree()

    def order_by_degree(self) -> dict:
        """gg_graph_order_by_degree: renumbers the device layout by descending
        degree; calls keep taking and returning uploaded ids."""
        st = Stats()
        _check(lib().gg_graph_order_by_degree(self._h, C.byref(st)))
        self.ordered = True
        self.order_stats = st.as_dict()
        return self.order_stats

    def vertex_order(self) -> np.ndarray:
        """new_id[v]: the device id of uploaded vertex v."""
        out = np.empty(self.num_nodes, dtype=np.int32)
        _check(lib().gg_graph_vertex_order(self._h, _p(out, C.c_int32)))
        return out

    @classmethod
    def _adopt(cls, handle: C.c_void_p, num_nodes: int) -> "DeviceGraph":
        """Wraps a handle the library created (undirected and unweighted unless
        the caller says otherwise)."""
        self = cls.__new__(cls)
        self.host = None
        self.ordered = False
        self.order_stats = None
        self.num_nodes = num_nodes
        self.weighted = False
        self.directed = False
        self._h = handle
        return self

    @property
    def handle(self):
        return self._h

    def copy_out(self) -> tuple[np.ndarray, np.ndarray]:
        """(offsets[n + 1], dests) of the device CSR (gg_graph_copy_out)."""
        off = np.empty(self.num_nodes + 1, dtype=np.int64)
        _check(lib().gg_graph_copy_out(self._h, _p(off, C.c_int64), None))
        dst = np.empty(int(off[-1]) * (2 if self.weighted else 1), dtype=np.int32)
        _check(lib().gg_graph_copy_out(self._h, None, _p(dst, C.c_int32)))
        return off, dst

    def copy_in(self) -> tuple[np.ndarray, np.ndarray]:
        """(offsets[n + 1], dests) of the device in-lists (gg_graph_copy_in)."""
        off = np.empty(self.num_nodes + 1, dtype=np.int64)
        _check(lib().gg_graph_copy_in(self._h, _p(off, C.c_int64), None))
        dst = np.empty(int(off[-1]) * (2 if self.weighted else 1), dtype=np.int32)
        _check(lib().gg_graph_copy_in(self._h, None, _p(dst, C.c_int32)))
        return off, dst

    def device_bytes(self) -> int:
        return lib().gg_graph_device_bytes(self._h)

    def profile(self, on: bool = True) -> None:
        """Time every kernel launch with HIP events (see gg_profile_enable)."""
        _check(lib().gg_profile_enable(self._h, int(on)))

    def profile_reset(self) -> None:
        _check(lib().gg_profile_reset(self._h))

    def profile_report(self) -> dict:
        """{kernel: {launches, ms, alg_bytes}} accumulated since the last reset."""
        out = {}
        name = C.create_string_buffer(128)
        idx = 0
        while True:
            n, ms, b = C.c_int64(), C.c_double(), C.c_double()
            rc = lib().gg_profile_get(self._h, idx, name, 128, C.byref(n), C.byref(ms), C.byref(b))
            if rc < 0:
                break
            out[name.value.decode()] = {"launches": n.value, "ms": ms.value, "alg_bytes": b.value}
            idx += 1
        return out

    def free(self) -> None:
        if getattr(self, "_h", None) is not None and _LIB is not None:
            _LIB.gg_graph_free(self._h)
            self._h = None

    def __del__(self):
        self.free()


@dataclass
class BFSResult:
    parent: np.ndarray
    depth: np.ndarray | None
    stats: dict


def bfs(dg: DeviceGraph, source: int, alpha: int = 15, beta: int = 18,
        want_depth: bool = True, want_parent: bool = True) -> BFSResult:
    """gg_bfs; with want_parent/want_depth False the result stays on the device."""
    n = dg.num_nodes
    parent = np.empty(n, dtype=np.int32) if want_parent else None
    depth = np.empty(n, dtype=np.int32) if want_depth else None
    st = Stats()
    _check(lib().gg_bfs(dg.handle, source, alpha, beta,
                        _p(parent, C.c_int32) if want_parent else None,
                        _p(depth, C.c_int32) if want_depth else None, C.byref(st)))
    return BFSResult(parent, depth, st.as_dict())


def DOBFS(dg: DeviceGraph, source: int, alpha: int = 15, beta: int = 18) -> np.ndarray:
    """Parent array of a direction-optimizing BFS (src/bfs.cc:122-179)."""
    return bfs(dg, source, alpha, beta, want_depth=False).parent


@dataclass
class BatchBFSResult:
    depths: np.ndarray | None  # [k, n] int32, -1 where unreached
    stats: dict


def bfs_batch(dg: DeviceGraph, roots, alpha: int = 15, want_depths: bool = True, out=None) -> BatchBFSResult:
    """gg_bfs_batch: up to 64 roots in one traversal; depths[r] equals the
    single-root BFS depths from roots[r] (the reference's per-trial loop,
    src/benchmark.h:106, over src/bfs.cc:122-179)."""
    r = np.ascontiguousarray(roots, dtype=np.int32)
    n = dg.num_nodes
    depths = _out((len(r), n), np.int32, out) if want_depths else None
    st = Stats()
    _check(lib().gg_bfs_batch(dg.handle, _p(r, C.c_int32), len(r), alpha,
                              _p(depths, C.c_int32) if want_depths else None, C.byref(st)))
    return BatchBFSResult(depths, st.as_dict())


@dataclass
class PackedBFSResult:
    depths: list | None  # per part: [R, |part|] int32, -1 where unreached
    stats: dict


def bfs_packed(dg: DeviceGraph, part, roots, alpha: int = 15, want_depths: bool = True) -> PackedBFSResult:
    """gg_bfs_packed: one traversal over a disjoint union of graphs (vertex
    ranges part[i]..part[i+1]); roots is [nparts, R] global vertex ids."""
    part = np.ascontiguousarray(part, dtype=np.int64)
    r = np.ascontiguousarray(roots, dtype=np.int32)
    if r.ndim != 2:
        raise ValueError("roots must be a [nparts, roots_per_part] array")
    nparts, R = r.shape
    if part.ndim != 1 or part.size != nparts + 1:
        raise ValueError(f"part needs nparts + 1 = {nparts + 1} bounds, got {part.size}")
    n = dg.num_nodes
    flat = np.empty(R * n, dtype=np.int32) if want_depths else None
    st = Stats()
    _check(lib().gg_bfs_packed(dg.handle, _p(part, C.c_int64), nparts, _p(r, C.c_int32), R, alpha,
                               _p(flat, C.c_int32) if want_depths else None, C.byref(st)))
    depths = None
    if want_depths:
        depths = [flat[R * part[i]:R * part[i + 1]].reshape(R, part[i + 1] - part[i]) for i in range(nparts)]
    return PackedBFSResult(depths, st.as_dict())


K_DIST_INF = (2**31 - 1) // 2  # src/sssp.cc:64


@dataclass
class SSSPResult:
    dist: np.ndarray | None
    stats: dict


def sssp(dg: DeviceGraph, source: int, delta: int, want_dist: bool = True, out=None) -> SSSPResult:
    """gg_sssp; with want_dist False the distances stay on the device."""
    dist = _out((dg.num_nodes,), np.int32, out) if want_dist else None
    st = Stats()
    _check(lib().gg_sssp(dg.handle, source, delta, _p(dist, C.c_int32) if want_dist else None,
                         C.byref(st)))
    return SSSPResult(dist, st.as_dict())


@dataclass
class SSSPBatchResult:
    dist: np.ndarray | None  # [k, n]
    stats: dict


def sssp_batch(dg: DeviceGraph, roots, delta: int, want_dist: bool = True, out=None) -> SSSPBatchResult:
    """gg_sssp_batch: up to 64 roots in one delta-stepping traversal;
    dist[i] equals sssp(dg, roots[i], delta).dist."""
    r = np.ascontiguousarray(roots, dtype=np.int32)
    if r.ndim != 1 or not 1 <= r.size <= 64:
        raise ValueError("1 to 64 roots")
    k, n = r.size, dg.num_nodes
    dist = _out((k, n), np.int32, out) if want_dist else None
    st = Stats()
    _check(lib().gg_sssp_batch(dg.handle, _p(r, C.c_int32), k, delta,
                               _p(dist, C.c_int32) if want_dist else None, C.byref(st)))
    return SSSPBatchResult(dist, st.as_dict())


def DeltaStep(dg: DeviceGraph, source: int, delta: int) -> np.ndarray:
    """Shortest distances by delta-stepping (src/sssp.cc:87-153)."""
    return sssp(dg, source, delta).dist


PR_FP32, PR_FP64, PR_GS = 0, 1, 2


@dataclass
class PRResult:
    scores: np.ndarray | None
    stats: dict


def pagerank(dg: DeviceGraph, max_iters: int = 20, epsilon: float = 1e-4, fp64: bool = False,
             want_scores: bool = True, gauss_seidel: bool = False, out=None) -> PRResult:
    """gg_pagerank: Jacobi with FP32 compensated (default) or FP64 accumulation,
    or the in-place Gauss-Seidel sweep (gauss_seidel=True, FP32)."""
    out = _out((dg.num_nodes,), np.float32, out) if want_scores else None
    st = Stats()
    mode = PR_GS if gauss_seidel else (PR_FP64 if fp64 else PR_FP32)
    _check(lib().gg_pagerank(dg.handle, max_iters, epsilon, mode,
                             _p(out, C.c_float) if want_scores else None, C.byref(st)))
    return PRResult(out, st.as_dict())


def pagerank_layout(dg: DeviceGraph, fp64: bool = False) -> dict:
    """gg_pagerank_layout: the source segments a Jacobi run gathers through
    (built on first use)."""
    info = (C.c_int64 * 6)()
    _check(lib().gg_pagerank_layout(dg.handle, PR_FP64 if fp64 else PR_FP32, info))
    return {"lds_segments": info[0], "l2_segments": info[1], "pairs": info[2], "entries": info[3],
            "source_bound": info[4], "build_seconds": info[5] / 1e6}


def pagerank_blocks(dg: DeviceGraph, fp64: bool = False) -> dict:
    """gg_pagerank_blocks: the propagation-block layout a Jacobi run of the
    precision uses (built on first use; width_log2 0 when it runs without)."""
    info = (C.c_int64 * 6)()
    _check(lib().gg_pagerank_blocks(dg.handle, PR_FP64 if fp64 else PR_FP32, info))
    return {"width_log2": info[0], "tiles": info[1], "padded_entries": info[2], "write_chunks": info[3],
            "sum_chunks": info[4], "build_seconds": info[5] / 1e6}


def PageRankPullGS(dg: DeviceGraph, max_iters: int = 20, epsilon: float = 1e-4) -> np.ndarray:
    """Scores of the Gauss-Seidel pull PageRank (src/pr.cc:34-61)."""
    return pagerank(dg, max_iters, epsilon, gauss_seidel=True).scores


def PageRankPull(dg: DeviceGraph, max_iters: int = 20, epsilon: float = 1e-4) -> np.ndarray:
    """Scores of Jacobi pull PageRank (src/pr_spmv.cc:35-61)."""
    return pagerank(dg, max_iters, epsilon).scores


@dataclass
class CCResult:
    comp: np.ndarray | None
    stats: dict


def cc(dg: DeviceGraph, neighbor_rounds: int = 2, want_comp: bool = True, out=None) -> CCResult:
    """gg_cc: Afforest labels (smallest vertex id per component)."""
    comp = _out((dg.num_nodes,), np.int32, out) if want_comp else None
    st = Stats()
    _check(lib().gg_cc(dg.handle, neighbor_rounds, _p(comp, C.c_int32) if want_comp else None,
                       C.byref(st)))
    return CCResult(comp, st.as_dict())


def cc_noalloc(dg: DeviceGraph, neighbor_rounds: int = 2, want_comp: bool = True) -> CCResult:
    """gg_cc_noalloc: Afforest with every device buffer in the library's
    fixed scratch array (no device allocation); labels equal gg_cc's."""
    comp = np.empty(dg.num_nodes, dtype=np.int32) if want_comp else None
    st = Stats()
    _check(lib().gg_cc_noalloc(dg.handle, neighbor_rounds, _p(comp, C.c_int32) if want_comp else None,
                               C.byref(st)))
    return CCResult(comp, st.as_dict())


def cc_sv(dg: DeviceGraph, want_comp: bool = True) -> CCResult:
    """gg_cc_sv: Shiloach-Vishkin labels (smallest vertex id per component)."""
    comp = np.empty(dg.num_nodes, dtype=np.int32) if want_comp else None
    st = Stats()
    _check(lib().gg_cc_sv(dg.handle, _p(comp, C.c_int32) if want_comp else None, C.byref(st)))
    return CCResult(comp, st.as_dict())


def ShiloachVishkin(dg: DeviceGraph) -> np.ndarray:
    """Component labels (src/cc_sv.cc:49-83)."""
    return cc_sv(dg).comp


def Afforest(dg: DeviceGraph, neighbor_rounds: int = 2) -> np.ndarray:
    """Component labels (src/cc.cc:95-151)."""
    return cc(dg, neighbor_rounds).comp


BC_FP32, BC_FP64, BC_FP32_KAHAN = 0, 1, 2


@dataclass
class BCResult:
    scores: np.ndarray | None
    stats: dict


def bc(dg: DeviceGraph, roots, fp64: bool = False, kahan: bool = False,
       want_scores: bool = True, out=None) -> BCResult:
    """gg_bc: max-normalised Brandes scores from the given roots.

    Default: the reference's float arithmetic; kahan=True: compensated score
    sums; fp64=True: double dependencies and scores."""
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    out = _out((dg.num_nodes,), np.float32, out) if want_scores else None
    mode = BC_FP64 if fp64 else (BC_FP32_KAHAN if kahan else BC_FP32)
    st = Stats()
    _check(lib().gg_bc(dg.handle, _p(roots, C.c_int32), len(roots), mode,
                       _p(out, C.c_float) if want_scores else None, C.byref(st)))
    return BCResult(out, st.as_dict())


def bc_batch(dg: DeviceGraph, roots, fp64: bool = False, want_scores: bool = True, out=None) -> BCResult:
    """gg_bc_batch: up to 64 roots in one traversal; equals gg_bc over the
    same roots (bit for bit with the reference's float arithmetic)."""
    roots = np.ascontiguousarray(roots, dtype=np.int32)
    out = _out((dg.num_nodes,), np.float32, out) if want_scores else None
    st = Stats()
    _check(lib().gg_bc_batch(dg.handle, _p(roots, C.c_int32), len(roots), BC_FP64 if fp64 else BC_FP32,
                             _p(out, C.c_float) if want_scores else None, C.byref(st)))
    return BCResult(out, st.as_dict())


def Brandes(dg: DeviceGraph, roots) -> np.ndarray:
    """Betweenness scores from `roots`, divided by the largest (src/bc.cc:95-150)."""
    return bc(dg, roots).scores


def tc(dg: DeviceGraph) -> tuple[int, dict]:
    """gg_tc: ordered triangle count of the uploaded graph: (count, stats)."""
    c = C.c_uint64()
    st = Stats()
    _check(lib().gg_tc(dg.handle, C.byref(c), C.byref(st)))
    return int(c.value), st.as_dict()


def OrderedCount(dg: DeviceGraph) -> int:
    """Triangle count, each counted once at its largest vertex (src/tc.cc:52-71)."""
    return tc(dg)[0]


def Hybrid(g: Graph, device: int = 0) -> int:
    """Relabel by degree when the reference's heuristic says so, then count
    (src/tc.cc:75-100)."""
    dg = DeviceGraph(g, device)
    if not g.worth_relabelling():
        return OrderedCount(dg)
    dr, _ = relabel_by_degree(dg)  # on the device, as part of the count
    return OrderedCount(dr)


def autotune(dg: DeviceGraph, what: int, roots=()) -> dict:
    """gg_autotune: time candidate settings of the selected kernels (TUNE_*
    bits) on this graph from `roots` and keep the fastest in the handle."""
    r = np.ascontiguousarray(roots, dtype=np.int32)
    t = Tuning()
    _check(lib().gg_autotune(dg.handle, what, _p(r, C.c_int32) if len(r) else None, len(r), C.byref(t)))
    return t.as_dict()


def tuning(dg: DeviceGraph) -> dict:
    t = Tuning()
    _check(lib().gg_tuning_get(dg.handle, C.byref(t)))
    return t.as_dict()


def set_tuning(dg: DeviceGraph, values: dict) -> None:
    t = Tuning()
    _check(lib().gg_tuning_get(dg.handle, C.byref(t)))
    for k, v in values.items():
        if k == "best_ms":
            continue
        if k in ("pr_segment_log2", "pr_lds_log2"):
            getattr(t, k)[0], getattr(t, k)[1] = v
            continue
        setattr(t, k, v)
    _check(lib().gg_tuning_set(dg.handle, C.byref(t)))


def build_graph(u, v, w=None, symmetrize: bool = True, weighted: bool = False,
                device: int = 0) -> tuple[DeviceGraph, dict]:
    """gg_graph_build: the squished CSR of the edge list u[i] -> v[i] built on
    the device (the graph gh_build_edges / Graph.from_edges builds on the
    host).  Returns (handle, stats)."""
    init(device)
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    if u.shape != v.shape:
        raise ValueError("u and v differ in length")
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.int32)
        if w.shape != u.shape:
            raise ValueError("w and u differ in length")
        wp = _p(w, C.c_int32)
    h = C.c_void_p()
    n = C.c_int64()
    st = Stats()
    _check(lib().gg_graph_build(_p(u, C.c_int32), _p(v, C.c_int32), wp, len(u), int(symmetrize),
                                int(weighted), C.byref(h), C.byref(n), C.byref(st)))
    dg = DeviceGraph._adopt(h, n.value)
    dg.weighted = bool(weighted)
    dg.directed = not symmetrize
    return dg, st.as_dict()


def read_graph(path: str, device: int = 0) -> tuple[DeviceGraph, dict]:
    """gg_graph_read: a .sg / .wsg file streamed straight into HBM (weighted
    when the suffix is .wsg).  Returns (handle, stats)."""
    init(device)
    weighted = path.endswith(".wsg")
    h = C.c_void_p()
    n = C.c_int64()
    st = Stats()
    _check(lib().gg_graph_read(path.encode(), int(weighted), C.byref(h), C.byref(n), C.byref(st)))
    dg = DeviceGraph._adopt(h, n.value)
    dg.weighted = weighted
    with open(path, "rb") as f:  # the header's first byte: directed
        dg.directed = f.read(1) == b"\x01"
    return dg, st.as_dict()


def generate_graph(kind: int, scale: int, degree: int = 16, seed: int = 27491095, weighted: bool = False,
                   device: int = 0) -> tuple[DeviceGraph, dict]:
    """gg_graph_generate: the graph Graph.kronecker (kind 0) / Graph.uniform
    (kind 1) builds on the host, generated and built on the device; the
    Kronecker permutation comes from the host (gh_permutation).  Returns
    (handle, stats)."""
    from .graph import Graph as HostGraph
    init(device)
    perm = HostGraph.permutation(1 << scale, seed) if kind == 0 else None
    h = C.c_void_p()
    n = C.c_int64()
    st = Stats()
    _check(lib().gg_graph_generate(kind, scale, degree, seed, int(weighted),
                                   _p(perm, C.c_int32) if perm is not None else None, C.byref(h), C.byref(n),
                                   C.byref(st)))
    dg = DeviceGraph._adopt(h, n.value)
    dg.weighted = bool(weighted)
    return dg, st.as_dict()


def relabel_by_degree(dg: DeviceGraph) -> tuple[DeviceGraph, dict]:
    """gg_relabel_by_degree: a new device graph numbered by descending
    (degree, id) with sorted lists (src/builder.h:367-402), and the stats."""
    h = C.c_void_p()
    st = Stats()
    _check(lib().gg_relabel_by_degree(dg.handle, C.byref(h), C.byref(st)))
    return DeviceGraph._adopt(h, dg.num_nodes), st.as_dict()
