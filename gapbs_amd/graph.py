"""Host graphs for the Python side: generation, .sg/.wsg I/O, roots and verifiers.

Thin ctypes layer over libgaphost.so (include/gaphost.h), whose C++ restates
the reference's host code: Generator (src/generator.h), BuilderBase
(src/builder.h), Reader/Writer for .sg/.wsg (src/reader.h:254-307,
src/writer.h:39-69), SourcePicker (src/benchmark.h:46-68) and the per-kernel
verifiers (bfs.cc:201-247, sssp.cc:164-194, pr.cc:78-94, cc.cc:177-217,
bc.cc:169-230, tc.cc:109-128).
"""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KRONECKER, UNIFORM, GRID = 0, 1, 2
K_RAND_SEED = 27491095  # src/util.h:22
K_DIST_INF = (2**31 - 1) // 2  # src/sssp.cc:64

_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)


class HostError(RuntimeError):
    """A host-layer call failed; `code` is the reference's exit status for it."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        # GAPBS_CPU_LIB_SUFFIX=.asan loads the AddressSanitizer build
        # (make -C gapbs_amd sanitize), for the sanitized CPU test run.
        path = os.path.join(_HERE, "libgaphost" + os.environ.get("GAPBS_CPU_LIB_SUFFIX", "") + ".so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.gh_generate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int, C.POINTER(vp)]
        L.gh_build_edges.argtypes = [_i32p, _i32p, _i32p, C.c_int64, C.c_int, C.c_int, C.POINTER(vp)]
        L.gh_permutation.argtypes = [C.c_int64, C.c_int64, _i32p]
        L.gh_generate_edges.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int64, C.c_int, C.POINTER(C.c_int64),
                                        _i32p, _i32p, _i32p]
        L.gh_read.argtypes = [C.c_char_p, C.c_int, C.POINTER(vp)]
        L.gh_write.argtypes = [vp, C.c_char_p]
        L.gh_free.argtypes = [vp]
        L.gh_free.restype = None
        L.gh_shape.argtypes = [vp, _i64p, _i64p, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.gh_stats_line.argtypes = [vp, C.c_char_p, C.c_int]
        for f in ("gh_out_offsets", "gh_in_offsets"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = _i64p
        for f in ("gh_out_dests", "gh_in_dests"):
            getattr(L, f).argtypes = [vp]
            getattr(L, f).restype = vp
        L.gh_strip_weights.argtypes = [vp, C.POINTER(vp)]
        L.gh_relabel_by_degree.argtypes = [vp, C.POINTER(vp)]
        L.gh_worth_relabelling.argtypes = [vp]
        L.gh_pack.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp), _i64p]
        L.gh_pick_sources.argtypes = [vp, C.c_int64, C.c_int, _i32p]
        L.gh_verify_bfs.argtypes = [vp, C.c_int32, _i32p]
        L.gh_verify_sssp.argtypes = [vp, C.c_int32, _i32p]
        L.gh_verify_pr.argtypes = [vp, C.POINTER(C.c_float), C.c_double]
        L.gh_verify_cc.argtypes = [vp, _i32p]
        L.gh_verify_bc.argtypes = [vp, _i32p, C.c_int, C.POINTER(C.c_float)]
        L.gh_verify_tc.argtypes = [vp, C.c_uint64]
        L.gh_last_error.restype = C.c_char_p
        _LIB = L
    return _LIB


def _check(rc: int) -> int:
    if rc < 0:
        raise HostError(rc, lib().gh_last_error().decode())
    return rc


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Graph:
    """A host CSR graph (int64 offsets, int32 ids, optional int32 weights).

    Array views (`out_offsets`, `out_neighbors`, `out_weights`, `in_*`) are
    zero-copy numpy views into the C++ graph and live as long as this object.
    """

    def __init__(self, handle: int):
        self._h = C.c_void_p(handle)
        n, e = C.c_int64(), C.c_int64()
        d, w = C.c_int(), C.c_int()
        _check(lib().gh_shape(self._h, C.byref(n), C.byref(e), C.byref(d), C.byref(w)))
        self.num_nodes = n.value
        self.num_edges_directed = e.value
        self.directed = bool(d.value)
        self.weighted = bool(w.value)

    # -- construction -------------------------------------------------------
    @classmethod
    def _make(cls, fn, *args) -> "Graph":
        h = C.c_void_p()
        _check(fn(*args, C.byref(h)))
        return cls(h.value)

    @classmethod
    def kronecker(cls, scale: int, degree: int = 16, seed: int = K_RAND_SEED, weighted=False):
        return cls._make(lib().gh_generate, KRONECKER, scale, degree, seed, int(weighted))

    @classmethod
    def uniform(cls, scale: int, degree: int = 16, seed: int = K_RAND_SEED, weighted=False):
        return cls._make(lib().gh_generate, UNIFORM, scale, degree, seed, int(weighted))

    @classmethod
    def grid(cls, side: int, seed: int = 7, weighted=True):
        return cls._make(lib().gh_generate, GRID, side, 0, seed, int(weighted))

    @staticmethod
    def permutation(n: int, seed: int = K_RAND_SEED) -> np.ndarray:
        """Generator::PermuteIDs' vertex permutation (gh_permutation)."""
        perm = np.empty(n, dtype=np.int32)
        _check(lib().gh_permutation(n, seed, _ptr(perm, C.c_int32)))
        return perm

    @staticmethod
    def generate_edges(kind: int, scale: int, degree: int = 16, seed: int = K_RAND_SEED, weighted=False):
        """The generator's edge list before the build (gh_generate_edges):
        (u, v, w or None), the list kronecker()/uniform()/grid() build from."""
        m = C.c_int64(0)
        _check(lib().gh_generate_edges(kind, scale, degree, seed, int(weighted), C.byref(m), None, None, None))
        u = np.empty(m.value, dtype=np.int32)
        v = np.empty(m.value, dtype=np.int32)
        w = np.empty(m.value, dtype=np.int32) if weighted else None
        _check(lib().gh_generate_edges(kind, scale, degree, seed, int(weighted), C.byref(m), _ptr(u, C.c_int32),
                                       _ptr(v, C.c_int32), _ptr(w, C.c_int32) if weighted else None))
        return u, v, w

    @classmethod
    def from_edges(cls, u, v, w=None, symmetrize=True) -> "Graph":
        u = np.ascontiguousarray(u, dtype=np.int32)
        v = np.ascontiguousarray(v, dtype=np.int32)
        wp = None
        if w is not None:
            w = np.ascontiguousarray(w, dtype=np.int32)
            wp = _ptr(w, C.c_int32)
        return cls._make(lib().gh_build_edges, _ptr(u, C.c_int32), _ptr(v, C.c_int32), wp,
                         len(u), int(symmetrize), int(w is not None))

    @classmethod
    def read(cls, path: str) -> "Graph":
        weighted = path.endswith(".wsg")
        return cls._make(lib().gh_read, path.encode(), int(weighted))

    def write(self, path: str) -> None:
        _check(lib().gh_write(self._h, path.encode()))

    def strip_weights(self) -> "Graph":
        return Graph._make(lib().gh_strip_weights, self._h)

    def relabel_by_degree(self) -> "Graph":
        return Graph._make(lib().gh_relabel_by_degree, self._h)

    @classmethod
    def pack(cls, graphs) -> tuple["Graph", np.ndarray]:
        """Disjoint union of unweighted graphs: (union, part bounds); graph
        i's vertex v is bounds[i] + v in the union."""
        hs = (C.c_void_p * len(graphs))(*[g._h for g in graphs])
        bounds = np.empty(len(graphs) + 1, dtype=np.int64)
        h = C.c_void_p()
        _check(lib().gh_pack(hs, len(graphs), C.byref(h), _ptr(bounds, C.c_int64)))
        return cls(h.value), bounds

    def worth_relabelling(self) -> bool:
        return bool(_check(lib().gh_worth_relabelling(self._h)))

    def __del__(self):
        if getattr(self, "_h", None) and _LIB is not None:
            _LIB.gh_free(self._h)
            self._h = None

    # -- views ----------------------------------------------------------------
    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def _offsets(self, fn) -> np.ndarray:
        p = fn(self._h)
        return np.ctypeslib.as_array(p, shape=(self.num_nodes + 1,))

    def _dests(self, fn, entries) -> np.ndarray:
        p = fn(self._h)
        if entries == 0:
            return np.zeros(0, dtype=np.int32 if not self.weighted else np.int64)
        width = 2 if self.weighted else 1
        arr = np.ctypeslib.as_array(C.cast(p, _i32p), shape=(entries * width,))
        return arr.reshape(entries, width) if self.weighted else arr

    @property
    def out_offsets(self) -> np.ndarray:
        return self._offsets(lib().gh_out_offsets)

    @property
    def in_offsets(self) -> np.ndarray:
        return self._offsets(lib().gh_in_offsets)

    @property
    def out_dests_raw(self) -> np.ndarray:
        """int32 neighbours, or an (E, 2) int32 array of {v, w} pairs."""
        return self._dests(lib().gh_out_dests, int(self.out_offsets[-1]))

    @property
    def in_dests_raw(self) -> np.ndarray:
        return self._dests(lib().gh_in_dests, int(self.in_offsets[-1]))

    @property
    def out_neighbors(self) -> np.ndarray:
        d = self.out_dests_raw
        return d[:, 0] if self.weighted else d

    @property
    def out_weights(self) -> np.ndarray:
        if not self.weighted:
            raise ValueError("graph is unweighted")
        return self.out_dests_raw[:, 1]

    @property
    def in_neighbors(self) -> np.ndarray:
        d = self.in_dests_raw
        return d[:, 0] if self.weighted else d

    def out_degree(self) -> np.ndarray:
        return np.diff(self.out_offsets)

    def num_edges(self) -> int:
        """Reference semantics: undirected graphs count each edge once."""
        return self.num_edges_directed if self.directed else self.num_edges_directed // 2

    def stats_line(self) -> str:
        buf = C.create_string_buffer(256)
        _check(lib().gh_stats_line(self._h, buf, 256))
        return buf.value.decode()

    # -- roots and verifiers -------------------------------------------------
    def pick_sources(self, k: int, given: int = -1) -> np.ndarray:
        out = np.empty(k, dtype=np.int32)
        _check(lib().gh_pick_sources(self._h, given, k, _ptr(out, C.c_int32)))
        return out

    def verify_bfs(self, source: int, parent: np.ndarray) -> bool:
        parent = np.ascontiguousarray(parent, dtype=np.int32)
        return bool(_check(lib().gh_verify_bfs(self._h, source, _ptr(parent, C.c_int32))))

    def verify_sssp(self, source: int, dist: np.ndarray) -> bool:
        dist = np.ascontiguousarray(dist, dtype=np.int32)
        return bool(_check(lib().gh_verify_sssp(self._h, source, _ptr(dist, C.c_int32))))

    def verify_pr(self, scores: np.ndarray, tolerance: float = 1e-4) -> bool:
        scores = np.ascontiguousarray(scores, dtype=np.float32)
        return bool(_check(lib().gh_verify_pr(self._h, _ptr(scores, C.c_float), tolerance)))

    def verify_cc(self, comp: np.ndarray) -> bool:
        comp = np.ascontiguousarray(comp, dtype=np.int32)
        return bool(_check(lib().gh_verify_cc(self._h, _ptr(comp, C.c_int32))))

    def verify_bc(self, roots, scores: np.ndarray) -> bool:
        roots = np.ascontiguousarray(roots, dtype=np.int32)
        scores = np.ascontiguousarray(scores, dtype=np.float32)
        return bool(_check(lib().gh_verify_bc(self._h, _ptr(roots, C.c_int32), len(roots),
                                              _ptr(scores, C.c_float))))

    def verify_tc(self, count: int) -> bool:
        return bool(_check(lib().gh_verify_tc(self._h, count)))
