"""Device graph build (gg_graph_build, gg_graph_generate) against the host.

The host builder (gh_build_edges) restates BuilderBase::MakeGraphFromEL,
MakeCSR and SquishCSR (src/builder.h:300-341, 141-186) and is pinned
byte-for-byte against the reference's own builder and .sg files in
test_host.py.  The device build of the same edge list must give the same
bytes: offsets, sorted lists with duplicates (smallest weight kept) and
self-loops dropped, both directions when symmetrised, out- and in-lists when
directed.  Cases: generated Kronecker / uniform / grid edge lists (the lists
gh_generate builds), random directed lists with repeats, self-loops and
conflicting weights, the empty list, a list of only self-loops, weights
defaulting to 1, and the error paths.  Device generation (R-MAT and uniform
edges, weights, mt19937 streams per 2^18-edge block) must give the graph
gh_generate gives: Kronecker and uniform, weighted or not, one block or
several, and a weight stream that hits the bounded draw's rejection.
"""

import ctypes as C

import numpy as np
import pytest

from gapbs_amd import gpu
from gapbs_amd import graph as hg
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu


def host_build(u, v, w, symmetrize, weighted):
    u = np.ascontiguousarray(u, dtype=np.int32)
    v = np.ascontiguousarray(v, dtype=np.int32)
    wp = None
    if w is not None:
        w = np.ascontiguousarray(w, dtype=np.int32)
        wp = w.ctypes.data_as(C.POINTER(C.c_int32))
    return Graph._make(hg.lib().gh_build_edges, u.ctypes.data_as(C.POINTER(C.c_int32)),
                       v.ctypes.data_as(C.POINTER(C.c_int32)), wp, len(u), int(symmetrize), int(weighted))


def same(u, v, w=None, symmetrize=True, weighted=None):
    weighted = (w is not None) if weighted is None else weighted
    h = host_build(u, v, w, symmetrize, weighted)
    dg, st = gpu.build_graph(u, v, w, symmetrize=symmetrize, weighted=weighted)
    assert dg.num_nodes == h.num_nodes
    off, dst = dg.copy_out()
    assert np.array_equal(off, h.out_offsets)
    assert np.array_equal(dst, h.out_dests_raw.reshape(-1))
    if not symmetrize:
        ioff, idst = dg.copy_in()
        assert np.array_equal(ioff, h.in_offsets)
        assert np.array_equal(idst, h.in_dests_raw.reshape(-1))
    return h, dg, st


@pytest.mark.parametrize("kind,scale,weighted", [(hg.KRONECKER, 12, False), (hg.KRONECKER, 12, True),
                                                 (hg.UNIFORM, 12, True), (hg.GRID, 64, True),
                                                 (hg.KRONECKER, 20, False)])
def test_generated_lists(kind, scale, weighted):
    print(f"[build] generated kind {kind} scale {scale} weighted {weighted}", flush=True)
    u, v, w = Graph.generate_edges(kind, scale, 16, 27491095, weighted=weighted)
    h, dg, st = same(u, v, w)
    assert st["device_ms"] > 0
    if kind == hg.KRONECKER and not weighted:
        r = int(h.pick_sources(1)[0])
        assert np.array_equal(gpu.bfs(dg, r).depth, oracle.bfs_depth(h, r))


@pytest.mark.parametrize("weighted", [False, True])
def test_directed_repeats_loops_and_weights(weighted):
    print(f"[build] directed random, weighted {weighted}", flush=True)
    rng = np.random.default_rng(3)
    n, m = 5000, 60000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    u[:500] = v[:500]                      # self-loops
    u[500:1500], v[500:1500] = u[1500:2500], v[1500:2500]  # repeated edges
    w = rng.integers(-7, 300, m) if weighted else None     # repeats with other weights
    h, dg, _ = same(u, v, w, symmetrize=False)
    r = int(h.pick_sources(1)[0])
    if not weighted:
        assert np.array_equal(gpu.bfs(dg, r).depth, oracle.bfs_depth(h, r))
    same(u, v, w, symmetrize=True)


def test_small_and_degenerate_lists():
    print("[build] small lists", flush=True)
    same([0], [1])
    same([3, 3, 1], [1, 1, 3], [5, 2, 9])           # duplicate keeps the smallest weight, 2
    same([2, 2, 2], [2, 2, 2])                       # only self-loops: 3 vertices, no entries
    same([0, 5], [5, 0], None, symmetrize=True, weighted=True)  # weights default to 1
    dg, _ = gpu.build_graph([], [])
    assert dg.num_nodes == 0
    off, dst = dg.copy_out()
    assert off.tolist() == [0] and dst.size == 0


def test_errors():
    print("[build] errors", flush=True)
    with pytest.raises(gpu.GpuError) as e:
        gpu.build_graph([0, -1], [1, 2])
    assert e.value.code == -22
    # 2 * 31 id bits + a 3-bit weight range do not fit one 64-bit key.
    with pytest.raises(gpu.GpuError) as e:
        gpu.build_graph([0, 2**31 - 2], [1, 0], [1, 8], weighted=True)
    assert e.value.code == -95


def host_generated(kind, scale, seed, weighted):
    make = Graph.kronecker if kind == hg.KRONECKER else Graph.uniform
    return make(scale, 16, seed=seed, weighted=weighted)


@pytest.mark.parametrize("kind,scale,seed,weighted", [
    (hg.KRONECKER, 10, 27491095, False),
    (hg.KRONECKER, 16, 27491095, True),
    (hg.KRONECKER, 17, 8675309, False),   # several 2^18-edge blocks, each its own stream
    (hg.UNIFORM, 12, 1234, True),
    (hg.UNIFORM, 15, 27491095, False),
    # mt19937 seeded with 30376 draws 0xFFFFFFFF at position 36855 of its
    # stream: the bounded weight draw rejects it, so every later weight of
    # the block (scale 14 x 16 = one 2^18-edge block) comes one draw later.
    (hg.KRONECKER, 14, 30376, True),
])
def test_device_generation_matches_host(kind, scale, seed, weighted):
    print(f"[build] generate kind {kind} scale {scale} seed {seed} weighted {weighted}", flush=True)
    h = host_generated(kind, scale, seed, weighted)
    dg, st = gpu.generate_graph(kind, scale, 16, seed, weighted)
    assert dg.num_nodes == h.num_nodes
    off, dst = dg.copy_out()
    assert np.array_equal(off, h.out_offsets)
    assert np.array_equal(dst, h.out_dests_raw.reshape(-1))
    assert st["device_ms"] > 0


def test_generation_errors():
    print("[build] generation errors", flush=True)
    i32p = C.POINTER(C.c_int32)
    h, n = C.c_void_p(), C.c_int64()
    for args in ((2, 10, 16), (0, 0, 16), (0, 31, 16), (1, 10, 0)):
        rc = gpu.lib().gg_graph_generate(args[0], args[1], args[2], 1, 0, C.cast(None, i32p), C.byref(h),
                                         C.byref(n), None)
        assert rc in (-22, -95), (args, rc)
    rc = gpu.lib().gg_graph_generate(0, 10, 16, 1, 0, C.cast(None, i32p), C.byref(h), C.byref(n), None)
    assert rc == -22  # Kronecker without the permutation
