"""GPU PageRank parity: gg_pagerank against the FP64 oracle.

Tolerances (scores are O(1/N)): FP32 compensated path within 1e-6 max-abs
of the FP64 oracle (the stated target) and within 1e-5 relative per vertex;
FP64 path within 1e-11 relative.  Iteration counts must equal the oracle's,
and the reference's PRVerifier rule must pass.
"""

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu
_DG = {}


def dev(tag, make):
    if tag not in _DG:
        _DG[tag] = gpu.DeviceGraph(cached_graph(tag, make))
    return _DG[tag]


GRAPHS = [("g10", lambda: Graph.kronecker(10)), ("g16", lambda: Graph.kronecker(16)),
          ("u12", lambda: Graph.uniform(12)), ("k18s", lambda: Graph.kronecker(18, seed=8675309))]


@pytest.mark.parametrize("tag,make", GRAPHS)
def test_pr_fp32_and_fp64_against_fp64_oracle(tag, make):
    print(f"[pr] {tag}", flush=True)
    g = cached_graph(tag, make)
    dg = dev(tag, make)
    want, it, _ = oracle.pagerank(g, 20, 1e-4, fp64=True)
    r32 = gpu.pagerank(dg, 20, 1e-4)
    r64 = gpu.pagerank(dg, 20, 1e-4, fp64=True)
    assert r32.stats["iterations"] == it and r64.stats["iterations"] == it
    s32 = r32.scores.astype(np.float64)
    assert np.abs(s32 - want).max() < 1e-6
    assert (np.abs(s32 - want) / want).max() < 1e-5
    assert (np.abs(r64.scores.astype(np.float64) - want) / want).max() < 1e-6  # float output rounding
    assert g.verify_pr(r32.scores, 1e-4)


def test_pr_fixed_iterations_precision_against_reference_float():
    # With epsilon 0 both paths run exactly max_iters; the compensated FP32
    # path must be at least as close to FP64 as the reference's float sums.
    g = cached_graph("g16", GRAPHS[1][1])
    dg = dev("g16", GRAPHS[1][1])
    want, _, _ = oracle.pagerank(g, 20, 0.0, fp64=True)
    ref32, _, _ = oracle.pagerank(g, 20, 0.0)
    got = gpu.pagerank(dg, 20, 0.0).scores
    e_gpu = np.abs(got.astype(np.float64) - want).sum()
    e_ref = np.abs(ref32.astype(np.float64) - want).sum()
    assert e_gpu <= e_ref * 1.01


def test_pr_directed_and_dangling():
    print("[pr] directed", flush=True)
    rng = np.random.default_rng(9)
    n, m = 5000, 60000
    g = Graph.from_edges(rng.integers(0, n, m), rng.integers(0, n, m), symmetrize=False)
    dg = gpu.DeviceGraph(g)
    want, it, _ = oracle.pagerank(g, 15, 0.0, fp64=True)
    got = gpu.pagerank(dg, 15, 0.0, fp64=True)
    assert got.stats["iterations"] == 15
    np.testing.assert_allclose(got.scores, want.astype(np.float32), rtol=1e-6, atol=0)


def test_pr_hub_block_bin():
    print("[pr] hubs", flush=True)
    u = np.concatenate([np.zeros(5000, np.int32), np.arange(1, 5000)])
    v = np.concatenate([np.arange(1, 5001), np.arange(2, 5001)])
    g = Graph.from_edges(u, v)
    dg = gpu.DeviceGraph(g)
    want, it, _ = oracle.pagerank(g, 20, 1e-6, fp64=True)
    got = gpu.pagerank(dg, 20, 1e-6)
    assert got.stats["iterations"] == it
    assert (np.abs(got.scores - want) / want).max() < 1e-5


def test_pr_zero_iterations_and_errors():
    dg = dev("g10", GRAPHS[0][1])
    g = cached_graph("g10", GRAPHS[0][1])
    r = gpu.pagerank(dg, 0)
    assert r.stats["iterations"] == 0
    assert np.all(r.scores == np.float32(1.0) / np.float32(g.num_nodes))
    with pytest.raises(gpu.GpuError) as e:
        gpu.pagerank(dg, -1)
    assert e.value.code == -22


@pytest.mark.parametrize("tag,make", GRAPHS[1:3])
def test_pr_gauss_seidel(tag, make):
    # The in-place sweep of PageRankPullGS (src/pr.cc:34-61): schedule-
    # dependent like the reference's, so it is held to the reference's own
    # PRVerifier rule and to the Jacobi fixed point it converges to.
    print(f"[pr] gauss-seidel {tag}", flush=True)
    g = cached_graph(tag, make)
    dg = dev(tag, make)
    gs = gpu.pagerank(dg, 20, 1e-4, gauss_seidel=True)
    assert g.verify_pr(gs.scores, 1e-4)
    assert gs.stats["residual"] < 1e-4
    # Sweeps are not compared with Jacobi: in-place sweeps converge faster at
    # first but, on Kronecker graphs, more slowly near the end.
    fixed, _, _ = oracle.pagerank(g, 200, 1e-12, fp64=True)
    # L1 distance to the fixed point <= last change x d / (1 - d) (L1 contraction).
    assert np.abs(gs.scores.astype(np.float64) - fixed).sum() < 1e-4 * 0.85 / 0.15 + 1e-5


def _segmented(dg, sb, eps=1e-4, lds=-1):
    """Jacobi runs with the given L2 and LDS source-segment widths (log2; 0:
    the default, -1: off), restoring the defaults afterwards."""
    gpu.set_tuning(dg, {"pr_segment_log2": [sb, sb], "pr_lds_log2": [lds, lds]})
    try:
        return (gpu.pagerank(dg, 20, eps), gpu.pagerank(dg, 20, eps, fp64=True))
    finally:
        gpu.set_tuning(dg, {"pr_segment_log2": [0, 0], "pr_lds_log2": [0, 0]})


@pytest.mark.parametrize("ordered", [True, False])
@pytest.mark.parametrize("lds,sb", [(-1, 6), (-1, 13), (4, 8), (6, 10), (8, 12), (14, 16)])
def test_pr_source_segments_against_fp64_oracle(ordered, lds, sb):
    # Source segments (the in-entries regrouped by source range, one launch
    # per segment; the first ones staged in LDS) forced at small widths on a
    # graph whose contributions fit L2 anyway, so every part of the layout is
    # exercised: up to 128 LDS and 64 L2 segments, block / wavefront / thread
    # pairs, and the tail gather.  The segmented runs take more launches than
    # the plain pull; both meet the oracle.
    print(f"[pr] segments lds={lds} sb={sb} ordered={ordered}", flush=True)
    g = cached_graph("g16", GRAPHS[1][1])
    dg = gpu.DeviceGraph(g, order_by_degree=ordered)
    want, it, _ = oracle.pagerank(g, 20, 1e-4, fp64=True)
    p32, p64 = _segmented(dg, -1)
    s32, s64 = _segmented(dg, sb, lds=lds)
    assert s32.stats["launches"] > p32.stats["launches"] and s64.stats["launches"] > p64.stats["launches"]
    for r in (s32, s64):
        assert r.stats["iterations"] == it
    x32 = s32.scores.astype(np.float64)
    assert np.abs(x32 - want).max() < 1e-6
    assert (np.abs(x32 - want) / want).max() < 1e-5
    assert (np.abs(s64.scores.astype(np.float64) - want) / want).max() < 1e-6
    assert np.abs(x32 - p32.scores.astype(np.float64)).max() < 1e-7
    assert g.verify_pr(s32.scores, 1e-4)
    dg.free()


@pytest.mark.parametrize("lds,sb", [(-1, 6), (-1, 11), (6, 9), (11, 12)])
def test_pr_source_segments_directed_hub_and_isolated(lds, sb):
    # A directed graph with a hub of 6000 in-neighbours (wavefront pairs at
    # 64-vertex segments, block pairs at 2048), dangling and isolated
    # vertices; at 64-vertex segments the 64 candidate segments cover only
    # part of the ids and the tail gather finishes the rest.
    print(f"[pr] segments directed lds={lds} sb={sb}", flush=True)
    rng = np.random.default_rng(17)
    n = 9000
    u = np.concatenate([rng.integers(0, n - 50, 40000), np.arange(1, 6001), rng.integers(0, 200, 5000)])
    v = np.concatenate([rng.integers(0, n - 50, 40000), np.zeros(6000, np.int64), rng.integers(0, 40, 5000)])
    g = Graph.from_edges(u, v, symmetrize=False)
    dg = gpu.DeviceGraph(g)
    want, it, _ = oracle.pagerank(g, 20, 0.0, fp64=True)
    s32, s64 = _segmented(dg, sb, 0.0, lds)
    p32, _ = _segmented(dg, -1, 0.0)
    assert s32.stats["launches"] > p32.stats["launches"]
    assert s32.stats["iterations"] == it == 20
    assert (np.abs(s32.scores.astype(np.float64) - want) / want).max() < 1e-5
    np.testing.assert_allclose(s64.scores, want.astype(np.float32), rtol=1e-6, atol=0)
    dg.free()


def test_pr_source_segments_stay_off_for_unsorted_lists():
    # Segments need sorted in-lists; an upload whose lists are shuffled keeps
    # the plain pull (same launch count) and the same scores.
    print("[pr] segments unsorted", flush=True)
    import ctypes as C
    g = cached_graph("u12", GRAPHS[2][1])
    off = np.ascontiguousarray(g.out_offsets, dtype=np.int64)
    dst = np.array(g.out_neighbors, dtype=np.int32)
    rng = np.random.default_rng(3)
    for v in range(g.num_nodes):
        rng.shuffle(dst[off[v]:off[v + 1]])
    h = C.c_void_p()
    rc = gpu.lib().gg_graph_upload(g.num_nodes, off.ctypes.data_as(C.POINTER(C.c_int64)),
                                   dst.ctypes.data_as(C.POINTER(C.c_int32)), 0, 0, None, None, C.byref(h))
    assert rc == 0
    dg = gpu.DeviceGraph._adopt(h, g.num_nodes)
    want, it, _ = oracle.pagerank(g, 20, 1e-4, fp64=True)
    s32, _ = _segmented(dg, 6, lds=4)
    p32, _ = _segmented(dg, -1)
    assert s32.stats["launches"] == p32.stats["launches"]
    assert s32.stats["iterations"] == it
    assert (np.abs(s32.scores.astype(np.float64) - want) / want).max() < 1e-5
    dg.free()


def test_pr_segment_width_is_validated():
    dg = dev("g10", GRAPHS[0][1])
    for key, bad in (("pr_segment_log2", [3, 0]), ("pr_segment_log2", [0, 31]), ("pr_segment_log2", [-2, 0]),
                     ("pr_lds_log2", [16, 0]), ("pr_lds_log2", [0, 15]), ("pr_lds_log2", [3, 0])):
        with pytest.raises(gpu.GpuError) as e:
            gpu.set_tuning(dg, {key: bad})
        assert e.value.code == -22
    t = gpu.tuning(dg)
    assert t["pr_segment_log2"] == [0, 0] and t["pr_lds_log2"] == [0, 0]


def _blocked(dg, lg, eps=1e-4):
    """Jacobi runs through propagation blocks of width 2^lg (the layout's
    description too), restoring the default afterwards."""
    gpu.set_tuning(dg, {"pr_blocks_log2": lg})
    try:
        info = gpu.pagerank_blocks(dg)
        return gpu.pagerank(dg, 20, eps), gpu.pagerank(dg, 20, eps, fp64=True), info
    finally:
        gpu.set_tuning(dg, {"pr_blocks_log2": 0})


@pytest.mark.parametrize("ordered", [True, False])
@pytest.mark.parametrize("lg", [4, 6, 9, 14])
def test_pr_blocks_against_fp64_oracle(ordered, lg):
    # Propagation blocks forced on K16 at widths from 16 vertices (4096 tiles
    # and bins, most blocks holding one entry padded to eight) to 16384
    # (four tiles).  A sum chunk holds at most 16 * 2^lg octets, so at the
    # small widths the hub bins are summed in several chunks and merged.  Both precisions meet
    # the oracle, and FP32 stays within 1e-7 of the plain pull.
    print(f"[pr] blocks lg={lg} ordered={ordered}", flush=True)
    g = cached_graph("g16", GRAPHS[1][1])
    dg = gpu.DeviceGraph(g, order_by_degree=ordered)
    want, it, _ = oracle.pagerank(g, 20, 1e-4, fp64=True)
    assert gpu.pagerank_blocks(dg)["width_log2"] == 0  # default: no blocks below 2^20 vertices
    plain = gpu.pagerank(dg, 20, 1e-4)
    b32, b64, info = _blocked(dg, lg)
    k = (g.num_nodes + (1 << lg) - 1) >> lg
    assert info["width_log2"] == lg and info["tiles"] == k
    entries = int(g.in_offsets[-1])
    assert info["padded_entries"] % 8 == 0 and entries <= info["padded_entries"] < entries + 8 * k * k
    assert info["sum_chunks"] >= k and info["write_chunks"] >= 1
    if lg <= 6:
        assert info["sum_chunks"] > k  # some bins summed in several chunks
    for r in (b32, b64):
        assert r.stats["iterations"] == it
    x32 = b32.scores.astype(np.float64)
    assert np.abs(x32 - want).max() < 1e-6
    assert (np.abs(x32 - want) / want).max() < 1e-5
    assert (np.abs(b64.scores.astype(np.float64) - want) / want).max() < 1e-6
    assert np.abs(x32 - plain.scores.astype(np.float64)).max() < 1e-7
    assert g.verify_pr(b32.scores, 1e-4)
    dg.free()


@pytest.mark.parametrize("lg", [5, 8, 12])
@pytest.mark.parametrize("shuffled", [False, True])
def test_pr_blocks_directed_hub_isolated_and_unsorted(lg, shuffled):
    # A directed graph (8951 vertices, not a multiple of any width) with a hub of
    # 6000 in-neighbours, dangling and isolated vertices; shuffled: the
    # lists uploaded unsorted (blocks do not need sorted lists).
    print(f"[pr] blocks directed lg={lg} shuffled={shuffled}", flush=True)
    import ctypes as C
    rng = np.random.default_rng(17)
    n = 9000
    u = np.concatenate([rng.integers(0, n - 50, 40000), np.arange(1, 6001), rng.integers(0, 200, 5000)])
    v = np.concatenate([rng.integers(0, n - 50, 40000), np.zeros(6000, np.int64), rng.integers(0, 40, 5000)])
    g = Graph.from_edges(u, v, symmetrize=False)
    n = g.num_nodes  # the largest id used + 1
    if shuffled:
        arrays = []
        for off, dst in ((g.out_offsets, g.out_neighbors), (g.in_offsets, g.in_neighbors)):
            off = np.ascontiguousarray(off, dtype=np.int64)
            dst = np.array(dst, dtype=np.int32)
            for x in range(n):
                rng.shuffle(dst[off[x]:off[x + 1]])
            arrays += [off, dst]
        p64, p32 = C.POINTER(C.c_int64), C.POINTER(C.c_int32)
        h = C.c_void_p()
        rc = gpu.lib().gg_graph_upload(n, arrays[0].ctypes.data_as(p64), arrays[1].ctypes.data_as(p32), 0, 1,
                                       arrays[2].ctypes.data_as(p64), arrays[3].ctypes.data_as(p32), C.byref(h))
        assert rc == 0, gpu.lib().gg_last_error()
        dg = gpu.DeviceGraph._adopt(h, n)
    else:
        dg = gpu.DeviceGraph(g)
    want, it, _ = oracle.pagerank(g, 20, 0.0, fp64=True)
    b32, b64, info = _blocked(dg, lg, 0.0)
    assert info["width_log2"] == lg
    assert b32.stats["iterations"] == it == 20
    assert (np.abs(b32.scores.astype(np.float64) - want) / want).max() < 1e-5
    np.testing.assert_allclose(b64.scores, want.astype(np.float32), rtol=1e-6, atol=0)
    dg.free()


def test_pr_blocks_width_is_validated():
    dg = dev("g10", GRAPHS[0][1])
    for bad in (3, 15, -2):
        with pytest.raises(gpu.GpuError) as e:
            gpu.set_tuning(dg, {"pr_blocks_log2": bad})
        assert e.value.code == -22
    assert gpu.tuning(dg)["pr_blocks_log2"] == 0
    # 2^18 vertices at width 2^4 would need 16384 tiles (more than 8192).
    big = dev("k18s", GRAPHS[3][1])
    gpu.set_tuning(big, {"pr_blocks_log2": 4})
    try:
        with pytest.raises(gpu.GpuError) as e:
            gpu.pagerank(big, 20, 1e-4)
        assert e.value.code == -22
    finally:
        gpu.set_tuning(big, {"pr_blocks_log2": 0})
    assert gpu.pagerank(big, 20, 1e-4).stats["iterations"] > 0


def test_pr_and_batch_bfs_blocks_build_splits_hubs_and_items():
    # The layout build's work items: vertices 0 and 1 take about 75000
    # in-edges each (lists of 2^15 entries or more are cut into chunks of
    # 2^16, each its own build item: two each), and the other vertices of
    # bin 0 (2^14 vertices at
    # width 2^14) take ~100 in-edges each, 1.6 M entries in one bin: more
    # than one vertex-range item of ~2^20 entries.  PageRank (both
    # precisions) and the batched BFS through the blocks must meet the
    # oracle.
    print("[pr] blocks build: hub chunks and split bins", flush=True)
    rng = np.random.default_rng(23)
    n = 100000  # 140000 draws of a source keep ~75000 distinct ones per hub
    hubs_u = rng.integers(0, n, 280000)
    hubs_v = np.repeat(np.array([0, 1]), 140000)
    mid_v = np.repeat(np.arange(2, 16384), 100)
    mid_u = rng.integers(0, n, mid_v.size)
    tail_u = rng.integers(0, n, 200000)
    tail_v = rng.integers(16384, n, 200000)
    u = np.concatenate([hubs_u, mid_u, tail_u])
    v = np.concatenate([hubs_v, mid_v, tail_v])
    g = Graph.from_edges(u, v, symmetrize=False)
    assert np.diff(g.in_offsets)[:2].min() > 2 ** 16
    dg = gpu.DeviceGraph(g)
    want, it, _ = oracle.pagerank(g, 20, 0.0, fp64=True)
    b32, b64, info = _blocked(dg, 14, 0.0)
    assert info["width_log2"] == 14 and info["tiles"] == 7
    assert b32.stats["iterations"] == it == 20
    assert (np.abs(b32.scores.astype(np.float64) - want) / want).max() < 1e-5
    np.testing.assert_allclose(b64.scores, want.astype(np.float32), rtol=1e-6, atol=0)
    # The batched BFS's blocks level over the same layout (forced for any
    # number of roots and any unfinished share).
    gpu.set_tuning(dg, {"pr_blocks_log2": 14, "bfs_blocks_pct": 1})
    roots = [0, 1, 5, 100, 20000, 99999, 12345, 7]
    res = gpu.bfs_batch(dg, roots, 1000000)
    dg.profile(True)
    again = gpu.bfs_batch(dg, roots, 1000000)
    rep = dg.profile_report()
    dg.profile(False)
    assert "msbfs_blocks_or" in rep and rep["msbfs_blocks_or"]["alg_bytes"] > 0
    for i, r in enumerate(roots):
        assert np.array_equal(res.depths[i], oracle.bfs_depth(g, r)), r
        assert np.array_equal(again.depths[i], res.depths[i]), r
    dg.free()
