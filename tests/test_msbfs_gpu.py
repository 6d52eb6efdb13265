"""GPU batched BFS parity: gg_bfs_batch (through the C-ABI) against the oracle.

The reference runs one BFS per trial (src/benchmark.h:106 over
src/bfs.cc:122-179); depths from a root are unique, so every root's row of
the batched result must equal the serial BFS depths of that root alone
(oracle.bfs_depth, restating src/bfs.cc:204-215) bit for bit.  Stats are
checked against sums over the per-root serial runs.
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


def check_batch(g, dg, roots, alpha=15):
    res = gpu.bfs_batch(dg, roots, alpha)
    deg = np.diff(g.out_offsets)
    reached = reached_edges = 0
    for i, r in enumerate(roots):
        want = oracle.bfs_depth(g, int(r))
        assert np.array_equal(res.depths[i], want), f"root #{i} = {r}"
        reached += int((want >= 0).sum())
        reached_edges += int(deg[want >= 0].sum())
    assert res.stats["reached"] == reached
    assert res.stats["reached_edges"] == reached_edges
    return res


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u10", lambda: Graph.uniform(10))])
def test_batch_64_sourcepicker_roots(tag, make):
    print(f"[msbfs] {tag} 64 roots", flush=True)
    g = cached_graph(tag, make)
    check_batch(g, dev(tag, make), g.pick_sources(64))


def test_batch_seeded_kronecker_partial_batches():
    print("[msbfs] k18 seed 8675309: 1, 7, 63, 64 roots", flush=True)
    make = lambda: Graph.kronecker(18, seed=8675309)  # noqa: E731
    g = cached_graph("k18s", make)
    dg = dev("k18s", make)
    roots = g.pick_sources(64)
    for k in (1, 7, 63, 64):
        res = check_batch(g, dg, roots[:k])
        assert res.stats["td_steps"] + res.stats["bu_steps"] == res.stats["iterations"]


@pytest.mark.parametrize("alpha", [1, 1000000])
def test_batch_direction_extremes(alpha):
    # Bottom-up when the frontier's out-edges exceed unfinished in-edges /
    # alpha: alpha 1 keeps most levels top-down, alpha 1e6 turns bottom-up
    # from level 2 on (level 1 is always top-down).
    print(f"[msbfs] alpha={alpha}", flush=True)
    make = lambda: Graph.kronecker(16)  # noqa: E731
    g = cached_graph("g16", make)
    res = check_batch(g, dev("g16", make), g.pick_sources(64), alpha)
    assert res.stats["td_steps"] >= 1
    if alpha == 1000000:
        assert res.stats["bu_steps"] > res.stats["td_steps"]


def test_batch_repeated_and_isolated_roots():
    print("[msbfs] repeated / isolated roots", flush=True)
    g = Graph.from_edges([0, 1, 3, 5, 9, 2], [1, 2, 4, 6, 9, 0])  # 9: self-loop only -> degree 0
    dg = gpu.DeviceGraph(g)
    roots = [0, 0, 9, 3, 2, 0, 7]
    res = check_batch(g, dg, roots)
    assert res.depths[2].tolist() == [-1] * 9 + [0]
    assert np.array_equal(res.depths[0], res.depths[1])


@pytest.mark.parametrize("alpha", [15, 1000000])
def test_batch_roots_in_small_components(alpha):
    # A Kronecker graph plus a separate 12-vertex path: the path roots' bits
    # die out after a few levels while the others are still spreading, so
    # bottom-up levels must stop looking for them (the live-root mask).
    print(f"[msbfs] small components alpha={alpha}", flush=True)
    base = cached_graph("g16", lambda: Graph.kronecker(16))
    off, nb = base.out_offsets, base.out_neighbors
    u = np.repeat(np.arange(base.num_nodes), np.diff(off))
    n0 = base.num_nodes
    pu = np.arange(n0, n0 + 11)
    g = Graph.from_edges(np.concatenate([u, pu]), np.concatenate([nb, pu + 1]))
    dg = gpu.DeviceGraph(g)
    roots = np.concatenate([base.pick_sources(60), [n0, n0 + 5, n0 + 11, n0 + 3]])
    check_batch(g, dg, roots, alpha)


def test_batch_directed_graph_uses_in_edges_bottom_up():
    # A directed uniform graph: bottom-up levels scan in-neighbours, so a
    # wrong direction shows up as wrong depths.
    print("[msbfs] directed", flush=True)
    rng = np.random.default_rng(5)
    n, m = 1 << 14, 1 << 18
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    g = Graph.from_edges(u, v, symmetrize=False)
    dg = gpu.DeviceGraph(g)
    roots = rng.choice(n, 64, replace=False)
    for alpha in (1, 15):
        check_batch(g, dg, roots, alpha)


def test_batch_closed_forms():
    print("[msbfs] closed forms", flush=True)
    side = 96
    r, c = np.divmod(np.arange(side * side), side)
    ids = np.arange(side * side)
    u = np.concatenate([ids[c < side - 1], ids[r < side - 1]])
    w = np.concatenate([ids[c < side - 1] + 1, ids[r < side - 1] + side])
    dg = gpu.DeviceGraph(Graph.from_edges(u, w))
    rng = np.random.default_rng(3)
    roots = rng.choice(side * side, 64, replace=False)
    for alpha in (1, 15):
        res = gpu.bfs_batch(dg, roots, alpha)
        for i, s in enumerate(roots):
            rs, cs = divmod(int(s), side)
            assert np.array_equal(res.depths[i], np.abs(r - rs) + np.abs(c - cs)), f"root {s}"
    n = 3000
    path = gpu.DeviceGraph(Graph.from_edges(np.arange(n - 1), np.arange(1, n)))
    res = gpu.bfs_batch(path, [0, n - 1, 1500])
    assert res.depths[0].tolist() == list(range(n))
    assert res.depths[1].tolist() == list(range(n - 1, -1, -1))
    assert res.depths[2].tolist() == [abs(i - 1500) for i in range(n)]


def test_batch_hub_chunks_star_of_stars():
    print("[msbfs] star of stars", flush=True)
    a = np.zeros(9000, np.int32)
    b = np.full(5000, 1, np.int32)
    u = np.concatenate([a, b, [0]])
    v = np.concatenate([np.arange(2, 9002), np.arange(9002, 14002), [1]])
    g = Graph.from_edges(u, v)
    dg = gpu.DeviceGraph(g)
    for alpha in (15, 1000000):
        check_batch(g, dg, [0, 5, 9500, 1, 13000], alpha)


def test_batch_matches_single_root_gpu_and_leaves_no_state():
    # Two different batches in a row on one handle, then single-root BFS.
    print("[msbfs] consecutive batches", flush=True)
    make = lambda: Graph.kronecker(16)  # noqa: E731
    g = cached_graph("g16", make)
    dg = dev("g16", make)
    roots = g.pick_sources(64)
    a = gpu.bfs_batch(dg, roots[:40]).depths
    b = gpu.bfs_batch(dg, roots[40:]).depths
    for i in (0, 17, 39):
        assert np.array_equal(a[i], gpu.bfs(dg, int(roots[i])).depth)
    for i in (0, 23):
        assert np.array_equal(b[i], gpu.bfs(dg, int(roots[40 + i])).depth)


def test_batch_errors():
    print("[msbfs] errors", flush=True)
    g = Graph.kronecker(10)
    dg = gpu.DeviceGraph(g)
    for roots in ([], list(range(65)), [0, g.num_nodes]):
        with pytest.raises(gpu.GpuError) as e:
            gpu.bfs_batch(dg, roots)
        assert e.value.code == -22
    with pytest.raises(gpu.GpuError) as e:
        gpu.bfs_batch(dg, [0], alpha=0)
    assert e.value.code == -22
    wdg = gpu.DeviceGraph(Graph.kronecker(10, weighted=True))
    with pytest.raises(gpu.GpuError) as e:
        gpu.bfs_batch(wdg, [0])
    assert e.value.code == -95


def test_graph_replayed_levels_match_host_loop():
    # Without per-kernel profiling the levels run from a captured hipGraph
    # (two levels, replayed until done); with profiling on, the host drives
    # each level.  Both must give the same depths and statistics, over
    # repeated calls (replays of the cached graph with eager work between
    # them), different root sets, partial batches, both layouts and a
    # graph with many levels.
    print("[msbfs] graph replay vs host loop", flush=True)
    cases = [("g16", lambda: Graph.kronecker(16), False), ("g16", lambda: Graph.kronecker(16), True),
             ("path", lambda: Graph.from_edges(np.arange(3000), np.arange(1, 3001)), False)]
    for tag, make, ordered in cases:
        g = cached_graph(tag, make)
        dg = gpu.DeviceGraph(g, order_by_degree=ordered)
        rng = np.random.default_rng(5)
        for k in (64, 64, 17, 1):
            roots = rng.choice(g.num_nodes, size=k, replace=False).astype(np.int32)
            a = gpu.bfs_batch(dg, roots, 15)
            assert a.stats["launches"] > 0
            dg.profile(True)
            b = gpu.bfs_batch(dg, roots, 15)
            dg.profile(False)
            assert np.array_equal(a.depths, b.depths), (tag, k)
            for f in ("iterations", "td_steps", "bu_steps", "reached", "reached_edges"):
                assert a.stats[f] == b.stats[f], (tag, k, f)
            # The heavy tier's early exit reads other chunks' partial results
            # as they land, so the entries read vary slightly run to run.
            assert abs(a.stats["edges_examined"] - b.stats["edges_examined"]) <= 0.01 * b.stats["edges_examined"]
            for i in (0, k - 1):
                assert np.array_equal(a.depths[i], oracle.bfs_depth(g, int(roots[i]))), (tag, k, i)


def _force_blocks(dg, width_log2, pct=1):
    """Propagation blocks of 2^width_log2 vertices, taken by the first
    bottom-up level after a top-down one whenever more than pct percent of
    the in-edges are unfinished (pct 1: practically always)."""
    gpu.set_tuning(dg, {"pr_blocks_log2": width_log2, "bfs_blocks_pct": pct})


@pytest.mark.parametrize("ordered", [True, False])
@pytest.mark.parametrize("width_log2", [4, 7, 10, 14])
def test_batch_blocks_pass_matches_oracle(ordered, width_log2):
    # The first bottom-up level ORs every in-entry's frontier mask through
    # the propagation blocks (at 2^4 vertices: 4096 tiles, bins summed in
    # several chunks and merged; at 2^14: four tiles), then the light and
    # heavy tiers finish it from acc; depths must equal the oracle's for
    # every root, through the captured level loop and the profiled host loop.
    print(f"[msbfs] blocks pass ordered={ordered} width=2^{width_log2}", flush=True)
    g = cached_graph("g16", lambda: Graph.kronecker(16))
    dg = gpu.DeviceGraph(g, order_by_degree=ordered)
    _force_blocks(dg, width_log2)
    roots = g.pick_sources(64)
    for alpha in (15, 1000000):
        res = check_batch(g, dg, roots, alpha)
        assert res.stats["bu_steps"] > 0
        dg.profile(True)
        again = gpu.bfs_batch(dg, roots, alpha)
        rep = dg.profile_report()
        dg.profile(False)
        assert "msbfs_blocks_or" in rep and rep["msbfs_blocks_or"]["alg_bytes"] > 0, rep.keys()
        assert np.array_equal(again.depths, res.depths)
        assert again.stats["edges_examined"] >= int(g.in_offsets[-1])  # the pass reads every in-entry once
    dg.free()


def test_batch_blocks_pass_hubs_and_components():
    # Star of stars (heavy vertices, finished from acc by the heavy tier's
    # second pass), isolated vertices, a second component, repeated roots.
    print("[msbfs] blocks pass hubs", flush=True)
    a = np.zeros(9000, np.int32)
    b = np.full(5000, 1, np.int32)
    u = np.concatenate([a, b, [0, 20000, 20001]])
    v = np.concatenate([np.arange(2, 9002), np.arange(9002, 14002), [1, 20001, 20002]])
    g = Graph.from_edges(u, v)
    dg = gpu.DeviceGraph(g)
    _force_blocks(dg, 8)
    check_batch(g, dg, [0, 5, 9500, 1, 13000, 20000, 14500, 5, 0], 1000000)
    dg.free()


def test_batch_blocks_pass_in_bc_and_default_off():
    # The batched BC's traversal (keep mode) takes the same path; on a graph
    # below 2^20 vertices the default takes no blocks pass (and gives the
    # same depths); -1 turns it off on any graph.
    print("[msbfs] blocks pass in bc", flush=True)
    g = cached_graph("g16", lambda: Graph.kronecker(16))
    plain = gpu.DeviceGraph(g, order_by_degree=True)
    roots = g.pick_sources(16)
    want = gpu.bc_batch(plain, roots).scores
    plain.profile(True)
    gpu.bfs_batch(plain, roots)
    assert "msbfs_blocks_or" not in plain.profile_report()
    plain.profile(False)
    blk = gpu.DeviceGraph(g, order_by_degree=True)
    _force_blocks(blk, 10)
    assert np.array_equal(gpu.bc_batch(blk, roots).scores, want)
    assert np.array_equal(gpu.bfs_batch(blk, roots).depths, gpu.bfs_batch(plain, roots).depths)
    gpu.set_tuning(blk, {"bfs_blocks_pct": -1})
    blk.profile(True)
    gpu.bfs_batch(blk, roots)
    assert "msbfs_blocks_or" not in blk.profile_report()
    blk.profile(False)
    with pytest.raises(gpu.GpuError) as e:
        gpu.set_tuning(blk, {"bfs_blocks_pct": 101})
    assert e.value.code == -22
    plain.free()
    blk.free()
