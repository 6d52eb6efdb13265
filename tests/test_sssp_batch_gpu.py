"""Batched delta-stepping (gg_sssp_batch) against the CPU oracle.

Each root of a batch must get exactly its shortest distances: the serial
Dijkstra of the oracle (restating src/sssp.cc:164-194) and the single-root
gg_sssp, bit for bit, for every batch size (1 to 64 roots, repeated roots),
bucket width, graph family (Kronecker, uniform, the grid with shortcuts,
random directed graphs, hubs whose lists are cut into chunks), on the
uploaded and the degree-ordered layout, and on both loop drivers (the
captured hipGraph rounds and the host-driven rounds used while profiling).
The golden roots are the reference's SourcePicker roots on its own
Kronecker s16 weighted graph.
"""

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu
INF = (2**31 - 1) // 2
_DG = {}


def dev(tag, make, order=False):
    key = (tag, order)
    if key not in _DG:
        _DG[key] = gpu.DeviceGraph(cached_graph(tag, make), order_by_degree=order)
    return _DG[key]


def check_batch(g, dg, roots, delta, oracle_roots=None):
    res = gpu.sssp_batch(dg, roots, delta)
    assert res.dist.shape == (len(roots), g.num_nodes)
    for i, r in enumerate(roots):
        if oracle_roots is None or i in oracle_roots:
            assert np.array_equal(res.dist[i], oracle.dijkstra(g, int(r))), f"root {r} (slot {i})"
    reached = int((res.dist < INF).sum())
    assert res.stats["reached"] == reached
    deg = np.diff(g.out_offsets)
    assert res.stats["reached_edges"] == int(sum(deg[res.dist[i] < INF].sum() for i in range(len(roots))))
    return res


def test_golden_roots_g16w(golden):
    """The reference's roots on its own g16 weighted graph, in one batch."""
    print("[sssp_batch] g16w golden roots", flush=True)
    make = lambda: Graph.kronecker(16, weighted=True)  # noqa: E731
    g = cached_graph("g16w", make)
    roots = [rec["root"] for rec in golden["sssp"]["g16w"]]
    res = check_batch(g, dev("g16w", make), roots, 2)
    for i, rec in enumerate(golden["sssp"]["g16w"]):
        assert int((res.dist[i] < INF).sum()) == rec["reached"]


@pytest.mark.parametrize("k", [1, 2, 5, 17, 32, 64])
def test_batch_sizes_match_single_root(k):
    print(f"[sssp_batch] k {k}", flush=True)
    make = lambda: Graph.kronecker(14, seed=8675309, weighted=True)  # noqa: E731
    g = cached_graph("k14s", make)
    dg = dev("k14s", make)
    roots = [int(r) for r in g.strip_weights().pick_sources(k)]
    res = check_batch(g, dg, roots, 16, oracle_roots={0, k - 1})
    for i in range(k):
        assert np.array_equal(res.dist[i], gpu.sssp(dg, roots[i], 16).dist), roots[i]


@pytest.mark.parametrize("delta", [1, 7, 64, 255, 4096, 1 << 29])
def test_delta_sweep(delta):
    print(f"[sssp_batch] delta {delta}", flush=True)
    make = lambda: Graph.kronecker(12, seed=5, weighted=True)  # noqa: E731
    g = cached_graph("k12w5", make)
    roots = [int(r) for r in g.strip_weights().pick_sources(40)]
    check_batch(g, dev("k12w5", make), roots, delta)


def test_repeated_roots_and_unreachable():
    print("[sssp_batch] repeated roots, isolated vertices", flush=True)
    make = lambda: Graph.uniform(11, weighted=True)  # noqa: E731
    g = cached_graph("u11w", make)
    roots = [3, 3, 100, 3, 2047, 100, 0]
    check_batch(g, dev("u11w", make), roots, 32)
    iso = Graph.from_edges([0, 2], [1, 3], w=[5, 6])
    res = gpu.sssp_batch(gpu.DeviceGraph(iso), [0, 3, 0], 1)
    assert res.dist.tolist() == [[0, 5, INF, INF], [INF, INF, 6, 0], [0, 5, INF, INF]]


@pytest.mark.parametrize("delta", [1, 100, 1000, 5000])
def test_grid_with_shortcuts(delta):
    print(f"[sssp_batch] grid 96 delta {delta}", flush=True)
    make = lambda: Graph.grid(96, seed=7)  # noqa: E731
    g = cached_graph("grid96", make)
    roots = [0, 4000, 9215] + [int(r) for r in g.strip_weights().pick_sources(29)]
    check_batch(g, dev("grid96", make), roots, delta)


def test_directed_graphs():
    print("[sssp_batch] directed", flush=True)
    rng = np.random.default_rng(11)
    n, m = 4000, 30000
    g = Graph.from_edges(rng.integers(0, n, m), rng.integers(0, n, m), w=rng.integers(1, 1000, m),
                         symmetrize=False)
    dg = gpu.DeviceGraph(g)
    for delta in (1, 50, 2000):
        check_batch(g, dg, list(range(0, 64 * 53, 53)), delta)


def test_hub_chunks():
    """Two hubs of 9000 and 5000 edges: their lists are cut into chunks that
    groups across the grid relax, for every root the hub is active for."""
    print("[sssp_batch] hubs", flush=True)
    rng = np.random.default_rng(2)
    a = np.zeros(9000, np.int32)
    u = np.concatenate([a, np.full(5000, 1, np.int32), [0]])
    v = np.concatenate([np.arange(2, 9002), np.arange(9002, 14002), [1]])
    g = Graph.from_edges(u, v, w=rng.integers(1, 255, len(u)))
    dg = gpu.DeviceGraph(g)
    for delta in (1, 64):
        check_batch(g, dg, [0, 5, 13000, 1, 9001, 0, 77], delta)


def test_degree_ordered_layout():
    print("[sssp_batch] degree-ordered", flush=True)
    make = lambda: Graph.kronecker(13, seed=3, weighted=True)  # noqa: E731
    g = cached_graph("k13w3", make)
    roots = [int(r) for r in g.strip_weights().pick_sources(20)]
    check_batch(g, dev("k13w3", make, order=True), roots, 8)


def test_profiled_loop_equals_captured_loop():
    """With per-kernel profiling on, the rounds are driven from the host one
    at a time; the distances and the work counters equal the captured loop's,
    and the relax kernel is credited with bytes."""
    print("[sssp_batch] profiled loop", flush=True)
    make = lambda: Graph.kronecker(12, seed=5, weighted=True)  # noqa: E731
    g = cached_graph("k12w5", make)
    dg = gpu.DeviceGraph(g)
    roots = [int(r) for r in g.strip_weights().pick_sources(24)]
    a = gpu.sssp_batch(dg, roots, 32)
    dg.profile(True)
    dg.profile_reset()
    b = gpu.sssp_batch(dg, roots, 32)
    rep = dg.profile_report()
    dg.profile(False)
    assert np.array_equal(a.dist, b.dist)
    assert a.stats["iterations"] == b.stats["iterations"]
    assert rep["sssp_batch_relax"]["alg_bytes"] > 0
    assert rep["sssp_batch_relax"]["launches"] >= b.stats["td_steps"]  # idle launches included


def test_errors():
    dg = dev("g10w", lambda: Graph.kronecker(10, weighted=True))
    for roots, delta in (([0], 0), ([-1], 2), ([1 << 20], 2)):
        with pytest.raises(gpu.GpuError) as e:
            gpu.sssp_batch(dg, roots, delta)
        assert e.value.code == -22
    with pytest.raises(ValueError):
        gpu.sssp_batch(dg, list(range(65)), 2)
    import ctypes as C
    r = np.zeros(65, np.int32)
    assert gpu.lib().gg_sssp_batch(dg.handle, r.ctypes.data_as(C.POINTER(C.c_int32)), 65, 2, None, None) == -22
    assert gpu.lib().gg_sssp_batch(dg.handle, r.ctypes.data_as(C.POINTER(C.c_int32)), 0, 2, None, None) == -22
    ug = gpu.DeviceGraph(Graph.kronecker(8))
    with pytest.raises(gpu.GpuError) as e:
        gpu.sssp_batch(ug, [0], 2)
    assert e.value.code == -95
