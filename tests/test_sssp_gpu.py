"""GPU SSSP parity: gg_sssp (through the C-ABI) against the CPU oracle.

Distances must be bit-exact (shortest distances are unique), for every
bucket width delta, on Kronecker, uniform, lattice and directed graphs.
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


def dev(tag, make):
    if tag not in _DG:
        _DG[tag] = gpu.DeviceGraph(cached_graph(tag, make))
    return _DG[tag]


def grid_graph(side, seed=7):
    return Graph.grid(side, seed=seed)


@pytest.mark.parametrize("tag,make", [("g10w", lambda: Graph.kronecker(10, weighted=True)),
                                      ("g16w", lambda: Graph.kronecker(16, weighted=True)),
                                      ("u12w", lambda: Graph.uniform(12, weighted=True))])
def test_sssp_matches_golden_roots(golden, tag, make):
    print(f"[sssp] {tag}", flush=True)
    g = cached_graph(tag, make)
    dg = dev(tag, make)
    for rec in golden["sssp"][tag]:
        res = gpu.sssp(dg, rec["root"], 2)
        want = oracle.dijkstra(g, rec["root"])
        assert np.array_equal(res.dist, want), rec["root"]
        assert res.stats["reached"] == rec["reached"]


@pytest.mark.parametrize("delta", [1, 3, 16, 255, 4096, 1 << 29])
def test_sssp_delta_sweep_kronecker(delta):
    print(f"[sssp] delta {delta}", flush=True)
    g = cached_graph("k16s", lambda: Graph.kronecker(16, seed=8675309, weighted=True))
    dg = dev("k16s", lambda: Graph.kronecker(16, seed=8675309, weighted=True))
    for r in g.strip_weights().pick_sources(3):
        assert np.array_equal(gpu.sssp(dg, int(r), delta).dist, oracle.dijkstra(g, int(r)))


@pytest.mark.parametrize("delta", [1, 100, 1000, 5000])
def test_sssp_grid_with_shortcuts(delta):
    print(f"[sssp] grid delta {delta}", flush=True)
    g = cached_graph("grid96", lambda: grid_graph(96))
    dg = dev("grid96", lambda: grid_graph(96))
    for r in (0, 4000, 9215):
        assert np.array_equal(gpu.sssp(dg, r, delta).dist, oracle.dijkstra(g, r))


def test_sssp_directed_and_unreachable():
    print("[sssp] directed", flush=True)
    rng = np.random.default_rng(11)
    n, m = 4000, 30000
    g = Graph.from_edges(rng.integers(0, n, m), rng.integers(0, n, m),
                         w=rng.integers(1, 1000, m), symmetrize=False)
    dg = gpu.DeviceGraph(g)
    for src in (0, 5, 77):
        for delta in (1, 50, 2000):
            assert np.array_equal(gpu.sssp(dg, src, delta).dist, oracle.dijkstra(g, src))
    iso = Graph.from_edges([0, 2], [1, 3], w=[5, 6])
    res = gpu.sssp(gpu.DeviceGraph(iso), 0, 1)
    assert res.dist.tolist() == [0, 5, INF, INF]


def test_sssp_hub_chunks():
    print("[sssp] hubs", flush=True)
    rng = np.random.default_rng(2)
    a = np.zeros(9000, np.int32)
    u = np.concatenate([a, np.full(5000, 1, np.int32), [0]])
    v = np.concatenate([np.arange(2, 9002), np.arange(9002, 14002), [1]])
    g = Graph.from_edges(u, v, w=rng.integers(1, 255, len(u)))
    dg = gpu.DeviceGraph(g)
    for src in (0, 5, 13000):
        for delta in (1, 64):
            assert np.array_equal(gpu.sssp(dg, src, delta).dist, oracle.dijkstra(g, src))


def test_sssp_errors():
    dg = dev("g10w", lambda: Graph.kronecker(10, weighted=True))
    with pytest.raises(gpu.GpuError) as e:
        gpu.sssp(dg, 0, 0)
    assert e.value.code == -22
    with pytest.raises(gpu.GpuError) as e:
        gpu.sssp(dg, -1, 2)
    assert e.value.code == -22
    ug = gpu.DeviceGraph(Graph.kronecker(8))
    with pytest.raises(gpu.GpuError) as e:
        gpu.sssp(ug, 0, 2)
    assert e.value.code == -95
