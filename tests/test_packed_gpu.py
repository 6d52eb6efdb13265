"""GPU packed batched BFS: many graphs in one traversal (gg_bfs_packed).

No reference counterpart: the reference runs one graph and one root per
trial (src/benchmark.h:106).  Every root's depths must equal the serial BFS
depths (oracle.bfs_depth, restating src/bfs.cc:204-215) on its own graph.
The packing config: 256 Kronecker graphs, edge factor 16, scale drawn
uniformly from 10..16 (numpy default_rng(0).integers(10, 17, 256)),
generator seed = graph index, 8 SourcePicker roots each.
"""

import numpy as np
import pytest

from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu


def packed_case(scales, roots_per_graph, directed=False):
    if directed:
        rng = np.random.default_rng(7)
        graphs = []
        for s in scales:
            n = 1 << s
            graphs.append(Graph.from_edges(rng.integers(0, n, 8 * n), rng.integers(0, n, 8 * n),
                                           symmetrize=False))
    else:
        graphs = [Graph.kronecker(int(s), 16, seed=i) for i, s in enumerate(scales)]
    union, bounds = Graph.pack(graphs)
    roots = np.stack([g.pick_sources(roots_per_graph) + bounds[i] for i, g in enumerate(graphs)])
    return graphs, union, bounds, roots


def check_packed(graphs, union, bounds, roots, alpha=15):
    dg = gpu.DeviceGraph(union)
    res = gpu.bfs_packed(dg, bounds, roots, alpha)
    reached = 0
    for i, g in enumerate(graphs):
        for j, r in enumerate(roots[i]):
            want = oracle.bfs_depth(g, int(r - bounds[i]))
            assert np.array_equal(res.depths[i][j], want), f"graph {i} root {j}"
            reached += int((want >= 0).sum())
    assert res.stats["reached"] == reached
    return res


@pytest.mark.parametrize("alpha", [15, 1000000])
def test_packed_small_kronecker(alpha):
    print(f"[packed] 40 graphs s6..11 alpha={alpha}", flush=True)
    scales = np.random.default_rng(1).integers(6, 12, 40)
    check_packed(*packed_case(scales, 8), alpha)


def test_packed_64_roots_and_directed():
    print("[packed] 64 roots per graph; directed", flush=True)
    check_packed(*packed_case([7, 9, 8], 64))
    check_packed(*packed_case([8, 10, 6, 9], 16, directed=True))


def test_packed_256_kronecker_s10_16():
    print("[packed] 256 graphs s10..16, 8 roots each", flush=True)
    scales = np.random.default_rng(0).integers(10, 17, 256)
    graphs, union, bounds, roots = packed_case(scales, 8)
    res = check_packed(graphs, union, bounds, roots)
    print(f"[packed] {union.num_nodes} vertices, {res.stats['iterations']} levels, "
          f"{res.stats['device_ms']:.2f} ms", flush=True)


def test_packed_rejects_bad_input():
    print("[packed] errors", flush=True)
    a, b = Graph.kronecker(6, seed=1), Graph.kronecker(6, seed=2)
    union, bounds = Graph.pack([a, b])
    dg = gpu.DeviceGraph(union)
    good = np.stack([a.pick_sources(2), b.pick_sources(2) + bounds[1]])
    gpu.bfs_packed(dg, bounds, good)
    with pytest.raises(gpu.GpuError) as e:  # root outside its part
        gpu.bfs_packed(dg, bounds, good[::-1].copy())
    assert e.value.code == -22
    with pytest.raises(gpu.GpuError) as e:  # parts not covering the vertices
        gpu.bfs_packed(dg, bounds[:-1], good[:1])
    assert e.value.code == -22
    # An edge joining the two parts.
    off, nb = union.out_offsets, union.out_neighbors
    u = np.repeat(np.arange(union.num_nodes), np.diff(off))
    joined = Graph.from_edges(np.concatenate([u, [0]]), np.concatenate([nb, [union.num_nodes - 1]]))
    with pytest.raises(gpu.GpuError) as e:
        gpu.bfs_packed(gpu.DeviceGraph(joined), bounds, good)
    assert e.value.code == -22
