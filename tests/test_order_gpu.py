"""Degree-ordered device layout (gg_graph_order_by_degree).

The handle is renumbered in place by descending (degree, id); every entry
point still takes and returns uploaded ids.  Checked here:
  * the device layout is byte-identical to the host restatement of
    BuilderBase::RelabelByDegree (src/builder.h:367-402) and the vertex
    order is its ranking;
  * BFS depths (src/bfs.cc:122-179, single and batched), SSSP distances
    (src/sssp.cc:156-181 restated as Dijkstra), component labels (smallest
    uploaded id per component) and triangle counts equal the oracle on the
    uploaded graph; BFS parents pass the reference's verifier
    (src/bfs.cc:204-241) in uploaded ids;
  * PageRank stays within 1e-6 of the oracle's FP64 scores;
  * BC scores, single-root and batched, are bit-exact with the oracle's
    float Brandes (src/bc.cc:95-150) on the uploaded graph: the dependency
    folds read the lists in uploaded order, so the sums are the reference's
    in its own successor order;
  * packed traversal, directed graphs and a second ordering are rejected.
"""

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu


def host_rank(g):
    deg = np.diff(g.out_offsets)
    order = np.lexsort((np.arange(g.num_nodes), deg))[::-1]
    new_id = np.empty(g.num_nodes, np.int32)
    new_id[order] = np.arange(g.num_nodes, dtype=np.int32)
    return new_id


def ordered(tag, make):
    key = tag + ".ordered"
    g = cached_graph(tag, make)
    dg = cached_graph(key, lambda: gpu.DeviceGraph(g, order_by_degree=True))
    return g, dg


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u12", lambda: Graph.uniform(12))])
def test_layout_matches_host_relabel(tag, make):
    print(f"[order] layout {tag}", flush=True)
    g, dg = ordered(tag, make)
    assert dg.ordered and dg.order_stats["device_ms"] > 0
    assert np.array_equal(dg.vertex_order(), host_rank(g))
    off, dst = dg.copy_out()
    h = g.relabel_by_degree()
    assert np.array_equal(off, h.out_offsets)
    assert np.array_equal(dst, h.out_neighbors)


@pytest.mark.parametrize("tag,make", [("g16", lambda: Graph.kronecker(16)),
                                      ("u12", lambda: Graph.uniform(12))])
def test_bfs_single_and_batched(tag, make):
    print(f"[order] bfs {tag}", flush=True)
    g, dg = ordered(tag, make)
    roots = g.pick_sources(64)
    for r in roots[:4]:
        res = gpu.bfs(dg, int(r))
        assert np.array_equal(res.depth, oracle.bfs_depth(g, int(r))), int(r)
        assert g.verify_bfs(int(r), res.parent)
    b = gpu.bfs_batch(dg, roots)
    for i in (0, 17, 63):
        assert np.array_equal(b.depths[i], oracle.bfs_depth(g, int(roots[i])))
    with pytest.raises(gpu.GpuError) as e:
        gpu.bfs_packed(dg, [0, g.num_nodes], [[int(roots[0])]])
    assert e.value.code == -95


def test_sssp_weighted():
    print("[order] sssp weighted g14", flush=True)
    wg = cached_graph("g14w", lambda: Graph.kronecker(14, weighted=True))
    dwg = gpu.DeviceGraph(wg, order_by_degree=True)
    assert np.array_equal(dwg.vertex_order(), host_rank(wg))
    for r in wg.pick_sources(3):
        for delta in (1, 32):
            assert np.array_equal(gpu.sssp(dwg, int(r), delta).dist, oracle.dijkstra(wg, int(r))), (int(r), delta)


def test_pagerank_cc_tc():
    print("[order] pr / cc / tc g16", flush=True)
    g, dg = ordered("g16", lambda: Graph.kronecker(16))
    want, _, _ = oracle.pagerank(g, 20, 1e-4, fp64=True)
    for fp64 in (False, True):
        assert np.abs(gpu.pagerank(dg, 20, 1e-4, fp64=fp64).scores - want).max() < 1e-6
    assert np.abs(gpu.pagerank(dg, 20, 1e-4, gauss_seidel=True).scores - want).max() < 1e-4
    assert np.array_equal(gpu.cc(dg).comp, oracle.cc_min_labels(g))
    count, _ = oracle.ordered_count(g.relabel_by_degree())
    assert gpu.tc(dg)[0] == count


def test_cc_many_components():
    # Many small components plus isolated vertices: every label must be the
    # smallest uploaded id of its component.
    print("[order] cc components", flush=True)
    rng = np.random.default_rng(5)
    n = 50000
    u = rng.integers(0, n, 30000)
    v = np.minimum(u + rng.integers(1, 40, 30000), n - 1)
    g = Graph.from_edges(u, v)
    dg = gpu.DeviceGraph(g, order_by_degree=True)
    assert np.array_equal(gpu.cc(dg).comp, oracle.cc_min_labels(g))


@pytest.mark.parametrize("tag,make", [("g12", lambda: Graph.kronecker(12)), ("u12", lambda: Graph.uniform(12)),
                                      ("g15", lambda: Graph.kronecker(15, seed=9))])
def test_bc_bit_exact_in_uploaded_order(tag, make):
    print(f"[order] bc {tag}", flush=True)
    g, dg = ordered(tag, make)
    roots = g.pick_sources(8)
    want = oracle.brandes(g, roots)
    assert np.array_equal(gpu.bc(dg, roots).scores, want)
    assert np.array_equal(gpu.bc_batch(dg, roots).scores, want)
    # The same handle's FP64 paths against the FP64 oracle.
    w64 = oracle.brandes(g, roots, fp64=True)
    assert np.abs(gpu.bc(dg, roots, fp64=True).scores - w64).max() < 1e-6
    assert np.abs(gpu.bc_batch(dg, roots, fp64=True).scores - w64).max() < 1e-6


def test_rejects_directed_and_repeat():
    print("[order] rejects", flush=True)
    gd = Graph.from_edges([0, 1, 2], [1, 2, 0], symmetrize=False)
    with pytest.raises(gpu.GpuError) as e:
        gpu.DeviceGraph(gd, order_by_degree=True)
    assert e.value.code == -95
    dg = gpu.DeviceGraph(Graph.kronecker(8), order_by_degree=True)
    with pytest.raises(gpu.GpuError) as e:
        dg.order_by_degree()
    assert e.value.code == -22


def test_autotune_on_ordered():
    print("[order] autotune g14", flush=True)
    g = cached_graph("g14", lambda: Graph.kronecker(14))
    dg = gpu.DeviceGraph(g, order_by_degree=True)
    roots = g.pick_sources(8)
    t = gpu.autotune(dg, gpu.TUNE_BFS | gpu.TUNE_BATCH, roots)
    res = gpu.bfs(dg, int(roots[1]), t["bfs_alpha"], t["bfs_beta"])
    assert np.array_equal(res.depth, oracle.bfs_depth(g, int(roots[1])))
