"""GPU connected components parity: gg_cc (Afforest, src/cc.cc:95-151) and
gg_cc_sv (Shiloach-Vishkin, src/cc_sv.cc:49-83) labels must equal the
reference's exactly (smallest vertex id per component), on undirected and
directed graphs (weak components), in uploaded or degree order, and for any
number of Afforest's sampled rounds."""

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


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u14d2", lambda: Graph.uniform(14, degree=2)),
                                      ("k18s", lambda: Graph.kronecker(18, seed=8675309))])
def test_cc_labels_exact(golden, tag, make):
    print(f"[cc] {tag}", flush=True)
    g = cached_graph(tag, make)
    dg = dev(tag, make)
    want = oracle.cc_min_labels(g)
    got = gpu.cc(dg).comp
    assert np.array_equal(got, want)
    if tag in golden["cc"]:
        assert len(np.unique(got)) == golden["cc"][tag]["components"]
    for rounds in (0, 1, 3):
        assert np.array_equal(gpu.cc(dg, rounds).comp, want)


def test_cc_directed_uses_inverse():
    print("[cc] directed", flush=True)
    rng = np.random.default_rng(6)
    n, m = 20000, 18000
    g = Graph.from_edges(rng.integers(0, n, m), rng.integers(0, n, m), symmetrize=False)
    got = gpu.cc(gpu.DeviceGraph(g)).comp
    assert np.array_equal(got, oracle.afforest(g)[0])
    assert g.verify_cc(got)


def test_cc_grid_and_hub():
    print("[cc] grid + hub", flush=True)
    g = Graph.grid(128, seed=3, weighted=False)
    assert np.array_equal(gpu.cc(gpu.DeviceGraph(g)).comp, oracle.cc_min_labels(g))
    u = np.concatenate([np.zeros(9000, np.int32), [20000]])
    v = np.concatenate([np.arange(1, 9001), [20001]])
    h = Graph.from_edges(u, v)
    assert np.array_equal(gpu.cc(gpu.DeviceGraph(h)).comp, oracle.cc_min_labels(h))


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u14d2", lambda: Graph.uniform(14, degree=2)),
                                      ("k18s", lambda: Graph.kronecker(18, seed=8675309))])
def test_sv_labels_exact(tag, make):
    # Hooks only ever move a label to a lower id of the same component, so
    # the smallest id stays a root and every label converges to it.
    print(f"[cc_sv] {tag}", flush=True)
    g = cached_graph(tag, make)
    r = gpu.cc_sv(dev(tag, make))
    assert np.array_equal(r.comp, oracle.cc_min_labels(g))
    assert r.stats["iterations"] >= 1
    assert g.verify_cc(r.comp)


def test_sv_directed_grid_hub_and_ordered():
    print("[cc_sv] directed / grid / hub / ordered", flush=True)
    rng = np.random.default_rng(6)
    n, m = 20000, 18000
    d = Graph.from_edges(rng.integers(0, n, m), rng.integers(0, n, m), symmetrize=False)
    assert np.array_equal(gpu.cc_sv(gpu.DeviceGraph(d)).comp, oracle.cc_min_labels(d))
    g = Graph.grid(128, seed=3, weighted=False)
    assert np.array_equal(gpu.cc_sv(gpu.DeviceGraph(g)).comp, oracle.cc_min_labels(g))
    u = np.concatenate([np.zeros(9000, np.int32), [20000]])   # a hub past the chunk cut-off
    v = np.concatenate([np.arange(1, 9001), [20001]])
    h = Graph.from_edges(u, v)
    assert np.array_equal(gpu.cc_sv(gpu.DeviceGraph(h)).comp, oracle.cc_min_labels(h))
    k = cached_graph("g16", lambda: Graph.kronecker(16))
    assert np.array_equal(gpu.cc_sv(gpu.DeviceGraph(k, order_by_degree=True)).comp, oracle.cc_min_labels(k))
