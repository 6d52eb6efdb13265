"""GPU triangle counting parity: gg_tc counts must equal the reference's
exactly, on the graph as built and relabelled by degree (the reference's
Hybrid path), including hubs above the chunk threshold."""

import math

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u12", lambda: Graph.uniform(12)),
                                      ("k18s", lambda: Graph.kronecker(18, seed=8675309))])
def test_tc_exact(golden, tag, make):
    print(f"[tc] {tag}", flush=True)
    g = cached_graph(tag, make)
    want = golden["tc"][tag]["triangles"] if tag in golden["tc"] else oracle.ordered_count(g)[0]
    assert gpu.OrderedCount(gpu.DeviceGraph(g)) == want
    assert gpu.OrderedCount(gpu.DeviceGraph(g.relabel_by_degree())) == want
    assert gpu.Hybrid(g) == want


def test_tc_closed_forms():
    print("[tc] closed forms", flush=True)
    iu, iv = np.triu_indices(300, 1)
    assert gpu.OrderedCount(gpu.DeviceGraph(Graph.from_edges(iu, iv))) == math.comb(300, 3)
    n = 6000  # wheel: hub n has n forward edges (> 4096-edge chunks)
    u = np.concatenate([np.arange(n), np.arange(n)])
    v = np.concatenate([(np.arange(n) + 1) % n, np.full(n, n)])
    assert gpu.OrderedCount(gpu.DeviceGraph(Graph.from_edges(u, v))) == n


def test_tc_rejects_directed():
    g = Graph.from_edges([0, 1], [1, 2], symmetrize=False)
    with pytest.raises(gpu.GpuError) as e:
        gpu.tc(gpu.DeviceGraph(g))
    assert e.value.code == -95
