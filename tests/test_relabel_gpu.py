"""Device relabelling by degree (gg_relabel_by_degree) against the host
restatement of BuilderBase::RelabelByDegree (src/builder.h:367-402): the
offsets and destinations must be byte-identical, on graphs whose rows
exercise every sort tier (<= 32 entries: one thread; <= 2048: a block
bitonic sort; longer: the (row, value) radix sort) and multi-tile radix
passes; and TC on the device-relabelled graph must give the reference's
count (its Hybrid path)."""

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd import gpu
from gapbs_amd.graph import Graph

pytestmark = pytest.mark.gpu


def check(g):
    dg = gpu.DeviceGraph(g)
    dr, st = gpu.relabel_by_degree(dg)
    off, dst = dr.copy_out()
    h = g.relabel_by_degree()
    assert np.array_equal(off, h.out_offsets)
    assert np.array_equal(dst, h.out_neighbors)
    assert st["device_ms"] > 0
    return dr, h


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u12", lambda: Graph.uniform(12)),
                                      ("k18s", lambda: Graph.kronecker(18, seed=8675309))])
def test_relabel_matches_host(golden, tag, make):
    print(f"[relabel] {tag}", flush=True)
    g = cached_graph(tag, make)
    dr, _ = check(g)
    if tag in golden["tc"]:
        assert gpu.OrderedCount(dr) == golden["tc"][tag]["triangles"]
        assert gpu.Hybrid(g) == golden["tc"][tag]["triangles"]


def test_relabel_hubs_and_ties():
    # Two hubs above 2048 entries, many equal degrees (ties go to the larger
    # old id first), isolated vertices and a 3000-vertex path.
    print("[relabel] hubs / ties", flush=True)
    a = np.zeros(9000, np.int32)
    b = np.full(3000, 1, np.int32)
    p = np.arange(20000, 22999)
    u = np.concatenate([a, b, [0], p])
    v = np.concatenate([np.arange(2, 9002), np.arange(9002, 12002), [1], p + 1])
    check(Graph.from_edges(u, v))


def test_relabel_scale_20():
    print("[relabel] kronecker s20", flush=True)
    check(Graph.kronecker(20, seed=3))


def test_relabel_rejects_directed_and_weighted():
    g = Graph.from_edges([0, 1], [1, 2], symmetrize=False)
    with pytest.raises(gpu.GpuError) as e:
        gpu.relabel_by_degree(gpu.DeviceGraph(g))
    assert e.value.code == -95
    with pytest.raises(gpu.GpuError) as e:
        gpu.relabel_by_degree(gpu.DeviceGraph(Graph.kronecker(8, weighted=True)))
    assert e.value.code == -95
