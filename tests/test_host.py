"""Host graph layer: generator/builder parity, .sg I/O, roots, library exports.

Mirrors the reference's test-generate / test-load smoke tests
(test/test.mk:42-66) and adds byte-level comparisons with the reference's
own builder where it was built from source (oracle/_ref).
"""

import ctypes
import hashlib
import os
import re

import numpy as np
import pytest

from conftest import ROOT, cached_graph
from gapbs_amd.graph import Graph, HostError
from oracle import reference as ref

needs_ref = pytest.mark.skipif(not ref.available("build"), reason="reference not built here")


def test_generate_stats_lines_match_reference_fixtures(golden):
    # test/reference/graph-g10.out and graph-u10.out
    assert Graph.kronecker(10).stats_line() == golden["stats_lines"]["g10"]
    assert Graph.uniform(10).stats_line() == golden["stats_lines"]["u10"]


@pytest.mark.parametrize("tag,make", [
    ("g10", lambda: Graph.kronecker(10)),
    ("u10", lambda: Graph.uniform(10)),
    ("g16", lambda: Graph.kronecker(16)),
    ("g10w", lambda: Graph.kronecker(10, weighted=True)),
])
def test_serialized_bytes_match_reference_converter(tmp_path, golden, tag, make):
    g = make()
    path = str(tmp_path / (tag + (".wsg" if g.weighted else ".sg")))
    g.write(path)
    with open(path, "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == golden["sg_sha256"][tag]


@needs_ref
@pytest.mark.parametrize("flags,kw", [
    (["-g", "12"], dict(kind="kronecker", scale=12)),
    (["-u", "12", "-k", "8"], dict(kind="uniform", scale=12, degree=8)),
    (["-g", "11", "-k", "4"], dict(kind="kronecker", scale=11, degree=4)),
])
@pytest.mark.parametrize("weighted", [False, True])
def test_builder_matches_reference_builder(flags, kw, weighted):
    kw = dict(kw)
    kind = kw.pop("kind")
    g = getattr(Graph, kind)(**kw, weighted=weighted)
    off, dst = ref.build(flags, weighted)
    assert np.array_equal(g.out_offsets, off)
    assert np.array_equal(g.out_dests_raw.reshape(-1), dst)


def test_seeded_generator_differs_and_is_deterministic():
    a = Graph.kronecker(10, seed=8675309)
    b = Graph.kronecker(10, seed=8675309)
    c = Graph.kronecker(10)
    assert np.array_equal(a.out_neighbors, b.out_neighbors)
    assert not (a.num_edges_directed == c.num_edges_directed
                and np.array_equal(a.out_neighbors, c.out_neighbors))


def test_csr_invariants_sorted_unique_no_self_loops():
    g = cached_graph("g12", lambda: Graph.kronecker(12))
    off, nbr = g.out_offsets, g.out_neighbors
    for u in range(0, g.num_nodes, 7):
        row = nbr[off[u]:off[u + 1]]
        assert np.all(np.diff(row) > 0)
        assert u not in row
    # symmetric: u in N(v) for v in N(u) (sampled)
    for u in range(0, g.num_nodes, 97):
        for v in nbr[off[u]:off[u + 1]]:
            assert u in nbr[off[v]:off[v + 1]]


def test_weighted_dedup_keeps_smallest_weight():
    # Duplicate edge 0-1 with weights 9 and 4 -> 4 kept in both directions
    # (sort by (v, w) then unique by v, src/graph.h:38-45, builder.h:152-154).
    g = Graph.from_edges([0, 0, 1, 2], [1, 1, 2, 2], w=[9, 4, 7, 3])
    off, d = g.out_offsets, g.out_dests_raw
    assert off.tolist() == [0, 1, 3, 4]
    assert d.tolist() == [[1, 4], [0, 4], [2, 7], [1, 7]]


def test_directed_build_has_inverse():
    g = Graph.from_edges([0, 0, 1, 3], [1, 2, 2, 0], symmetrize=False)
    assert g.directed
    assert g.out_offsets.tolist() == [0, 2, 3, 3, 4]
    assert g.in_offsets.tolist() == [0, 1, 2, 4, 4]
    assert g.in_neighbors.tolist() == [3, 0, 0, 1]
    assert g.stats_line() == "Graph has 4 nodes and 4 directed edges for degree: 1"


def test_sg_roundtrip_directed_and_weighted(tmp_path):
    g = Graph.from_edges([0, 0, 1, 3], [1, 2, 2, 0], symmetrize=False)
    p = str(tmp_path / "d.sg")
    g.write(p)
    h = Graph.read(p)
    assert h.directed and np.array_equal(h.in_neighbors, g.in_neighbors)
    w = Graph.kronecker(8, weighted=True)
    p2 = str(tmp_path / "w.wsg")
    w.write(p2)
    w2 = Graph.read(p2)
    assert np.array_equal(w2.out_dests_raw, w.out_dests_raw)


def test_source_picker_matches_golden(golden):
    g = cached_graph("g16", lambda: Graph.kronecker(16))
    assert g.pick_sources(16).tolist() == golden["roots"]["g16"]
    assert g.pick_sources(3, given=5).tolist() == [5, 5, 5]


@needs_ref
def test_source_picker_matches_reference_on_seeded_graph():
    g = Graph.kronecker(14, seed=42)
    assert np.array_equal(g.pick_sources(64), ref.pick_sources(g, 64))


def test_host_errors_carry_reference_exit_codes(tmp_path):
    with pytest.raises(HostError) as e:
        Graph.read(str(tmp_path / "missing.sg"))
    assert e.value.code == -6  # src/reader.h:274-276
    with pytest.raises(HostError) as e:
        Graph.kronecker(31)
    assert e.value.code == -31  # src/generator.h:84-90
    g = Graph.from_edges([0], [1], symmetrize=False)
    with pytest.raises(HostError) as e:
        g.relabel_by_degree()
    assert e.value.code == -11  # src/builder.h:370-373


def _declared(header):
    with open(os.path.join(ROOT, "include", header)) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b((?:gg|gh)_\w+)\s*\(", text, re.M)))


@pytest.mark.parametrize("header,libpath", [
    ("gapgpu.h", "gapbs_amd/libgapgpu.so"),
    ("gaphost.h", "gapbs_amd/libgaphost.so"),
])
def test_library_exports_every_declared_symbol(header, libpath):
    names = _declared(header)
    assert len(names) >= 5
    lib = ctypes.CDLL(os.path.join(ROOT, libpath))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_gpu_library_fails_loudly_without_device():
    from gapbs_amd import gpu
    if not os.path.exists("/dev/kfd"):
        with pytest.raises(gpu.GpuError) as e:
            gpu.DeviceGraph(Graph.kronecker(6))
        assert e.value.code in (-19, -5)
    else:
        pytest.skip("a GPU driver is present")


@pytest.mark.parametrize("kind,scale,weighted", [(0, 10, False), (0, 10, True), (1, 10, True), (2, 32, True)])
def test_generated_edge_list_rebuilds_the_generated_graph(kind, scale, weighted):
    # gh_generate_edges hands out the list gh_generate builds (the device
    # builder's input): building it on the host gives the same bytes.
    from gapbs_amd.graph import Graph
    u, v, w = Graph.generate_edges(kind, scale, 16, 8675309, weighted=weighted)
    make = {0: Graph.kronecker, 1: Graph.uniform}.get(kind)
    want = make(scale, 16, seed=8675309, weighted=weighted) if make else Graph.grid(scale, seed=8675309,
                                                                                    weighted=weighted)
    got = Graph.from_edges(u, v, w)
    assert np.array_equal(got.out_offsets, want.out_offsets)
    assert np.array_equal(got.out_dests_raw, want.out_dests_raw)


def test_permutation_is_a_seeded_permutation():
    from gapbs_amd.graph import Graph
    p = Graph.permutation(1 << 12, 42)
    assert np.array_equal(np.sort(p), np.arange(1 << 12))
    assert np.array_equal(p, Graph.permutation(1 << 12, 42))
    assert not np.array_equal(p, Graph.permutation(1 << 12, 43))
