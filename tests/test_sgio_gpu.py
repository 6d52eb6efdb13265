"""Serialized graphs straight into HBM (gg_graph_read) against the host reader.

The host reader/writer (gh_read / gh_write) restate Reader::ReadSerializedGraph
and WriterBase::WriteSerializedGraph (src/reader.h:254-307, writer.h:39-69)
and are pinned against the reference's own .sg/.wsg files in test_host.py.
A file the host writes must arrive in HBM byte for byte: offsets and lists,
both sides of a directed graph, weighted {v, w} pairs, files larger than one
64 MiB staging buffer, and the empty graph; the reference's suffix rules and
short or corrupt files are errors, never a partial graph.
"""

import numpy as np
import pytest

from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu


def roundtrip(g, path):
    g.write(str(path))
    dg, st = gpu.read_graph(str(path))
    assert dg.num_nodes == g.num_nodes
    off, dst = dg.copy_out()
    assert np.array_equal(off, g.out_offsets)
    assert np.array_equal(dst, g.out_dests_raw.reshape(-1))
    ioff, idst = dg.copy_in()
    assert np.array_equal(ioff, g.in_offsets)
    assert np.array_equal(idst, g.in_dests_raw.reshape(-1))
    return dg, st


def test_undirected_weighted_and_large(tmp_path):
    print("[sgio] kronecker 12, 12 weighted, 21 (multi-chunk)", flush=True)
    roundtrip(Graph.kronecker(12), tmp_path / "g12.sg")
    roundtrip(Graph.kronecker(12, weighted=True), tmp_path / "g12.wsg")
    g = Graph.kronecker(21)  # ~ 0.27 GB of lists: several staging chunks
    dg, st = roundtrip(g, tmp_path / "g21.sg")
    r = int(g.pick_sources(1)[0])
    assert np.array_equal(gpu.bfs(dg, r).depth, oracle.bfs_depth(g, r))
    assert st["total_ms"] > 0


def test_directed_and_empty(tmp_path):
    print("[sgio] directed and empty", flush=True)
    rng = np.random.default_rng(5)
    g = Graph.from_edges(rng.integers(0, 3000, 40000), rng.integers(0, 3000, 40000), symmetrize=False)
    roundtrip(g, tmp_path / "d.sg")
    gw = Graph.from_edges(rng.integers(0, 500, 4000), rng.integers(0, 500, 4000),
                          rng.integers(1, 50, 4000), symmetrize=False)
    roundtrip(gw, tmp_path / "d.wsg")
    roundtrip(Graph.from_edges([0], [0]), tmp_path / "loop.sg")  # one vertex, no entries


def test_errors(tmp_path):
    print("[sgio] errors", flush=True)
    g = Graph.kronecker(8)
    good = tmp_path / "g.sg"
    g.write(str(good))
    data = good.read_bytes()
    cases = {"g.el": data, "g.wsg": data, "short.sg": data[: len(data) // 2],
             "header.sg": data[:10]}
    bad = bytearray(data)
    bad[17 + 8 * (g.num_nodes + 1)] = 0xFF  # first destination -> far out of range
    bad[17 + 8 * (g.num_nodes + 1) + 3] = 0x7F
    cases["dest.sg"] = bytes(bad)
    for name, blob in cases.items():
        (tmp_path / name).write_bytes(blob)
        with pytest.raises(gpu.GpuError) as e:
            gpu.read_graph(str(tmp_path / name))
        assert e.value.code == -22, name
    # A crafted 33-byte file whose header announces 2^62 entries (one
    # vertex, offsets {0, 2^62}): the byte count would wrap to 0, so the
    # header is refused on the host before any allocation or kernel.
    import struct
    for suffix in (".sg", ".wsg"):
        for entries in (1 << 62, 1 << 61, 1 << 39):
            blob = struct.pack("<?qq", False, entries, 1) + struct.pack("<qq", 0, entries)
            path = tmp_path / f"wrap{entries}{suffix}"
            path.write_bytes(blob)
            with pytest.raises(gpu.GpuError) as e:
                gpu.read_graph(str(path))
            assert e.value.code == -22, (entries, suffix)
    (tmp_path / "trailing.sg").write_bytes(data + b"\0" * 8)  # longer than its header says
    with pytest.raises(gpu.GpuError) as e:
        gpu.read_graph(str(tmp_path / "trailing.sg"))
    assert e.value.code == -22
    with pytest.raises(gpu.GpuError) as e:
        gpu.read_graph(str(tmp_path / "missing.sg"))
    assert e.value.code == -22
    # a .sg read as weighted is refused as the reference refuses it
    import ctypes as C
    h, n = C.c_void_p(), C.c_int64()
    assert gpu.lib().gg_graph_read(str(good).encode(), 1, C.byref(h), C.byref(n), None) == -22
