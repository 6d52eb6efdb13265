"""GPU BFS parity: gg_bfs (through the C-ABI) against the CPU oracle.

Depths must be bit-exact; parents are validated by the reference's verifier
rules (gh_verify_bfs, restating src/bfs.cc:201-247).  Because the GPU takes
the reference's direction schedule, the number of top-down / bottom-up
steps and the adjacency entries examined must also equal the oracle's.
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


def check_root(g, dg, r, alpha=15, beta=18):
    res = gpu.bfs(dg, int(r), alpha, beta)
    want = oracle.bfs_depth(g, int(r))
    assert np.array_equal(res.depth, want), f"root {r}"
    assert g.verify_bfs(int(r), res.parent), f"root {r}"
    assert np.array_equal(oracle.depth_from_parent(res.parent, int(r)), want)
    _, _, st = oracle.dobfs(g, int(r), alpha, beta)
    s = res.stats
    assert (s["td_steps"], s["bu_steps"]) == (st["td_steps"], st["bu_steps"]), f"root {r}"
    assert s["edges_examined"] == st["edges_examined"], f"root {r}"
    assert s["vertices_examined"] == st["vertices_examined"], f"root {r}"
    assert s["reached"] == st["reached"] and s["reached_edges"] == st["reached_edges"]
    return res


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u10", lambda: Graph.uniform(10))])
def test_bfs_matches_oracle_sourcepicker_roots(golden, tag, make):
    print(f"[bfs] {tag}", flush=True)
    g = cached_graph(tag, make)
    dg = dev(tag, make)
    roots = golden["roots"][tag]
    assert g.pick_sources(16).tolist() == roots
    for r in roots:
        check_root(g, dg, r)


def test_bfs_seeded_kronecker_and_uniform():
    print("[bfs] k18 seed 8675309 / u16 seed 1234", flush=True)
    for tag, make in (("k18s", lambda: Graph.kronecker(18, seed=8675309)),
                      ("u16s", lambda: Graph.uniform(16, seed=1234))):
        g = cached_graph(tag, make)
        dg = dev(tag, make)
        for r in g.pick_sources(4):
            check_root(g, dg, r)


@pytest.mark.parametrize("alpha,beta", [(1, 18), (15, 1), (1000000, 18), (2, 2)])
def test_bfs_alpha_beta_variants(alpha, beta):
    print(f"[bfs] alpha={alpha} beta={beta}", flush=True)
    g = cached_graph("g16", lambda: Graph.kronecker(16))
    dg = dev("g16", lambda: Graph.kronecker(16))
    for r in g.pick_sources(3):
        check_root(g, dg, r, alpha, beta)


def test_bfs_closed_forms():
    print("[bfs] closed forms", flush=True)
    n = 2000
    path = Graph.from_edges(np.arange(n - 1), np.arange(1, n))
    res = gpu.bfs(gpu.DeviceGraph(path), 0)
    assert res.depth.tolist() == list(range(n))
    side = 96
    r, c = np.divmod(np.arange(side * side), side)
    ids = np.arange(side * side)
    u = np.concatenate([ids[c < side - 1], ids[r < side - 1]])
    w = np.concatenate([ids[c < side - 1] + 1, ids[r < side - 1] + side])
    lat = Graph.from_edges(u, w)
    src = 17 * side + 40
    res = gpu.bfs(gpu.DeviceGraph(lat), src)
    assert np.array_equal(res.depth, np.abs(r - 17) + np.abs(c - 40))
    assert lat.verify_bfs(src, res.parent)
    iu, iv = np.triu_indices(300, 1)  # complete graph: one level, hubs > 256 edges
    k = Graph.from_edges(iu, iv)
    res = gpu.bfs(gpu.DeviceGraph(k), 11)
    assert res.depth[11] == 0 and set(res.depth.tolist()) == {0, 1}
    assert k.verify_bfs(11, res.parent)


@pytest.mark.parametrize("alpha", [15, 1000000])
def test_bfs_hub_chunks_star_of_stars(alpha):
    # Two hubs of 9000 and 5000 leaves (> 4096-edge chunks) joined by an edge,
    # so the top-down step from a hub spreads its edges over the grid.
    print(f"[bfs] star of stars alpha={alpha}", flush=True)
    a = np.zeros(9000, np.int32)
    b = np.full(5000, 1, np.int32)
    u = np.concatenate([a, b, [0]])
    v = np.concatenate([np.arange(2, 9002), np.arange(9002, 14002), [1]])
    g = Graph.from_edges(u, v)
    dg = gpu.DeviceGraph(g)
    for r in (0, 5, 9500):
        check_root(g, dg, r, alpha)


def test_bfs_disconnected_and_isolated_source():
    print("[bfs] disconnected", flush=True)
    g = Graph.from_edges([0, 1, 3, 5, 9], [1, 2, 4, 6, 9])  # 9-9 self-loop is squished away
    dg = gpu.DeviceGraph(g)
    res = gpu.bfs(dg, 0)
    assert res.parent.tolist() == [0, 0, 1, -1, -1, -1, -1, -1, -1, -1]
    assert res.depth.tolist() == [0, 1, 2, -1, -1, -1, -1, -1, -1, -1]
    res = gpu.bfs(dg, 9)  # degree-0 root (reachable with -r)
    assert res.parent.tolist() == [-1] * 9 + [9]


def test_bfs_directed_graph_uses_inverse_for_bottom_up():
    print("[bfs] directed", flush=True)
    rng = np.random.default_rng(5)
    n, m = 3000, 40000
    u = rng.integers(0, n, m)
    v = rng.integers(0, n, m)
    g = Graph.from_edges(u, v, symmetrize=False)
    dg = gpu.DeviceGraph(g)
    for r in g.pick_sources(4):
        for alpha in (15, 1):
            check_root(g, dg, r, alpha)


def test_bfs_errors():
    g = cached_graph("g10", lambda: Graph.kronecker(10))
    dg = dev("g10", lambda: Graph.kronecker(10))
    with pytest.raises(gpu.GpuError) as e:
        gpu.bfs(dg, g.num_nodes)
    assert e.value.code == -22
    with pytest.raises(gpu.GpuError) as e:
        gpu.bfs(dg, 0, alpha=0)
    assert e.value.code == -22
    w = gpu.DeviceGraph(Graph.kronecker(8, weighted=True))
    with pytest.raises(gpu.GpuError) as e:
        gpu.bfs(w, 0)
    assert e.value.code == -95


def _raw_upload(n, off, dst, directed=False, ioff=None, idst=None):
    import ctypes as C
    h = C.c_void_p()
    p64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int64)) if a is not None else None  # noqa: E731
    p32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32)) if a is not None else None  # noqa: E731
    rc = gpu.lib().gg_graph_upload(n, p64(off), p32(dst), 0, int(directed), p64(ioff), p32(idst), C.byref(h))
    if rc == 0:
        gpu.lib().gg_graph_free(h)
    return rc


def test_upload_errors():
    """Malformed CSR arrays are refused on the host or by the device
    validation pass, never handed to a kernel."""
    gpu.init(0)
    off = np.array([0, 1, 2], dtype=np.int64)
    dst = np.array([1, 0], dtype=np.int32)
    assert _raw_upload(2, off, dst) == 0
    assert _raw_upload(2, off, np.array([1, 2], dtype=np.int32)) == -22  # destination == n
    assert _raw_upload(2, np.array([0, 2, 1], dtype=np.int64), dst) == -22  # decreasing
    assert _raw_upload(2, np.array([1, 1, 2], dtype=np.int64), dst) == -22  # out side not from 0
    # A directed upload whose inverse offsets start below 0 (they would index
    # in_dests before its start) or above it.
    assert _raw_upload(2, off, dst, True, off, dst) == 0
    for bad in ([-4, 1, 2], [1, 1, 2]):
        assert _raw_upload(2, off, dst, True, np.array(bad, dtype=np.int64), dst) == -22, bad


def test_memory_cap(monkeypatch):
    """GAPGPU_MEM_CAP_BYTES caps the library's device allocations: an upload
    beyond it returns GG_ENOMEM (a host-side error return) and leaves nothing
    allocated, and uploads work again once the cap is lifted."""
    gpu.init(0)
    g = Graph.kronecker(10)
    monkeypatch.setenv("GAPGPU_MEM_CAP_BYTES", "4096")
    with pytest.raises(gpu.GpuError) as e:
        gpu.DeviceGraph(g)
    assert e.value.code == -12
    assert "GAPGPU_MEM_CAP_BYTES" in str(e.value)
    monkeypatch.delenv("GAPGPU_MEM_CAP_BYTES")
    dg = gpu.DeviceGraph(g)
    r = int(g.pick_sources(1)[0])
    assert np.array_equal(gpu.bfs(dg, r).depth, oracle.bfs_depth(g, r))


def test_bfs_parent_only_matches_depth_run():
    g = cached_graph("g16", lambda: Graph.kronecker(16))
    dg = dev("g16", lambda: Graph.kronecker(16))
    r = int(g.pick_sources(1)[0])
    p = gpu.DOBFS(dg, r)
    assert np.array_equal(oracle.depth_from_parent(p, r), oracle.bfs_depth(g, r))
