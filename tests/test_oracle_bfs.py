"""Pins the CPU oracle's BFS (oracle/oracle.cc) before it is trusted.

Checks: the reference's DOBFS built from source (oracle/_ref) where present,
the golden digests of the reference's depths (tests/golden/golden.json), and
closed forms (path, star, complete graph, lattice, disconnected parts) that a
dropped level, an off-by-one or a wrong direction would fail.
"""

import hashlib

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd.graph import Graph
from oracle import oracle, reference as ref


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u10", lambda: Graph.uniform(10))])
def test_oracle_depths_match_reference_digests(golden, tag, make):
    g = cached_graph(tag, make)
    for rec in golden["bfs"][tag]:
        r = rec["root"]
        d = oracle.bfs_depth(g, r)
        assert sha(d) == rec["depth_sha256"]
        p, _, st = oracle.dobfs(g, r)
        assert sha(oracle.depth_from_parent(p, r)) == rec["depth_sha256"]
        assert st["reached"] == rec["reached"]
        assert g.verify_bfs(r, p)


@pytest.mark.skipif(not ref.available("bfs"), reason="reference not built here")
def test_oracle_dobfs_matches_reference_build_on_seeded_graph():
    g = Graph.kronecker(14, seed=8675309)
    roots = g.pick_sources(8)
    parents, _ = ref.bfs(ref.RefGraph("bfs", g), roots)
    for r, pr in zip(roots, parents):
        p, _, _ = oracle.dobfs(g, int(r))
        assert np.array_equal(oracle.depth_from_parent(p, int(r)),
                              oracle.depth_from_parent(pr, int(r)))


def _path(n):
    return Graph.from_edges(np.arange(n - 1), np.arange(1, n))


def test_closed_form_path():
    g = _path(300)
    d = oracle.bfs_depth(g, 0)
    assert d.tolist() == list(range(300))
    p, _, st = oracle.dobfs(g, 0)
    assert p[1:].tolist() == list(range(299))
    assert st["td_steps"] + st["bu_steps"] == 300  # one step per level, last one empty


def test_closed_form_star_and_complete():
    n = 50
    star = Graph.from_edges(np.zeros(n - 1, np.int32), np.arange(1, n))
    assert oracle.bfs_depth(star, 7).tolist() == [1] + [2] * 6 + [0] + [2] * (n - 8)
    iu, iv = np.triu_indices(40, 1)
    k = Graph.from_edges(iu, iv)
    p, _, _ = oracle.dobfs(k, 3)
    d = oracle.depth_from_parent(p, 3)
    assert d[3] == 0 and set(d.tolist()) == {0, 1}


def test_closed_form_lattice_manhattan():
    side = 64
    r, c = np.divmod(np.arange(side * side), side)
    h = (c < side - 1)
    v = (r < side - 1)
    ids = np.arange(side * side)
    u = np.concatenate([ids[h], ids[v]])
    w = np.concatenate([ids[h] + 1, ids[v] + side])
    g = Graph.from_edges(u, w)
    src = 5 * side + 9
    d = oracle.bfs_depth(g, src)
    assert np.array_equal(d, np.abs(r - 5) + np.abs(c - 9))
    p, _, _ = oracle.dobfs(g, src)
    assert np.array_equal(oracle.depth_from_parent(p, src), d)


def test_disconnected_vertices_unreached():
    g = Graph.from_edges([0, 1, 3, 5], [1, 2, 4, 6])
    d = oracle.bfs_depth(g, 0)
    assert d.tolist() == [0, 1, 2, -1, -1, -1, -1]
    p, _, _ = oracle.dobfs(g, 0)
    assert p.tolist() == [0, 0, 1, -1, -1, -1, -1]


def test_planted_parent_mistake_is_caught():
    g = cached_graph("g10", lambda: Graph.kronecker(10))
    r = int(g.pick_sources(1)[0])
    p, _, _ = oracle.dobfs(g, r)
    bad = p.copy()
    v = int(np.flatnonzero(bad >= 0)[5])
    if v == r:
        v = int(np.flatnonzero(bad >= 0)[6])
    bad[v] = v  # a self-parent is neither an edge nor one level up
    assert not g.verify_bfs(r, bad)
    with pytest.raises(ValueError):
        oracle.depth_from_parent(bad, r)
