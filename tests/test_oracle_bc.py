"""Pins the CPU oracle's Brandes (orc_brandes_f32 / orc_brandes_f64).

Checks: golden digests of the reference's Brandes scores (the float oracle
must be bit-identical), the reference built from source where present, the
reference's BCVerifier rule, and closed forms: a path from an end root
(dependency of vertex i is n-1-i, root included), a star (all paths through
the centre), and brute-force shortest-path counting on a small graph.
"""

import hashlib
import itertools

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd.graph import Graph
from oracle import oracle, reference as ref


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


GRAPHS = {"g10": lambda: Graph.kronecker(10), "g16": lambda: Graph.kronecker(16),
          "u12": lambda: Graph.uniform(12)}


@pytest.mark.parametrize("tag", sorted(GRAPHS))
def test_float_oracle_bit_identical_to_reference(golden, tag):
    g = cached_graph(tag, GRAPHS[tag])
    rec = golden["bc"][tag]
    assert g.pick_sources(rec["k"]).tolist() == rec["roots"]
    s = oracle.brandes(g, rec["roots"])
    assert sha(s) == rec["scores_sha256"]
    assert g.verify_bc(rec["roots"], s)


@pytest.mark.skipif(not ref.available("bc"), reason="reference not built here")
def test_float_oracle_matches_reference_build_seeded():
    g = Graph.kronecker(13, seed=77)
    r, _ = ref.brandes(g, 6)
    assert np.array_equal(r, oracle.brandes(g, g.pick_sources(6)))


def test_closed_form_path_from_end():
    n = 50
    g = Graph.from_edges(np.arange(n - 1), np.arange(1, n))
    s = oracle.brandes(g, [0], fp64=True)
    want = (n - 1 - np.arange(n)).astype(np.float64)
    np.testing.assert_allclose(s, want / want.max(), rtol=1e-15)


def test_closed_form_star_all_roots():
    n = 30
    g = Graph.from_edges(np.zeros(n - 1, np.int32), np.arange(1, n))
    s = oracle.brandes(g, np.arange(n), fp64=True)
    # Centre: (n-1)(n-2) pairs through it plus its own dependency n-1 as a
    # root; each leaf only its own dependency (n-1) as a root.
    centre = (n - 1) * (n - 2) + (n - 1)
    leaf = n - 1
    np.testing.assert_allclose(s[0], 1.0)
    np.testing.assert_allclose(s[1:], leaf / centre, rtol=1e-15)


def test_brute_force_small_graph():
    rng = np.random.default_rng(12)
    n = 12
    pairs = [(a, b) for a, b in itertools.combinations(range(n), 2) if rng.random() < 0.3]
    g = Graph.from_edges([p[0] for p in pairs], [p[1] for p in pairs])
    nn = g.num_nodes
    off, nbr = g.out_offsets, g.out_neighbors
    adj = [set(nbr[off[u]:off[u + 1]]) for u in range(nn)]
    roots = list(range(nn))
    # dependency by enumerating all shortest paths (includes the root's own
    # dependency on its targets, the reference's convention)
    score = np.zeros(nn)
    for s in roots:
        dist = {s: 0}
        frontier = [s]
        while frontier:
            nxt = []
            for u in frontier:
                for v in adj[u]:
                    if v not in dist:
                        dist[v] = dist[u] + 1
                        nxt.append(v)
            frontier = nxt
        paths = {s: [[s]]}
        for v in sorted(dist, key=dist.get):
            if v == s:
                continue
            paths[v] = [p + [v] for u in adj[v] if dist.get(u, -9) == dist[v] - 1 for p in paths[u]]
        for t in dist:
            if t == s:
                continue
            for p in paths[t]:
                for w in p[:-1]:
                    score[w] += 1.0 / len(paths[t])
    want = score / score.max()
    np.testing.assert_allclose(oracle.brandes(g, roots, fp64=True), want, rtol=1e-12, atol=1e-15)


def test_planted_score_mistake_fails_verifier():
    g = cached_graph("g10", GRAPHS["g10"])
    roots = g.pick_sources(4)
    s = oracle.brandes(g, roots)
    bad = s.copy()
    bad[np.argsort(s)[-2]] += 1e-4
    assert g.verify_bc(roots, s) and not g.verify_bc(roots, bad)
