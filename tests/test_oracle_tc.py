"""Pins the CPU oracle's triangle counts (orc_ordered_count, orc_tc_intersect).

Checks: the reference's counts and relabelling decision (golden), the
reference built from source where present, closed forms (complete graph
C(n,3), wheel n, bipartite lattice 0), invariance under relabelling, and
the reference's TCVerifier rule.
"""

import math

import numpy as np
import pytest

from conftest import cached_graph
from gapbs_amd.graph import Graph
from oracle import oracle, reference as ref

GRAPHS = {"g10": lambda: Graph.kronecker(10), "g16": lambda: Graph.kronecker(16),
          "u12": lambda: Graph.uniform(12)}


@pytest.mark.parametrize("tag", sorted(GRAPHS))
def test_counts_match_reference(golden, tag):
    g = cached_graph(tag, GRAPHS[tag])
    want = golden["tc"][tag]["triangles"]
    assert oracle.ordered_count(g)[0] == want
    assert oracle.tc_intersect(g) == want
    assert oracle.ordered_count(g.relabel_by_degree())[0] == want
    assert g.worth_relabelling() == golden["tc"][tag]["worth_relabelling"]
    assert g.verify_tc(want) and not g.verify_tc(want + 1)


@pytest.mark.skipif(not ref.available("tc"), reason="reference not built here")
def test_counts_match_reference_build_seeded():
    g = Graph.kronecker(13, seed=99)
    assert oracle.ordered_count(g)[0] == ref.tc(g)[0] == ref.tc(g, hybrid=False)[0]


def wheel(n):
    u = np.concatenate([np.arange(n), np.arange(n)])
    v = np.concatenate([(np.arange(n) + 1) % n, np.full(n, n)])
    return Graph.from_edges(u, v)


def test_closed_forms():
    iu, iv = np.triu_indices(60, 1)
    assert oracle.ordered_count(Graph.from_edges(iu, iv))[0] == math.comb(60, 3)
    assert oracle.ordered_count(wheel(500))[0] == 500
    side = 30
    ids = np.arange(side * side)
    r, c = np.divmod(ids, side)
    lat = Graph.from_edges(np.concatenate([ids[c < side - 1], ids[r < side - 1]]),
                           np.concatenate([ids[c < side - 1] + 1, ids[r < side - 1] + side]))
    assert oracle.ordered_count(lat)[0] == 0


def test_sampled_counts_partition_the_full_count(golden):
    # The bounded CPU sample times OrderedCount's outer loop over one residue
    # class of u; the classes must partition the count exactly.
    g = cached_graph("g16", GRAPHS["g16"]).relabel_by_degree()
    want = golden["tc"]["g16"]["triangles"]
    parts = [oracle.ordered_count_sample(g, 7, first)[0] for first in range(7)]
    assert sum(parts) == want and all(p > 0 for p in parts)
