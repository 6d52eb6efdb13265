"""The allocation-free Afforest entry point (gg_cc_noalloc) against gg_cc.

gg_cc_noalloc runs gg_cc's kernels with every device buffer carved out of a
fixed scratch array per GPU, so two calls in a row share (and inherit) its
contents.  Each case calls it twice, the second time on a smaller graph,
and checks both results against gg_cc on the same handle bit for bit, and
gg_cc against the oracle's Afforest (src/cc.cc:95-151).  Calls go through
the C-ABI (ctypes) as a caller that cannot allocate would make them.
"""

import ctypes as C

import numpy as np
import pytest

from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu


def _noalloc(dg, rounds=2):
    comp = np.empty(dg.num_nodes, dtype=np.int32)
    st = gpu.Stats()
    rc = gpu.lib().gg_cc_noalloc(dg.handle, rounds, comp.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st))
    return rc, comp


def _check_pair(graphs, order_by_degree=False, rounds=2):
    for g in graphs:
        dg = gpu.DeviceGraph(g, order_by_degree=order_by_degree)
        rc, got = _noalloc(dg, rounds)
        assert rc == 0, gpu.lib().gg_last_error().decode()
        want = gpu.cc(dg, rounds).comp
        assert np.array_equal(got, want)
        assert np.array_equal(want, oracle.afforest(g)[0])  # labels do not depend on the rounds
        dg.free()


def test_two_calls_second_smaller():
    print("[noalloc] kronecker 16 then 12", flush=True)
    _check_pair([Graph.kronecker(16), Graph.kronecker(12)])


def test_degree_ordered_and_rounds():
    print("[noalloc] degree-ordered kronecker 15 then uniform 11; rounds 0 and 5", flush=True)
    _check_pair([Graph.kronecker(15), Graph.uniform(11)], order_by_degree=True)
    _check_pair([Graph.kronecker(14), Graph.kronecker(10)], rounds=0)
    _check_pair([Graph.kronecker(14), Graph.kronecker(10)], rounds=5)


def test_directed_and_hub_chunks():
    """A directed graph (in-lists linked too) and a star whose hub list is
    cut into 4096-entry chunks held in the scratch array."""
    print("[noalloc] directed, then star of 20000", flush=True)
    rng = np.random.default_rng(3)
    d = Graph.from_edges(rng.integers(0, 5000, 30000), rng.integers(0, 5000, 30000), symmetrize=False)
    n = 20000
    star = Graph.from_edges(np.zeros(n - 1, dtype=np.int64), np.arange(1, n))
    _check_pair([d, star])


def test_oversize_is_refused():
    """A graph whose arrays exceed the fixed scratch is refused on the host
    with GG_EINVAL and a message naming the bytes it needs."""
    print("[noalloc] oversize", flush=True)
    n = 20_000_000  # 80 MB of labels > 64 MiB
    g = Graph.from_edges(np.array([0, 5]), np.array([n - 1, 7]))
    dg = gpu.DeviceGraph(g)
    rc, _ = _noalloc(dg)
    assert rc == -22
    assert "bytes of scratch" in gpu.lib().gg_last_error().decode()
    # The handle still works through the allocating entry point, and the
    # allocation-free one afterwards on a small graph.
    comp = gpu.cc(dg).comp
    assert comp[n - 1] == 0 and comp[7] == 5 and comp[1] == 1
    _check_pair([Graph.kronecker(10)])
