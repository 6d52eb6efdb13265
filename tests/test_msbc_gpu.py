"""GPU batched Brandes parity: gg_bc_batch (through the C-ABI) against the oracles.

The reference runs Brandes one root at a time and adds each root's
dependencies to the float scores in turn (src/bc.cc:108-140).  In FP32 mode
the batched kernel folds every root's dependencies with the reference's
arithmetic and adds them in root order, so its scores must equal the float
oracle (oracle.brandes, bit-identical to the reference at these sizes) and
the single-root GPU path bit for bit; FP64 mode stays within 1e-6 relative
of the FP64 oracle.
"""

import ctypes

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


def check(g, dg, roots, single=False):
    roots = np.asarray(roots, dtype=np.int32)
    r32 = gpu.bc_batch(dg, roots)
    want32 = oracle.brandes(g, roots)
    assert np.array_equal(r32.scores, want32), "fp32 differs from the reference arithmetic"
    if single:
        assert np.array_equal(r32.scores, gpu.bc(dg, roots).scores)
    want64 = oracle.brandes(g, roots, fp64=True)
    r64 = gpu.bc_batch(dg, roots, fp64=True)
    nz = want64 > 1e-3
    if nz.any():
        assert (np.abs(r64.scores - want64)[nz] / want64[nz]).max() < 1e-6
    assert np.abs(r64.scores - want64).max() < 1e-7
    reached = sum(int((oracle.bfs_depth(g, int(r)) >= 0).sum()) for r in roots)
    assert r32.stats["reached"] == reached
    return r32


@pytest.mark.parametrize("tag,make", [("g10", lambda: Graph.kronecker(10)),
                                      ("g16", lambda: Graph.kronecker(16)),
                                      ("u12", lambda: Graph.uniform(12))])
def test_batch_bc_64_roots(tag, make):
    print(f"[msbc] {tag} 64 roots", flush=True)
    g = cached_graph(tag, make)
    check(g, dev(tag, make), g.pick_sources(64), single=(tag == "g10"))


def test_batch_bc_partial_batches_seeded():
    print("[msbc] k18 seed 8675309: 1, 5, 33 roots", flush=True)
    make = lambda: Graph.kronecker(18, seed=8675309)  # noqa: E731
    g = cached_graph("k18s", make)
    dg = dev("k18s", make)
    roots = g.pick_sources(64)
    for k in (1, 5, 33):
        check(g, dg, roots[:k], single=(k <= 5))


def test_batch_bc_repeated_isolated_and_heavy():
    # Roots repeated (their dependencies count twice, as in the reference's
    # loop), a degree-0 root, and two hubs above the heavy-chunk threshold.
    print("[msbc] repeated / isolated / heavy", flush=True)
    a = np.zeros(3000, np.int32)
    b = np.full(2500, 1, np.int32)
    u = np.concatenate([a, b, [0, 7000]])
    v = np.concatenate([np.arange(2, 3002), np.arange(3002, 5502), [1, 7000]])
    g = Graph.from_edges(u, v)
    dg = gpu.DeviceGraph(g)
    check(g, dg, [5, 5, 3500, 0, 7000, 3001, 1, 5], single=True)


def test_batch_bc_directed():
    print("[msbc] directed", flush=True)
    rng = np.random.default_rng(11)
    n, m = 1 << 12, 1 << 15
    g = Graph.from_edges(rng.integers(0, n, m), rng.integers(0, n, m), symmetrize=False)
    dg = gpu.DeviceGraph(g)
    check(g, dg, rng.choice(n, 40, replace=False), single=True)


def test_batch_bc_path_closed_form():
    print("[msbc] path", flush=True)
    n = 300
    g = Graph.from_edges(np.arange(n - 1), np.arange(1, n))
    s = gpu.bc_batch(gpu.DeviceGraph(g), [0], fp64=True).scores
    want = (n - 1 - np.arange(n)) / (n - 1)  # dependency of i on root 0, the root's own included
    assert np.abs(s - want).max() < 1e-6


def test_batch_bc_errors():
    print("[msbc] errors", flush=True)
    dg = gpu.DeviceGraph(Graph.kronecker(10))
    for roots in ([], list(range(65))):
        with pytest.raises(gpu.GpuError) as e:
            gpu.bc_batch(dg, roots)
        assert e.value.code == -22
    one = np.zeros(1, np.int32)
    rc = gpu.lib().gg_bc_batch(dg.handle, one.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), 1,
                               gpu.BC_FP32_KAHAN, None, None)
    assert rc == -95
    wdg = gpu.DeviceGraph(Graph.kronecker(10, weighted=True))
    with pytest.raises(gpu.GpuError) as e:
        gpu.bc_batch(wdg, [0])
    assert e.value.code == -95


@pytest.mark.parametrize("ordered", [False, True])
def test_batch_bc_hub_folds(ordered):
    # Vertices with 2048+ successors have their dependency folds split over
    # a block (terms computed in parallel, folded in list order by one
    # wavefront), in wide blocks from 65536 successors: four such hubs
    # sharing neighbours (one wide), partial last rounds, and roots placed
    # so the hubs sit at several levels.
    print(f"[msbc] hub folds ordered={ordered}", flush=True)
    rng = np.random.default_rng(17)
    n = 150000
    hubs = [0, 1, 2, 3]
    degs = [70001, 20000, 5000, 2048]
    u, v = [], []
    for h, d in zip(hubs, degs):
        nb = rng.choice(np.arange(4, n), size=d, replace=False)
        u.append(np.full(d, h))
        v.append(nb)
    u.append(rng.integers(4, n, 300000))
    v.append(rng.integers(4, n, 300000))
    g = Graph.from_edges(np.concatenate(u), np.concatenate(v))
    dg = gpu.DeviceGraph(g, order_by_degree=ordered)
    roots = np.concatenate([hubs, rng.choice(np.arange(4, n), 28, replace=False)]).astype(np.int32)
    # Degree-ordered or not, the folds read each vertex's list in uploaded
    # order, so the scores equal the oracle's float Brandes bit for bit.
    check(g, dg, roots)
    dg.profile(True)
    gpu.bc_batch(dg, roots)
    rep = dg.profile_report()
    assert "msbc_delta_hub" in rep and "msbc_delta_hub_wide" in rep
