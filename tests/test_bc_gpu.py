"""GPU Brandes parity: gg_bc against the oracles.

FP32 mode reproduces the reference's float arithmetic: bit-identical to the
float oracle when path counts are exact integers (always at these sizes),
hence the reference's BCVerifier rule (|diff| <= FLT_EPSILON) holds.  The
Kahan mode stays within 2e-6 max-abs of the FP64 oracle and at least as
close to it as the reference's float result; FP64 mode within 1e-6
relative of the FP64 oracle (the output is float).
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


GRAPHS = [("g10", lambda: Graph.kronecker(10), 4), ("g16", lambda: Graph.kronecker(16), 4),
          ("u12", lambda: Graph.uniform(12), 8), ("k18s", lambda: Graph.kronecker(18, seed=8675309), 4)]


@pytest.mark.parametrize("tag,make,k", GRAPHS)
def test_bc_fp32_and_fp64(tag, make, k):
    print(f"[bc] {tag}", flush=True)
    g = cached_graph(tag, make)
    dg = dev(tag, make)
    roots = g.pick_sources(k)
    want = oracle.brandes(g, roots, fp64=True)
    ref32 = oracle.brandes(g, roots)
    r32 = gpu.bc(dg, roots)
    assert np.array_equal(r32.scores, ref32)
    if g.num_nodes <= 70000:
        assert g.verify_bc(roots, r32.scores)
    rk = gpu.bc(dg, roots, kahan=True)
    e_k = np.abs(rk.scores.astype(np.float64) - want)
    assert e_k.max() < 2e-6
    assert e_k.sum() <= np.abs(ref32.astype(np.float64) - want).sum() * 1.01
    r64 = gpu.bc(dg, roots, fp64=True)
    nz = want > 1e-3
    assert (np.abs(r64.scores - want)[nz] / want[nz]).max() < 1e-6
    assert np.abs(r64.scores - want).max() < 1e-7


def test_bc_closed_forms():
    print("[bc] closed forms", flush=True)
    n = 400
    path = Graph.from_edges(np.arange(n - 1), np.arange(1, n))
    s = gpu.bc(gpu.DeviceGraph(path), [0], fp64=True).scores
    want = (n - 1 - np.arange(n)) / (n - 1)
    np.testing.assert_allclose(s, want, rtol=1e-6)
    m = 3000  # star with a hub above the chunk threshold
    star = Graph.from_edges(np.zeros(m - 1, np.int32), np.arange(1, m))
    roots = np.arange(0, m, 37)
    got = gpu.bc(gpu.DeviceGraph(star), roots).scores
    assert np.array_equal(got, oracle.brandes(star, roots))


def test_bc_repeated_root_and_disconnected():
    g = Graph.from_edges([0, 1, 2, 5, 6], [1, 2, 3, 6, 7])
    dg = gpu.DeviceGraph(g)
    roots = [0, 0, 5]
    assert np.array_equal(gpu.bc(dg, roots).scores, oracle.brandes(g, roots))
    with pytest.raises(gpu.GpuError):
        gpu.bc(dg, [g.num_nodes])


def test_bc_hub_folds():
    # Vertices with 16384+ successors fold on a block (terms in parallel,
    # folded in edge order): two such hubs sharing neighbours, a partial
    # last round, roots around them; FP32 bit-exact with the reference's
    # float arithmetic, Kahan and FP64 within the stated tolerances.
    print("[bc] hub folds", flush=True)
    rng = np.random.default_rng(23)
    n = 40000
    u, v = [], []
    for h, d in ((0, 30000), (1, 16384)):
        u.append(np.full(d, h))
        v.append(rng.choice(np.arange(2, n), size=d, replace=False))
    u.append(rng.integers(2, n, 50000))
    v.append(rng.integers(2, n, 50000))
    g = Graph.from_edges(np.concatenate(u), np.concatenate(v))
    dg = gpu.DeviceGraph(g)
    roots = np.array([0, 1, 7, 123, 4567, 30001], dtype=np.int32)
    assert np.array_equal(gpu.bc(dg, roots).scores, oracle.brandes(g, roots))
    want = oracle.brandes(g, roots, fp64=True)
    assert np.abs(gpu.bc(dg, roots, fp64=True).scores - want).max() < 1e-6
    assert np.abs(gpu.bc(dg, roots, kahan=True).scores - want).max() < 1e-5
    dg.profile(True)
    gpu.bc(dg, roots[:1])
    assert "bc_delta_hub" in dg.profile_report()


def test_bc_deep_graph_uses_level_by_level_discovery():
    # More levels than the depth buckets hold (4096): discovery falls back to
    # the top-down level loop; results as the reference's.
    print("[bc] path of 5000", flush=True)
    g = Graph.from_edges(np.arange(4999), np.arange(1, 5000))
    dg = gpu.DeviceGraph(g)
    roots = np.array([0, 2500, 4999], dtype=np.int32)
    assert np.array_equal(gpu.bc(dg, roots).scores, oracle.brandes(g, roots))
