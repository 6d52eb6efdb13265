"""GPU autotuner (gg_autotune): picks settings from its candidate sets, keeps
them in the handle, and every kernel stays exact under what it picks (the
settings move only the time).  Also the drivers' -q flag end to end."""

import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, cached_graph
from gapbs_amd import gpu
from gapbs_amd.graph import Graph
from oracle import oracle

pytestmark = pytest.mark.gpu

ALPHAS = {2, 4, 8, 15, 30, 60, 120, 240}
BETAS = {2, 6, 18, 54, 162}
BATCH_ALPHAS = {1, 2, 4, 8, 15, 30, 60, 120, 1000}


def test_tune_bfs_and_batch():
    print("[tune] bfs / batch g16", flush=True)
    g = cached_graph("g16", lambda: Graph.kronecker(16))
    dg = gpu.DeviceGraph(g)
    roots = g.pick_sources(64)
    t = gpu.autotune(dg, gpu.TUNE_BFS | gpu.TUNE_BATCH, roots)
    assert t["bfs_alpha"] in ALPHAS and t["bfs_beta"] in BETAS and t["batch_alpha"] in BATCH_ALPHAS
    assert t["tuned"] == gpu.TUNE_BFS | gpu.TUNE_BATCH and t["best_ms"][0] > 0 and t["best_ms"][1] > 0
    assert gpu.tuning(dg) == t
    for r in roots[:4]:
        res = gpu.bfs(dg, int(r), t["bfs_alpha"], t["bfs_beta"])
        assert np.array_equal(res.depth, oracle.bfs_depth(g, int(r)))
        assert g.verify_bfs(int(r), res.parent)
    b = gpu.bfs_batch(dg, roots, t["batch_alpha"])
    for i in (0, 31, 63):
        assert np.array_equal(b.depths[i], oracle.bfs_depth(g, int(roots[i])))


def test_tune_sssp():
    print("[tune] sssp weighted g14", flush=True)
    wg = Graph.kronecker(14, weighted=True)
    dwg = gpu.DeviceGraph(wg)
    roots = wg.pick_sources(2)
    t = gpu.autotune(dwg, gpu.TUNE_SSSP, roots)
    d = t["sssp_delta"]
    assert d & (d - 1) == 0 and 1 <= d <= 4096
    for r in roots:
        assert np.array_equal(gpu.sssp(dwg, int(r), d).dist, oracle.dijkstra(wg, int(r)))
    with pytest.raises(gpu.GpuError) as e:  # SSSP tuning needs weights
        gpu.autotune(gpu.DeviceGraph(Graph.kronecker(10)), gpu.TUNE_SSSP, [0])
    assert e.value.code == -95


def test_tune_pagerank_and_explicit_bins():
    print("[tune] pagerank g16", flush=True)
    g = cached_graph("g16", lambda: Graph.kronecker(16))
    dg = gpu.DeviceGraph(g)
    want, _, _ = oracle.pagerank(g, 20, 1e-4, fp64=True)
    t = gpu.autotune(dg, gpu.TUNE_PR)
    assert t["pr_small_max"] in {8, 16, 32, 64, 128} and t["pr_medium_max"] in {256, 512, 1024, 2048, 8192}
    assert t["pr_blocks_per_cu"] in {4, 8, 16, 32}
    assert np.abs(gpu.pagerank(dg, 20, 1e-4).scores - want).max() < 1e-6
    # Extreme cut-offs: every vertex with 2+ in-edges takes the block path.
    gpu.set_tuning(dg, {"pr_small_max": 1, "pr_medium_max": 2, "pr_blocks_per_cu": 1})
    assert np.abs(gpu.pagerank(dg, 20, 1e-4).scores - want).max() < 1e-6
    gpu.set_tuning(dg, {"pr_small_max": 4096, "pr_medium_max": 1 << 30})  # all thread-per-vertex
    assert np.abs(gpu.pagerank(dg, 20, 1e-4).scores - want).max() < 1e-6
    with pytest.raises(gpu.GpuError) as e:
        gpu.set_tuning(dg, {"pr_small_max": 64, "pr_medium_max": 64})
    assert e.value.code == -22


def test_tune_pagerank_block_width():
    # A graph of 2^20 vertices runs Jacobi PageRank through propagation
    # blocks (default width 2^10): the tuner times that width and its
    # neighbours, keeps one, and the scores stay within 1e-7 of the plain
    # pull's at any of them.
    print("[tune] pagerank blocks k20", flush=True)
    dg, _ = gpu.generate_graph(0, 20, 16, 27491095)
    assert gpu.pagerank_blocks(dg)["width_log2"] == 10
    gpu.set_tuning(dg, {"pr_blocks_log2": -1, "pr_segment_log2": [-1, -1]})
    plain = gpu.pagerank(dg, 20, 1e-4)
    gpu.set_tuning(dg, {"pr_blocks_log2": 0, "pr_segment_log2": [0, 0]})
    t = gpu.autotune(dg, gpu.TUNE_PR)
    assert t["pr_blocks_log2"] in {0, 9, 11} and t["tuned"] & gpu.TUNE_PR and t["best_ms"][3] > 0
    assert gpu.pagerank_blocks(dg)["width_log2"] == {0: 10}.get(t["pr_blocks_log2"], t["pr_blocks_log2"])
    r = gpu.pagerank(dg, 20, 1e-4)
    assert r.stats["iterations"] == plain.stats["iterations"]
    assert np.abs(r.scores.astype(np.float64) - plain.scores).max() < 1e-7
    dg.free()


@pytest.mark.parametrize("kernel,extra", [("bfs", ["-n2"]), ("sssp", ["-n2"]), ("pr", ["-n1"])])
def test_driver_autotune_flag(kernel, extra):
    print(f"[tune] driver {kernel} -q", flush=True)
    p = subprocess.run([os.path.join(ROOT, "gapbs_amd", "bin", kernel), "-g12", "-v", "-q"] + extra,
                       capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stdout
    assert "Autotune:" in p.stdout and "Verification:           PASS" in p.stdout, p.stdout
    assert "FAIL" not in p.stdout
