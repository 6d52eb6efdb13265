"""Benchmark of the GPU kernel path on one MI355X.

Workload: a Kronecker graph (Graph500 R-MAT A=.57 B=.19 C=.19, edge factor
16, IDs permuted, symmetrised; default scale 22, generator seed 8675309)
built on the host and uploaded to HBM before timing.  One "step" runs every
kernel once on it, results left on the device:
  bfs   direction-optimizing BFS from each of R SourcePicker roots
  sssp  delta-stepping from the same R roots (weights uniform in [1, 255])
  pr    PageRank, damping 0.85, tolerance 1e-4, at most 20 iterations (FP32)
  cc    Afforest
  bc    Brandes from the first 4 roots (the reference's float arithmetic)
  tc    ordered triangle count on the graph relabelled by degree

Prints one JSON line.  value = aggregate BFS GTEPS (sum over roots of m_r,
the undirected edges of the reached component, over the BFS wall time);
every kernel's own throughput, time and roofline are under "kernels".  The
roofline is measured live: the library times each kernel launch with HIP
events in the dispatch packet and credits it with its algorithmic bytes.
The CPU baseline is the reference compiled from source (oracle/_ref) when
present, else this repo's OpenMP restatement, timed on the host cores on a
bounded sample (first roots only, one run of the others).
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BFS/SSSP/BC GTEPS, PR/CC/TC seconds, Kronecker s26 ef16; % HBM roofline"
HBM_PEAK_GBS = 8000.0
KERNELS = ("bfs", "sssp", "pr", "cc", "bc", "tc")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Timed:
    """Wall time of the enclosed calls, with the library's per-kernel profile."""

    def __init__(self, *dgs):
        self.dgs = dgs

    def __enter__(self):
        for dg in self.dgs:
            dg.profile(True)
            dg.profile_reset()
        self.t = time.perf_counter()
        return self

    def __exit__(self, *exc):
        self.seconds = time.perf_counter() - self.t
        self.prof = {}
        for dg in self.dgs:
            for k, v in dg.profile_report().items():
                p = self.prof.setdefault(k, {"launches": 0, "ms": 0.0, "alg_bytes": 0.0})
                for f in p:
                    p[f] += v[f]
            dg.profile(False)


def roofline(prof):
    """Roofline of the kernel with the most device time among those credited with bytes."""
    cands = {k: v for k, v in prof.items() if v["alg_bytes"] > 0 and v["ms"] > 0}
    if not cands:
        return None
    name, v = max(cands.items(), key=lambda kv: kv[1]["ms"])
    achieved = v["alg_bytes"] / (v["ms"] / 1e3) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
            "avg_launch_us": round(1e3 * v["ms"] / v["launches"], 2), "launches": v["launches"]}


def cpu_baselines(g, wg, rg_tc, roots, m_per_root, budget_s):
    """Reference (or port) CPU timings on bounded samples of the same workload."""
    import numpy as np
    from oracle import oracle, reference as ref
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {}
    use_ref = ref.available("bfs")
    kind = "reference" if use_ref else "port"
    # BFS / SSSP: the first roots until the budget is spent.
    secs, done, edges = 0.0, 0, 0
    rg = ref.RefGraph("bfs", g) if use_ref else None
    for r, m in zip(roots, m_per_root):
        secs += ref.bfs_time(rg, [r]) if use_ref else oracle.dobfs(g, r)[1]
        edges += m
        done += 1
        if secs > budget_s:
            break
    out["bfs"] = {"value": edges / secs / 1e9, "unit": "GTEPS", "sample": f"{done} roots", "seconds": secs}
    secs, done, edges = 0.0, 0, 0
    rgw = ref.RefGraph("sssp", wg) if use_ref else None
    for r, m in zip(roots, m_per_root):
        secs += ref.sssp_time(rgw, [r], 2) if use_ref else oracle.delta_step(wg, r, 2)[1]
        edges += m
        done += 1
        if secs > budget_s:
            break
    out["sssp"] = {"value": edges / secs / 1e9, "unit": "GTEPS", "sample": f"{done} roots, delta 2",
                   "seconds": secs}
    if use_ref:
        _, s = ref.pagerank("pr_spmv", g, 20, 1e-4)
        out["pr"] = {"value": s, "unit": "s", "sample": "one run (PageRankPull, Jacobi)"}
        _, s = ref.afforest(g)
        out["cc"] = {"value": s, "unit": "s", "sample": "one run"}
        _, s = ref.brandes(g, 4)
        out["bc"] = {"value": s, "unit": "s", "sample": "4 roots, one run"}
        _, s = ref.tc(rg_tc, hybrid=False)
        out["tc"] = {"value": s, "unit": "s", "sample": "OrderedCount on the relabelled graph, one run"}
    else:
        _, _, _ = oracle.pagerank(g, 20, 1e-4)
        t = time.perf_counter()
        oracle.pagerank(g, 20, 1e-4)
        out["pr"] = {"value": time.perf_counter() - t, "unit": "s", "sample": "one run"}
        out["cc"] = {"value": oracle.afforest(g)[1], "unit": "s", "sample": "one run"}
        out["tc"] = {"value": oracle.ordered_count(rg_tc)[1], "unit": "s",
                     "sample": "OrderedCount on the relabelled graph, one run"}
    for v in out.values():
        v["kind"] = kind
        v["cores"] = cores
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=22)
    ap.add_argument("--seed", type=int, default=8675309)
    ap.add_argument("--roots", type=int, default=16)
    ap.add_argument("--delta", type=int, default=32, help="SSSP bucket width on the GPU")
    ap.add_argument("--kernels", default=",".join(KERNELS))
    ap.add_argument("--cpu-budget", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one GPU, and the kernels "
              "run on a single device (multi-GPU sharding is a design note only).", file=sys.stderr)
        sys.exit(2)
    kernels = [k for k in args.kernels.split(",") if k]

    from gapbs_amd import gpu
    from gapbs_amd.graph import Graph

    t0 = time.time()
    g = Graph.kronecker(args.scale, 16, seed=args.seed)
    log(f"{g.stats_line()} built in {time.time() - t0:.1f}s")
    roots = [int(r) for r in g.pick_sources(args.roots)]
    dg = gpu.DeviceGraph(g)
    wg = dwg = rg_tc = drg = None
    if "sssp" in kernels:
        t0 = time.time()
        wg = Graph.kronecker(args.scale, 16, seed=args.seed, weighted=True)
        dwg = gpu.DeviceGraph(wg)
        log(f"weighted graph built and uploaded in {time.time() - t0:.1f}s")
    if "tc" in kernels:
        t0 = time.time()
        rg_tc = g.relabel_by_degree() if g.worth_relabelling() else g
        drg = gpu.DeviceGraph(rg_tc)
        log(f"relabelled graph built and uploaded in {time.time() - t0:.1f}s")
    log(f"device bytes {sum(d.device_bytes() for d in (dg, dwg, drg) if d) / 1e9:.2f} GB")

    m_per_root = [gpu.bfs(dg, r, want_depth=False, want_parent=False).stats["reached_edges"] // 2
                  for r in roots]
    m_total = sum(m_per_root)
    bc_roots = roots[:4]

    def run_kernel(k):
        if k == "bfs":
            for r in roots:
                gpu.bfs(dg, r, want_depth=False, want_parent=False)
        elif k == "sssp":
            for r in roots:
                gpu.sssp(dwg, r, args.delta, want_dist=False)
        elif k == "pr":
            return gpu.pagerank(dg, 20, 1e-4, want_scores=False).stats
        elif k == "cc":
            gpu.cc(dg, want_comp=False)
        elif k == "bc":
            return gpu.bc(dg, bc_roots, want_scores=False).stats
        elif k == "tc":
            return gpu.tc(drg)[1]
        return None

    for _ in range(args.warmup):
        for k in kernels:
            run_kernel(k)
    results = {}
    total_wall = 0.0
    for k in kernels:
        devs = [d for d in (dg, dwg, drg) if d is not None]
        with Timed(*devs) as t:
            last = None
            for _ in range(args.steps):
                last = run_kernel(k)
        per = t.seconds / args.steps
        total_wall += t.seconds
        r = {"seconds": round(per, 6), "roofline": roofline(t.prof),
             "kernels": {name: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2)}
                         for name, v in sorted(t.prof.items(), key=lambda kv: -kv[1]["ms"])[:6]}}
        if k in ("bfs", "sssp"):
            r["gteps"] = round(m_total / per / 1e9, 3)
        if k == "bc":
            bc_m = sum(m_per_root[:4])
            r["gteps"] = round(bc_m / per / 1e9, 3)
        if k == "pr" and last:
            r["iterations"] = last["iterations"]
        if k == "tc" and last is not None:
            r["merge_steps"] = last["edges_examined"]
        results[k] = r
        log(f"{k}: {r}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(g, wg, rg_tc, roots, m_per_root, args.cpu_budget)
        for k, v in cpu.items():
            if k in results:
                results[k]["cpu"] = v
                log(f"cpu {k}: {v}")
    head = results.get("bfs", next(iter(results.values())))
    all_prof_roof = max((v["roofline"] for v in results.values() if v["roofline"]),
                        key=lambda rl: rl["avg_launch_us"] * rl["launches"], default=None)
    out = {
        "metric": METRIC,
        "value": head.get("gteps"),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_wall * 1e3 / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"kronecker s{args.scale} ef16 seed {args.seed}: bfs and sssp from "
                               f"{len(roots)} SourcePicker roots, pr, cc, bc from 4 roots, tc",
                   "num_nodes": g.num_nodes, "num_edges_directed": g.num_edges_directed,
                   "sssp_delta": args.delta},
        "roofline": results.get("bfs", {}).get("roofline") or all_prof_roof,
        "kernels": results,
        "cpu_baseline": ({**cpu["bfs"], "sample": "BFS from the first " + cpu["bfs"]["sample"] +
                          " on the same graph"} if cpu and "bfs" in cpu else None),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
