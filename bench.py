"""Benchmark of the GPU kernel path on one MI355X.

Workload (default): a Kronecker graph, scale 26, edge factor 16 (Graph500
R-MAT A=.57 B=.19 C=.19, IDs permuted, symmetrised, generator seed 42; 67.1M
vertices, 2.1B directed edges, int64 offsets), built on the host and
uploaded to HBM before any timing.  By default the device graphs are then
renumbered by descending degree in HBM (gg_graph_order_by_degree: hubs
together, so the per-vertex arrays the kernels gather through the lists
stay cache-resident); the renumbering's device time is reported under
config.order_ms and not counted, every call still takes and returns the
generator's ids, and --layout upload keeps the generator's order.

Headline step: one batched BFS of 64 SourcePicker roots (64-bit per-vertex
visit masks, gg_bfs_batch).  value = aggregate GTEPS = sum over roots of m_r
(undirected edges of the root's component, the reference's TEPS convention)
over the wall time of the step.  The same 64 roots run as 64 sequential
single-root direction-optimizing BFS calls ("bfs"), and the other kernels
run on the same graph, each timed in its own loop with results left on the
device:
  sssp  delta-stepping from the first --sssp-roots roots (weights in [1, 255])
  pr    PageRank, damping 0.85, tolerance 1e-4, at most 20 iterations (FP32,
        compensated per-vertex sums)
  pr_fp64  the same in FP64; the FP32 result's max-abs and L1 error against it
        and the FP64/FP32 time ratio are reported (the FP64 path is held to
        the FP64 oracle in the tests)
  cc    Afforest
  bc    Brandes from the first --bc-roots roots (the reference's float arithmetic);
        the FP32 and compensated (Kahan) scores' max-abs and L1 error against
        the FP64 scores of the same roots are reported
  bc_batch  Brandes from all 64 roots in one batched traversal (same arithmetic)
  tc    Hybrid: relabel by degree on the device (when the reference's
        heuristic says so), then ordered triangle count

Before timing, gg_autotune picks this graph's BFS alpha/beta, batched
alpha, SSSP delta and PageRank bin cut-offs from the same roots (its time is
reported under config.autotune and not counted; --no-autotune keeps the
reference's alpha 15 / beta 18 and --delta).

The roofline figure is measured live: the library times each kernel launch
with HIP events carried by the dispatch packet and credits it with its
algorithmic bytes.  HBM traffic per launch (rocprofv3 FETCH_SIZE and
WRITE_SIZE, scaled by the same counters on a copy of known size) is read
from profiles/traffic.json when it holds an entry for this workload and
kernel.  The CPU baseline is the reference compiled from source
(oracle/_ref) when present, else this repo's OpenMP restatement, timed on
the host cores on bounded samples (first roots only; TC over a 1/stride
sample of its outer loop, scaled).

Prints one JSON line on stdout; progress goes to stderr.
"""

import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BFS/SSSP/BC GTEPS, PR/CC/TC seconds, Kronecker s26 ef16; % HBM roofline"
HBM_PEAK_GBS = 8000.0
KERNELS = ("bfs_batch", "bfs", "bc_batch", "sssp", "pr", "pr_fp64", "cc", "cc_sv", "bc", "tc", "build", "generate",
           "load", "packed")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


EXTRA_PROFILES = []  # per-kernel reports of device graphs freed inside a timed region
PROFILING = [False]  # inside a Timed region


def build_report(g, edges, last, per, host_build_s, no_cpu):
    """The device CSR build's line: device time (edge list already in HBM)
    against the compulsory bytes -- the edge list read once (8 B per edge),
    the CSR written once (8 B per offset, 4 B per entry) -- and a check that
    one more build equals the host builder's graph byte for byte."""
    import numpy as np
    from gapbs_amd import gpu
    m = len(edges[0])
    dev_s = last["device_ms"] / 1e3
    alg = 8.0 * m + 8.0 * (g.num_nodes + 1) + 4.0 * int(g.out_offsets[-1])
    db, _ = gpu.build_graph(*edges)
    off, dst = db.copy_out()
    db.free()
    same = bool(np.array_equal(off, g.out_offsets) and np.array_equal(dst, g.out_neighbors))
    assert same, "device build differs from the host builder"
    out = {"device_seconds": round(dev_s, 6), "edges": m, "matches_host_build": same,
           "roofline": {"bound": "hbm", "kernel": "gg_graph_build (device time)",
                        "achieved": round(alg / dev_s / 1e9, 2), "peak": 8000.0, "unit": "GB/s",
                        "frac": round(alg / dev_s / 8e12, 5), "traffic": None, "alg_bytes_per_launch": alg},
           "note": "seconds include copying the edge list from host memory; device_seconds start in HBM"}
    if not no_cpu and host_build_s is not None:
        out["cpu"] = {"value": round(host_build_s, 3), "unit": "s", "kind": "port",
                      "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                      "sample": "the whole edge list through the host builder (OpenMP restatement of "
                                "BuilderBase::MakeGraphFromEL + SquishCSR, src/builder.h:300-341)"}
    return out


def packed_report(packed, last, per, steps, warmup, run_kernel):
    """The packed line: traversed edges per second over all 2048 roots, and
    the same roots graph by graph (one batched call per graph)."""
    m = last["reached_edges"] // 2
    for _ in range(warmup):
        run_kernel("packed_per_graph")
    t = time.perf_counter()
    for _ in range(steps):
        run_kernel("packed_per_graph")
    per_graph = (time.perf_counter() - t) / steps
    return {"gteps": round(m / per / 1e9, 3), "graphs": len(packed[3]), "roots": int(packed[2].size),
            "per_graph_seconds": round(per_graph, 6), "speedup_vs_per_graph": round(per_graph / per, 3)}


def load_report(g, path, per, no_cpu):
    """The .sg load line: file bytes per second into HBM (page-cached file,
    read through pinned staging), a byte-for-byte check against the host
    graph, and the host path (gh_read, then gg_graph_upload) as the CPU
    baseline."""
    import numpy as np
    from gapbs_amd import gpu
    from gapbs_amd.graph import Graph
    size = os.path.getsize(path)
    db, _ = gpu.read_graph(path)
    off, dst = db.copy_out()
    db.free()
    same = bool(np.array_equal(off, g.out_offsets) and np.array_equal(dst, g.out_neighbors))
    assert same, "device load differs from the host graph"
    out = {"file_bytes": size, "gb_per_s": round(size / per / 1e9, 2), "matches_host_graph": same}
    if not no_cpu:
        t0 = time.time()
        hg = Graph.read(path)
        t1 = time.time()
        gpu.DeviceGraph(hg).free()
        t2 = time.time()
        out["cpu"] = {"value": round(t2 - t0, 3), "unit": "s", "kind": "port", "read_s": round(t1 - t0, 3),
                      "upload_s": round(t2 - t1, 3), "cores": 1,
                      "sample": "host reader (gh_read, Reader::ReadSerializedGraph restated) then gg_graph_upload"}
    return out


def generate_report(g, args, last, host_gen_s, host_build_s):
    """The device generation's line: generation + build device time, and a
    check that one more run equals the host generator's graph byte for byte.
    Seconds include the host's vertex permutation (a sequential shuffle)."""
    import numpy as np
    from gapbs_amd import gpu
    db, _ = gpu.generate_graph(0, args.scale, 16, args.seed)
    off, dst = db.copy_out()
    db.free()
    same = bool(np.array_equal(off, g.out_offsets) and np.array_equal(dst, g.out_neighbors))
    assert same, "device generation differs from the host generator"
    out = {"device_seconds": round(last["device_ms"] / 1e3, 6), "matches_host_graph": same,
           "note": "seconds include the host permutation and its copy; device_seconds = generation + build"}
    if not args.no_cpu and host_gen_s is not None:
        out["cpu"] = {"value": round(host_gen_s + host_build_s, 3), "unit": "s", "kind": "port",
                      "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
                      "sample": "host generator (R-MAT blocks, PermuteIDs, OpenMP) + host builder, whole graph"}
    return out


class Timed:
    """Wall time of the enclosed calls, with the library's per-kernel profile."""

    def __init__(self, *dgs):
        self.dgs = dgs

    def __enter__(self):
        for dg in self.dgs:
            dg.profile(True)
            dg.profile_reset()
        PROFILING[0] = True
        self.t = time.perf_counter()
        return self

    def __exit__(self, *exc):
        self.seconds = time.perf_counter() - self.t
        PROFILING[0] = False
        self.prof = {}
        reports = [dg.profile_report() for dg in self.dgs] + EXTRA_PROFILES
        EXTRA_PROFILES.clear()
        for rep in reports:
            for k, v in rep.items():
                p = self.prof.setdefault(k, {"launches": 0, "ms": 0.0, "alg_bytes": 0.0})
                for f in p:
                    p[f] += v[f]
        for dg in self.dgs:
            dg.profile(False)


def load_traffic(workload_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get(workload_key, {})


def roofline(prof, traffic):
    """Roofline of the kernel with the most device time among those credited with bytes."""
    cands = {k: v for k, v in prof.items() if v["alg_bytes"] > 0 and v["ms"] > 0}
    if not cands:
        return None
    name, v = max(cands.items(), key=lambda kv: kv[1]["ms"])
    achieved = v["alg_bytes"] / (v["ms"] / 1e3) / 1e9
    t = traffic.get(name)
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": t["bytes_per_launch"] if t else None,
            "alg_bytes_per_launch": round(v["alg_bytes"] / v["launches"]),
            "avg_launch_us": round(1e3 * v["ms"] / v["launches"], 2), "launches": v["launches"]}


def cpu_baselines(kernels, g, wg, relabel_tc, roots, m_per_root, sssp_roots, bc_roots, budget_s, tc_stride):
    """Reference (or port) CPU timings on bounded samples of the same workload."""
    from oracle import oracle, reference as ref
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    out = {}
    use_ref = ref.available("bfs")
    kind = "reference" if use_ref else "port"

    def sample_roots(run, rts):
        secs, done, edges = 0.0, 0, 0
        for r, m in zip(rts, m_per_root):
            secs += run(r)
            edges += m
            done += 1
            if secs > budget_s:
                break
        return secs, done, edges

    if "bfs" in kernels or "bfs_batch" in kernels:
        rg = ref.RefGraph("bfs", g) if use_ref else None
        secs, done, edges = sample_roots(
            lambda r: ref.bfs_time(rg, [r]) if use_ref else oracle.dobfs(g, r)[1], roots)
        del rg
        out["bfs"] = {"value": edges / secs / 1e9, "unit": "GTEPS", "seconds": secs,
                      "sample": f"DOBFS from the first {done} of the {len(roots)} roots, one at a time"}
        log(f"cpu bfs: {out['bfs']}")
    if "sssp" in kernels and wg is not None:
        rgw = ref.RefGraph("sssp", wg) if use_ref else None
        secs, done, edges = sample_roots(
            lambda r: ref.sssp_time(rgw, [r], 2) if use_ref else oracle.delta_step(wg, r, 2)[1],
            roots[:sssp_roots])
        del rgw
        out["sssp"] = {"value": edges / secs / 1e9, "unit": "GTEPS", "seconds": secs,
                       "sample": f"DeltaStep (delta 2, the reference's setting) from the first {done} roots"}
        log(f"cpu sssp: {out['sssp']}")
    if "pr" in kernels:
        if use_ref:
            _, s = ref.pagerank("pr_spmv", g, 20, 1e-4)
        else:
            t = time.perf_counter()
            oracle.pagerank(g, 20, 1e-4)
            s = time.perf_counter() - t
        out["pr"] = {"value": s, "unit": "s", "sample": "one PageRankPull run (Jacobi), tol 1e-4, <= 20 iters"}
        log(f"cpu pr: {out['pr']}")
    if "cc" in kernels:
        s = ref.afforest(g)[1] if use_ref else oracle.afforest(g)[1]
        out["cc"] = {"value": s, "unit": "s", "sample": "one Afforest run"}
        log(f"cpu cc: {out['cc']}")
    if "bc" in kernels and use_ref:
        _, s = ref.brandes(g, 1)
        bc_m = m_per_root[0]
        out["bc"] = {"value": bc_m / s / 1e9, "unit": "GTEPS", "seconds": s,
                     "sample": "Brandes from the first root only (its SourcePicker)"}
        log(f"cpu bc: {out['bc']}")
    if "tc" in kernels:
        t = time.perf_counter()
        rg = g.relabel_by_degree() if relabel_tc else g  # the host builder's parallel relabel
        relabel_s = time.perf_counter() - t if relabel_tc else 0.0
        _, s = oracle.ordered_count_sample(rg, tc_stride, 0)
        del rg
        out["tc"] = {"value": relabel_s + s * tc_stride, "unit": "s", "kind": "port", "relabel_seconds": relabel_s,
                     "sampled_seconds": s,
                     "sample": f"RelabelByDegree (host builder, OpenMP) + OrderedCount (OpenMP restatement) over "
                               f"u = 0 mod {tc_stride} of the relabelled graph, count time x {tc_stride}"}
        log(f"cpu tc: {out['tc']}")
    for v in out.values():
        v.setdefault("kind", kind)
        v["cores"] = cores
    return out


def copy_probe(nbytes):
    """One device-to-device copy of nbytes (rocprofv3 counter calibration)."""
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert hip.hipDeviceSynchronize() == 0
    kind_d2d = 3
    assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), kind_d2d) == 0
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scale", type=int, default=26)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--roots", type=int, default=64, help="batched / sequential BFS roots (1..64)")
    ap.add_argument("--sssp-roots", type=int, default=4)
    ap.add_argument("--bc-roots", type=int, default=4)
    ap.add_argument("--delta", type=int, default=32, help="SSSP bucket width on the GPU")
    ap.add_argument("--alpha", type=int, default=15, help="direction switch for both BFS paths")
    ap.add_argument("--tc-steps", type=int, default=1, help="timed TC runs (one is several seconds at s26)")
    ap.add_argument("--kernels", default=",".join(KERNELS))
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds per sampled CPU baseline")
    ap.add_argument("--tc-stride", type=int, default=64, help="CPU TC sample: one vertex in stride")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use the reference's alpha/beta (15/18), delta --delta and default PageRank bins")
    ap.add_argument("--copy-probe", type=int, default=0, help="only run one device copy of this many bytes")
    ap.add_argument("--layout", choices=("degree", "upload"), default="degree",
                    help="device vertex order: descending degree (default) or as generated")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one GPU, and the kernels "
              "run on a single device (multi-GPU sharding is a design note only).", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    kernels = [k for k in args.kernels.split(",") if k]
    unknown = set(kernels) - set(KERNELS)
    if unknown:
        sys.exit(f"unknown kernels {sorted(unknown)}")

    from gapbs_amd import gpu
    from gapbs_amd.graph import KRONECKER, Graph

    workload = f"kronecker s{args.scale} ef16 seed {args.seed}"
    if args.layout == "degree":
        workload += ", degree-ordered layout"
    ordered = args.layout == "degree"
    order_ms = {}
    t0 = time.time()
    build_edges, host_build_s, host_gen_s = None, None, None
    if "build" in kernels or "generate" in kernels:
        # The generator's edge list, kept for the device build; the host
        # builds the same graph from it (timed: the CPU baselines of the
        # device build and, with the generation, of the device generation).
        build_edges = Graph.generate_edges(KRONECKER, args.scale, 16, args.seed)[:2]
        host_gen_s = time.time() - t0
        t1 = time.time()
        g = Graph.from_edges(*build_edges)
        host_build_s = time.time() - t1
        if "build" not in kernels:
            build_edges = None
    else:
        g = Graph.kronecker(args.scale, 16, seed=args.seed)
    log(f"{g.stats_line()} built in {time.time() - t0:.1f}s")
    roots = [int(r) for r in g.pick_sources(args.roots)]
    dg = gpu.DeviceGraph(g, order_by_degree=ordered)
    if ordered:
        order_ms["graph"] = round(dg.order_stats["device_ms"], 3)
    wg = dwg = None
    relabel_tc = "tc" in kernels and g.worth_relabelling()
    if "sssp" in kernels:
        t0 = time.time()
        wg = Graph.kronecker(args.scale, 16, seed=args.seed, weighted=True)
        dwg = gpu.DeviceGraph(wg, order_by_degree=ordered)
        if ordered:
            order_ms["weighted_graph"] = round(dwg.order_stats["device_ms"], 3)
        log(f"weighted graph built and uploaded in {time.time() - t0:.1f}s")
    if ordered:
        log(f"degree ordering (device ms): {order_ms}")
    log(f"device bytes {sum(d.device_bytes() for d in (dg, dwg) if d) / 1e9:.2f} GB")

    # m_r per root, from the device's own reach statistics.
    m_per_root = [gpu.bfs(dg, r, want_depth=False, want_parent=False).stats["reached_edges"] // 2
                  for r in roots]
    m_total = sum(m_per_root)
    sssp_roots = roots[:args.sssp_roots]
    bc_roots = roots[:args.bc_roots]

    def run_kernel(k):
        if k == "bfs_batch":
            return gpu.bfs_batch(dg, roots, batch_alpha, want_depths=False).stats
        if k == "bfs":
            for r in roots:
                gpu.bfs(dg, r, alpha, beta, want_depth=False, want_parent=False)
        elif k == "sssp":
            for r in sssp_roots:
                gpu.sssp(dwg, r, delta, want_dist=False)
        elif k == "pr":
            return gpu.pagerank(dg, 20, 1e-4, want_scores=False).stats
        elif k == "pr_fp64":
            return gpu.pagerank(dg, 20, 1e-4, fp64=True, want_scores=False).stats
        elif k == "cc":
            gpu.cc(dg, want_comp=False)
        elif k == "cc_sv":
            return gpu.cc_sv(dg, want_comp=False).stats
        elif k == "bc":
            return gpu.bc(dg, bc_roots, want_scores=False).stats
        elif k == "bc_batch":
            return gpu.bc_batch(dg, roots, want_scores=False).stats
        elif k == "tc":
            # Hybrid (src/tc.cc:95-100): relabel by degree when worth it --
            # on the device, inside the step as in the reference's trial.
            if not relabel_tc:
                return gpu.tc(dg)[1]
            dr, rst = gpu.relabel_by_degree(dg)
            profiling = PROFILING[0]
            if profiling:
                dr.profile(True)
            c, st = gpu.tc(dr)
            if profiling:
                EXTRA_PROFILES.append(dr.profile_report())
            dr.free()
            return {**st, "relabel_ms": rst["device_ms"], "triangles": c}
        elif k == "build":
            db, st = gpu.build_graph(*build_edges)
            db.free()
            return st
        elif k == "generate":
            db, st = gpu.generate_graph(0, args.scale, 16, args.seed)
            db.free()
            return st
        elif k == "load":
            db, st = gpu.read_graph(sg_path[0])
            db.free()
            return st
        elif k == "packed":
            return gpu.bfs_packed(packed[0], packed[1], packed[2], batch_alpha, want_depths=False).stats
        elif k == "packed_per_graph":
            for i, (dgi, r) in enumerate(zip(packed[3], packed[2])):
                gpu.bfs_batch(dgi, r - packed[1][i], batch_alpha, want_depths=False)
        return None

    # Per-graph tuning (gg_autotune) before any timing; its own time is
    # reported, not included.  Results do not depend on the settings.
    alpha, beta, batch_alpha, delta = args.alpha, 18, args.alpha, args.delta
    tuned = None
    if not args.no_autotune:
        t0 = time.time()
        what = 0
        if "bfs" in kernels:
            what |= gpu.TUNE_BFS
        if "bfs_batch" in kernels:
            what |= gpu.TUNE_BATCH
        if "pr" in kernels or "pr_fp64" in kernels:
            what |= gpu.TUNE_PR
        if what:
            tuned = gpu.autotune(dg, what, roots)
        if "sssp" in kernels:
            tw = gpu.autotune(dwg, gpu.TUNE_SSSP, sssp_roots)
            tuned = {**(tuned or {}), "sssp_delta": tw["sssp_delta"], "sssp_best_ms": tw["best_ms"][2]}
        if tuned is not None:
            tuned["seconds"] = round(time.time() - t0, 2)
            alpha, beta = tuned.get("bfs_alpha", alpha), tuned.get("bfs_beta", beta)
            batch_alpha, delta = tuned.get("batch_alpha", batch_alpha), tuned.get("sssp_delta", delta)
            log(f"autotune: {tuned}")

    # The graph as a .sg file for the load line (written once, untimed; the
    # reads come from the page cache).
    sg_path = [None]
    if "load" in kernels:
        import tempfile
        sg_dir = tempfile.mkdtemp(dir=os.environ.get("TMPDIR", "/tmp"))
        sg_path[0] = os.path.join(sg_dir, "graph.sg")
        t0 = time.time()
        g.write(sg_path[0])
        log(f"wrote {os.path.getsize(sg_path[0]) / 1e9:.2f} GB .sg in {time.time() - t0:.1f}s")

    # The packed workload: 256 Kronecker graphs of scale 10..16 (edge factor
    # 16, generator seed = graph index), 8 roots each, in one traversal; the
    # same roots graph by graph (one batched call per graph) for comparison.
    packed = None
    if "packed" in kernels:
        import numpy as np
        t0 = time.time()
        scales = np.random.default_rng(0).integers(10, 17, 256)
        pgraphs = [Graph.kronecker(int(sc), 16, seed=i) for i, sc in enumerate(scales)]
        union, bounds = Graph.pack(pgraphs)
        proots = np.stack([pg.pick_sources(8) + bounds[i] for i, pg in enumerate(pgraphs)])
        packed = (gpu.DeviceGraph(union), bounds, proots, [gpu.DeviceGraph(pg) for pg in pgraphs])
        log(f"packed workload: {union.num_nodes} vertices, {union.num_edges()} edges, built in "
            f"{time.time() - t0:.1f}s")

    traffic = load_traffic(workload)
    results = {}
    devs = [d for d in (dg, dwg) if d is not None]
    for k in kernels:
        steps = args.tc_steps if k == "tc" else args.steps
        for _ in range(args.warmup):
            run_kernel(k)
        with Timed(*devs) as t:
            last = None
            for _ in range(steps):
                last = run_kernel(k)
        per = t.seconds / steps
        r = {"seconds": round(per, 6), "steps": steps, "roofline": roofline(t.prof, traffic),
             "kernels": {name: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2)}
                         for name, v in sorted(t.prof.items(), key=lambda kv: -kv[1]["ms"])[:6]}}
        if k in ("bfs_batch", "bfs"):
            r["gteps"] = round(m_total / per / 1e9, 3)
        if k == "bfs_batch" and last:
            r["td_steps"], r["bu_steps"] = last["td_steps"], last["bu_steps"]
            assert last["reached_edges"] // 2 == m_total, "batched reach differs from the single-root runs"
        if k == "sssp":
            r["gteps"] = round(sum(m_per_root[:len(sssp_roots)]) / per / 1e9, 3)
        if k == "bc":
            r["gteps"] = round(sum(m_per_root[:len(bc_roots)]) / per / 1e9, 3)
        if k == "bc_batch":
            r["gteps"] = round(m_total / per / 1e9, 3)
        if k in ("pr", "pr_fp64", "cc_sv") and last:
            r["iterations"] = last["iterations"]
        if k == "tc" and last is not None:
            r["probes"] = last["edges_examined"]
            if "relabel_ms" in last:
                r["relabel_ms_last"] = round(last["relabel_ms"], 3)
        if k == "build" and last is not None:
            r.update(build_report(g, build_edges, last, per, host_build_s, args.no_cpu))
            build_edges = None
        if k == "generate" and last is not None:
            r.update(generate_report(g, args, last, host_gen_s, host_build_s))
        if k == "packed" and last is not None:
            r.update(packed_report(packed, last, per, args.steps, args.warmup, run_kernel))
        if k == "load":
            r.update(load_report(g, sg_path[0], per, args.no_cpu))
            os.remove(sg_path[0])
            os.rmdir(os.path.dirname(sg_path[0]))
        results[k] = r
        log(f"{k}: {r}")
    if "bfs_batch" in results and "bfs" in results:
        results["bfs_batch"]["speedup_vs_sequential"] = round(results["bfs"]["seconds"] /
                                                              results["bfs_batch"]["seconds"], 3)
    if "pr" in results and "pr_fp64" in results:
        import numpy as np
        s32 = gpu.pagerank(dg, 20, 1e-4).scores.astype(np.float64)
        s64 = gpu.pagerank(dg, 20, 1e-4, fp64=True).scores.astype(np.float64)
        results["pr"]["vs_fp64"] = {"max_abs": float(np.abs(s32 - s64).max()),
                                    "l1": float(np.abs(s32 - s64).sum()),
                                    "fp64_over_fp32_time": round(results["pr_fp64"]["seconds"] /
                                                                 results["pr"]["seconds"], 3)}
        log(f"pr precision: {results['pr']['vs_fp64']}")
    if "bc" in results:
        import numpy as np
        s64 = gpu.bc(dg, bc_roots, fp64=True).scores.astype(np.float64)
        err = {}
        for name, kw in (("fp32", {}), ("kahan", {"kahan": True})):
            d = np.abs(gpu.bc(dg, bc_roots, **kw).scores.astype(np.float64) - s64)
            err[name] = {"max_abs": float(d.max()), "l1": float(d.sum())}
        results["bc"]["vs_fp64"] = err
        log(f"bc precision: {err}")
    if "bc_batch" in results and "bc" in results:  # per-root GTEPS: bc runs fewer roots
        results["bc_batch"]["speedup_vs_sequential"] = round(results["bc_batch"]["gteps"] /
                                                             results["bc"]["gteps"], 3)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(kernels, g, wg, relabel_tc, roots, m_per_root, len(sssp_roots), len(bc_roots),
                            args.cpu_budget, args.tc_stride)
        for k, v in cpu.items():
            if k in results:
                results[k]["cpu"] = v
    head_name = "bfs_batch" if "bfs_batch" in results else next(iter(results))
    head = results[head_name]
    out = {
        "metric": METRIC,
        "value": head.get("gteps"),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["seconds"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{workload}: batched BFS of {len(roots)} SourcePicker roots per step "
                               f"(headline); per-kernel lines for 64 sequential BFS, sssp "
                               f"({len(sssp_roots)} roots), pr, cc, bc ({len(bc_roots)} roots), batched bc "
                               f"({len(roots)} roots), tc",
                   "num_nodes": g.num_nodes, "num_edges_directed": g.num_edges_directed,
                   "roots": len(roots), "bfs_alpha_beta": [alpha, beta], "batch_alpha": batch_alpha,
                   "sssp_delta": delta, "autotune": tuned, "layout": args.layout,
                   "order_ms": order_ms or None},
        "roofline": head["roofline"],
        "kernels": results,
        "cpu_baseline": ({**cpu["bfs"], "sample": cpu["bfs"]["sample"] + " (the reference has no batched "
                          "BFS: its trial loop runs one root at a time)"} if cpu and "bfs" in cpu else None),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
