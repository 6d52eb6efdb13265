"""Benchmark of the GPU kernel path on one MI355X.

Workloads (--workload; the default is the headline):
  kron26     Kronecker scale 26, edge factor 16 (Graph500 R-MAT A=.57 B=.19
             C=.19, IDs permuted, symmetrised, generator seed 42; 67.1M
             vertices, 2.1B directed edges, int64 offsets).  Headline step:
             one batched BFS of 64 SourcePicker roots (gg_bfs_batch).  Other
             lines: the same 64 roots one BFS call at a time, batched and
             single-root SSSP, PageRank FP32 / FP64, Afforest and
             Shiloach-Vishkin CC, BC from 64 roots one root at a time (the
             reference's Brandes loop) and batched, TC (Hybrid: device
             relabel + count), and the graph build, generation, .sg load and
             the packed traversal of 256 small Kronecker graphs.
  kron22     Kronecker scale 22, seed 8675309: the six kernels beside the
             reference's own binaries on the same graph and roots.
  uniform25  uniform scale 25, degree 16, seed 1234: the precision pattern
             (PageRank FP32 against FP64, BC from 16 roots with FP32 scores
             and FP64 path counts against FP64) and SSSP.
  grid4096   the 4096 x 4096 grid (5% of the lattice edges deleted, 1%
             random shortcuts, weights in [1, 1000], seed 7): BFS and SSSP
             from 32 roots, one at a time against batched, and a sweep of the
             bucket width delta.
The graphs are built on the host and uploaded to HBM before any timing; by
default the device copies are renumbered by descending degree in HBM
(gg_graph_order_by_degree; the renumbering's device time is reported under
config.order_ms and not counted; every call still takes and returns the
generator's ids); --layout upload keeps the generator's order.

Each line times its step with the library as a caller uses it: results left
in HBM, batched traversals replaying their captured hipGraph level loops.
That is "seconds" (and GTEPS = sum over roots of m_r, the undirected edges of
the root's component, over it).  "e2e_seconds" times the same step with the
results copied back to host memory through the C-ABI, the reference's trial
boundary (src/benchmark.h:107-109).  A separate profiled pass of the same
step (per-kernel HIP events carried by the dispatch packets; levels driven
from the host, since events cannot be captured) gives the per-kernel table
and the roofline: achieved = algorithmic bytes credited to the kernel with
the most device time / its average launch time, against 8 TB/s and against
the device-to-device copy bandwidth measured at start.  HBM traffic per
launch (rocprofv3 FETCH_SIZE and WRITE_SIZE, scaled by the same counters on
a copy of known size) comes from profiles/traffic.json when it holds an
entry for this workload and kernel.

CPU baselines: the reference compiled from source (oracle/_ref) when
present, else this repo's OpenMP restatement, on the host cores, on bounded
samples (first roots only); TC at scale 26 runs this repo's restatement of
OrderedCount over a 1/stride sample of its outer loop, scaled (the
reference's loop has no sample hook), and the reference's own Hybrid in full
at scale 22.

Before timing, gg_autotune picks the graph's BFS alpha/beta, batched alpha,
SSSP delta and PageRank bin cut-offs from the same roots (reported under
config.autotune, not counted).  Prints one JSON line on stdout; progress goes
to stderr.
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
WORKLOADS = {
    "kron26": ("bfs_batch", "bfs", "sssp_batch", "sssp", "pr", "pr_fp64", "cc", "cc_sv", "bc", "bc_batch", "tc",
               "build", "generate", "load", "packed"),
    "kron22": ("bfs", "bfs_batch", "sssp", "sssp_batch", "pr", "pr_fp64", "cc", "bc", "bc_batch", "tc"),
    "uniform25": ("pr", "pr_fp64", "bc", "bc_batch", "sssp", "sssp_batch"),
    "grid4096": ("bfs", "bfs_batch", "sssp", "sssp_batch", "delta_sweep"),
}
# Longest timed steps: at most this many steps (and warmups) per line, so
# that a run with many steps still finishes in a few minutes.
STEP_CAP = {"bc": 1, "tc": 1, "sssp": 4, "build": 3, "generate": 3, "load": 3, "delta_sweep": 1}
WARM_CAP = {"bc": 0, "tc": 1, "build": 1, "generate": 1, "load": 1, "delta_sweep": 0}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


EXTRA_PROFILES = []  # per-kernel reports of device graphs freed inside a profiled pass
PROFILING = [False]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def hip():
    return ctypes.CDLL("libamdhip64.so")


def copy_bandwidth(nbytes=4 << 30, reps=5):
    """Device-to-device copy rate of the HBM (bytes read + bytes written per
    second), the practical ceiling the kernels' fractions are also quoted
    against."""
    h = hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert h.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert h.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    kind_d2d = 3
    assert h.hipMemcpy(b, a, ctypes.c_size_t(nbytes), kind_d2d) == 0
    assert h.hipDeviceSynchronize() == 0
    t = time.perf_counter()
    for _ in range(reps):
        assert h.hipMemcpy(b, a, ctypes.c_size_t(nbytes), kind_d2d) == 0
    assert h.hipDeviceSynchronize() == 0
    secs = (time.perf_counter() - t) / reps
    h.hipFree(a)
    h.hipFree(b)
    return 2.0 * nbytes / secs / 1e9


def copy_probe(nbytes):
    """One device-to-device copy of nbytes (rocprofv3 counter calibration)."""
    h = hip()
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert h.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert h.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    assert h.hipDeviceSynchronize() == 0
    assert h.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0
    assert h.hipDeviceSynchronize() == 0
    h.hipFree(a)
    h.hipFree(b)
    print(json.dumps({"copy_probe_bytes": nbytes}), flush=True)


def merge_reports(reports):
    prof = {}
    for rep in reports:
        for k, v in rep.items():
            p = prof.setdefault(k, {"launches": 0, "ms": 0.0, "alg_bytes": 0.0})
            for f in p:
                p[f] += v[f]
    return prof


def profiled(devs, fn, steps):
    """Runs fn `steps` times with per-kernel profiling on: ({kernel: totals}, seconds per step)."""
    for dg in devs:
        dg.profile(True)
        dg.profile_reset()
    PROFILING[0] = True
    t = time.perf_counter()
    try:
        for _ in range(steps):
            fn()
    finally:
        secs = (time.perf_counter() - t) / steps
        PROFILING[0] = False
        reports = [dg.profile_report() for dg in devs] + EXTRA_PROFILES
        EXTRA_PROFILES.clear()
        for dg in devs:
            dg.profile(False)
    return merge_reports(reports), secs


def load_traffic(workload_key):
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get(workload_key, {})


def roofline(prof, traffic, copy_gbs):
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
            "avg_launch_us": round(1e3 * v["ms"] / v["launches"], 2), "launches": v["launches"],
            "copy_gbs": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 5)}


def kernel_table(prof, top=8):
    return {name: {"launches": v["launches"], "avg_us": round(1e3 * v["ms"] / v["launches"], 2),
                   "alg_gbs": round(v["alg_bytes"] / (v["ms"] / 1e3) / 1e9, 2) if v["ms"] > 0 else None}
            for name, v in sorted(prof.items(), key=lambda kv: -kv[1]["ms"])[:top]}


# ---------------------------------------------------------------------------
# Lines that need more than one timed call.

def build_report(g, edges, last, host_build_s, no_cpu):
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
           "compulsory": {"kernel": "gg_graph_build (device time)", "achieved": round(alg / dev_s / 1e9, 2),
                          "frac": round(alg / dev_s / 8e12, 5), "bytes": alg},
           "note": "seconds include copying the edge list from host memory; device_seconds start in HBM"}
    if not no_cpu and host_build_s is not None:
        out["cpu"] = {"value": round(host_build_s, 3), "unit": "s", "kind": "port", "cores": cpu_threads(),
                      "sample": "the whole edge list through the host builder (OpenMP restatement of "
                                "BuilderBase::MakeGraphFromEL + SquishCSR, src/builder.h:300-341)"}
    return out


def generate_report(g, scale, seed, last, host_gen_s, host_build_s, no_cpu):
    """The device generation's line: generation + build device time, and a
    check that one more run equals the host generator's graph byte for byte.
    Seconds include the host's vertex permutation (a sequential shuffle)."""
    import numpy as np
    from gapbs_amd import gpu
    db, _ = gpu.generate_graph(0, scale, 16, seed)
    off, dst = db.copy_out()
    db.free()
    same = bool(np.array_equal(off, g.out_offsets) and np.array_equal(dst, g.out_neighbors))
    assert same, "device generation differs from the host generator"
    out = {"device_seconds": round(last["device_ms"] / 1e3, 6), "matches_host_graph": same,
           "note": "seconds include the host permutation and its copy; device_seconds = generation + build"}
    if not no_cpu and host_gen_s is not None:
        out["cpu"] = {"value": round(host_gen_s + host_build_s, 3), "unit": "s", "kind": "port",
                      "cores": cpu_threads(),
                      "sample": "host generator (R-MAT blocks, PermuteIDs, OpenMP) + host builder, whole graph"}
    return out


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


# ---------------------------------------------------------------------------
# CPU baselines.

def cpu_baselines(W, kernels, budget_s, tc_stride):
    """Reference (or port) CPU timings on bounded samples of the workload."""
    from oracle import oracle, reference as ref
    out = {}
    use_ref = ref.available("bfs")
    kind = "reference" if use_ref else "port"
    g, wg, roots, m_per_root = W["g"], W["wg"], W["roots"], W["m_per_root"]

    def sample_roots(run, rts):
        secs, done, edges = 0.0, 0, 0
        for r, m in zip(rts, m_per_root):
            secs += run(r)
            edges += m
            done += 1
            if secs > budget_s:
                break
        return secs, done, edges

    if {"bfs", "bfs_batch"} & set(kernels):
        rg = ref.RefGraph("bfs", g) if use_ref else None
        secs, done, edges = sample_roots(
            lambda r: ref.bfs_time(rg, [r]) if use_ref else oracle.dobfs(g, r)[1], roots)
        del rg
        out["bfs"] = {"value": edges / secs / 1e9, "unit": "GTEPS", "seconds": secs,
                      "sample": f"DOBFS from the first {done} of the {len(roots)} roots, one at a time"}
        log(f"cpu bfs: {out['bfs']}")
    if {"sssp", "sssp_batch"} & set(kernels) and wg is not None:
        delta = W["cpu_delta"]
        rgw = ref.RefGraph("sssp", wg) if use_ref else None
        secs, done, edges = sample_roots(
            lambda r: ref.sssp_time(rgw, [r], delta) if use_ref else oracle.delta_step(wg, r, delta)[1],
            roots)
        del rgw
        out["sssp"] = {"value": edges / secs / 1e9, "unit": "GTEPS", "seconds": secs,
                       "sample": f"DeltaStep (delta {delta}, {W['cpu_delta_why']}) from the first {done} roots"}
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
    if {"cc", "cc_sv"} & set(kernels):
        s = ref.afforest(g)[1] if use_ref else oracle.afforest(g)[1]
        out["cc"] = {"value": s, "unit": "s", "sample": "one Afforest run"}
        log(f"cpu cc: {out['cc']}")
    if {"bc", "bc_batch"} & set(kernels) and use_ref:
        k = W["bc_cpu_roots"]
        _, s = ref.brandes(g, k)
        bc_m = sum(m_per_root[:k])
        out["bc"] = {"value": bc_m / s / 1e9, "unit": "GTEPS", "seconds": s,
                     "sample": f"Brandes from the first {k} roots (its SourcePicker)"}
        log(f"cpu bc: {out['bc']}")
    if "tc" in kernels:
        if use_ref and W["tc_cpu"] == "reference":
            c, s = ref.tc(g, hybrid=True)
            out["tc"] = {"value": s, "unit": "s", "kind": "reference", "triangles": c,
                         "sample": "the reference's Hybrid (relabel when worth it, then OrderedCount), whole graph"}
        else:
            t = time.perf_counter()
            rg = g.relabel_by_degree() if W["relabel_tc"] else g  # the host builder's parallel relabel
            relabel_s = time.perf_counter() - t if W["relabel_tc"] else 0.0
            _, s = oracle.ordered_count_sample(rg, tc_stride, 0)
            del rg
            out["tc"] = {"value": relabel_s + s * tc_stride, "unit": "s", "kind": "port",
                         "relabel_seconds": relabel_s, "sampled_seconds": s,
                         "sample": f"RelabelByDegree (host builder, OpenMP) + OrderedCount (OpenMP restatement, "
                                   f"same loop as src/tc.cc:52-71) over u = 0 mod {tc_stride} of the relabelled "
                                   f"graph, count time x {tc_stride}; the reference's OrderedCount has no sample "
                                   f"hook (its whole run is timed at scale 22)"}
        log(f"cpu tc: {out['tc']}")
    for v in out.values():
        v.setdefault("kind", kind)
        v["cores"] = cpu_threads()
    return out


# ---------------------------------------------------------------------------

def make_workload(args, kernels):
    """Graphs, device handles and roots of the chosen workload."""
    from gapbs_amd import gpu
    from gapbs_amd.graph import KRONECKER, Graph
    name = args.workload
    W = {"build_edges": None, "host_gen_s": None, "host_build_s": None, "order_ms": {}}
    t0 = time.time()
    weighted_needed = bool({"sssp", "sssp_batch", "delta_sweep"} & set(kernels))
    if name == "kron26" or name == "kron22":
        scale, seed = (args.scale or 26, args.seed if args.seed is not None else 42) if name == "kron26" else \
            (args.scale or 22, args.seed if args.seed is not None else 8675309)
        W["scale"], W["seed"] = scale, seed
        W["desc"] = f"kronecker s{scale} ef16 seed {seed}"
        if "build" in kernels or "generate" in kernels:
            edges = Graph.generate_edges(KRONECKER, scale, 16, seed)[:2]
            W["host_gen_s"] = time.time() - t0
            t1 = time.time()
            g = Graph.from_edges(*edges)
            W["host_build_s"] = time.time() - t1
            W["build_edges"] = edges if "build" in kernels else None
        else:
            g = Graph.kronecker(scale, 16, seed=seed)
        wg = Graph.kronecker(scale, 16, seed=seed, weighted=True) if weighted_needed else None
        W["cpu_delta"], W["cpu_delta_why"] = 2, "the reference's setting for Kronecker graphs, benchmark/bench.mk:136"
    elif name == "uniform25":
        scale, seed = args.scale or 25, args.seed if args.seed is not None else 1234
        W["scale"], W["seed"] = scale, seed
        W["desc"] = f"uniform s{scale} deg 16 seed {seed}"
        g = Graph.uniform(scale, 16, seed=seed)
        wg = Graph.uniform(scale, 16, seed=seed, weighted=True) if weighted_needed else None
        W["cpu_delta"], W["cpu_delta_why"] = 2, "the reference's setting for uniform graphs, benchmark/bench.mk:140"
    elif name == "grid4096":
        side, seed = args.scale or 4096, args.seed if args.seed is not None else 7
        W["scale"], W["seed"] = side, seed
        W["desc"] = f"grid {side}x{side} (5% edges deleted, 1% shortcuts, weights 1..1000) seed {seed}"
        wg = Graph.grid(side, seed=seed)
        g = wg.strip_weights()
        W["cpu_delta"], W["cpu_delta_why"] = None, "the GPU's tuned width"
    else:
        raise SystemExit(f"unknown workload {name}")
    log(f"{g.stats_line()} built in {time.time() - t0:.1f}s")
    ordered = args.layout == "degree"
    W["g"], W["wg"] = g, wg
    nroots = args.roots or {"kron26": 64, "kron22": 16, "uniform25": 16, "grid4096": 32}[name]
    W["roots"] = [int(r) for r in g.pick_sources(nroots)]
    W["dg"] = gpu.DeviceGraph(g, order_by_degree=ordered)
    if ordered:
        W["order_ms"]["graph"] = round(W["dg"].order_stats["device_ms"], 3)
    W["dwg"] = None
    if wg is not None:
        W["dwg"] = gpu.DeviceGraph(wg, order_by_degree=ordered)
        if ordered:
            W["order_ms"]["weighted_graph"] = round(W["dwg"].order_stats["device_ms"], 3)
    W["relabel_tc"] = "tc" in kernels and g.worth_relabelling()
    W["tc_cpu"] = "reference" if name == "kron22" or args.tc_cpu == "reference" else "port"
    W["bc_cpu_roots"] = {"kron26": 1, "kron22": 4, "uniform25": 2}.get(name, 1)
    W["workload"] = W["desc"] + (", degree-ordered layout" if ordered else "")
    # m_r per root, from the device's own reach statistics.
    W["m_per_root"] = [gpu.bfs(W["dg"], r, want_depth=False, want_parent=False).stats["reached_edges"] // 2
                       for r in W["roots"]]
    log(f"device bytes {sum(d.device_bytes() for d in (W['dg'], W['dwg']) if d) / 1e9:.2f} GB")
    return W


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="kron26")
    ap.add_argument("--scale", type=int, default=None, help="graph scale (grid: side) instead of the workload's")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--roots", type=int, default=None, help="roots per batch (1..64); workload default")
    ap.add_argument("--sssp-roots", type=int, default=4, help="single-root SSSP calls per step (Kronecker)")
    ap.add_argument("--bc-roots", type=int, default=4, help="roots of the BC precision report")
    ap.add_argument("--delta", type=int, default=32, help="SSSP bucket width without autotuning")
    ap.add_argument("--batch-delta", type=int, default=None,
                    help="batched SSSP bucket width (default: searched, or --delta without autotuning)")
    ap.add_argument("--alpha", type=int, default=15, help="direction switch for both BFS paths")
    ap.add_argument("--kernels", default=None, help="comma-separated lines (default: the workload's)")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds per sampled CPU baseline")
    ap.add_argument("--tc-stride", type=int, default=64, help="CPU TC sample: one vertex in stride")
    ap.add_argument("--tc-cpu", choices=("sample", "reference"), default="sample",
                    help="CPU TC baseline at scale 26: a sampled OrderedCount (default, seconds) or the "
                         "reference's whole Hybrid (about 150 s on 16 cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-e2e", action="store_true", help="skip the timings with host result copies")
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
    kernels = list(WORKLOADS[args.workload]) if args.kernels is None else [k for k in args.kernels.split(",") if k]
    unknown = set(kernels) - set(WORKLOADS[args.workload])
    if unknown:
        sys.exit(f"unknown lines {sorted(unknown)} for workload {args.workload}")

    import numpy as np
    from gapbs_amd import gpu
    from gapbs_amd.graph import Graph

    gpu.init(0)
    copy_gbs = copy_bandwidth()
    log(f"device-to-device copy: {copy_gbs:.0f} GB/s (read + write)")
    W = make_workload(args, kernels)
    g, dg, dwg, roots, m_per_root = W["g"], W["dg"], W["dwg"], W["roots"], W["m_per_root"]
    m_total = sum(m_per_root)
    sssp_roots = roots[:args.sssp_roots] if args.workload != "grid4096" else roots
    bc_roots = roots[:args.bc_roots]
    bc_seq_roots = roots if args.workload == "kron26" else roots[:{"kron22": 4, "uniform25": 16}.get(
        args.workload, len(roots))]
    bc_batch_roots = bc_seq_roots

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
        if dwg is not None:
            tw = gpu.autotune(dwg, gpu.TUNE_SSSP, roots[:4])
            tuned = {**(tuned or {}), "sssp_delta": tw["sssp_delta"], "sssp_best_ms": tw["best_ms"][2]}
        if tuned is not None:
            tuned["seconds"] = round(time.time() - t0, 2)
            alpha, beta = tuned.get("bfs_alpha", alpha), tuned.get("bfs_beta", beta)
            batch_alpha, delta = tuned.get("batch_alpha", batch_alpha), tuned.get("sssp_delta", delta)
            log(f"autotune: {tuned}")
    if W["cpu_delta"] is None:
        W["cpu_delta"] = delta
    # Batched SSSP shares its buckets between roots, so its best width is
    # picked separately: from 16 times the single-root width, widths four
    # times larger, then four times smaller, while they get faster (one
    # untimed call each; results do not depend on the width).
    batch_delta = args.batch_delta or delta
    if "sssp_batch" in kernels and not args.no_autotune and args.batch_delta is None:
        t0 = time.time()
        times = {}

        def probe(d):
            if d not in times:
                times[d] = gpu.sssp_batch(dwg, roots, d, want_dist=False).stats["device_ms"]
            return times[d]

        start = min(16 * delta, 1 << 20)
        gpu.sssp_batch(dwg, roots, start, want_dist=False)  # allocations
        for step in (4, 0.25):
            d = start
            while 1 <= d * step <= (1 << 20) and probe(int(d * step)) < probe(d):
                d = int(d * step)
        batch_delta = min(times, key=times.get)
        tuned = {**(tuned or {}), "sssp_batch_delta": batch_delta,
                 "sssp_batch_ms": {str(d): round(v, 3) for d, v in times.items()},
                 "sssp_batch_seconds": round(time.time() - t0, 2)}
        log(f"batched SSSP width: {batch_delta} ({times})")

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
    packed_alpha = batch_alpha
    if "packed" in kernels:
        t0 = time.time()
        scales = np.random.default_rng(0).integers(10, 17, 256)
        pgraphs = [Graph.kronecker(int(sc), 16, seed=i) for i, sc in enumerate(scales)]
        union, bounds = Graph.pack(pgraphs)
        proots = np.stack([pg.pick_sources(8) + bounds[i] for i, pg in enumerate(pgraphs)])
        packed = (gpu.DeviceGraph(union), bounds, proots, [gpu.DeviceGraph(pg) for pg in pgraphs])
        log(f"packed workload: {union.num_nodes} vertices, {union.num_edges()} edges, built in "
            f"{time.time() - t0:.1f}s")
        if not args.no_autotune:  # the packed traversal's own direction switch, from one call per candidate
            gpu.bfs_packed(packed[0], bounds, proots, batch_alpha, want_depths=False)  # allocations, part check
            ptimes = {a: gpu.bfs_packed(packed[0], bounds, proots, a, want_depths=False).stats["device_ms"]
                      for a in (2, 4, 8, 15, 30, 60, 120, 240)}
            packed_alpha = min(ptimes, key=ptimes.get)
            tuned = {**(tuned or {}), "packed_batch_alpha": packed_alpha,
                     "packed_ms": {str(a): round(v, 3) for a, v in ptimes.items()}}
            log(f"packed batch alpha: {packed_alpha} ({ptimes})")

    n = g.num_nodes
    host_out = {}

    def out_buf(key, shape, dtype):
        """A reused, pre-touched host result array (first touch outside timing)."""
        a = host_out.get(key)
        if a is None or a.shape != tuple(shape) or a.dtype != np.dtype(dtype):
            a = np.empty(shape, dtype=dtype)
            a.fill(0)
            host_out[key] = a
        return a

    def run_kernel(k, e2e=False):
        if k == "bfs_batch":
            o = out_buf("bfs_batch", (len(roots), n), np.int32) if e2e else None
            return gpu.bfs_batch(dg, roots, batch_alpha, want_depths=e2e, out=o).stats
        if k == "bfs":
            for r in roots:
                gpu.bfs(dg, r, alpha, beta, want_depth=e2e, want_parent=False)
        elif k == "sssp":
            o = out_buf("sssp", (n,), np.int32) if e2e else None
            for r in sssp_roots:
                gpu.sssp(dwg, r, delta, want_dist=e2e, out=o)
        elif k == "sssp_batch":
            o = out_buf("sssp_batch", (len(roots), n), np.int32) if e2e else None
            return gpu.sssp_batch(dwg, roots, batch_delta, want_dist=e2e, out=o).stats
        elif k == "pr":
            o = out_buf("pr", (n,), np.float32) if e2e else None
            return gpu.pagerank(dg, 20, 1e-4, want_scores=e2e, out=o).stats
        elif k == "pr_fp64":
            o = out_buf("pr", (n,), np.float32) if e2e else None
            return gpu.pagerank(dg, 20, 1e-4, fp64=True, want_scores=e2e, out=o).stats
        elif k == "cc":
            o = out_buf("cc", (n,), np.int32) if e2e else None
            return gpu.cc(dg, want_comp=e2e, out=o).stats
        elif k == "cc_sv":
            return gpu.cc_sv(dg, want_comp=e2e).stats
        elif k == "bc":
            o = out_buf("bc", (n,), np.float32) if e2e else None
            return gpu.bc(dg, bc_seq_roots, want_scores=e2e, out=o).stats
        elif k == "bc_batch":
            o = out_buf("bc", (n,), np.float32) if e2e else None
            return gpu.bc_batch(dg, bc_batch_roots, want_scores=e2e, out=o).stats
        elif k == "tc":
            # Hybrid (src/tc.cc:95-100): relabel by degree when worth it --
            # on the device, inside the step as in the reference's trial.
            if not W["relabel_tc"]:
                return gpu.tc(dg)[1]
            dr, rst = gpu.relabel_by_degree(dg)
            if PROFILING[0]:
                dr.profile(True)
            c, st = gpu.tc(dr)
            if PROFILING[0]:
                EXTRA_PROFILES.append(dr.profile_report())
            dr.free()
            return {**st, "relabel_ms": rst["device_ms"], "triangles": c}
        elif k == "build":
            db, st = gpu.build_graph(*W["build_edges"])
            db.free()
            return st
        elif k == "generate":
            db, st = gpu.generate_graph(0, W["scale"], 16, W["seed"])
            db.free()
            return st
        elif k == "load":
            db, st = gpu.read_graph(sg_path[0])
            db.free()
            return st
        elif k == "packed":
            return gpu.bfs_packed(packed[0], packed[1], packed[2], packed_alpha, want_depths=False).stats
        elif k == "packed_per_graph":
            for i, (dgi, r) in enumerate(zip(packed[3], packed[2])):
                gpu.bfs_batch(dgi, r - packed[1][i], packed_alpha, want_depths=False)
        elif k == "delta_sweep":
            return delta_sweep()
        return None

    def delta_sweep():
        """SSSP device time per bucket width from one root (single-root) and
        from all roots (batched), every width's distances checked equal."""
        res = {"single_root_ms": {}, "batched_ms": {}}
        ref_d = None
        for d in (16, 64, 256, 1024, 4096):
            s = gpu.sssp(dwg, roots[0], d)
            b = gpu.sssp_batch(dwg, roots, d, want_dist=(d == 4096))
            res["single_root_ms"][d] = round(s.stats["device_ms"], 3)
            res["batched_ms"][d] = round(b.stats["device_ms"], 3)
            if ref_d is None:
                ref_d = s.dist
            assert np.array_equal(s.dist, ref_d), f"delta {d} changed the distances"
            if b.dist is not None:
                assert np.array_equal(b.dist[0], ref_d), "batched distances differ"
        return res

    traffic = load_traffic(W["workload"])
    devs = [d for d in (dg, dwg) if d is not None]
    results = {}
    for k in kernels:
        steps = min(args.steps, STEP_CAP.get(k, args.steps))
        warm = min(args.warmup, WARM_CAP.get(k, args.warmup))
        for _ in range(warm):
            run_kernel(k)
        if k == "bc" and warm == 0:  # allocations and the fold lists, outside the timed step
            gpu.bc(dg, roots[:1], want_scores=False)
        if k in ("pr", "pr_fp64"):  # the Jacobi layout is built once per graph, outside timing
            gpu.pagerank_blocks(dg, fp64=(k == "pr_fp64"))
            gpu.pagerank_layout(dg, fp64=(k == "pr_fp64"))
        last = None
        t = time.perf_counter()
        for _ in range(steps):
            last = run_kernel(k)
        per = (time.perf_counter() - t) / steps
        r = {"seconds": round(per, 6), "steps": steps}
        if k != "delta_sweep":
            pdevs = devs + ([packed[0]] if k == "packed" else [])
            prof, psec = profiled(pdevs, lambda: run_kernel(k), 1 if k in STEP_CAP else min(steps, 2))
            r["profiled_seconds"] = round(psec, 6)
            # (the packed graphs are another workload: no counter traffic for them)
            r["roofline"] = roofline(prof, {} if k == "packed" else traffic, copy_gbs)
            r["kernels"] = kernel_table(prof)
        if not args.no_e2e and k in ("bfs_batch", "bfs", "sssp", "sssp_batch", "pr", "pr_fp64", "cc", "bc",
                                     "bc_batch"):
            run_kernel(k, e2e=True)  # allocate and touch the host arrays
            esteps = 1 if k in STEP_CAP else min(steps, 2)
            t = time.perf_counter()
            for _ in range(esteps):
                run_kernel(k, e2e=True)
            r["e2e_seconds"] = round((time.perf_counter() - t) / esteps, 6)
        m = {"bfs_batch": m_total, "bfs": m_total, "sssp_batch": m_total,
             "sssp": sum(m_per_root[:len(sssp_roots)]), "bc": sum(m_per_root[:len(bc_seq_roots)]),
             "bc_batch": sum(m_per_root[:len(bc_batch_roots)])}.get(k)
        if m is not None:
            r["gteps"] = round(m / per / 1e9, 3)
            if "e2e_seconds" in r:
                r["e2e_gteps"] = round(m / r["e2e_seconds"] / 1e9, 3)
            r["roots"] = {"bfs_batch": len(roots), "bfs": len(roots), "sssp_batch": len(roots),
                          "sssp": len(sssp_roots), "bc": len(bc_seq_roots), "bc_batch": len(bc_batch_roots)}[k]
        if k in ("bfs_batch", "sssp_batch") and last:
            r["iterations"] = last["iterations"]
            assert last["reached_edges"] // 2 == m_total, f"{k}: reach differs from the single-root runs"
        if k in ("pr", "pr_fp64", "cc_sv") and last:
            r["iterations"] = last["iterations"]
        if k in ("pr", "pr_fp64"):  # the layout the Jacobi iterations went through (built before timing)
            blocks = gpu.pagerank_blocks(dg, fp64=(k == "pr_fp64"))
            r["layout"] = ({"propagation_blocks": blocks} if blocks["width_log2"] else
                           {"source_segments": gpu.pagerank_layout(dg, fp64=(k == "pr_fp64"))})
        if k == "tc" and last is not None:
            r["probes"] = last["edges_examined"]
            r["triangles"] = last.get("triangles")
            if "relabel_ms" in last:
                r["relabel_ms_last"] = round(last["relabel_ms"], 3)
        if k == "build" and last is not None:
            r.update(build_report(g, W["build_edges"], last, W["host_build_s"], args.no_cpu))
            W["build_edges"] = None
        if k == "generate" and last is not None:
            r.update(generate_report(g, W["scale"], W["seed"], last, W["host_gen_s"], W["host_build_s"],
                                     args.no_cpu))
        if k == "packed" and last is not None:
            mp = last["reached_edges"] // 2
            for _ in range(args.warmup):
                run_kernel("packed_per_graph")
            t = time.perf_counter()
            for _ in range(steps):
                run_kernel("packed_per_graph")
            per_graph = (time.perf_counter() - t) / steps
            r.update({"gteps": round(mp / per / 1e9, 3), "graphs": len(packed[3]), "roots": int(packed[2].size),
                      "per_graph_seconds": round(per_graph, 6), "speedup_vs_per_graph": round(per_graph / per, 3)})
        if k == "load":
            r.update(load_report(g, sg_path[0], per, args.no_cpu))
            os.remove(sg_path[0])
            os.rmdir(os.path.dirname(sg_path[0]))
        if k == "delta_sweep":
            r.update(last)
        results[k] = r
        log(f"{k}: {r}")

    if "bfs_batch" in results and "bfs" in results:
        results["bfs_batch"]["speedup_vs_sequential"] = round(results["bfs"]["seconds"] /
                                                              results["bfs_batch"]["seconds"], 3)
    if "sssp_batch" in results and "sssp" in results:  # per-root rate: the single-root line may run fewer roots
        results["sssp_batch"]["speedup_vs_sequential"] = round(results["sssp_batch"]["gteps"] /
                                                               results["sssp"]["gteps"], 3)
    if "bc_batch" in results and "bc" in results:
        results["bc_batch"]["speedup_vs_sequential"] = round(results["bc"]["seconds"] /
                                                             results["bc_batch"]["seconds"], 3)
    if "pr" in results and "pr_fp64" in results:
        s32 = gpu.pagerank(dg, 20, 1e-4).scores.astype(np.float64)
        s64 = gpu.pagerank(dg, 20, 1e-4, fp64=True).scores.astype(np.float64)
        results["pr"]["vs_fp64"] = {"max_abs": float(np.abs(s32 - s64).max()),
                                    "l1": float(np.abs(s32 - s64).sum()),
                                    "fp64_over_fp32_time": round(results["pr_fp64"]["seconds"] /
                                                                 results["pr"]["seconds"], 3)}
        log(f"pr precision: {results['pr']['vs_fp64']}")
    if "bc" in results:
        s64 = gpu.bc(dg, bc_roots, fp64=True).scores.astype(np.float64)
        err = {"roots": len(bc_roots)}
        for name, kw in (("fp32", {}), ("kahan", {"kahan": True})):
            d = np.abs(gpu.bc(dg, bc_roots, **kw).scores.astype(np.float64) - s64)
            err[name] = {"max_abs": float(d.max()), "l1": float(d.sum())}
        results["bc"]["vs_fp64"] = err
        log(f"bc precision: {err}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baselines(W, kernels, args.cpu_budget, args.tc_stride)
        for k, v in cpu.items():
            for line in (k, k + "_batch"):
                if line in results:
                    results[line]["cpu"] = v
        if "cc" in cpu and "cc_sv" in results:
            results["cc_sv"]["cpu"] = cpu["cc"]
    head_name = next((k for k in ("bfs_batch", "sssp_batch", "bfs", "sssp", "bc") if k in results),
                     next(iter(results)))
    head = results[head_name]
    try:
        import psutil
        kids = [f"{p.pid}:{p.name()}" for p in psutil.Process().children(recursive=True)]
    except Exception as e:  # noqa: BLE001
        kids = [f"unknown ({e})"]
    out = {
        "metric": METRIC,
        "value": head.get("gteps"),
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["seconds"] * 1e3, 3),
        "e2e_ms_per_step": round(head["e2e_seconds"] * 1e3, 3) if "e2e_seconds" in head else None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic",
        "config": {"workload": f"{W['workload']}: {head_name} of {head.get('roots')} SourcePicker roots per step "
                               f"(headline); other lines: {', '.join(k for k in results if k != head_name)}",
                   "name": args.workload,
                   "num_nodes": g.num_nodes, "num_edges_directed": g.num_edges_directed,
                   "roots": len(roots), "bfs_alpha_beta": [alpha, beta], "batch_alpha": batch_alpha,
                   "sssp_delta": delta, "sssp_batch_delta": batch_delta, "autotune": tuned, "layout": args.layout,
                   "order_ms": W["order_ms"] or None, "copy_gbs": round(copy_gbs, 1),
                   "host": {"cpu_model": cpu_model(), "cpu_threads": cpu_threads()},
                   "timing": "seconds: results left in HBM, captured level loops replayed (no per-kernel "
                             "events); e2e_seconds: results copied to host memory through the C-ABI; "
                             "roofline and kernels: a separate profiled pass"},
        "roofline": head.get("roofline"),
        "kernels": results,
        "cpu_baseline": ({**cpu["bfs"], "model": cpu_model(),
                          "sample": cpu["bfs"]["sample"] + " (the reference has no batched BFS: its trial loop "
                          "runs one root at a time)"} if cpu and "bfs" in cpu and head_name.startswith("bfs")
                         else (cpu.get(head_name.replace("_batch", "")) if cpu else None)),
        "child_processes_at_end": kids,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
