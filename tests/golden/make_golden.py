"""Writes tests/golden/golden.json from the reference (built from its own
sources by `make -C oracle ref`) and the reference's own test fixtures.

Run in a container that has /root/reference:  python tests/golden/make_golden.py
The JSON holds data only: expected stats lines, hashes of the reference's
serialized graphs, SourcePicker root sequences and per-root result digests
of the reference kernels, so a machine without the reference sources can pin
the oracle and the GPU path against them.
"""

import hashlib
import json
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gapbs_amd.graph import Graph  # noqa: E402
from oracle import oracle, reference as ref  # noqa: E402

REF_TEST = "/root/reference/test/reference"
REF_BIN = os.path.join(ROOT, "oracle", "_ref")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    out = {"stats_lines": {}, "sg_sha256": {}, "roots": {}, "bfs": {}, "sssp": {}, "pr": {}, "cc": {}, "bc": {}, "tc": {}}
    # The reference's own expected stats lines (test/test.mk:44-49).
    for name in ("g10", "u10"):
        with open(os.path.join(REF_TEST, f"graph-{name}.out")) as f:
            out["stats_lines"][name] = f.read().strip()
    # Bytes of the reference converter's serialized graphs.
    with tempfile.TemporaryDirectory() as td:
        for flags, tag, suffix in ((["-g10"], "g10", ".sg"), (["-u10"], "u10", ".sg"),
                                   (["-g16"], "g16", ".sg"), (["-g10", "-w"], "g10w", ".wsg")):
            path = os.path.join(td, tag + suffix)
            subprocess.run([os.path.join(REF_BIN, "converter"), *flags, "-b", path], check=True,
                           stdout=subprocess.DEVNULL)
            with open(path, "rb") as f:
                out["sg_sha256"][tag] = hashlib.sha256(f.read()).hexdigest()
    # SourcePicker sequences and reference DOBFS depth digests.
    for tag, g in (("g10", Graph.kronecker(10)), ("g16", Graph.kronecker(16)),
                   ("u10", Graph.uniform(10))):
        roots = ref.pick_sources(g, 16)
        out["roots"][tag] = [int(r) for r in roots]
        rg = ref.RefGraph("bfs", g)
        parents, _ = ref.bfs(rg, roots)
        digests = []
        for r, p in zip(roots, parents):
            d = oracle.depth_from_parent(p, int(r))
            digests.append({"root": int(r), "depth_sha256": sha(d),
                            "reached": int((d >= 0).sum()), "max_depth": int(d.max())})
        out["bfs"][tag] = digests
    # Reference DeltaStep distance digests (delta 2, as benchmark/bench.mk:136-147).
    for tag, g in (("g10w", Graph.kronecker(10, weighted=True)),
                   ("g16w", Graph.kronecker(16, weighted=True)),
                   ("u12w", Graph.uniform(12, weighted=True))):
        roots = ref.pick_sources(g.strip_weights(), 16)
        dists, _ = ref.sssp(ref.RefGraph("sssp", g), roots, 2)
        out["sssp"][tag] = [{"root": int(r), "dist_sha256": sha(d),
                             "reached": int((d < (2**31 - 1) // 2).sum()),
                             "max_dist": int(d[d < (2**31 - 1) // 2].max())}
                            for r, d in zip(roots, dists)]
    # Reference PageRankPull (pr_spmv) float scores, tolerance 1e-4, <= 20 iterations.
    for tag, g in (("g10", Graph.kronecker(10)), ("g16", Graph.kronecker(16)),
                   ("u12", Graph.uniform(12))):
        scores, _ = ref.pagerank("pr_spmv", g, 20, 1e-4)
        top = np.argsort(-scores, kind="stable")[:5]
        out["pr"][tag] = {"scores_sha256": sha(scores), "sum": float(scores.astype(np.float64).sum()),
                          "top5": [[int(i), float(scores[i])] for i in top]}
    # Reference Afforest labels.
    for tag, g in (("g10", Graph.kronecker(10)), ("g16", Graph.kronecker(16)),
                   ("u14d2", Graph.uniform(14, degree=2))):
        comp, _ = ref.afforest(g)
        out["cc"][tag] = {"comp_sha256": sha(comp), "components": int(len(np.unique(comp)))}
    # Reference Brandes scores (its own SourcePicker roots), -i4 as in
    # benchmark/bench.mk:158-159.
    for tag, g, k in (("g10", Graph.kronecker(10), 4), ("g16", Graph.kronecker(16), 4),
                      ("u12", Graph.uniform(12), 8)):
        scores, _ = ref.brandes(g, k)
        out["bc"][tag] = {"k": k, "roots": [int(r) for r in ref.pick_sources(g, k)],
                          "scores_sha256": sha(scores),
                          "argmax": int(np.argmax(scores))}
    # Reference triangle counts (Hybrid) and its relabelling decision.
    for tag, g in (("g10", Graph.kronecker(10)), ("g16", Graph.kronecker(16)),
                   ("u12", Graph.uniform(12))):
        count, _ = ref.tc(g)
        out["tc"][tag] = {"triangles": count, "worth_relabelling": ref.worth_relabelling(g)}
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote golden.json")


if __name__ == "__main__":
    main()
