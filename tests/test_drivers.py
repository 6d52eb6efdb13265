"""The command-line drivers, tested the way the reference's test.mk tests its
binaries: the generate/load stats lines (test/test.mk:42-66, CPU only: -n0
never touches the GPU) and `-vn1` verification for every kernel
(test/test.mk:73-86, on the GPU)."""

import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "gapbs_amd", "bin")
KERNELS = ["bc", "bfs", "cc", "cc_sv", "pr", "sssp", "tc"]


def run(args, timeout=300):
    p = subprocess.run([os.path.join(BIN, args[0])] + args[1:], capture_output=True, text=True,
                       timeout=timeout)
    return p.returncode, p.stdout


@pytest.mark.parametrize("flag", ["g10", "u10"])
def test_generate(golden, flag):
    rc, out = run(["bfs", f"-{flag}", "-n0"])
    assert rc == 0 and golden["stats_lines"][flag] in out


def test_load_serialized_and_reject_bad_suffix(tmp_path):
    p = str(tmp_path / "g.sg")
    rc, out = run(["converter", "-g10", "-b", p])
    assert rc == 0
    rc, out = run(["bfs", "-f", p, "-n0"])
    assert rc == 0 and "1024 nodes and 10496 undirected edges" in out
    rc, out = run(["bfs", "-f", str(tmp_path / "g.xyz"), "-n0"])
    assert rc == 256 - 3 and "Unrecognized suffix" in out
    rc, out = run(["bfs", "-f", str(tmp_path / "missing.sg"), "-n0"])
    assert rc == 256 - 6


def test_no_input_and_directed_tc(tmp_path):
    rc, out = run(["bfs"])
    assert rc == 255 and "No graph input specified" in out
    el = tmp_path / "d.sg"
    # a directed graph for tc: build with the Python host layer
    from gapbs_amd.graph import Graph
    Graph.from_edges([0, 1], [1, 2], symmetrize=False).write(str(el))
    rc, out = run(["tc", "-f", str(el), "-n1"])
    assert rc == 256 - 2 and "requires undirected" in out


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", KERNELS)
def test_verify_g10(kernel):
    print(f"[driver] {kernel} -g10 -vn1", flush=True)
    rc, out = run([kernel, "-g10", "-vn1"])
    assert rc == 0, out
    assert "Verification:           PASS" in out, out


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,extra", [("bfs", ["-n4"]), ("sssp", ["-n4", "-d2"]),
                                          ("pr", ["-n2", "-p", "fp64"]), ("cc", ["-n2"]), ("cc_sv", ["-n2"]),
                                          ("bc", ["-n2", "-i4"]), ("bc", ["-n1", "-i4", "-p", "kahan"]),
                                          ("tc", ["-n1"])])
def test_verify_g14(kernel, extra):
    print(f"[driver] {kernel} -g14 -v {extra}", flush=True)
    rc, out = run([kernel, "-g14", "-v"] + extra)
    assert rc == 0, out
    n = out.count("Verification:           PASS")
    if "kahan" in extra:  # compensated sums differ from the reference's float rounding
        assert n + out.count("Verification:           FAIL") >= 1
    else:
        assert n >= 1 and "FAIL" not in out, out


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,extra", [("bfs", ["-n2", "-j", "16"]), ("bc", ["-n1", "-i", "8", "-j", "8"]),
                                          ("bc", ["-n1", "-i", "8", "-j", "8", "-p", "fp64"]),
                                          ("pr", ["-n1", "-p", "gs"])])
def test_verify_batched_and_gauss_seidel(kernel, extra):
    print(f"[driver] {kernel} -g12 -v {extra}", flush=True)
    rc, out = run([kernel, "-g12", "-v"] + extra)
    assert rc == 0, out
    if "fp64" in extra:  # double arithmetic differs from the float reference by more than FLT_EPSILON
        assert out.count("Verification:") == 1, out
    else:
        assert "Verification:           PASS" in out and "FAIL" not in out, out


@pytest.mark.gpu
@pytest.mark.parametrize("kernel,extra", [("bfs", ["-n2"]), ("bfs", ["-n1", "-j", "16"]), ("sssp", ["-n2", "-d2"]),
                                          ("pr", ["-n1"]), ("cc", ["-n2"]), ("tc", ["-n1"]),
                                          ("bc", ["-n1", "-i4"])])
def test_verify_degree_ordered(kernel, extra):
    # -o renumbers the device copy by degree; results keep the graph's ids
    # and must pass the reference's verifiers (BC sums its dependencies in
    # the uploaded successor order, so its scores meet FLT_EPSILON too).
    print(f"[driver] {kernel} -g12 -v -o {extra}", flush=True)
    rc, out = run([kernel, "-g12", "-v", "-o"] + extra)
    assert rc == 0, out
    assert "Degree Ordering:" in out, out
    assert "Verification:           PASS" in out and "FAIL" not in out, out
