"""The C++ drop-in layer (include/hc/*.h): reference-named drivers and
registries checked against the oracle by the programs in tests/cpp/.

  verify.cpp                 registries, SourcePicker, segmenter, drivers
  bfs_verify.cpp / sssp_verify.cpp / pr_verify.cpp
                             the reference's verifiers (tests/{bfs,sssp,pr}/
                             *_verify.cu) in shape: a graph file argument and
                             every driver called by its reference name

The CPU variants run the host-core drivers (the GPU sections report SKIP
without a device); the GPU variants run everything and must skip nothing."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "bin")


def _binary(name):
    # make rebuilds a program when a header it includes has changed.
    subprocess.run(["make", "-C", CPP], check=True, capture_output=True)
    return os.path.join(BIN, name)


def _run(args, timeout=600):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    return r.stdout


@pytest.fixture(scope="module")
def graph_files(hc, tmp_path_factory):
    d = tmp_path_factory.mktemp("graphs")
    g = hc.HostGraph.generate("kron", 12, 16, weighted=True, seed=91)
    paths = {"wsg": str(d / "k12.wsg"), "sg": str(d / "k12.sg"), "skd": str(d / "k12.skd"),
             "cpu_skd": str(d / "k12cpu.skd")}
    g.write(paths["wsg"], True)
    g.write(paths["sg"], False)
    with open(paths["skd"], "w") as f:
        f.write("NVIDIA Tesla K80\n0 2\nSSSP GPU block-red 64 256\n"
                "Intel i7-9700K\n3 4\nSSSP CPU one-to-one\n"
                "AMD Instinct MI355X\n5 7\nSSSP GPU bucketed\n")
    with open(paths["cpu_skd"], "w") as f:
        f.write("Intel Xeon E5-2686\n0 3\nSSSP CPU one-to-one\n"
                "Host CPU\n4 7\nSSSP CPU one-to-one\n")
    return paths


def test_cpp_cpu_drivers(hc):
    out = _run([_binary("verify"), "cpu"])
    assert "PASS SourcePicker" in out and "PASS bfs_do_cpu" in out


def test_cpp_schedule_dump(hc, tmp_path):
    p = tmp_path / "s.skd"
    p.write_text("NVIDIA Tesla K80\n0 1\nSSSP GPU block-red 256 1024\n"
                 "Intel i7-9700K\n2 2\nSSSP CPU one-to-one\n")
    r = subprocess.run([_binary("verify"), "skd", str(p)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert "segments 3 blocks 2" in r.stdout
    assert "NVIDIA Tesla K80#0 [0,1] SSSP GPU block-red 256 1024" in r.stdout


def test_reference_verifiers_cpu(hc, graph_files):
    out = _run([_binary("bfs_verify"), graph_files["sg"]])
    for k in ("CPU push by node", "CPU push by edge", "CPU pull", "CPU DO"):
        assert f"Verifying BFS {k} kernel ... PASS" in out, k
    out = _run([_binary("sssp_verify"), graph_files["wsg"], graph_files["cpu_skd"]])
    assert "Verifying SSSP CPU kernel ... PASS" in out
    assert "Verifying SSSP heterogeneous kernel (loaded schedule) ... PASS" in out
    assert "SSSP CPU Block" in out and " > Number of segments: 8" in out
    out = _run([_binary("pr_verify"), graph_files["wsg"]])
    assert "Verifying PR CPU kernel ... PASS" in out


@pytest.mark.gpu
def test_cpp_gpu_drivers(hc, gpu):
    out = _run([_binary("verify"), "gpu"])
    assert "PASS sssp_pull_heterogeneous" in out and "PASS BFS GPU bucketed parents" in out


@pytest.mark.gpu
def test_reference_verifiers_gpu(hc, gpu, graph_files):
    out = _run([_binary("bfs_verify"), graph_files["sg"]])
    assert "SKIP" not in out and out.count("PASS") == 9, out
    out = _run([_binary("sssp_verify"), graph_files["wsg"], graph_files["skd"]])
    assert "SKIP" not in out and out.count("PASS") == 9, out
    assert "SSSP GPU Block" in out and " > block count:   64" in out
    out = _run([_binary("pr_verify"), graph_files["wsg"]])
    assert "SKIP" not in out and out.count("PASS") == 6, out
