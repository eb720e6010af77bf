"""The C++ drop-in layer (include/hc/*.h): reference-named drivers and
registries checked by tests/cpp/verify.cpp against the oracle.  The CPU
variant runs the host-core drivers only; the GPU variant runs every GPU and
heterogeneous driver."""
from __future__ import annotations

import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")
BIN = os.path.join(CPP, "bin", "verify")


def _verify_binary():
    # make rebuilds the program when a header it includes has changed.
    subprocess.run(["make", "-C", CPP], check=True, capture_output=True)
    return BIN


def _run(mode):
    r = subprocess.run([_verify_binary(), mode], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    return r.stdout


def test_cpp_cpu_drivers(hc):
    out = _run("cpu")
    assert "PASS SourcePicker" in out and "PASS bfs_do_cpu" in out


def test_cpp_schedule_dump(hc, tmp_path):
    p = tmp_path / "s.skd"
    p.write_text("NVIDIA Tesla K80\n0 1\nSSSP GPU block-red 256 1024\n"
                 "Intel i7-9700K\n2 2\nSSSP CPU one-to-one\n")
    r = subprocess.run([_verify_binary(), "skd", str(p)], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stderr
    assert "segments 3 blocks 2" in r.stdout
    assert "NVIDIA Tesla K80#0 [0,1] SSSP GPU block-red 256 1024" in r.stdout


@pytest.mark.gpu
def test_cpp_gpu_drivers(hc, gpu):
    out = _run("gpu")
    assert "PASS sssp_pull_heterogeneous" in out and "PASS BFS GPU bucketed parents" in out
