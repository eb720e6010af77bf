"""The C ABI boundary (CPU only): the shared libraries load and export
every symbol their headers declare; argument errors come back as status
codes with a message (the reference prints and exits, src/util.h:16-22)."""
from __future__ import annotations

import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:hc|or)_[a-z0-9_]+)\s*\(", text)))


def exported(lib_path):
    L = C.CDLL(lib_path)
    return L


def test_library_exports_every_declared_symbol(hc):
    L = exported(hc.LIB_PATH)
    headers = [os.path.join(ROOT, "include", f) for f in os.listdir(os.path.join(ROOT, "include"))
               if f.endswith(".h")]
    names = [n for h in headers for n in declared(h)]
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_oracle_exports_every_declared_symbol(oracle):
    L = exported(oracle.LIB_PATH)
    names = declared(os.path.join(ROOT, "oracle", "oracle.h"))
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing and len(names) >= 12


def test_error_codes_without_gpu(hc):
    L = hc.lib()
    assert L.hc_version() >= 10000
    rc = L.hc_set_tuning(b"no.such.knob", 1)
    assert rc == -7 and b"unknown tuning key" in L.hc_last_error()
    assert L.hc_set_tuning(b"grid_per_cu", 0) == -1
    s = hc.HostGraphStruct()
    assert L.hc_host_graph_generate(0, 40, 16, 1, 1, C.byref(s)) == -1
    assert b"scale" in L.hc_last_error()
    assert L.hc_equal_edge_ranges(None, 4, 1, None) == -1
    hc.reset_tuning()


def test_every_tuning_key_round_trips(hc):
    # Each advertised knob accepts its current value and one in its range.
    try:
        for key in hc.TUNING_KEYS:
            v = hc.get_tuning(key)
            hc.set_tuning(key, v)
            assert hc.get_tuning(key) == v
            alt = 1 - v if v in (0, 1) else v // 2
            hc.set_tuning(key, alt)
            assert hc.get_tuning(key) == alt, key
    finally:
        hc.reset_tuning()


def test_device_calls_fail_cleanly_without_gpu(hc):
    import subprocess
    import sys
    # A fresh process: on a machine without a GPU every device entry point
    # must return HC_ERR_NO_DEVICE instead of crashing.
    code = (
        "import sys; sys.path.insert(0, %r)\n"
        "import importlib; hc = importlib.import_module('hetero-compute_amd')\n"
        "import ctypes as C\n"
        "rc = hc.lib().hc_init(0)\n"
        "if rc == 0: print('gpu'); raise SystemExit(0)\n"
        "assert rc == -8, rc\n"
        "g = hc.HostGraph.generate('kron', 6, 4)\n"
        "h = C.c_void_p()\n"
        "assert hc.lib().hc_graph_upload(g.struct_ref(), C.byref(h)) == -8\n"
        "p = C.c_void_p()\n"
        "assert hc.lib().hc_dmalloc(C.byref(p), 16) == -8\n"
        "print('nodevice')\n" % ROOT)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                         timeout=120)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() in ("gpu", "nodevice")
