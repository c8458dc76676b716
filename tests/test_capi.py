"""CPU checks of the C-ABI library: it loads and exports exactly the entry
points include/ecg.h declares, and the host-side wrapper declares them all.
No compute call is made (there is no GPU here)."""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

from ecg_nonbo import hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    text = open(os.path.join(ROOT, "include", "ecg.h")).read()
    return sorted(set(re.findall(r"^int (ecg_\w+)\(", text, flags=re.M)))


def test_library_exports_every_declared_entry_point():
    assert os.path.exists(hip.LIB_PATH), "build libecg_hip.so first (python atom-mol-nonbo_amd/build.py)"
    lib = ctypes.CDLL(hip.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", hip.LIB_PATH], capture_output=True, text=True).stdout
    exported = sorted(set(re.findall(r" T (ecg_\w+)$", out, flags=re.M)))
    assert exported == declared


def test_wrapper_binds_every_declared_entry_point():
    assert sorted(hip.EXPORTS) == _declared()
    hip.load()   # declares prototypes; touches no device


def test_library_is_gfx950():
    data = open(hip.LIB_PATH, "rb").read()
    assert b"gfx950" in data
