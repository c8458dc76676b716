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


def test_fortran_interfaces_cover_every_entry_point():
    """The ISO_C_BINDING module binds every entry point of include/ecg.h
    under its C name, and the host routines keep the reference's names."""
    src = open(os.path.join(ROOT, "atom-mol-nonbo_amd", "fortran", "ecg_gpu.f90")).read()
    bound = sorted(set(re.findall(r'bind\(C, name="(ecg_\w+)"\)', src)))
    assert bound == _declared()
    for routine in ("MatrixElements", "ComputeMatElem", "ComputeMatElemAndDeriv", "GHEPIIS", "EnergyGA",
                    "EnergyIA", "EnergyGAM", "EnergyIAM", "EnergyGB", "EnergyIB"):
        assert re.search(r"^\s*(subroutine|real\(dprec\) function) %s\(" % routine, src, flags=re.M), routine


def test_fortran_layer_compiles_and_links():
    import pytest
    sys_path = os.path.join(ROOT, "atom-mol-nonbo_amd")
    import importlib.util
    spec = importlib.util.spec_from_file_location("ecg_build", os.path.join(sys_path, "build.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    if not os.path.exists(os.environ.get("FC", mod.FLANG)):
        pytest.skip("no Fortran compiler")
    obj = mod.build_fortran()
    assert obj and os.path.exists(obj) and os.path.exists(mod.DRIVER)
    out = subprocess.run(["nm", "-u", mod.DRIVER], capture_output=True, text=True).stdout
    for name in ("ecg_ctx_create", "ecg_compute_mat_elem", "ecg_ghepiis", "ecg_energy", "ecg_energy_grad"):
        assert name in out, name
