"""The C-ABI library loads and exports every entry point include/*.h declares."""
import ctypes
import os
import re

import pytest

import svmi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"\b(svmi_[a-z_0-9]+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(svmi.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 20
    missing = [n for n in declared if not hasattr(L, n)]
    assert not missing, missing
    assert set(svmi.EXPORTS) == declared


def test_abi_version_and_struct_layout():
    assert svmi.lib().svmi_abi_version() == 2
    assert svmi.GATE_DTYPE.itemsize == 48 and svmi.PAULI_DTYPE.itemsize == 24
    assert ctypes.sizeof(svmi.BatchItem) == 56 and ctypes.sizeof(svmi.Stats) == 88


def test_pybind_module_surface():
    import libsvsim
    S = libsvsim.Simulation
    for name in svmi.OP_NAMES:
        assert hasattr(S, name), name
    for m in ("append", "upload", "clear_circuit", "run", "reset", "measure"):
        assert hasattr(S, m)
    g = S.U3(0.1, 0.2, 0.3, 0)
    assert isinstance(g, libsvsim.Gate)
    assert int(libsvsim.OP.RYY) == 37


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="checks the no-GPU error path")
def test_create_without_gpu_is_an_error():
    with pytest.raises(svmi.SvmiError) as e:
        svmi.Engine(4)
    assert e.value.status == 5
