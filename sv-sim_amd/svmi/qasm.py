"""OpenQASM 2.0 front end (svmi_qasm_parse) and the translator command.

    from svmi import qasm
    n, gates, info = qasm.load("qft_n15.qasm")    # GATE_DTYPE records for Engine.upload
    with svmi.Engine(n) as e:
        e.upload(gates); e.run()

The command line mirrors the reference translator svsim/QASMBench/svsim_qasm.py:

    python -m svmi.qasm -i adder.qasm -o svsim_circuit.py

writes a script for the `libsvsim` module (`python svsim_circuit.py n_qubits
gpu`) with one `sim.append(sim.OP(...))` per gate, and prints the qubit,
basic-gate and CX counts as the reference does.  With an output ending in
.cpp/.cc/.cxx (or --cplus) it writes, as svsim/tool/svsim_qasm_cplus.py
does, a C++ program for the `SVSim::Simulation` facade (svsim_amd.hpp): one
`Simulation::OP(...)` append per gate in prepare_circuit, then a main that
uploads, runs and prints 5 measured states:

    python -m svmi.qasm -i adder.qasm -o adder.cpp
    g++ -std=c++17 -O2 adder.cpp -I include -I sv-sim_amd/csrc -L sv-sim_amd -lsvmi
"""
import argparse
import ctypes
import sys

import numpy as np

from . import GATE_DTYPE, OP_NAMES, _check, _ptr, lib


class QasmInfo(ctypes.Structure):
    _fields_ = [("n_qubits", ctypes.c_int32), ("reserved", ctypes.c_int32), ("n_gates", ctypes.c_uint64),
                ("basic_gates", ctypes.c_uint64), ("cx_gates", ctypes.c_uint64), ("skipped", ctypes.c_uint64)]


def _bind():
    L = lib()
    if not getattr(L, "_qasm_bound", False):
        L.svmi_qasm_parse.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.POINTER(QasmInfo)]
        L._qasm_bound = True
    return L


def parse(text):
    """OpenQASM 2.0 source -> (n_qubits, GATE_DTYPE array, info dict)."""
    L = _bind()
    data = text.encode() if isinstance(text, str) else bytes(text)
    info = QasmInfo()
    _check(L.svmi_qasm_parse(data, len(data), None, 0, ctypes.byref(info)))
    gates = np.zeros(info.n_gates, dtype=GATE_DTYPE)
    if info.n_gates:
        _check(L.svmi_qasm_parse(data, len(data), _ptr(gates), len(gates), ctypes.byref(info)))
    return info.n_qubits, gates, {k: getattr(info, k) for k, _ in QasmInfo._fields_ if k != "reserved"}


def load(path):
    with open(path, "rb") as f:
        return parse(f.read())


# factory argument lists of the reference (svsim/src/py_wrapper.cpp:53-109):
# angle slots first, then the gate's qubits
_ANGLES = {"U3": ("theta", "phi", "lambda"), "U2": ("phi", "lambda"), "U1": ("lambda",), "RX": ("theta",),
           "RY": ("theta",), "RZ": ("phi",), "CRX": ("lambda",), "CRY": ("lambda",), "CRZ": ("lambda",),
           "CU1": ("lambda",), "CU3": ("theta", "phi", "lambda"), "RXX": ("theta",), "RZZ": ("theta",)}
_ARITY = {"CX": 2, "CZ": 2, "CY": 2, "SWAP": 2, "CH": 2, "CRX": 2, "CRY": 2, "CRZ": 2, "CU1": 2, "CU3": 2,
          "RXX": 2, "RZZ": 2, "CCX": 3, "CSWAP": 3, "RCCX": 3, "RC3X": 4, "C3X": 4, "C3SQRTX": 4, "C4X": 5}


def to_script(gates, comment=""):
    """Python source for the libsvsim module, one sim.append per gate."""
    out = ["import sys", "import libsvsim as svsim", ""]
    if comment:
        out += ["# " + line for line in comment.splitlines()] + [""]
    out += ["if len(sys.argv) != 3:", "    print('$python circuit.py n_qubits gpu')", "    exit()", "",
            "sim = svsim.Simulation(int(sys.argv[1]), int(sys.argv[2]))", ""]
    for g in gates:
        name, args = _gate_args(g)
        out.append(f"sim.append(sim.{name}({', '.join(args)}))")
    out += ["", "sim.upload()", "sim.run()", ""]
    return "\n".join(out)


def _gate_args(g):
    name = OP_NAMES[int(g["op"])]
    args = [repr(float(g[a])) for a in _ANGLES.get(name, ())]
    args += [str(int(q)) for q in g["qb"][:_ARITY.get(name, 1)]]
    return name, args


def to_cplus(n_qubits, gates, comment=""):
    """C++ source for the SVSim::Simulation facade, one append per gate
    (the program svsim/tool/svsim_qasm_cplus.py:346-372 writes)."""
    out = ["#include <cstdio>", '#include "svsim_amd.hpp"', ""]
    if comment:
        out += ["// " + line for line in comment.splitlines()] + [""]
    out += ["// Use the SVSim namespace to enable the C++ API", "using namespace SVSim;", "",
            "void prepare_circuit(Simulation& sim)", "{",
            "    auto add = [&sim](Gate* g) { sim.append(g); delete g; };"]
    for g in gates:
        name, args = _gate_args(g)
        out.append(f"    add(Simulation::{name}({', '.join(args)}));")
    out += ["}", "", "int main()", "{", f"    int n_qubits = {int(n_qubits)};", "    int i_gpu = 0;",
            "    Simulation sim(n_qubits, i_gpu);", "    prepare_circuit(sim);", "    sim.upload();", "    sim.sim();",
            "    auto* res = sim.measure(5);", "    print_measurement(res, n_qubits, 5);", "    delete[] res;",
            "    return 0;", "}", ""]
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description="OpenQASM 2.0 -> libsvsim circuit script or C++ program "
                                             "(svsim_qasm.py's and svsim_qasm_cplus.py's interface)")
    ap.add_argument("--input", "-i", required=True, help="input OpenQASM file, such as adder.qasm")
    ap.add_argument("--output", "-o", default=None,
                    help="output script (default: svsim_circuit.py, or svsim_circuit.cpp with --cplus)")
    ap.add_argument("--cplus", action="store_true", help="write a C++ program (implied by a .cpp/.cc/.cxx output)")
    a = ap.parse_args(argv)
    out = a.output or ("svsim_circuit.cpp" if a.cplus else "svsim_circuit.py")
    cplus = a.cplus or out.endswith((".cpp", ".cc", ".cxx"))
    n, gates, info = load(a.input)
    text = f"translated from {a.input}: {n} qubits"
    with open(out, "w") as f:
        f.write(to_cplus(n, gates, text) if cplus else to_script(gates, text))
    print(f"== SV-Sim: Translating {a.input} to {out} ==")
    print(f"Number of qubits: {n}")
    print(f"Number of basic gates: {info['basic_gates']}")
    print(f"Number of cnot gates: {info['cx_gates']}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
