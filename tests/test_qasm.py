"""OpenQASM 2.0 front end (svmi_qasm_parse, sv-sim_amd/csrc/svmi_qasm.cpp).

Pinned against what the reference publishes and against closed forms (the
reference translator, svsim/QASMBench/svsim_qasm.py, is a Python program that
writes Python; it is read, not run):
  * every QASMBench circuit the reference ships (tests/golden/qasm/, copied by
    tests/golden/make_qasm_golden.py) parses to the qubit, gate and CNOT
    counts its README lists (counted with the translator's tables,
    svsim_qasm.py:30-88);
  * circuits with known results, simulated by the oracle: cat state, the
    Bernstein-Vazirani hidden string, the half adder, the QFT network, and
    inverseqft_n4 whose `if`-guarded gates the translator drops;
  * the parser's own rules: expressions, parameterised user gates, register
    broadcast, ignored statements, error messages with line numbers.
"""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import qasm

GOLD = os.path.join(os.path.dirname(__file__), "golden")
QDIR = os.path.join(GOLD, "qasm")
ALIAS = {"simon_n6": "simons_n6"}   # README row name differs from the file name


def _files():
    return sorted(f[:-5] for f in os.listdir(QDIR) if f.endswith(".qasm"))


def _run(name):
    n, g, info = qasm.load(os.path.join(QDIR, name + ".qasm"))
    re, im = O.run(n, g, nthreads=4)
    return n, re + 1j * im, info


def test_qasmbench_counts_match_the_published_tables():
    with open(os.path.join(GOLD, "qasmbench_counts.json")) as f:
        want = json.load(f)
    checked = 0
    for name in _files():
        n, g, info = qasm.load(os.path.join(QDIR, name + ".qasm"))
        assert info["n_gates"] == len(g) and info["n_qubits"] == n
        row = want.get(ALIAS.get(name, name))
        if row is None:
            continue
        assert (n, info["basic_gates"], info["cx_gates"]) == (row["qubits"], row["gates"], row["cnot"]), name
        checked += 1
    assert checked >= 56


def test_cat_state_and_bernstein_vazirani():
    n, psi, _ = _run("cat_state_n4")
    want = np.zeros(16, complex)
    want[0] = want[15] = 1 / np.sqrt(2)
    assert np.allclose(psi, want, atol=1e-15)
    n, psi, _ = _run("bv_n14")                   # "Hidden string is 1111111111111" (file header)
    p = np.abs(psi) ** 2
    data = np.arange(1 << n) & ((1 << 13) - 1)
    assert abs(p[data == (1 << 13) - 1].sum() - 1.0) < 1e-12


def test_half_adder():
    # x q0; x q1 (a = b = 1), q2 = 0; the circuit leaves q1 = a xor b, q3 = a and b
    n, psi, _ = _run("adder_n4")
    assert abs(abs(psi[0b1001]) - 1.0) < 1e-12


def test_qft_network():
    # H on q0 first, then phases controlled by higher qubits, no final swaps:
    # the DFT of the bit-reversed input, x = |q0 q2 set> -> rev(5) = 10
    n, psi, _ = _run("qft_n4")
    k = np.arange(16)
    assert np.allclose(psi, np.exp(2j * np.pi * 10 * k / 16) / 4, atol=1e-14)


def test_if_statements_are_dropped_like_the_translator():
    # h q; then h on every qubit with the `if`-guarded u1 gates dropped: |0000>
    n, psi, info = _run("inverseqft_n4")
    assert abs(psi[0] - 1.0) < 1e-14 and info["skipped"] >= 10


PROGRAM = """
OPENQASM 2.0;
include "qelib1.inc";
// parameterised user gates, nested, with expressions
gate rot(a, b) q { u3(a, b / 2, -a) q; rz(2^3 * b) q; }
gate pair(t) x, y { rot(t, pi) x; cu1(sin(pi/6) + t) x, y; cx y, x; }
qreg a[2]; qreg b[2];
creg c[2];
U(0.1, 0.2, 0.3) a[0]; CX a[0],
   b[1];                       // a statement over two lines
pair(-0.25) a[1], b[0];
h b;  x a[1]; barrier a, b; measure a -> c; reset b[0];
if (c == 1) x a[0];
rzz(sqrt(4) * 0.5) a[0], a[1]; ry(1e-1) b[1]; u2(.5, ln(exp(1))) b[0]; id a[0];
"""


def test_parser_rules():
    n, g, info = qasm.parse(PROGRAM)
    assert n == 4 and info["skipped"] == 4
    got = [(svmi.OP_NAMES[int(r["op"])], tuple(int(q) for q in r["qb"]), float(r["theta"]), float(r["phi"]),
            float(r["lambda"])) for r in g]
    pi = np.pi
    want = [
        ("U3", (0, 0, 0, 0, 0), 0.1, 0.2, 0.3),
        ("CX", (0, 3, 0, 0, 0), 0, 0, 0),
        ("U3", (1, 0, 0, 0, 0), -0.25, pi / 2, 0.25),        # rot(t, pi) on a[1]
        ("RZ", (1, 0, 0, 0, 0), 0, 8 * pi, 0),
        ("CU1", (1, 2, 0, 0, 0), 0, 0, 0.5 - 0.25),
        ("CX", (2, 1, 0, 0, 0), 0, 0, 0),
        ("H", (2, 0, 0, 0, 0), 0, 0, 0), ("H", (3, 0, 0, 0, 0), 0, 0, 0),   # broadcast over b
        ("X", (1, 0, 0, 0, 0), 0, 0, 0),
        ("RZZ", (0, 1, 0, 0, 0), 1.0, 0, 0),
        ("RY", (3, 0, 0, 0, 0), 0.1, 0, 0),
        ("U2", (2, 0, 0, 0, 0), 0, 0.5, 1.0),
        ("ID", (0, 0, 0, 0, 0), 0, 0, 0),
    ]
    assert len(got) == len(want)
    for a, b in zip(got, want):
        assert a[:2] == b[:2] and np.allclose(a[2:], b[2:], atol=1e-15), (a, b)
    # u3 + rz + cu1 + cx from `pair`, U, CX ... by the translator's tables
    assert info["cx_gates"] == 1 + 2 + 1 + 2


@pytest.mark.parametrize("src,msg", [
    ("qreg q[2];\nfoo q[0];", "line 2: unknown gate 'foo'"),
    ("qreg q[2];\ncx q[0];", "line 2: gate 'cx' acts on 2 qubits"),
    ("qreg q[2];\nrx q[0];", "line 2: gate 'rx' takes 1 parameters"),
    ("qreg q[2];\ncx q[0], q[0];", "line 2: gate 'cx' repeats a qubit"),
    ("qreg q[2];\nh q[2];", "line 2: q[2] is out of range"),
    ("qreg q[2]; qreg r[3];\ncx q, r;", "line 2: registers of different sizes"),
    ("qreg q[2];\ngate g a {\nh a;\n", "unterminated body of gate 'g'"),
    ("qreg q[1];\nrx(pi/) q[0];", "line 2: incomplete expression"),
    ("qreg q[1];\nrx(theta) q[0];", "line 2: unknown name 'theta'"),
    ("qreg q[1];\nopaque magic a;\nmagic q[0];", "line 3: opaque gate 'magic'"),
])
def test_errors_name_the_line(src, msg):
    with pytest.raises(svmi.SvmiError) as e:
        qasm.parse(src)
    assert msg in str(e.value), str(e.value)


def test_translator_command_writes_a_libsvsim_script(tmp_path):
    out = tmp_path / "circuit.py"
    r = subprocess.run([sys.executable, "-m", "svmi.qasm", "-i", os.path.join(QDIR, "wstate_n3.qasm"), "-o", str(out)],
                       capture_output=True, text=True, cwd=os.path.join(os.path.dirname(GOLD), "..", "sv-sim_amd"))
    assert r.returncode == 0, r.stderr
    assert "Number of qubits: 3" in r.stdout and "Number of basic gates: 30" in r.stdout
    assert "Number of cnot gates: 9" in r.stdout
    src = out.read_text()
    assert "sim.append(sim.U3(1.91063, 0.0, 0.0, 0))" in src and "sim.append(sim.CCX(0, 1, 2))" in src
    compile(src, str(out), "exec")


ALL_GATES_QASM = """OPENQASM 2.0;
include "qelib1.inc";
qreg q[5];
u3(0.1,0.2,0.3) q[0]; u2(0.4,0.5) q[1]; u1(0.6) q[2]; cx q[0],q[1]; id q[3]; x q[4]; y q[0]; z q[1];
h q[2]; s q[3]; sdg q[4]; t q[0]; tdg q[1]; rx(0.7) q[2]; ry(0.8) q[3]; rz(0.9) q[4]; cz q[0],q[2];
cy q[1],q[3]; swap q[2],q[4]; ch q[3],q[0]; ccx q[0],q[1],q[2]; cswap q[1],q[2],q[3]; crx(1.1) q[2],q[3];
cry(1.2) q[3],q[4]; crz(1.3) q[4],q[0]; cu1(1.4) q[0],q[1]; cu3(1.5,1.6,1.7) q[1],q[2]; rxx(1.8) q[2],q[3];
rzz(1.9) q[3],q[4]; rccx q[0],q[1],q[2]; rc3x q[0],q[1],q[2],q[3]; c3x q[1],q[2],q[3],q[4];
c3sqrtx q[0],q[2],q[3],q[4]; c4x q[0],q[1],q[2],q[3],q[4];
"""


def test_translator_writes_a_cplus_program_that_builds(tmp_path):
    """The C++ output (svsim/tool/svsim_qasm_cplus.py's program, on the
    SVSim::Simulation facade): every append maps back to the parsed record,
    and the program compiles and links against the library."""
    src = tmp_path / "all.qasm"
    src.write_text(ALL_GATES_QASM)
    out = tmp_path / "all.cpp"
    pkg = os.path.join(os.path.dirname(GOLD), "..", "sv-sim_amd")
    r = subprocess.run([sys.executable, "-m", "svmi.qasm", "-i", str(src), "-o", str(out)],
                       capture_output=True, text=True, cwd=pkg)
    assert r.returncode == 0, r.stderr
    assert "Number of qubits: 5" in r.stdout
    n, gates, _ = qasm.load(str(src))
    assert len({svmi.OP_NAMES[int(g["op"])] for g in gates}) >= 30
    lines = [ln.strip() for ln in out.read_text().splitlines() if ln.strip().startswith("add(Simulation::")]
    assert len(lines) == len(gates)
    for ln, g in zip(lines, gates):
        name, rest = ln[len("add(Simulation::"):].split("(", 1)
        args = [a.strip() for a in rest.rsplit("));", 1)[0].split(",")]
        assert name == svmi.OP_NAMES[int(g["op"])]
        angles = qasm._ANGLES.get(name, ())
        assert [float(a) for a in args[:len(angles)]] == [float(g[a]) for a in angles]
        assert [int(a) for a in args[len(angles):]] == [int(q) for q in g["qb"][:qasm._ARITY.get(name, 1)]]
    assert "int n_qubits = 5;" in out.read_text()
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = tmp_path / "all_bin"
    r = subprocess.run(["g++", "-std=c++17", "-O1", str(out), "-I", os.path.join(root, "include"),
                        "-I", os.path.join(root, "sv-sim_amd", "csrc"), "-L", os.path.join(root, "sv-sim_amd"),
                        "-lsvmi", "-o", str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
