"""GPU parity of the QASMBench circuits the reference ships, parsed by the
OpenQASM front end and run by the engine, against the oracle running the same
gate records (amplitudes within 1e-10); larger ones through closed forms; the
translator's script and Simulation.append_qasm through libsvsim."""
import os
import subprocess
import sys

import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import qasm

pytestmark = pytest.mark.gpu

QDIR = os.path.join(os.path.dirname(__file__), "golden", "qasm")
TOL = 1e-10


def _names(max_qubits):
    out = []
    for f in sorted(os.listdir(QDIR)):
        if f.endswith(".qasm"):
            n, _, _ = qasm.load(os.path.join(QDIR, f))
            if n <= max_qubits:
                out.append(f[:-5])
    return out


def test_qasmbench_up_to_20_qubits_vs_oracle():
    names = _names(20)
    assert len(names) >= 45
    for name in names:
        n, g, _ = qasm.load(os.path.join(QDIR, name + ".qasm"))
        print(f"case: {name} ({n} qubits, {len(g)} gates)", flush=True)
        with svmi.Engine(n) as e:
            e.upload(g)
            e.run()
            re, im = e.get_state()
        ore, oim = O.run(n, g, nthreads=8)
        err = max(np.abs(re - ore).max(), np.abs(im - oim).max())
        assert err < TOL, (name, err)


def test_large_qasmbench_closed_forms():
    for name, n_want in (("cat_state_n22", 22), ("ghz_state_n23", 23)):
        n, g, _ = qasm.load(os.path.join(QDIR, name + ".qasm"))
        assert n == n_want
        print(f"case: {name}", flush=True)
        with svmi.Engine(n) as e:
            e.upload(g)
            e.run()
            re, im = e.get_amplitudes([0, (1 << n) - 1])
            tot = e.measure(np.array([]))[1]
        assert np.allclose(np.abs(re + 1j * im), 1 / np.sqrt(2), atol=1e-12) and abs(tot - 1) < 1e-12


def test_translator_script_and_append_qasm_through_libsvsim(tmp_path):
    import libsvsim
    src = os.path.join(QDIR, "adder_n10.qasm")
    out = tmp_path / "circuit.py"
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sv-sim_amd")
    r = subprocess.run([sys.executable, "-m", "svmi.qasm", "-i", src, "-o", str(out)], cwd=pkg,
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    # the generated script appends its gates to a libsvsim.Simulation
    ns = {"__name__": "circuit"}
    sys_argv = sys.argv
    try:
        sys.argv = ["circuit.py", "10", "0"]
        exec(compile(out.read_text(), str(out), "exec"), ns)
    finally:
        sys.argv = sys_argv
    re1, im1 = ns["sim"].get_sv()
    sim = libsvsim.Simulation(10, 0)
    with open(src) as f:
        assert sim.append_qasm(f.read()) > 0
    sim.upload()
    sim.run()
    re2, im2 = sim.get_sv()
    assert np.array_equal(re1, re2) and np.array_equal(im1, im2)
    n, g, _ = qasm.load(src)
    ore, oim = O.run(n, g)
    assert max(np.abs(re2 - ore).max(), np.abs(im2 - oim).max()) < TOL


def test_cplus_translator_program_runs(tmp_path):
    """The C++ program the translator writes (svsim_qasm_cplus.py's main:
    upload, sim, measure(5), print_measurement) built against the facade and
    run: adder_n10 ends in one basis state, so all 5 measured states are the
    oracle's."""
    n, g, _ = qasm.load(os.path.join(QDIR, "adder_n10.qasm"))
    ore, oim = O.run(n, g)
    p = ore ** 2 + oim ** 2
    want = int(np.argmax(p))
    assert abs(p[want] - 1) < 1e-12
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    pkg = os.path.join(root, "sv-sim_amd")
    cpp, exe = tmp_path / "adder.cpp", tmp_path / "adder_bin"
    r = subprocess.run([sys.executable, "-m", "svmi.qasm", "-i", os.path.join(QDIR, "adder_n10.qasm"), "-o", str(cpp)],
                       cwd=pkg, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    r = subprocess.run(["g++", "-std=c++17", "-O1", str(cpp), "-I", os.path.join(root, "include"),
                        "-I", os.path.join(pkg, "csrc"), "-L", pkg, "-lsvmi", "-Wl,-rpath," + pkg, "-o", str(exe)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print("case: adder_n10 C++ program", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=90, env=dict(os.environ, SVSIM_QUIET="1"))
    assert r.returncode == 0, r.stderr
    tests = [ln.split(": ")[1] for ln in r.stdout.splitlines() if ln.startswith("Test-")]
    assert "Measurement (tests=5)" in r.stdout and tests == [format(want, f"0{n}b")] * 5
