"""GPU parity of the QIR gate set: multi-controlled gates and mid-circuit
Pauli measurement with collapse (svmi_qir_run) against the oracle
(oracle/qir_oracle.c, pinned bit-exact to svsim_qsharp/svsim_cpu_omp.hpp)
and the reference's fixtures.

Bar: amplitudes and P(1) within 1e-10 max abs error of the FP64 oracle;
outcomes equal (the programs' draws are chosen at least 1e-9 away from
every P(1), where the two reductions' rounding cannot flip the comparison).
At sizes the oracle cannot reach in seconds the checks are properties: the
norm after every collapse, a second Z measurement of a measured qubit
(P(1) exactly 0 or 1), and measurement of a GHZ state.
"""
import os
import sys

import numpy as np
import pytest

import svmi
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLD)
import make_qir_golden as MG  # noqa: E402

pytestmark = pytest.mark.gpu

TOL = 1e-10
Q = svmi.QIR_OP


def _err(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))))


def _safe_program(n, ng, nm, seed):
    while True:
        rng = np.random.default_rng(seed)
        g = MG.random_program(n, ng, nm, rng)
        re, im = MG.random_state(n, rng)
        r, i = re.copy(), im.copy()
        p, _, _ = O.qir_apply(n, r, i, g, nthreads=8)
        if np.all(np.abs(g["theta"][g["op"] == Q["Measure"]] - p) > 1e-9):
            return g, (re, im), (r, i, p)
        seed += 1000


def test_golden_program_n8():
    d = np.load(os.path.join(GOLD, "qir_program_n8.npz"))
    print("case: qir_program_n8", flush=True)
    with svmi.Engine(8) as e:
        p, o, last = e.qir_run(8, d["gates"])
        re, im = e.get_state()
    assert _err(re, d["re"]) < TOL and _err(im, d["im"]) < TOL
    assert _err(p, d["prob"]) < TOL and np.array_equal(o, d["outcome"])
    assert abs(last - d["prob"][-1]) < TOL


def test_golden_kats_every_op():
    d = np.load(os.path.join(GOLD, "qir_kats.npz"))
    cap = int(d["capacity"])
    print("case: qir_kats", flush=True)
    with svmi.Engine(cap) as e:
        for i, g in enumerate(d["gates"]):
            e.set_state(d["in_re"][i], d["in_im"][i])
            p, o, _ = e.qir_run(int(d["n"][i]), np.array([g], dtype=svmi.QIR_GATE_DTYPE))
            re, im = e.get_state()
            assert _err(re, d["out_re"][i]) < TOL and _err(im, d["out_im"][i]) < TOL, (O.QIR_OPS[g["op"]], i)
            if g["op"] == Q["Measure"]:
                assert abs(p[0] - d["prob"][i]) < TOL and o[0] == d["outcome"][i]


@pytest.mark.parametrize("n,tile", [(3, 12), (6, 4), (10, 6), (14, 12), (16, 8), (18, 12)])
def test_random_programs_vs_oracle(n, tile):
    g, st, (r, i, p) = _safe_program(n, 400, 16, seed=500 + n)
    print(f"case: random QIR program n={n} tile={tile}", flush=True)
    with svmi.Engine(n) as e:
        e.set_option(svmi.OPT_TILE_QUBITS, tile)
        e.set_state(*st)
        gp, go, last = e.qir_run(n, g)
        re, im = e.get_state()
    assert _err(re, r) < TOL and _err(im, i) < TOL
    assert _err(gp, p) < TOL and abs(last - p[-1]) < TOL
    assert np.array_equal(go, (g["theta"][g["op"] == Q["Measure"]] < p).astype(np.int32))


def test_prefix_of_a_larger_state_and_last_prob():
    # 9 allocated qubits inside a 12-qubit handle: only the first 2^9
    # amplitudes move; last_prob persists across calls until reset
    n, cap = 9, 12
    g, st, (r, i, p) = _safe_program(n, 200, 6, seed=77)
    rng = np.random.default_rng(1)
    fre = rng.standard_normal(1 << cap) * 0.01
    fim = rng.standard_normal(1 << cap) * 0.01
    fre[: 1 << n], fim[: 1 << n] = st
    print("case: prefix n=9 in 12-qubit handle", flush=True)
    with svmi.Engine(cap) as e:
        e.set_state(fre, fim)
        gp, go, last = e.qir_run(n, g)
        re, im = e.get_state()
        assert _err(re[: 1 << n], r) < TOL and _err(im[: 1 << n], i) < TOL
        assert np.array_equal(re[1 << n:], fre[1 << n:]) and np.array_equal(im[1 << n:], fim[1 << n:])
        _, _, last2 = e.qir_run(n, [("H", 0)])          # no Measure: unchanged
        assert last2 == last
        e.reset_state()
        _, _, last3 = e.qir_run(n, [])
        assert last3 == 0.0


def test_few_qubits_in_a_larger_handle():
    # n < 4 allocated qubits run inside a 4-qubit tile; the rest stays zero
    for n in (1, 2, 3):
        rng = np.random.default_rng(n)
        g = [("H", 0), ("RY", n - 1, 0, 0.3), ("Measure", 0, 1, 0.4), ("EX", 0, 0, 0.2)]
        if n > 1:
            g += [("ControlledX", n - 1, 1, 0.0), ("Swap", 0, n - 1, 0.0), ("Measure", n - 1, 2, 0.7)]
        st = MG.random_state(n, rng)
        r, i = st[0].copy(), st[1].copy()
        p, o, _ = O.qir_apply(n, r, i, g, nthreads=1)
        assert np.all(np.abs(np.array([x[3] for x in g if x[0] == "Measure"]) - p) > 1e-9)
        print(f"case: {n}-qubit program in 6-qubit handle", flush=True)
        with svmi.Engine(6) as e:
            full_re = np.zeros(64)
            full_im = np.zeros(64)
            full_re[: 1 << n], full_im[: 1 << n] = st
            e.set_state(full_re, full_im)
            gp, go, _ = e.qir_run(n, g)
            re, im = e.get_state()
        assert _err(re[: 1 << n], r) < TOL and _err(im[: 1 << n], i) < TOL
        assert np.all(re[1 << n:] == 0) and np.all(im[1 << n:] == 0)
        assert _err(gp, p) < TOL and np.array_equal(go, o)


def test_errors_leave_state_alone():
    with svmi.Engine(5) as e:
        e.set_random_state(3)
        before = e.get_state()
        for g in ([(99, 0, 0, 0.0)], [("X", 5, 0, 0.0)], [("Swap", 2, 2, 0.0)]):
            with pytest.raises(svmi.SvmiError):
                e.qir_run(5, g)
        with pytest.raises(svmi.SvmiError):
            e.qir_run(6, [])
        after = e.get_state()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])


def test_large_state_properties():
    # 26 qubits: GHZ preparation with multi-controlled gates, then mid-circuit
    # measurements; checked through closed forms
    n = 26
    g = [("H", 0)] + [("ControlledX", q, 1 << (q - 1), 0.0) for q in range(1, n)]
    g += [("Measure", 13, 0, 0.3)]                       # GHZ: P(1) = 1/2, outcome 1
    g += [("Measure", 25, 0, 0.5), ("Measure", 0, 1, 0.2)]
    g += [("ControlledX", 0, ((1 << n) - 1) & ~1, 0.0)]  # 25 controls
    print("case: 26-qubit GHZ with mid-circuit measurements", flush=True)
    with svmi.Engine(n) as e:
        p, o, last = e.qir_run(n, g)
        tot = e.measure(np.array([]))[1]
        amps = e.get_amplitudes([0, (1 << n) - 1, (1 << n) - 2])
    assert abs(p[0] - 0.5) < 1e-12 and o[0] == 1
    assert abs(p[1] - 1.0) < 1e-12 and o[1] == 1       # GHZ collapsed: all ones
    assert abs(p[2] - 0.5) < 1e-12                      # X basis of a Z eigenstate
    assert abs(tot - 1.0) < 1e-12
    # after the X-basis outcome the state is |1..1> (x) |+/-> on qubit 0;
    # the 25-control X on qubit 0 maps |-> -> -|->, |+> -> |+>
    s = 1 / np.sqrt(2)
    a_last = amps[0][1] + 1j * amps[1][1]
    a_prev = amps[0][2] + 1j * amps[1][2]
    assert abs(abs(a_last) - s) < 1e-12 and abs(abs(a_prev) - s) < 1e-12
    assert abs(amps[0][0]) < 1e-12


def test_qir_mixed_workload_n20():
    # the benchmark's program shape at 20 qubits: (n-1)-control X/Z, 1-3
    # control gates, Swap and 16 measurements, from |0..0>
    from svmi import workloads as W
    n = 20
    g = W.qir_mixed(n=n)
    re, im = O.zero_state(n)
    p, o, _ = O.qir_apply(n, re, im, g, nthreads=8)
    assert np.min(np.abs(g["theta"][g["op"] == Q["Measure"]] - p)) > 1e-9
    print("case: qir_mixed n=20", flush=True)
    with svmi.Engine(n) as e:
        gp, go, _ = e.qir_run(n, g)
        gre, gim = e.get_state()
    assert _err(gre, re) < TOL and _err(gim, im) < TOL
    assert _err(gp, p) < TOL and np.array_equal(go, o)


def test_runtime_driver_on_gpu_matches_oracle_engine():
    # the same QIR-runtime call sequence through libsvsim.QirSimulator on the
    # GPU engine and on the oracle (tests/test_qir_driver.py's backend)
    import libsvsim
    from test_qir_driver import OracleBackend
    P = libsvsim.PauliId
    cap = 10
    draws = [0.37, 0.81, 0.05, 0.66, 0.5, 0.12, 0.93]

    def program(sim):
        it = iter(draws)
        sim.set_draw(lambda: next(it))
        qs = [sim.AllocateQubit() for _ in range(8)]
        res = []
        for q in qs:
            sim.H(q)
        sim.ControlledX(qs[:5], qs[7])
        sim.Exp([P.PauliId_X, P.PauliId_Y, P.PauliId_Z], [qs[1], qs[3], qs[6]], 0.42)
        sim.ControlledExp([qs[0], qs[2]], [P.PauliId_Z, P.PauliId_X], [qs[4], qs[5]], -0.7)
        res.append(sim.Measure([P.PauliId_Z], [qs[2]]))
        sim.ControlledR([qs[1]], P.PauliId_Y, qs[3], 1.1)
        sim.T(qs[4])
        sim.AdjointS(qs[5])
        res.append(sim.Measure([P.PauliId_X], [qs[3]]))
        sim.ReleaseQubit(qs[2])          # measured: swap into the last slot first
        extra = sim.AllocateQubit()
        sim.ControlledAdjointT([qs[0], extra], qs[6])
        res.append(sim.Measure([P.PauliId_Y], [qs[6]]))
        res.append(sim.Measure([P.PauliId_Z], [qs[0]]))
        sim.flush()
        return [int(sim.GetResultValue(r)) for r in res], sim.last_prob(), sim.num_qubits()

    print("case: QIR runtime driver, GPU vs oracle engine", flush=True)
    gpu = libsvsim.QirSimulator(cap, 0)
    g_res = program(gpu)
    g_re, g_im = gpu.get_sv()
    be = OracleBackend(cap)
    c_res = program(libsvsim.QirSimulator(backend=be))
    assert g_res[0] == c_res[0] and g_res[2] == c_res[2]
    assert abs(g_res[1] - c_res[1]) < TOL
    assert _err(g_re, be.re) < TOL and _err(g_im, be.im) < TOL
