"""GPU: svmi_dm_qir_run (the density-matrix engine, svsim_qsharp_noise) on the
device against the oracle restatement (bit-exact against the reference
engine on the CPU) -- states within 1e-10, measurement outcomes equal -- and
at 14 qubits (a 28-qubit state) against the state-vector engine through
rho = |psi><psi| on the diagonal."""
import os
import sys

import numpy as np
import pytest

import svmi
from oracle import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLD)
from make_qir_golden import random_program  # noqa: E402

TOL = 1e-10


def test_golden_and_random_programs_vs_oracle():
    d = np.load(os.path.join(GOLD, "dm_program_n4.npz"))
    with svmi.Engine(8) as e:
        probs, outs = e.dm_qir_run(4, d["gates"])
        re, im = e.get_state()
    assert max(np.abs(re - d["re"]).max(), np.abs(im - d["im"]).max()) < TOL
    assert np.abs(probs - d["probs"]).max() < TOL and np.array_equal(outs, d["outcomes"])
    rng = np.random.default_rng(808)
    for n in (1, 2, 3, 5, 8, 10):
        g = random_program(n, 60, 3, rng) if n > 1 else [(O.QIR["H"], 0, 0, 0.0), (O.QIR["Measure"], 0, 1, 0.4),
                                                           (O.QIR["RY"], 0, 0, 0.7)]
        print(f"case: n={n}", flush=True)
        with svmi.Engine(2 * n) as e:
            probs, outs = e.dm_qir_run(n, g)
            re, im = e.get_state()
        ore, oim, oprobs, oouts = O.dm_qir_apply(n, g, nthreads=8)
        assert max(np.abs(re - ore).max(), np.abs(im - oim).max()) < TOL, n
        assert np.abs(probs - oprobs).max() < TOL and np.array_equal(outs, oouts), n


def test_handle_too_small_is_an_error():
    with svmi.Engine(6) as e:
        with pytest.raises(svmi.SvmiError) as err:
            e.dm_qir_run(4, [(O.QIR["H"], 0, 0, 0.0)])
    assert "2 * n_qubits" in str(err.value)


def test_n14_diagonal_matches_state_vector_engine():
    # no Swap (row-only in the density-matrix engine) and no Measure: rho =
    # |psi><psi|, so rho[i, i] = |psi_i|^2 with psi from svmi_qir_run
    n = 14
    rng = np.random.default_rng(14)
    g = O.as_qir_gates(random_program(n, 300, 0, rng))
    g = g[g["op"] != O.QIR["Swap"]]
    with svmi.Engine(n) as e:
        e.qir_run(n, g)
        pre, pim = e.get_state()
    idx = np.array([(i << n) | i for i in range(1 << n)], dtype=np.uint64)
    with svmi.Engine(2 * n) as e:
        e.dm_qir_run(n, g)
        dre, dim_ = e.get_amplitudes(idx)
        off = e.get_amplitudes(np.array([(5 << n) | 3], dtype=np.uint64))
    assert np.abs(dre - (pre ** 2 + pim ** 2)).max() < 1e-12 and np.abs(dim_).max() < 1e-12
    want = (pre[3] + 1j * pim[3]) * np.conj(pre[5] + 1j * pim[5])   # rho[row 3, col 5]
    assert abs(off[0][0] + 1j * off[1][0] - want) < 1e-12
    assert abs(dre.sum() - 1.0) < 1e-10


def test_qir_noise_driver_on_the_device_vs_the_oracle_engine():
    """libsvsim.QirNoiseSimulator on the GPU engine against the same driver on
    the oracle engine: the same calls and draws, including a release and a
    re-allocation that grow the engine's qubit count (rho reinterpreted over
    the same amplitudes, as in the reference)."""
    import libsvsim
    sys.path.insert(0, os.path.dirname(__file__))
    from test_qir_noise_driver import OracleDmBackend, random_calls
    P = libsvsim.PauliId
    cap = 10
    be = OracleDmBackend(cap)
    sims = [libsvsim.QirNoiseSimulator(capacity=cap, device=0), libsvsim.QirNoiseSimulator(backend=be)]
    results = [[], []]
    for i, sim in enumerate(sims):
        draws = iter(np.random.default_rng(31).random(400))
        sim.set_draw(lambda d=draws: float(next(d)))
        qs = [sim.AllocateQubit() for _ in range(4)]
        random_calls(sim, qs, np.random.default_rng(32), 120, measure=True)
        results[i].append(sim.GetResultValue(sim.Measure([P.PauliId_Y], [qs[1]])))
        sim.ReleaseQubit(qs[3])
        qs[3] = sim.AllocateQubit()
        assert sim.engine_qubits() == 5
        random_calls(sim, qs, np.random.default_rng(33), 60, measure=True)
        results[i].append(sim.GetResultValue(sim.Measure([P.PauliId_X], [qs[0]])))
        results[i].append(sim.last_prob())
    print("case: noise driver, 4 live qubits, engine 4 then 5", flush=True)
    re, im = sims[0].get_sv()
    assert results[0][:2] == results[1][:2] and abs(results[0][2] - results[1][2]) < TOL
    assert max(np.abs(re - be.re).max(), np.abs(im - be.im).max()) < TOL
