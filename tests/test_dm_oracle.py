"""The oracle's restatement of the reference density-matrix engine
(svsim_qsharp_noise/svsim_cpu_omp.hpp; oracle/qir_oracle.c
orc_dm_qir_apply), pinned by:
  * the reference engine's own output on a committed program
    (tests/golden/dm_program_n4.npz, tests/golden/make_dm_golden.py) and, where
    oracle/_ref is built, live on random programs (bit-exact);
  * the closed form rho = |psi><psi| for programs without Swap and Measure,
    psi from the state-vector QIR oracle (pinned by its own tests);
  * a hand-computed measurement with the reference's row-only collapse.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _random_program(n, n_gates, n_meas, rng):
    import sys
    sys.path.insert(0, GOLD)
    from make_qir_golden import random_program
    return random_program(n, n_gates, n_meas, rng)


def test_reference_golden_program():
    d = np.load(os.path.join(GOLD, "dm_program_n4.npz"))
    re, im, probs, outs = O.dm_qir_apply(4, d["gates"], nthreads=2)
    assert np.array_equal(re, d["re"]) and np.array_equal(im, d["im"])
    assert np.array_equal(probs, d["probs"]) and np.array_equal(outs, d["outcomes"])
    assert len(outs) == 3


@pytest.mark.skipif(not O.ref_dm_available(), reason="oracle/_ref not built (reference sources absent)")
def test_live_reference_random_programs():
    rng = np.random.default_rng(77)
    for _ in range(20):
        n = int(rng.integers(2, 6))
        g = _random_program(n, 40, 3, rng)
        r = O.ref_dm_run(n, g)
        o = O.dm_qir_apply(n, g, nthreads=3)
        for a, b in zip(r, o):
            assert np.array_equal(a, b)


def test_closed_form_pure_state():
    rng = np.random.default_rng(12)
    for n in (1, 2, 4):
        g = _random_program(n, 30, 0, rng) if n > 1 else [(O.QIR["H"], 0, 0, 0.0), (O.QIR["RX"], 0, 0, 0.4)]
        g = O.as_qir_gates(g)
        g = g[g["op"] != O.QIR["Swap"]]
        re, im, _, _ = O.dm_qir_apply(n, g)
        sre = np.zeros(1 << n)
        sim = np.zeros(1 << n)
        sre[0] = 1.0
        O.qir_apply(n, sre, sim, g)
        psi = sre + 1j * sim
        rho = np.outer(psi, psi.conj())                  # rho[row, col]
        got = (re + 1j * im).reshape(1 << n, 1 << n).T   # index row | (col << n)
        assert np.abs(got - rho).max() < 1e-13
        assert abs(np.trace(got).real - 1) < 1e-13


def test_row_only_collapse():
    # |+><+| measured in Z with draw 0.3 < P(1) = 0.5: the reference keeps the
    # entries whose row bit is 1 and divides them by P(1) (the column is not
    # projected): rho' = [[0, 0], [1, 1]]
    g = [(O.QIR["H"], 0, 0, 0.0), (O.QIR["Measure"], 0, 0, 0.3)]
    re, im, probs, outs = O.dm_qir_apply(1, g)
    assert abs(probs[0] - 0.5) < 1e-15 and outs[0] == 1
    got = (re + 1j * im).reshape(2, 2).T
    assert np.allclose(got, [[0, 0], [1, 1]], atol=1e-15)
