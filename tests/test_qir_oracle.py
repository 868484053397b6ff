"""Pins the QIR-gate-set oracle (oracle/qir_oracle.c) to the reference.

The reference here is the QIR-facing engine svsim_qsharp/svsim_cpu_omp.hpp
(multi-controlled gates through masks, 328-394 and OP_HEAD_MASK 606-615;
Pauli measure-and-collapse, 1406-1500).  Checks, in order of strength:
  * bit-exact agreement with fixtures written by the reference itself
    (tests/golden/make_qir_golden.py via oracle/_ref);
  * bit-exact agreement with the live reference build where oracle/_ref exists;
  * closed forms that do not go through the reference: every op's 2x2
    matrix against the matrix exponential of its Pauli generator, controlled
    ops against the block-diagonal brute-force unitary, and measurement
    against P(1) = (1 - <P>)/2 and the projector (I -+ P)/2 for P = Z, X, Y.
"""
import json
import os
import sys

import numpy as np
import pytest
from scipy.linalg import expm

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLD)
import make_qir_golden as MG  # noqa: E402

Q = O.QIR
I2 = np.eye(2, dtype=complex)
PX = np.array([[0, 1], [1, 0]], dtype=complex)
PY = np.array([[0, -1j], [1j, 0]])
PZ = np.diag([1.0 + 0j, -1.0])


def _state(n, seed):
    rng = np.random.default_rng(seed)
    return MG.random_state(n, rng)


def test_kats_bit_exact():
    d = np.load(os.path.join(GOLD, "qir_kats.npz"))
    for i, g in enumerate(d["gates"]):
        re, im = d["in_re"][i].copy(), d["in_im"][i].copy()
        p, o, _ = O.qir_apply(int(d["n"][i]), re, im, np.array([g], dtype=O.QIR_GATE_DTYPE), nthreads=2)
        assert np.array_equal(re, d["out_re"][i]) and np.array_equal(im, d["out_im"][i]), \
            f"{O.QIR_OPS[g['op']]} case {i}"
        if g["op"] == Q["Measure"]:
            assert p[0] == d["prob"][i] and o[0] == d["outcome"][i]


def test_program_n8_bit_exact():
    d = np.load(os.path.join(GOLD, "qir_program_n8.npz"))
    re = np.zeros(256)
    im = np.zeros(256)
    re[0] = 1.0
    p, o, last = O.qir_apply(8, re, im, d["gates"], nthreads=3)
    assert np.array_equal(re, d["re"]) and np.array_equal(im, d["im"])
    assert np.array_equal(p, d["prob"]) and np.array_equal(o, d["outcome"])
    assert last == d["prob"][-1]
    assert 0 < o.sum() < len(o)            # both outcomes occur


def test_hashes_bit_exact():
    with open(os.path.join(GOLD, "qir_hashes.json")) as f:
        want = json.load(f)
    for n, ng, nm, seed in [(12, 600, 16, 31), (14, 800, 20, 37)]:
        rng = np.random.default_rng(seed)
        g = MG.random_program(n, ng, nm, rng)
        re, im = MG.random_state(n, rng)
        p, o, _ = O.qir_apply(n, re, im, g, nthreads=4)
        w = want[f"qir_random_n{n}"]
        assert MG.sha(re, im, p) == w["sha256"]
        assert [int(x) for x in o] == w["outcomes"]


@pytest.mark.skipif(not O.ref_qir_available(), reason="oracle/_ref QIR engine not built (no reference sources)")
def test_live_reference_random_programs():
    for n, seed in ((3, 1), (7, 2), (10, 3)):
        rng = np.random.default_rng(seed)
        g = MG.random_program(n, 300, 10, rng)
        re, im = MG.random_state(n + 1, rng)
        re[1 << n:] = 0.0   # one unallocated slot
        im[1 << n:] = 0.0
        a_re, a_im = re.copy(), im.copy()
        p, o, _ = O.qir_apply(n, a_re, a_im, g, nthreads=4)
        b_re, b_im, bp, bo = O.ref_qir_run(n + 1, n, g, re, im, nthreads=3)
        assert np.array_equal(a_re, b_re) and np.array_equal(a_im, b_im)
        assert np.array_equal(p, bp) and np.array_equal(o, bo)


def _matrix_of(op, theta):
    """2x2 action of a one-qubit op, read off the oracle on basis states."""
    M = np.zeros((2, 2), dtype=complex)
    for col in range(2):
        re = np.zeros(2)
        im = np.zeros(2)
        re[col] = 1.0
        O.qir_apply(1, re, im, [(op, 0, 0, theta)], nthreads=1)
        M[:, col] = re + 1j * im
    return M


def test_single_qubit_closed_forms():
    t = 0.7312
    s2 = 1 / np.sqrt(2)
    want = {
        "X": PX, "Y": PY, "Z": PZ, "H": s2 * np.array([[1, 1], [1, -1]], dtype=complex),
        "S": np.diag([1, 1j]), "T": np.diag([1, np.exp(1j * np.pi / 4)]),
        "AdjointS": np.diag([1, -1j]), "AdjointT": np.diag([1, np.exp(-1j * np.pi / 4)]),
        "RI": expm(-0.5j * t * I2), "RX": expm(-0.5j * t * PX), "RY": expm(-0.5j * t * PY),
        "RZ": expm(-0.5j * t * PZ),
        "EI": expm(1j * t * I2), "EX": expm(1j * t * PX), "EY": expm(1j * t * PY), "EZ": expm(1j * t * PZ),
    }
    for name, M in want.items():
        got = _matrix_of(Q[name], t)
        assert np.allclose(got, M, atol=1e-15), name
        # the controlled form with an empty mask is the same gate
        if "Controlled" + name in Q:
            assert np.allclose(_matrix_of(Q["Controlled" + name], t), M, atol=1e-15), name


def _embed(n, target, controls, U):
    """Brute-force 2^n unitary of U on `target`, applied where all controls are 1."""
    D = 1 << n
    full = np.zeros((D, D), dtype=complex)
    for col in range(D):
        on = all((col >> c) & 1 for c in controls)
        b = (col >> target) & 1
        if not on:
            full[col, col] = 1.0
            continue
        for r in range(2):
            row = (col & ~(1 << target)) | (r << target)
            full[row, col] += U[r, b]
    return full


def test_controlled_ops_brute_force():
    n = 4
    t = -1.234
    for name in ["X", "Y", "Z", "H", "S", "T", "RI", "RX", "RY", "RZ", "EI", "EX", "EY", "EZ",
                 "AdjointS", "AdjointT"]:
        U = _matrix_of(Q[name], t)
        for target, controls in [(0, [1]), (2, [0, 3]), (3, [0, 1, 2]), (1, [])]:
            mask = sum(1 << c for c in controls)
            re, im = _state(n, 5)
            psi = re + 1j * im
            O.qir_apply(n, re, im, [(Q["Controlled" + name], target, mask, t)], nthreads=2)
            assert np.allclose(re + 1j * im, _embed(n, target, controls, U) @ psi, atol=1e-14), (name, controls)
    # a mask naming the target, or a qubit above n, leaves the state alone
    for mask in (0b0101, 1 << 7):
        re, im = _state(n, 6)
        r0, i0 = re.copy(), im.copy()
        O.qir_apply(n, re, im, [(Q["ControlledX"], 0, mask, 0.0)], nthreads=1)
        assert np.array_equal(re, r0) and np.array_equal(im, i0)


def test_swap_is_a_permutation():
    n = 5
    re, im = _state(n, 9)
    psi = re + 1j * im
    O.qir_apply(n, re, im, [(Q["Swap"], 1, 3, 0.0)], nthreads=2)
    idx = np.arange(1 << n)
    src = (idx & ~0b01010) | (((idx >> 1) & 1) << 3) | (((idx >> 3) & 1) << 1)
    assert np.array_equal(re + 1j * im, psi[src])


def _pauli_full(n, q, P):
    M = np.array([[1.0 + 0j]])
    for b in range(n - 1, -1, -1):
        M = np.kron(M, P if b == q else I2)
    return M


@pytest.mark.parametrize("basis,P", [(0, PZ), (1, PX), (2, PY)])
def test_measure_projector_closed_form(basis, P):
    n = 4
    for q in range(n):
        for draw in (0.02, 0.98):
            re, im = _state(n, 40 + q)
            psi = re + 1j * im
            Pf = _pauli_full(n, q, P)
            p1 = 0.5 * (1.0 - np.real(np.vdot(psi, Pf @ psi)))
            p, o, last = O.qir_apply(n, re, im, [(Q["Measure"], q, basis, draw)], nthreads=1)
            assert abs(p[0] - p1) < 1e-14 and last == p[0]
            assert o[0] == (1 if draw < p1 else 0)
            sign = -1.0 if o[0] == 1 else 1.0
            post = 0.5 * (np.eye(1 << n) + sign * Pf) @ psi
            post /= np.sqrt(p1 if o[0] == 1 else 1.0 - p1)
            assert np.allclose(re + 1j * im, post, atol=1e-14)
            assert abs(np.sum(re * re + im * im) - 1.0) < 1e-13


def test_rejects_bad_records():
    re, im = _state(3, 1)
    for g in [(99, 0, 0, 0.0), (Q["X"], 3, 0, 0.0), (Q["Swap"], 1, 1, 0.0), (Q["Swap"], 1, 3, 0.0)]:
        with pytest.raises(ValueError):
            O.qir_apply(3, re, im, [g])
