"""The QIR runtime driver (sv-sim_amd/csrc/svsim_qir.hpp, bound as
libsvsim.QirSimulator) on the CPU, with the oracle as its engine.

The driver mirrors the reference's SVSimSimulator
(svsim_qsharp/svsim_qsharp_wrapper.cpp), which needs the Microsoft QIR
runtime headers and so cannot be built here; its behaviour is pinned by
closed forms instead: R / Exp / ControlledExp against the matrix exponential
of the Pauli string (the reference's CX-chain + RZ(-2 theta) decomposition,
including that a multi-target Exp treats an I on its first target as Z),
qubit release against the same circuit on the remaining qubits, and
measurement against the projector rule.
"""
from functools import reduce

import numpy as np
import pytest
from scipy.linalg import expm

import libsvsim
from oracle import oracle as O

P = libsvsim.PauliId
MAT = {P.PauliId_I: np.eye(2, dtype=complex), P.PauliId_X: np.array([[0, 1], [1, 0]], dtype=complex),
       P.PauliId_Y: np.array([[0, -1j], [1j, 0]]), P.PauliId_Z: np.diag([1.0 + 0j, -1.0])}
TOL = 1e-12


class OracleBackend(libsvsim.QirBackend):
    """The CPU oracle as the driver's engine (state of 2^cap amplitudes)."""

    def __init__(self, cap, state=None):
        super().__init__()
        self.cap = cap
        self.reset()
        if state is not None:
            self.re[:], self.im[:] = state
        self.calls = []

    def run(self, n, gates):
        arr = O.as_qir_gates([(g.op, g.qubit, g.mask, g.theta) for g in gates])
        self.calls.append((n, arr))
        p, o, self.last = O.qir_apply(n, self.re, self.im, arr, last_prob=self.last, nthreads=1)
        return list(p), [int(x) for x in o], self.last

    def reset(self):
        self.re, self.im = O.zero_state(self.cap)
        self.last = 0.0

    def capacity(self):
        return self.cap

    def psi(self):
        return self.re + 1j * self.im


def _rand_state(n, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    v /= np.linalg.norm(v)
    return v


def _op_on(n, mats):
    """Tensor product with mats[q] acting on qubit (bit) q."""
    return reduce(np.kron, [mats.get(q, np.eye(2)) for q in range(n - 1, -1, -1)])


def _driver(n, seed):
    psi = _rand_state(n, seed)
    be = OracleBackend(n, (psi.real.copy(), psi.imag.copy()))
    sim = libsvsim.QirSimulator(backend=be)
    qs = [sim.AllocateQubit() for _ in range(n)]
    assert [sim.slot_of(q) for q in qs] == list(range(n))
    return sim, be, qs, psi


def test_rotations_closed_form():
    t = 0.913
    for axis in (P.PauliId_I, P.PauliId_X, P.PauliId_Y, P.PauliId_Z):
        sim, be, qs, psi = _driver(3, 1)
        sim.R(axis, qs[1], t)
        sim.ControlledR([qs[0], qs[2]], axis, qs[1], -t)
        sim.flush()
        U = _op_on(3, {1: expm(-0.5j * t * MAT[axis])})
        Vc = expm(0.5j * t * MAT[axis])
        C = np.eye(8, dtype=complex)
        for col in range(8):
            if (col & 1) and (col >> 2) & 1:
                C[:, col] = 0
                for r in range(2):
                    row = (col & ~2) | (r << 1)
                    C[row, col] = Vc[r, (col >> 1) & 1]
        assert np.allclose(be.psi(), C @ U @ psi, atol=TOL), axis


@pytest.mark.parametrize("paulis", [[P.PauliId_X, P.PauliId_Z, P.PauliId_Y], [P.PauliId_Y, P.PauliId_I, P.PauliId_X],
                                    [P.PauliId_Z, P.PauliId_Z], [P.PauliId_I, P.PauliId_X, P.PauliId_Y]])
def test_exp_closed_form(paulis):
    t = -0.37
    n = 4
    sim, be, qs, psi = _driver(n, 2)
    targets = [qs[3], qs[0], qs[2]][: len(paulis)]
    sim.Exp(paulis, targets, t)
    sim.flush()
    # the decomposition rotates about Z on targets[0] whatever its Pauli is
    # when there are several targets (an I there acts as Z)
    eff = list(paulis)
    if len(eff) > 1 and eff[0] == P.PauliId_I:
        eff[0] = P.PauliId_Z
    G = _op_on(n, {sim.slot_of(q): MAT[p] for q, p in zip(targets, eff)})
    assert np.allclose(be.psi(), expm(1j * t * G) @ psi, atol=TOL)


def test_single_target_exp_and_controlled_exp():
    t = 1.21
    n = 4
    for p in (P.PauliId_I, P.PauliId_X, P.PauliId_Y, P.PauliId_Z):
        sim, be, qs, psi = _driver(n, 3)
        sim.Exp([p], [qs[2]], t)
        sim.flush()
        assert np.allclose(be.psi(), expm(1j * t * _op_on(n, {2: MAT[p]})) @ psi, atol=TOL)
    for paulis in ([P.PauliId_X], [P.PauliId_Y, P.PauliId_Z]):
        sim, be, qs, psi = _driver(n, 4)
        targets = [qs[1], qs[3]][: len(paulis)]
        sim.ControlledExp([qs[0], qs[2]], paulis, targets, t)
        sim.flush()
        G = _op_on(n, {sim.slot_of(q): MAT[p] for q, p in zip(targets, paulis)})
        E = expm(1j * t * G)
        on = np.array([(x & 1) and (x >> 2) & 1 for x in range(1 << n)], bool)
        want = psi.copy()
        sub = np.where(on)[0]
        want[sub] = (E @ psi)[sub]   # E acts within the control-on subspace
        assert np.allclose(be.psi(), want, atol=TOL), paulis


def test_release_swaps_into_the_last_slot():
    n = 4
    sim, be, qs, psi = _driver(n, 5)
    be.reset()
    a, b, c, d = qs
    sim.H(a)
    sim.ControlledX([a], d)
    sim.R(P.PauliId_Y, d, 0.4)
    sim.ControlledZ([a, d], b)
    sim.ReleaseQubit(c)              # c in |0>: slot 2 <-> slot 3, d now in slot 2
    assert sim.num_qubits() == 3 and sim.slot_of(d) == 2
    sim.H(b)
    sim.flush()
    # the same circuit on 3 qubits (a, b, d) in slots (0, 1, 2)
    ref_re, ref_im = O.zero_state(3)
    O.qir_apply(3, ref_re, ref_im, [("H", 0), ("ControlledX", 2, 1), ("RY", 2, 0, 0.4), ("ControlledZ", 1, 0b101),
                                    ("H", 1)])
    assert np.allclose(be.re[:8], ref_re, atol=TOL) and np.allclose(be.im[:8], ref_im, atol=TOL)
    assert np.allclose(be.re[8:], 0) and np.allclose(be.im[8:], 0)
    # a freed logical id is reused by the next allocation, in the last slot
    e = sim.AllocateQubit()
    assert sim.slot_of(e) == 3 and sim.num_qubits() == 4
    for q in (a, b, d, e):
        sim.ReleaseQubit(q)
    assert sim.num_qubits() == 0
    assert be.re[0] == 1.0 and np.count_nonzero(be.re) == 1 and not np.any(be.im)


def test_measure_draws_and_results():
    sim, be, qs, _ = _driver(2, 6)
    be.reset()
    draws = iter([0.3, 0.9, 0.1])
    sim.set_draw(lambda: next(draws))
    sim.H(qs[0])
    sim.ControlledX([qs[0]], qs[1])
    r0 = sim.Measure([P.PauliId_Z], [qs[0]])            # P(1) = 1/2, draw 0.3 -> One
    assert sim.GetResultValue(r0) == libsvsim.ResultValue.Result_One and abs(sim.last_prob() - 0.5) < TOL
    r1 = sim.Measure([P.PauliId_Z], [qs[1]])            # correlated: P(1) = 1
    assert sim.AreEqualResults(r0, r1) and r1 == sim.UseOne()
    r2 = sim.Measure([P.PauliId_X], [qs[1]])            # X basis of |1>: P(1) = 1/2, draw 0.1 -> One
    assert r2 == sim.UseOne() and abs(sim.last_prob() - 0.5) < TOL
    # the queue ran at each Measure, one Measure record per run
    assert [int((arr["op"] == O.QIR["Measure"]).sum()) for _, arr in be.calls] == [1, 1, 1]
    with pytest.raises(Exception):
        sim.Measure([P.PauliId_Z, P.PauliId_Z], [qs[0], qs[1]])
    with pytest.raises(Exception):
        sim.QubitToString(qs[0])


def test_default_draw_is_libc_rand():
    sim, be, qs, _ = _driver(1, 7)
    be.reset()
    sim.H(qs[0])
    outs = [sim.GetResultValue(sim.Measure([P.PauliId_X], [qs[0]])) for _ in range(8)]
    # |+> measured in X is always Zero (P(1) = 0, draw >= 0)
    assert all(o == libsvsim.ResultValue.Result_Zero for o in outs)


def test_too_many_qubits():
    sim = libsvsim.QirSimulator(backend=OracleBackend(3))
    for _ in range(3):
        sim.AllocateQubit()
    with pytest.raises(Exception):
        sim.AllocateQubit()
    assert sim.num_qubits() == 3
