"""The density-matrix QIR runtime driver (sv-sim_amd/csrc/svsim_qir.hpp,
noise::SVSimSimulator, bound as libsvsim.QirNoiseSimulator) on the CPU, with
the density-matrix oracle as its engine.

The driver mirrors the reference's SVSimSimulator of
svsim_qsharp_noise/svsim_qsharp_wrapper.cpp, which needs the Microsoft QIR
runtime headers and so cannot be built here.  Its gate set is the
state-vector driver's (itself pinned by closed forms in test_qir_driver.py):
the same calls on both drivers must hand the engines the same records, and
without measurements the density matrix must be rho = psi psi^dagger of the
state-vector driver's psi.  Its own bookkeeping is pinned against the
wrapper's rules: handles are indices, ReleaseQubit runs nothing, the
engine's qubit count only grows until the last release resets.
"""
import numpy as np
import pytest

import libsvsim
from oracle import oracle as O

P = libsvsim.PauliId
PAULIS = [P.PauliId_I, P.PauliId_X, P.PauliId_Y, P.PauliId_Z]
TOL = 1e-12


class OracleDmBackend(libsvsim.QirBackend):
    """The density-matrix oracle on the first 2^(2n) entries of a state of
    2^cap amplitudes, as the device engine holds rho."""

    def __init__(self, cap):
        super().__init__()
        self.cap = cap
        self.reset()
        self.calls = []

    def run(self, n, gates):
        arr = O.as_qir_gates([(g.op, g.qubit, g.mask, g.theta) for g in gates])
        self.calls.append((n, arr))
        d = 1 << (2 * n)
        re, im, p, o = O.dm_qir_apply(n, arr, state=(self.re[:d], self.im[:d]), nthreads=1)
        self.re[:d], self.im[:d] = re, im
        if len(p):
            self.last = float(p[-1])
        return list(p), [int(x) for x in o], self.last

    def reset(self):
        self.re, self.im = O.zero_state(self.cap)
        self.last = 0.0

    def capacity(self):
        return self.cap

    def rho(self, n):
        d = 1 << (2 * n)
        v = self.re[:d] + 1j * self.im[:d]
        return v.reshape(1 << n, 1 << n).T     # [row, col], index row | col << n


class OracleSvBackend(libsvsim.QirBackend):
    def __init__(self, cap):
        super().__init__()
        self.cap = cap
        self.reset()
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


def random_calls(sim, qs, rng, count, measure=False):
    """Random driver calls over the whole gate set on the qubits qs."""
    one = ["X", "Y", "Z", "H", "S", "T", "AdjointS", "AdjointT"]
    ctl = ["Controlled" + g for g in one]
    n = len(qs)
    for _ in range(count):
        kind = rng.integers(0, 7 if measure else 6)
        t = float(rng.uniform(-3, 3))
        perm = [qs[i] for i in rng.permutation(n)]
        if kind == 0:
            getattr(sim, one[rng.integers(len(one))])(perm[0])
        elif kind == 1 and n > 1:
            nc = int(rng.integers(1, n))
            getattr(sim, ctl[rng.integers(len(ctl))])(perm[1:1 + nc], perm[0])
        elif kind == 2:
            sim.R(PAULIS[rng.integers(4)], perm[0], t)
        elif kind == 3 and n > 1:
            sim.ControlledR(perm[1:2], PAULIS[rng.integers(4)], perm[0], t)
        elif kind == 4:
            k = int(rng.integers(1, n + 1))
            sim.Exp([PAULIS[i] for i in rng.integers(0, 4, k)], perm[:k], t)
        elif kind == 5 and n > 1:
            k = int(rng.integers(1, n))
            sim.ControlledExp(perm[k:k + 1], [PAULIS[i] for i in rng.integers(0, 4, k)], perm[:k], t)
        elif kind == 6:
            sim.Measure([PAULIS[rng.integers(4)]], [perm[0]])


def _pair(n):
    be_sv, be_dm = OracleSvBackend(n), OracleDmBackend(2 * n)
    return libsvsim.QirSimulator(backend=be_sv), be_sv, libsvsim.QirNoiseSimulator(backend=be_dm), be_dm


def _recs(gates):
    return [(g.op, g.qubit, g.mask, g.theta) for g in gates]


@pytest.mark.parametrize("n,seed", [(1, 1), (2, 2), (3, 3), (4, 4)])
def test_gate_set_gives_rho_of_the_state_vector_drivers_psi(n, seed):
    sv, be_sv, dm, be_dm = _pair(n)
    qsv = [sv.AllocateQubit() for _ in range(n)]
    qdm = [dm.AllocateQubit() for _ in range(n)]
    assert [dm.slot_of(q) for q in qdm] == list(range(n)) and dm.engine_qubits() == n
    random_calls(sv, qsv, np.random.default_rng(seed), 60)
    random_calls(dm, qdm, np.random.default_rng(seed), 60)
    # the gate set is shared: the same records in the same order
    assert _recs(sv.pending()) == _recs(dm.pending()) and len(dm.pending()) >= 30
    sv.flush()
    dm.flush()
    psi = be_sv.re + 1j * be_sv.im
    assert np.allclose(be_dm.rho(n), np.outer(psi, psi.conj()), atol=TOL)


def test_measure_runs_the_queue_with_the_state_vector_drivers_records():
    n = 3
    sv, be_sv, dm, be_dm = _pair(n)
    qsv = [sv.AllocateQubit() for _ in range(n)]
    qdm = [dm.AllocateQubit() for _ in range(n)]
    for sim in (sv, dm):
        draws = iter(np.random.default_rng(5).random(100))
        sim.set_draw(lambda d=draws: float(next(d)))
    random_calls(sv, qsv, np.random.default_rng(6), 80, measure=True)
    random_calls(dm, qdm, np.random.default_rng(6), 80, measure=True)
    assert len(be_dm.calls) >= 5 and len(be_dm.calls) == len(be_sv.calls)
    for (n1, a1), (n2, a2) in zip(be_sv.calls, be_dm.calls):
        assert n1 == n2 == n and np.array_equal(a1, a2)
    # the first measurement sees a pure state: P(1) is the state vector's
    first = [O.QIR["Measure"] == op for op in be_dm.calls[0][1]["op"]].index(True)
    g = be_dm.calls[0][1][: first + 1]
    _, _, p_dm, _ = O.dm_qir_apply(n, g)
    re, im = O.zero_state(n)
    p_sv, _, _ = O.qir_apply(n, re, im, g)
    assert abs(p_dm[0] - p_sv[0]) < TOL


def test_results_and_draws():
    dm = libsvsim.QirNoiseSimulator(backend=OracleDmBackend(4))
    a, b = dm.AllocateQubit(), dm.AllocateQubit()
    draws = iter([0.3, 0.9, 0.1])
    dm.set_draw(lambda: next(draws))
    dm.H(a)
    dm.ControlledX([a], b)
    r0 = dm.Measure([P.PauliId_Z], [a])            # P(1) = 1/2, draw 0.3 -> One
    assert dm.GetResultValue(r0) == libsvsim.ResultValue.Result_One and abs(dm.last_prob() - 0.5) < TOL
    r1 = dm.Measure([P.PauliId_Z], [b])            # correlated: P(1) = 1
    assert dm.AreEqualResults(r0, r1) and r1 == dm.UseOne() and abs(dm.last_prob() - 1) < TOL
    r2 = dm.Measure([P.PauliId_I], [b])            # I measures in Z: P(1) = 1, draw 0.1 -> One
    assert r2 == dm.UseOne()
    with pytest.raises(Exception):
        dm.Measure([P.PauliId_Z, P.PauliId_Z], [a, b])
    with pytest.raises(Exception):
        dm.QubitToString(a)


def test_release_runs_nothing_and_the_engine_count_only_grows():
    cap = 8
    be = OracleDmBackend(cap)
    dm = libsvsim.QirNoiseSimulator(backend=be)
    qs = [dm.AllocateQubit() for _ in range(3)]
    dm.H(qs[0])
    dm.ControlledX([qs[0]], qs[2])
    dm.Measure([P.PauliId_X], [qs[1]])
    assert [n for n, _ in be.calls] == [3]
    dm.R(P.PauliId_Y, qs[2], 0.7)
    dm.ReleaseQubit(qs[2])                       # runs nothing, keeps the queue
    assert len(be.calls) == 1 and len(dm.pending()) == 1
    assert dm.num_qubits() == 2 and dm.engine_qubits() == 3
    q = dm.AllocateQubit()                       # the n-th live qubit has handle n - 1
    assert q == qs[2] and dm.slot_of(q) == 2
    assert dm.num_qubits() == 3 and dm.engine_qubits() == 4
    before = (be.re.copy(), be.im.copy())
    dm.Measure([P.PauliId_Z], [qs[0]])
    # the queue runs on rho of 4 qubits over the same amplitudes
    assert [n for n, _ in be.calls] == [3, 4]
    re, im, _, _ = O.dm_qir_apply(4, be.calls[1][1], state=(before[0][:256], before[1][:256]))
    assert np.allclose(be.re[:256], re, atol=TOL) and np.allclose(be.im[:256], im, atol=TOL)
    # a gate on a qubit at or above the engine count is rejected
    beyond = (4 + 1) << 4                        # handle of index 4
    with pytest.raises(Exception, match="out of range"):
        dm.X(beyond)
    with pytest.raises(Exception, match="out of range"):
        dm.ControlledX([beyond], qs[0])


def test_last_release_drops_the_queue_and_resets():
    be = OracleDmBackend(6)
    dm = libsvsim.QirNoiseSimulator(backend=be)
    a, b = dm.AllocateQubit(), dm.AllocateQubit()
    dm.H(a)
    dm.Measure([P.PauliId_Z], [a])
    dm.X(b)
    dm.ReleaseQubit(b)
    dm.ReleaseQubit(a)
    assert dm.num_qubits() == 0 and dm.engine_qubits() == 0 and len(dm.pending()) == 0
    assert len(be.calls) == 1 and dm.last_prob() == 0.0
    assert be.re[0] == 1.0 and np.count_nonzero(be.re) == 1 and not np.any(be.im)
    with pytest.raises(Exception):
        dm.ReleaseQubit(a)
    # allocation starts over at handle 0
    assert dm.AllocateQubit() == a and dm.engine_qubits() == 1


def test_capacity_is_half_the_state():
    dm = libsvsim.QirNoiseSimulator(backend=OracleDmBackend(7))
    for _ in range(3):
        dm.AllocateQubit()
    with pytest.raises(Exception, match="more than slots"):
        dm.AllocateQubit()
    assert dm.num_qubits() == 3 and dm.engine_qubits() == 3
    # a release and a new allocation push the engine count past the state:
    # the run is rejected instead of touching memory beyond it
    q0 = (0 + 1) << 4                           # handle of index 0
    dm.ReleaseQubit(q0)
    dm.AllocateQubit()
    dm.H(q0)
    with pytest.raises(Exception, match="exceeds"):
        dm.flush()
