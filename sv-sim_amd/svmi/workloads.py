"""Seeded synthetic workloads: the reference-gate circuits the engine is
measured and checked on.

All circuits are GATE_DTYPE arrays in the reference slot layout
(svsim/src/svsim_cpu_omp.hpp:377-564: RZ's angle in `phi`, CU1's in
`lambda`, ...).  Angles come from numpy's PCG64 generator, so the same seed
gives the same circuit on every machine.

  random_layers   12-qubit random circuit, 40 layers of {H,RX,RY,RZ} on every
                  qubit + CX on a random perfect pairing (+1024 shots)
  u3_brickwall    20 layers of U3 on every qubit + brick-wall CZ (30 qubits)
  qft             H + CP(pi/2^k) on all pairs (i, i+k) + reversal SWAPs (33 q)
  hea_ising       hardware-efficient ansatz RY+RZ layers + CX ladder, Ising
                  observable sum J_i Z_i Z_{i+1} + sum h_i X_i (20 q, 8 layers)
  batch_ising     512 ansatz circuits, n ~ U{10..22}, layers ~ U{2..10}
  qir_mixed       QIR gate-set program (svsim_qsharp engine records): single-
                  qubit gates, 1-3-control gates, (n-1)-control X/Z, Swap and
                  mid-circuit measurements in the Z, X and Y bases (30 q)
"""
import numpy as np

from . import GATE_DTYPE, OP, PAULI_DTYPE, QIR_GATE_DTYPE, QIR_OP

RAND_MAX = 2147483647  # glibc; the reference draws rand()/RAND_MAX

SEEDS = {"random_layers": 1001, "u3_brickwall": 1002, "qft": 1003, "hea_ising": 1004, "batch_ising": 1005,
         "qir_mixed": 1006}


class Circuit:
    def __init__(self):
        self.rows = []

    def add(self, name, qubits, theta=0.0, phi=0.0, lam=0.0):
        self.rows.append((OP[name], tuple(int(q) for q in qubits), float(theta), float(phi), float(lam)))

    def array(self):
        arr = np.zeros(len(self.rows), dtype=GATE_DTYPE)
        for i, (op, qs, t, p, l) in enumerate(self.rows):
            arr[i]["op"] = op
            arr[i]["qb"][: len(qs)] = qs
            arr[i]["theta"], arr[i]["phi"], arr[i]["lambda"] = t, p, l
        return arr


def random_layers(n=12, layers=40, seed=SEEDS["random_layers"]):
    """Each layer: one of {H, RX, RY, RZ} per qubit, angle U[0, 2pi); then CX
    on a seeded random perfect pairing (an odd qubit count leaves one out)."""
    rng = np.random.default_rng(seed)
    c = Circuit()
    for _ in range(layers):
        kinds = rng.integers(0, 4, size=n)
        angles = rng.random(n) * 2 * np.pi
        for q in range(n):
            k = int(kinds[q])
            if k == 0:
                c.add("H", [q])
            elif k == 1:
                c.add("RX", [q], theta=angles[q])
            elif k == 2:
                c.add("RY", [q], theta=angles[q])
            else:
                c.add("RZ", [q], phi=angles[q])
        perm = rng.permutation(n)
        for i in range(0, n - 1, 2):
            c.add("CX", [perm[i], perm[i + 1]])
    return c.array()


def u3_brickwall(n=30, layers=20, seed=SEEDS["u3_brickwall"]):
    """Layer l: U3(theta, phi, lambda ~ U[0, 2pi)) on every qubit, then CZ on
    (i, i+1) for i = l mod 2, l mod 2 + 2, ..."""
    rng = np.random.default_rng(seed)
    c = Circuit()
    for l in range(layers):
        ang = rng.random((n, 3)) * 2 * np.pi
        for q in range(n):
            c.add("U3", [q], theta=ang[q, 0], phi=ang[q, 1], lam=ang[q, 2])
        for i in range(l % 2, n - 1, 2):
            c.add("CZ", [i, i + 1])
    return c.array()


def qft(n=33):
    """H(i), then CP(pi/2^k) = CU1 on (i+k, i) for k = 1..n-1-i, for each i;
    then SWAP(i, n-1-i) for i < n/2."""
    c = Circuit()
    for i in range(n):
        c.add("H", [i])
        for k in range(1, n - i):
            c.add("CU1", [i + k, i], lam=np.pi / 2 ** k)
    for i in range(n // 2):
        c.add("SWAP", [i, n - 1 - i])
    return c.array()


def random_state(n, seed):
    """Seeded N(0,1) + i N(0,1) vector, normalised (re, im)."""
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(1 << n)
    im = rng.standard_normal(1 << n)
    s = np.sqrt(np.sum(re * re) + np.sum(im * im))
    return re / s, im / s


def hea(n, layers, rng):
    """Ansatz: per layer RY(theta) on every qubit, RZ(theta) on every qubit,
    then CX(0,1), CX(1,2), ..., CX(n-2, n-1).  theta ~ U[-pi, pi).
    Returns (gates, param_slots[g, 3], n_params)."""
    c = Circuit()
    slots = []
    p = 0
    for _ in range(layers):
        th = rng.random(2 * n) * 2 * np.pi - np.pi
        for q in range(n):
            c.add("RY", [q], theta=th[q])
            slots.append((p, -1, -1))
            p += 1
        for q in range(n):
            c.add("RZ", [q], phi=th[n + q])
            slots.append((-1, p, -1))
            p += 1
        for q in range(n - 1):
            c.add("CX", [q, q + 1])
            slots.append((-1, -1, -1))
    return c.array(), np.array(slots, dtype=np.int32).reshape(-1, 3), p


def ising_terms(n, rng):
    """sum_i J_i Z_i Z_{i+1} + sum_i h_i X_i, J, h ~ N(0, 1)."""
    J = rng.standard_normal(n - 1)
    h = rng.standard_normal(n)
    t = np.zeros(2 * n - 1, dtype=PAULI_DTYPE)
    for i in range(n - 1):
        t[i] = (J[i], 0, (1 << i) | (1 << (i + 1)))
    for i in range(n):
        t[n - 1 + i] = (h[i], 1 << i, 0)
    return t


def hea_ising(n=20, layers=8, seed=SEEDS["hea_ising"]):
    rng = np.random.default_rng(seed)
    gates, slots, p = hea(n, layers, rng)
    terms = ising_terms(n, rng)
    return gates, slots, p, terms


def batch_ising(count=512, nmin=10, nmax=22, lmin=2, lmax=10, seed=SEEDS["batch_ising"]):
    """List of dicts {n, gates, param_slots, n_params, terms}."""
    rng = np.random.default_rng(seed)
    items = []
    for _ in range(count):
        n = int(rng.integers(nmin, nmax + 1))
        L = int(rng.integers(lmin, lmax + 1))
        sub = np.random.default_rng(int(rng.integers(0, 2 ** 63)))
        gates, slots, p = hea(n, L, sub)
        items.append({"n": n, "layers": L, "gates": gates, "param_slots": slots, "n_params": p,
                      "terms": ising_terms(n, sub)})
    return items


def draws(shots, seed):
    """Integer draws k in [0, RAND_MAX] and the doubles k / RAND_MAX the
    reference's measure() computes from rand() (svsim_cpu_omp.hpp:351)."""
    rng = np.random.default_rng(seed)
    k = rng.integers(0, RAND_MAX, size=shots, endpoint=True).astype(np.int64)
    return k, k.astype(np.float64) / float(RAND_MAX)


def qir_mixed(n=30, n_gates=1200, n_measures=16, seed=SEEDS["qir_mixed"]):
    """QIR gate-set program (QIR_GATE_DTYPE records): 60% uncontrolled
    single-qubit gates of every kind, 30% gates with 1-3 random controls, 5%
    X or Z with all n-1 other qubits as controls (Grover-style), 5% Swap; a
    measurement every n_gates/n_measures records, bases cycling Z, X, Y, with
    draws k/RAND_MAX."""
    rng = np.random.default_rng(seed)
    single = [QIR_OP[x] for x in ("X", "Y", "Z", "H", "S", "T", "RI", "RX", "RY", "RZ", "EI", "EX", "EY", "EZ",
                                  "AdjointS", "AdjointT")]
    ctl = [QIR_OP["Controlled" + x] for x in ("X", "Y", "Z", "H", "S", "T", "RI", "RX", "RY", "RZ", "EI", "EX",
                                              "EY", "EZ", "AdjointS", "AdjointT")]
    every = max(1, n_gates // max(n_measures, 1))
    rows = []
    m = 0
    for i in range(n_gates):
        if n_measures and (i + 1) % every == 0 and m < n_measures:
            k = int(rng.integers(0, RAND_MAX, endpoint=True))
            rows.append((QIR_OP["Measure"], int(rng.integers(n)), m % 3, k / float(RAND_MAX)))
            m += 1
            continue
        u = rng.random()
        q = int(rng.integers(n))
        theta = float(rng.uniform(-2 * np.pi, 2 * np.pi))
        if u < 0.60:
            rows.append((int(rng.choice(single)), q, 0, theta))
        elif u < 0.90:
            others = np.delete(np.arange(n), q)
            mask = 0
            for c in rng.choice(others, size=int(rng.integers(1, 4)), replace=False):
                mask |= 1 << int(c)
            rows.append((int(rng.choice(ctl)), q, mask, theta))
        elif u < 0.95:
            op = QIR_OP["ControlledX"] if rng.random() < 0.5 else QIR_OP["ControlledZ"]
            rows.append((op, q, ((1 << n) - 1) & ~(1 << q), 0.0))
        else:
            rows.append((QIR_OP["Swap"], q, int(rng.choice(np.delete(np.arange(n), q))), 0.0))
    return np.array(rows, dtype=QIR_GATE_DTYPE)
