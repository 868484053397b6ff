"""Pins the CPU oracle (oracle/svsim_oracle.c) to the reference engine.

Checks, in order of strength:
  * bit-exact agreement with fixtures written by the reference itself
    (tests/golden/make_golden.py via oracle/_ref);
  * bit-exact agreement with the live reference build where oracle/_ref exists;
  * the reference's own 14 known-answer tests
    (svsim/unit_test/unit_test_cpu_omp.cpp:44-302, tolerance 1e-3);
  * closed forms (textbook matrices, QFT = DFT, Bell/GHZ), invariants and the
    measure() interval rule against brute force.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import oracle as O
from svmi import workloads as W

GOLD = os.path.join(os.path.dirname(__file__), "golden")
PI = np.pi


def _sha(re, im):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(re, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(im, dtype="<f8").tobytes())
    return h.hexdigest()


def test_gate_kats_bit_exact():
    d = np.load(os.path.join(GOLD, "gate_kats.npz"))
    n = int(d["n"])
    for i, g in enumerate(d["gates"]):
        re, im = O.run(n, np.array([g], dtype=O.GATE_DTYPE), (d["init_re"][i], d["init_im"][i]), nthreads=2)
        assert np.array_equal(re, d["out_re"][i]) and np.array_equal(im, d["out_im"][i]), f"op {g['op']} case {i}"


def test_random_layers_bit_exact_and_measure():
    d = np.load(os.path.join(GOLD, "random_layers_n12.npz"))
    n = int(d["n"])
    re, im = O.run(n, d["gates"], nthreads=3)
    assert np.array_equal(re, d["re"]) and np.array_equal(im, d["im"])
    u = d["draws"].astype(np.float64) / float(O.RAND_MAX)
    assert np.array_equal(O.measure(n, re, im, u), d["shots"])


def test_hashes_bit_exact():
    with open(os.path.join(GOLD, "hashes.json")) as f:
        want = json.load(f)
    for n in (10, 14):
        assert _sha(*O.run(n, W.u3_brickwall(n=n, layers=20))) == want[f"u3_brickwall_n{n}"]
        st = W.random_state(n, seed=3000 + n)
        assert _sha(*O.run(n, W.qft(n=n), st)) == want[f"qft_n{n}_random_input"]
    for n, L in ((12, 8), (16, 4)):
        g, _, _, _ = W.hea_ising(n=n, layers=L)
        assert _sha(*O.run(n, g)) == want[f"hea_n{n}_l{L}"]


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference sources)")
def test_live_reference_all_ops():
    rng = np.random.default_rng(11)
    for n in (3, 6, 9):
        gates = []
        for _ in range(3):
            for op in range(38):
                k = O.lib().orc_gate_arity(op)
                if k > n:
                    continue
                gates.append((op, list(rng.choice(n, size=k, replace=False)), *(rng.random(3) * 6 - 3)))
        st = W.random_state(n, seed=n)
        a = O.run(n, gates, st, nthreads=4)
        b = O.ref_run(n, gates, st, nthreads=2)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        k, u = W.draws(500, seed=n)
        assert np.array_equal(O.measure(n, *a, u), O.ref_measure(n, *b, k.astype(np.int32)))


# reference unit tests: svsim/unit_test/unit_test_cpu_omp.cpp:44-302
KATS = [
    ("X", 1, [("X", [0])], [0., 1.], [0., 0.]),
    ("Y", 1, [("Y", [0])], [0., 0.], [0., 1.]),
    ("Z", 1, [("H", [0]), ("Z", [0])], [0.707, -0.707], [0., 0.]),
    ("H", 1, [("H", [0])], [0.707, 0.707], [0., 0.]),
    ("S", 1, [("H", [0]), ("S", [0])], [0.707, 0.], [0., 0.707]),
    ("SDG", 1, [("H", [0]), ("SDG", [0])], [0.707, 0.], [0., -0.707]),
    ("T", 1, [("H", [0]), ("T", [0])], [0.707, 0.5], [0., 0.5]),
    ("TDG", 1, [("H", [0]), ("TDG", [0])], [0.707, 0.5], [0., -0.5]),
    ("RX", 1, [("RX", [0], PI / 3.)], [0.866, 0.], [0., -0.5]),
    ("RY", 1, [("RY", [0], PI / 6.)], [0.966, 0.259], [0., 0.]),
    ("U1", 1, [("H", [0]), ("U1", [0], 0, 0, PI / 5.)], [0.707, 0.572], [0., 0.416]),
    ("U2", 1, [("H", [0]), ("U2", [0], 0, PI / 4., PI / 5.)], [0.095, 0.432], [-0.294, 0.847]),
    ("U3", 1, [("H", [0]), ("U3", [0], PI / 4., PI / 5., PI / 6.)], [0.419, 0.485], [-0.135, 0.756]),
    ("CX", 2, [("H", [0]), ("CX", [0, 1])], [0.707, 0, 0, 0.707], [0, 0, 0, 0]),
]


@pytest.mark.parametrize("name,n,seq,er,ei", KATS, ids=[k[0] for k in KATS])
def test_reference_unit_kats(name, n, seq, er, ei):
    from svmi import OP
    gates = [(OP[s[0]], s[1], *(list(s[2:]) + [0, 0, 0])[:3]) for s in seq]
    re, im = O.run(n, gates)
    assert np.abs(re - er).max() <= 1e-3 and np.abs(im - ei).max() <= 1e-3


def _mat1(name, th=0.0, ph=0.0, la=0.0):
    c, s = np.cos, np.sin
    r2 = 1 / np.sqrt(2)
    return {
        "X": np.array([[0, 1], [1, 0]]), "Y": np.array([[0, -1j], [1j, 0]]), "Z": np.diag([1, -1]),
        "H": r2 * np.array([[1, 1], [1, -1]]), "S": np.diag([1, 1j]), "SDG": np.diag([1, -1j]),
        "T": np.diag([1, np.exp(1j * PI / 4)]), "TDG": np.diag([1, np.exp(-1j * PI / 4)]),
        "RX": np.array([[c(th / 2), -1j * s(th / 2)], [-1j * s(th / 2), c(th / 2)]]),
        "RY": np.array([[c(th / 2), -s(th / 2)], [s(th / 2), c(th / 2)]]),
        "RZ": np.diag([1, np.exp(1j * ph)]),      # reference RZ == U1 (cpu_omp:2065-2069)
        "U1": np.diag([1, np.exp(1j * la)]),
        "U2": r2 * np.array([[1, -np.exp(1j * la)], [np.exp(1j * ph), np.exp(1j * (ph + la))]]),
        "U3": np.array([[c(th / 2), -np.exp(1j * la) * s(th / 2)],
                        [np.exp(1j * ph) * s(th / 2), np.exp(1j * (ph + la)) * c(th / 2)]]),
        "W": r2 * np.array([[1, -1j], [-1j, 1]]), "ID": np.eye(2),
    }[name]


def _apply1(M, q, re, im):
    n = int(np.log2(re.size))
    psi = (re + 1j * im).reshape([2] * n)
    ax = n - 1 - q
    psi = np.moveaxis(np.tensordot(M, np.moveaxis(psi, ax, 0), axes=1), 0, ax)
    psi = psi.reshape(-1)
    return psi.real, psi.imag


@pytest.mark.parametrize("name", ["X", "Y", "Z", "H", "S", "SDG", "T", "TDG", "RX", "RY", "RZ", "U1", "U2",
                                  "U3", "W", "ID"])
def test_closed_form_single_qubit(name):
    from svmi import OP
    n = 4
    rng = np.random.default_rng(3)
    for q in range(n):
        th, ph, la = rng.random(3) * 6 - 3
        st = W.random_state(n, seed=q + 10)
        got = O.run(n, [(OP[name], [q], th, ph, la)], st)
        want = _apply1(_mat1(name, th, ph, la), q, *st)
        assert np.abs(got[0] - want[0]).max() < 1e-14 and np.abs(got[1] - want[1]).max() < 1e-14


def test_closed_form_two_qubit_diag_and_perm():
    from svmi import OP
    n = 3
    rng = np.random.default_rng(4)
    idx = np.arange(1 << n)
    for a, b in [(0, 1), (2, 0), (1, 2)]:
        st = W.random_state(n, seed=a * 3 + b)
        psi = st[0] + 1j * st[1]
        ba, bb = (idx >> a) & 1, (idx >> b) & 1
        lam = rng.random() * 6
        cases = {
            "CZ": psi * np.where(ba & bb, -1, 1),
            "CU1": psi * np.where(ba & bb, np.exp(1j * lam), 1),
            "CX": psi[np.where(ba == 1, idx ^ (1 << b), idx)],
            "SWAP": psi[np.where(ba != bb, idx ^ (1 << a) ^ (1 << b), idx)],
            "RZZ": psi * np.exp(1j * lam * (ba ^ bb)),    # CX.U1.CX (cpu_omp:2227-2233)
        }
        for name, want in cases.items():
            got = O.run(n, [(OP[name], [a, b], lam, 0, lam)], st)
            assert np.abs(got[0] + 1j * got[1] - want).max() < 1e-14, name


def test_qft_is_dft():
    # the circuit maps |x> to sum_y exp(2 pi i x y / N) |y> / sqrt(N) where x
    # and y read qubit 0 as the most significant bit: R . ifft . R * sqrt(N)
    # with R the bit reversal of the index
    n = 8
    st = W.random_state(n, seed=5)
    got = O.run(n, W.qft(n=n), st)
    psi = st[0] + 1j * st[1]
    R = np.array([int(format(x, f"0{n}b")[::-1], 2) for x in range(1 << n)])
    want = (np.fft.ifft(psi[R]) * np.sqrt(1 << n))[R]
    assert np.abs(got[0] + 1j * got[1] - want).max() < 1e-13


def test_measure_rule_matches_linear_search_and_edges():
    n = 6
    re, im = O.run(n, W.random_layers(n=n, layers=5))
    _, u = W.draws(3000, seed=1)
    u = np.concatenate([u, [0.0, 1.0, 1.0 - 1e-17]])
    assert np.array_equal(O.measure(n, re, im, u), O.measure(n, re, im, u, linear=True))
    # r beyond the total probability -> 0 (no interval matches)
    scan = O.prob_scan(n, re, im)
    assert O.measure(n, re, im, [scan[-1]])[0] == 0
    # zero-probability basis states are never returned
    z = np.zeros(1 << n)
    z[5] = 1.0
    assert set(O.measure(n, z, np.zeros(1 << n), u[:200]).tolist()) <= {5, 0}


def test_norm_and_inverse_round_trip():
    from svmi import OP
    n = 9
    g = W.u3_brickwall(n=n, layers=6)
    st = W.random_state(n, seed=9)
    re, im = O.run(n, g, st)
    assert abs(np.sum(re * re + im * im) - 1) < 1e-12
    inv = []
    for row in g[::-1]:
        if row["op"] == OP["U3"]:
            inv.append((OP["U3"], [row["qb"][0]], -row["theta"], -row["lambda"], -row["phi"]))
        else:
            inv.append((row["op"], list(row["qb"][:2]), 0, 0, 0))
    back = O.run(n, inv, (re, im))
    assert np.abs(back[0] - st[0]).max() < 1e-12 and np.abs(back[1] - st[1]).max() < 1e-12


def test_expval_and_pauli_sum():
    n = 5
    st = W.random_state(n, seed=2)
    psi = st[0] + 1j * st[1]
    X = np.array([[0, 1], [1, 0]]); Y = np.array([[0, -1j], [1j, 0]]); Z = np.diag([1, -1]); I2 = np.eye(2)
    terms = [(0.3, 0b00101, 0b00100), (-1.1, 0, 0b11000), (0.7, 0b10010, 0b10000), (0.25, 0b01000, 0b01001)]
    H = np.zeros((32, 32), complex)
    for c, x, z in terms:
        P = np.array([[1]])
        for q in reversed(range(n)):
            xb, zb = (x >> q) & 1, (z >> q) & 1
            P = np.kron(P, Y if xb and zb else X if xb else Z if zb else I2)
        H += c * P
    want = np.vdot(psi, H @ psi).real
    assert abs(O.expval(n, *st, terms) - want) < 1e-13
    hre, him = O.apply_pauli_sum(n, *st, terms)
    assert np.abs(hre + 1j * him - H @ psi).max() < 1e-13
