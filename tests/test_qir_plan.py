"""QIR gate set on the host: the planner's segmented program (controlled ops
inside sweeps, measurement steps, collapse factors written at run time),
executed by svmi_debug_emulate_qir with the kernels' own op code, against
the oracle (oracle/qir_oracle.c) and the reference's fixtures.

Tolerance: amplitudes and P(1) within 1e-12 (FP64 with a different
summation order and fused matrices); outcomes equal wherever the draw is not
within 1e-9 of P(1).
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

Q = svmi.QIR_OP
TOL = 1e-12


def _zero(n):
    re = np.zeros(1 << n)
    im = np.zeros(1 << n)
    re[0] = 1.0
    return re, im


def _close(a, b):
    return np.max(np.abs(a - b)) < TOL


def _safe_program(n, ng, nm, seed):
    """Seeded random program whose draws stay 1e-9 away from every P(1)."""
    while True:
        rng = np.random.default_rng(seed)
        g = MG.random_program(n, ng, nm, rng)
        re, im = MG.random_state(n, rng)
        r, i = re.copy(), im.copy()
        p, _, _ = O.qir_apply(n, r, i, g, nthreads=2)
        draws = g["theta"][g["op"] == Q["Measure"]]
        if np.all(np.abs(draws - p) > 1e-9):
            return g, (re, im), (r, i, p)
        seed += 1000


def test_names_match_oracle():
    assert svmi.QIR_OP_NAMES == O.QIR_OPS
    assert svmi.QIR_GATE_DTYPE == O.QIR_GATE_DTYPE


def test_golden_program_every_tile():
    d = np.load(os.path.join(GOLD, "qir_program_n8.npz"))
    for tile in (4, 5, 6, 8):
        re, im, p, o = svmi.emulate_qir(8, d["gates"], _zero(8), tile_qubits=tile, low_qubits=min(4, tile))
        assert _close(re, d["re"]) and _close(im, d["im"]), tile
        assert _close(p, d["prob"]) and np.array_equal(o, d["outcome"]), tile


def test_golden_kats():
    d = np.load(os.path.join(GOLD, "qir_kats.npz"))
    for i, g in enumerate(d["gates"]):
        n = int(d["n"][i])
        dim = 1 << n
        st = (d["in_re"][i][:dim], d["in_im"][i][:dim])
        re, im, p, o = svmi.emulate_qir(n, np.array([g], dtype=svmi.QIR_GATE_DTYPE), st, tile_qubits=4)
        assert _close(re, d["out_re"][i][:dim]) and _close(im, d["out_im"][i][:dim]), (O.QIR_OPS[g["op"]], i)
        # slots above the allocated qubits are the reference's untouched input
        assert np.array_equal(d["out_re"][i][dim:], d["in_re"][i][dim:])
        if g["op"] == Q["Measure"]:
            assert abs(p[0] - d["prob"][i]) < TOL and o[0] == d["outcome"][i]


def test_one_qubit_programs():
    rng = np.random.default_rng(8)
    for _ in range(20):
        g = [(int(rng.integers(0, 14)), 0, 0, float(rng.uniform(-6, 6))) for _ in range(6)]
        g += [("Measure", 0, int(rng.integers(3)), float(rng.random()))] + g[:3]
        st = MG.random_state(1, rng)
        r, i = st[0].copy(), st[1].copy()
        p, o, _ = O.qir_apply(1, r, i, g, nthreads=1)
        if abs(p[0] - g[6][3]) < 1e-9:
            continue
        re, im, ep, eo = svmi.emulate_qir(1, g, st)
        assert _close(re, r) and _close(im, i) and _close(ep, p) and np.array_equal(eo, o)


@pytest.mark.parametrize("n,tile", [(2, 4), (3, 4), (5, 4), (7, 5), (10, 6), (12, 8), (13, 12)])
def test_random_programs_vs_oracle(n, tile):
    g, st, (r, i, p) = _safe_program(n, 250, 12, seed=100 + n)
    re, im, ep, eo = svmi.emulate_qir(n, g, st, tile_qubits=tile, low_qubits=min(4, tile))
    assert _close(re, r) and _close(im, i)
    assert _close(ep, p)
    assert np.array_equal(eo, (g["theta"][g["op"] == Q["Measure"]] < p).astype(np.int32))


def test_many_controls_need_no_tile_bits():
    # a 15-control X and Z on 16 qubits join the last sweep of the H layer:
    # their controls outside the tile cost no tile bits
    n = 16
    ctrl = ((1 << n) - 1) & ~1
    gates = [("H", q, 0, 0.0) for q in range(n)] + [("ControlledX", 0, ctrl, 0.0), ("ControlledZ", 0, ctrl, 0.0)]
    sweeps, nm, desc = svmi.plan_qir(n, gates, tile_qubits=12)
    assert nm == 0 and sweeps == 2, desc
    assert "CD" in desc.splitlines()[-1], desc
    rng = np.random.default_rng(3)
    st = MG.random_state(n, rng)
    re, im, _, _ = svmi.emulate_qir(n, gates, st, tile_qubits=6)
    r, i = st[0].copy(), st[1].copy()
    O.qir_apply(n, r, i, gates, nthreads=4)
    assert _close(re, r) and _close(im, i)


def test_measurement_is_a_barrier():
    # ops after a measurement never move before it, even on other qubits
    gates = [("H", 0), ("ControlledX", 1, 1), ("Measure", 0, 0, 0.3), ("H", 1), ("Measure", 1, 1, 0.6),
             ("X", 2)]
    sweeps, nm, desc = svmi.plan_qir(4, gates, tile_qubits=4)
    assert nm == 2
    lines = [x.strip() for x in desc.splitlines()]
    kinds = [x.split()[0] for x in lines[1:] if x.startswith(("sweep", "measure"))]
    assert kinds == ["sweep", "measure", "sweep", "measure", "sweep"], desc


def test_measure_outcomes_and_collapse():
    # Bell pair: P(1) = 1/2; after the first outcome the second qubit agrees
    for draw, want in ((0.25, 1), (0.75, 0)):
        gates = [("H", 0), ("ControlledX", 1, 1), ("Measure", 0, 0, draw), ("Measure", 1, 0, 0.5)]
        re, im, p, o = svmi.emulate_qir(2, gates, _zero(2))
        assert abs(p[0] - 0.5) < 1e-15 and o[0] == want
        assert abs(p[1] - want) < 1e-15 and o[1] == want
        assert abs(re[3 if want else 0] - 1.0) < 1e-15


def test_rejects_bad_records():
    for g, status in [((99, 0, 0, 0.0), "ERR_OP"), (("X", 4, 0, 0.0), "ERR_QUBIT"), (("Swap", 1, 1, 0.0), "ERR_QUBIT"),
                      (("Swap", 1, 9, 0.0), "ERR_QUBIT")]:
        with pytest.raises(svmi.SvmiError) as e:
            svmi.emulate_qir(4, [g], _zero(4))
        assert status in str(e.value)
