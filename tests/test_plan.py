"""Host logic of the engine, checked on the CPU.

svmi.emulate runs the engine's lowering + scheduler and then executes the
device program on the host with the same op code the HIP kernels use
(sv-sim_amd/csrc/svmi_exec.hpp), so these tests pin the planner, the gate
lowering and the adjoint recurrences against the oracle without a GPU.
"""
import os

import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import workloads as W


def _all_ops_circuit(n, reps, rng):
    gates = []
    for _ in range(reps):
        for op in range(38):
            k = svmi.ARITY[svmi.OP_NAMES[op]]
            if k > n:
                continue
            gates.append((op, list(rng.choice(n, size=k, replace=False)), *(rng.random(3) * 4 * np.pi - 2 * np.pi)))
    return gates


def _err(a, b):
    return max(np.abs(a[0] - b[0]).max(), np.abs(a[1] - b[1]).max())


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 8, 11, 13, 15])
@pytest.mark.parametrize("fusion,tile,low", [(0, 13, 4), (1, 13, 4), (1, 12, 4), (1, 6, 2), (1, 4, 0), (1, 9, 3)])
def test_emulated_program_matches_oracle(n, fusion, tile, low):
    rng = np.random.default_rng(100 + n)
    gates = _all_ops_circuit(n, 3, rng)
    st = W.random_state(n, seed=n)
    assert _err(svmi.emulate(n, gates, st, fusion=fusion, tile_qubits=tile, low_qubits=low), O.run(n, gates, st)) < 1e-12


def test_sweep_budget_splits_but_keeps_result():
    n = 12
    g = W.u3_brickwall(n=n, layers=6)
    st = W.random_state(n, seed=1)
    p0, _ = svmi.plan(n, g)
    p1, _ = svmi.plan(n, g, sweep_budget=64)
    assert p1 > p0
    assert _err(svmi.emulate(n, g, st, sweep_budget=64), O.run(n, g, st)) < 1e-12


def test_pass_counts_of_the_measured_configs():
    # one pass per op without fusion; tiled sweeps collapse the circuits
    assert svmi.plan(12, W.random_layers())[0] == 1
    assert svmi.plan(30, W.u3_brickwall(), fusion=False)[0] == 890
    assert svmi.plan(30, W.u3_brickwall())[0] <= 30
    assert svmi.plan(33, W.qft())[0] <= 10
    g, _, _, _ = W.hea_ising()
    assert svmi.plan(20, g)[0] <= 20


def test_srn_is_an_ordering_barrier():
    # the reference SRN is real-linear only: phases on other qubits must not
    # move across it
    n = 4
    gates = [(svmi.OP["T"], [2]), (svmi.OP["SRN"], [0]), (svmi.OP["S"], [0]), (svmi.OP["RX"], [2], 0.3)]
    st = W.random_state(n, seed=3)
    assert _err(svmi.emulate(n, gates, st), O.run(n, gates, st)) < 1e-14


@pytest.mark.parametrize("n,layers,tile", [(4, 2, 12), (7, 3, 5), (10, 2, 11), (13, 2, 8)])
def test_adjoint_matches_parameter_shift(n, layers, tile):
    rng = np.random.default_rng(n)
    gates, slots, P = W.hea(n, layers, rng)
    terms = W.ising_terms(n, rng)
    st = W.random_state(n, seed=n + 1)
    E, g = svmi.emulate_adjoint(n, gates, terms, slots, P, st, tile_qubits=tile)
    assert abs(E - O.expval(n, *O.run(n, gates, st), terms)) < 1e-12
    assert np.abs(g - O.param_shift_grad(n, gates, slots, P, terms, state=st)).max() < 1e-12


def _rotations_and_permutations(n, layers, rng):
    """Per layer a rotation with a gradient slot on every qubit (RX/RY theta,
    RZ phi), then n two-qubit gates on random pairs: CX (either orientation),
    SWAP and CZ."""
    gates, slots, p = [], [], 0
    for _ in range(layers):
        for q in range(n):
            name = ("RX", "RY", "RZ")[int(rng.integers(3))]
            ang = float(rng.random() * 6 - 3)
            if name == "RZ":
                gates.append((svmi.OP[name], [q], 0.0, ang, 0.0))
                slots.append((-1, p, -1))
            else:
                gates.append((svmi.OP[name], [q], ang, 0.0, 0.0))
                slots.append((p, -1, -1))
            p += 1
        for _ in range(n):
            a, b = (int(x) for x in rng.choice(n, size=2, replace=False))
            gates.append((svmi.OP[("CX", "CX", "SWAP", "CZ")[int(rng.integers(4))]], [a, b]))
            slots.append((-1, -1, -1))
    return gates, np.array(slots, dtype=np.int32), p


@pytest.mark.parametrize("n,tile", [(9, 6), (12, 8), (13, 11)])
def test_tile_map_adjoint_matches_parameter_shift(n, tile):
    # CX and SWAP gates on tile qubits outside a phase's registers relabel the
    # LDS indices of the two-state tile ([m] phases, the store through the
    # final map); the gradient must still be the parameter-shift one
    rng = np.random.default_rng(100 + n)
    gates, slots, P = _rotations_and_permutations(n, 3, rng)
    terms = W.ising_terms(n, rng)
    st = W.random_state(n, seed=n)
    _, desc = svmi.plan_adjoint(n, gates, slots, P, tile_qubits=tile)
    assert "[m]" in desc
    E, g = svmi.emulate_adjoint(n, gates, terms, slots, P, st, tile_qubits=tile)
    assert abs(E - O.expval(n, *O.run(n, gates, st), terms)) < 1e-12
    assert np.abs(g - O.param_shift_grad(n, gates, slots, P, terms, state=st)).max() < 1e-12


@pytest.mark.parametrize("n,tile,low", [(12, 12, 4), (14, 8, 2), (16, 6, 2), (17, 12, 4)])
def test_one_state_tile_map_matches_oracle(n, tile, low):
    # forward sweeps fold CX and SWAP gates into the tile map as well; the
    # store writes the tile back through the final map
    gates = W.random_layers(n=n, layers=8)
    rng = np.random.default_rng(n)
    swaps = [(svmi.OP["SWAP"], [int(a), int(b)]) for a, b in rng.integers(0, n, size=(6, 2)) if a != b]
    gates = np.concatenate([svmi.as_gates(gates), svmi.as_gates(swaps)])
    _, desc = svmi.plan(n, gates, tile_qubits=tile, low_qubits=low)
    assert "[m]" in desc
    st = W.random_state(n, seed=n + 7)
    assert _err(svmi.emulate(n, gates, st, tile_qubits=tile, low_qubits=low), O.run(n, gates, st)) < 1e-12


def test_tile_map_across_coefficient_budget_parts():
    # a sweep cut into parts by the coefficient budget reloads its tile in the
    # plain layout at every part, so each part's maps are rebased (HEA-21: the
    # first tile set is cut in two, both parts mapped)
    n = 21
    gates = W.hea_ising(n=n, layers=8)[0]
    _, desc = svmi.plan(n, gates)
    sweeps = [l for l in desc.splitlines() if l.strip().startswith("sweep")]
    assert sweeps[0].split("ext=")[1] == sweeps[1].split("ext=")[1]
    assert "[m]" in sweeps[0] and "[m]" in sweeps[1]
    st = W.random_state(n, seed=5)
    assert _err(svmi.emulate(n, gates, st), O.run(n, gates, st)) < 1e-12


def test_tile_map_folds_the_hea_ladder():
    # the HEA backward program's CX ladder inside a tile costs no phase
    gates, slots, P, _ = W.hea_ising(n=20, layers=8)
    _, desc = svmi.plan_adjoint(20, gates, slots, P)
    assert "[m]" in desc
    phases = int(desc.split("phases=")[1].split()[0])
    assert phases <= 70, phases


def test_adjoint_all_parametric_gates_vs_fd():
    n = 5
    rng = np.random.default_rng(9)
    gates, slots, p = [], [], 0
    for name in ["U3", "U2", "U1", "RX", "RY", "RZ", "CRX", "CRY", "CRZ", "CU1", "CU3", "RXX", "RZZ", "RYY",
                 "H", "CX", "CCX", "SWAP", "C4X"]:
        k = svmi.ARITY[name]
        gates.append((svmi.OP[name], list(rng.choice(n, size=k, replace=False)), *(rng.random(3) * 6 - 3)))
        slots.append([p, p + 1, p + 2])
        p += 3
    slots = np.array(slots)
    terms = [(0.7, 0b00011, 0b00010), (-1.3, 0, 0b10100), (0.4, 0b11000, 0b01000), (0.9, 0b00100, 0)]
    st = W.random_state(n, seed=4)
    E, g = svmi.emulate_adjoint(n, gates, terms, slots, p, st)
    G = O.as_gates(gates)
    fd = np.zeros(p)
    names = ("theta", "phi", "lambda")
    for gi in range(len(G)):
        for s in range(3):
            vals = []
            for h in (1e-5, -1e-5, 2e-5, -2e-5):
                G2 = G.copy()
                G2[gi][names[s]] += h
                vals.append(O.expval(n, *O.run(n, G2, st), terms))
            fd[slots[gi, s]] = (8 * (vals[0] - vals[1]) - (vals[2] - vals[3])) / 12e-5
    assert abs(E - O.expval(n, *O.run(n, G, st), terms)) < 1e-13
    assert np.abs(g - fd).max() < 1e-9


def test_adjoint_rejects_non_unitary():
    n = 3
    gates = [(svmi.OP["R"], [0], 0.5)]
    with pytest.raises(svmi.SvmiError) as e:
        svmi.emulate_adjoint(n, gates, [(1.0, 0, 1)], [[0, -1, -1]], 1, W.random_state(n, 1))
    assert e.value.status == 7


def test_invalid_gates_are_errors_not_aborts():
    st = W.random_state(3, 1)
    for bad, status in [([(svmi.OP["CX"], [1, 1])], 2), ([(svmi.OP["H"], [3])], 2), ([(99, [0])], 3)]:
        with pytest.raises(svmi.SvmiError) as e:
            svmi.emulate(3, bad, st)
        assert e.value.status == status


def test_batch_queue_emulation_matches_parameter_shift():
    # the batched task queue (interleaved circuits, rebased programs,
    # dependency chain, gradient reduction) executed on the host level by level,
    # tasks of a level in reverse order
    items = W.batch_ising(count=14, nmin=1, nmax=13, lmin=1, lmax=4, seed=21)
    ev, grads = svmi.emulate_batch(items)
    for it, e, g in zip(items, ev, grads):
        n = it["n"]
        assert abs(e - O.expval(n, *O.run(n, it["gates"]), it["terms"])) < 1e-12
        gp = O.param_shift_grad(n, it["gates"], it["param_slots"], it["n_params"], it["terms"])
        assert np.abs(g - gp).max() < 1e-12


def test_batch_rejects_five_qubit_gates():
    item = {"n": 6, "gates": [(svmi.OP["C4X"], [0, 1, 2, 3, 4])], "param_slots": [[-1, -1, -1]], "n_params": 0,
            "terms": [(1.0, 0, 1)]}
    with pytest.raises(svmi.SvmiError) as e:
        svmi.emulate_batch([item])
    assert e.value.status == 7


@pytest.mark.parametrize("n,tile,low", [(16, 6, 2), (18, 8, 4), (17, 12, 4)])
def test_phase_factors_match_oracle(n, tile, low):
    # QFT-pattern circuits put many controlled phases on qubits outside the
    # tile; their product is formed once per tile from nibble tables over the
    # external bits (PT..x.. ops), the ones on group bits from tables over the
    # group index -- check the plan really uses that and that the result is
    # unchanged
    gates = W.qft(n=n)
    _, desc = svmi.plan(n, gates, tile_qubits=tile, low_qubits=low)
    assert "x" in "".join(tok for tok in desc.split() if tok.startswith("PT"))
    st = W.random_state(n, seed=n + 40)
    assert _err(svmi.emulate(n, gates, st, tile_qubits=tile, low_qubits=low), O.run(n, gates, st)) < 1e-12


def test_specialised_sweep_source_compiles_without_a_device():
    # hipRTC compiles for gfx950 on the host; tiles above 12 qubits get no kernel
    rng = np.random.default_rng(31)
    gates = []
    for op in range(38):
        k = svmi.ARITY[svmi.OP_NAMES[op]]
        gates.append((op, list(rng.choice(9, size=k, replace=False)), *(rng.random(3) * 8 - 4)))
    src, ms, nbytes = svmi.jit_source(9, gates, tile_qubits=8, low_qubits=2, compile=True)
    assert src.count("__global__") >= 1 and nbytes > 0 and ms > 0
    assert "svmi_jit_0" in src and "#error" not in src
    src13, _, _ = svmi.jit_source(14, gates, tile_qubits=13, compile=False)
    assert "__global__" not in src13


def test_batch_tuning_table_and_overrides():
    # the autotuned table (sv-sim_amd/csrc/svmi_tuning.hpp) covers 4..30 qubits
    # with valid settings; per-item overrides plan differently but give the
    # same expectations and gradients; out-of-range overrides are errors
    import re
    src = open(os.path.join(os.path.dirname(__file__), "..", "sv-sim_amd", "csrc", "svmi_tuning.hpp")).read()
    rows = [tuple(map(int, m)) for m in re.findall(r"\{(\d+), (\d+), (\d+), (\d+)\}", src)]
    assert [r[0] for r in rows] == list(range(4, 31))
    for n, fwd, adj, cost in rows:
        assert 4 <= fwd <= 13 and 4 <= adj <= 11 and 0 <= cost <= 32, (n, fwd, adj, cost)
    items = W.batch_ising(count=4, nmin=6, nmax=11, lmin=1, lmax=3, seed=21)
    e0, g0 = svmi.batch_adjoint(items, _emulate=True)
    e1, g1 = svmi.batch_adjoint([dict(it, fwd_tile=6, adj_tile=7, op_cost=8) for it in items], _emulate=True)
    assert np.abs(e0 - e1).max() < 1e-12 and max(np.abs(a - b).max() for a, b in zip(g0, g1)) < 1e-12
    with pytest.raises(svmi.SvmiError) as e:
        svmi.batch_adjoint([dict(items[0], adj_tile=12)], _emulate=True)
    assert "adj_tile" in str(e.value)


@pytest.mark.parametrize("tile", [6, 8, 12])
def test_register_coordinate_rewrites_match_oracle(tile):
    """Quarter turns kept in register coordinates ([q] phases: three-shear
    rotations as two shears, later tables, lift scales, signs and controlled
    phases rewritten, scatter map XORed) and leading permutations folded into
    gather maps ([g] phases), in circuits whose plans use them, against the
    oracle; the adjoint's folded CX ladders against parameter shift."""
    n = 13
    for gates in (W.u3_brickwall(n, 10), W.random_layers(n, 30)):
        st = W.random_state(n, seed=tile)
        assert _err(svmi.emulate(n, gates, st, tile_qubits=tile), O.run(n, gates, st)) < 1e-12
    assert "[q]" in svmi.plan(n, W.u3_brickwall(n, 10), tile_qubits=tile)[1]
    assert "[g]" in svmi.plan(n, W.random_layers(n, 30), tile_qubits=tile)[1]
    rng = np.random.default_rng(tile)
    gates, slots, P = W.hea(n, 3, rng)
    terms = W.ising_terms(n, rng)
    adj_tile = min(tile, 11)
    assert "[g]" in svmi.plan_adjoint(n, gates, slots, P, tile_qubits=adj_tile)[1]
    st = W.random_state(n, seed=tile + 1)
    E, g = svmi.emulate_adjoint(n, gates, terms, slots, P, st, tile_qubits=adj_tile)
    assert abs(E - O.expval(n, *O.run(n, gates, st), terms)) < 1e-12
    assert np.abs(g - O.param_shift_grad(n, gates, slots, P, terms, state=st)).max() < 1e-12
