"""GPU parity of the tile map of sweeps (SVMI_OPT_TILE_MAP).

With the map on (the default), CX and SWAP gates on two tile qubits that a
phase does not hold in registers are not run: they relabel the LDS indices
of the tile, later phases gather and scatter through the map and the tile's
store writes through the final one.  Adjoint (two-state) sweeps: gradients
and <H> agree with the map off (a different phase schedule, so to rounding,
1e-12) and with the oracle's parameter-shift gradient.  Forward (one-state)
sweeps, which run on k_sweep1's mapped build: states agree with the map off
and with the oracle (1e-10).
"""
import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import workloads as W

pytestmark = pytest.mark.gpu


def _adjoint(n, gates, slots, P, terms, tile_map, state=None):
    with svmi.Engine(n) as e:
        e.set_option(svmi.OPT_TILE_MAP, tile_map)
        e.upload(gates)
        if state is not None:
            e.set_state(*state)
        return e.adjoint(terms, slots, P)


def _rotations_and_permutations(n, layers, rng):
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


@pytest.mark.parametrize("n", [16, 20])
def test_hea_tile_map_on_off_agree(n):
    gates, slots, P, terms = W.hea_ising(n=n, layers=8)
    print("case hea", n, flush=True)
    e1, g1 = _adjoint(n, gates, slots, P, terms, 1)
    e0, g0 = _adjoint(n, gates, slots, P, terms, 0)
    assert abs(e1 - e0) < 1e-12
    assert np.abs(g1 - g0).max() < 1e-12


@pytest.mark.parametrize("n", [14, 17])
def test_permutation_mix_matches_parameter_shift(n):
    rng = np.random.default_rng(300 + n)
    gates, slots, P = _rotations_and_permutations(n, 3, rng)
    terms = W.ising_terms(n, rng)
    st = W.random_state(n, seed=n)
    print("case mix", n, len(gates), "gates", P, "params", flush=True)
    e1, g1 = _adjoint(n, gates, slots, P, terms, 1, st)
    e0, g0 = _adjoint(n, gates, slots, P, terms, 0, st)
    assert abs(e1 - O.expval(n, *O.run(n, gates, st), terms)) < 1e-12
    gp = O.param_shift_grad(n, gates, slots, P, terms, state=st)
    assert np.abs(g1 - gp).max() < 1e-11
    assert np.abs(g1 - g0).max() < 1e-12


def _state(n, gates, tile_map, state=None, tile=None):
    with svmi.Engine(n) as e:
        e.set_option(svmi.OPT_TILE_MAP, tile_map)
        if tile:
            e.set_option(svmi.OPT_TILE_QUBITS, tile)
        e.upload(gates)
        if state is not None:
            e.set_state(*state)
        e.run()
        return e.get_state()


@pytest.mark.parametrize("n,name,tile", [(12, "random_layers", None), (20, "random_layers", None),
                                         (21, "random_layers", 13), (21, "qft", None), (18, "qft", 8),
                                         (24, "random_layers", None), (24, "qft", None), (21, "hea", None)])
def test_forward_tile_map_matches_oracle(n, name, tile):
    # (21, "hea"): its first tile set is cut into two sweeps by the coefficient
    # budget, each reloading the tile in the plain layout with rebased maps
    state = None
    if name == "random_layers":
        gates = W.random_layers(n=n, layers=12)
    elif name == "hea":
        gates = W.hea_ising(n=n, layers=8)[0]
        state = W.random_state(n, seed=n)
    else:
        gates = W.qft(n)
        state = W.random_state(n, seed=n)
    _, desc = svmi.plan(n, gates, tile_qubits=tile or 12)
    assert "[m]" in desc
    print("case", n, name, tile, len(gates), "gates", flush=True)
    re1, im1 = _state(n, gates, 1, state, tile)
    re0, im0 = _state(n, gates, 0, state, tile)
    ore, oim = O.run(n, gates, state)
    assert max(np.abs(re1 - ore).max(), np.abs(im1 - oim).max()) < 1e-10
    assert max(np.abs(re1 - re0).max(), np.abs(im1 - im0).max()) < 1e-12
