"""GPU parity of the double-buffered sweep kernel (LDS DMA prefetch).

With SVMI_OPT_PREFETCH set, one-state sweeps of 12-qubit tiles over states
of at least two resident grids of tiles (21+ qubits on a 256-CU MI355X) run
persistent with two LDS
tile buffers, the next tile streaming in by LDS DMA while the current one's
phases run.  The phases are the same code as the one-tile-per-workgroup
kernel, so the states must be bit-identical with the prefetch on and off;
both are also compared with the oracle (1e-10).  Circuits: the mixed
random-layer circuit (dense, permutation, diagonal and phase ops), the u3
brick-wall circuit (lifting rotations and register tables) and the QFT on a
random state (products of controlled phases and swaps).
"""
import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _random_state(n, seed):
    rng = np.random.default_rng(seed)
    re = rng.standard_normal(1 << n)
    im = rng.standard_normal(1 << n)
    nrm = np.sqrt((re * re + im * im).sum())
    return re / nrm, im / nrm


def _run(n, gates, prefetch, state=None):
    with svmi.Engine(n) as e:
        # without tile maps: sweeps with maps run on the one-tile-per-workgroup
        # kernel whatever the prefetch setting
        e.set_option(svmi.OPT_TILE_MAP, 0)
        e.set_option(svmi.OPT_PREFETCH, prefetch)
        e.upload(gates)
        if state is not None:
            e.set_state(*state)
        e.run()
        return e.get_state()


@pytest.mark.parametrize("n,name", [(21, "random_layers"), (22, "u3_brickwall"), (21, "qft"), (23, "u3_brickwall")])
def test_prefetch_sweeps_bit_identical_and_match_oracle(n, name):
    state = None
    if name == "random_layers":
        gates = W.random_layers(n=n, layers=12)
    elif name == "u3_brickwall":
        gates = W.u3_brickwall(n=n, layers=6)
    else:
        gates = W.qft(n)
        state = _random_state(n, 77)
    print("case", n, name, len(gates), "gates", flush=True)
    re1, im1 = _run(n, gates, 1, state)
    re0, im0 = _run(n, gates, 0, state)
    assert np.array_equal(re1, re0) and np.array_equal(im1, im0)
    ore, oim = O.run(n, gates, state)
    err = max(np.abs(re1 - ore).max(), np.abs(im1 - oim).max())
    print("max |err| vs oracle", err, flush=True)
    assert err < TOL


def test_prefetch_runs_twice_on_one_handle():
    """Consecutive runs and uploads on one handle (the kernel's LDS buffers
    and row table start fresh every launch)."""
    n = 21
    g1 = W.u3_brickwall(n=n, layers=3)
    g2 = W.random_layers(n=n, layers=6)
    with svmi.Engine(n) as e:
        e.set_option(svmi.OPT_TILE_MAP, 0)
        e.set_option(svmi.OPT_PREFETCH, 1)
        e.upload(g1)
        e.run()
        e.upload(g2)
        e.run()
        re, im = e.get_state()
    ore, oim = O.run(n, g1)
    ore, oim = O.run(n, g2, (ore, oim))
    assert max(np.abs(re - ore).max(), np.abs(im - oim).max()) < TOL


def _run_spec(n, gates, prefetch, state=None):
    with svmi.Engine(n) as e:
        e.set_option(svmi.OPT_PREFETCH, prefetch)
        e.set_option(svmi.OPT_SPECIALIZE, 1)
        e.upload(gates)
        assert e.stats()["specialized_sweeps"] > 0
        if state is not None:
            e.set_state(*state)
        e.run()
        return e.get_state()


@pytest.mark.parametrize("n,name", [(22, "u3_brickwall"), (21, "qft"), (21, "random_layers")])
def test_specialised_prefetch_sweeps_bit_identical_and_match_oracle(n, name):
    """Specialised kernels in the persistent double-buffered form (SVMI_OPT_SPECIALIZE
    with SVMI_OPT_PREFETCH) against the one-tile-per-workgroup specialised kernels
    (same arithmetic: equal states) and the oracle."""
    state = None
    if name == "random_layers":
        gates = W.random_layers(n=n, layers=12)
    elif name == "u3_brickwall":
        gates = W.u3_brickwall(n=n, layers=6)
    else:
        gates = W.qft(n)
        state = _random_state(n, 78)
    print("case", n, name, len(gates), "gates", flush=True)
    re1, im1 = _run_spec(n, gates, 1, state)
    re0, im0 = _run_spec(n, gates, 0, state)
    assert np.array_equal(re1, re0) and np.array_equal(im1, im0)
    ore, oim = O.run(n, gates, state)
    assert max(np.abs(re1 - ore).max(), np.abs(im1 - oim).max()) < TOL
