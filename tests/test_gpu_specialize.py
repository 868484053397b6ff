"""Specialised sweeps (SVMI_OPT_SPECIALIZE): the hipRTC-compiled
straight-line kernels against the oracle (1e-10) and against the interpreted
sweeps of the same program, on circuits that use every forward op type the
planner emits, at several tile shapes; sampling after a specialised run; the
option switched on and off after upload."""
import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _err(a, b):
    return max(np.abs(a[0] - b[0]).max(), np.abs(a[1] - b[1]).max())


def _all_op_circuit(n, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(3):
        for op in range(38):
            k = svmi.ARITY[svmi.OP_NAMES[op]]
            if k <= n:
                gates.append((op, list(rng.choice(n, size=k, replace=False)), *(rng.random(3) * 8 - 4)))
    return gates


def _run(n, gates, state, spec, tile=None, low=None):
    with svmi.Engine(n) as e:
        e.set_option(svmi.OPT_SPECIALIZE, spec)
        if tile is not None:
            e.set_option(svmi.OPT_TILE_QUBITS, tile)
        if low is not None:
            e.set_option(svmi.OPT_LOW_QUBITS, low)
        e.set_state(*state)
        e.upload(gates)
        e.run()
        return e.get_state(), e.stats()


@pytest.mark.parametrize("n,tile,low", [(4, None, None), (9, None, None), (15, None, None), (15, 6, 2), (14, 8, 0),
                                        (15, 13, None)])
def test_all_op_circuits_vs_oracle_and_interpreter(n, tile, low):
    gates = _all_op_circuit(n, 900 + n)
    st = W.random_state(n, seed=n)
    print(f"case: n={n} tile={tile} low={low}", flush=True)
    got, stats = _run(n, gates, st, 1, tile, low)
    interp, _ = _run(n, gates, st, 0, tile, low)
    assert _err(got, O.run(n, gates, st)) < TOL
    assert _err(got, interp) < 1e-13
    if tile == 13:
        assert stats["specialized_sweeps"] == 0    # tiles above 12 qubits stay interpreted
    else:
        assert stats["specialized_sweeps"] > 0 and stats["compile_ms"] > 0


def test_brickwall_and_qft_n20_and_sampling():
    n = 20
    for g in (W.u3_brickwall(n=n, layers=6), W.qft(n=n)):
        st = W.random_state(n, seed=77)
        u = W.draws(4096, seed=78)[1]
        out = []
        for spec in (1, 0):
            with svmi.Engine(n) as e:
                e.set_option(svmi.OPT_SPECIALIZE, spec)
                e.set_state(*st)
                e.upload(g)
                e.run()
                if spec:
                    assert e.stats()["specialized_sweeps"] > 0
                out.append((e.get_state(), e.measure(u)[0]))
        assert _err(out[0][0], O.run(n, g, st, nthreads=8)) < TOL
        assert _err(out[0][0], out[1][0]) < 1e-13
        assert np.array_equal(out[0][1], out[1][1])


def test_option_toggles_after_upload():
    n = 12
    g = W.random_layers(n=n, layers=10)
    want = O.run(n, g)
    with svmi.Engine(n) as e:
        e.upload(g)
        for spec in (1, 0, 1):
            e.set_option(svmi.OPT_SPECIALIZE, spec)
            assert (e.stats()["specialized_sweeps"] > 0) == bool(spec)
            e.reset_state()
            e.run()
            assert _err(e.get_state(), want) < TOL
        # re-uploading new angles of the same structure recompiles (hipRTC's cache serves it)
        g2 = W.random_layers(n=n, layers=10, seed=4242)
        e.upload(g2)
        e.reset_state()
        e.run()
        assert e.stats()["specialized_sweeps"] > 0
        assert _err(e.get_state(), O.run(n, g2)) < TOL
