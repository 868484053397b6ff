"""GPU parity: the HIP engine through its C ABI against the oracle.

Bar (from the workload definitions): amplitudes and expectations within
1e-10 max abs error of the FP64 CPU oracle; sampled indices bit-exact for the
same fed draws except draws closer than DELTA to a scan boundary, where the
two scans' rounding may legitimately differ (counted and bounded).
"""
import json
import os

import numpy as np
import pytest

import svmi
from oracle import oracle as O
from svmi import workloads as W

pytestmark = pytest.mark.gpu

TOL = 1e-10
GOLD = os.path.join(os.path.dirname(__file__), "golden")


def _err(a, b):
    return max(np.abs(a[0] - b[0]).max(), np.abs(a[1] - b[1]).max())


def _gpu_run(n, gates, state=None, **opts):
    with svmi.Engine(n) as e:
        for k, v in opts.items():
            e.set_option({"fusion": svmi.OPT_FUSION, "tile": svmi.OPT_TILE_QUBITS, "low": svmi.OPT_LOW_QUBITS,
                          "budget": svmi.OPT_SWEEP_FLOPS, "coef": svmi.OPT_COEF_BUDGET,
                          "opcost": svmi.OPT_OP_COST, "pipeline": svmi.OPT_PIPELINE}[k], v)
        if state is not None:
            e.set_state(*state)
        e.upload(gates)
        e.run()
        return e.get_state()


def _sample_check(n, re, im, u, got):
    """Bit-exact except draws within DELTA of an oracle scan boundary."""
    want = O.measure(n, re, im, u)
    scan = O.prob_scan(n, re, im)
    delta = 64 * np.finfo(float).eps * max(1.0, scan[-1]) * np.log2(len(scan))
    pos = np.searchsorted(scan, u)
    near = np.zeros(len(u), bool)
    for d in (-1, 0, 1):
        p = np.clip(pos + d, 0, len(scan) - 1)
        near |= np.abs(scan[p] - u) <= delta
    bad = (got != want) & ~near
    return int(bad.sum()), int(near.sum())


def test_random_layers_n12_amplitudes_and_shots():
    d = np.load(os.path.join(GOLD, "random_layers_n12.npz"))
    n = int(d["n"])
    u = d["draws"].astype(np.float64) / float(W.RAND_MAX)
    for fusion in (1, 0):
        with svmi.Engine(n) as e:
            e.set_option(svmi.OPT_FUSION, fusion)
            e.upload(d["gates"])
            e.run()
            st = e.get_state()
            assert _err(st, (d["re"], d["im"])) < TOL
            shots, total = e.measure(u)
        assert abs(total - 1.0) < 1e-12
        bad, near = _sample_check(n, d["re"], d["im"], u, shots)
        assert bad == 0 and near == 0
        assert np.array_equal(shots, d["shots"])


def test_gate_kats_all_ops():
    d = np.load(os.path.join(GOLD, "gate_kats.npz"))
    n = int(d["n"])
    for fusion in (1, 0):
        for i, g in enumerate(d["gates"]):
            got = _gpu_run(n, np.array([g], dtype=svmi.GATE_DTYPE), (d["init_re"][i], d["init_im"][i]), fusion=fusion)
            assert _err(got, (d["out_re"][i], d["out_im"][i])) < 1e-13, (fusion, int(g["op"]), i)


@pytest.mark.parametrize("n", [1, 3, 4, 6, 9, 13, 15])
@pytest.mark.parametrize("opts", [{}, {"fusion": 0}, {"tile": 6, "low": 2}, {"tile": 8, "low": 0}, {"tile": 11}, {"tile": 12},
                                  {"coef": 2000}, {"opcost": 0}, {"tile": 13}])
def test_random_all_op_circuits(n, opts):
    rng = np.random.default_rng(500 + n)
    gates = []
    for _ in range(3):
        for op in range(38):
            k = svmi.ARITY[svmi.OP_NAMES[op]]
            if k <= n:
                gates.append((op, list(rng.choice(n, size=k, replace=False)), *(rng.random(3) * 8 - 4)))
    st = W.random_state(n, seed=n)
    assert _err(_gpu_run(n, gates, st, **opts), O.run(n, gates, st)) < TOL


@pytest.mark.parametrize("n", [14, 20])
def test_u3_brickwall_and_qft_vs_oracle(n):
    g = W.u3_brickwall(n=n, layers=20)
    want = O.run(n, g)
    assert _err(_gpu_run(n, g), want) < TOL
    assert _err(_gpu_run(n, g, coef=3000, tile=8), want) < TOL  # >= 512 tiles: sweeps split by coefficient bytes
    assert _err(_gpu_run(n, g, pipeline=1), want) < TOL       # persistent prefetching sweeps
    st = W.random_state(n, seed=3000 + n)
    q = W.qft(n=n)
    assert _err(_gpu_run(n, q, st), O.run(n, q, st)) < TOL


def test_reference_hash_circuits():
    # circuits whose reference amplitudes are pinned by SHA-256 in hashes.json;
    # the oracle reproduces them bit-exactly (test_oracle), the GPU within TOL
    with open(os.path.join(GOLD, "hashes.json")) as f:
        assert len(json.load(f)) == 6
    g, _, _, _ = W.hea_ising(n=16, layers=4)
    assert _err(_gpu_run(16, g), O.run(16, g)) < TOL


def test_qft_closed_form_n24():
    n = 24
    st = W.random_state(n, seed=77)
    got = _gpu_run(n, W.qft(n=n), st)
    psi = st[0] + 1j * st[1]
    R = np.arange(1 << n)
    rev = np.zeros_like(R)
    for b in range(n):
        rev |= ((R >> b) & 1) << (n - 1 - b)
    want = (np.fft.ifft(psi[rev]) * np.sqrt(1 << n))[rev]
    assert np.abs(got[0] + 1j * got[1] - want).max() < TOL


def test_random_state_generator_matches_host():
    n = 14
    with svmi.Engine(n) as e:
        e.set_random_state(1234)
        re, im = e.get_state()
        sub = e.get_amplitudes([0, 5, (1 << n) - 1])
    hr, hi = svmi.random_amplitudes(1234, np.arange(1 << n))
    nrm = np.sqrt(np.sum(hr * hr + hi * hi))
    assert np.abs(re - hr / nrm).max() < 1e-12 and np.abs(im - hi / nrm).max() < 1e-12
    assert abs(np.sum(re * re + im * im) - 1.0) < 1e-12
    assert np.array_equal(sub[0], re[[0, 5, (1 << n) - 1]])


def _inverse_qft(n):
    g = W.qft(n=n)
    inv = []
    for row in g[::-1]:
        name = svmi.OP_NAMES[int(row["op"])]
        qs = [int(q) for q in row["qb"][: svmi.ARITY[name]]]
        inv.append((int(row["op"]), qs, 0.0, 0.0, -row["lambda"]))
    return inv


def test_qft_n33_closed_form_and_round_trip():
    # 33 qubits, 128 GiB state: QFT of a basis state has the closed form
    # exp(2 pi i rev(x) rev(y) / N) / sqrt(N) (qubit 0 read as the most
    # significant bit); QFT followed by its inverse returns a random input
    n = 33
    N = 1 << n
    rng = np.random.default_rng(33)
    ys = np.unique(rng.integers(0, N, size=2048, dtype=np.uint64))
    x = 0b101100111000011110000011111000001

    def rev(v):
        return int(format(int(v), f"0{n}b")[::-1], 2)

    with svmi.Engine(n) as e:
        e.upload([(svmi.OP["X"], [q]) for q in range(n) if (x >> q) & 1])
        e.run()
        e.upload(W.qft(n=n))
        e.run()
        re, im = e.get_amplitudes(ys)
        rx = rev(x)
        ang = np.array([2 * np.pi * ((rx * rev(y)) % N) / N for y in ys])
        want = np.exp(1j * ang) / np.sqrt(N)
        assert np.abs(re + 1j * im - want).max() < 1e-13
        e.set_random_state(2033)
        r0, i0 = e.get_amplitudes(ys)
        e.run()
        e.upload(_inverse_qft(n))
        e.run()
        r1, i1 = e.get_amplitudes(ys)
    assert max(np.abs(r1 - r0).max(), np.abs(i1 - i0).max()) < TOL


def test_large_state_round_trip_n28():
    # n = 28 (4 GiB): circuit followed by its inverse returns the input state
    n = 28
    g = W.u3_brickwall(n=n, layers=4)
    inv = []
    for row in g[::-1]:
        if row["op"] == svmi.OP["U3"]:
            inv.append((svmi.OP["U3"], [int(row["qb"][0])], -row["theta"], -row["lambda"], -row["phi"]))
        else:
            inv.append((int(row["op"]), [int(row["qb"][0]), int(row["qb"][1])]))
    st = W.random_state(n, seed=28)
    with svmi.Engine(n) as e:
        e.set_state(*st)
        e.upload(g)
        e.run()
        mid = e.get_state()
        assert abs(np.sum(mid[0] ** 2 + mid[1] ** 2) - 1.0) < 1e-12
        e.upload(inv)
        e.run()
        back = e.get_state()
    assert _err(back, st) < TOL


def test_sampling_random_state_n20():
    n = 20
    st = W.random_state(n, seed=20)
    _, u = W.draws(4096, seed=21)
    with svmi.Engine(n) as e:
        e.set_state(*st)
        got, total = e.measure(u)
    assert abs(total - 1.0) < 1e-12
    bad, near = _sample_check(n, st[0], st[1], u, got)
    assert bad == 0 and near <= 2


def test_expval_vs_oracle():
    n = 12
    st = W.random_state(n, seed=5)
    rng = np.random.default_rng(6)
    terms = [(rng.standard_normal(), int(rng.integers(0, 1 << n)), int(rng.integers(0, 1 << n))) for _ in range(12)]
    with svmi.Engine(n) as e:
        e.set_state(*st)
        got = e.expval(terms)
    assert abs(got - O.expval(n, *st, terms)) < TOL


@pytest.mark.parametrize("n,layers", [(6, 3), (12, 4)])
def test_adjoint_vs_oracle_parameter_shift(n, layers):
    rng = np.random.default_rng(n)
    gates, slots, P = W.hea(n, layers, rng)
    terms = W.ising_terms(n, rng)
    with svmi.Engine(n) as e:
        e.upload(gates)
        E, g = e.adjoint(terms, slots, P)
    assert abs(E - O.expval(n, *O.run(n, gates), terms)) < TOL
    assert np.abs(g - O.param_shift_grad(n, gates, slots, P, terms)).max() < TOL


@pytest.mark.parametrize("n,tile", [(3, 12), (5, 12), (9, 6), (13, 12), (14, 8)])
def test_adjoint_every_parametric_gate_vs_oracle(n, tile):
    # two-state sweeps: every op kind the backward program can hold (dense 1-
    # and 2-qubit blocks, diagonals, permutations, phase products, gradient
    # ops), at tiles from 2 threads (shuffle reduction) to 256 (DPP)
    rng = np.random.default_rng(700 + n)
    names = ["U3", "U2", "U1", "RX", "RY", "RZ", "CRX", "CRY", "CRZ", "CU1", "CU3", "RXX", "RZZ", "RYY",
             "H", "CX", "CZ", "SWAP", "CCX", "T", "S"]
    gates, slots, p = [], [], 0
    for _ in range(2):
        for name in names:
            k = svmi.ARITY[name]
            if k > n:
                continue
            gates.append((svmi.OP[name], list(rng.choice(n, size=k, replace=False)), *(rng.random(3) * 6 - 3)))
            slots.append([p, p + 1, p + 2])
            p += 3
    slots = np.array(slots)
    terms = W.ising_terms(n, rng)
    st = W.random_state(n, seed=n + 40)
    with svmi.Engine(n) as e:
        e.set_option(svmi.OPT_TILE_QUBITS, tile)
        e.set_state(*st)
        e.upload(gates)
        E, g = e.adjoint(terms, slots, p)
    Eh, gh = svmi.emulate_adjoint(n, gates, terms, slots, p, st, tile_qubits=tile)
    assert abs(E - O.expval(n, *O.run(n, gates, st), terms)) < TOL
    assert abs(E - Eh) < 1e-12 and np.abs(g - gh).max() < 1e-12
    # parameter shift is exact for the slots of its list; finite differences
    # (5-point, h = 1e-5) for the rest
    G = O.as_gates(gates)
    fd = np.zeros(p)
    nm = ("theta", "phi", "lambda")
    for gi in range(len(G)):
        for s_ in range(3):
            vals = []
            for h in (1e-5, -1e-5, 2e-5, -2e-5):
                G2 = G.copy()
                G2[gi][nm[s_]] += h
                vals.append(O.expval(n, *O.run(n, G2, st), terms))
            fd[slots[gi, s_]] = (8 * (vals[0] - vals[1]) - (vals[2] - vals[3])) / 12e-5
    assert np.abs(g - fd).max() < 1e-8


def test_adjoint_reuses_and_replans_backward_program():
    # one engine: repeated call, a different parameter map, then new angles
    n = 9
    rng = np.random.default_rng(77)
    gates, slots, P = W.hea(n, 3, rng)
    terms = W.ising_terms(n, rng)
    shared = slots.copy()
    shared[shared >= 0] %= 5          # 5 parameters, each feeding several slots
    gates2, _, _ = W.hea(n, 3, rng)   # same structure, new angles
    with svmi.Engine(n) as e:
        e.upload(gates)
        for g_, s_, p_ in ((gates, slots, P), (gates, slots, P), (gates, shared, 5)):
            e.reset_state()
            E, g = e.adjoint(terms, s_, p_)
            assert abs(E - O.expval(n, *O.run(n, g_), terms)) < TOL
            assert np.abs(g - O.param_shift_grad(n, g_, s_, p_, terms)).max() < TOL
        e.upload(gates2)
        e.reset_state()
        E, g = e.adjoint(terms, slots, P)
        assert abs(E - O.expval(n, *O.run(n, gates2), terms)) < TOL
        assert np.abs(g - O.param_shift_grad(n, gates2, slots, P, terms)).max() < TOL


def test_adjoint_hea_n20_spot_check():
    gates, slots, P, terms = W.hea_ising()
    n = 20
    with svmi.Engine(n) as e:
        e.upload(gates)
        E, g = e.adjoint(terms, slots, P)
        # GPU parameter shift for every parameter
        G = gates.copy()
        ps = np.zeros(P)
        names = ("theta", "phi", "lambda")
        for gi in range(len(G)):
            for s in range(3):
                p = slots[gi, s]
                if p < 0:
                    continue
                vals = []
                for sh in (np.pi / 2, -np.pi / 2):
                    G2 = G.copy()
                    G2[gi][names[s]] += sh
                    e.reset_state()
                    e.upload(G2)
                    e.run()
                    vals.append(e.expval(terms))
                ps[p] = 0.5 * (vals[0] - vals[1])
    assert abs(E - O.expval(n, *O.run(n, gates), terms)) < TOL
    assert np.abs(g - ps).max() < TOL
    # oracle parameter shift on a few parameters
    for p in (0, 159, 319):
        gi, s = np.argwhere(slots == p)[0]
        vals = []
        for sh in (np.pi / 2, -np.pi / 2):
            G2 = gates.copy()
            G2[gi][names[s]] += sh
            vals.append(O.expval(n, *O.run(n, G2), terms))
        assert abs(g[p] - 0.5 * (vals[0] - vals[1])) < TOL


def test_batch_matches_single():
    items = W.batch_ising(count=6, nmin=4, nmax=12, lmin=1, lmax=3, seed=9)
    ev, grads = svmi.batch_adjoint(items)
    for it, e, g in zip(items, ev, grads):
        n = it["n"]
        assert abs(e - O.expval(n, *O.run(n, it["gates"]), it["terms"])) < TOL
        assert np.abs(g - O.param_shift_grad(n, it["gates"], it["param_slots"], it["n_params"], it["terms"])).max() < TOL


def test_batch_with_planning_overrides():
    # per-item tiles and fusion op cost (instead of the autotuned table)
    items = W.batch_ising(count=5, nmin=5, nmax=13, lmin=1, lmax=3, seed=33)
    for ov in ({"fwd_tile": 6, "adj_tile": 6, "op_cost": 0}, {"fwd_tile": 13, "adj_tile": 9, "op_cost": 8}):
        ev, grads = svmi.batch_adjoint([dict(it, **ov) for it in items])
        for it, e, g in zip(items, ev, grads):
            n = it["n"]
            assert abs(e - O.expval(n, *O.run(n, it["gates"]), it["terms"])) < TOL
            want = O.param_shift_grad(n, it["gates"], it["param_slots"], it["n_params"], it["terms"])
            assert np.abs(g - want).max() < TOL, ov


def test_errors_are_reported():
    with svmi.Engine(3) as e:
        with pytest.raises(svmi.SvmiError) as ex:
            e.run()
        assert ex.value.status == 4
        with pytest.raises(svmi.SvmiError) as ex:
            e.upload([(svmi.OP["CX"], [0, 0])])
        assert ex.value.status == 2
        e.upload([(svmi.OP["R"], [0], 0.3)])
        with pytest.raises(svmi.SvmiError) as ex:
            e.adjoint([(1.0, 0, 1)], [[0, -1, -1]], 1)
        assert ex.value.status == 7
    with pytest.raises(svmi.SvmiError):
        svmi.Engine(4, device=64)


def test_second_constructor_argument_is_the_device_ordinal():
    """Simulation(n, i) selects HIP device i, as the AMD single-GPU engine's
    constructor does (svsim/src/svsim_amdgpu_sin.hpp:207-220); an ordinal
    past the visible devices (e.g. a core count meant for the CPU build) is
    a clear error, never a silent device change."""
    import libsvsim as svsim
    import torch
    ndev = torch.cuda.device_count()
    sim = svsim.Simulation(3, 0)
    sim.append(sim.X(1))
    sim.upload()
    sim.run()
    assert sim.measure(3) == [2, 2, 2]
    with pytest.raises(RuntimeError, match="HIP ordinal"):
        svsim.Simulation(3, ndev)


def test_libsvsim_dropin_adder():
    # svsim/example/adder_n10.py with the reference API, then the measured
    # basis state is checked: 1 + 15 on the Cuccaro adder -> deterministic
    import libsvsim as svsim
    n = 10
    sim = svsim.Simulation(n, 0)

    def majority(a, b, c):
        sim.append(sim.CX(c, b))
        sim.append(sim.CX(c, a))
        sim.append(sim.CCX(a, b, c))

    def unmaj(a, b, c):
        sim.append(sim.CCX(a, b, c))
        sim.append(sim.CX(c, a))
        sim.append(sim.CX(a, b))

    for q in (1, 5, 6, 7, 8):
        sim.append(sim.X(q))
    sim.upload()
    sim.run()
    sim.clear_circuit()
    majority(0, 5, 1)
    majority(1, 6, 2)
    majority(2, 7, 3)
    majority(3, 8, 4)
    sim.append(sim.CX(4, 9))
    unmaj(3, 8, 4)
    unmaj(2, 7, 3)
    unmaj(1, 6, 2)
    unmaj(0, 5, 1)
    sim.upload()
    sim.run()
    res = sim.measure(10)
    assert len(res) == 10 and len(set(res)) == 1
    # same circuit on the oracle from |0>
    gates = [(svmi.OP["X"], [q]) for q in (1, 5, 6, 7, 8)]
    re0, im0 = O.run(n, gates)
    re, im = sim.get_sv()
    assert np.abs(re - O.run(n, [g for g in _adder_gates()], (re0, im0))[0]).max() < TOL
    assert np.argmax(re ** 2 + im ** 2) == res[0]
    sim.reset()
    re, im = sim.get_sv()
    assert re[0] == 1.0 and np.abs(re[1:]).max() == 0.0


def _adder_gates():
    out = []

    def cx(c, t):
        out.append((svmi.OP["CX"], [c, t]))

    def ccx(a, b, c):
        out.append((svmi.OP["CCX"], [a, b, c]))

    for a, b, c in ((0, 5, 1), (1, 6, 2), (2, 7, 3), (3, 8, 4)):
        cx(c, b), cx(c, a), ccx(a, b, c)
    cx(4, 9)
    for a, b, c in ((3, 8, 4), (2, 7, 3), (1, 6, 2), (0, 5, 1)):
        ccx(a, b, c), cx(c, a), cx(a, b)
    return out
