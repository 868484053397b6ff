"""GPU parity at the sizes of the measured workloads.

The other GPU tests compare against the oracle at up to 20 qubits.  Here:
  - batched gradients of ansatz circuits of 13-22 qubits (the upper half of
    the 512-circuit batch's range, where the batch plans 13-qubit forward
    tiles from the tuning table), against the oracle's energy and parameter
    shift and against the single-circuit adjoint;
  - the 20-qubit, 320-parameter ansatz gradient through the
    torch.autograd.Function, against central finite differences
    (eps = 1e-6) of GPU energies and the oracle's parameter shift;
  - the 30-qubit u3 brick-wall circuit (20 layers), which the oracle does not
    run in test time: total probability, sampled indices against a
    higher-precision scan of the amplitudes read back from the device, and
    the circuit followed by its inverse returning |0..0>.
Bars: energies, gradients and amplitudes 1e-10 (finite differences 1e-7,
their own truncation and rounding error being ~1e-9); a sample is correct
when its draw lies in its index's scan interval to within 1e-12.
"""
import numpy as np
import pytest
import torch

import svmi
from oracle import oracle as O
from svmi import workloads as W
from svmi.autograd import ExpvalCircuit, bind_params

pytestmark = pytest.mark.gpu

TOL = 1e-10
NAMES = ("theta", "phi", "lambda")


def _slot_values(gates, slots, n_params):
    """Parameter values the circuit's bound angles hold."""
    v = np.zeros(n_params)
    for gi in range(len(gates)):
        for s in range(3):
            if slots[gi, s] >= 0:
                v[slots[gi, s]] = gates[gi][NAMES[s]]
    return v


def _oracle_shift(n, gates, slots, terms, p):
    """Oracle parameter-shift derivative for parameter p (RY theta / RZ phi
    slots: two-term rule with shift pi/2, summed over the slots p feeds)."""
    d = 0.0
    for gi, s in np.argwhere(slots == p):
        vals = []
        for sh in (np.pi / 2, -np.pi / 2):
            g2 = gates.copy()
            g2[gi][NAMES[s]] += sh
            vals.append(O.expval(n, *O.run(n, g2), terms))
        d += 0.5 * (vals[0] - vals[1])
    return d


def test_batch_13_to_22_qubits_vs_oracle():
    rng = np.random.default_rng(2024)
    items = []
    for n, layers in ((13, 3), (14, 2), (16, 3), (18, 2), (20, 2), (22, 2)):
        sub = np.random.default_rng(int(rng.integers(1 << 62)))
        gates, slots, p = W.hea(n, layers, sub)
        items.append({"n": n, "gates": gates, "param_slots": slots, "n_params": p, "terms": W.ising_terms(n, sub)})
    print("batch of", [it["n"] for it in items], flush=True)
    ev, grads = svmi.batch_adjoint(items)
    for it, e, g in zip(items, ev, grads):
        n, gates, slots, P, terms = it["n"], it["gates"], it["param_slots"], it["n_params"], it["terms"]
        print("circuit n =", n, "params =", P, flush=True)
        assert abs(e - O.expval(n, *O.run(n, gates), terms)) < TOL
        with svmi.Engine(n) as eng:
            eng.upload(gates)
            e1, g1 = eng.adjoint(terms, slots, P)
        assert abs(e - e1) < TOL and np.abs(g - g1).max() < TOL
        sel = range(P) if n <= 14 else np.linspace(0, P - 1, 5).astype(int)
        for p in sel:
            assert abs(g[p] - _oracle_shift(n, gates, slots, terms, p)) < TOL, (n, p)


def test_c4_ansatz_n20_through_autograd_vs_central_fd_and_oracle():
    gates, slots, P, terms = W.hea_ising()
    n = 20
    v0 = _slot_values(gates, slots, P)
    circ = ExpvalCircuit(n, gates, slots, P, terms)
    params = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    E = circ(params)
    E.backward()
    g = params.grad.numpy()
    circ.close()
    assert abs(E.item() - O.expval(n, *O.run(n, gates), terms)) < TOL
    eps = 1e-6
    sel = np.linspace(0, P - 1, 12).astype(int)
    with svmi.Engine(n) as eng:
        def energy(v):
            eng.upload(bind_params(gates, slots, v))
            eng.reset_state()
            eng.run()
            return eng.expval(terms)
        for p in sel:
            vp, vm = v0.copy(), v0.copy()
            vp[p] += eps
            vm[p] -= eps
            fd = (energy(vp) - energy(vm)) / (2 * eps)
            assert abs(g[p] - fd) < 1e-7, (p, g[p], fd)
    for p in (0, 161, 319):
        assert abs(g[p] - _oracle_shift(n, gates, slots, terms, p)) < TOL, p


def _inverse_u3_brickwall(g):
    inv = []
    for row in g[::-1]:
        if row["op"] == svmi.OP["U3"]:
            inv.append((svmi.OP["U3"], [int(row["qb"][0])], -row["theta"], -row["lambda"], -row["phi"]))
        else:
            inv.append((int(row["op"]), [int(row["qb"][0]), int(row["qb"][1])]))
    return inv


def test_u3_brickwall_n30_probability_sampling_and_inverse():
    n = 30
    g = W.u3_brickwall(n=n, layers=20)
    _, u = W.draws(256, seed=3030)
    with svmi.Engine(n) as e:
        e.upload(g)
        e.run()
        got, total = e.measure(u)
        print("n = 30 forward run, total probability", total, flush=True)
        assert abs(total - 1.0) < 1e-12
        re, im = e.get_state()
        # probabilities in blocks of 2^12, block sums pairwise, the prefix over
        # blocks and inside a block in extended precision
        p = re * re
        p += im * im
        del re, im
        blocks = p.reshape(-1, 1 << 12)
        bsum = blocks.sum(axis=1)
        bpre = np.concatenate([[0.0], np.cumsum(bsum.astype(np.longdouble))])
        assert abs(float(bpre[-1]) - 1.0) < 1e-12
        for r, j in zip(u, got):
            b = int(j) >> 12
            scan = bpre[b] + np.concatenate([[0.0], np.cumsum(blocks[b].astype(np.longdouble))])
            lo, hi = float(scan[int(j) & 4095]), float(scan[(int(j) & 4095) + 1])
            assert lo - 1e-12 <= r < hi + 1e-12, (r, int(j), lo, hi)
        del p, blocks
        e.upload(_inverse_u3_brickwall(g))
        e.run()
        re0, im0 = e.get_amplitudes([0, 1, 1 << 29, (1 << 30) - 1])
        _, total2 = e.measure([])
    print("inverse run: amplitude 0 =", re0[0], im0[0], flush=True)
    assert abs(re0[0] - 1.0) < TOL and abs(im0[0]) < TOL
    assert np.abs(re0[1:]).max() < TOL and np.abs(im0[1:]).max() < TOL
    assert abs(total2 - 1.0) < 1e-12
