"""torch.autograd binding of the adjoint gradient (svmi.autograd).

CPU tests drive the autograd plumbing with the host execution of the
engine's adjoint program; GPU tests run ExpvalCircuit / batch_expval on the
device.  Both are checked against the oracle: CPU parameter shift (exact for
the RY/RZ ansatz gates), central finite differences of the oracle's <H>, and
torch.autograd.gradcheck."""
import numpy as np
import pytest
import torch

import svmi
from oracle import oracle as O
from svmi import workloads as W
from svmi.autograd import ExpvalCircuit, _Expval, batch_expval, bind_params

TOL = 1e-10


def _problem(n, layers, seed):
    rng = np.random.default_rng(seed)
    gates, slots, p = W.hea(n, layers, rng)
    return gates, slots, p, W.ising_terms(n, rng)


def _oracle_energy(n, gates, slots, terms, v):
    g = bind_params(gates, slots, v)
    return O.expval(n, *O.run(n, g), terms)


def _central_fd(n, gates, slots, terms, v, h=1e-5):
    out = np.zeros(len(v))
    for i in range(len(v)):
        vp, vm = v.copy(), v.copy()
        vp[i] += h
        vm[i] -= h
        out[i] = (_oracle_energy(n, gates, slots, terms, vp) - _oracle_energy(n, gates, slots, terms, vm)) / (2 * h)
    return out


def test_bind_params_sets_only_mapped_slots():
    n = 3
    gates, slots, p, _ = _problem(n, 2, 1)
    v = np.arange(p, dtype=np.float64) + 0.25
    g = bind_params(gates, slots, v)
    for i, row in enumerate(slots):
        for s, f in enumerate(("theta", "phi", "lambda")):
            want = v[row[s]] if row[s] >= 0 else gates[i][f]
            assert g[i][f] == want
    with pytest.raises(ValueError):
        bind_params(gates, slots, v[:-1])


def test_autograd_plumbing_on_host_program():
    n = 5
    gates, slots, p, terms = _problem(n, 2, 7)
    re, im = O.zero_state(n)

    def runner(v):
        return svmi.emulate_adjoint(n, bind_params(gates, slots, v), terms, slots, p, (re, im))

    v0 = np.random.default_rng(3).uniform(-np.pi, np.pi, p)
    params = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    E = _Expval.apply(params, runner)
    assert abs(E.item() - _oracle_energy(n, gates, slots, terms, v0)) < 1e-12
    (3.0 * E).backward()
    gp = O.param_shift_grad(n, bind_params(gates, slots, v0), slots, p, terms)
    assert np.abs(params.grad.numpy() - 3.0 * gp).max() < 1e-11
    assert np.abs(params.grad.numpy() / 3.0 - _central_fd(n, gates, slots, terms, v0)).max() < 1e-7
    assert torch.autograd.gradcheck(lambda x: _Expval.apply(x, runner), (params.detach().requires_grad_(),),
                                    eps=1e-6, atol=1e-6)


@pytest.mark.gpu
def test_expval_circuit_gradient_on_gpu():
    n = 10
    gates, slots, p, terms = _problem(n, 3, 11)
    circ = ExpvalCircuit(n, gates, slots, p, terms)
    v0 = np.random.default_rng(5).uniform(-np.pi, np.pi, p)
    params = torch.tensor(v0, dtype=torch.float64, requires_grad=True)
    E = circ(params)
    assert abs(E.item() - _oracle_energy(n, gates, slots, terms, v0)) < TOL
    E.backward()
    g = params.grad.numpy()
    assert np.abs(g - O.param_shift_grad(n, bind_params(gates, slots, v0), slots, p, terms)).max() < TOL
    assert np.abs(g - _central_fd(n, gates, slots, terms, v0)).max() < 1e-7
    small = ExpvalCircuit(4, *_problem(4, 1, 2))
    x = torch.tensor(np.random.default_rng(1).uniform(-1, 1, small.n_params), dtype=torch.float64,
                     requires_grad=True)
    assert torch.autograd.gradcheck(small, (x,), eps=1e-6, atol=1e-6)
    circ.close()
    small.close()


@pytest.mark.gpu
def test_batch_expval_gradients_on_gpu():
    rng = np.random.default_rng(17)
    circuits, params, ref_e, ref_g = [], [], [], []
    for n, layers in ((4, 1), (9, 2), (12, 3), (7, 4)):
        gates, slots, p, terms = _problem(n, layers, int(rng.integers(1 << 30)))
        circuits.append({"n": n, "gates": gates, "param_slots": slots, "n_params": p, "terms": terms})
        v = rng.uniform(-np.pi, np.pi, p)
        params.append(torch.tensor(v, dtype=torch.float64, requires_grad=True))
        ref_e.append(_oracle_energy(n, gates, slots, terms, v))
        ref_g.append(O.param_shift_grad(n, bind_params(gates, slots, v), slots, p, terms))
    E = batch_expval(circuits, params)
    assert np.abs(E.detach().numpy() - np.array(ref_e)).max() < TOL
    w = torch.tensor([1.0, -2.0, 0.5, 3.0], dtype=torch.float64)
    (w * E).sum().backward()
    for i, prm in enumerate(params):
        assert np.abs(prm.grad.numpy() - w[i].item() * ref_g[i]).max() < 1e-9


@pytest.mark.gpu
def test_cuda_tensor_params_and_engine_share_the_process():
    # torch's HIP runtime and the engine's are both live in this process
    n = 6
    gates, slots, p, terms = _problem(n, 2, 23)
    circ = ExpvalCircuit(n, gates, slots, p, terms)
    v0 = np.random.default_rng(9).uniform(-np.pi, np.pi, p)
    params = torch.tensor(v0, dtype=torch.float64, device="cuda", requires_grad=True)
    E = circ(params)
    assert E.device.type == "cuda"
    (E * E).backward()
    e_ref = _oracle_energy(n, gates, slots, terms, v0)
    g_ref = O.param_shift_grad(n, bind_params(gates, slots, v0), slots, p, terms)
    assert abs(E.item() - e_ref) < TOL
    assert np.abs(params.grad.cpu().numpy() - 2 * e_ref * g_ref).max() < 1e-9
    circ.close()
