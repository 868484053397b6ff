"""torch.autograd bindings of the engine's adjoint gradient.

    circ = ExpvalCircuit(n, gates, param_slots, n_params, terms)
    params = torch.zeros(n_params, dtype=torch.float64, requires_grad=True)
    E = circ(params)            # <psi(params)| H |psi(params)>, psi from |0..0>
    E.backward()                # params.grad = dE/dparams (adjoint sweep)

A gate's angle slot (theta, phi, lambda) that param_slots maps to parameter p
takes the value params[p]; a parameter may feed several slots, and the
gradient sums over them.  The forward call runs the circuit and the adjoint
sweep together on the GPU (one svmi_adjoint call), so backward only scales
the stored gradient.  batch_expval does the same for many circuits through
svmi_batch_adjoint.  Higher derivatives are not provided.
"""
import numpy as np
import torch
from torch.autograd.function import once_differentiable

from . import Engine, as_gates, batch_adjoint

_SLOT_FIELDS = ("theta", "phi", "lambda")


def bind_params(gates, param_slots, values):
    """Copy of the gate array with every trainable slot set to values[p]."""
    g = as_gates(gates).copy()
    ps = np.asarray(param_slots, dtype=np.int32).reshape(-1, 3)
    if len(ps) != len(g):
        raise ValueError("param_slots needs one row of 3 entries per gate")
    values = np.asarray(values, dtype=np.float64)
    for s, f in enumerate(_SLOT_FIELDS):
        m = ps[:, s] >= 0
        if np.any(ps[m, s] >= len(values)):
            raise ValueError("param_slots refers to a parameter beyond n_params")
        g[f][m] = values[ps[m, s]]
    return g


def _host_values(params):
    if params.dim() != 1:
        raise ValueError("params must be a 1-D tensor")
    return params.detach().to(device="cpu", dtype=torch.float64).numpy()


class _Expval(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, runner):
        e, grad = runner(_host_values(params))
        ctx.save_for_backward(torch.as_tensor(np.asarray(grad), dtype=params.dtype).to(params.device))
        return torch.tensor(float(e), dtype=params.dtype, device=params.device)

    @staticmethod
    @once_differentiable
    def backward(ctx, g_out):
        (grad,) = ctx.saved_tensors
        return g_out * grad, None


class ExpvalCircuit:
    """<H> of a parameterised circuit as a differentiable function of its
    parameters, on one GPU state of n qubits."""

    def __init__(self, n, gates, param_slots, n_params, terms, device=0):
        self.n = int(n)
        self.gates = as_gates(gates)
        self.param_slots = np.ascontiguousarray(param_slots, dtype=np.int32).reshape(-1, 3)
        self.n_params = int(n_params)
        self.terms = terms
        self.engine = Engine(self.n, device)

    def run(self, values):
        """(E, dE/dparams) for parameter values (numpy), on the GPU."""
        self.engine.upload(bind_params(self.gates, self.param_slots, values))
        self.engine.reset_state()
        return self.engine.adjoint(self.terms, self.param_slots, self.n_params)

    def __call__(self, params):
        if params.numel() != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {params.numel()}")
        return _Expval.apply(params, self.run)

    def close(self):
        self.engine.close()


class _BatchExpval(torch.autograd.Function):
    @staticmethod
    def forward(ctx, runner, *params):
        ev, grads = runner([_host_values(p) for p in params])
        ctx.save_for_backward(*[torch.as_tensor(g, dtype=p.dtype).to(p.device) for g, p in zip(grads, params)])
        ref = params[0]
        return torch.as_tensor(np.asarray(ev), dtype=ref.dtype).to(ref.device)

    @staticmethod
    @once_differentiable
    def backward(ctx, g_out):
        grads = ctx.saved_tensors
        return (None, *[g_out[i] * g for i, g in enumerate(grads)])


def batch_expval(circuits, params, device=0):
    """Expectations of many circuits at once (one svmi_batch_adjoint call).

    circuits: list of dicts {n, gates, param_slots, n_params, terms};
    params: list of 1-D tensors, one per circuit.  Returns a tensor of the
    expectations, differentiable with respect to every params tensor."""
    if len(circuits) != len(params) or not circuits:
        raise ValueError("one params tensor per circuit is required")

    def runner(values):
        items = []
        for c, v in zip(circuits, values):
            if v.size != int(c["n_params"]):
                raise ValueError("params size does not match n_params")
            items.append({"n": c["n"], "gates": bind_params(c["gates"], c["param_slots"], v),
                          "param_slots": c["param_slots"], "n_params": c["n_params"], "terms": c["terms"]})
        return batch_adjoint(items, device=device)

    return _BatchExpval.apply(runner, *params)
