"""Python access to the MI355X state-vector engine through its C ABI.

``svmi.Engine`` wraps one ``svmi_handle`` of ``libsvmi.so`` (include/svmi.h):
device state, uploaded circuit, run, measure with fed draws, Pauli
expectations and adjoint gradients.  Gate records use the reference layout
(svsim/src/svsim_cpu_omp.hpp:103-194).  The library must be built
(``python sv-sim_amd/build.py``); nothing here falls back to the CPU.
"""
import ctypes
import os
import sys

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libsvmi.so")

# Reference enum OP (svsim_cpu_omp.hpp:46-52).
OP_NAMES = ["U3", "U2", "U1", "CX", "ID", "X", "Y", "Z", "H", "S",
            "SDG", "T", "TDG", "RX", "RY", "RZ", "CZ", "CY", "SWAP", "CH",
            "CCX", "CSWAP", "CRX", "CRY", "CRZ", "CU1", "CU3", "RXX", "RZZ", "RCCX",
            "RC3X", "C3X", "C3SQRTX", "C4X", "R", "SRN", "W", "RYY"]
OP = {name: i for i, name in enumerate(OP_NAMES)}
ARITY = {name: 1 for name in OP_NAMES}
ARITY.update({n: 2 for n in ("CX", "CZ", "CY", "SWAP", "CH", "CRX", "CRY", "CRZ", "CU1", "CU3",
                             "RXX", "RZZ", "RYY")})
ARITY.update({"CCX": 3, "CSWAP": 3, "RCCX": 3, "RC3X": 4, "C3X": 4, "C3SQRTX": 4, "C4X": 5})

GATE_DTYPE = np.dtype([("op", "<i4"), ("qb", "<i4", (5,)),
                       ("theta", "<f8"), ("phi", "<f8"), ("lambda", "<f8")])
PAULI_DTYPE = np.dtype([("coeff", "<f8"), ("xmask", "<u8"), ("zmask", "<u8")])
# svmi_qir_gate: the reference QIR Gate record (svsim_qsharp/svsim_cpu_omp.hpp:113-147).
QIR_GATE_DTYPE = np.dtype([("op", "<i4"), ("qubit", "<i4"), ("mask", "<u8"), ("theta", "<f8")])
# QIR op numbering (svsim_qsharp/svsim_cpu_omp.hpp:45-59).
QIR_OP_NAMES = ["X", "Y", "Z", "H", "S", "T", "RI", "RX", "RY", "RZ", "EI", "EX", "EY", "EZ",
                "ControlledX", "ControlledY", "ControlledZ", "ControlledH", "ControlledS", "ControlledT",
                "ControlledRI", "ControlledRX", "ControlledRY", "ControlledRZ",
                "ControlledEI", "ControlledEX", "ControlledEY", "ControlledEZ",
                "AdjointS", "AdjointT", "ControlledAdjointS", "ControlledAdjointT", "Swap", "Measure"]
QIR_OP = {name: i for i, name in enumerate(QIR_OP_NAMES)}

STATUS = {0: "OK", 1: "ERR_ARG", 2: "ERR_QUBIT", 3: "ERR_OP", 4: "ERR_STATE",
          5: "ERR_DEVICE", 6: "ERR_OOM", 7: "ERR_UNSUPPORTED"}

(OPT_FUSION, OPT_TILE_QUBITS, OPT_LOW_QUBITS, OPT_SWEEP_FLOPS, OPT_PROFILE, OPT_PIPELINE, OPT_OP_COST,
 OPT_COEF_BUDGET, OPT_SPECIALIZE, OPT_MERGE_QUBITS, OPT_PREFETCH, OPT_TILE_MAP) = range(12)

# Every exported entry point (include/svmi.h, include/svmi_debug.h).
EXPORTS = ["svmi_create", "svmi_destroy", "svmi_num_qubits", "svmi_reset_state", "svmi_set_random_state",
           "svmi_set_state", "svmi_get_state", "svmi_get_amplitudes", "svmi_upload", "svmi_clear_circuit",
           "svmi_run", "svmi_measure", "svmi_expval", "svmi_adjoint", "svmi_batch_adjoint",
           "svmi_set_option", "svmi_get_stats", "svmi_last_error", "svmi_abi_version", "svmi_qir_run", "svmi_qasm_parse",
           "svmi_debug_emulate", "svmi_debug_emulate_adjoint", "svmi_debug_emulate_batch", "svmi_debug_plan",
           "svmi_debug_plan_batch", "svmi_debug_plan_adjoint", "svmi_debug_emulate_qir", "svmi_debug_plan_qir",
           "svmi_debug_jit", "svmi_dm_qir_run", "svmi_debug_emulate_dm_qir", "svmi_debug_step_times", "svmi_debug_fp64_rates", "svmi_expval_static"]


class SvmiError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Stats(ctypes.Structure):
    _fields_ = [("sweeps", ctypes.c_uint64), ("kernel_launches", ctypes.c_uint64),
                ("ops", ctypes.c_uint64), ("gates", ctypes.c_uint64), ("device_ms", ctypes.c_double),
                ("sweep_launches", ctypes.c_uint64), ("sweep_ms", ctypes.c_double),
                ("other_launches", ctypes.c_uint64), ("other_ms", ctypes.c_double),
                ("specialized_sweeps", ctypes.c_uint64), ("compile_ms", ctypes.c_double)]


class BatchItem(ctypes.Structure):
    _fields_ = [("n_qubits", ctypes.c_int32), ("fwd_tile", ctypes.c_uint8), ("adj_tile", ctypes.c_uint8),
                ("op_cost", ctypes.c_uint8), ("reserved", ctypes.c_uint8),
                ("gates", ctypes.c_void_p), ("n_gates", ctypes.c_size_t),
                ("param_slots", ctypes.c_void_p), ("n_params", ctypes.c_size_t),
                ("terms", ctypes.c_void_p), ("n_terms", ctypes.c_size_t)]


_lib = None
_vp = ctypes.c_void_p
_sz = ctypes.c_size_t


def _torch_runtime_first():
    """PyTorch-ROCm wheels carry their own HIP runtime; libsvmi.so uses the
    system one.  Both work in one process when torch's runtime initialises
    the device first (the other order leaves torch without GPUs), so if torch
    is loaded, initialise its device runtime before libsvmi.so."""
    t = sys.modules.get("torch")
    if t is not None and t.cuda.is_available():
        t.cuda.init()


def lib():
    """Load libsvmi.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} not built; run python sv-sim_amd/build.py")
        _torch_runtime_first()
        L = ctypes.CDLL(LIB_PATH)
        st = ctypes.c_int
        L.svmi_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
        L.svmi_destroy.argtypes = [_vp]
        L.svmi_num_qubits.argtypes = [_vp]
        L.svmi_num_qubits.restype = ctypes.c_int
        L.svmi_reset_state.argtypes = [_vp]
        L.svmi_set_state.argtypes = [_vp, _vp, _vp]
        L.svmi_set_random_state.argtypes = [_vp, ctypes.c_uint64]
        L.svmi_get_state.argtypes = [_vp, _vp, _vp]
        L.svmi_get_amplitudes.argtypes = [_vp, _vp, _sz, _vp, _vp]
        L.svmi_upload.argtypes = [_vp, _vp, _sz]
        L.svmi_clear_circuit.argtypes = [_vp]
        L.svmi_run.argtypes = [_vp]
        L.svmi_measure.argtypes = [_vp, _vp, _sz, _vp, _vp]
        L.svmi_expval.argtypes = [_vp, _vp, _sz, _vp]
        L.svmi_expval_static.argtypes = [_vp, _vp, _sz, _vp]
        L.svmi_adjoint.argtypes = [_vp, _vp, _sz, _vp, _sz, _vp, _vp]
        L.svmi_batch_adjoint.argtypes = [ctypes.c_int, _vp, _sz, _vp, _vp]
        L.svmi_set_option.argtypes = [_vp, ctypes.c_int, ctypes.c_int64]
        L.svmi_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
        L.svmi_last_error.restype = ctypes.c_char_p
        L.svmi_debug_emulate.argtypes = [ctypes.c_int, _vp, _sz, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int64, _vp, _vp]
        L.svmi_debug_emulate_adjoint.argtypes = [ctypes.c_int, _vp, _sz, _vp, _sz, _vp, _sz, ctypes.c_int,
                                                 ctypes.c_int, _vp, _vp, _vp, _vp]
        L.svmi_debug_emulate_batch.argtypes = [_vp, _sz, _vp, _vp]
        L.svmi_debug_plan.argtypes = [ctypes.c_int, _vp, _sz, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int64, ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, _sz]
        L.svmi_debug_plan_adjoint.argtypes = [ctypes.c_int, _vp, _sz, _vp, _sz, ctypes.c_int, ctypes.c_int,
                                              ctypes.POINTER(ctypes.c_uint64), ctypes.c_char_p, _sz]
        L.svmi_qir_run.argtypes = [_vp, ctypes.c_int, _vp, _sz, _vp, _vp, _vp]
        L.svmi_dm_qir_run.argtypes = [_vp, ctypes.c_int, _vp, _sz, _vp, _vp]
        L.svmi_debug_emulate_dm_qir.argtypes = [ctypes.c_int, _vp, _sz, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                                _vp]
        L.svmi_debug_jit.argtypes = [ctypes.c_int, _vp, _sz, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_uint64),
                                     ctypes.c_char_p, _sz]
        L.svmi_debug_emulate_qir.argtypes = [ctypes.c_int, _vp, _sz, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp,
                                             _vp]
        L.svmi_debug_plan_qir.argtypes = [ctypes.c_int, _vp, _sz, ctypes.c_int, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_uint64),
                                          ctypes.c_char_p, _sz]
        L.svmi_debug_step_times.argtypes = [_vp, _vp, _sz, ctypes.POINTER(ctypes.c_uint64)]
        L.svmi_debug_fp64_rates.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                            ctypes.POINTER(ctypes.c_double)]
        for name in EXPORTS:
            f = getattr(L, name)
            if name not in ("svmi_num_qubits", "svmi_last_error", "svmi_abi_version"):
                f.restype = st
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise SvmiError(status, lib().svmi_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(_vp) if a is not None and a.size else None


def as_gates(gates):
    """GATE_DTYPE array from a GATE_DTYPE array or (op, qubits, theta, phi, lambda) tuples."""
    if isinstance(gates, np.ndarray) and gates.dtype == GATE_DTYPE:
        return np.ascontiguousarray(gates)
    arr = np.zeros(len(gates), dtype=GATE_DTYPE)
    for i, g in enumerate(gates):
        op = OP[g[0]] if isinstance(g[0], str) else int(g[0])
        qs = list(g[1])
        arr[i]["op"] = op
        arr[i]["qb"][: len(qs)] = qs
        vals = (list(g[2:5]) + [0.0, 0.0, 0.0])[:3]
        arr[i]["theta"], arr[i]["phi"], arr[i]["lambda"] = vals
    return arr


def as_qir_gates(gates):
    """Accept a QIR_GATE_DTYPE array or a list of (op, qubit, mask, theta)
    with op a name or number."""
    if isinstance(gates, np.ndarray) and gates.dtype == QIR_GATE_DTYPE:
        return np.ascontiguousarray(gates)
    arr = np.zeros(len(gates), dtype=QIR_GATE_DTYPE)
    for i, g in enumerate(gates):
        op = QIR_OP[g[0]] if isinstance(g[0], str) else int(g[0])
        arr[i] = (op, int(g[1]), int(g[2]) if len(g) > 2 else 0, float(g[3]) if len(g) > 3 else 0.0)
    return arr


def as_terms(terms):
    if isinstance(terms, np.ndarray) and terms.dtype == PAULI_DTYPE:
        return np.ascontiguousarray(terms)
    arr = np.zeros(len(terms), dtype=PAULI_DTYPE)
    for i, (c, x, z) in enumerate(terms):
        arr[i] = (c, x, z)
    return arr


class Engine:
    """One device state vector of n qubits on HIP device `device`."""

    def __init__(self, n_qubits, device=0):
        h = _vp()
        _check(lib().svmi_create(int(n_qubits), int(device), ctypes.byref(h)))
        self._h = h
        self.n = int(n_qubits)
        self.n_gates = 0

    def close(self):
        if getattr(self, "_h", None):
            lib().svmi_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def reset_state(self):
        _check(lib().svmi_reset_state(self._h))

    def set_random_state(self, seed):
        """Seeded random unit state generated on the device (see random_amplitudes)."""
        _check(lib().svmi_set_random_state(self._h, int(seed) & 0xFFFFFFFFFFFFFFFF))

    def set_state(self, re, im):
        re = np.ascontiguousarray(re, dtype=np.float64)
        im = np.ascontiguousarray(im, dtype=np.float64)
        if re.size != (1 << self.n) or im.size != (1 << self.n):
            raise ValueError("state must have 2^n entries")
        _check(lib().svmi_set_state(self._h, _ptr(re), _ptr(im)))

    def get_state(self):
        re = np.empty(1 << self.n)
        im = np.empty(1 << self.n)
        _check(lib().svmi_get_state(self._h, _ptr(re), _ptr(im)))
        return re, im

    def get_amplitudes(self, idx):
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        re = np.empty(len(idx))
        im = np.empty(len(idx))
        _check(lib().svmi_get_amplitudes(self._h, _ptr(idx), len(idx), _ptr(re), _ptr(im)))
        return re, im

    def upload(self, gates):
        g = as_gates(gates)
        _check(lib().svmi_upload(self._h, _ptr(g), len(g)))
        self.n_gates = len(g)

    def clear_circuit(self):
        _check(lib().svmi_clear_circuit(self._h))

    def run(self):
        _check(lib().svmi_run(self._h))

    def measure(self, uniforms):
        u = np.ascontiguousarray(uniforms, dtype=np.float64)
        out = np.zeros(len(u), dtype=np.uint64)
        tot = ctypes.c_double(0.0)
        _check(lib().svmi_measure(self._h, _ptr(u), len(u), _ptr(out), ctypes.byref(tot)))
        return out, tot.value

    def expval(self, terms):
        t = as_terms(terms)
        out = ctypes.c_double(0.0)
        _check(lib().svmi_expval(self._h, _ptr(t), len(t), ctypes.byref(out)))
        return out.value

    def adjoint(self, terms, param_slots, n_params):
        t = as_terms(terms)
        ps = np.ascontiguousarray(param_slots, dtype=np.int32).reshape(-1)
        if ps.size != 3 * self.n_gates:
            raise ValueError("param_slots needs 3 entries per uploaded gate")
        grad = np.zeros(n_params)
        e = ctypes.c_double(0.0)
        _check(lib().svmi_adjoint(self._h, _ptr(t), len(t), _ptr(ps), int(n_params), ctypes.byref(e),
                                  _ptr(grad) if n_params else None))
        return e.value, grad

    def qir_run(self, n_qubits, gates):
        """QIR gates on the first 2^n_qubits amplitudes (svmi_qir_run).
        Returns (probs, outcomes, last_prob): P(1) and outcome per Measure."""
        g = as_qir_gates(gates)
        nm = int(np.count_nonzero(g["op"] == QIR_OP["Measure"]))
        probs = np.zeros(max(nm, 1))
        outs = np.zeros(max(nm, 1), dtype=np.int32)
        last = ctypes.c_double(0.0)
        _check(lib().svmi_qir_run(self._h, int(n_qubits), _ptr(g), len(g), _ptr(probs), _ptr(outs),
                                  ctypes.byref(last)))
        return probs[:nm], outs[:nm], last.value

    def dm_qir_run(self, n_qubits, gates):
        """QIR gates on the density matrix of n_qubits held in the first
        2^(2 n_qubits) amplitudes (svmi_dm_qir_run; the engine needs 2 n_qubits
        qubits).  Returns (probs, outcomes) per Measure."""
        g = as_qir_gates(gates)
        nm = int(np.count_nonzero(g["op"] == QIR_OP["Measure"]))
        probs = np.zeros(max(nm, 1))
        outs = np.zeros(max(nm, 1), dtype=np.int32)
        _check(lib().svmi_dm_qir_run(self._h, int(n_qubits), _ptr(g), len(g), _ptr(probs), _ptr(outs)))
        return probs[:nm], outs[:nm]

    def set_option(self, option, value):
        _check(lib().svmi_set_option(self._h, int(option), int(value)))

    def stats(self):
        s = Stats()
        _check(lib().svmi_get_stats(self._h, ctypes.byref(s)))
        return {k: getattr(s, k) for k, _ in Stats._fields_}

    def step_times(self):
        """Per-step device times (ms) of the last program run with profiling on."""
        n = ctypes.c_uint64(0)
        _check(lib().svmi_debug_step_times(self._h, None, 0, ctypes.byref(n)))
        ms = np.zeros(max(n.value, 1))
        _check(lib().svmi_debug_step_times(self._h, _ptr(ms), n.value, ctypes.byref(n)))
        return ms[:n.value]


def _splitmix64(z):
    z = (z + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(0xFFFFFFFFFFFFFFFF)
    return z ^ (z >> np.uint64(31))


def random_amplitudes(seed, idx):
    """Unnormalised amplitudes of Engine.set_random_state(seed) at indices idx
    (the same counter-based generator, evaluated on the host)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        s = np.uint64(int(seed) & 0xFFFFFFFFFFFFFFFF)
        u0 = ((_splitmix64(s + np.uint64(2) * idx + np.uint64(1)) >> np.uint64(11)) + np.uint64(1)).astype(np.float64)
        u1 = ((_splitmix64(s + np.uint64(2) * idx + np.uint64(2)) >> np.uint64(11)) + np.uint64(1)).astype(np.float64)
    u0 *= 2.0 ** -53
    u1 *= 2.0 ** -53
    r = np.sqrt(-2.0 * np.log(u0))
    return r * np.cos(2 * np.pi * u1), r * np.sin(2 * np.pi * u1)


def _batch_items(items):
    keep = []
    arr = (BatchItem * len(items))()
    total = 0
    for i, it in enumerate(items):
        g = as_gates(it["gates"])
        ps = np.ascontiguousarray(it["param_slots"], dtype=np.int32).reshape(-1)
        t = as_terms(it["terms"])
        keep += [g, ps, t]
        arr[i].n_qubits = int(it["n"])
        arr[i].gates = g.ctypes.data
        arr[i].n_gates = len(g)
        arr[i].param_slots = ps.ctypes.data if ps.size else None
        arr[i].n_params = int(it["n_params"])
        arr[i].terms = t.ctypes.data if t.size else None
        arr[i].n_terms = len(t)
        # planning overrides (0 = the autotuned table)
        arr[i].fwd_tile = int(it.get("fwd_tile", 0))
        arr[i].adj_tile = int(it.get("adj_tile", 0))
        arr[i].op_cost = int(it["op_cost"]) + 1 if it.get("op_cost") is not None else 0
        total += int(it["n_params"])
    return arr, keep, total


def batch_adjoint(items, device=0, _emulate=False):
    """items: list of dicts {n, gates, param_slots, n_params, terms} and,
    optionally, fwd_tile / adj_tile / op_cost (else the autotuned table).
    Returns (expvals, list of gradient arrays).  One launch per dependency
    level runs every circuit's forward pass, <H> and adjoint sweep."""
    arr, keep, total = _batch_items(items)
    ev = np.zeros(len(items))
    gr = np.zeros(max(total, 1))
    if _emulate:
        _check(lib().svmi_debug_emulate_batch(ctypes.cast(arr, _vp), len(items), _ptr(ev), _ptr(gr)))
    else:
        _check(lib().svmi_batch_adjoint(int(device), ctypes.cast(arr, _vp), len(items), _ptr(ev), _ptr(gr)))
    out, off = [], 0
    for it in items:
        out.append(gr[off:off + int(it["n_params"])].copy())
        off += int(it["n_params"])
    return ev, out


# ---- host checks of the planner (no device; see include/svmi_debug.h) ----

def fp64_rates(waves_per_simd, device=0):
    """(vector FMA, matrix-core) FP64 TFLOP/s measured on the device."""
    v, m = ctypes.c_double(), ctypes.c_double()
    _check(lib().svmi_debug_fp64_rates(int(device), int(waves_per_simd), ctypes.byref(v), ctypes.byref(m)))
    return v.value, m.value


def plan_batch(items):
    """{levels, launches, tasks, state_bytes} of the batch plan."""
    arr, keep, _ = _batch_items(items)
    v = [ctypes.c_uint64() for _ in range(4)]
    _check(lib().svmi_debug_plan_batch(ctypes.cast(arr, _vp), len(items), *[ctypes.byref(x) for x in v]))
    return dict(zip(("levels", "launches", "tasks", "state_bytes"), (x.value for x in v)))

def emulate(n, gates, state, fusion=True, tile_qubits=12, low_qubits=4, sweep_budget=0):
    g = as_gates(gates)
    re = np.array(state[0], dtype=np.float64, copy=True)
    im = np.array(state[1], dtype=np.float64, copy=True)
    _check(lib().svmi_debug_emulate(int(n), _ptr(g), len(g), int(fusion), int(tile_qubits), int(low_qubits),
                                    int(sweep_budget), _ptr(re), _ptr(im)))
    return re, im


def emulate_adjoint(n, gates, terms, param_slots, n_params, state, tile_qubits=12, low_qubits=4):
    g = as_gates(gates)
    t = as_terms(terms)
    ps = np.ascontiguousarray(param_slots, dtype=np.int32).reshape(-1)
    re = np.ascontiguousarray(state[0], dtype=np.float64)
    im = np.ascontiguousarray(state[1], dtype=np.float64)
    grad = np.zeros(n_params)
    e = ctypes.c_double(0.0)
    _check(lib().svmi_debug_emulate_adjoint(int(n), _ptr(g), len(g), _ptr(t), len(t), _ptr(ps), int(n_params),
                                            int(tile_qubits), int(low_qubits), _ptr(re), _ptr(im),
                                            ctypes.byref(e), _ptr(grad) if n_params else None))
    return e.value, grad


def emulate_batch(items):
    """Host execution of batch_adjoint's task queue (planner check)."""
    return batch_adjoint(items, _emulate=True)


def plan_adjoint(n, gates, param_slots, n_params, tile_qubits=11, low_qubits=4):
    """(passes, description) of the backward program of svmi_adjoint."""
    g = as_gates(gates)
    ps = np.ascontiguousarray(param_slots, dtype=np.int32).reshape(-1)
    buf = ctypes.create_string_buffer(1 << 20)
    passes = ctypes.c_uint64(0)
    _check(lib().svmi_debug_plan_adjoint(int(n), _ptr(g), len(g), _ptr(ps), int(n_params), int(tile_qubits),
                                         int(low_qubits), ctypes.byref(passes), buf, len(buf)))
    return passes.value, buf.value.decode()


def plan(n, gates, fusion=True, tile_qubits=12, low_qubits=4, sweep_budget=0):
    g = as_gates(gates)
    passes = ctypes.c_uint64(0)
    buf = ctypes.create_string_buffer(1 << 16)
    _check(lib().svmi_debug_plan(int(n), _ptr(g), len(g), int(fusion), int(tile_qubits), int(low_qubits),
                                 int(sweep_budget), ctypes.byref(passes), buf, len(buf)))
    return passes.value, buf.value.decode()


def emulate_qir(n, gates, state, tile_qubits=12, low_qubits=4):
    """Host execution of svmi_qir_run's program (planner check).
    Returns (re, im, probs, outcomes)."""
    g = as_qir_gates(gates)
    re = np.array(state[0], dtype=np.float64, copy=True)
    im = np.array(state[1], dtype=np.float64, copy=True)
    nm = int(np.count_nonzero(g["op"] == QIR_OP["Measure"]))
    probs = np.zeros(max(nm, 1))
    outs = np.zeros(max(nm, 1), dtype=np.int32)
    _check(lib().svmi_debug_emulate_qir(int(n), _ptr(g), len(g), int(tile_qubits), int(low_qubits), _ptr(re),
                                        _ptr(im), _ptr(probs), _ptr(outs)))
    return re, im, probs[:nm], outs[:nm]


def emulate_dm_qir(n, gates, state=None, tile_qubits=12, low_qubits=4):
    """Host execution of svmi_dm_qir_run's program on a density matrix of n
    qubits (2^(2n) entries, default |0..0><0..0|).  Returns (re, im, probs,
    outcomes)."""
    g = as_qir_gates(gates)
    dim = 1 << (2 * n)
    if state is None:
        re = np.zeros(dim)
        im = np.zeros(dim)
        re[0] = 1.0
    else:
        re = np.array(state[0], dtype=np.float64, copy=True)
        im = np.array(state[1], dtype=np.float64, copy=True)
    nm = int(np.count_nonzero(g["op"] == QIR_OP["Measure"]))
    probs = np.zeros(max(nm, 1))
    outs = np.zeros(max(nm, 1), dtype=np.int32)
    _check(lib().svmi_debug_emulate_dm_qir(int(n), _ptr(g), len(g), int(tile_qubits), int(low_qubits), _ptr(re),
                                           _ptr(im), _ptr(probs), _ptr(outs)))
    return re, im, probs[:nm], outs[:nm]


def plan_qir(n, gates, tile_qubits=12, low_qubits=4):
    """(sweeps, measurements, description) of svmi_qir_run's program."""
    g = as_qir_gates(gates)
    sw = ctypes.c_uint64(0)
    nm = ctypes.c_uint64(0)
    buf = ctypes.create_string_buffer(1 << 20)
    _check(lib().svmi_debug_plan_qir(int(n), _ptr(g), len(g), int(tile_qubits), int(low_qubits), ctypes.byref(sw),
                                     ctypes.byref(nm), buf, len(buf)))
    return sw.value, nm.value, buf.value.decode()


def jit_source(n, gates, tile_qubits=12, low_qubits=4, compile=False):
    """(source, compile_ms, code_bytes) of the specialised sweep kernels
    (compile=2: with the persistent double-buffered kernels of SVMI_OPT_PREFETCH)."""
    g = as_gates(gates)
    buf = ctypes.create_string_buffer(1 << 24)
    ms = ctypes.c_double(0.0)
    nb = ctypes.c_uint64(0)
    _check(lib().svmi_debug_jit(int(n), _ptr(g), len(g), int(tile_qubits), int(low_qubits), int(compile),
                                ctypes.byref(ms), ctypes.byref(nb), buf, len(buf)))
    return buf.value.decode(), ms.value, nb.value
