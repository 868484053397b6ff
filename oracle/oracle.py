"""ctypes bindings for the CPU oracle (test infrastructure only).

Loads ``oracle/libsvsim_oracle.so`` (the plain-C restatement of the reference
CPU-OpenMP engine, svsim/src/svsim_cpu_omp.hpp) and, when it was built here,
``oracle/_ref/libsvsim_ref.so`` (the reference engine itself, compiled from
its sources by ``oracle/Makefile``).

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg use
this module, always as the checker; the product never imports it.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "libsvsim_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libsvsim_ref.so")
REF_QIR_SO = os.path.join(HERE, "_ref", "libsvsim_qir_ref.so")
REF_DM_SO = os.path.join(HERE, "_ref", "libsvsim_dm_ref.so")

# Gate record layout shared by the oracle, the reference driver and svmi.h:
# {int32 op; int32 qb[5]; double theta, phi, lambda;}  -> 48 bytes.
GATE_DTYPE = np.dtype([("op", "<i4"), ("qb", "<i4", (5,)),
                       ("theta", "<f8"), ("phi", "<f8"), ("lambda", "<f8")])
PAULI_DTYPE = np.dtype([("coeff", "<f8"), ("xmask", "<u8"), ("zmask", "<u8")])
# QIR gate record (svsim_qsharp/svsim_cpu_omp.hpp:113-147; svmi_qir_gate):
# {int32 op; int32 qubit; uint64 mask; double theta;} -> 24 bytes.
QIR_GATE_DTYPE = np.dtype([("op", "<i4"), ("qubit", "<i4"), ("mask", "<u8"), ("theta", "<f8")])

# QIR op numbering (svsim_qsharp/svsim_cpu_omp.hpp:45-59).
QIR_OPS = ["X", "Y", "Z", "H", "S", "T", "RI", "RX", "RY", "RZ", "EI", "EX", "EY", "EZ",
           "ControlledX", "ControlledY", "ControlledZ", "ControlledH", "ControlledS", "ControlledT",
           "ControlledRI", "ControlledRX", "ControlledRY", "ControlledRZ",
           "ControlledEI", "ControlledEX", "ControlledEY", "ControlledEZ",
           "AdjointS", "AdjointT", "ControlledAdjointS", "ControlledAdjointT", "Swap", "Measure"]
QIR = {name: i for i, name in enumerate(QIR_OPS)}

# libc RAND_MAX on glibc; the reference draws r = rand()/RAND_MAX (cpu_omp:351).
RAND_MAX = 2147483647

_dp = ctypes.POINTER(ctypes.c_double)
_vp = ctypes.c_void_p


def build(quiet=True):
    """Compile the oracle (and oracle/_ref where the reference is present)."""
    import subprocess
    out = subprocess.run(["make", "-s", "-C", HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout, end="")


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        L.orc_apply_gates.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int]
        L.orc_apply_gates.restype = ctypes.c_int
        L.orc_measure.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp]
        L.orc_measure.restype = ctypes.c_double
        L.orc_measure_linear.argtypes = L.orc_measure.argtypes
        L.orc_measure_linear.restype = ctypes.c_double
        L.orc_prob_scan.argtypes = [ctypes.c_int, _vp, _vp, _vp]
        L.orc_prob_scan.restype = None
        L.orc_expval.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, ctypes.c_int]
        L.orc_expval.restype = ctypes.c_double
        L.orc_apply_pauli_sum.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp]
        L.orc_apply_pauli_sum.restype = None
        L.orc_gate_arity.argtypes = [ctypes.c_int]
        L.orc_gate_arity.restype = ctypes.c_int
        L.orc_qir_apply.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp, _vp, ctypes.c_int]
        L.orc_qir_apply.restype = ctypes.c_int
        _lib = L
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref_qir_available():
    return os.path.exists(REF_QIR_SO)


def ref():
    """The reference engine compiled from /root/reference (None if absent)."""
    global _ref
    if _ref is None:
        if not ref_available():
            return None
        L = ctypes.CDLL(REF_SO)
        L.ref_run.argtypes = [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_size_t, _vp, _vp, _vp, _vp]
        L.ref_run.restype = ctypes.c_int
        L.ref_measure.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp]
        L.ref_measure.restype = ctypes.c_int
        L.ref_rand_max.restype = ctypes.c_int
        L.ref_sim_times.argtypes = [ctypes.c_int, ctypes.c_int, _vp, ctypes.c_size_t, ctypes.c_int, _vp]
        L.ref_sim_times.restype = ctypes.c_int
        _ref = L
    return _ref


def _ptr(a):
    return a.ctypes.data_as(_vp)


def as_gates(gates):
    """Accept a GATE_DTYPE array or a list of (op, qubits, theta, phi, lambda)."""
    if isinstance(gates, np.ndarray) and gates.dtype == GATE_DTYPE:
        return np.ascontiguousarray(gates)
    arr = np.zeros(len(gates), dtype=GATE_DTYPE)
    for i, g in enumerate(gates):
        op, qs = g[0], list(g[1])
        arr[i]["op"] = op
        arr[i]["qb"][: len(qs)] = qs
        arr[i]["theta"], arr[i]["phi"], arr[i]["lambda"] = (list(g[2:5]) + [0.0, 0.0, 0.0])[:3]
    return arr


def zero_state(n):
    re = np.zeros(1 << n)
    im = np.zeros(1 << n)
    re[0] = 1.0
    return re, im


def run(n, gates, state=None, nthreads=0):
    """Apply gates with the oracle; returns (re, im) (new arrays)."""
    g = as_gates(gates)
    if state is None:
        re, im = zero_state(n)
    else:
        re = np.array(state[0], dtype=np.float64, copy=True)
        im = np.array(state[1], dtype=np.float64, copy=True)
    rc = lib().orc_apply_gates(n, _ptr(re), _ptr(im), _ptr(g), len(g), int(nthreads))
    if rc != 0:
        raise ValueError("oracle rejected the circuit (bad op or qubit)")
    return re, im


def measure(n, re, im, uniforms, linear=False):
    re = np.ascontiguousarray(re, dtype=np.float64)
    im = np.ascontiguousarray(im, dtype=np.float64)
    u = np.ascontiguousarray(uniforms, dtype=np.float64)
    out = np.zeros(len(u), dtype=np.uint64)
    f = lib().orc_measure_linear if linear else lib().orc_measure
    f(n, _ptr(re), _ptr(im), _ptr(u), len(u), _ptr(out))
    return out


def prob_scan(n, re, im):
    re = np.ascontiguousarray(re, dtype=np.float64)
    im = np.ascontiguousarray(im, dtype=np.float64)
    scan = np.zeros((1 << n) + 1)
    lib().orc_prob_scan(n, _ptr(re), _ptr(im), _ptr(scan))
    return scan


def as_terms(terms):
    if isinstance(terms, np.ndarray) and terms.dtype == PAULI_DTYPE:
        return np.ascontiguousarray(terms)
    arr = np.zeros(len(terms), dtype=PAULI_DTYPE)
    for i, (c, x, z) in enumerate(terms):
        arr[i] = (c, x, z)
    return arr


def expval(n, re, im, terms, nthreads=0):
    t = as_terms(terms)
    re = np.ascontiguousarray(re, dtype=np.float64)
    im = np.ascontiguousarray(im, dtype=np.float64)
    return lib().orc_expval(n, _ptr(re), _ptr(im), _ptr(t), len(t), int(nthreads))


def apply_pauli_sum(n, re, im, terms):
    t = as_terms(terms)
    re = np.ascontiguousarray(re, dtype=np.float64)
    im = np.ascontiguousarray(im, dtype=np.float64)
    ore = np.zeros(1 << n)
    oim = np.zeros(1 << n)
    lib().orc_apply_pauli_sum(n, _ptr(re), _ptr(im), _ptr(t), len(t), _ptr(ore), _ptr(oim))
    return ore, oim


def param_shift_grad(n, gates, param_slots, n_params, terms, state=None, nthreads=0):
    """Parameter-shift gradient of <H> w.r.t. the bound angle slots.

    param_slots[g, s] = parameter index of slot s (0=theta, 1=phi, 2=lambda)
    of gate g, or -1.  Exact (two-term rule, shift pi/2) for gates whose
    generator has two eigenvalues one apart in the bound slot: RX/RY/U3 theta
    (angle theta/2), RZ/U1/U2/U3 phi and lambda, CU1 lambda, RZZ theta.
    """
    g0 = as_gates(gates)
    slots = np.asarray(param_slots).reshape(len(g0), 3)
    names = ("theta", "phi", "lambda")
    grad = np.zeros(n_params)
    for gi in range(len(g0)):
        for s in range(3):
            p = slots[gi, s]
            if p < 0:
                continue
            vals = []
            for sh in (np.pi / 2, -np.pi / 2):
                g = g0.copy()
                g[gi][names[s]] = g[gi][names[s]] + sh
                re, im = run(n, g, state, nthreads)
                vals.append(expval(n, re, im, terms, nthreads))
            grad[p] += 0.5 * (vals[0] - vals[1])
    return grad


# ---- the reference engine itself (oracle/_ref) ----

def ref_run(n, gates, state=None, nthreads=1):
    L = ref()
    if L is None:
        raise RuntimeError("oracle/_ref not built (reference sources absent)")
    g = as_gates(gates)
    ore = np.zeros(1 << n)
    oim = np.zeros(1 << n)
    if state is None:
        rc = L.ref_run(n, nthreads, _ptr(g), len(g), None, None, _ptr(ore), _ptr(oim))
    else:
        ire = np.ascontiguousarray(state[0], dtype=np.float64)
        iim = np.ascontiguousarray(state[1], dtype=np.float64)
        rc = L.ref_run(n, nthreads, _ptr(g), len(g), _ptr(ire), _ptr(iim), _ptr(ore), _ptr(oim))
    if rc != 0:
        raise ValueError("reference driver rejected the circuit")
    return ore, oim


def ref_sim_times(n, gates, reps, nthreads=1):
    """Milliseconds of Simulation::sim() alone for `reps` runs from |0..0>
    (what the reference itself times, svsim/src/svsim_cpu_omp.hpp:288-319)."""
    L = ref()
    if L is None:
        raise RuntimeError("oracle/_ref not built (reference sources absent)")
    g = as_gates(gates)
    ms = np.zeros(int(reps))
    if L.ref_sim_times(n, int(nthreads), _ptr(g), len(g), int(reps), _ptr(ms)) != 0:
        raise ValueError("reference driver rejected the circuit")
    return ms


def ref_measure(n, re, im, draws):
    """Reference Simulation::measure with integer draws fed to rand()."""
    L = ref()
    if L is None:
        raise RuntimeError("oracle/_ref not built")
    re = np.ascontiguousarray(re, dtype=np.float64)
    im = np.ascontiguousarray(im, dtype=np.float64)
    d = np.ascontiguousarray(draws, dtype=np.int32)
    out = np.zeros(len(d), dtype=np.uint64)
    rc = L.ref_measure(n, _ptr(re), _ptr(im), _ptr(d), len(d), _ptr(out))
    if rc != 0:
        raise ValueError("reference measure failed")
    return out


# ---- QIR gate set: multi-controlled gates and measure-and-collapse ----

def as_qir_gates(gates):
    """Accept a QIR_GATE_DTYPE array or a list of (op, qubit, mask, theta)."""
    if isinstance(gates, np.ndarray) and gates.dtype == QIR_GATE_DTYPE:
        return np.ascontiguousarray(gates)
    arr = np.zeros(len(gates), dtype=QIR_GATE_DTYPE)
    for i, g in enumerate(gates):
        op = QIR[g[0]] if isinstance(g[0], str) else int(g[0])
        arr[i] = (op, int(g[1]), int(g[2]) if len(g) > 2 else 0, float(g[3]) if len(g) > 3 else 0.0)
    return arr


def n_measures(gates):
    return int(np.count_nonzero(as_qir_gates(gates)["op"] == QIR["Measure"]))


def qir_apply(n, re, im, gates, last_prob=0.0, nthreads=0):
    """Apply QIR gates to the first 2^n amplitudes of (re, im) in place with
    the oracle.  Returns (probs, outcomes, last_prob) for the Measure records."""
    g = as_qir_gates(gates)
    assert re.dtype == np.float64 and im.dtype == np.float64 and re.flags.c_contiguous and im.flags.c_contiguous
    assert len(re) >= (1 << n) and len(im) >= (1 << n)
    nm = int(np.count_nonzero(g["op"] == QIR["Measure"]))
    probs = np.zeros(max(nm, 1))
    outs = np.zeros(max(nm, 1), dtype=np.int32)
    lp = np.array([last_prob], dtype=np.float64)
    rc = lib().orc_qir_apply(n, _ptr(re), _ptr(im), _ptr(g), len(g), _ptr(lp), _ptr(probs), _ptr(outs),
                             int(nthreads))
    if rc != 0:
        raise ValueError("oracle rejected the QIR gate list (bad op, target or swap qubit)")
    return probs[:nm], outs[:nm], float(lp[0])


def ref_qir_run(capacity, n, gates, re, im, nthreads=1):
    """The reference QIR engine (oracle/_ref/libsvsim_qir_ref.so) on a copy of
    the 2^capacity-amplitude state with n allocated qubits.
    Returns (re, im, probs, outcomes)."""
    global _ref_qir
    if _ref_qir is None:
        if not ref_qir_available():
            raise RuntimeError("oracle/_ref QIR engine not built (reference sources absent)")
        L = ctypes.CDLL(REF_QIR_SO)
        L.ref_qir_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_size_t, _vp, _vp,
                                  _vp, _vp]
        L.ref_qir_run.restype = ctypes.c_int
        _ref_qir = L
    g = as_qir_gates(gates)
    ore = np.array(re, dtype=np.float64, copy=True)
    oim = np.array(im, dtype=np.float64, copy=True)
    assert len(ore) == 1 << capacity
    nm = int(np.count_nonzero(g["op"] == QIR["Measure"]))
    probs = np.zeros(max(nm, 1))
    outs = np.zeros(max(nm, 1), dtype=np.int32)
    rc = _ref_qir.ref_qir_run(capacity, n, nthreads, _ptr(g), len(g), _ptr(ore), _ptr(oim), _ptr(probs),
                              _ptr(outs))
    if rc != 0:
        raise ValueError("reference QIR driver rejected the gate list")
    return ore, oim, probs[:nm], outs[:nm]


_ref_qir = None


# ---- density-matrix engine (svsim_qsharp_noise/svsim_cpu_omp.hpp) ----

def dm_qir_apply(n, gates, state=None, nthreads=0):
    """The density-matrix engine on n qubits with the oracle: rho starts as
    |0..0><0..0| (or the given 2^(2n)-entry (re, im), index row | (col << n)).
    Returns (re, im, probs, outcomes)."""
    g = as_qir_gates(gates)
    dim = 1 << (2 * n)
    if state is None:
        re = np.zeros(dim)
        im = np.zeros(dim)
        re[0] = 1.0
    else:
        re = np.array(state[0], dtype=np.float64, copy=True)
        im = np.array(state[1], dtype=np.float64, copy=True)
        assert len(re) == dim and len(im) == dim
    nm = int(np.count_nonzero(g["op"] == QIR["Measure"]))
    probs = np.zeros(max(nm, 1))
    outs = np.zeros(max(nm, 1), dtype=np.int32)
    L = lib()
    L.orc_dm_qir_apply.argtypes = [ctypes.c_int, _vp, _vp, _vp, ctypes.c_size_t, _vp, _vp, ctypes.c_int]
    L.orc_dm_qir_apply.restype = ctypes.c_int
    rc = L.orc_dm_qir_apply(n, _ptr(re), _ptr(im), _ptr(g), len(g), _ptr(probs), _ptr(outs), int(nthreads))
    if rc != 0:
        raise ValueError("oracle rejected the QIR gate list (bad op, target or swap qubit)")
    return re, im, probs[:nm], outs[:nm]


def ref_dm_available():
    return os.path.exists(REF_DM_SO)


_ref_dm = None


def ref_dm_run(n, gates, nthreads=1):
    """The reference density-matrix engine (oracle/_ref/libsvsim_dm_ref.so)
    on n qubits from |0..0><0..0| (arrays of 2^(2n) entries).
    Returns (re, im, probs, outcomes)."""
    global _ref_dm
    if _ref_dm is None:
        if not ref_dm_available():
            raise RuntimeError("oracle/_ref density-matrix engine not built (reference sources absent)")
        L = ctypes.CDLL(REF_DM_SO)
        L.ref_qir_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_size_t, _vp, _vp,
                                  _vp, _vp]
        L.ref_qir_run.restype = ctypes.c_int
        _ref_dm = L
    g = as_qir_gates(gates)
    cap = 2 * n
    ore = np.zeros(1 << cap)
    oim = np.zeros(1 << cap)
    ore[0] = 1.0
    nm = int(np.count_nonzero(g["op"] == QIR["Measure"]))
    probs = np.zeros(max(nm, 1))
    outs = np.zeros(max(nm, 1), dtype=np.int32)
    rc = _ref_dm.ref_qir_run(cap, n, nthreads, _ptr(g), len(g), _ptr(ore), _ptr(oim), _ptr(probs), _ptr(outs))
    if rc != 0:
        raise ValueError("reference density-matrix driver rejected the gate list")
    return ore, oim, probs[:nm], outs[:nm]

