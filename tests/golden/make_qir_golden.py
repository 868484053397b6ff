"""Write the QIR-gate-set golden fixtures in tests/golden/ from the REFERENCE.

Runs the reference QIR-facing engine (svsim_qsharp/svsim_cpu_omp.hpp),
compiled from its sources into oracle/_ref/libsvsim_qir_ref.so by
oracle/Makefile, on seeded inputs and stores inputs + outputs:

  qir_kats.npz   every QIR op (34) at several placements and control masks
                 (no control, 1-3 controls, a mask naming the target, a mask
                 bit above the allocated qubits) on seeded random states of
                 6 slots with 6 or 5 allocated qubits; Measure in the Z, X
                 and Y bases with draws on both sides of P(1): final
                 amplitudes (all 64 slots), P(1) and the outcome
  qir_program_n8.npz   a 240-gate 8-qubit program with 12 mid-circuit
                 measurements from |0..0>: final amplitudes, every P(1) and
                 outcome
  qir_hashes.json      SHA-256 of the reference amplitudes (float64 re||im) and
                 the P(1) list for larger seeded programs (n = 12, 14)

Usage (in a container that has /root/reference):
    make -C oracle && python tests/golden/make_qir_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

# ops whose `mask` field is a control mask
CONTROLLED = set(range(O.QIR["ControlledX"], O.QIR["ControlledEZ"] + 1)) | {
    O.QIR["ControlledAdjointS"], O.QIR["ControlledAdjointT"]}


def random_state(n, rng):
    re = rng.standard_normal(1 << n)
    im = rng.standard_normal(1 << n)
    s = np.sqrt(np.sum(re * re + im * im))
    return re / s, im / s


def sha(re, im, probs):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(re, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(im, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(probs, dtype="<f8").tobytes())
    return h.hexdigest()


def random_program(n, n_gates, n_meas, rng):
    """Random QIR program: every op kind, 0-3 controls, measurements at
    random positions in random bases with uniform draws."""
    gates = []
    meas_at = set(rng.choice(np.arange(5, n_gates), size=n_meas, replace=False).tolist())
    for i in range(n_gates):
        if i in meas_at:
            gates.append((O.QIR["Measure"], int(rng.integers(n)), int(rng.integers(3)), float(rng.random())))
            continue
        op = int(rng.integers(0, O.QIR["Swap"] + 1))
        q = int(rng.integers(n))
        mask = 0
        if op in CONTROLLED:
            others = [x for x in range(n) if x != q]
            for c in rng.choice(others, size=min(int(rng.integers(1, 4)), len(others)), replace=False):
                mask |= 1 << int(c)
        if op == O.QIR["Swap"]:
            mask = int(rng.choice([x for x in range(n) if x != q]))
        gates.append((op, q, mask, float(rng.uniform(-2 * np.pi, 2 * np.pi))))
    return O.as_qir_gates(gates)


def kats():
    rng = np.random.default_rng(11)
    cap = 6
    rows, ns, in_re, in_im, out_re, out_im, probs, outs = [], [], [], [], [], [], [], []
    for op in range(len(O.QIR_OPS)):
        variants = 6 if op in CONTROLLED else 4
        for v in range(variants):
            n = cap if v % 2 == 0 else cap - 1
            q = int(rng.integers(n))
            mask = 0
            theta = float(rng.uniform(-2 * np.pi, 2 * np.pi))
            if op in CONTROLLED:
                others = [x for x in range(n) if x != q]
                if v == 4:
                    mask = (1 << q) | (1 << others[0])          # names the target: identity
                elif v == 5:
                    mask = (1 << (cap - 1)) | (1 << others[0])  # bit above the 5 allocated
                    n = cap - 1
                    q = min(q, n - 1)
                else:
                    for c in rng.choice(others, size=v % 3 + 1, replace=False):
                        mask |= 1 << int(c)
            elif op == O.QIR["Swap"]:
                mask = int(rng.choice([x for x in range(n) if x != q]))
            elif op == O.QIR["Measure"]:
                mask = v % 3
                theta = [0.05, 0.95, 0.5, 0.3][v]
            g = O.as_qir_gates([(op, q, mask, theta)])
            re, im = random_state(cap, rng)
            if n < cap:   # unallocated slots hold zeros, as after a release of a |0> qubit
                re[1 << n:] = 0.0
                im[1 << n:] = 0.0
            ore, oim, p, o = O.ref_qir_run(cap, n, g, re, im, nthreads=2)
            rows.append(g[0])
            ns.append(n)
            in_re.append(re)
            in_im.append(im)
            out_re.append(ore)
            out_im.append(oim)
            probs.append(p[0] if len(p) else np.nan)
            outs.append(o[0] if len(o) else -1)
    np.savez_compressed(os.path.join(HERE, "qir_kats.npz"), gates=np.array(rows, dtype=O.QIR_GATE_DTYPE),
                        n=np.array(ns, dtype=np.int32), capacity=np.int32(cap), in_re=np.array(in_re),
                        in_im=np.array(in_im), out_re=np.array(out_re), out_im=np.array(out_im),
                        prob=np.array(probs), outcome=np.array(outs, dtype=np.int32))


def program_n8():
    rng = np.random.default_rng(23)
    n = 8
    g = random_program(n, 240, 12, rng)
    re = np.zeros(1 << n)
    im = np.zeros(1 << n)
    re[0] = 1.0
    ore, oim, p, o = O.ref_qir_run(n, n, g, re, im, nthreads=2)
    np.savez_compressed(os.path.join(HERE, "qir_program_n8.npz"), gates=g, re=ore, im=oim, prob=p, outcome=o)


def hashes():
    out = {}
    for n, ng, nm, seed in [(12, 600, 16, 31), (14, 800, 20, 37)]:
        rng = np.random.default_rng(seed)
        g = random_program(n, ng, nm, rng)
        re, im = random_state(n, rng)
        ore, oim, p, o = O.ref_qir_run(n, n, g, re, im, nthreads=4)
        out[f"qir_random_n{n}"] = {"seed": seed, "n_gates": ng, "n_measure": nm, "sha256": sha(ore, oim, p),
                                   "outcomes": [int(x) for x in o]}
    with open(os.path.join(HERE, "qir_hashes.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    if not O.ref_qir_available():
        sys.exit("oracle/_ref/libsvsim_qir_ref.so missing: run `make -C oracle` where /root/reference exists")
    kats()
    program_n8()
    hashes()
    print("wrote qir_kats.npz, qir_program_n8.npz, qir_hashes.json")
