"""Write the golden fixtures in tests/golden/ from the REFERENCE engine.

Runs the reference CPU-OpenMP engine (svsim/src/svsim_cpu_omp.hpp), compiled
from its sources into oracle/_ref/libsvsim_ref.so by oracle/Makefile, on
seeded inputs and stores inputs + outputs:

  gate_kats.npz      every reference op (38) at several qubit placements on
                     seeded random 5-qubit states: final amplitudes
  random_layers_n12.npz   the 12-qubit 40-layer random circuit from |0..0>:
                     all 4096 amplitudes, and 1024 measure() outcomes for fed
                     integer draws (the reference computes r = k/RAND_MAX)
  hashes.json        SHA-256 of the reference amplitudes (float64 re||im bytes)
                     for larger seeded circuits (u3 brick-wall, QFT on a random
                     input, hardware-efficient ansatz) at n <= 16

Usage (in a container that has /root/reference):
    make -C oracle && python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "sv-sim_amd"))

from oracle import oracle as O  # noqa: E402
from svmi import workloads as W  # noqa: E402

ARITY = [1, 1, 1, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 2, 2,
         3, 3, 2, 2, 2, 2, 2, 2, 2, 3, 4, 4, 4, 5, 1, 1, 1, 2]


def sha(re, im):
    h = hashlib.sha256()
    h.update(np.ascontiguousarray(re, dtype="<f8").tobytes())
    h.update(np.ascontiguousarray(im, dtype="<f8").tobytes())
    return h.hexdigest()


def gate_kats():
    rng = np.random.default_rng(7)
    n = 5
    rows, inits_re, inits_im, outs_re, outs_im = [], [], [], [], []
    for op in range(38):
        for _ in range(4):
            qs = list(rng.choice(n, size=ARITY[op], replace=False)) + [0] * (5 - ARITY[op])
            ang = rng.random(3) * 4 * np.pi - 2 * np.pi
            g = np.zeros(1, dtype=O.GATE_DTYPE)
            g[0]["op"] = op
            g[0]["qb"] = qs
            g[0]["theta"], g[0]["phi"], g[0]["lambda"] = ang
            re, im = W.random_state(n, int(rng.integers(0, 2 ** 31)))
            ore, oim = O.ref_run(n, g, (re, im), nthreads=2)
            rows.append(g[0])
            inits_re.append(re)
            inits_im.append(im)
            outs_re.append(ore)
            outs_im.append(oim)
    np.savez_compressed(os.path.join(HERE, "gate_kats.npz"), n=n, gates=np.array(rows, dtype=O.GATE_DTYPE),
                        init_re=np.array(inits_re), init_im=np.array(inits_im),
                        out_re=np.array(outs_re), out_im=np.array(outs_im))


def random_layers():
    n = 12
    g = W.random_layers(n=n, layers=40)
    re, im = O.ref_run(n, g, None, nthreads=4)
    k, _ = W.draws(1024, seed=2001)
    shots = O.ref_measure(n, re, im, k.astype(np.int32))
    np.savez_compressed(os.path.join(HERE, "random_layers_n12.npz"), n=n, gates=g, re=re, im=im,
                        draws=k, shots=shots)


def hashes():
    out = {}
    for n in (10, 14):
        g = W.u3_brickwall(n=n, layers=20)
        out[f"u3_brickwall_n{n}"] = sha(*O.ref_run(n, g, None, nthreads=4))
    for n in (10, 14):
        g = W.qft(n=n)
        st = W.random_state(n, seed=3000 + n)
        out[f"qft_n{n}_random_input"] = sha(*O.ref_run(n, g, st, nthreads=4))
    for n, L in ((12, 8), (16, 4)):
        g, _, _, _ = W.hea_ising(n=n, layers=L)
        out[f"hea_n{n}_l{L}"] = sha(*O.ref_run(n, g, None, nthreads=4))
    with open(os.path.join(HERE, "hashes.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    if not O.ref_available():
        sys.exit("oracle/_ref/libsvsim_ref.so missing: run make -C oracle where /root/reference exists")
    gate_kats()
    random_layers()
    hashes()
    print("fixtures written to", HERE)
