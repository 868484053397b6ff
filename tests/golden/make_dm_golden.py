"""Golden vectors of the reference density-matrix engine
(svsim_qsharp_noise/svsim_cpu_omp.hpp, compiled from its sources into
oracle/_ref/libsvsim_dm_ref.so by oracle/Makefile).

Writes tests/golden/dm_program_n4.npz: a seeded random QIR program on 4
qubits (every gate kind, 1-3-qubit controls, Swap, 3 measurements in bases
Z/X/Y) and what the reference returns for it: rho as 2^8 entries (index
row | (col << 4)), P(1) and the outcome of each measurement.

Usage:  python tests/golden/make_dm_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.dirname(os.path.dirname(HERE)), HERE]
from oracle import oracle as O  # noqa: E402
from make_qir_golden import random_program  # noqa: E402


def main():
    if not O.ref_dm_available():
        sys.exit("oracle/_ref/libsvsim_dm_ref.so not built (reference sources absent)")
    rng = np.random.default_rng(4040)
    g = random_program(4, 60, 3, rng)
    re, im, probs, outs = O.ref_dm_run(4, g)
    np.savez(os.path.join(HERE, "dm_program_n4.npz"), gates=g, re=re, im=im, probs=probs, outcomes=outs)
    print("measurements", probs, outs)


if __name__ == "__main__":
    main()
