"""svmi_dm_qir_run's program (density-matrix engine on the sweep machinery),
executed on the host by svmi_debug_emulate_dm_qir with the kernels' op
code, against the oracle restatement (itself bit-exact against the
reference engine, tests/test_dm_oracle.py) and the reference's golden
program."""
import os
import sys

import numpy as np
import pytest

import svmi
from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(__file__), "golden")
sys.path.insert(0, GOLD)
from make_qir_golden import random_program  # noqa: E402


def test_golden_program():
    d = np.load(os.path.join(GOLD, "dm_program_n4.npz"))
    re, im, probs, outs = svmi.emulate_dm_qir(4, d["gates"])
    assert max(np.abs(re - d["re"]).max(), np.abs(im - d["im"]).max()) < 1e-13
    assert np.abs(probs - d["probs"]).max() < 1e-13 and np.array_equal(outs, d["outcomes"])


@pytest.mark.parametrize("tile,low", [(12, 4), (6, 2), (9, 0)])
def test_random_programs_vs_oracle(tile, low):
    rng = np.random.default_rng(60 + tile)
    for _ in range(10):
        n = int(rng.integers(2, 7))
        g = random_program(n, 50, 3, rng)
        re, im, probs, outs = svmi.emulate_dm_qir(n, g, tile_qubits=tile, low_qubits=low)
        ore, oim, oprobs, oouts = O.dm_qir_apply(n, g)
        assert max(np.abs(re - ore).max(), np.abs(im - oim).max()) < 1e-12
        assert np.abs(probs - oprobs).max() < 1e-12 and np.array_equal(outs, oouts)


def test_limits():
    with pytest.raises(svmi.SvmiError):
        svmi.emulate_dm_qir(14, [(O.QIR["H"], 0, 0, 0.0)])
    with pytest.raises(svmi.SvmiError):
        svmi.emulate_dm_qir(2, [(O.QIR["Swap"], 0, 2, 0.0)])   # second qubit not allocated


def test_control_bit_on_a_column_qubit():
    """A control bit in [n, 2n) acts on the row half as a control on column
    qubit c - n and leaves the column half alone (the reference engine's
    OP_HEAD_MASK over the 2n-qubit state, svsim_qsharp_noise/svsim_cpu_omp.hpp:
    619-628): the result matches the oracle and is not Hermitian."""
    n = 3
    g = [(O.QIR["H"], 0, 0, 0.0), (O.QIR["H"], 1, 0, 0.0),
         (O.QIR["ControlledX"], 2, 1 << n, 0.0),            # control: column qubit 0
         (O.QIR["ControlledRY"], 1, (1 << n) | 1, 0.7)]     # controls: row 0 and column 0
    re, im, _, _ = svmi.emulate_dm_qir(n, g)
    ore, oim, _, _ = O.dm_qir_apply(n, g)
    assert max(np.abs(re - ore).max(), np.abs(im - oim).max()) < 1e-13
    rho = (ore + 1j * oim).reshape(1 << n, 1 << n).T       # rho[row, col], index row | col << n
    assert np.abs(rho - rho.conj().T).max() > 0.1
