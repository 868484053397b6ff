"""The C++ Simulation facade (sv-sim_amd/csrc/svsim_amd.hpp) on the GPU.

measure(rep) seeds with srand(RAND_SEED) as the reference does
(svsim/src/svsim_cpu_omp.hpp:347, config.hpp:44), so a program built with
-DRAND_SEED=<constant> draws the same uniforms on every call: two measure(5)
calls return the same states, and those states are the oracle's measure rule
applied to the uniforms rand()/RAND_MAX of that seed.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "sv-sim_amd")
sys.path.insert(0, ROOT)
sys.path.insert(0, PKG)

from oracle import oracle as O  # noqa: E402
from svmi import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

PROGRAM = r"""
#include <cstdio>
#include <cstdlib>
#include "svsim_amd.hpp"
using namespace SVSim;
int main()
{
    const int n = 6;
    Simulation sim(n, 0);
    for (int q = 0; q < n; q++) sim.append(Simulation::H(q));
    sim.append(Simulation::RY(0.7, 2));
    sim.append(Simulation::CX(2, 5));
    sim.append(Simulation::U3(0.3, 1.1, -0.4, 4));
    sim.upload();
    sim.sim();
    for (int call = 0; call < 2; call++) {
        IdxType* res = sim.measure(5);
        printf("call%d:", call);
        for (int i = 0; i < 5; i++) printf(" %llu", (unsigned long long)res[i]);
        printf("\n");
        delete[] res;
    }
    srand(unsigned(RAND_SEED));
    printf("draws:");
    for (int i = 0; i < 5; i++) printf(" %.17g", double(rand()) / double(RAND_MAX));
    printf("\n");
    return 0;
}
"""


def test_measure_uses_rand_seed(tmp_path):
    cpp, exe = tmp_path / "seeded.cpp", tmp_path / "seeded_bin"
    cpp.write_text(PROGRAM)
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-DRAND_SEED=1", str(cpp), "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(PKG, "csrc"), "-L", PKG, "-lsvmi", "-Wl,-rpath," + PKG, "-o", str(exe)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    print("case: facade measure with RAND_SEED=1", flush=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=90, env=dict(os.environ, SVSIM_QUIET="1"))
    assert r.returncode == 0, r.stderr
    lines = {ln.split(":")[0]: ln.split(":")[1].split() for ln in r.stdout.splitlines() if ":" in ln}
    c0 = [int(x) for x in lines["call0"]]
    c1 = [int(x) for x in lines["call1"]]
    assert c0 == c1
    u = np.array([float(x) for x in lines["draws"]])
    OP = W.OP
    g = np.zeros(9, dtype=W.GATE_DTYPE)
    for q in range(6):
        g[q]["op"] = OP["H"]
        g[q]["qb"][0] = q
    g[6]["op"], g[6]["qb"][0], g[6]["theta"] = OP["RY"], 2, 0.7
    g[7]["op"], g[7]["qb"][0], g[7]["qb"][1] = OP["CX"], 2, 5
    g[8]["op"], g[8]["qb"][0] = OP["U3"], 4
    g[8]["theta"], g[8]["phi"], g[8]["lambda"] = 0.3, 1.1, -0.4
    ore, oim = O.run(6, g)
    assert c0 == [int(x) for x in O.measure(6, ore, oim, u)]
