"""Drop-in check of the C++ facade (sv-sim_amd/csrc/svsim_amd.hpp): the
reference's single-node C++ examples (svsim/example/adder_n10_{amdgpu_sin,
amdgpu_omp,cpu_sin,cpu_omp}.cpp) build against it with only their engine
#includes replaced -- Simulation(n), Simulation::OP factories, append,
upload, sim, `auto* res = sim.measure(5)`, print_measurement, RAND_SEED.

The examples are read where the reference tree is present (this test skips
elsewhere) and compiled from a temporary copy; nothing of them is kept.  The
MPI / SHMEM examples need the multi-node engines, which this library does
not provide.  Running the same adder on the GPU through the facade is
tests/test_gpu_qasm.py::test_cplus_translator_program_runs.
"""
import os
import re
import subprocess

import pytest

REF_EXAMPLES = "/root/reference/svsim/example"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NAMES = ["adder_n10_amdgpu_sin", "adder_n10_amdgpu_omp", "adder_n10_cpu_sin", "adder_n10_cpu_omp"]


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference tree not present")
@pytest.mark.parametrize("name", NAMES)
def test_reference_example_builds_against_the_facade(name, tmp_path):
    with open(os.path.join(REF_EXAMPLES, name + ".cpp")) as f:
        src = f.read()
    includes = re.findall(r'#include "\.\./src/[^"]+"', src)
    assert includes, "the example includes the reference engine"
    for inc in includes:
        src = src.replace(inc, '#include "svsim_amd.hpp"' if "/svsim_" in inc else "")
    cpp = tmp_path / (name + ".cpp")
    cpp.write_text(src)
    r = subprocess.run(["g++", "-std=c++17", "-O1", str(cpp), "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "sv-sim_amd", "csrc"), "-L", os.path.join(ROOT, "sv-sim_amd"),
                        "-lsvmi", "-o", str(tmp_path / name)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
