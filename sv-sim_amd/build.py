"""Build the engine in-tree for gfx950.

Outputs (next to this file):
  libsvmi.so                      HIP engine + C ABI (include/svmi.h)
  libsvsim.<ext-suffix>.so        pybind11 module with the reference's
                                  libsvsim surface (svsim/src/py_wrapper.cpp)

Usage: python sv-sim_amd/build.py [--jobs N] [--verbose]
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SVMI_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # branches on uniform (scalar) conditions -- the sweep kernels' op
    # dispatch -- stay unstructured: fewer register copies at the joins
    # (u3 sweep at 30 qubits 17.1 -> 14.8 ms on MI355X)
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
ENGINE_SOURCES = ["svmi_api.hip", "svmi_kernels.hip", "svmi_lower.cpp", "svmi_plan.cpp", "svmi_qir.cpp",
                  "svmi_qasm.cpp"]
LIB = os.path.join(HERE, "libsvmi.so")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile(src, verbose):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    _run([HIPCC, *HIP_FLAGS, *lang, "-c", path, "-o", obj], verbose)
    return obj


def build_engine(jobs=4, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), ENGINE_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    return LIB


def module_path():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(HERE, "libsvsim" + suffix)


def build_module(verbose=False):
    import pybind11
    out = module_path()
    tmp = out + ".tmp"
    py_inc = sysconfig.get_paths()["include"]
    cmd = ["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-I", pybind11.get_include(), "-I", py_inc, "-I", os.path.join(ROOT, "include"),
           os.path.join(CSRC, "py_wrapper.cpp"), "-o", tmp,
           "-L", HERE, "-lsvmi", "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build_all(jobs=4, verbose=False):
    build_engine(jobs, verbose)
    build_module(verbose)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("--verbose", action="store_true")
    a = ap.parse_args()
    try:
        build_all(a.jobs, a.verbose)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
    print("built", LIB, module_path())
