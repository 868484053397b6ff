"""Copy the QASMBench circuits and their published counts into tests/golden/.

Inputs (read where /root/reference exists):
  svsim/QASMBench/{small,medium,large}/<name>/<name>.qasm -- the benchmark
      circuits the reference's translator (svsim/QASMBench/svsim_qasm.py) is
      run on; copied unchanged to tests/golden/qasm/ (bwt_n21, 1.4 MB, is left
      out);
  svsim/QASMBench/README.md -- its tables list qubits, gates and CNOTs per
      benchmark as counted by the translator's tables; written to
      tests/golden/qasmbench_counts.json as {"<name>_n<qubits>": {qubits,
      gates, cnot}}.

Usage:  python tests/golden/make_qasm_golden.py
"""
import json
import os
import re
import shutil
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/svsim/QASMBench"
SKIP = {"bwt_n21"}


def main():
    if not os.path.isdir(SRC):
        sys.exit("QASMBench sources not present")
    dst = os.path.join(HERE, "qasm")
    os.makedirs(dst, exist_ok=True)
    names = []
    for size in ("small", "medium", "large"):
        for name in sorted(os.listdir(os.path.join(SRC, size))):
            f = os.path.join(SRC, size, name, name + ".qasm")
            if name in SKIP or not os.path.exists(f):
                continue
            shutil.copyfile(f, os.path.join(dst, name + ".qasm"))
            names.append(name)
    counts = {}
    row = re.compile(r"^\|\s*([A-Za-z0-9_]+)\s*\|.*\|\s*(\d+)\s*\|\s*(\d+)\s*\|\s*(\d+)\s*\|[^|]*\|\s*$")
    for line in open(os.path.join(SRC, "README.md"), encoding="utf-8"):
        m = row.match(line.strip())
        if m:
            key = f"{m.group(1)}_n{m.group(2)}"
            counts[key] = {"qubits": int(m.group(2)), "gates": int(m.group(3)), "cnot": int(m.group(4))}
    with open(os.path.join(HERE, "qasmbench_counts.json"), "w") as f:
        json.dump(counts, f, indent=1, sort_keys=True)
    print(f"copied {len(names)} circuits; {len(counts)} README rows")


if __name__ == "__main__":
    main()
