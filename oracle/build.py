"""Build the oracle's C restatement (oracle/cpu_rl.c) into oracle/libmlf_cpu.so.

Test infrastructure: loaded only by tests/ and bench.py's cpu_baseline leg.
Plain gcc with OpenMP; AVX2/FMA code generation (any x86-64 host of an
MI355X node has them).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cpu_rl.c")
OUT = os.path.join(HERE, "libmlf_cpu.so")


def build(verbose: bool = False) -> str:
    tmp = OUT + ".tmp"
    cmd = ["gcc", "-O3", "-mavx2", "-mfma", "-fopenmp", "-fPIC", "-shared", "-std=c11", "-Wall",
           SRC, "-o", tmp, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gcc failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
