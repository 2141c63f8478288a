"""Builds the CPU checker (liboracle.so) with g++ (-O2, -ffp-contract=off, no fast-math).

Test infrastructure only: loaded by tests/, the smoke check and the CPU-baseline leg of bench.py.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-pthread",
           "-o", OUT + ".tmp", os.path.join(HERE, "oracle.cpp")]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("g++ failed:\n" + res.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
