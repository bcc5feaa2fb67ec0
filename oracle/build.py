"""Compile oracle/pf5d_oracle_c.c into oracle/_build/liboracle_pf5d.so (test infrastructure).

The reference is pure Python (Numba), so there is no reference C source to
compile into oracle/_ref; this builds the project's own C restatement.
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "pf5d_oracle_c.c")
OUT_DIR = os.path.join(HERE, "_build")
OUT = os.path.join(OUT_DIR, "liboracle_pf5d.so")
FLAGS = ["-O3", "-fno-fast-math", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared"]


def build():
    os.makedirs(OUT_DIR, exist_ok=True)
    tmp = OUT + ".tmp"
    subprocess.run(["gcc"] + FLAGS + ["-o", tmp, SRC, "-lm"], check=True)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build())
