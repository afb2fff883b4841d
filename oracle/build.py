"""Build recipe for the oracle's C restatement (test infrastructure only).

Compiles ``oracle/philox_cms.c`` with gcc into ``oracle/liboracle.so``.
The reference is pure Python (Numba JIT kernels, no C/C++ sources), so there
is no ``oracle/_ref`` build: nothing of the reference is compiled here.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SOURCE = os.path.join(HERE, "philox_cms.c")
TARGET = os.path.join(HERE, "liboracle.so")


def build(verbose=False):
    cmd = ["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-fno-fast-math",
           "-ffp-contract=off", SOURCE, "-o", TARGET + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(TARGET + ".tmp", TARGET)
    return TARGET


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
