"""Build recipe for the oracle (test infrastructure): gcc -O2 -fopenmp on nfx_oracle.c.

Run ``python -m oracle.build`` from the repository root, or call :func:`build`.
"""

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libnfx_oracle.so")


def build(verbose: bool = False) -> str:
    src = os.path.join(HERE, "nfx_oracle.c")
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-Wall", "-o", LIB + ".tmp", src, "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
