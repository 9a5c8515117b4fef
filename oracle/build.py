"""Build the C oracle into oracle/_build/liboracle.so (test infrastructure only).

    python oracle/build.py

gcc, -O2, no fast-math and no FP contraction, so each operation rounds once as in
the reference's eager float64 path.  OpenMP parallelises over batch members only;
every member's arithmetic is serial and identical to the single-thread order.
"""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
OUT = HERE / "_build" / "liboracle.so"


def build(verbose: bool = False) -> Path:
    OUT.parent.mkdir(exist_ok=True)
    cmd = ["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-fopenmp",
           str(HERE / "swe_oracle.c"), "-o", str(OUT) + ".tmp", "-lm"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    Path(str(OUT) + ".tmp").replace(OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
