"""Build recipe of the CPU oracle (test infrastructure only).

gcc, -O2, no FMA contraction, no fast-math: the f32/f64 operation sequence
of mvs_oracle.c is executed exactly as written, like the reference's Rust.
The reference itself is Rust and cannot be built here (no Rust toolchain),
so there is no oracle/_ref build.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "mvs_oracle.c")
LIB = os.path.join(HERE, "libmvs_oracle.so")
CFLAGS = ["-O2", "-std=gnu11", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def build_oracle() -> str:
    tmp = LIB + ".tmp"
    cmd = ["gcc"] + CFLAGS + [SRC, "-o", tmp, "-lm", "-lpthread"]
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + r.stdout)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build_oracle())
