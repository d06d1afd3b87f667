"""In-tree build of the HIP library (gfx950) and of the CPU oracle.

    python -m rustcolmap_amd.build          # both
    python -m rustcolmap_amd.build --lib    # HIP library only

The HIP sources are compiled with hipcc for gfx950 only.  Parity-relevant
flags: -ffp-contract=off (no FMA contraction, as Rust), IEEE f32 denormals,
and the default correctly rounded f32 divide/sqrt.  Packed-FP32 instructions
are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "rustcolmap_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libmvs_hip.so")
OBJDIR = os.path.join(PKG, "_build")
HIP_SOURCES = ["stereo.hip", "fuse.hip", "image.hip", "match.hip", "net_kernels.hip", "net.hip", "capi.hip", "plan.cpp"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"), "-Wno-unused-command-line-argument",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_library(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=len(cmds)) as ex:
        outs = list(ex.map(_run, cmds))
    if verbose:
        for o in outs:
            if o:
                print(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    return LIB


def build_all(verbose: bool = False):
    sys.path.insert(0, ROOT)
    from oracle import build as obuild   # test infrastructure: compiled, never linked into the library
    lib = build_library(verbose)
    orc = obuild.build_oracle()
    return lib, orc


if __name__ == "__main__":
    if "--lib" in sys.argv:
        print(build_library(verbose=True))
    else:
        print(build_all(verbose=True))
