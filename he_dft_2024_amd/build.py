"""Build the HIP library libhedft.so (gfx950) in-tree with hipcc.

Usage: python -m he_dft_2024_amd.build  (or he_dft_2024_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(OUT, "libhedft.so")
SOURCES = ["fft.hip", "stencil.hip", "meanfield.hip", "steppers.hip", "statics.hip", "observables.hip", "hedft.hip",
           "impurity.hip", "serve.hip", "denio.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HEDFT_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 instructions are generated (they are not used on this path)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _compile(src: str, out: str = OUT, extra: tuple = ()) -> str:
    obj = os.path.join(out, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        errs = [l for l in r.stderr.splitlines() if "error" in l]
        raise RuntimeError(f"hipcc failed for {src}:\n" + "\n".join(errs[:20]) + "\n" + r.stderr[-4000:])
    return obj


def build(verbose: bool = True, out: str = OUT, extra: tuple = ()) -> str:
    """Compile and link libhedft.so into `out`; `extra` adds hipcc flags
    (e.g. -D tuning macros)."""
    os.makedirs(out, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(lambda s: _compile(s, out, tuple(extra)), SOURCES))
    lib = os.path.join(out, "libhedft.so")
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-pthread", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-8000:]}")
    os.replace(tmp, lib)
    if verbose:
        print(f"built {lib}")
    return lib


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
