"""Build recipe of liblsdem.so (hipcc, gfx950)."""
from __future__ import annotations

import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SOURCES = [HERE / "csrc" / "kernels.hip", HERE / "csrc" / "api.hip", HERE / "csrc" / "apps.hip",
           HERE / "csrc" / "surrogate.hip"]
OUT = HERE / "liblsdem.so"
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # bit-exact FP64 arithmetic: no contraction of a*b+c into fma
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


APPS = sorted((HERE / "apps").glob("*.cpp"))
INCLUDE = HERE.parent / "include"


def _run(cmd, verbose):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    if verbose:
        print(" ".join(cmd))


def build(verbose: bool = False) -> Path:
    _run([HIPCC, *FLAGS, *map(str, SOURCES), "-o", str(OUT)], verbose)
    # example apps written against the LSSolver drop-in (include/kernel/LSSolver.h)
    for app in APPS:
        _run([HIPCC, "-std=c++17", "-O2", f"-I{INCLUDE}", str(app), f"-L{HERE}", "-llsdem",
              "-Wl,-rpath,$ORIGIN/..", "-o", str(app.with_suffix(""))], verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
