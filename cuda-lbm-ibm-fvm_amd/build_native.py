"""Build the native pieces in-tree.

* ``lib/libvof_hip.so`` — gfx950 HIP kernels + C ABI (hipcc).
* ``lib/3dvof``         — the drop-in CLI, linked against the library.
* ``oracle/liboracle.so`` — the CPU oracle used by tests (test infrastructure,
  built here because ``build()`` builds every checker).
* ``oracle/liboracle_fast.so`` — the same source built for speed, timed as
  bench.py's CPU baseline only.

Run ``python cuda-lbm-ibm-fvm_amd/build_native.py`` or call :func:`build_all`.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
ORACLE = os.path.join(ROOT, "oracle")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off: no a*b+c contraction, so every source operation is one
# IEEE rounding, as in the oracle.  Packed-FP32 instructions are disabled.
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip() -> str:
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libvof_hip.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, f) for f in ("vof_kernels.hip", "vof_adjoint.hip", "vof_host.hip", "vof_vtk.cpp")]
    _run([HIPCC, *HIP_FLAGS, "-shared", "-pthread", *srcs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_cli() -> str:
    out = os.path.join(LIB, "3dvof")
    _run(["g++", "-O2", "-std=c++17", os.path.join(CSRC, "vof_cli.cpp"), "-L" + LIB, "-lvof_hip",
          "-Wl,-rpath,$ORIGIN", "-o", out])
    return out


def build_oracle() -> str:
    out = os.path.join(ORACLE, "liboracle.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-march=x86-64-v3", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
          os.path.join(ORACLE, "vof_oracle.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle_fast() -> str:
    """Timing build of the oracle for bench.py's CPU baseline: the same source
    at -O3 for the AVX-512 baseline (x86-64-v4) with FMA contraction allowed
    (GCC's default), so its bits may differ from the parity build's; never
    used as the checker."""
    out = os.path.join(ORACLE, "liboracle_fast.so")
    tmp = out + ".tmp"
    _run(["g++", "-O3", "-march=x86-64-v4", "-std=c++17", "-fopenmp", "-fPIC", "-shared",
          os.path.join(ORACLE, "vof_oracle.cpp"), "-o", tmp])
    os.replace(tmp, out)
    return out


def build_all(hip: bool = True) -> None:
    build_oracle()
    build_oracle_fast()
    if hip:
        build_hip()
        build_cli()


if __name__ == "__main__":
    build_all(hip="--oracle-only" not in sys.argv)
