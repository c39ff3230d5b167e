"""Build libpsc_hip.so for gfx950 with hipcc, in-tree.

    python particuology_hydraulic-penalty-_amd/csrc/build.py

Four translation units:
  psc_exact.hip  FP64 kernels in the reference's operation order, compiled
                 with -ffp-contract=off (no fused multiply-add), plus the
                 voxelizer, particle loads, moments and f import/export;
  psc_fast.hip   the FP32 deviation step, -ffp-contract=fast-honor-pragmas
                 (fused multiply-add wherever the source does not switch it
                 off with #pragma clang fp contract(off); plain
                 -ffp-contract=fast would ignore the pragma);
  psc_adjoint.hip  smooth solid fraction and the reverse-mode step,
                 -ffp-contract=fast-honor-pragmas;
  psc_abi.hip    the extern "C" layer.
Packed FP32 instructions are disabled on every device compile.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)
LIB_NAME = "libpsc_hip.so"
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}",
]
UNITS = [
    ("psc_exact.hip", ["-ffp-contract=off"]),
    ("psc_fast.hip", ["-ffp-contract=fast-honor-pragmas"]),
    ("psc_adjoint.hip", ["-ffp-contract=fast-honor-pragmas"]),
    ("psc_abi.hip", ["-ffp-contract=off"]),
]


def hipcc() -> str:
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return cand


# Host-only AddressSanitizer and UndefinedBehaviorSanitizer instrumentation (each flag
# applies to the host pass only; device code is built as usual), for the CPU checks of
# the C ABI's argument handling (tests/test_abi_sanitize.py).
HOST_SANITIZE = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                 "-Xarch_host", "-fno-omit-frame-pointer", "-g"]


def build(verbose: bool = False, out_dir: str | None = None, sanitize: bool = False,
          lib_name: str = LIB_NAME) -> str:
    """Compile every unit and link the shared library; return its path.
    sanitize=True instruments the host code (a separate library for tests)."""
    cc = hipcc()
    out_dir = out_dir or PKG
    target = os.path.join(out_dir, lib_name)
    extra_all = HOST_SANITIZE if sanitize else []
    with tempfile.TemporaryDirectory(prefix="psc_build_") as tmp:
        objs = []
        for src, extra in UNITS:
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            cmd = [cc, "-c", os.path.join(HERE, src), "-o", obj] + COMMON + extra + extra_all
            if verbose:
                print(" ".join(cmd), flush=True)
            subprocess.run(cmd, check=True)
            objs.append(obj)
        tmp_lib = os.path.join(tmp, lib_name)
        cmd = [cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp_lib] + objs + extra_all
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)
        shutil.copyfile(tmp_lib, target + ".tmp")
        os.replace(target + ".tmp", target)
    return target


if __name__ == "__main__":
    path = build(verbose="-q" not in sys.argv)
    print("built", path)
