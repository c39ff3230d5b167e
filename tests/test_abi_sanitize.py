"""The C ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

The library is rebuilt with host-only instrumentation into a temporary
directory (device code is compiled as usual and never runs here), and a small
C++ driver (tests/sanitize/abi_checks.cpp) calls every entry point with the
arguments it must reject -- null pointers, lattices whose extent products
overflow 32 bits, bad flags, sizes and plans -- and the ones that complete
without a device.  The test passes when every check holds and neither
sanitizer reports anything.
"""

import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT

HERE = os.path.dirname(os.path.abspath(__file__))


def _cxx():
    for c in ("/opt/rocm/llvm/bin/clang++", shutil.which("clang++")):
        if c and os.path.exists(c):
            return c
    return None


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    import importlib.util
    cxx = _cxx()
    if cxx is None or not (shutil.which("hipcc") or os.path.exists("/opt/rocm/bin/hipcc")):
        pytest.skip("the ROCm compilers are not installed")
    spec = importlib.util.spec_from_file_location(
        "psc_build_sanitize", os.path.join(ROOT, "particuology_hydraulic-penalty-_amd", "csrc", "build.py"))
    build = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(build)
    out = str(tmp_path_factory.mktemp("sanitize"))
    build.build(out_dir=out, sanitize=True, lib_name="libpsc_hip_sanitize.so")
    exe = os.path.join(out, "abi_checks")
    # the checking program is host C++ only; its sanitizer runtime serves the library's instrumentation
    cmd = [cxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
           "-I", os.path.join(ROOT, "include"), os.path.join(HERE, "sanitize", "abi_checks.cpp"),
           "-L", out, "-lpsc_hip_sanitize", f"-Wl,-rpath,{out}", "-o", exe]
    subprocess.run(cmd, check=True)
    return exe


def test_abi_argument_checks_are_sanitizer_clean(driver):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", HIP_VISIBLE_DEVICES="")
    r = subprocess.run([driver], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-3000:])
    print(r.stderr[-3000:])
    assert r.returncode == 0, r.stderr[-2000:]
    assert "ALL OK" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert r.stdout.count("ok  ") >= 50
