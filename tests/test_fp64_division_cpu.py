"""The FP64 step's divisions (csrc/psc_div.h): a multiply and four FMAs from
a reciprocal known in advance must give IEEE division's quotient, bit for
bit, signed zeros included (CPU).

The header is compiled here with g++ (host FMA, no contraction of the
multiply) and run over the step's divisors -- tau, cs^2 = 1/3 and cs^4 as the
kernel forms them, rho near 1 -- and over random divisors of both signs, with
numerators of the lattice's magnitudes, random doubles over a wide exponent
range, and numerators whose quotient lies within 1e-15 ulp of a rounding
midpoint (constructed from the divisor's integer significand B: odd N in
[2^53, 2^54) with N B = A 2^(53+j) -+ s 2^t, so A / B is s 2^t 2^-106
away from the midpoint N 2^-(53+j)).  The test also checks that it can
fail: a reciprocal multiply, and a correction with the wrong sign, differ
from IEEE division on the same inputs.
"""

import ctypes as ct
import math
import os
import random
import shutil
import subprocess

import numpy as np
import pytest

from tests.conftest import ROOT

HEADER_DIR = os.path.join(ROOT, "particuology_hydraulic-penalty-_amd", "csrc")

SRC = r"""
#include "psc_div.h"
extern "C" {
void div_pos_n(const double* a, const double* b, const double* y, double* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = psc::div_pos(a[i], b[i], y[i]);
}
void div_any_n(const double* a, const double* b, const double* y, double* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = psc::div_any(a[i], b[i], y[i]);
}
void reciprocal_n(const double* a, const double* b, const double* y, double* out, long n) {
  for (long i = 0; i < n; ++i) out[i] = a[i] * y[i];  // planted: no correction
}
void wrong_sign_n(const double* a, const double* b, const double* y, double* out, long n) {
  for (long i = 0; i < n; ++i) {  // planted: residual a + b q instead of a - b q
    const double q0 = a[i] * y[i];
    const double q1 = __builtin_fma(__builtin_fma(b[i], q0, a[i]), y[i], q0);
    out[i] = __builtin_fma(__builtin_fma(b[i], q1, a[i]), y[i], q1);
  }
}
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("div")
    src = d / "div.cpp"
    src.write_text(SRC)
    so = d / "libdiv.so"
    subprocess.run([gxx, "-O2", "-ffp-contract=off", "-shared", "-fPIC", f"-I{HEADER_DIR}", str(src), "-o", str(so)],
                   check=True)
    h = ct.CDLL(str(so))
    for name in ("div_pos_n", "div_any_n", "reciprocal_n", "wrong_sign_n"):
        getattr(h, name).argtypes = [ct.c_void_p] * 4 + [ct.c_long]
    return h


def _run(h, name, a, b):
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(np.broadcast_to(b, a.shape), dtype=np.float64)
    y = 1.0 / b  # IEEE division: RN(1/b)
    out = np.empty_like(a)
    getattr(h, name)(a.ctypes.data, b.ctypes.data, y.ctypes.data, out.ctypes.data, a.size)
    return out


def _near_midpoints(b, count, rng):
    """Numerators a with a / b within s 2^t 2^-106 (relative) of a rounding
    midpoint, for odd s = 1, 3, 5, ...; both signs, random binades."""
    m, e = math.frexp(abs(b))
    B = int(m * 2 ** 53)                  # b = B 2^(e - 53), B < 2^53
    t = (B & -B).bit_length() - 1
    pick = random.Random(int(rng.integers(1 << 30)))
    out = []
    s = 1
    while len(out) < count:
        for j in (0, 1):                  # quotient in [1, 2) or [1/2, 1)
            mod = 1 << (53 + j - t)
            inv = pow(B >> t, -1, mod)
            for sign in (1, -1):
                n0 = (sign * s * inv) % mod   # N B = sign s 2^t (mod 2^(53+j)), N odd
                k_lo, k_hi = -(-((1 << 53) - n0) // mod), ((1 << 54) - 1 - n0) // mod
                if k_lo > k_hi:
                    continue
                for k in {pick.randint(k_lo, k_hi) for _ in range(3)}:
                    N = n0 + k * mod
                    A, rem = divmod(N * B - sign * s * (1 << t), 1 << (53 + j))
                    if N & 1 and not rem and (1 << 52) <= A < (1 << 53):
                        out.append(math.ldexp(A, e - 53 + int(rng.integers(-30, 30))) * rng.choice([-1.0, 1.0]))
        s += 2
        if s > 200_001:                   # divisors with short significands have few such numerators
            break
    return np.array(out[:count])


def _numerators(rng, b, n):
    """Lattice-sized values, wide-range random doubles, zeros, and numerators
    whose quotient by b sits next to a rounding midpoint."""
    lattice = rng.uniform(-0.5, 0.5, n) * 10.0 ** rng.integers(-16, 1, n)
    wide = (rng.uniform(1.0, 2.0, n) * np.exp2(rng.integers(-200, 200, n).astype(np.float64))
            * rng.choice([-1.0, 1.0], n))
    return np.concatenate([lattice, wide, _near_midpoints(b, 4000, rng), [0.0, -0.0, 1.0, -1.0, b, -b]])


def _same_bits(x, y):
    return np.array_equal(x.view(np.uint64), y.view(np.uint64))


def test_quotients_equal_ieee_division_for_the_step_divisors(lib):
    rng = np.random.default_rng(1)
    cs2 = 1.0 / 3.0
    divisors = [0.8, 0.72, 0.6, 0.5 + 1e-9, 1.0, 1.25, 1.7, 2.0, cs2, cs2 * cs2] + list(rng.uniform(0.5, 3.0, 6))
    for b in divisors:
        a = _numerators(rng, b, 400_000)
        got = _run(lib, "div_pos_n", a, b)
        assert _same_bits(got, a / b), b
    # rho: one divisor per cell near 1, both forms
    n = 2_000_000
    rho = rng.uniform(0.9, 1.1, n)
    a = np.concatenate([rng.uniform(-0.05, 0.05, n) * 10.0 ** rng.integers(-12, 1, n)])
    assert _same_bits(_run(lib, "div_any_n", a, rho), a / rho)
    assert _same_bits(_run(lib, "div_pos_n", a, rho), a / rho)


def test_quotients_equal_ieee_division_for_random_divisors_of_both_signs(lib):
    rng = np.random.default_rng(2)
    n = 1_000_000
    b = rng.uniform(1.0, 2.0, n) * np.exp2(rng.integers(-300, 300, n).astype(np.float64)) * rng.choice([-1.0, 1.0], n)
    a = rng.uniform(1.0, 2.0, n) * np.exp2(rng.integers(-300, 300, n).astype(np.float64)) * rng.choice([-1.0, 1.0], n)
    a[:1000] = 0.0
    a[1000:2000] = -0.0
    assert _same_bits(_run(lib, "div_any_n", a, b), a / b)
    for bb in (-0.8, -1.0 / 3.0, -1.7):
        aa = _numerators(rng, -bb, 200_000)
        assert _same_bits(_run(lib, "div_any_n", aa, bb), aa / bb), bb


def test_the_check_fails_on_planted_mistakes(lib):
    rng = np.random.default_rng(3)
    for b in (0.8, 1.0 / 3.0, (1.0 / 3.0) ** 2):
        a = _numerators(rng, b, 100_000)
        for planted in ("reciprocal_n", "wrong_sign_n"):
            assert not _same_bits(_run(lib, planted, a, b), a / b), (planted, b)
