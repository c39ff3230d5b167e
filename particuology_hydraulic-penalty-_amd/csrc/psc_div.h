// Correctly rounded FP64 division from a reciprocal known in advance.
//
// The FP64 step reproduces NumPy's IEEE divisions (round to nearest even) bit
// for bit.  The divisors of its ~100 divisions per cell are few: tau, s^3, the
// constants cs^2 and cs^4, and per cell rho (and a moving cell's solid
// fraction).  With y = RN(1/b) computed once (by the host for tau and s^3, by
// the compiler for the constants, by one IEEE division per cell otherwise),
// each quotient takes a multiply and four FMAs instead of the full division
// sequence with its quarter-rate reciprocal:
//
//   q0 = RN(a y)                 within 1.5 ulp of a/b
//   q1 = RN(q0 + RN(a - b q0) y) one Newton-Raphson correction: within 1 ulp
//   r1 = a - b q1                exact (an FMA), as q1 is within 1 ulp of a/b
//   q  = RN(q1 + r1 y)           = RN(a/b)
//
// The last line is Markstein's theorem (P. Markstein, "Computation of
// elementary functions on the IBM RISC System/6000 processor", IBM J. Res.
// Dev. 34, 1990; restated in the Handbook of Floating-Point Arithmetic,
// Muller et al.): if y is within half an ulp of 1/b and q1 within one ulp of
// a/b, RN(q1 + r1 y) is the correctly rounded quotient.  The conditions
// exclude overflow and underflow of the quotient and of r1, far from the
// lattice's values (|a/b| between 2^-900 and 2^900 suffices).  A zero
// numerator gets the sign IEEE division gives it.  An infinite numerator gives
// NaN instead of an infinity; both mark a diverged state.
//
// Host and device code: the CPU tests compile this header with g++ and check
// it against IEEE division.
#pragma once

#if defined(__HIPCC__)
#define PSC_DIV_HD __host__ __device__ __forceinline__
#else
#define PSC_DIV_HD inline
#endif

namespace psc {

// a / b for b > 0 with y = RN(1/b).
PSC_DIV_HD double div_pos(double a, double b, double y) {
  const double q0 = a * y;
  const double q1 = __builtin_fma(__builtin_fma(-b, q0, a), y, q0);
  const double q = __builtin_fma(__builtin_fma(-b, q1, a), y, q1);
  return __builtin_copysign(q, a);  // a zero numerator: the sign of a (b > 0)
}

// a / b for b of either sign with y = RN(1/b).
PSC_DIV_HD double div_any(double a, double b, double y) {
  const double q0 = a * y;
  const double q1 = __builtin_fma(__builtin_fma(-b, q0, a), y, q0);
  const double q = __builtin_fma(__builtin_fma(-b, q1, a), y, q1);
  return a == 0.0 ? q0 : q;  // a zero numerator: q0 = a y has the sign of a / b
}

}  // namespace psc
