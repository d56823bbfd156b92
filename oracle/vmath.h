/*
 * vmath.h -- deterministic exp/log/pow for the CPU oracle.
 *
 * TEST INFRASTRUCTURE: part of the oracle.  Only tests/, the smoke check and
 * bench.py's CPU baseline use anything under oracle/.
 *
 * The reference evaluates its s95p equation of state (NEWEST10/s95p.h:13-123)
 * with libm exp/pow.  libm and the GPU's math library may round the last bit
 * differently, which would make Newton iteration counts (that stop on two
 * equal iterates, NEWEST10/root.h:346-363) differ between CPU and GPU.  The
 * oracle therefore evaluates exp, log and pow with the algorithms below, built
 * only from IEEE-754 operations that round identically on x86-64 SSE2 and on
 * gfx950 (add, mul, div, rint, frexp, ldexp; compiled without FP contraction).
 * The GPU kernels carry their own implementation of the same algorithms.
 *
 *   exp(x): x = k ln2 + r with a Cody-Waite split of ln2 (hi part has 32
 *           trailing zero bits so k*ln2_hi is exact), |r| <= ln2/2, then the
 *           Taylor series of e^r to degree 15 in Horner form, scaled by 2^k.
 *   log(x): x = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
 *           log m = 2 atanh(s) = 2 s sum_{n>=0} s^{2n}/(2n+1) to n = 12,
 *           plus e ln2 in the same hi/lo split.
 *   pow(x,y) = exp(y log x) for x > 0.
 * Accuracy against glibc is pinned by tests/test_oracle_math.py (relative
 * error well below 1e-13 over the ranges the EOS uses).
 */
#ifndef VIRAL_ORACLE_VMATH_H
#define VIRAL_ORACLE_VMATH_H
#include <math.h>

#define VM_LN2_HI 6.93147180369123816490e-01 /* 0x3fe62e42fee00000 */
#define VM_LN2_LO 1.90821492927058770002e-10 /* 0x3dea39ef35793c76 */
#define VM_INV_LN2 1.44269504088896338700e+00

static inline double vm_exp(double x)
{
  if (x != x) return x;
  if (x > 709.0) return HUGE_VAL;
  if (x < -745.0) return 0.0;
  double k = rint(x * VM_INV_LN2);
  double r = (x - k * VM_LN2_HI) - k * VM_LN2_LO;
  /* Horner form of sum_{n=0}^{15} r^n/n! */
  double p = 1.0 / 1307674368000.0;             /* 1/15! */
  p = p * r + 1.0 / 87178291200.0;              /* 1/14! */
  p = p * r + 1.0 / 6227020800.0;               /* 1/13! */
  p = p * r + 1.0 / 479001600.0;                /* 1/12! */
  p = p * r + 1.0 / 39916800.0;                 /* 1/11! */
  p = p * r + 1.0 / 3628800.0;                  /* 1/10! */
  p = p * r + 1.0 / 362880.0;                   /* 1/9!  */
  p = p * r + 1.0 / 40320.0;                    /* 1/8!  */
  p = p * r + 1.0 / 5040.0;                     /* 1/7!  */
  p = p * r + 1.0 / 720.0;                      /* 1/6!  */
  p = p * r + 1.0 / 120.0;                      /* 1/5!  */
  p = p * r + 1.0 / 24.0;                       /* 1/4!  */
  p = p * r + 1.0 / 6.0;                        /* 1/3!  */
  p = p * r + 0.5;                              /* 1/2!  */
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, (int)k);
}

static inline double vm_log(double x)
{
  if (x != x || x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int e;
  double m = frexp(x, &e);          /* m in [0.5, 1) */
  if (m < 0.70710678118654752440) { /* bring m into [sqrt(1/2), sqrt(2)) */
    m = m * 2.0;
    e = e - 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double z = s * s;
  double p = 1.0 / 25.0;
  p = p * z + 1.0 / 23.0;
  p = p * z + 1.0 / 21.0;
  p = p * z + 1.0 / 19.0;
  p = p * z + 1.0 / 17.0;
  p = p * z + 1.0 / 15.0;
  p = p * z + 1.0 / 13.0;
  p = p * z + 1.0 / 11.0;
  p = p * z + 1.0 / 9.0;
  p = p * z + 1.0 / 7.0;
  p = p * z + 1.0 / 5.0;
  p = p * z + 1.0 / 3.0;
  double lm = 2.0 * s + 2.0 * s * (z * p);
  double de = (double)e;
  return de * VM_LN2_HI + (de * VM_LN2_LO + lm);
}

static inline double vm_pow(double x, double y)
{
  return vm_exp(y * vm_log(x));
}

#endif
