/*
 * exo_oracle_math.c -- TEST INFRASTRUCTURE (oracle).  Deterministic exp / log /
 * pow used by the oracle's hydrostatic decomposition.
 *
 * The reference evaluates log, pow and exp from libm in
 * src/snap/decomposition/change_to_perturbation.cpp:63-65,109-115,181-182 and
 * src/snap/decomposition/decomposition.cpp:357-360.  libm and the GPU math
 * library differ in the last bit, which would make GPU-vs-oracle parity a
 * tolerance question instead of an identity.  So both sides evaluate the same
 * published method, written out operation by operation:
 *   exp : Cody-Waite reduction x = k ln2 + r, |r| <= ln2/2, degree-13 Taylor
 *         polynomial in Horner form, exact scaling by 2^k.
 *   log : x = 2^k m, m in [sqrt(1/2), sqrt(2)), f = (m-1)/(m+1),
 *         log m = 2 atanh f = 2 f (1 + f^2/3 + f^4/5 + ...) to f^24, plus k ln2.
 *   pow : exp(y log x) for x > 0.
 * tests/test_oracle_math.py pins them against libm (exp/log within 2 ulp,
 * pow within 64 ulp on the decomposition's argument range).
 * Built with -ffp-contract=off so no fused multiply-add changes a rounding.
 */
#include <math.h>

#include "exo_oracle.h"

/* ln 2 split so that k * OR_LN2_HI is exact for |k| < 2^11 */
static const double OR_LN2_HI = 6.93147180369123816490e-01;
static const double OR_LN2_LO = 1.90821492927058770002e-10;
static const double OR_INV_LN2 = 1.44269504088896338700e+00;

double or_exp(double x) {
  if (x != x) return x;
  if (x > 709.782712893384) return HUGE_VAL;
  if (x < -745.1332191019412) return 0.0;
  double kd = floor(x * OR_INV_LN2 + 0.5);
  int k = (int)kd;
  double r = (x - kd * OR_LN2_HI) - kd * OR_LN2_LO;
  /* sum_{n=0}^{13} r^n / n! */
  double p = 1.0 / 6227020800.0;            /* 1/13! */
  p = 1.0 / 479001600.0 + r * p;           /* 1/12! */
  p = 1.0 / 39916800.0 + r * p;            /* 1/11! */
  p = 1.0 / 3628800.0 + r * p;             /* 1/10! */
  p = 1.0 / 362880.0 + r * p;              /* 1/9!  */
  p = 1.0 / 40320.0 + r * p;               /* 1/8!  */
  p = 1.0 / 5040.0 + r * p;                /* 1/7!  */
  p = 1.0 / 720.0 + r * p;                 /* 1/6!  */
  p = 1.0 / 120.0 + r * p;                 /* 1/5!  */
  p = 1.0 / 24.0 + r * p;                  /* 1/4!  */
  p = 1.0 / 6.0 + r * p;                   /* 1/3!  */
  p = 0.5 + r * p;                         /* 1/2!  */
  p = 1.0 + r * p;                         /* 1/1!  */
  p = 1.0 + r * p;                         /* 1/0!  */
  return ldexp(p, k);
}

double or_log(double x) {
  if (x != x) return x;
  if (x < 0.0) return NAN;
  if (x == 0.0) return -HUGE_VAL;
  if (x == HUGE_VAL) return x;
  int k;
  double m = frexp(x, &k); /* x = m 2^k, m in [0.5, 1) */
  if (m < 0.70710678118654752440) {
    m = m * 2.0;
    k = k - 1;
  }
  double f = (m - 1.0) / (m + 1.0);
  double s = f * f;
  /* sum_{n=0}^{12} s^n / (2n+1) */
  double p = 1.0 / 25.0;
  p = 1.0 / 23.0 + s * p;
  p = 1.0 / 21.0 + s * p;
  p = 1.0 / 19.0 + s * p;
  p = 1.0 / 17.0 + s * p;
  p = 1.0 / 15.0 + s * p;
  p = 1.0 / 13.0 + s * p;
  p = 1.0 / 11.0 + s * p;
  p = 1.0 / 9.0 + s * p;
  p = 1.0 / 7.0 + s * p;
  p = 1.0 / 5.0 + s * p;
  p = 1.0 / 3.0 + s * p;
  double lm = 2.0 * f + (2.0 * f) * (s * p);
  double kd = (double)k;
  return kd * OR_LN2_HI + (lm + kd * OR_LN2_LO);
}

double or_pow(double x, double y) { return or_exp(y * or_log(x)); }
