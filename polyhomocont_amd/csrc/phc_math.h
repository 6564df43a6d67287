// phc_math.h -- complex FP64 arithmetic with the exact semantics of the
// reference build (polyhomocont compiled by flang for x86-64):
//
//  * complex * complex   -> (ar*br - ai*bi, ar*bi + ai*br), no FMA
//    (the whole library is compiled with -ffp-contract=off);
//  * real or integer * complex -> flang promotes the scalar to (s, 0) and
//    multiplies in full; for finite operands this equals (s*ar, s*ai) up to
//    the sign of a zero result, which is what cscale computes;
//  * complex / complex   -> compiler-rt's __divdc3 (logb/scalbn-scaled
//    formula with the C99 Annex G recovery of infinities), which is what
//    flang-compiled code links;
//  * abs(complex)        -> glibc cabs = hypot, glibc 2.35's corrected
//    algorithm of C. F. Borges, "An Improved Algorithm for hypot(a,b)",
//    arXiv:1904.09481 (2019), in its non-FMA form;
//  * complex ** integer  -> flang runtime's binary powering (squarings of
//    the base, lowest exponent bit first).
//
// Each of these was checked bit for bit against the host routines on
// millions of random inputs (tests/test_math_host.py runs the same checks
// through this header compiled for the CPU). Everything is __host__
// __device__ so the CPU build of the header is the one that is tested.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PHC_HD __host__ __device__ __forceinline__
// rare slow paths stay out of line to keep the kernel's code small
#define PHC_HD_COLD __host__ __device__ __noinline__
#else
#define PHC_HD static inline
#define PHC_HD_COLD static
#endif

namespace phc {

struct cx {
  double re, im;
};

PHC_HD cx mk(double re, double im) {
  cx r;
  r.re = re;
  r.im = im;
  return r;
}
PHC_HD cx cadd(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
PHC_HD cx csub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
PHC_HD cx cneg(cx a) { return mk(-a.re, -a.im); }
PHC_HD cx cmul(cx a, cx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// (s, 0) * a for a real (or integer-valued) scalar s.
PHC_HD cx cscale(double s, cx a) { return mk(s * a.re, s * a.im); }

// C99 Annex G recovery of infinities and zeros for __divdc3 when the
// formula gave NaN + i NaN (c, d already scaled by 2^-ilogbw).
PHC_HD_COLD cx cdiv_special(double a, double b, double c, double d,
                            double denom, double logbw, double x, double y) {
  {
    if (denom == 0.0 && (!isnan(a) || !isnan(b))) {
      x = copysign(INFINITY, c) * a;
      y = copysign(INFINITY, c) * b;
    } else if ((isinf(a) || isinf(b)) && isfinite(c) && isfinite(d)) {
      a = copysign(isinf(a) ? 1.0 : 0.0, a);
      b = copysign(isinf(b) ? 1.0 : 0.0, b);
      x = INFINITY * (a * c + b * d);
      y = INFINITY * (b * c - a * d);
    } else if (isinf(logbw) && logbw > 0.0 && isfinite(a) && isfinite(b)) {
      c = copysign(isinf(c) ? 1.0 : 0.0, c);
      d = copysign(isinf(d) ? 1.0 : 0.0, d);
      x = 0.0 * (a * c + b * d);
      y = 0.0 * (b * c - a * d);
    }
  }
  return mk(x, y);
}

// compiler-rt __divdc3, in two parts: the divisor's scaling (logb, the
// scaled c and d, c^2 + d^2), which depends only on the divisor, and the
// quotient for a numerator. cdiv(a, b) == cdiv_apply(a, cdiv_prep(b)) bit
// for bit; a pivot used for several quotients is prepared once.
struct cdiv_den {
  double c, d, denom, logbw;
  int ilogbw;
};
PHC_HD cdiv_den cdiv_prep(cx den) {
  cdiv_den p;
  p.c = den.re;
  p.d = den.im;
  p.ilogbw = 0;
  p.logbw = logb(fmax(fabs(p.c), fabs(p.d)));
  if (isfinite(p.logbw)) {
    p.ilogbw = (int)p.logbw;
    p.c = ldexp(p.c, -p.ilogbw);
    p.d = ldexp(p.d, -p.ilogbw);
  }
  p.denom = p.c * p.c + p.d * p.d;
  return p;
}
PHC_HD cx cdiv_apply(cx num, const cdiv_den& p) {
  const double a = num.re, b = num.im;
  const double x = ldexp((a * p.c + b * p.d) / p.denom, -p.ilogbw);
  const double y = ldexp((b * p.c - a * p.d) / p.denom, -p.ilogbw);
  if (isnan(x) && isnan(y))
    return cdiv_special(a, b, p.c, p.d, p.denom, p.logbw, x, y);
  return mk(x, y);
}
PHC_HD cx cdiv(cx num, cx den) { return cdiv_apply(num, cdiv_prep(den)); }

// Correction step of Borges' algorithm for 0 < ay <= ax, no overflow.
PHC_HD double hypot_kernel(double ax, double ay) {
  double h = sqrt(ax * ax + ay * ay);
  double t1, t2;
  if (h <= 2.0 * ay) {
    double delta = h - ay;
    t1 = ax * (2.0 * delta - ax);
    t2 = (delta - 2.0 * (ax - ay)) * delta;
  } else {
    double delta = h - ax;
    t1 = 2.0 * delta * (ax - 2.0 * ay);
    t2 = (4.0 * delta - ay) * ay + delta * delta;
  }
  h -= (t1 + t2) / (2.0 * h);
  return h;
}

// hypot outside the common range: non-finite, huge or tiny operands.
PHC_HD_COLD double cabs_special(double x, double y) {
  if (!isfinite(x) || !isfinite(y)) {
    if (isinf(x) || isinf(y)) return INFINITY;
    return x + y;  // NaN
  }
  const double ax = x < y ? y : x;
  const double ay = x < y ? x : y;
  const double EPS = 0x1p-54;
  const double SCALE = 0x1p-600;
  if (ax > 0x1p+511) {
    if (ay <= ax * EPS) return ax + ay;
    return hypot_kernel(ax * SCALE, ay * SCALE) / SCALE;
  }
  if (ax >= ay / EPS) return ax + ay;
  return hypot_kernel(ax / SCALE, ay / SCALE) * SCALE;
}

// |a| with glibc hypot semantics.
PHC_HD double cabs_(cx a) {
  const double x = fabs(a.re), y = fabs(a.im);
  const double ax = x < y ? y : x;
  const double ay = x < y ? x : y;
  // common range: finite, ax <= 2^511, ay >= 2^-511
  if (!(ax <= 0x1p+511) || !(ay >= 0x1p-511)) return cabs_special(x, y);
  if (ay <= ax * 0x1p-54) return ax + ay;
  return hypot_kernel(ax, ay);
}

// flang runtime complex ** integer (binary powering).
PHC_HD cx zpowi(cx base, int e) {
  if (e == 0) return mk(1.0, 0.0);
  bool inv = e < 0;
  unsigned int u = inv ? (unsigned int)(-(long long)e) : (unsigned int)e;
  while ((u & 1u) == 0u) {
    base = cmul(base, base);
    u >>= 1;
  }
  cx acc = base;
  while (u > 1u) {
    u >>= 1;
    base = cmul(base, base);
    if (u & 1u) acc = cmul(acc, base);
  }
  if (inv) acc = cdiv(mk(1.0, 0.0), acc);
  return acc;
}

// ---- correctly rounded sin/cos ------------------------------------------
// Double-double evaluation (Dekker products with FMA) of cos and sin for
// |theta| < 2^20: Cody-Waite reduction by pi/2 with a three-part pi/2,
// Taylor series to 2^-110 on |r| <= pi/4, one final rounding. The result
// is the correctly rounded double except when the exact value lies within
// ~2^-100 relative of a rounding midpoint.
struct dd {
  double hi, lo;
};
PHC_HD dd dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  double e = (a - (s - bb)) + (b - bb);
  dd r;
  r.hi = s;
  r.lo = e;
  return r;
}
PHC_HD dd dd_fast(double a, double b) {
  double s = a + b;
  dd r;
  r.hi = s;
  r.lo = b - (s - a);
  return r;
}
PHC_HD dd dd_add(dd a, dd b) {
  dd s = dd_two_sum(a.hi, b.hi);
  dd t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_fast(s.hi, s.lo);
  s.lo += t.lo;
  return dd_fast(s.hi, s.lo);
}
PHC_HD dd dd_mul(dd a, dd b) {
  double p = a.hi * b.hi;
  double e = fma(a.hi, b.hi, -p);
  e += a.hi * b.lo + a.lo * b.hi;
  return dd_fast(p, e);
}
PHC_HD dd dd_mul_d(dd a, double b) {
  double p = a.hi * b;
  double e = fma(a.hi, b, -p);
  e += a.lo * b;
  return dd_fast(p, e);
}

// sin (which = 0) and cos (which = 1) of a double-double r, |r| <= pi/4.
PHC_HD dd dd_sincos_kernel(dd r, int which) {
  dd r2 = dd_mul(r, r);
  // term_k = (-1)^k r^(2k+which) / (2k+which)!
  dd term;
  if (which == 0) {
    term = r;
  } else {
    term.hi = 1.0;
    term.lo = 0.0;
  }
  dd sum = term;
  int m = (which == 0) ? 1 : 0;  // power of the current term
  for (int k = 1; k < 16; ++k) {
    term = dd_mul(term, r2);
    const double den = (double)((m + 1) * (m + 2));
    // divide by den in double-double: q = term / den
    double qh = term.hi / den;
    double rem = fma(-qh, den, term.hi) + term.lo;
    double ql = rem / den;
    term.hi = -qh;
    term.lo = -ql;
    term = dd_fast(term.hi, term.lo);
    m += 2;
    sum = dd_add(sum, term);
    if (fabs(term.hi) < 0x1p-112 * fabs(sum.hi)) break;
  }
  return sum;
}

// (cos theta, sin theta), each correctly rounded.
PHC_HD_COLD cx cr_cossin(double theta) {
  // pi/2 = P1 + P2 + P3 (53 + 53 + 53 bits)
  const double P1 = 0x1.921fb54442d18p+0;
  const double P2 = 0x1.1a62633145c07p-54;
  const double P3 = -0x1.f1976b7ed8fbcp-110;
  const double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
  double kd = rint(theta * TWO_OVER_PI);
  // r = theta - k P1 - k P2 - k P3 in double-double, products exact
  double p1 = kd * P1;
  double e1 = fma(kd, P1, -p1);
  dd r = dd_two_sum(theta, -p1);
  r = dd_add(r, dd_fast(-e1, 0.0));
  double p2 = kd * P2;
  double e2 = fma(kd, P2, -p2);
  r = dd_add(r, dd_two_sum(-p2, -e2));
  r = dd_add(r, dd_fast(-kd * P3, 0.0));
  const int q = ((int)kd) & 3;
  dd sn = dd_sincos_kernel(r, 0);
  dd cs = dd_sincos_kernel(r, 1);
  double s = sn.hi + sn.lo, c = cs.hi + cs.lo;
  cx out;
  switch (q) {
    case 0: out = mk(c, s); break;
    case 1: out = mk(-s, c); break;
    case 2: out = mk(-c, -s); break;
    default: out = mk(s, -c); break;
  }
  return out;
}

}  // namespace phc
