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
// Provenance and licences of the methods restated here (no upstream code is
// included; the algorithms and published constants are restated so that
// results match those libraries bit for bit):
//  * complex division: the algorithm of LLVM compiler-rt's __divdc3
//    (compiler-rt is Apache-2.0 WITH LLVM-exception);
//  * hypot: C. F. Borges, arXiv:1904.09481, as used by glibc 2.35
//    (glibc is LGPL-2.1-or-later);
//  * sin/cos: the IBM Accurate Mathematical Library method of glibc 2.35's
//    sysdeps/ieee754/dbl-64 (LGPL-2.1-or-later), coefficients as published
//    there; the 1/128 table (phc_sincos_table.h) is generated from its
//    definition by gen_sincos_table.py.
// Each of these was checked bit for bit against the host routines on
// millions of random inputs (tests/test_math_host.py runs the same checks
// through this header compiled for the CPU). Everything is __host__
// __device__ so the CPU build of the header is the one that is tested.
#pragma once

#include <math.h>
#include <stdint.h>

#include "phc_sincos_table.h"

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

// Fortran MAX and MIN as flang compiles them for x86-64 (maxsd, minsd):
// the second argument unless the first compares greater (less). A NaN
// second argument is returned, a NaN first argument is not -- unlike
// fmax/fmin, which drop NaNs. Paths that overflow to NaN take the
// reference's steps only with these.
PHC_HD double ftn_max(double a, double b) { return a > b ? a : b; }
PHC_HD double ftn_min(double a, double b) { return a < b ? a : b; }

// C99 Annex G recovery of infinities and zeros for __divdc3 when the
// formula gave NaN + i NaN (c, d already scaled by 2^-ilogbw). The divisor
// was infinite (logb = +inf) exactly when c or d is: a finite divisor's
// c and d are scaled to finite values, an infinite one is not scaled.
PHC_HD_COLD cx cdiv_special(double a, double b, double c, double d,
                            double denom, double x, double y) {
  {
    if (denom == 0.0 && (!isnan(a) || !isnan(b))) {
      x = copysign(INFINITY, c) * a;
      y = copysign(INFINITY, c) * b;
    } else if ((isinf(a) || isinf(b)) && isfinite(c) && isfinite(d)) {
      a = copysign(isinf(a) ? 1.0 : 0.0, a);
      b = copysign(isinf(b) ? 1.0 : 0.0, b);
      x = INFINITY * (a * c + b * d);
      y = INFINITY * (b * c - a * d);
    } else if ((isinf(c) || isinf(d)) && isfinite(a) && isfinite(b)) {
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
// quotient for a numerator. cdiv(a, b) == cdiv_apply(a, cdiv_scale(b)) bit
// for bit; a pivot used for several quotients is scaled once.
//
// logb(m), m = max(|c|, |d|), is finite exactly when m is finite and
// nonzero; it is then the exponent of m (subnormals as if normalized),
// which is what frexp's exponent minus one gives (the device's own logb is
// formed that way). Otherwise the divisor is not scaled.
//
// On the device the divisor also carries `rden`, the reciprocal of
// c^2 + d^2 as the compiler's IEEE division sequence refines it (v_rcp_f64
// and two Newton steps), so that a quotient only runs the sequence's last
// three operations (see qdiv_fast, cdiv_apply_fast). A divisor that is not
// scaled has c^2 + d^2 = 0, inf or NaN, whose refined reciprocal is NaN:
// its quotients fail qdiv_fast_ok and take the full division.
struct cdiv_den {
  double c, d, denom, rden;
  int ilogbw;
};
PHC_HD int exponent_of(double m) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_frexp_exp(m) - 1;
#else
  return ilogb(m);
#endif
}
// The scaling alone (rden not formed).
PHC_HD cdiv_den cdiv_scale(cx den) {
  cdiv_den p;
  const double m = fmax(fabs(den.re), fabs(den.im));
  const bool scale = isfinite(m) && m != 0.0;
  p.ilogbw = scale ? exponent_of(m) : 0;
  // ldexp by 0 returns its argument, zeros, infinities and NaNs included
  p.c = ldexp(den.re, -p.ilogbw);
  p.d = ldexp(den.im, -p.ilogbw);
  p.denom = p.c * p.c + p.d * p.d;
  p.rden = 0.0;
  return p;
}
// rden of a scaled divisor (a pivot searched by all lanes is scaled by
// every candidate and gets its reciprocal once, after the broadcast).
PHC_HD void cdiv_recip(cdiv_den& p) {
#if defined(__HIP_DEVICE_COMPILE__)
  // LLVM's f64 division: r = rcp(y); e = fma(-y, r, 1); r = fma(r, e, r);
  // e = fma(-y, r, 1); r = fma(r, e, r)
  double r = __builtin_amdgcn_rcp(p.denom);
  double e = fma(-p.denom, r, 1.0);
  r = fma(r, e, r);
  e = fma(-p.denom, r, 1.0);
  p.rden = fma(r, e, r);
#else
  (void)p;
#endif
}
// Scaling and reciprocal, for cdiv_apply_fast.
PHC_HD cdiv_den cdiv_prep(cx den) {
  cdiv_den p = cdiv_scale(den);
  cdiv_recip(p);
  return p;
}

// n / denom on the device, bit for bit what the compiler's division gives
// where its scaling steps do nothing. The compiler divides x by y as
//   s = div_scale(y); r = the refined reciprocal of s (cdiv_recip);
//   u = div_scale(x); q = u r; e = fma(-s, q, u);
//   q = div_fmas(e, r, q); result = div_fixup(q, y, x).
// For y = denom in [1, 8) (a scaled divisor) and a numerator with
// 2^-969 <= |x| < 2^768, div_scale returns its operand and clears VCC, so
// div_fmas is a plain fma, and div_fixup returns q for a normal quotient:
// the result is fma(e, r, x r) with e = fma(-y, x r, x). This forms the
// same with the residual's sign flipped, -e = fma(y, x r, -x) (exact), so
// that a zero numerator keeps its sign as div_fixup would give it
// (+0 / y = +0, -0 / y = -0). qdiv_fast_ok tells where it applies.
PHC_HD double qdiv_fast(double n, const cdiv_den& p) {
  const double q = n * p.rden;
  const double me = fma(p.denom, q, -n);
  return fma(-me, p.rden, q);
}
PHC_HD bool qdiv_fast_ok(double n, double q) {
#if defined(__HIP_DEVICE_COMPILE__)
  // frexp exponent in [-968, 768], or 0 for a zero; infinite or NaN
  // numerators (exponent 0 too) and unscaled divisors give a NaN q
  const int e = __builtin_amdgcn_frexp_exp(n);
  return (unsigned int)(e + 968) <= 1736u && !isnan(q);
#else
  (void)n;
  (void)q;
  return false;
#endif
}

// The __divdc3 quotient: IEEE divisions, then Annex G.
PHC_HD cx cdiv_apply(cx num, const cdiv_den& p) {
  const double a = num.re, b = num.im;
  const double x = ldexp((a * p.c + b * p.d) / p.denom, -p.ilogbw);
  const double y = ldexp((b * p.c - a * p.d) / p.denom, -p.ilogbw);
  if (isnan(x) && isnan(y))
    return cdiv_special(a, b, p.c, p.d, p.denom, x, y);
  return mk(x, y);
}
PHC_HD cx cdiv(cx num, cx den) { return cdiv_apply(num, cdiv_scale(den)); }

// The same quotient through qdiv_fast (the divisor's rden formed), for
// the linearization unit's LU and solves: it clears `ok` where qdiv_fast
// does not apply to either part, and the unit is then redone with
// cdiv_apply. Keeping the full division out of this path keeps it out of
// the unit's hot code. (Host: cdiv_apply.)
PHC_HD cx cdiv_apply_fast(cx num, const cdiv_den& p, bool& ok) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double a = num.re, b = num.im;
  const double nr = a * p.c + b * p.d;
  const double ni = b * p.c - a * p.d;
  const double qr = qdiv_fast(nr, p), qi = qdiv_fast(ni, p);
  ok = ok & qdiv_fast_ok(nr, qr) & qdiv_fast_ok(ni, qi);
  return mk(ldexp(qr, -p.ilogbw), ldexp(qi, -p.ilogbw));
#else
  (void)ok;
  return cdiv_apply(num, p);
#endif
}

// Correction step of Borges' algorithm for 0 < ay <= ax, no overflow.
// Both cases of the correction are formed and one is selected, so lanes
// that fall in different cases do not serialize on the GPU.
PHC_HD double hypot_kernel(double ax, double ay) {
  const double h = sqrt(ax * ax + ay * ay);
  const double dn = h - ay;  // h <= 2 ay
  const double t1n = ax * (2.0 * dn - ax);
  const double t2n = (dn - 2.0 * (ax - ay)) * dn;
  const double df = h - ax;  // otherwise
  const double t1f = 2.0 * df * (ax - 2.0 * ay);
  const double t2f = (4.0 * df - ay) * ay + df * df;
  const bool near = h <= 2.0 * ay;
  const double t1 = near ? t1n : t1f, t2 = near ? t2n : t2f;
  return h - (t1 + t2) / (2.0 * h);
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
  const double hk = hypot_kernel(ax, ay);
  return (ay <= ax * 0x1p-54) ? ax + ay : hk;
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

// z^e and w^e for 0 <= e < 2^nbits with exactly the products of zpowi
// (the squarings of the base, and the product of the squares of the set
// bits taken from the lowest set bit up, the accumulated product on the
// left), but without branches: every lane runs the nbits - 1 squarings and
// the nbits products and keeps its own by selects, so lanes with different
// exponents do not serialize. nbits is the same for all lanes; e = 0 gives
// (1, 0). Squarings past the highest set bit are formed and not used.
PHC_HD void zpowi_pair(cx z, cx w, int e, int nbits, cx& ze, cx& we) {
  double zr = 1.0, zi = 0.0, wr = 1.0, wi = 0.0;  // the products so far
  bool have = false;
  for (int b = 0; b < nbits; ++b) {
    if (b) {
      z = cmul(z, z);
      w = cmul(w, w);
    }
    const bool bit = ((e >> b) & 1) != 0;
    const cx pz = cmul(mk(zr, zi), z), pw = cmul(mk(wr, wi), w);
    // selects on the doubles themselves (a select between whole structs
    // can turn into a stack slot picked by address)
    const bool mul = bit && have, first = bit && !have;
    zr = mul ? pz.re : (first ? z.re : zr);
    zi = mul ? pz.im : (first ? z.im : zi);
    wr = mul ? pw.re : (first ? w.re : wr);
    wi = mul ? pw.im : (first ? w.im : wi);
    have = have || bit;
  }
  ze = mk(zr, zi);
  we = mk(wr, wi);
}

// ---- sin and cos as glibc's sincos computes them ------------------------
// The reference's arc angles (src/tracker.f90:159) reach glibc 2.35's
// sincos (flang -O2 merges the cos/sin pair into one call). Its dbl-64
// implementation is the IBM Accurate Mathematical Library method: the
// argument is reduced by pi/2 (hp0 + hp1 on 0.855 <= |x| < 2.426, a
// four-part pi/2 beyond), sin and cos of the reduced x + dx come from the
// table value at the nearest multiple of 1/128 combined with short
// polynomials for the remainder, or from a Taylor polynomial for |x| <
// 0.126. glibc's x86-64 sincos is built without FMA, and so is this. The
// constants below are that method's published coefficients; the table is
// phc_sincos_table.h. tests/test_math_host.py checks libm_sincos against
// the host's sincos bit for bit on every branch.
namespace sincos_detail {
constexpr double BIG = 0x1.8p45;  // big + |x| rounds |x| to a multiple of 1/128
constexpr double HP0 = 0x1.921fb54442d18p0, HP1 = 0x1.1a62633145c07p-54;
constexpr double SN3 = -1.66666666666664880952546298448555E-01;
constexpr double SN5 = 8.33333214285722277379541354343671E-03;
constexpr double CS2 = 4.99999999999999999999950396842453E-01;
constexpr double CS4 = -4.16666666666664434524222570944589E-02;
constexpr double CS6 = 1.38888874007937613028114285595617E-03;
constexpr double T1 = -0x1.5555555555555p-3, T2 = 0.0083333333333323288;
constexpr double T3 = -0.00019841269834414642, T4 = 2.755729806860771e-06;
constexpr double T5 = -2.502201484831839e-08;
constexpr double HPINV = 0x1.45f306dc9c883p-1, TOINT = 0x1.8p52;
constexpr double MP1 = 0x1.921fb58p0, MP2 = -0x1.dde973cp-27;
constexpr double PP3 = -0x1.cb3b398p-55, PP4 = -0x1.d747f23e32ed7p-83;

PHC_HD uint64_t bits_of(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, sizeof b);
  return b;
}

// sin(a + da) for |a| < 0.126
PHC_HD double taylor_sin(double a, double da) {
  const double xx = a * a;
  const double p = (((T5 * xx + T4) * xx + T3) * xx + T2) * xx + T1;
  const double t = (p * a - 0.5 * da) * xx + da;
  return a + t;
}

// cos(x + dx)
PHC_HD double cos_part(double x, double dx) {
  if (x < 0) dx = -dx;
  const double u = BIG + fabs(x);
  x = fabs(x) - (u - BIG) + dx;
  const double xx = x * x;
  const double s = x + x * xx * (SN3 + xx * SN5);
  const double c = xx * (CS2 + xx * (CS4 + xx * CS6));
  const double* e = SINCOS_TABLE + 4 * (uint32_t)bits_of(u);
  const double cor = (e[3] - s * e[1] - e[2] * c) - e[0] * s;
  return e[2] + cor;
}

// sin(x + dx)
PHC_HD double sin_part(double x, double dx) {
  if (fabs(x) < 0.126) return taylor_sin(x, dx);
  const double xold = x;
  if (x <= 0) dx = -dx;
  const double u = BIG + fabs(x);
  x = fabs(x) - (u - BIG);
  const double xx = x * x;
  const double s = x + (dx + x * xx * (SN3 + xx * SN5));
  const double c = x * dx + xx * (CS2 + xx * (CS4 + xx * CS6));
  const double* e = SINCOS_TABLE + 4 * (uint32_t)bits_of(u);
  const double cor = (e[1] + s * e[3] - e[0] * c) + e[2] * s;
  return copysign(e[0] + cor, xold);
}
}  // namespace sincos_detail

// (cos x, sin x) for finite |x| < 105414350 (arc angles lie in [0, 2 pi]).
PHC_HD cx sincos_body(double x) {
  using namespace sincos_detail;
  const uint32_t k = (uint32_t)(bits_of(x) >> 32) & 0x7fffffffu;
  if (k < 0x3e400000u) return mk(1.0, x);  // |x| < 2^-27
  if (k < 0x3feb6000u)                     // |x| < 0.855469
    return mk(cos_part(x, 0.0), sin_part(x, 0.0));
  if (k < 0x400368fdu) {  // |x| < 2.426265: x = +-(pi/2 - (a + da))
    const double y = HP0 - fabs(x);
    const double a = y + HP1;
    const double da = (y - a) + HP1;
    return mk(sin_part(a, da), copysign(cos_part(a, da), x));
  }
  if (k >= 0x7ff00000u) return mk(x - x, x - x);
  // x = n pi/2 + (a + da)
  const double t = x * HPINV + TOINT;
  const double xn = t - TOINT;
  const double y = (x - xn * MP1) - xn * MP2;
  const int n = (int)(bits_of(t) & 3u);
  double t1 = xn * PP3;
  const double t2 = y - t1;
  double da = (y - t2) - t1;
  t1 = xn * PP4;
  double a = t2 - t1;
  da += (t2 - a) - t1;
  if (n == 1 || n == 2) {
    a = -a;
    da = -da;
  }
  const double sv = sin_part(a, da), cv = cos_part(a, da);
  // with r = a + da: (cos x, sin x) = (cos r, sin r), (-sin r, cos r),
  // (-cos r, -sin r), (sin r, -cos r) for n = 0..3; r is negated above for
  // n = 1, 2, so sv and cv already carry those signs
  switch (n) {
    case 0: return mk(cv, sv);
    case 1: return mk(sv, cv);
    case 2: return mk(-cv, sv);
    default: return mk(sv, -cv);
  }
}

// Out of line: lanes may hold different angles.
PHC_HD_COLD cx libm_sincos(double x) { return sincos_body(x); }

// The same for an angle that every lane of the wavefront holds (one path
// per wavefront): taken into scalar registers first, so the branches are
// scalar and the table rows come through the scalar cache. Out of line
// too: inlined, it costs the 16-lane grid kernels their second wave per
// SIMD.
PHC_HD_COLD cx libm_sincos_uniform(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t b = sincos_detail::bits_of(x);
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)b);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(b >> 32));
  const uint64_t u = ((uint64_t)hi << 32) | lo;
  __builtin_memcpy(&x, &u, sizeof x);
#endif
  return sincos_body(x);
}

}  // namespace phc
