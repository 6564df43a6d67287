/*
 * phc_oracle.c -- TEST INFRASTRUCTURE ONLY: the checker, never the product.
 *
 * A CPU restatement in plain C of the reference path tracker
 * (EVADEVacStab/polyhomocont), function by function, with the floating-point
 * semantics of the reference as compiled by flang for x86-64 (no FMA):
 *   complex * complex    (ar*br - ai*bi, ar*bi + ai*br)
 *   real * complex       the real promoted to (s, 0), full product
 *   complex / complex    compiler-rt's __divdc3 (logb/scalbn-scaled)
 *   abs(complex)         libm cabs
 *   complex ** integer   binary powering (flang runtime)
 *   sum(array)           sequential from zero
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * load it (through oracle/__init__.py). It exports the entry points of
 * oracle/ref_driver.f90 with the prefix phc_orc_ over the structs of
 * include/phc.h and include/polyhomocont.h.
 *
 * Pinned by: tests/test_oracle.py (bit-for-bit against the compiled
 * reference oracle/_ref on every entry point, and against closed forms /
 * known answers of the reference's own tests).
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/polyhomocont.h"

typedef struct {
  double re, im;
} cx;

static cx mk(double re, double im) {
  cx r = {re, im};
  return r;
}
static cx cadd(cx a, cx b) { return mk(a.re + b.re, a.im + b.im); }
static cx csub(cx a, cx b) { return mk(a.re - b.re, a.im - b.im); }
static cx cneg(cx a) { return mk(-a.re, -a.im); }
static cx cmul(cx a, cx b) {
  return mk(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
/* (s, 0) * a, as flang evaluates real*complex and integer*complex */
static cx cmulr(double s, cx a) {
  return mk(a.re * s - a.im * 0.0, a.im * s + a.re * 0.0);
}
static double cabs_(cx a) { return cabs(a.re + I * a.im); }

/* compiler-rt __divdc3 */
static cx cdiv(cx num, cx den) {
  double a = num.re, b = num.im, c = den.re, d = den.im;
  int ilogbw = 0;
  double logbw = logb(fmax(fabs(c), fabs(d)));
  if (isfinite(logbw)) {
    ilogbw = (int)logbw;
    c = scalbn(c, -ilogbw);
    d = scalbn(d, -ilogbw);
  }
  double denom = c * c + d * d;
  double x = scalbn((a * c + b * d) / denom, -ilogbw);
  double y = scalbn((b * c - a * d) / denom, -ilogbw);
  if (isnan(x) && isnan(y)) {
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
static cx cdivr(cx a, double s) { return cdiv(a, mk(s, 0.0)); }

/* complex ** integer as the flang runtime computes it */
static cx zpowi(cx base, int e) {
  if (e == 0) return mk(1.0, 0.0);
  int inv = e < 0;
  unsigned u = inv ? (unsigned)(-(long long)e) : (unsigned)e;
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

static const double PI = 3.14159265358979323846;
static const double HUGE_ = 1.7976931348623157e308; /* huge(1.0_wp) */
static const double TINY_ = 2.2250738585072014e-308; /* tiny(1.0_wp) */

/* Fortran MAX and MIN as flang compiles them for x86-64 (maxsd/minsd): the
 * second argument unless the first compares greater (less). A NaN second
 * argument is returned, a NaN first argument is not; norm_inf's fold
 * r = max(r, abs(z(i))) is therefore NaN exactly when the last NaN component
 * is the last component, and is otherwise the largest modulus after the
 * last NaN. */
static double ftn_max(double a, double b) { return a > b ? a : b; }
static double ftn_min(double a, double b) { return a < b ? a : b; }
/* MAXVAL: NaN elements are ignored unless every element is NaN. */
static double maxval_fold(double r, double x, int *seen) {
  if (x != x) return *seen ? r : x;
  if (!*seen || r != r) {
    *seen = 1;
    return x;
  }
  return x > r ? x : r;
}

/* ------------------------------------------------- rng, src/random.f90:29-109 */
typedef struct {
  int64_t s1, s2;
} rng_t;
static const int64_t M1 = 2147483563, M2 = 2147483399, A1 = 40014, A2 = 40692;

static double rng_uniform(rng_t *g) { /* :64-78 */
  g->s1 = (A1 * g->s1) % M1;
  g->s2 = (A2 * g->s2) % M2;
  int64_t z = g->s1 - g->s2;
  if (z < 1) z += M1 - 1;
  return (double)z / (double)M1;
}
static void rng_seed(rng_t *g, int seed) { /* :43-62 */
  int64_t s = llabs((long long)seed);
  g->s1 = s % (M1 - 1) + 1;
  g->s2 = (s * 7919 + 104729) % (M2 - 1) + 1;
  for (int i = 0; i < 16; ++i) (void)rng_uniform(g);
}
static cx rng_unit_complex(rng_t *g) { /* :80-91 */
  double phi = 2.0 * PI * rng_uniform(g);
  return mk(cos(phi), sin(phi));
}
static cx rng_normal_complex(rng_t *g) { /* :93-109 */
  double r = sqrt(-log(rng_uniform(g)));
  double phi = 2.0 * PI * rng_uniform(g);
  return mk(r * cos(phi), r * sin(phi));
}

/* ----------------------------------------- sparse_system, src/sparse.f90 */
typedef struct {
  int n, neq, maxexp, nterms;
  int first[64];  /* 1-based offsets as in the reference: first(1) = 1 */
  int deg[64];
  int *hexps;     /* [nterms][n+1] */
  cx *coefs;      /* [nterms] */
} sys_t;

static void sys_free(sys_t *s) {
  free(s->hexps);
  free(s->coefs);
  s->hexps = NULL;
  s->coefs = NULL;
}

/* sparse_init + add_equation for each equation (src/sparse.f90:65-141):
 * invalid equations are rejected (not added), zero terms dropped, the w
 * exponent d - |a| appended. */
static void sys_build(const phc_user_system *u, sys_t *s) {
  memset(s, 0, sizeof(*s));
  s->n = u->n;
  s->first[0] = 1;
  int total = 0;
  for (int i = 0; i < u->neq; ++i) total += u->nterms[i];
  s->hexps = (int *)calloc((size_t)(total + 1) * (u->n + 1), sizeof(int));
  s->coefs = (cx *)calloc((size_t)total + 1, sizeof(cx));
  int off = 0;
  for (int i = 0; i < u->neq; ++i) {
    const int m = u->nterms[i];
    int ok = s->neq < s->n && u->n <= 63;
    for (int k = 0; k < m && ok; ++k)
      for (int j = 0; j < u->n; ++j)
        if (u->exps[(size_t)(off + k) * u->n + j] < 0) ok = 0;
    int d = 0;
    if (ok) {
      for (int k = 0; k < m; ++k) {
        cx c = mk(u->coefs[2 * (off + k)], u->coefs[2 * (off + k) + 1]);
        if (c.re == 0.0 && c.im == 0.0) continue;
        int sum = 0;
        for (int j = 0; j < u->n; ++j) sum += u->exps[(size_t)(off + k) * u->n + j];
        if (sum > d) d = sum;
      }
      if (d < 1) ok = 0;
    }
    if (ok) {
      int kept = 0;
      for (int k = 0; k < m; ++k) {
        cx c = mk(u->coefs[2 * (off + k)], u->coefs[2 * (off + k) + 1]);
        if (c.re == 0.0 && c.im == 0.0) continue;
        int *h = s->hexps + (size_t)s->nterms * (u->n + 1);
        int sum = 0;
        for (int j = 0; j < u->n; ++j) {
          h[j] = u->exps[(size_t)(off + k) * u->n + j];
          sum += h[j];
          if (h[j] > s->maxexp) s->maxexp = h[j];
        }
        h[u->n] = d - sum;
        if (h[u->n] > s->maxexp) s->maxexp = h[u->n];
        s->coefs[s->nterms] = c;
        s->nterms++;
        kept++;
      }
      s->deg[s->neq] = d;
      s->first[s->neq + 1] = s->first[s->neq] + kept;
      s->neq++;
    }
    off += m;
  }
}

/* eval_terms on the homogenized exponents (src/sparse.f90:223-280): power
 * table, prefix and suffix products over all m variables, f and jac. jac is
 * column-major n x m. */
static void eval_terms(const sys_t *s, const cx *v, cx *f, cx *jac) {
  const int m = s->n + 1, me = s->maxexp;
  cx *pw = (cx *)malloc(sizeof(cx) * (size_t)(me + 1) * m);
  cx pre[65], suf[66];
  for (int j = 0; j < m; ++j) pw[j] = mk(1.0, 0.0);
  for (int e = 1; e <= me; ++e)
    for (int j = 0; j < m; ++j) pw[e * m + j] = cmul(pw[(e - 1) * m + j], v[j]);
  for (int i = 0; i < s->neq; ++i) f[i] = mk(0.0, 0.0);
  for (int k = 0; k < s->neq * m; ++k) jac[k] = mk(0.0, 0.0);
  for (int i = 0; i < s->neq; ++i) {
    for (int k = s->first[i] - 1; k < s->first[i + 1] - 1; ++k) {
      const cx c = s->coefs[k];
      const int *ex = s->hexps + (size_t)k * m;
      pre[0] = mk(1.0, 0.0);
      for (int j = 1; j <= m; ++j) pre[j] = cmul(pre[j - 1], pw[ex[j - 1] * m + j - 1]);
      suf[m + 1] = mk(1.0, 0.0);
      for (int j = m; j >= 1; --j) suf[j] = cmul(pw[ex[j - 1] * m + j - 1], suf[j + 1]);
      f[i] = cadd(f[i], cmul(c, pre[m]));
      for (int j = 1; j <= m; ++j) {
        const int e = ex[j - 1];
        if (e > 0) {
          cx t = cmulr((double)e, c);
          t = cmul(t, pw[(e - 1) * m + j - 1]);
          t = cmul(t, pre[j - 1]);
          t = cmul(t, suf[j + 1]);
          jac[i + (j - 1) * s->n] = cadd(jac[i + (j - 1) * s->n], t);
        }
      }
    }
  }
  free(pw);
}

/* ------------------------- total_degree_homotopy, src/homotopy.f90:78-231 */
typedef struct {
  const sys_t *sys;
  int n, nstart;
  cx gamma, b[64], patch[65];
} hom_t;

static int hom_init(hom_t *h, const sys_t *s, rng_t *g) { /* :104-149 */
  h->sys = s;
  h->n = s->n;
  if (h->n < 1 || s->neq != h->n) return 1;
  int64_t ns = 1;
  for (int i = 0; i < h->n; ++i) {
    if (s->deg[i] < 1) return 1;
  }
  for (int i = 0; i < h->n; ++i) {
    ns *= s->deg[i];
    if (ns > 2147483647LL) return 1;
  }
  h->nstart = (int)ns;
  h->gamma = rng_unit_complex(g);
  for (int i = 0; i < h->n; ++i) h->b[i] = rng_unit_complex(g);
  for (int i = 0; i <= h->n; ++i) h->patch[i] = rng_normal_complex(g);
  double ss = 0.0; /* sum(abs(patch)**2) */
  for (int i = 0; i <= h->n; ++i) {
    const double a = cabs_(h->patch[i]);
    ss = ss + a * a;
  }
  const double nrm = sqrt(ss);
  for (int i = 0; i <= h->n; ++i) h->patch[i] = cdivr(h->patch[i], nrm);
  return 0;
}

static void hom_start(const hom_t *h, int k, cx *z) { /* :151-177 */
  int rest = k - 1;
  for (int i = 0; i < h->n; ++i) {
    const int d = h->sys->deg[i];
    const int digit = rest % d;
    rest = rest / d;
    const double phi =
        (atan2(h->b[i].im, h->b[i].re) + 2.0 * PI * (double)digit) / (double)d;
    z[i] = mk(cos(phi), sin(phi));
  }
  z[h->n] = mk(1.0, 0.0);
  cx s = mk(0.0, 0.0);
  for (int i = 0; i <= h->n; ++i) s = cadd(s, cmul(h->patch[i], z[i]));
  for (int i = 0; i <= h->n; ++i) z[i] = cdiv(z[i], s);
}

/* tdh_evaluate (src/homotopy.f90:188-231); hz column-major N x N */
static void hom_eval(const hom_t *h, const cx *z, cx t, cx *hv, cx *hz, cx *ht) {
  const int n = h->n, N = n + 1;
  cx f[64], jf[64 * 65];
  eval_terms(h->sys, z, f, jf);
  const cx w = z[n];
  const cx gt = cmul(h->gamma, t);
  const cx omt = mk(1.0 - t.re, 0.0 - t.im);
  for (int j = 0; j < N; ++j)
    for (int i = 0; i < n; ++i) hz[i + j * N] = cmul(omt, jf[i + j * n]);
  for (int i = 0; i < n; ++i) {
    const int d = h->sys->deg[i];
    const cx zd1 = zpowi(z[i], d - 1);
    const cx wd1 = zpowi(w, d - 1);
    const cx g = csub(cmul(zd1, z[i]), cmul(cmul(h->b[i], wd1), w));
    hv[i] = cadd(cmul(gt, g), cmul(omt, f[i]));
    ht[i] = csub(cmul(h->gamma, g), f[i]);
    hz[i + i * N] = cadd(hz[i + i * N], cmul(cmulr((double)d, gt), zd1));
    hz[i + n * N] =
        csub(hz[i + n * N], cmul(cmulr((double)d, cmul(gt, h->b[i])), wd1));
  }
  cx s = mk(0.0, 0.0);
  for (int j = 0; j < N; ++j) s = cadd(s, cmul(h->patch[j], z[j]));
  hv[n] = mk(s.re - 1.0, s.im - 0.0);
  for (int j = 0; j < N; ++j) hz[n + j * N] = h->patch[j];
  ht[n] = mk(0.0, 0.0);
}

/* ------------------------------------------------ linalg, src/linalg.f90 */
static int lu_factor(int n, cx *a, int *ipiv) { /* :26-78, column-major */
  cx tmp[65];
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = cabs_(a[k + k * n]);
    for (int i = k + 1; i < n; ++i) {
      if (cabs_(a[i + k * n]) > amax) {
        amax = cabs_(a[i + k * n]);
        p = i;
      }
    }
    ipiv[k] = p;
    if (amax == 0.0) return 2;
    if (p != k) {
      for (int j = 0; j < n; ++j) tmp[j] = a[k + j * n];
      for (int j = 0; j < n; ++j) a[k + j * n] = a[p + j * n];
      for (int j = 0; j < n; ++j) a[p + j * n] = tmp[j];
    }
    const cx piv = a[k + k * n];
    for (int i = k + 1; i < n; ++i) {
      a[i + k * n] = cdiv(a[i + k * n], piv);
      for (int j = k + 1; j < n; ++j)
        a[i + j * n] = csub(a[i + j * n], cmul(a[i + k * n], a[k + j * n]));
    }
  }
  return 0;
}

static void lu_solve(int n, const cx *a, const int *ipiv, cx *b) { /* :80-114 */
  for (int k = 0; k < n; ++k) {
    if (ipiv[k] != k) {
      cx t = b[k];
      b[k] = b[ipiv[k]];
      b[ipiv[k]] = t;
    }
  }
  for (int i = 1; i < n; ++i) {
    cx s = mk(0.0, 0.0);
    for (int j = 0; j < i; ++j) s = cadd(s, cmul(a[i + j * n], b[j]));
    b[i] = csub(b[i], s);
  }
  for (int i = n - 1; i >= 0; --i) {
    cx s = mk(0.0, 0.0);
    for (int j = i + 1; j < n; ++j) s = cadd(s, cmul(a[i + j * n], b[j]));
    b[i] = cdiv(csub(b[i], s), a[i + i * n]);
  }
}

static double cond1(int n, const cx *a) { /* :135-174 */
  double anorm = 0.0;
  for (int j = 0; j < n; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + cabs_(a[i + j * n]);
    anorm = ftn_max(anorm, s);
  }
  cx lu[65 * 65], col[65];
  int ipiv[65];
  memcpy(lu, a, sizeof(cx) * (size_t)n * n);
  if (lu_factor(n, lu, ipiv)) return HUGE_;
  double ainv = 0.0;
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < n; ++i) col[i] = mk(0.0, 0.0);
    col[j] = mk(1.0, 0.0);
    lu_solve(n, lu, ipiv, col);
    double s = 0.0;
    for (int i = 0; i < n; ++i) s = s + cabs_(col[i]);
    ainv = ftn_max(ainv, s);
  }
  return anorm * ainv;
}

static double norm_inf(int n, const cx *z) { /* :176-189 */
  double r = 0.0;
  for (int i = 0; i < n; ++i) r = ftn_max(r, cabs_(z[i]));
  return r;
}

/* --------------------------------------------- tracker, src/tracker.f90 */
typedef struct {
  double step_init, step_max, step_min, ctol, contraction;
  int maxit, grow_after, max_steps;
} topts_t;
typedef struct {
  int arc;
  cx t0, t1;
  double radius, theta0, dtheta;
} seg_t_;
typedef struct {
  int status, naccepted, nrejected;
  double s, step;
  cx t;
} info_t;

static cx seg_t(const seg_t_ *g, double s) { /* :142-162 */
  if (!g->arc) {
    if (s >= 1.0) return g->t1;
    return cadd(g->t0, cmulr(s, csub(g->t1, g->t0)));
  }
  const double th = g->theta0 + s * g->dtheta;
  return cmulr(g->radius, mk(cos(th), sin(th)));
}
static cx seg_dt(const seg_t_ *g, double s) { /* :164-177 */
  if (!g->arc) return csub(g->t1, g->t0);
  return cmul(mk(0.0, g->dtheta), seg_t(g, s));
}

static int tangent(const hom_t *h, const seg_t_ *g, const cx *z, double s,
                   cx *dz) { /* :295-317 */
  const int N = h->n + 1;
  cx hv[65], hz[65 * 65];
  int ipiv[65];
  hom_eval(h, z, seg_t(g, s), hv, hz, dz);
  if (lu_factor(N, hz, ipiv)) return 0;
  const cx dt = seg_dt(g, s);
  for (int i = 0; i < N; ++i) dz[i] = cneg(cmul(dz[i], dt));
  lu_solve(N, hz, ipiv, dz);
  return 1;
}

static int predict_rk4(const hom_t *h, const seg_t_ *g, const cx *z, double s,
                       double ds, cx *zp) { /* :266-293 */
  const int N = h->n + 1;
  cx k1[65], k2[65], k3[65], k4[65], y[65];
  if (!tangent(h, g, z, s, k1)) return 0;
  for (int i = 0; i < N; ++i) y[i] = cadd(z[i], cmulr(0.5 * ds, k1[i]));
  if (!tangent(h, g, y, s + 0.5 * ds, k2)) return 0;
  for (int i = 0; i < N; ++i) y[i] = cadd(z[i], cmulr(0.5 * ds, k2[i]));
  if (!tangent(h, g, y, s + 0.5 * ds, k3)) return 0;
  for (int i = 0; i < N; ++i) y[i] = cadd(z[i], cmulr(ds, k3[i]));
  if (!tangent(h, g, y, s + ds, k4)) return 0;
  for (int i = 0; i < N; ++i) {
    cx acc = cadd(cadd(cadd(k1[i], cmulr(2.0, k2[i])), cmulr(2.0, k3[i])), k4[i]);
    zp[i] = cadd(z[i], cmulr(ds / 6.0, acc));
  }
  return 1;
}

static int newton_correct(const hom_t *h, cx *z, cx t, const topts_t *o) {
  /* :319-360 */
  const int N = h->n + 1;
  cx hv[65], hz[65 * 65], ht[65];
  int ipiv[65];
  double dprev = HUGE_;
  for (int it = 1; it <= o->maxit; ++it) {
    hom_eval(h, z, t, hv, hz, ht);
    if (lu_factor(N, hz, ipiv)) return 0;
    for (int i = 0; i < N; ++i) hv[i] = cneg(hv[i]);
    lu_solve(N, hz, ipiv, hv);
    for (int i = 0; i < N; ++i) z[i] = cadd(z[i], hv[i]);
    const double dn = norm_inf(N, hv);
    if (dn <= o->ctol * norm_inf(N, z)) return 1;
    if (it > 1 && dn > o->contraction * dprev) return 0;
    dprev = dn;
  }
  return 0;
}

static void track_segment(const hom_t *h, cx *z, const seg_t_ *g,
                          const topts_t *o, info_t *info) { /* :196-264 */
  const int N = h->n + 1;
  cx zp[65];
  double s = 0.0, step = ftn_min(o->step_init, o->step_max);
  int nsucc = 0;
  info->naccepted = info->nrejected = 0;
  info->status = 0;
  while (s < 1.0) {
    if (info->naccepted + info->nrejected >= o->max_steps) {
      info->status = 2;
      break;
    }
    double ds = ftn_min(step, 1.0 - s);
    double sn = s + ds;
    if (1.0 - sn <= 1.0e-3 * step) {
      sn = 1.0;
      ds = 1.0 - s;
    }
    int ok = predict_rk4(h, g, z, s, ds, zp);
    if (ok) ok = newton_correct(h, zp, seg_t(g, sn), o);
    if (ok) {
      memcpy(z, zp, sizeof(cx) * N);
      s = sn;
      info->naccepted++;
      nsucc++;
      if (nsucc >= o->grow_after) {
        step = ftn_min(2.0 * step, o->step_max);
        nsucc = 0;
      }
    } else {
      info->nrejected++;
      nsucc = 0;
      step = 0.5 * step;
      if (step < o->step_min) {
        info->status = 1;
        break;
      }
    }
  }
  info->s = s;
  info->t = seg_t(g, s);
  info->step = step;
}

/* --------------------------------------------- endgame, src/endgame.f90 */
typedef struct {
  double t_start, radius_factor, radius_min, loop_radius, agreement, tol,
      residual_tol, closure_factor;
  int nsamples, max_winding;
} eopts_t;
typedef struct {
  int status, winding, naccepted, nrejected;
  double error, residual, radius;
  cx z[65];
} egres_t;

static double backward_error(const hom_t *h, const cx *z) { /* :269-285 */
  const int N = h->n + 1;
  cx hv[65], hz[65 * 65], ht[65];
  hom_eval(h, z, mk(0.0, 0.0), hv, hz, ht);
  double rmax = -HUGE_;
  int seen = 0;
  for (int i = 0; i < N; ++i) {
    double s = 0.0;
    for (int j = 0; j < N; ++j) s = s + cabs_(hz[i + j * N]);
    rmax = maxval_fold(rmax, s, &seen);
  }
  return norm_inf(N, hv) / ftn_max(rmax * norm_inf(N, z), TINY_);
}

static int better(double resid, double err, double best_resid, double best_err,
                  double rtol) { /* :287-311 */
  const int valid = resid <= rtol, best_valid = best_resid <= rtol;
  if (valid != best_valid) return valid;
  return err < best_err;
}

static int loops(const hom_t *h, const cx *z, double r, const eopts_t *eo,
                 int max_loops, topts_t *so, cx *est, int *c, egres_t *res) {
  /* :313-378 */
  const int N = h->n + 1;
  info_t info;
  cx zc[65], zprev[65], zsum[65], d[65];
  const double dtheta = 2.0 * PI / (double)eo->nsamples;
  const double floor_ = 10.0 * so->ctol * norm_inf(N, z);
  memcpy(zc, z, sizeof(cx) * N);
  for (int i = 0; i < N; ++i) zsum[i] = mk(0.0, 0.0);
  int nsamp = 0;
  double spacing = HUGE_;
  for (*c = 1; *c <= max_loops; ++*c) {
    for (int j = 1; j <= eo->nsamples; ++j) {
      for (int i = 0; i < N; ++i) zsum[i] = cadd(zsum[i], zc[i]);
      nsamp++;
      memcpy(zprev, zc, sizeof(cx) * N);
      seg_t_ g = {1, {0, 0}, {0, 0}, r, (double)(j - 1) * dtheta, dtheta};
      track_segment(h, zc, &g, so, &info);
      so->step_init = ftn_min(ftn_max(info.step, so->step_min), 1.0);
      res->naccepted += info.naccepted;
      res->nrejected += info.nrejected;
      if (info.status != 0) return 0;
      for (int i = 0; i < N; ++i) d[i] = csub(zc[i], zprev[i]);
      spacing = ftn_min(spacing, norm_inf(N, d));
    }
    for (int i = 0; i < N; ++i) d[i] = csub(zc[i], z[i]);
    if (norm_inf(N, d) <= ftn_max(eo->closure_factor * spacing, floor_)) {
      for (int i = 0; i < N; ++i) est[i] = cdivr(zsum[i], (double)nsamp);
      return 1;
    }
  }
  return 0;
}

static void cauchy_endgame(const hom_t *h, cx *z, const eopts_t *eo,
                           const topts_t *to, egres_t *res) { /* :124-267 */
  const int N = h->n + 1;
  topts_t so = *to;
  so.step_init = 1.0;
  so.step_max = 1.0;
  so.grow_after = 2;
  info_t info;
  cx est[65], prev[65], zold[65], d[65];
  res->status = 2;
  res->winding = 0;
  res->error = HUGE_;
  res->residual = HUGE_;
  res->radius = 0.0;
  res->naccepted = res->nrejected = 0;
  memcpy(res->z, z, sizeof(cx) * N);
  double r = eo->t_start, best_resid = HUGE_, dprev = -1.0, cest = -1.0,
         cest_prev = -1.0;
  int have_prev = 0, gate = 0, max_loops = eo->max_winding, c = 0, ok;

#define RECORD()                                                            \
  do {                                                                      \
    double err_ = HUGE_;                                                    \
    if (have_prev) {                                                        \
      for (int i = 0; i < N; ++i) d[i] = csub(est[i], prev[i]);             \
      err_ = norm_inf(N, d) / ftn_max(norm_inf(N, est), TINY_);                \
    }                                                                       \
    const double resid_ = backward_error(h, est);                           \
    if (res->status == 2 ||                                                 \
        better(resid_, err_, best_resid, res->error, eo->residual_tol)) {   \
      res->status = 1;                                                      \
      memcpy(res->z, est, sizeof(cx) * N);                                  \
      res->winding = c;                                                     \
      res->error = err_;                                                    \
      res->residual = resid_;                                               \
      res->radius = r;                                                      \
      best_resid = resid_;                                                  \
    }                                                                       \
    if (err_ <= eo->tol && resid_ <= eo->residual_tol) res->status = 0;     \
  } while (0)

  for (;;) {
    if (gate || r < eo->loop_radius) {
      if (!gate) max_loops = eo->max_winding;
      ok = loops(h, z, r, eo, max_loops, &so, est, &c, res);
    } else {
      ok = 0;
    }
    if (!ok) {
      have_prev = 0;
      gate = 0;
    }
    if (ok) {
      RECORD();
      if (res->status == 0) return;
      memcpy(prev, est, sizeof(cx) * N);
      have_prev = 1;
    }
    const double rnew = eo->radius_factor * r;
    if (rnew < eo->radius_min) return;
    memcpy(zold, z, sizeof(cx) * N);
    seg_t_ g = {0, {r, 0.0}, {rnew, 0.0}, 0, 0, 0};
    track_segment(h, z, &g, &so, &info);
    so.step_init = ftn_min(ftn_max(info.step, so.step_min), 1.0);
    res->naccepted += info.naccepted;
    res->nrejected += info.nrejected;
    if (info.status != 0) {
      memcpy(z, zold, sizeof(cx) * N);
      if (!(gate || r < eo->loop_radius)) {
        ok = loops(h, z, r, eo, eo->max_winding, &so, est, &c, res);
        have_prev = 0;
        if (ok) RECORD();
      }
      return;
    }
    r = rnew;
    for (int i = 0; i < N; ++i) d[i] = csub(z[i], zold[i]);
    const double dd = norm_inf(N, d);
    if (!gate && dprev > 0.0 && dd > 0.0) {
      cest_prev = cest;
      cest = -1.0;
      if (dprev > dd) cest = log(1.0 / eo->radius_factor) / log(dprev / dd);
      if (cest >= 0.5 && cest <= eo->max_winding + 0.5 && cest_prev > 0.0) {
        if (fabs(cest - cest_prev) <= eo->agreement * ftn_max(1.0, cest)) {
          gate = 1;
          int nc = (int)lround(cest);
          max_loops = eo->max_winding < 2 * nc + 2 ? eo->max_winding : 2 * nc + 2;
        }
      }
    }
    dprev = dd;
  }
#undef RECORD
}

/* ------------------------------------------------- solve, src/solve.f90 */
static void options(const phc_solve_options *c, topts_t *to, eopts_t *eo) {
  to->step_init = c->tracker.step_init;
  to->step_max = c->tracker.step_max;
  to->step_min = c->tracker.step_min;
  to->ctol = c->tracker.corrector_tol;
  to->contraction = c->tracker.contraction;
  to->maxit = c->tracker.corrector_maxit;
  to->grow_after = c->tracker.grow_after;
  to->max_steps = c->tracker.max_steps;
  eo->t_start = c->endgame.t_start;
  eo->radius_factor = c->endgame.radius_factor;
  eo->radius_min = c->endgame.radius_min;
  eo->loop_radius = c->endgame.loop_radius;
  eo->agreement = c->endgame.winding_agreement;
  eo->tol = c->endgame.tol;
  eo->residual_tol = c->endgame.residual_tol;
  eo->closure_factor = c->endgame.closure_factor;
  eo->nsamples = c->endgame.nsamples;
  eo->max_winding = c->endgame.max_winding;
}

typedef struct {
  int status, winding, eg_status, naccepted, nrejected;
  double t, eg_error, eg_residual;
} prec_t;

/* endpoint (src/solve.f90:217-252) */
static void endpoint(const hom_t *h, cx *z, const phc_solve_options *c,
                     prec_t *p) {
  topts_t to;
  eopts_t eo;
  options(c, &to, &eo);
  const double t_end = c->use_endgame ? eo.t_start : 0.0;
  seg_t_ g = {0, {1.0, 0.0}, {t_end, 0.0}, 0, 0, 0};
  info_t info;
  track_segment(h, z, &g, &to, &info);
  p->status = info.status;
  p->t = info.t.re;
  p->naccepted = info.naccepted;
  p->nrejected = info.nrejected;
  p->winding = 0;
  p->eg_status = 0;
  p->eg_error = 0.0;
  p->eg_residual = 0.0;
  if (info.status != 0 || !c->use_endgame) return;
  egres_t eg;
  cauchy_endgame(h, z, &eo, &to, &eg);
  p->naccepted += eg.naccepted;
  p->nrejected += eg.nrejected;
  if (eg.status == 2) {
    p->status = 3;
    return;
  }
  memcpy(z, eg.z, sizeof(cx) * (h->n + 1));
  p->t = 0.0;
  p->winding = eg.winding;
  p->eg_status = eg.status;
  p->eg_error = eg.error;
  p->eg_residual = eg.residual;
}

static void refine(const hom_t *h, cx *z, int maxit) { /* :254-286 */
  const int N = h->n + 1;
  cx hv[65], hz[65 * 65], ht[65];
  int ipiv[65];
  double dprev = HUGE_;
  for (int it = 1; it <= maxit; ++it) {
    hom_eval(h, z, mk(0.0, 0.0), hv, hz, ht);
    if (lu_factor(N, hz, ipiv)) return;
    for (int i = 0; i < N; ++i) hv[i] = cneg(hv[i]);
    lu_solve(N, hz, ipiv, hv);
    const double dn = norm_inf(N, hv);
    if (dn > dprev) return;
    for (int i = 0; i < N; ++i) z[i] = cadd(z[i], hv[i]);
    if (dn <= 4.0 * 2.220446049250313e-16 * norm_inf(N, z)) return;
    dprev = dn;
  }
}

typedef struct {
  int at_inf, singular, is_real, sol_idx, sing_idx;
  double cond, residual;
  cx x[64];
} cls_t;

static void classify(const hom_t *h, cx *z, const phc_solve_options *c,
                     const prec_t *p, cls_t *q) { /* :288-328 */
  const int n = h->n, N = n + 1;
  cx hv[65], hz[65 * 65], ht[65];
  hom_eval(h, z, mk(0.0, 0.0), hv, hz, ht);
  q->cond = cond1(N, hz);
  q->singular = p->winding > 1 || q->cond > c->singular_cond;
  if (!q->singular) {
    refine(h, z, c->refine_maxit);
    hom_eval(h, z, mk(0.0, 0.0), hv, hz, ht);
    q->cond = cond1(N, hz);
  }
  const double zn = norm_inf(N, z);
  double rmax = -HUGE_;
  int seen = 0;
  for (int i = 0; i < n; ++i) {
    /* abs(h(i)) / znorm**deg(i), real ** integer by repeated squaring */
    const int d = h->sys->deg[i];
    double pw = zn, acc = 1.0;
    int e = d, first = 1;
    while (e > 0) { /* llvm.powi order: square-and-multiply from the top bit */
      (void)first;
      if (e & 1) acc = acc * pw;
      e >>= 1;
      if (e) pw = pw * pw;
    }
    rmax = maxval_fold(rmax, cabs_(hv[i]) / acc, &seen);
  }
  q->residual = rmax;
  const double lim =
      ftn_max(c->infinity_tol, 10.0 * ftn_min(p->eg_error, 1.0e-4));
  q->at_inf = cabs_(z[n]) <= lim * zn;
  q->is_real = 0;
  if (!q->at_inf) {
    double im = -HUGE_;
    int seen = 0;
    for (int i = 0; i < n; ++i) {
      q->x[i] = cdiv(z[i], z[n]);
      im = maxval_fold(im, fabs(q->x[i].im), &seen);
    }
    q->is_real = im <= c->real_tol * ftn_max(1.0, norm_inf(n, q->x));
  }
}

static int find(int n, const cx *x, const cx *pts, int npts, double tol) {
  /* :379-400 */
  const double scale = ftn_max(1.0, norm_inf(n, x));
  cx d[64];
  for (int j = 0; j < npts; ++j) {
    for (int i = 0; i < n; ++i) d[i] = csub(x[i], pts[(size_t)j * n + i]);
    if (norm_inf(n, d) <= tol * scale) return j + 1;
  }
  return 0;
}

/* ------------------------------------------------------ exported entries */
static void put_z(double *p, size_t k, const cx *z, int n) {
  for (int j = 0; j < n; ++j) {
    p[2 * (k * n + j)] = z[j].re;
    p[2 * (k * n + j) + 1] = z[j].im;
  }
}

static int setup(const phc_user_system *u, int seed, sys_t *s, hom_t *h) {
  rng_t g;
  sys_build(u, s);
  rng_seed(&g, seed);
  return hom_init(h, s, &g);
}

int phc_orc_homotopy(const phc_user_system *u, int32_t seed, double *gamma,
                     double *b, double *patch, int64_t *nstart) {
  sys_t s;
  hom_t h;
  int st = setup(u, seed, &s, &h);
  if (!st) {
    if (gamma) put_z(gamma, 0, &h.gamma, 1);
    if (b) put_z(b, 0, h.b, h.n);
    if (patch) put_z(patch, 0, h.patch, h.n + 1);
    if (nstart) *nstart = h.nstart;
  }
  sys_free(&s);
  return st;
}

int phc_orc_start_points(const phc_user_system *u, int32_t seed, int64_t first,
                         int64_t count, double *start) {
  sys_t s;
  hom_t h;
  int st = setup(u, seed, &s, &h);
  cx z[65];
  for (int64_t k = 0; !st && k < count; ++k) {
    hom_start(&h, (int)(first + k + 1), z);
    put_z(start, (size_t)k, z, h.n + 1);
  }
  sys_free(&s);
  return st;
}

int phc_orc_endpoints(const phc_user_system *u, const phc_solve_options *c,
                      int64_t first, int64_t count, const phc_endpoints *o) {
  sys_t s;
  hom_t h;
  int st = setup(u, c->seed, &s, &h);
  cx z[65];
  prec_t p;
  for (int64_t k = 0; !st && k < count; ++k) {
    hom_start(&h, (int)(first + k + 1), z);
    endpoint(&h, z, c, &p);
    o->status[k] = p.status;
    o->t[k] = p.t;
    put_z(o->z, (size_t)k, z, h.n + 1);
    o->winding[k] = p.winding;
    o->eg_status[k] = p.eg_status;
    o->eg_error[k] = p.eg_error;
    if (o->eg_residual) o->eg_residual[k] = p.eg_residual;
    o->naccepted[k] = p.naccepted;
    o->nrejected[k] = p.nrejected;
  }
  sys_free(&s);
  return st;
}

int phc_orc_solve(const phc_user_system *u, const phc_solve_options *c,
                  phc_solve_summary *sm, const phc_path_results *pr) {
  /* solve (src/solve.f90:154-215) */
  sys_t s;
  hom_t h;
  int st = setup(u, c->seed, &s, &h);
  if (st) {
    sys_free(&s);
    return st;
  }
  const int n = h.n, N = n + 1, np = h.nstart;
  cx *sols = (cx *)calloc((size_t)np * n + 1, sizeof(cx));
  cx *ssols = (cx *)calloc((size_t)np * n + 1, sizeof(cx));
  int *mult = (int *)calloc((size_t)np + 1, sizeof(int));
  sm->n = n;
  sm->npaths = np;
  sm->nsolutions = sm->nreal = sm->nsingular = sm->ninfinite = 0;
  sm->nfailed = sm->nduplicates = 0;
  cx z[65];
  for (int k = 0; k < np; ++k) {
    prec_t p;
    cls_t q;
    memset(&q, 0, sizeof(q));
    hom_start(&h, k + 1, z);
    endpoint(&h, z, c, &p);
    if (p.status != 0) {
      sm->nfailed++;
    } else {
      classify(&h, z, c, &p, &q);
      if (q.at_inf) {
        sm->ninfinite++;
      } else if (q.singular) {
        int j = find(n, q.x, ssols, sm->nsingular, c->singular_dedup_tol);
        if (j > 0) {
          mult[j - 1]++;
        } else {
          memcpy(ssols + (size_t)sm->nsingular * n, q.x, sizeof(cx) * n);
          mult[sm->nsingular] = 1;
          j = ++sm->nsingular;
        }
        q.sing_idx = j;
      } else {
        int j = find(n, q.x, sols, sm->nsolutions, c->dedup_tol);
        if (j > 0) {
          sm->nduplicates++;
        } else {
          memcpy(sols + (size_t)sm->nsolutions * n, q.x, sizeof(cx) * n);
          j = ++sm->nsolutions;
        }
        q.sol_idx = j;
      }
    }
    if (pr) {
      if (pr->status) ((int32_t *)pr->status)[k] = p.status;
      if (pr->t) ((double *)pr->t)[k] = p.t;
      if (pr->z) put_z((double *)pr->z, (size_t)k, z, N);
      if (pr->x) {
        cx zero[64] = {{0, 0}};
        put_z((double *)pr->x, (size_t)k, (p.status == 0 && !q.at_inf) ? q.x : zero, n);
      }
      if (pr->at_infinity) ((int32_t *)pr->at_infinity)[k] = q.at_inf;
      if (pr->singular) ((int32_t *)pr->singular)[k] = q.singular;
      if (pr->is_real) ((int32_t *)pr->is_real)[k] = q.is_real;
      if (pr->winding) ((int32_t *)pr->winding)[k] = p.winding;
      if (pr->eg_status) ((int32_t *)pr->eg_status)[k] = p.eg_status;
      if (pr->eg_error) ((double *)pr->eg_error)[k] = p.eg_error;
      if (pr->cond) ((double *)pr->cond)[k] = q.cond;
      if (pr->residual) ((double *)pr->residual)[k] = q.residual;
      if (pr->solution_index) ((int32_t *)pr->solution_index)[k] = q.sol_idx;
      if (pr->singular_index) ((int32_t *)pr->singular_index)[k] = q.sing_idx;
      if (pr->naccepted) ((int32_t *)pr->naccepted)[k] = p.naccepted;
      if (pr->nrejected) ((int32_t *)pr->nrejected)[k] = p.nrejected;
    }
  }
  /* collect_real (src/solve.f90:402-426) */
  for (int j = 0; j < sm->nsolutions; ++j) {
    const cx *x = sols + (size_t)j * n;
    double im = -HUGE_;
    int seen = 0;
    for (int i = 0; i < n; ++i) im = maxval_fold(im, fabs(x[i].im), &seen);
    if (im <= c->real_tol * ftn_max(1.0, norm_inf(n, x))) {
      if (sm->real_solutions)
        for (int i = 0; i < n; ++i) sm->real_solutions[(size_t)sm->nreal * n + i] = x[i].re;
      sm->nreal++;
    }
  }
  if (sm->solutions) put_z(sm->solutions, 0, sols, n * sm->nsolutions);
  if (sm->singular_solutions) put_z(sm->singular_solutions, 0, ssols, n * sm->nsingular);
  if (sm->multiplicities) memcpy(sm->multiplicities, mult, sizeof(int) * sm->nsingular);
  free(sols);
  free(ssols);
  free(mult);
  sys_free(&s);
  return 0;
}

int phc_orc_eval(const phc_user_system *u, int32_t seed, const double *z,
                 const double *t, double *hv, double *hz, double *ht) {
  sys_t s;
  hom_t h;
  int st = setup(u, seed, &s, &h);
  if (!st) hom_eval(&h, (const cx *)z, mk(t[0], t[1]), (cx *)hv, (cx *)hz, (cx *)ht);
  sys_free(&s);
  return st;
}

int phc_orc_eval_sparse(const phc_user_system *u, const double *z, double *f,
                        double *jac) {
  sys_t s;
  sys_build(u, &s);
  eval_terms(&s, (const cx *)z, (cx *)f, (cx *)jac);
  sys_free(&s);
  return 0;
}

int phc_orc_track_segment(const phc_user_system *u, int32_t seed,
                          const phc_tracker_options *ct, int32_t kind,
                          const double *t0, const double *t1, double radius,
                          double theta0, double dtheta, double *z,
                          int32_t *st_out, double *s_out, double *t_out,
                          double *step_out, int32_t *counts) {
  sys_t s;
  hom_t h;
  int st = setup(u, seed, &s, &h);
  if (!st) {
    topts_t to = {ct->step_init, ct->step_max, ct->step_min, ct->corrector_tol,
                  ct->contraction, ct->corrector_maxit, ct->grow_after,
                  ct->max_steps};
    seg_t_ g = {kind == 2, mk(t0[0], t0[1]), mk(t1[0], t1[1]), radius, theta0, dtheta};
    info_t info;
    track_segment(&h, (cx *)z, &g, &to, &info);
    *st_out = info.status;
    *s_out = info.s;
    t_out[0] = info.t.re;
    t_out[1] = info.t.im;
    *step_out = info.step;
    counts[0] = info.naccepted;
    counts[1] = info.nrejected;
  }
  sys_free(&s);
  return st;
}

void phc_orc_rng(int32_t seed, int32_t kind, int64_t count, double *out) {
  rng_t g;
  rng_seed(&g, seed);
  for (int64_t k = 0; k < count; ++k) {
    if (kind == 0) {
      out[k] = rng_uniform(&g);
    } else {
      cx z = kind == 1 ? rng_unit_complex(&g) : rng_normal_complex(&g);
      out[2 * k] = z.re;
      out[2 * k + 1] = z.im;
    }
  }
}

int phc_orc_lu(int32_t n, double *a, int32_t *ipiv, double *b) {
  int piv[65];
  int st = lu_factor(n, (cx *)a, piv);
  for (int i = 0; i < n; ++i) ipiv[i] = piv[i] + 1;
  if (st) return st;
  lu_solve(n, (const cx *)a, piv, (cx *)b);
  return 0;
}

double phc_orc_cond1(int32_t n, const double *a) { return cond1(n, (const cx *)a); }
