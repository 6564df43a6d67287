// phc_wave.h -- the linearization unit on a grid of lanes: LP = min(G*G,
// 64) lanes per path (N <= G): four paths per wavefront for G = 4, one path
// per wavefront for G = 8 and G = 16.
//
// Layout. The path's lanes form R = LP / G groups of G lanes; lane
// q = r G + i (group r, lane i) holds the entries (i, j) of the N x N
// system for the E = G / R columns j = r + R e, e = 0..E-1 (E = 1 for
// G <= 8, 4 for G = 16). Row i lives in lane i of every group. Each group
// is a full replica of the path state machine (lane i of every group holds
// component i of the path's vectors), which is how the rest of the
// tracker sees it.
//
// Unit, at zq and t (src/homotopy.f90:488-511, src/sparse.f90:229-278,
// src/linalg.f90:26-114):
//   1. group 0 writes the power table T[e G + j] = zq_j^e to LDS;
//   2. all lanes of the path evaluate the term values c_k prod(...) of all
//      terms, round robin, into V[k];
//   3. the slot of entry (i, j) sums dF_i/dz_j over its job list (the
//      terms of equation i with a nonzero exponent of z_j, in term order),
//      each job a product of the term's factors in the reference's
//      association; the slot of (G-1, e) sums F_e from V in term order;
//   4. the lanes assemble H_z in place, H and H_t into LDS;
//   5. LU with partial pivoting on the grid: the group holding column k
//      finds its pivot (DPP arg-max, first maximum in logical order), the
//      pivot and multipliers are broadcast, every lane updates its own
//      entries;
//   6. the triangular solves with a row per lane (lane i of each group
//      reads row i of the factors back from LDS); each step's value is
//      broadcast from the lane of the row at that logical position.
// Every product, sum and quotient is the one of the row-per-lane unit,
// in the same order, so both units give identical bits.
#pragma once

#include <type_traits>
#include <utility>

namespace phc {

// A term for the grid unit: coefficient; the LDS offsets e G + j of the
// factors zq_j^e of its nonzero exponents, ascending j, in bytes 0..7 of
// `off` (`nnz` of them; unused bytes are 0, the offset of pw(0, 0) = 1);
// `slot`: where its value goes in V.
struct WTerm {
  double re, im;
  uint64_t off;
  uint32_t nnz, slot;
};
static_assert(sizeof(WTerm) == 32, "staged copies move terms in 16-byte units");

// Table records through either kind of table pointer (staged in LDS, or
// global-typed): the records' copy constructors take generic references,
// so a global-typed record is read as plain 16-byte vectors.
typedef unsigned int phc_u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned long long phc_u64x2 __attribute__((ext_vector_type(2)));
PHC_D uint4 ld(const uint4* p, size_t i) { return p[i]; }
PHC_D uint4 ld(gptr<uint4> p, size_t i) {
  const phc_u32x4 v = ((gptr<phc_u32x4>)p)[i];
  return make_uint4(v.x, v.y, v.z, v.w);
}
PHC_D WTerm ld(const WTerm* p, size_t i) { return p[i]; }
PHC_D WTerm ld(gptr<WTerm> p, size_t i) {
  const gptr<phc_u64x2> q = (gptr<phc_u64x2>)p + 2 * i;
  const phc_u64x2 a = q[0], b = q[1];
  WTerm w;
  w.re = __longlong_as_double((long long)a.x);
  w.im = __longlong_as_double((long long)a.y);
  w.off = b.x;
  w.nnz = (uint32_t)b.y;
  w.slot = (uint32_t)(b.y >> 32);
  return w;
}

// Job records of lane (i, j): the terms of equation i with a nonzero
// exponent e of z_j, in term order. Two layouts (uint4):
//  short (every term has at most 4 nonzero exponents):
//    x = term | e << 20; y = offsets [pw(e-1, j), prefix 0..2];
//    z = offsets [suffix 0..2, -] with the suffix in right-fold order;
//    absent factors are 0 (pw(0, 0) = 1); rows past a lane's last job
//    point at the zero-coefficient term `nterms` with e = 1;
//  general: x = term | e << 20; past the last job the zero-coefficient
//    term (one slot per lane) or WJOB_NONE (16-lane grid);
//    (y, z) = offsets [pw(e-1, j), the other factors ascending];
//    w = prefix count | suffix count << 4.
// The 16-lane grid with the short layout deals whole sums to lanes instead
// (run layout, jac_runs): x, y, z as above, w = V index of the step's four
// term values | sum index << 16 | last step of the sum << 31.
constexpr uint32_t WJOB_NONE = 0xFFFFFFFFu;

// Request fields of the grid unit, packed in Req::gfold: largest number of
// nonzero exponents of a term, of a job's prefix and suffix factors, and
// the product length FL of the short layout (0: general layout).
PHC_D int gf_maxp(int g) { return (g >> 8) & 255; }
PHC_D int gf_maxs(int g) { return (g >> 16) & 255; }
PHC_D int gf_fl(int g) { return (g >> 24) & 15; }
// Run layout (16-lane grid, short layout): see jac_runs.
PHC_D bool gf_runs(int g) { return (g >> 28) & 1; }

// LDS of one path. V holds the term values of equation e at
// V[e * vstride ...], zero-padded to vstride, then four zeros that the
// Jacobian lanes add instead of values (they read the same four every
// step); it is cleared once per launch, and when a slot takes a new path.
template <int G>
struct GridShape {
  static constexpr int LP = (G * G < 64) ? G * G : 64;  // lanes per path
  static constexpr int R = LP / G;                       // column groups
  static constexpr int E = G / R;                        // columns per lane
};

// Job records in a request's table: [njobs][LP] in the run layout,
// [njobs][LP][E] otherwise.
template <int G>
PHC_D int job_records(const Req& R) {
  return R.njobs * GridShape<G>::LP * (gf_runs(R.gfold) ? 1 : GridShape<G>::E);
}

//
// The term values are dead once the job loop has read them, so with one
// slot per lane (G = 4, 8) the arrays that the LU and the solves fill
// afterwards, the factors A and the right-hand sides RB, live inside V when
// V is large enough. A unit that used V so clears that span before it
// returns, because V's padding must read as zero in the next unit's job
// loop. (The 16-lane grid keeps them apart: its A holds the run sums
// during the job loop, and its registers are at the two-wave limit.) Host
// and device size the block with these same functions.
PHC_HD constexpr int wave_slots(int G) {
  return G / ((G * G < 64 ? G * G : 64) / G);
}
// Entries of V that A and RB take over, 0 when they are separate.
PHC_HD constexpr int alias_entries(int G, int vmax) {
  return wave_slots(G) == 1 && vmax >= 2 * G * G ? 2 * G * G : 0;
}
// Complex entries of a path's unit block: T, V, A and RB (unless inside V),
// H and HT.
PHC_HD constexpr size_t wave_lds_entries(int G, int expmax, int vmax) {
  return (size_t)(expmax + 1) * G + (size_t)vmax + (size_t)2 * G * G -
         (size_t)alias_entries(G, vmax) + 2 * G;
}

template <int G>
struct WaveLds {
  cx* T;   // [(expmax + 1) G] power table, T[e G + j] = zq_j^e, T[j] = 1
  cx* V;   // [vmax]
  cx* A;   // [G G] factors, A[i G + j]
  cx* H;   // [G] H_i
  cx* HT;  // [G] H_t,i
  cx* RB;  // [G G] right-hand side by slot (lane q, column e: q E + e)
  PHC_D static size_t bytes(int expmax, int vmax) {
    // the 16-lane grid never aliases; its size keeps the expression with
    // which that kernel stays within two waves' registers
    if constexpr (GridShape<G>::E > 1)
      return ((size_t)(expmax + 1) * G + (size_t)vmax + (size_t)2 * G * G +
              2 * G) *
             sizeof(cx);
    else
      return wave_lds_entries(G, expmax, vmax) * sizeof(cx);
  }
  // T | V | A | H | HT | RB, or T | V | H | HT with A and RB inside V
  PHC_D void bind(char* base, int expmax, int vmax) {
    T = (cx*)base;
    V = T + (size_t)(expmax + 1) * G;
    if (alias_entries(G, vmax) > 0) {
      A = V;
      RB = V + G * G;
      H = V + vmax;
      HT = H + G;
    } else {
      A = V + vmax;
      H = A + (size_t)G * G;
      HT = H + G;
      RB = HT + G;
    }
  }
};

// Broadcast from lane `src` of the path's LP lanes. With one path per
// wavefront src is wavefront-uniform and the value goes through a scalar
// register; otherwise through ds_bpermute.
template <int LP>
PHC_D int pbcast_i(int v, int src) {
  if (LP == 64) return __builtin_amdgcn_readlane(v, src);
  return __shfl(v, src, LP);
}
template <int LP>
PHC_D double pbcast_d(double v, int src) {
  if (LP == 64) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
  return __shfl(v, src, LP);
}
template <int LP>
PHC_D cx pbcast_c(cx v, int src) {
  return mk(pbcast_d<LP>(v.re, src), pbcast_d<LP>(v.im, src));
}
// Lane-varying source within the path.
template <int LP>
PHC_D cx pshfl_c(cx v, int src) {
  return mk(__shfl(v.re, src, LP), __shfl(v.im, src, LP));
}
// A divisor scaled by lane `src` (cdiv_scale of its own entry); the fast
// LU forms its reciprocal once, after the broadcast.
template <int LP>
PHC_D cdiv_den pbcast_den(const cdiv_den& v, int src) {
  cdiv_den r;
  r.c = pbcast_d<LP>(v.c, src);
  r.d = pbcast_d<LP>(v.d, src);
  r.denom = pbcast_d<LP>(v.denom, src);
  r.ilogbw = pbcast_i<LP>(v.ilogbw, src);
  r.rden = 0.0;
  return r;
}

// Slots e of a lane that hold a column of an N-unknown system: columns
// r + R e < N for some group r.
template <int G>
PHC_D int slots_used(int N) {
  constexpr int R = GridShape<G>::R, E = GridShape<G>::E;
  return E == 1 ? 1 : (N + R - 1) / R;
}

// f(std::integral_constant<int, e>) for e = 0..E-1: the index of the
// lane's slot arrays stays a compile-time constant even where the body
// holds loops of its own (a plain loop around such a body is not always
// unrolled, and a runtime index sends the arrays to scratch memory).
template <typename F, int... I>
PHC_D void static_for_seq(F&& f, std::integer_sequence<int, I...>) {
  (f(std::integral_constant<int, I>{}), ...);
}
template <int E, typename F>
PHC_D void static_for(F&& f) {
  static_for_seq(f, std::make_integer_sequence<int, E>{});
}

PHC_D int woff(uint64_t off, int a) { return (int)((off >> (8 * a)) & 255u); }
PHC_D int byte_of(uint32_t w, int b) { return (int)((w >> (8 * b)) & 255u); }

// Term values c_k pre(m) (src/sparse.f90:260-264, :270), round robin over
// the path's lanes, products of FV factors padded with pw(0, 0) = 1; two
// terms per step so that their loads and products overlap.
template <int FV>
PHC_D cx term_value(const WTerm& w, const cx* __restrict__ T) {
  cx f[FV];
#pragma unroll
  for (int a = 0; a < FV; ++a) f[a] = T[woff(w.off, a)];
  cx P = f[0];
#pragma unroll
  for (int a = 1; a < FV; ++a) P = cmul(P, f[a]);
  return cmul(mk(w.re, w.im), P);
}
template <int LP, int FV, class WP>
PHC_D void term_values(const Req& R, WP wt,
                       const cx* __restrict__ T, cx* __restrict__ V, int q) {
  const int nt = R.nterms;
  int k = q;
  if (k + LP < nt) {
    // the next turn's two terms are loaded while this turn's compute
    WTerm w0 = ld(wt, k), w1 = ld(wt, k + LP);
    for (;;) {
      const int kn = k + 2 * LP;
      const bool more = kn + LP < nt;
      WTerm n0 = w0, n1 = w1;
      if (more) {
        n0 = ld(wt, kn);
        n1 = ld(wt, kn + LP);
      }
      const cx v0 = term_value<FV>(w0, T);
      const cx v1 = term_value<FV>(w1, T);
      V[w0.slot] = v0;
      V[w1.slot] = v1;
      k = kn;
      if (!more) break;
      w0 = n0;
      w1 = n1;
    }
  }
  if (k < nt) {
    const WTerm w0 = ld(wt, k);
    V[w0.slot] = term_value<FV>(w0, T);
  }
}

// Term values of the one-slot general layout (jac_general_ps), which
// also keep the term's prefix products P_a = f_1 ... f_a (a < m, left fold)
// and suffix products S_b = f_b (f_{b+1} ( ... f_m)) (2 <= b <= m, right
// fold) at V[ps + a - 1] and V[ps + m - 1 + b - 2], ps = WTerm::nnz >> 8
// (src/sparse.f90:259-267). Factors past the m-th are pw(0, 0) = 1, as in
// term_value. Lane 0 of the path writes the one that empty products name.
template <int LP, class WP>
PHC_D void term_values_ps(const Req& R, WP wt, const cx* __restrict__ T,
                          cx* __restrict__ V, int q, int one_slot) {
  if (q == 0) V[one_slot] = mk(1.0, 0.0);
  for (int k = q; k < R.nterms; k += LP) {
    const WTerm w = ld(wt, k);
    const int m = (int)(w.nnz & 255u), ps = (int)(w.nnz >> 8);
    cx f[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) f[a] = T[woff(w.off, a)];
    cx P = f[0];
#pragma unroll
    for (int a = 1; a < 8; ++a) {
      if (a < m) V[ps + a - 1] = P;  // P_a
      P = cmul(P, f[a]);
    }
    V[w.slot] = cmul(mk(w.re, w.im), P);
    cx S = f[7];  // S_8
    if (m == 8) V[ps + 7 + 6] = S;
#pragma unroll
    for (int b = 7; b >= 2; --b) {
      S = cmul(f[b - 1], S);  // S_b
      if (b <= m) V[ps + m - 1 + b - 2] = S;
    }
  }
}

// The coefficient of a job record's term.
template <class WP>
PHC_D cx job_coef(const uint4 job, WP wt) {
  const auto& w = wt[job.x & 0xFFFFFu];
  return mk(w.re, w.im);
}

// One derivative c e pw(e-1, j) pre(j-1) suf(j+1) (src/sparse.f90:273) of
// a short-layout job record with its term's coefficient c, FL factors per
// prefix and suffix.
template <int FL>
PHC_D cx job_product(const uint4 job, const cx c, const cx* __restrict__ T) {
  const cx B = T[byte_of(job.y, 0)];
  cx Pf[FL], Sf[FL];
#pragma unroll
  for (int b = 0; b < FL; ++b) {
    Pf[b] = T[byte_of(job.y, 1 + b)];
    Sf[b] = T[byte_of(job.z, b)];
  }
  cx d = cmul(cscale((double)(job.x >> 20), c), B);
  cx P = Pf[0], S = Sf[0];
#pragma unroll
  for (int b = 1; b < FL; ++b) {
    P = cmul(P, Pf[b]);
    S = cmul(Sf[b], S);
  }
  return cmul(cmul(d, P), S);
}

// Jacobian entry of the lane's one slot (src/sparse.f90:271-276) or, in
// the value slots, F_e = sum of its equation's values in term order, four
// per job. Short job layout, no branches; two jobs per turn (the job rows
// are padded to an even count), the next turn's records loaded while this
// one computes. Job table [m][LP]. Grids with several slots per lane use
// the run layout (jac_runs) instead.
template <int G, int FL, class WP, class JP>
PHC_D void jac_short(const Req& R, WP wt,
                     JP jobs, const cx* __restrict__ T,
                     const cx* __restrict__ V, int q, int vbase, int vstep,
                     cx& acc) {
  static_assert(GridShape<G>::E == 1, "one slot per lane");
  constexpr int LP = GridShape<G>::LP, U = 2;
  const int M = R.njobs;
  acc = mk(0.0, 0.0);
  uint4 jn[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (u < M) jn[u] = ld(jobs, (size_t)u * LP + q);
  for (int m = 0; m < M; m += U) {
    uint4 jc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jc[u] = jn[u];
    if (m + U < M) {
#pragma unroll
      for (int u = 0; u < U; ++u) jn[u] = ld(jobs, (size_t)(m + U + u) * LP + q);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const cx d = job_product<FL>(jc[u], job_coef(jc[u], wt), T);
      const int vb = vbase + vstep * (m + u);
      const cx v0 = V[vb], v1 = V[vb + 1], v2 = V[vb + 2], v3 = V[vb + 3];
      acc = cadd(acc, d);
      acc = cadd(acc, v0);
      acc = cadd(acc, v1);
      acc = cadd(acc, v2);
      acc = cadd(acc, v3);
    }
  }
}

// Jacobian entries and values F_i of the 16-lane grid in the run layout:
// each sum is a run of consecutive job steps on one lane, and the runs are
// dealt over all 64 lanes, so no lane adds the jobs of several slots in
// lock step while others idle. A lane adds its steps in order into one
// running sum -- per step the job's derivative, then four term values, as
// jac_short does -- and at a run's last step (record w bit 31) stores the
// sum to S[w bits 16..30] and starts the next run from +0. Table
// [m][LP], two steps per turn with the next turn's records prefetched.
template <int FL, class WP, class JP>
PHC_D void jac_runs(const Req& R, WP wt,
                    JP jobs, const cx* __restrict__ T,
                    const cx* __restrict__ V, int q, cx* __restrict__ S) {
  constexpr int LP = 64, U = 2;
  const int M = R.njobs;
  cx cur = mk(0.0, 0.0);
  uint4 jn[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (u < M) jn[u] = ld(jobs, (size_t)u * LP + q);
  for (int m = 0; m < M; m += U) {
    uint4 jc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jc[u] = jn[u];
    if (m + U < M) {
#pragma unroll
      for (int u = 0; u < U; ++u) jn[u] = ld(jobs, (size_t)(m + U + u) * LP + q);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint4 job = jc[u];
      const cx d = job_product<FL>(job, job_coef(job, wt), T);
      const int vb = (int)(job.w & 0xFFFFu);
      const cx v0 = V[vb], v1 = V[vb + 1], v2 = V[vb + 2], v3 = V[vb + 3];
      cur = cadd(cur, d);
      cur = cadd(cur, v0);
      cur = cadd(cur, v1);
      cur = cadd(cur, v2);
      cur = cadd(cur, v3);
      if (job.w >> 31) {
        S[(job.w >> 16) & 0x7FFFu] = cur;
        cur = mk(0.0, 0.0);
      }
    }
  }
}

// One derivative of a general-layout job record, prefix and suffix
// products padded with pw(0, 0) = 1 to FLG factors each (the short layout
// pads the same way); FLG = max(maxp, maxs) is a template argument, so the
// factor loops carry no runtime bound.
template <int FLG>
PHC_D cx job_product_general(const uint4 job, const cx c,
                             const cx* __restrict__ T) {
  const int ex = (int)((job.x >> 20) & 255u);
  const uint64_t off = ((uint64_t)job.z << 32) | job.y;
  const int np = (int)(job.w & 15u), ns = (int)(job.w >> 4);
  const cx B = T[woff(off, 0)];
  cx Pf[FLG], Sf[FLG];
#pragma unroll
  for (int b = 0; b < FLG; ++b) {
    Pf[b] = T[b < np ? woff(off, 1 + b) : 0];
    Sf[b] = T[b < ns ? woff(off, np + ns - b) : 0];
  }
  cx d = cmul(cscale((double)ex, c), B);
  cx P = Pf[0], S = Sf[0];
#pragma unroll
  for (int b = 1; b < FLG; ++b) {
    P = cmul(P, Pf[b]);
    S = cmul(Sf[b], S);
  }
  return cmul(cmul(d, P), S);
}

// General job layout with one slot per lane (G = 4, 8; terms with more
// than four nonzero exponents). The term pass (term_values_ps) keeps each
// term's prefix products P_1..P_{m-1} and suffix products S_2..S_m in V,
// so the job of the term's q-th factor, c e pw(e-1, j) pre(j-1) suf(j+1)
// (src/sparse.f90:271-273), is ((c e) pw(e-1, j)) P_{q-1} S_{q+1}: three
// products, in the reference's association (an empty product is the one
// in V). The loop of jac_short around it: two jobs per turn, the next
// turn's records prefetched, four term values per step for the F lanes.
// Rows past a lane's last job are the zero-coefficient term, adding +0.
template <int G, class WP, class JP>
PHC_D void jac_general_ps(const Req& R, WP wt, JP jobs,
                          const cx* __restrict__ T, const cx* __restrict__ V,
                          int q, int vbase, int vstep, cx& acc) {
  constexpr int LP = GridShape<G>::LP, U = 2;
  const int M = R.njobs;
  acc = mk(0.0, 0.0);
  uint4 jn[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (u < M) jn[u] = ld(jobs, (size_t)u * LP + q);
  for (int m = 0; m < M; m += U) {
    uint4 jc[U];
#pragma unroll
    for (int u = 0; u < U; ++u) jc[u] = jn[u];
    if (m + U < M) {
#pragma unroll
      for (int u = 0; u < U; ++u) jn[u] = ld(jobs, (size_t)(m + U + u) * LP + q);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint4 j = jc[u];
      const cx c = job_coef(j, wt);
      const cx B = T[j.y];
      const cx pre = V[j.z & 0xFFFFu], suf = V[j.z >> 16];
      const int ex = (int)((j.x >> 20) & 255u);
      const cx d = cmul(cmul(cmul(cscale((double)ex, c), B), pre), suf);
      const int vb = vbase + vstep * (m + u);
      const cx v0 = V[vb], v1 = V[vb + 1], v2 = V[vb + 2], v3 = V[vb + 3];
      acc = cadd(acc, d);
      acc = cadd(acc, v0);
      acc = cadd(acc, v1);
      acc = cadd(acc, v2);
      acc = cadd(acc, v3);
    }
  }
}

// The same with the general job layout (any number of factors up to 8).
template <int G, class WP, class JP>
PHC_D void jac_general(const Req& R, WP wt,
                       JP jobs, const cx* __restrict__ T,
                       const cx* __restrict__ V, int q,
                       const int (&vbase)[GridShape<G>::E],
                       const int (&vcount)[GridShape<G>::E],
                       cx (&acc)[GridShape<G>::E]) {
  constexpr int LP = GridShape<G>::LP, E = GridShape<G>::E;
  constexpr int FM = 8;
  const int maxp = gf_maxp(R.gfold), maxs = gf_maxs(R.gfold);
  const int M = R.njobs;
  const int eu = slots_used<G>(R.N);
  static_for<E>([&](auto ec) {
    constexpr int e = decltype(ec)::value;
    acc[e] = mk(0.0, 0.0);
    if (e >= eu) return;
    int fk = 0;
    for (int m = 0; m < M; ++m) {
      const uint4 job = ld(jobs, ((size_t)m * LP + q) * E + e);
      if (job.x != WJOB_NONE) {
        const int ex = (int)((job.x >> 20) & 255u);
        const uint64_t off = ((uint64_t)job.z << 32) | job.y;
        const int np = (int)(job.w & 15u), ns = (int)(job.w >> 4);
        const auto& w = wt[job.x & 0xFFFFFu];
        const cx c = mk(w.re, w.im);
        const cx B = T[woff(off, 0)];
        cx Pf[FM], Sf[FM];
#pragma unroll
        for (int b = 0; b < FM; ++b) {
          if (b < maxp) Pf[b] = T[b < np ? woff(off, 1 + b) : 0];
          if (b < maxs) Sf[b] = T[b < ns ? woff(off, np + ns - b) : 0];
        }
        cx d = cmul(cscale((double)ex, c), B);
        if (maxp > 0) {
          cx P = Pf[0];
#pragma unroll
          for (int b = 1; b < FM; ++b)
            if (b < maxp) P = cmul(P, Pf[b]);
          d = cmul(d, P);
        }
        if (maxs > 0) {
          cx S = Sf[0];
#pragma unroll
          for (int b = 1; b < FM; ++b)
            if (b < maxs) S = cmul(Sf[b], S);
          d = cmul(d, S);
        }
        acc[e] = cadd(acc[e], d);
      }
      for (int r = 0; r < 4 && fk < vcount[e]; ++r, ++fk)
        acc[e] = cadd(acc[e], V[vbase[e] + fk]);
    }
    for (; fk < vcount[e]; ++fk) acc[e] = cadd(acc[e], V[vbase[e] + fk]);
  });
}

// The |a|^2 screen of the fast LUs decides only for a largest |a|^2 in
// [2^-930, 2^930), well inside the normal range: a test of its exponent
// field (a negative value or NaN fails it).
PHC_D bool screen_range(double sv) {
  const unsigned int hi = (unsigned int)((unsigned long long)__double_as_longlong(sv) >> 32);
  return (hi >> 20) - (1023u - 930u) < 1860u;
}

// Pivot of one LU step among the candidates `cand` of the column's group
// (src/linalg.f90:52-58: largest abs, first maximum in logical order),
// broadcast from lane `src`. |a|^2 screens first: when exactly one
// candidate has |a|^2 within 1e-12 of the largest, its hypot is strictly
// the largest too (|a|^2 and hypot are both within a few units in the last
// place), so the hypot values are not needed. Near ties, NaN entries and
// |a|^2 outside the normal range fall back to comparing the hypot values
// themselves, with the reference's rule for NaN moduli (pivot_argmax).
template <int G, int LP>
PHC_D void pivot_search(bool cand, cx ak, int pos, int k, int lane, int src,
                        double& bv, int& bp, int& bl) {
  const double sq = ak.re * ak.re + ak.im * ak.im;
  double sv = cand ? sq : -1.0;
  int sp = pos, sl = lane;
  gargmax<G>(sv, sp, sl);
  sv = pbcast_d<LP>(sv, src);
  // a NaN candidate counts as a near tie: the hypot scan decides
  unsigned long long near =
      __ballot(cand && (sq >= sv * (1.0 - 1e-12) || sq != sq));
  if (LP < 64) near = (near >> ((__lane_id() / LP) * LP)) & ((1ull << LP) - 1ull);
  if (sv > 1e-280 && sv < 1e280 && __popcll(near) == 1) {
    bv = sv;
    bp = pbcast_i<LP>(sp, src);
    bl = pbcast_i<LP>(sl, src);
    return;
  }
  pivot_argmax<G>(cand, cabs_(ak), pos, k, lane, bv, bp, bl);
  bv = pbcast_d<LP>(bv, src);
  bp = pbcast_i<LP>(bp, src);
  bl = pbcast_i<LP>(bl, src);
}

// LU steps of a grid with several slots per lane (src/linalg.f90:49-76).
// Step k works on slot ke = k / RG of group kr = k % RG; the steps of one
// slot share a compile-time ke, so the slot arrays stay in registers.
// EU = ceil(N / RG) slots hold columns of the system; it is a template
// argument so that the slot loops carry no runtime bound. Returns 0 on a
// zero pivot.
template <int G, int EU>
PHC_D int lu_slots(int N, cx (&a)[GridShape<G>::E],
                   const bool (&in)[GridShape<G>::E],
                   const int (&col)[GridShape<G>::E], int rep, int lane,
                   int& pos, uint64_t& plb) {
  constexpr int LP = GridShape<G>::LP, RG = GridShape<G>::R,
                E = GridShape<G>::E;
  int ok = 1;
  static_for<E>([&](auto ec) {
    constexpr int ke = decltype(ec)::value;
    if constexpr (ke < EU) {
#pragma unroll
      for (int kr = 0; kr < RG; ++kr) {
        const int k = ke * RG + kr;
        if (!ok || k >= N) break;
        const cx aik = pshfl_c<LP>(a[ke], kr * G + lane);
        // every lane prepares its own entry as a divisor while the pivot is
        // searched; the winner's is broadcast (the pivot's scaling)
        const cdiv_den mine = cdiv_scale(a[ke]);
        double bv;
        int bp, bl;
        pivot_search<G, LP>(in[ke] && rep == kr && pos >= k, a[ke], pos, k,
                            lane, kr * G, bv, bp, bl);
        if (bv == 0.0) {
          ok = 0;
          break;
        }
        plb |= (uint64_t)bl << (4 * k);
        const cdiv_den pd = pbcast_den<LP>(mine, kr * G + bl);
        // slots left of ke hold finished columns (col[e] < k)
        cx pr[E];
#pragma unroll
        for (int e = ke; e < EU; ++e) pr[e] = pshfl_c<LP>(a[e], rep * G + bl);
        int np = pos;
        if (pos == k) np = bp;
        if (lane == bl) np = k;
        pos = np;
        const cx li = cdiv_apply(aik, pd);
        if (in[ke] && rep == kr && pos > k) a[ke] = li;
#pragma unroll
        for (int e = ke; e < EU; ++e) {
          const cx upd = csub(a[e], cmul(li, pr[e]));
          if (in[e] && col[e] > k && pos > k) a[e] = upd;
        }
      }
    }
  });
  return ok;
}

// LU steps of a grid with one column per lane (E = 1: G = 4 and G = 8;
// src/linalg.f90:49-76). Lane (rep, lane) holds entry (lane, rep) of the
// matrix in `a`; column k is in group k, so group k searches the pivot.
// Returns 0 on a zero pivot. This is the exact rule in every case (hypot
// values compared when the |a|^2 screen is inconclusive); lu_one_slot runs
// it only when its own screen fails.
template <int G>
PHC_D int lu_one_slot_exact(int N, cx& a, const bool in, const int col,
                            int rep, int lane, int& pos, uint64_t& plb) {
  constexpr int LP = GridShape<G>::LP;
  static_assert(GridShape<G>::E == 1, "one column per lane");
  int ok = 1;
#pragma unroll
  for (int k = 0; k < G; ++k) {
    if (k >= N) break;
    // column k of this lane's row, fetched while the pivot is searched:
    // every group forms its row's multiplier itself
    const cx aik = pshfl_c<LP>(a, k * G + lane);
    const cdiv_den mine = cdiv_scale(a);  // the winner's is the pivot's
    double bv;
    int bp, bl;
    pivot_search<G, LP>(in && rep == k && pos >= k, a, pos, k, lane, k * G,
                        bv, bp, bl);
    if (bv == 0.0) {
      ok = 0;
      break;
    }
    plb |= (uint64_t)bl << (4 * k);
    const cdiv_den pd = pbcast_den<LP>(mine, k * G + bl);
    const cx pr = pshfl_c<LP>(a, rep * G + bl);
    int np = pos;
    if (pos == k) np = bp;
    if (lane == bl) np = k;
    pos = np;
    const cx li = cdiv_apply(aik, pd);
    if (in && rep == k && pos > k) a = li;
    const cx upd = csub(a, cmul(li, pr));
    if (in && col > k && pos > k) a = upd;
  }
  return ok;
}

// The same LU, fast. Each pivot search is the |a|^2 screen alone: the
// largest |a|^2 of the column's candidates (a value-only DPP max), the
// winner found from a ballot of the candidates that reach it. The screen
// decides a step when exactly one candidate lies within 1e-12 of that
// maximum, no candidate is NaN and the maximum is in [1e-280, 1e280]: its
// hypot is then strictly the largest too, so the step is the exact rule's.
// The steps run without branches and only record whether every screen was
// conclusive: returns 1, or 2 when one was not (near ties, NaN, zero
// pivots, extreme ranges) -- the caller then redoes the unit with
// lu_one_slot_exact. The entries travel through the path's G x G tile `T`
// (LDS, free during the LU): T[rep G + lane] = entry (lane, rep).
template <int G>
PHC_D int lu_one_slot(int N, cx& a, const bool in, const int col, int rep,
                      int lane, int& pos, uint64_t& plb, cx* __restrict__ T) {
  constexpr int LP = GridShape<G>::LP;
  static_assert(GridShape<G>::E == 1, "one column per lane");
  // the path's lanes in a ballot (several paths share a wavefront if G = 4)
  const int shift = LP == 64 ? 0 : (int)(__lane_id() / LP) * LP;
  const unsigned long long pmask = LP == 64 ? ~0ull : ((1ull << LP) - 1ull);
  bool sure = true;
  bool dok = true;  // every quotient within cdiv_apply_fast's range
#pragma unroll
  for (int k = 0; k < G; ++k) {
    if (k >= N) break;
    wave_lds_fence();
    T[rep * G + lane] = a;
    wave_lds_fence();
    const cx aik = T[k * G + lane];
    const int gbase = k * G;
    const bool cand = in && rep == k && pos >= k;
    const cdiv_den mine = cdiv_scale(a);  // the winner's is the pivot's
    const double sq = a.re * a.re + a.im * a.im;
    const double sv = pbcast_d<LP>(gmax<G>(cand ? sq : -1.0), gbase);
    const unsigned long long hit =
        (__builtin_amdgcn_ballot_w64(cand && sq == sv) >> shift) & pmask;
    const unsigned long long near =
        (__builtin_amdgcn_ballot_w64(cand &&
                                     (sq >= sv * (1.0 - 1e-12) || sq != sq)) >>
         shift) &
        pmask;
    sure = sure && screen_range(sv) && __popcll(near) == 1;
    // the winner's lane in the path (any lane of the group if none: the
    // step is then redone exactly)
    const int wl = hit ? __ffsll(hit) - 1 : gbase;
    const int bl = wl - gbase;
    const int bp = pbcast_i<LP>(pos, wl);
    plb |= (uint64_t)(bl & 15) << (4 * k);
    cdiv_den pd = pbcast_den<LP>(mine, wl);
    cdiv_recip(pd);
    const cx pr = T[rep * G + bl];
    int np = pos;
    if (pos == k) np = bp;
    if (lane == bl) np = k;
    pos = np;
    const cx li = cdiv_apply_fast(aik, pd, dok);
    if (in && rep == k && pos > k) a = li;
    const cx upd = csub(a, cmul(li, pr));
    if (in && col > k && pos > k) a = upd;
  }
  // a quotient outside qdiv_fast's range: redone exactly too
  sure = sure && ((__ballot(!dok) >> shift) & pmask) == 0;
  return sure ? 1 : 2;
}

// lu_slots with the screen of lu_one_slot: the steps run branch-free on
// the |a|^2 screen, the entries through the path's G x G tile `T`
// (T[col G + row], LDS, free during the LU). Returns 1, or 2 when a
// step's screen was inconclusive (the caller then redoes the unit with
// lu_slots).
template <int G, int EU>
PHC_D int lu_slots_fast(int N, cx (&a)[GridShape<G>::E],
                        const bool (&in)[GridShape<G>::E],
                        const int (&col)[GridShape<G>::E], int rep, int lane,
                        int& pos, uint64_t& plb, cx* __restrict__ T) {
  constexpr int LP = GridShape<G>::LP, RG = GridShape<G>::R,
                E = GridShape<G>::E;
  static_assert(LP == 64, "one path per wavefront");
  bool sure = true;
  bool dok = true;  // every quotient within cdiv_apply_fast's range
  static_for<E>([&](auto ec) {
    constexpr int ke = decltype(ec)::value;
    if constexpr (ke < EU) {
#pragma unroll
      for (int kr = 0; kr < RG; ++kr) {
        const int k = ke * RG + kr;
        if (k >= N) break;
        wave_lds_fence();
#pragma unroll
        for (int e = ke; e < EU; ++e) T[col[e] * G + lane] = a[e];
        wave_lds_fence();
        const cx aik = T[k * G + lane];
        const int gbase = kr * G;
        const bool cand = in[ke] && rep == kr && pos >= k;
        const cdiv_den mine = cdiv_scale(a[ke]);  // the winner's: the pivot's
        const double sq = a[ke].re * a[ke].re + a[ke].im * a[ke].im;
        const double sv = pbcast_d<LP>(gmax<G>(cand ? sq : -1.0), gbase);
        const unsigned long long hit =
            __builtin_amdgcn_ballot_w64(cand && sq == sv);
        const unsigned long long near = __builtin_amdgcn_ballot_w64(
            cand && (sq >= sv * (1.0 - 1e-12) || sq != sq));
        sure = sure && screen_range(sv) && __popcll(near) == 1;
        const int wl = hit ? __ffsll(hit) - 1 : gbase;
        const int bl = wl - gbase;
        const int bp = pbcast_i<LP>(pos, wl);
        plb |= (uint64_t)(bl & 15) << (4 * k);
        cdiv_den pd = pbcast_den<LP>(mine, wl);
        cdiv_recip(pd);
        int np = pos;
        if (pos == k) np = bp;
        if (lane == bl) np = k;
        pos = np;
        const cx li = cdiv_apply_fast(aik, pd, dok);
        if (in[ke] && rep == kr && pos > k) a[ke] = li;
#pragma unroll
        for (int e = ke; e < EU; ++e) {
          const cx upd = csub(a[e], cmul(li, T[col[e] * G + bl]));
          if (in[e] && col[e] > k && pos > k) a[e] = upd;
        }
      }
    }
  });
  // a quotient outside qdiv_fast's range: redone exactly too
  sure = sure && __ballot(!dok) == 0;
  return sure ? 1 : 2;
}

// Forward and back substitution (src/linalg.f90:93-111) for NC rows known
// at compile time, row per lane: lane i of every group holds row i of the
// factors and runs the recurrence of that row; the value of the row at
// logical position k is broadcast from its lane (lane PL[k] of the group,
// PL packed in plb). Sums in the reference's order, as in the
// row-per-lane unit. With NC a constant the steps carry no row-count
// tests (such tests inside the unrolled steps cost as much as the steps).
//
// FAST: the quotients through cdiv_apply_fast (pdiag carries its rden);
// `dok` is cleared when the quotient of this lane's own row (the one its
// step broadcasts) fell outside that range, and the unit is then redone.
template <int G, int NC, bool FAST>
PHC_D cx solve_rows(const cx (&arow)[G], const cx bown, const cdiv_den pdiag,
                    int pos, bool row, int lane, uint64_t plb, bool& dok) {
  constexpr int LP = GridShape<G>::LP;
  auto rowcast = [&](cx v, int r) -> cx {
    if (LP == 64) return pbcast_c<LP>(v, r);  // group 0, lane r
    return gshfl_c<G>(v, r);
  };
  // Fortran's SUM starts from +0: row i's sum is 0 + p_1 + ... (which is
  // p_1 unless p_1 has a -0 part); row 0 keeps its right-hand side
  cx sacc = mk(0.0, 0.0);
  static_for<NC>([&](auto kc) {
    constexpr int k = decltype(kc)::value;
    const cx cand = pos == 0 ? bown : csub(bown, sacc);
    const cx yk = rowcast(cand, (int)((plb >> (4 * k)) & 15u));
    if (row && pos > k) sacc = cadd(sacc, cmul(arow[k], yk));
  });
  const cx yfin = pos == 0 ? bown : csub(bown, sacc);
  // back substitution: the row at position i needs
  // sum_{j > i} a(i, j) x_j in ascending j, from +0; each product
  // a(i, j) x_j is formed once, when x_j is known (px[j]), and step i adds
  // them in order
  cx px[NC];
  cx xmine = mk(0.0, 0.0);
  static_for<NC>([&](auto ic) {
    constexpr int i = NC - 1 - decltype(ic)::value;
    cx num = yfin;
    if constexpr (i + 1 < NC) {
      cx sm = cadd(mk(0.0, 0.0), px[i + 1]);
      static_for<NC>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        if constexpr (j >= i + 2) sm = cadd(sm, px[j]);
      });
      num = csub(yfin, sm);
    }
    cx cand;
    if constexpr (FAST) {
      bool qok = true;
      cand = cdiv_apply_fast(num, pdiag, qok);
      dok = dok & (qok | (pos != i));
    } else {
      cand = cdiv_apply(num, pdiag);
    }
    const cx xi = rowcast(cand, (int)((plb >> (4 * i)) & 15u));
    if (lane == i) xmine = xi;
    if constexpr (i > 0) px[i] = cmul(arow[i], xi);
  });
  return xmine;
}

// solve_rows for the path's N among NC .. G (a lanes class of G serves
// N from G / 2 + 1 to G; the 4-lane class from 2).
template <int G, int NC, bool FAST>
PHC_D cx solve_dispatch(int N, const cx (&arow)[G], const cx bown,
                        const cdiv_den pdiag, int pos, bool row, int lane,
                        uint64_t plb, bool& dok) {
  if constexpr (NC < G) {
    if (N != NC)
      return solve_dispatch<G, NC + 1, FAST>(N, arow, bown, pdiag, pos, row,
                                             lane, plb, dok);
  }
  return solve_rows<G, NC, FAST>(arow, bown, pdiag, pos, row, lane, plb, dok);
}

// Factors that are exactly one: the reference multiplies by pw(0, j) = 1
// for every zero exponent (src/sparse.f90:262, :266); this unit multiplies
// by 1 only to pad a lane's product to the length shared by the
// wavefront, and skips it elsewhere. Either way the result is the same up
// to the sign of a zero component. Zero padding of sums is exact: the
// sums start at +0 and so are never -0.
template <int G, bool EXACT, class WP, class JP>
PHC_D void unit_wave(const Req& R, WP wt,
                     JP jobs, const int lane,
                     const int rep, const int mode, const cx zq, const cx t,
                     const cx dt, const WaveLds<G>& W, cx& xout, int& okout,
                     double& be) {
  constexpr int LP = GridShape<G>::LP, RG = GridShape<G>::R,
                E = GridShape<G>::E;
  const int N = R.N, n = R.n;
  const int q = rep * G + lane;
  cx* __restrict__ T = W.T;
  // restrict except where A and RB may live inside V (one slot per lane)
  using VPtr = std::conditional_t<(E > 1), cx* __restrict__, cx*>;
  VPtr V = W.V;
  // the span of V that A and RB used, cleared before the unit returns (its
  // padding must read as zero in the next unit's job loop)
  auto release_v = [&]() {
    if (E == 1 && W.A == W.V) {  // A and RB live in V's first 2 G^2 entries
      wave_lds_fence();
#pragma unroll
      for (int i = q; i < 2 * G * G; i += LP) V[i] = mk(0.0, 0.0);
    }
  };
  const int fl_len = gf_fl(R.gfold);
  int col[E];
#pragma unroll
  for (int e = 0; e < E; ++e) col[e] = rep + RG * e;

  // 1. power table (src/sparse.f90:251-255); row 0 holds pw(0, j) = 1
  wave_lds_fence();
  if (rep == 0) {
    T[lane] = mk(1.0, 0.0);
    if (lane < N) {
      cx p = zq;
      T[G + lane] = p;
      for (int e = 2; e <= R.maxexp; ++e) {
        p = cmul(p, zq);
        T[e * G + lane] = p;
      }
    }
  }
  wave_lds_fence();

  // 2. term values
  if (fl_len != 0)
    term_values<LP, 4>(R, wt, T, V, q);
  else if constexpr (E == 1)
    term_values_ps<LP>(R, wt, T, V, q, R.n * R.vstride + 4);
  else
    term_values<LP, 8>(R, wt, T, V, q);
  wave_lds_fence();

  // 3. Jacobian entries in slots (i < n, j < N), F_e in slot (G-1, e)
  const int vs = R.vstride;  // V entries per equation
  bool jl[E], fl[E];
  int vbase[E], vcount[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    jl[e] = lane < n && col[e] < N;
    fl[e] = lane == G - 1 && col[e] < n;
    vbase[e] = (fl[e] ? col[e] : n) * vs;
    vcount[e] = fl[e] ? R.first[col[e] + 1] - R.first[col[e]] : 0;
  }
  cx acc[E];
  if constexpr (E > 1) {
    if (gf_runs(R.gfold)) {
      // sums into S = W.A (free until the factors are stored), then each
      // slot takes its own: entry (i, j) at i N + j, F_i at n N + i
      if (fl_len == 1)
        jac_runs<1>(R, wt, jobs, T, V, q, W.A);
      else if (fl_len == 2)
        jac_runs<2>(R, wt, jobs, T, V, q, W.A);
      else
        jac_runs<3>(R, wt, jobs, T, V, q, W.A);
      wave_lds_fence();
#pragma unroll
      for (int e = 0; e < E; ++e)
        acc[e] = jl[e]   ? W.A[lane * N + col[e]]
                 : fl[e] ? W.A[n * N + col[e]]
                         : mk(0.0, 0.0);
    } else {
      jac_general<G>(R, wt, jobs, T, V, q, vbase, vcount, acc);
    }
  } else {
    // F lanes walk their equation's values four per step; the others add
    // the four zeros after the last equation, every step
    const int vstep = fl[0] ? 4 : 0;
    if (fl_len == 1)
      jac_short<G, 1>(R, wt, jobs, T, V, q, vbase[0], vstep, acc[0]);
    else if (fl_len == 2)
      jac_short<G, 2>(R, wt, jobs, T, V, q, vbase[0], vstep, acc[0]);
    else if (fl_len == 3)
      jac_short<G, 3>(R, wt, jobs, T, V, q, vbase[0], vstep, acc[0]);
    else
      jac_general_ps<G>(R, wt, jobs, T, V, q, vbase[0], vstep, acc[0]);
  }

  // 4. homotopy (src/homotopy.f90:491-511) and the patch row; the slot
  // that computes H_i and H_t,i keeps the right-hand side of row i in rhs:
  // tangent -(H_t dt/ds), Newton -H (src/tracker.f90:314, :348)
  cx a[E], rhs[E];
  bool in[E];
  {
    const cx gt = cmul(R.gamma, t);
    const cx omt = mk(1.0 - t.re, 0.0 - t.im);
    const cx w = T[G + n];
    // start-system terms of row `lane` (every Jacobian slot of a lane is in
    // that row), formed once instead of per slot
    const int il = lane < n ? lane : 0;
    const int dl = R.deg[il];
    const cx zl = T[G + il];
    cx zdl, wdl;
    zpowi_pair(zl, w, dl - 1, R.dbits, zdl, wdl);
    const cx bl = R.b[il];
    const cx diag_add = cmul(cscale((double)dl, gt), zdl);
    const cx wcol_sub = cmul(cscale((double)dl, cmul(gt, bl)), wdl);
    // each equation's start-system value g_i = z_i^d - b_i w^d is formed
    // once, by row lane i of group 0, and passed to the F slot through HT
    // (which that slot overwrites after reading)
    if (rep == 0 && lane < n)
      W.HT[lane] = csub(cmul(zdl, zl), cmul(cmul(bl, wdl), w));
    // The patch row's sum(a*z) in index order; every lane forms it and lane
    // n keeps it. Every load and product is formed, entries past N
    // included (the patch is zero there; the table row may hold anything),
    // and the sum keeps the first N by selects, so no branch holds the
    // loads back. Unrolled by four: a full unroll of the 16-lane grid's
    // sum would hold all its loads in registers at once.
    cx psum = mk(0.0, 0.0);
#pragma unroll 4
    for (int jj = 0; jj < G; ++jj) {
      const cx pj = cmul(R.patch[jj], T[G + jj]);
      const cx sj = cadd(psum, pj);
      psum.re = jj < N ? sj.re : psum.re;
      psum.im = jj < N ? sj.im : psum.im;
    }
    wave_lds_fence();
    static_for<E>([&](auto ec) {
      constexpr int e = decltype(ec)::value;
      const int j = col[e];
      a[e] = mk(0.0, 0.0);
      rhs[e] = mk(0.0, 0.0);
      if (jl[e]) {
        a[e] = cmul(omt, acc[e]);
        if (j == lane) a[e] = cadd(a[e], diag_add);
        if (j == n) a[e] = csub(a[e], wcol_sub);
      } else if (fl[e]) {
        const int i = j;
        const cx g = W.HT[i];
        const cx h = cadd(cmul(gt, g), cmul(omt, acc[e]));
        const cx ht = csub(cmul(R.gamma, g), acc[e]);
        W.H[i] = h;
        W.HT[i] = ht;
        rhs[e] = (mode == PC_TAN) ? cneg(cmul(ht, dt)) : cneg(h);
      }
      if (lane == n && j < N) a[e] = R.patch[j];
      if (lane == n && j == n) {
        const cx h = mk(psum.re - 1.0, psum.im - 0.0);
        const cx ht = mk(0.0, 0.0);
        W.H[n] = h;
        W.HT[n] = ht;
        rhs[e] = (mode == PC_TAN) ? cneg(cmul(ht, dt)) : cneg(h);
      }
      in[e] = lane < N && j < N;
    });
  }
  if (mode == PC_BE) {
#pragma unroll
    for (int e = 0; e < E; ++e)
      if (in[e]) W.A[lane * G + col[e]] = a[e];
    wave_lds_fence();
    // ||H(z,0)|| / max(max_i sum_j |H_z(i,j)| * ||z||, tiny)
    // (src/endgame.f90:269-285); every group computes it
    double hv = 0.0, rs = 0.0;
    if (lane < N) {
      hv = cabs_(W.H[lane]);
      for (int j = 0; j < N; ++j) rs = rs + cabs_(W.A[lane * G + j]);
    }
    const double hn = gnorm_abs<G>(hv, lane, N);
    const double rmax = gmaxval_nonneg<G>(rs, lane, N);
    const double zn = gnorm<G>(zq, lane, N);
    be = hn / ftn_max(rmax * zn, DBL_MIN);
    okout = 1;
    xout = mk(0.0, 0.0);
    release_v();
    return;
  }

  // 5. LU (src/linalg.f90:49-76) with logical row swaps: pos is the
  // logical position of this lane's row; pivot ties go to the smallest
  // position, as in the reference's first-maximum scan. Column k is in
  // group k % R, slot k / R.
  int pos = lane < N ? lane : (1 << 20);
  int ok = 1;
  // pivot lanes of the steps: PL[k] (< 16) in bits 4k..4k+3
  uint64_t plb = 0;
  if constexpr (E == 1) {
    if constexpr (EXACT)
      ok = lu_one_slot_exact<G>(N, a[0], in[0], col[0], rep, lane, pos, plb);
    else
      ok = lu_one_slot<G>(N, a[0], in[0], col[0], rep, lane, pos, plb, W.A);
  } else {
    // the slot count eu = ceil(N / RG) takes two values per class
    const int eu = slots_used<G>(N);
    constexpr int E1 = E > 1 ? E - 1 : 1;
    if constexpr (EXACT) {
      if (eu == E)
        ok = lu_slots<G, E>(N, a, in, col, rep, lane, pos, plb);
      else
        ok = lu_slots<G, E1>(N, a, in, col, rep, lane, pos, plb);
    } else {
      if (eu == E)
        ok = lu_slots_fast<G, E>(N, a, in, col, rep, lane, pos, plb, W.A);
      else
        ok = lu_slots_fast<G, E1>(N, a, in, col, rep, lane, pos, plb, W.A);
    }
  }
  okout = ok;
  if (ok != 1) {  // 0: zero pivot; 2: the fast LU's screen was inconclusive
    xout = mk(0.0, 0.0);
    release_v();
    return;
  }

  // 6. forward and back substitution (src/linalg.f90:93-111): solve_rows,
  // for this path's N (the host gives a lanes class of G only paths with
  // G / 2 < N <= G, or 2 <= N <= 4 for G = 4)
#pragma unroll
  for (int e = 0; e < E; ++e) {
    if (in[e]) W.A[lane * G + col[e]] = a[e];
    W.RB[q * E + e] = rhs[e];
  }
  wave_lds_fence();
  const bool row = lane < N;
  cx arow[G];
#pragma unroll
  for (int j = 0; j < G; ++j)
    arow[j] = row ? W.A[lane * G + j] : mk(1.0, 0.0);
  // The quotient of back-substitution step i is taken from the lane whose
  // row is at logical position i, where arow[i] is that row's diagonal
  // entry: its divisor scaling (and, for the fast quotients, its
  // reciprocal) is prepared once, off the step chain.
  const cx diag = row ? W.A[lane * G + (pos < G ? pos : 0)] : mk(1.0, 0.0);
  const cdiv_den pdiag = EXACT ? cdiv_scale(diag) : cdiv_prep(diag);
  cx bown = mk(0.0, 0.0);
  if (row) {
    // the slot that holds row `lane`'s right-hand side
    const int oc = lane < n ? lane : n;             // its column
    const int orow = lane < n ? G - 1 : n;          // its row
    bown = W.RB[((oc % RG) * G + orow) * E + oc / RG];
  }
  bool dok = true;
  xout = solve_dispatch<G, (G == 4 ? 2 : G / 2 + 1), !EXACT>(
      N, arow, bown, pdiag, pos, row, lane, plb, dok);
  if constexpr (!EXACT) {
    // a quotient outside cdiv_apply_fast's range: the caller redoes the
    // unit with the exact LU and solves (per path, as for the LU's screen)
    const int shift = LP == 64 ? 0 : (int)(__lane_id() / LP) * LP;
    const unsigned long long pmask = LP == 64 ? ~0ull : ((1ull << LP) - 1ull);
    if (((__ballot(!dok) >> shift) & pmask) != 0) okout = 2;
  }
  release_v();
}

}  // namespace phc
