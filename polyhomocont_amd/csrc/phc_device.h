// phc_device.h -- the path tracker on the device.
//
// Mapping. A path is owned by a group of G lanes (G = 4, 8 or 16, a power
// of two >= N = n + 1). Lane r of the group owns component r of every
// N-vector. Scalars of the path state machine are replicated in all lanes
// of the group and updated identically, so control flow is uniform inside
// a group. The linearization unit either runs on the group itself, row r
// in lane r (`unit` below, 64/G paths per wavefront), or on a grid of
// min(G*G, 64) lanes of which the group is one replica (`unit_wave`,
// phc_wave.h, the default; see the MODE list at the kernel).
//
// Row unit (the grid unit does the same work with an entry per lane): at a
// point zq and parameter t
//   1. lane j writes the powers zq_j^e, e = 1..maxexp, to the group's table
//      in LDS (src/sparse.f90:251-255);
//   2. lane i < n walks the terms of equation i and accumulates F^h_i and
//      row i of dF^h/dz with the prefix/suffix products of
//      src/sparse.f90:259-278, skipping factors that are exactly one;
//   3. lane i assembles row i of H, H_z, H_t (src/homotopy.f90:488-511),
//      lane n the patch row;
//   4. LU with partial pivoting, row per lane; pivot search is a group
//      max-reduction with the reference's first-maximum tie rule, rows are
//      swapped logically (src/linalg.f90:26-78);
//   5. forward/back substitution (src/linalg.f90:80-114) with the sums in
//      the reference's order.
// Arithmetic follows the reference operation by operation (see
// phc_math.h), so device trajectories reproduce the reference bit for bit.
//
// State machine. `endpoint` (src/solve.f90:217-252) with track_segment
// (src/tracker.f90:196-264), predict_rk4/tangent/newton_correct
// (src/tracker.f90:266-360) and cauchy_endgame/loops/record
// (src/endgame.f90:124-378) is flattened into a resumable state machine
// whose only blocking points are linearization units (tangent, Newton
// step, backward error). Between units a group runs cheap scalar logic;
// every group of a wavefront then executes the next unit together.
#pragma once

#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

#include "phc_math.h"
#include "../../include/phc.h"

namespace phc {

#define PHC_D __device__ __forceinline__

constexpr int WG = 256;  // threads per workgroup

// One term of a packed system: coefficient and the exponent of unknown j in
// byte j (j < 16).
struct TermRec {
  double re, im;
  uint32_t e[4];
};

// One solve request: the packed system, its random constants, options and
// output arrays (all device pointers).
struct alignas(16) Req {
  int32_t n, N, maxexp, nterms;
  int32_t maxterms;  // terms of the largest equation
  int32_t njobs;     // grid unit: job rows (0: no grid tables)
  int32_t gfold;     // grid unit: max nonzero factors of a term (bits
                     // 0-7), of a job's prefix (8-15) and suffix (16-23),
                     // product length of the short job layout (24-27)
  int32_t vstride;   // grid unit: V entries per equation
  int32_t vtail;     // grid unit: V entries after the equations' (four
                     // zeros; for the one-slot general layout also a one
                     // and the terms' prefix and suffix products)
  int32_t dbits;     // bits of the largest deg[i] - 1 (start-system powers)
  int32_t first[PHC_MAX_UNKNOWNS + 1];
  int32_t deg[PHC_MAX_UNKNOWNS];
  cx gamma;
  cx b[PHC_MAX_UNKNOWNS];
  cx patch[PHC_MAX_UNKNOWNS];
  const TermRec* terms;
  const struct WTerm* wterms;  // grid unit: compact terms
  const uint4* jobs;           // grid unit: [njobs][G G] job records
  const cx* start;  // [npaths][N]
  int64_t npaths;
  // tracker_options
  double step_init, step_max, step_min, ctol, contraction;
  int32_t maxit, grow_after, max_steps, use_endgame;
  // endgame_options
  double t_start, rf, rmin, loop_radius, agreement, egtol, rtol, closure;
  int32_t nsamples, max_winding;
  double dtheta, log_inv_rf;
  // outputs
  int32_t* o_status;
  double* o_t;
  cx* o_z;
  int32_t* o_wind;
  int32_t* o_egs;
  double* o_egerr;
  double* o_egres;
  int32_t* o_acc;
  int32_t* o_rej;
  int32_t* o_nlin;
  int32_t* o_nbe;
  uint32_t* o_done;  // paths of this request finished
};

// What a group waits for.
enum : int32_t { PC_IDLE = 0, PC_TAN = 1, PC_NEWT = 2, PC_BE = 3 };
// Where track_segment returns to.
enum : int32_t { RET_MAIN = 0, RET_RADIAL = 1, RET_ARC = 2 };
// Caller of loops / record inside cauchy_endgame.
enum : int32_t { CTX_MAIN = 0, CTX_FALLBACK = 1 };

// Scalars of a path's state machine: one copy per path, in LDS, read and
// written by every lane of the path (all lanes hold the same values, so
// their simultaneous writes store the same bits).
struct PathScalars {
  int32_t pc, req, path, ret;
  // track_segment (src/tracker.f90:207-221)
  int32_t sk, sj, stage, nit, nsucc, sacc, srej, sstatus, o_grow;
  double sa, sb, sdth, s, ds, sn, step, se, dnp, o_step_max;
  // endpoint (src/solve.f90:226-251)
  int32_t pstatus, nacc, nrej, nlin, nbe;
  double pt;
  // cauchy_endgame (src/endgame.f90:135-174)
  int32_t have_prev, gate, max_loops, eg_status, eg_wind, eg_acc, eg_rej;
  int32_t lc, lj, lmax, lctx, nsamp, rctx, lok;
  double r, rnew, best_resid, dprev, cest, cest_prev, eg_err, eg_resid,
      seg_init, rec_err, spacing, floor;
};

// The path's N-vectors, [vector][component] in LDS after the scalars; lane
// j of every lane group of the path works on component j:
//   ZT point tracked by the current segment, ZQ evaluation point / Newton
//   iterate, ACC RK4 accumulator k1 + 2 k2 + 2 k3 + k4, ZEG end game point
//   on the real axis, then the end game's estimates and loop sums.
enum : int { V_ZT, V_ZQ, V_ACC, V_ZEG, V_EST, V_PREV, V_ZOLD, V_RESZ,
             V_ZPREV, V_ZSUM, NVEC };

// LDS of one path's state: the scalars (16-byte aligned) and NVEC x G
// components.
PHC_HD constexpr size_t state_lds_bytes(int G) {
  return ((sizeof(PathScalars) + 15) & ~size_t(15)) +
         (size_t)NVEC * G * sizeof(cx);
}

// Between launches of the resumable kernel (serving engine) a path slot
// keeps its LDS state block, state_lds_bytes(G) bytes, in HBM as it is in
// LDS: the path's lanes copy it 16 bytes each. An idle slot saves only
// its program counter (its next path starts from scratch).

// A queued path of the serving engine: request slot and path index.
struct QItem {
  int32_t req, path;
};

struct Launch {
  const Req* reqs;
  const int64_t* req_base;  // [nreq + 1] prefix sums of path counts
  int32_t nreq;
  int32_t expmax;           // largest maxexp of the requests
  int32_t vmax;             // grid unit: largest term count of the requests
  int32_t max_ticks;        // resumable launches: a slot starts no unit
                            // after this many wall-clock ticks of its
                            // launch (0: no time bound; see max_units)
  unsigned long long* next; // work-queue head
  int64_t total;            // paths in the queue (fixed-size runs)
  // serving: items[i % cap] for head <= i < *tail, appended by the host
  // between launches
  const QItem* items;
  unsigned long long cap;
  const unsigned long long* tail;
  char* slots;              // [slots][state_lds_bytes(G)] saved states
                            // (resumable launches)
  int32_t resume;           // load saved states at entry
  int32_t max_units;        // units per group per launch, 0 = unlimited
  unsigned int* active;     // groups still holding a path at exit
  unsigned long long* counters;  // [4]: units redone exactly, units, unit
                                 // slots, most units run by one path slot
};

// ---------------------------------------------------------------- group ops
// Broadcasts from a runtime lane of the group use ds_bpermute (__shfl);
// reductions over the group use DPP lane swizzles inside a row of 16
// lanes: row_ror:8 (16-lane groups), row_half_mirror, then quad_perm xor 1
// and xor 2 leave every lane of the group with the reduced value.
template <int G>
PHC_D double gshfl(double v, int src) {
  return __shfl(v, src, G);
}
template <int G>
PHC_D int gshfl_i(int v, int src) {
  return __shfl(v, src, G);
}
template <int G>
PHC_D cx gshfl_c(cx v, int src) {
  return mk(__shfl(v.re, src, G), __shfl(v.im, src, G));
}

template <int CTRL>
PHC_D int dpp_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, CTRL, 0xF, 0xF, true);
}
template <int CTRL>
PHC_D double dpp_d(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i<CTRL>((int)(b & 0xffffffffll));
  const int hi = dpp_i<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
constexpr int DPP_XOR1 = 0xB1;         // quad_perm [1,0,3,2]
constexpr int DPP_XOR2 = 0x4E;         // quad_perm [2,3,0,1]
constexpr int DPP_HALF_MIRROR = 0x141; // lane i <-> 7 - i within 8 lanes
constexpr int DPP_ROR8 = 0x128;        // lane i <- (i + 8) mod 16

template <int G>
PHC_D double gmax(double v) {
  if (G >= 16) v = fmax(v, dpp_d<DPP_ROR8>(v));
  if (G >= 8) v = fmax(v, dpp_d<DPP_HALF_MIRROR>(v));
  v = fmax(v, dpp_d<DPP_XOR1>(v));
  v = fmax(v, dpp_d<DPP_XOR2>(v));
  return v;
}

// Arg-max over the group: largest v, ties to the smallest p (then any l).
template <int CTRL>
PHC_D void argmax_step(double& v, int& p, int& l) {
  const double ov = dpp_d<CTRL>(v);
  const int op = dpp_i<CTRL>(p);
  const int ol = dpp_i<CTRL>(l);
  // selects, not a branch: the comparison runs on every pivot step
  const bool take = (ov > v) | ((ov == v) & (op < p));
  v = take ? ov : v;
  p = take ? op : p;
  l = take ? ol : l;
}
template <int G>
PHC_D void gargmax(double& v, int& p, int& l) {
  if (G >= 16) argmax_step<DPP_ROR8>(v, p, l);
  if (G >= 8) argmax_step<DPP_HALF_MIRROR>(v, p, l);
  argmax_step<DPP_XOR1>(v, p, l);
  argmax_step<DPP_XOR2>(v, p, l);
}

// A ballot restricted to the calling lane's G-lane group (bit i: lane i of
// the group).
template <int G>
PHC_D unsigned long long group_ballot(bool p) {
  const unsigned long long b = __ballot(p);
  if constexpr (G == 64) return b;
  return (b >> ((__lane_id() / G) * G)) & ((1ull << G) - 1ull);
}

// norm_inf (src/linalg.f90:176-189) of the moduli `a` of the N components
// held by the group (lane < N; 0 elsewhere): the fold r = max(r, a_i) from
// r = 0 with Fortran MAX (ftn_max). Without NaNs that is the largest
// modulus; with NaNs it is NaN when the last component is NaN and otherwise
// the largest modulus after the last NaN component.
template <int G>
PHC_D double gnorm_abs(double a, int lane, int N) {
  const unsigned long long nan = group_ballot<G>(lane < N && a != a);
  const int last = nan ? 63 - __clzll(nan) : -1;
  const double r = gmax<G>(lane > last ? a : 0.0);
  return last == N - 1 ? __longlong_as_double(0x7ff8000000000000ll) : r;
}
template <int G>
PHC_D double gnorm(cx v, int lane, int N) {
  return gnorm_abs<G>(lane < N ? cabs_(v) : 0.0, lane, N);
}

// MAXVAL over the values `a` of the group's lanes < N: NaNs are ignored
// unless every value is NaN (flang's MAXVAL). Non-negative values only
// (row sums of moduli): other lanes contribute 0.
template <int G>
PHC_D double gmaxval_nonneg(double a, int lane, int N) {
  const unsigned long long nan = group_ballot<G>(lane < N && a != a);
  const double r = gmax<G>(lane < N && a == a ? a : 0.0);
  return nan == (N >= 64 ? ~0ull : (1ull << N) - 1ull)
             ? __longlong_as_double(0x7ff8000000000000ll)
             : r;
}

// The pivot of column k (src/linalg.f90:51-58) among the group's candidate
// rows, `av` = |a(i, k)| of the row at logical position pos: the scan starts
// at position k and takes a later row only if its modulus is strictly
// larger, so a NaN at position k stays the pivot and a later NaN never
// wins; ties go to the smaller position. bv, bp, bl: the pivot's modulus,
// position and lane in the group.
template <int G>
PHC_D void pivot_argmax(bool cand, double av, int pos, int k, int lane,
                        double& bv, int& bp, int& bl) {
  bv = (cand && av == av) ? av : -1.0;
  bp = pos;
  bl = lane;
  gargmax<G>(bv, bp, bl);
  const unsigned long long nk = group_ballot<G>(cand && pos == k && av != av);
  if (nk) {
    bv = __longlong_as_double(0x7ff8000000000000ll);
    bp = k;
    bl = __ffsll(nk) - 1;
  }
}

// A pointer read from memory (a request's tables) is generic to the
// compiler, so its loads become flat loads, which count on the LDS wait
// counter too: every LDS wait in the unit would then also wait for them,
// the job records prefetched for the next step included. Passing the
// tables as global-typed pointers (the unit's table arguments are template
// parameters) makes them global loads.
template <class T>
using gptr = const __attribute__((address_space(1))) T*;
// A unit's table: the copy staged in LDS, or the request's own in global
// memory.
template <bool STAGED, class T>
PHC_D auto table(T* staged, const T* global) {
  if constexpr (STAGED)
    return (const T*)staged;
  else
    return (gptr<T>)global;
}

PHC_D void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ segment
// seg_t (src/tracker.f90:142-162; seg_dt, :164-177, is formed where the
// tangent needs it). Lines are real (all lines the
// tracker uses are); arcs evaluate cos and sin of theta0 + s dtheta as
// glibc's sincos does. UNIFORM: every lane of the wavefront tracks this
// path (one path per wavefront), so the angle is taken as a scalar.
template <bool UNIFORM = false>
PHC_D cx seg_t(const PathScalars& S, double s) {
  if (S.sk == 0) {
    if (s >= 1.0) return mk(S.sb, 0.0);
    return mk(S.sa + s * (S.sb - S.sa), 0.0);
  }
  const double th = S.sa + s * S.sdth;
  const cx cs = UNIFORM ? libm_sincos_uniform(th) : libm_sincos(th);
  return cscale(S.sb, cs);  // radius in sb
}

// seg_t with the last angle's cos and sin kept by the caller: RK4's second
// and third stages share their s, the corrector's iterations share the
// fourth stage's, and the next step starts where the last one ended, so
// most units of an arc repeat the previous unit's angle. The pair (kth,
// kcs) holds theta and sincos(theta) for whichever path last computed
// them; a NaN kth (the start value) matches no angle.
template <bool UNIFORM = false>
PHC_D cx seg_t_cached(const PathScalars& S, double s, double& kth, cx& kcs) {
  if (S.sk == 0) return seg_t<UNIFORM>(S, s);
  const double th = S.sa + s * S.sdth;
  if (!(th == kth)) {
    kcs = UNIFORM ? libm_sincos_uniform(th) : libm_sincos(th);
    kth = th;
  }
  return cscale(S.sb, kcs);
}

// -------------------------------------------------------------------- unit
// LDS of one group: the power table pw(e, j) at T[e * G + j] (e >= 1), the
// N x N matrix at A[r * G + j], pivot lanes at PL[k] and the solution at
// X[j]. All indices inside a group are runtime values, so the LU and the
// triangular solves are compact loops rather than unrolled code.
template <int G>
struct GroupLds {
  cx* T;
  cx* A;
  cx* X;
  int* PL;
  PHC_D static size_t bytes(int expmax) {
    return ((size_t)(expmax + 1) * G + (size_t)G * G + G) * sizeof(cx) +
           (size_t)G * sizeof(int);
  }
  PHC_D void bind(char* base, int expmax) {
    T = (cx*)base;
    A = T + (size_t)(expmax + 1) * G;
    X = A + (size_t)G * G;
    PL = (int*)(X + G);
  }
};

// Split mode: one path per wavefront, R = 64 / G replicas of the G-lane
// group. Per chunk c, replica g evaluates term c R + g of every equation
// and writes its value and derivatives to CB; role lanes then add them in
// term order: lane (g, r < n) owns dF_r/dz_g, lane (g, n) owns F_g.
template <int G>
struct SplitLds {
  cx* CB;  // [R][G][G + 1]: value, then derivatives by nonzero position
  cx* AE;  // [G][G] exchanged Jacobian rows
  cx* FE;  // [G] exchanged values
  static constexpr int R = 64 / G;
  PHC_D static size_t bytes() {
    return ((size_t)R * G * (G + 1) + (size_t)G * G + G) * sizeof(cx);
  }
  PHC_D void bind(char* base) {
    CB = (cx*)base;
    AE = CB + (size_t)R * G * (G + 1);
    FE = AE + (size_t)G * G;
  }
};

// Mask of the unknowns with nonzero exponent in a term (bit j = unknown j),
// stored by the host in the top 16 bits of e[3] when N <= 8.
PHC_D unsigned term_mask(const TermRec& t) { return t.e[3] >> 16; }

// Products of one term (src/sparse.f90:261-276): the value c * prod and,
// for every unknown j with a nonzero exponent, c e_j v_j^(e_j - 1) times
// the prefix and suffix products of the other factors. Without SPLIT the
// derivatives are added to the row accumulator hz[j]; with SPLIT the a-th
// derivative (a-th nonzero exponent) is stored in cb[1 + a].
template <int G, bool SPLIT>
PHC_D void term_products(const TermRec& tr, const int N,
                         const cx* __restrict__ T, cx (&hz)[G], cx& val,
                         cx* __restrict__ cb) {
  const cx c = mk(tr.re, tr.im);
  int ex[G];
#pragma unroll
  for (int j = 0; j < G; ++j) ex[j] = (tr.e[j >> 2] >> (8 * (j & 3))) & 255;
  if (G > 8) {  // bytes past the last unknown may hold the term mask
#pragma unroll
    for (int j = 8; j < G; ++j) ex[j] = j < N ? ex[j] : 0;
  }
  // suffix products over the nonzero exponents
  cx suf[G];
  bool sfs[G];
  cx S = mk(1.0, 0.0);
  bool sset = false;
#pragma unroll
  for (int j = G - 1; j >= 0; --j) {
    sfs[j] = sset;
    suf[j] = S;
    if (ex[j] > 0) {
      const cx pv = T[ex[j] * G + j];
      S = sset ? cmul(pv, S) : pv;
      sset = true;
    }
  }
  // prefix products, derivatives, value
  cx P = mk(1.0, 0.0);
  bool pset = false;
  int a = 1;
#pragma unroll
  for (int j = 0; j < G; ++j) {
    if (ex[j] > 0) {
      cx d = cscale((double)ex[j], c);
      if (ex[j] >= 2) d = cmul(d, T[(ex[j] - 1) * G + j]);
      if (pset) d = cmul(d, P);
      if (sfs[j]) d = cmul(d, suf[j]);
      if (SPLIT) {
        cb[a] = d;
        ++a;
      } else {
        hz[j] = cadd(hz[j], d);
      }
      const cx pv = T[ex[j] * G + j];
      P = pset ? cmul(P, pv) : pv;
      pset = true;
    }
  }
  val = cmul(c, P);
}

// Evaluates H, H_z, H_t at (zq, t) and, for PC_TAN / PC_NEWT, solves
// H_z x = -(H_t dt) / -H. Returns x_r in lane r. For PC_BE returns the
// backward error of src/endgame.f90:269-285 in *be.
template <int G, bool SPLIT>
PHC_D void unit(const Req& R, const TermRec* __restrict__ terms,
                const int lane, const int rep, const int mode, const cx zq,
                const cx t, const cx dt, const GroupLds<G>& L,
                const SplitLds<G>& SL, cx& xout, int& okout, double& be) {
  const int N = R.N, n = R.n;
  const bool row = lane < N;
  cx* __restrict__ T = L.T;

  // 1. power table pw(e, j) = zq_j^e, e = 1..maxexp (src/sparse.f90:252-255)
  wave_lds_fence();
  if (row) {
    cx p = zq;
    T[G + lane] = p;
    for (int e = 2; e <= R.maxexp; ++e) {
      p = cmul(p, zq);
      T[e * G + lane] = p;
    }
  }
  wave_lds_fence();

  // 2. F^h_i and row i of the Jacobian (src/sparse.f90:257-278)
  cx f = mk(0.0, 0.0);
  cx hz[G];
#pragma unroll
  for (int j = 0; j < G; ++j) hz[j] = mk(0.0, 0.0);
  if (!SPLIT) {
    int k0 = 0, k1 = 0;
    if (lane < n) {
      k0 = R.first[lane];
      k1 = R.first[lane + 1];
    }
    TermRec nxt;
    if (k0 < k1) nxt = terms[k0];
    for (int k = k0; k < k1; ++k) {
      const TermRec tr = nxt;
      if (k + 1 < k1) nxt = terms[k + 1];  // prefetch the next term
      cx val;
      term_products<G, false>(tr, N, T, hz, val, nullptr);
      f = cadd(f, val);
    }
  } else {
    // chunks of R terms per equation; replica `rep` takes term c R + rep
    constexpr int RR = SplitLds<G>::R;
    const int nt_row = (lane < n) ? R.first[lane + 1] - R.first[lane] : 0;
    const int row_first = (lane < n) ? R.first[lane] : 0;
    cx* const cb_mine = SL.CB + ((size_t)rep * G + lane) * (G + 1);
    // role of this lane in the ordered sums
    const bool jac_role = lane < n && rep < N;
    const bool f_role = lane == n && rep < n;
    const int f_nt = f_role ? R.first[rep + 1] - R.first[rep] : 0;
    const unsigned below = (1u << rep) - 1u;
    cx acc = mk(0.0, 0.0);
    bool aset = false;
    const int nchunks = (R.maxterms + RR - 1) / RR;
    for (int c = 0; c < nchunks; ++c) {
      const int m = c * RR + rep;
      if (lane < n && m < nt_row) {
        cx val;
        term_products<G, true>(terms[row_first + m], N, T, hz, val, cb_mine);
        cb_mine[0] = val;
      }
      wave_lds_fence();
#pragma unroll
      for (int t = 0; t < RR; ++t) {
        const int mt = c * RR + t;
        if (jac_role && mt < nt_row) {
          const unsigned mask = term_mask(terms[row_first + mt]);
          if ((mask >> rep) & 1u) {
            const int a = __popc(mask & below);
            const cx v = SL.CB[((size_t)t * G + lane) * (G + 1) + 1 + a];
            acc = aset ? cadd(acc, v) : v;
            aset = true;
          }
        }
        if (f_role && mt < f_nt) {
          const cx v = SL.CB[((size_t)t * G + rep) * (G + 1)];
          acc = aset ? cadd(acc, v) : v;
          aset = true;
        }
      }
      wave_lds_fence();
    }
    if (jac_role) SL.AE[lane * G + rep] = acc;
    if (f_role) SL.FE[rep] = acc;
    wave_lds_fence();
    if (lane < n) {
#pragma unroll
      for (int j = 0; j < G; ++j)
        if (j < N) hz[j] = SL.AE[lane * G + j];
      f = SL.FE[lane];
    }
  }

  // 3. homotopy rows (src/homotopy.f90:491-511)
  const cx w = gshfl_c<G>(zq, n);
  cx h = mk(0.0, 0.0), ht = mk(0.0, 0.0);
  if (lane < n) {
    const int d = R.deg[lane];
    const cx gt = cmul(R.gamma, t);
    const cx omt = mk(1.0 - t.re, 0.0 - t.im);
#pragma unroll
    for (int j = 0; j < G; ++j) hz[j] = cmul(omt, hz[j]);
    cx zd1, wd1;
    zpowi_pair(zq, w, d - 1, R.dbits, zd1, wd1);
    const cx bi = R.b[lane];
    const cx g = csub(cmul(zd1, zq), cmul(cmul(bi, wd1), w));
    h = cadd(cmul(gt, g), cmul(omt, f));
    ht = csub(cmul(R.gamma, g), f);
    const cx dg = cmul(cscale((double)d, gt), zd1);
    const cx dw = cmul(cscale((double)d, cmul(gt, bi)), wd1);
#pragma unroll
    for (int j = 0; j < G; ++j) {
      if (j == lane) hz[j] = cadd(hz[j], dg);
      if (j == n) hz[j] = csub(hz[j], dw);
    }
  }
  // patch row: h_{n+1} = sum(a*z) - 1 with the sum in index order
  {
    const cx pz = row ? cmul(R.patch[lane], zq) : mk(0.0, 0.0);
    cx ssum = mk(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < G; ++j) {
      const cx pj = gshfl_c<G>(pz, j);
      const cx sj = cadd(ssum, pj);
      ssum.re = j < N ? sj.re : ssum.re;  // selects: no branch per term
      ssum.im = j < N ? sj.im : ssum.im;
    }
    if (lane == n) {
      h = mk(ssum.re - 1.0, ssum.im - 0.0);
#pragma unroll
      for (int j = 0; j < G; ++j) hz[j] = (j < N) ? R.patch[j] : mk(0.0, 0.0);
    }
  }

  if (mode == PC_BE) {
    // ||H(z,0)|| / max(max_i sum_j |H_z(i,j)| * ||z||, tiny)
    const double hn = gnorm<G>(h, lane, N);
    double rs = 0.0;
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (j < N) rs = rs + cabs_(hz[j]);
    const double rmax = gmaxval_nonneg<G>(rs, lane, N);
    const double zn = gnorm<G>(zq, lane, N);
    be = hn / ftn_max(rmax * zn, DBL_MIN);
    okout = 1;
    xout = mk(0.0, 0.0);
    return;
  }

  // right-hand side: tangent -(H_t dt/ds), Newton -H (src/tracker.f90:314,348)
  const cx b = (mode == PC_TAN) ? cneg(cmul(ht, dt)) : cneg(h);

  // 4. LU with partial pivoting in LDS, logical row swaps
  //    (src/linalg.f90:49-76): lane r keeps physical row r; pos is its
  //    logical position, pivot ties go to the smallest logical position.
  cx* __restrict__ A = L.A;
  cx* const arow = A + lane * G;
  if (row) {
#pragma unroll
    for (int j = 0; j < G; ++j)
      if (j < N) arow[j] = hz[j];
  }
  int pos = row ? lane : 1 << 20;
  int ok = 1;
  wave_lds_fence();
  for (int k = 0; k < N; ++k) {
    const cx ak = row ? arow[k] : mk(0.0, 0.0);
    double bv;
    int bp, bl;
    pivot_argmax<G>(row && pos >= k, cabs_(ak), pos, k, lane, bv, bp, bl);
    if (bv == 0.0) {
      ok = 0;
      break;
    }
    if (lane == 0) L.PL[k] = bl;
    int np = pos;
    if (pos == k) np = bp;
    if (lane == bl) np = k;
    pos = np;
    const cx* const prow = A + bl * G;
    if (row && pos > k) {
      const cx l = cdiv(ak, prow[k]);
      arow[k] = l;
      for (int j = k + 1; j < N; ++j) arow[j] = csub(arow[j], cmul(l, prow[j]));
    }
    wave_lds_fence();
  }
  okout = ok;
  if (!ok) {
    xout = mk(0.0, 0.0);
    return;
  }

  // 5. forward substitution with unit L in logical row order; Fortran's
  // SUM starts from +0
  cx sacc = mk(0.0, 0.0);
  for (int k = 0; k < N; ++k) {
    const cx cand = pos == 0 ? b : csub(b, sacc);
    const cx yk = gshfl_c<G>(cand, L.PL[k]);
    if (row && pos > k) sacc = cadd(sacc, cmul(arow[k], yk));
  }
  const cx yfin = pos == 0 ? b : csub(b, sacc);
  // back substitution, sums in increasing column order; the row at
  // logical position i is owned by lane PL[i]
  cx* __restrict__ X = L.X;
  cx xmine = mk(0.0, 0.0);
  for (int i = N - 1; i >= 0; --i) {
    cx cand = mk(0.0, 0.0);
    if (row) {
      cx num = yfin;
      if (i + 1 < N) {
        cx s = cadd(mk(0.0, 0.0), cmul(arow[i + 1], X[i + 1]));
        for (int j = i + 2; j < N; ++j) s = cadd(s, cmul(arow[j], X[j]));
        num = csub(yfin, s);
      }
      cand = cdiv(num, arow[i]);
    }
    const cx xi = gshfl_c<G>(cand, L.PL[i]);
    if (lane == 0) X[i] = xi;
    if (lane == i) xmine = xi;
    wave_lds_fence();
  }
  xout = xmine;
}

}  // namespace phc
#include "phc_wave.h"
namespace phc {

// ------------------------------------------------------------ state machine
// Internal states: a transition never calls another transition, it only
// sets S.pc; Machine::run() dispatches until the group needs a unit
// (PC_TAN, PC_NEWT, PC_BE) or the path is done (PC_IDLE).
enum : int32_t {
  X_SEG_HEAD = 16,
  X_SEG_FINISH,
  X_MAIN_DONE,
  X_EG_TOP,
  X_EG_AFTER_LOOPS_MAIN,
  X_EG_AFTER_RECORD_MAIN,
  X_EG_RADIAL,
  X_EG_AFTER_RADIAL,
  X_EG_AFTER_LOOPS_FALLBACK,
  X_EG_RETURN,
  X_LOOPS_BEGIN,
  X_LOOPS_ARC,
  X_LOOPS_AFTER_ARC,
  X_LOOPS_END,
  X_RECORD_BEGIN,
};

// All functions below run with group-uniform control flow.
template <int G>
struct Machine {
  const Req& R;
  PathScalars& S;
  cx* const V;  // the path's vectors, [NVEC][G] (LDS)
  const int lane;
  bool writer;  // the replica that stores results (split mode: replica 0)

  // this lane's component of each vector
  PHC_D cx& zt() const { return V[V_ZT * G + lane]; }
  PHC_D cx& zq() const { return V[V_ZQ * G + lane]; }
  PHC_D cx& acc() const { return V[V_ACC * G + lane]; }
  PHC_D cx& zeg() const { return V[V_ZEG * G + lane]; }
  PHC_D cx& est() const { return V[V_EST * G + lane]; }
  PHC_D cx& prev() const { return V[V_PREV * G + lane]; }
  PHC_D cx& zold() const { return V[V_ZOLD * G + lane]; }
  PHC_D cx& resz() const { return V[V_RESZ * G + lane]; }
  PHC_D cx& zprev() const { return V[V_ZPREV * G + lane]; }
  PHC_D cx& zsum() const { return V[V_ZSUM * G + lane]; }

  PHC_D double norm(cx v) const { return gnorm<G>(v, lane, R.N); }

  // --- track_segment -------------------------------------------------------
  PHC_D void seg_begin(int kind, double a, double b, int j0, int ret,
                       double step_init, double step_max, int grow) {
    S.sk = kind;
    S.sj = j0;
    S.sa = a;  // line: t0; arc: theta0 = (j - 1) dtheta
    S.sb = b;  // line: t1; arc: radius
    S.sdth = kind ? R.dtheta : 0.0;
    S.ret = ret;
    S.s = 0.0;
    S.step = ftn_min(step_init, step_max);
    S.o_step_max = step_max;
    S.o_grow = grow;
    S.nsucc = 0;
    S.sacc = 0;
    S.srej = 0;
    S.sstatus = PHC_PATH_SUCCESS;
    S.pc = X_SEG_HEAD;
  }

  // Loop head of track_segment (src/tracker.f90:222-236).
  PHC_D void seg_head() {
    if (!(S.s < 1.0)) {
      S.pc = X_SEG_FINISH;
      return;
    }
    if (S.sacc + S.srej >= R.max_steps) {
      S.sstatus = PHC_PATH_FAILED_MAX_STEPS;
      S.pc = X_SEG_FINISH;
      return;
    }
    S.ds = ftn_min(S.step, 1.0 - S.s);
    S.sn = S.s + S.ds;
    if (1.0 - S.sn <= 1.0e-3 * S.step) {
      S.sn = 1.0;
      S.ds = 1.0 - S.s;
    }
    S.stage = 0;
    zq() = zt();
    S.se = S.s;
    S.pc = PC_TAN;
  }

  // src/tracker.f90:240-257
  PHC_D void seg_accept() {
    zt() = zq();
    S.s = S.sn;
    S.sacc += 1;
    S.nsucc += 1;
    if (S.nsucc >= S.o_grow) {
      S.step = ftn_min(2.0 * S.step, S.o_step_max);
      S.nsucc = 0;
    }
    S.pc = X_SEG_HEAD;
  }

  PHC_D void seg_reject() {
    S.srej += 1;
    S.nsucc = 0;
    S.step = 0.5 * S.step;
    if (S.step < R.step_min) {
      S.sstatus = PHC_PATH_FAILED_MIN_STEP;
      S.pc = X_SEG_FINISH;
      return;
    }
    S.pc = X_SEG_HEAD;
  }

  // predict_rk4 (src/tracker.f90:283-291)
  PHC_D void on_tangent(cx k, int ok) {
    S.nlin += 1;
    if (!ok) {
      seg_reject();
      return;
    }
    if (S.stage == 0) {
      acc() = k;
      zq() = cadd(zt(), cscale(0.5 * S.ds, k));
      S.se = S.s + 0.5 * S.ds;
      S.stage = 1;
      S.pc = PC_TAN;
    } else if (S.stage == 1) {
      acc() = cadd(acc(), cscale(2.0, k));
      zq() = cadd(zt(), cscale(0.5 * S.ds, k));
      S.se = S.s + 0.5 * S.ds;
      S.stage = 2;
      S.pc = PC_TAN;
    } else if (S.stage == 2) {
      acc() = cadd(acc(), cscale(2.0, k));
      zq() = cadd(zt(), cscale(S.ds, k));
      S.se = S.s + S.ds;
      S.stage = 3;
      S.pc = PC_TAN;
    } else {
      acc() = cadd(acc(), k);
      zq() = cadd(zt(), cscale(S.ds / 6.0, acc()));
      S.nit = 1;
      S.dnp = DBL_MAX;
      if (R.maxit < 1)
        seg_reject();
      else
        S.pc = PC_NEWT;
    }
  }

  // newton_correct (src/tracker.f90:342-358)
  PHC_D void on_newton(cx dz, int ok) {
    S.nlin += 1;
    if (!ok) {
      seg_reject();
      return;
    }
    zq() = cadd(zq(), dz);
    const double dn = norm(dz);
    const double zn = norm(zq());
    if (dn <= R.ctol * zn) {
      seg_accept();
      return;
    }
    if (S.nit > 1 && dn > R.contraction * S.dnp) {
      seg_reject();
      return;
    }
    S.dnp = dn;
    if (S.nit >= R.maxit) {
      seg_reject();
      return;
    }
    S.nit += 1;
    S.pc = PC_NEWT;
  }

  PHC_D void seg_finish() {
    S.pc = (S.ret == RET_MAIN)     ? X_MAIN_DONE
           : (S.ret == RET_RADIAL) ? X_EG_AFTER_RADIAL
                                   : X_LOOPS_AFTER_ARC;
  }

  // --- endpoint (src/solve.f90:230-251) -------------------------------------
  PHC_D void start_path(const cx zstart) {
    zt() = zstart;
    S.nacc = S.nrej = S.nlin = S.nbe = 0;
    const double t_end = R.use_endgame ? R.t_start : 0.0;
    seg_begin(0, 1.0, t_end, 0, RET_MAIN, R.step_init, R.step_max,
              R.grow_after);
  }

  PHC_D void main_done() {
    S.pstatus = S.sstatus;
    S.pt = seg_t(S, S.s).re;
    S.nacc = S.sacc;
    S.nrej = S.srej;
    if (S.sstatus != PHC_PATH_SUCCESS || !R.use_endgame) {
      finish(zt(), 0, 0, 0.0, 0.0);
      return;
    }
    // cauchy_endgame set-up (src/endgame.f90:160-174)
    S.seg_init = 1.0;  // segopts: step_init = step_max = 1, grow_after = 2
    S.eg_status = PHC_ENDGAME_FAILED;
    resz() = zt();
    zeg() = zt();
    S.r = R.t_start;
    S.have_prev = 0;
    S.best_resid = DBL_MAX;
    S.gate = 0;
    S.dprev = -1.0;
    S.cest = -1.0;
    S.cest_prev = -1.0;
    S.max_loops = R.max_winding;
    S.eg_wind = 0;
    S.eg_err = DBL_MAX;
    S.eg_resid = DBL_MAX;
    S.eg_acc = 0;
    S.eg_rej = 0;
    S.pc = X_EG_TOP;
  }

  // --- cauchy_endgame main loop (src/endgame.f90:176-236) --------------------
  PHC_D void eg_top() {
    if (S.gate || S.r < R.loop_radius) {
      if (!S.gate) S.max_loops = R.max_winding;
      S.lmax = S.max_loops;
      S.lctx = CTX_MAIN;
      S.pc = X_LOOPS_BEGIN;
      return;
    }
    S.lok = 0;
    S.pc = X_EG_AFTER_LOOPS_MAIN;
  }

  PHC_D void eg_after_loops_main() {
    if (!S.lok) {
      S.have_prev = 0;
      S.gate = 0;
      S.pc = X_EG_RADIAL;
      return;
    }
    S.rctx = CTX_MAIN;
    S.pc = X_RECORD_BEGIN;
  }

  PHC_D void eg_after_record_main() {
    if (S.eg_status == PHC_ENDGAME_CONVERGED) {
      S.pc = X_EG_RETURN;
      return;
    }
    prev() = est();
    S.have_prev = 1;
    S.pc = X_EG_RADIAL;
  }

  PHC_D void eg_radial() {
    S.rnew = R.rf * S.r;
    if (S.rnew < R.rmin) {
      S.pc = X_EG_RETURN;
      return;
    }
    zold() = zeg();
    zt() = zeg();
    seg_begin(0, S.r, S.rnew, 0, RET_RADIAL, S.seg_init, 1.0, 2);
  }

  PHC_D void eg_after_radial() {
    zeg() = zt();
    S.seg_init = ftn_min(ftn_max(S.step, R.step_min), 1.0);
    S.eg_acc += S.sacc;
    S.eg_rej += S.srej;
    if (S.sstatus != PHC_PATH_SUCCESS) {
      zeg() = zold();
      if (!(S.gate || S.r < R.loop_radius)) {
        S.lmax = R.max_winding;
        S.lctx = CTX_FALLBACK;
        S.pc = X_LOOPS_BEGIN;
        return;
      }
      S.pc = X_EG_RETURN;
      return;
    }
    S.r = S.rnew;
    const double d = norm(csub(zeg(), zold()));
    if (!S.gate && S.dprev > 0.0 && d > 0.0) {
      S.cest_prev = S.cest;
      S.cest = -1.0;
      if (S.dprev > d) S.cest = R.log_inv_rf / log(S.dprev / d);
      if (S.cest >= 0.5 && S.cest <= (double)R.max_winding + 0.5 &&
          S.cest_prev > 0.0) {
        if (fabs(S.cest - S.cest_prev) <= R.agreement * ftn_max(1.0, S.cest)) {
          S.gate = 1;
          S.max_loops = min(R.max_winding, 2 * (int)round(S.cest) + 2);
        }
      }
    }
    S.dprev = d;
    S.pc = X_EG_TOP;
  }

  PHC_D void eg_after_loops_fallback() {
    S.have_prev = 0;
    if (S.lok) {
      S.rctx = CTX_FALLBACK;
      S.pc = X_RECORD_BEGIN;
      return;
    }
    S.pc = X_EG_RETURN;
  }

  // record (src/endgame.f90:240-265)
  PHC_D void record_begin() {
    S.rec_err = DBL_MAX;
    if (S.have_prev)
      S.rec_err = norm(csub(est(), prev())) / ftn_max(norm(est()), DBL_MIN);
    zq() = est();
    S.pc = PC_BE;
  }

  PHC_D void on_backward_error(double resid) {
    S.nbe += 1;
    const double err = S.rec_err;
    bool take = S.eg_status == PHC_ENDGAME_FAILED;
    if (!take) {  // better (src/endgame.f90:287-311)
      const bool valid = resid <= R.rtol;
      const bool best_valid = S.best_resid <= R.rtol;
      take = (valid != best_valid) ? valid : (err < S.eg_err);
    }
    if (take) {
      S.eg_status = PHC_ENDGAME_NOT_CONVERGED;
      resz() = est();
      S.eg_wind = S.lc;
      S.eg_err = err;
      S.eg_resid = resid;
      S.best_resid = resid;
    }
    if (err <= R.egtol && resid <= R.rtol) S.eg_status = PHC_ENDGAME_CONVERGED;
    S.pc = (S.rctx == CTX_MAIN) ? X_EG_AFTER_RECORD_MAIN : X_EG_RETURN;
  }

  PHC_D void eg_return() {
    S.nacc += S.eg_acc;
    S.nrej += S.eg_rej;
    if (S.eg_status == PHC_ENDGAME_FAILED) {
      S.pstatus = PHC_PATH_ENDGAME_FAILED;
      finish(zeg(), 0, 0, 0.0, 0.0);
      return;
    }
    S.pt = 0.0;
    finish(resz(), S.eg_wind, S.eg_status, S.eg_err, S.eg_resid);
  }

  // loops (src/endgame.f90:349-376)
  PHC_D void loops_begin() {
    S.floor = 10.0 * R.ctol * norm(zeg());
    zt() = zeg();
    zsum() = mk(0.0, 0.0);
    S.nsamp = 0;
    S.spacing = DBL_MAX;
    S.lc = 1;
    S.lj = 1;
    if (S.lc > S.lmax) {
      S.lok = 0;
      S.pc = X_LOOPS_END;
      return;
    }
    S.pc = X_LOOPS_ARC;
  }

  PHC_D void loops_arc() {
    zsum() = cadd(zsum(), zt());
    S.nsamp += 1;
    zprev() = zt();
    seg_begin(1, (double)(S.lj - 1) * R.dtheta, S.r, S.lj - 1, RET_ARC,
              S.seg_init, 1.0, 2);
  }

  PHC_D void loops_after_arc() {
    S.seg_init = ftn_min(ftn_max(S.step, R.step_min), 1.0);
    S.eg_acc += S.sacc;
    S.eg_rej += S.srej;
    if (S.sstatus != PHC_PATH_SUCCESS) {
      S.lok = 0;
      S.pc = X_LOOPS_END;
      return;
    }
    S.spacing = ftn_min(S.spacing, norm(csub(zt(), zprev())));
    S.lj += 1;
    if (S.lj <= R.nsamples) {
      S.pc = X_LOOPS_ARC;
      return;
    }
    if (norm(csub(zt(), zeg())) <= ftn_max(R.closure * S.spacing, S.floor)) {
      est() = cdiv(zsum(), mk((double)S.nsamp, 0.0));
      S.lok = 1;
      S.pc = X_LOOPS_END;
      return;
    }
    S.lc += 1;
    S.lj = 1;
    if (S.lc > S.lmax) {
      S.lok = 0;
      S.pc = X_LOOPS_END;
      return;
    }
    S.pc = X_LOOPS_ARC;
  }

  // --- results -------------------------------------------------------------
  PHC_D void finish(cx z, int wind, int egs, double egerr, double egres) {
    const int64_t k = S.path;
    if (writer && lane < R.N) R.o_z[k * R.N + lane] = z;
    if (writer && lane == 0) {
      R.o_status[k] = S.pstatus;
      R.o_t[k] = S.pt;
      R.o_wind[k] = wind;
      R.o_egs[k] = egs;
      R.o_egerr[k] = egerr;
      if (R.o_egres) R.o_egres[k] = egres;
      R.o_acc[k] = S.nacc;
      R.o_rej[k] = S.nrej;
      if (R.o_nlin) R.o_nlin[k] = S.nlin;
      if (R.o_nbe) R.o_nbe[k] = S.nbe;
      __threadfence();
      if (R.o_done) atomicAdd(R.o_done, 1u);
    }
    S.pc = PC_IDLE;
  }

  // Dispatch until a unit is needed or the path is finished.
  PHC_D void run() {
    for (;;) {
      switch (S.pc) {
        case X_SEG_HEAD: seg_head(); break;
        case X_SEG_FINISH: seg_finish(); break;
        case X_MAIN_DONE: main_done(); break;
        case X_EG_TOP: eg_top(); break;
        case X_EG_AFTER_LOOPS_MAIN: eg_after_loops_main(); break;
        case X_EG_AFTER_RECORD_MAIN: eg_after_record_main(); break;
        case X_EG_RADIAL: eg_radial(); break;
        case X_EG_AFTER_RADIAL: eg_after_radial(); break;
        case X_EG_AFTER_LOOPS_FALLBACK: eg_after_loops_fallback(); break;
        case X_EG_RETURN: eg_return(); break;
        case X_LOOPS_BEGIN: loops_begin(); break;
        case X_LOOPS_ARC: loops_arc(); break;
        case X_LOOPS_AFTER_ARC: loops_after_arc(); break;
        case X_LOOPS_END:
          S.pc = (S.lctx == CTX_MAIN) ? X_EG_AFTER_LOOPS_MAIN
                                      : X_EG_AFTER_LOOPS_FALLBACK;
          break;
        case X_RECORD_BEGIN: record_begin(); break;
        default: return;  // PC_IDLE, PC_TAN, PC_NEWT, PC_BE
      }
    }
  }
};

// --------------------------------------------------------------------- kernel
// Lane mappings (MODE):
//   MODE_ROW    one path per G-lane group, row per lane;
//   MODE_SPLIT  one path per wavefront: 64/G identical replicas of the group
//               split the terms of each equation in the evaluation;
//   MODE_GRID   one path per min(G*G, 64) lanes (phc_wave.h), entries of
//               the linear system spread over the lanes.
// STAGED: the launch has one request; its descriptor and its term program
// are copied into LDS once per workgroup and every path reads them from
// there. Otherwise requests and programs are read from global memory (L2).
constexpr int MODE_ROW = 0, MODE_SPLIT = 1, MODE_GRID = 2;

constexpr size_t staged_req_bytes() { return (sizeof(Req) + 15) & ~size_t(15); }

template <int G, int MODE>
constexpr int lanes_per_path() {
  return MODE == MODE_ROW ? G
                          : (MODE == MODE_SPLIT ? 64 : GridShape<G>::LP);
}

template <int G, int MODE>
PHC_D size_t unit_lds_bytes(int expmax, int vmax) {
  return state_lds_bytes(G) +
         (MODE == MODE_GRID
              ? WaveLds<G>::bytes(expmax, vmax)
              : (MODE == MODE_SPLIT
                     ? GroupLds<G>::bytes(expmax) + SplitLds<G>::bytes()
                     : GroupLds<G>::bytes(expmax)));
}

// The 16-lane grid kernel reading its program from global memory holds
// the fast and the exact unit in 266 VGPRs unbounded; it is held to two
// waves per SIMD (256), which spills 10 of them to scratch and still runs
// its units 4 % faster than the exact unit alone at two waves.
template <int G, bool STAGED, int MODE>
__global__ void __launch_bounds__(WG, (G == 16 && !STAGED && MODE == MODE_GRID) ? 2 : 1)
    track_kernel(Launch L) {
  extern __shared__ __align__(16) char lds_raw[];
  constexpr bool SPLIT = MODE == MODE_SPLIT;
  constexpr bool GRID = MODE == MODE_GRID;
  constexpr int LP = lanes_per_path<G, MODE>();
  const int lane = threadIdx.x % G;
  const int rep = (int)(threadIdx.x % LP) / G;
  const int owner = (int)(threadIdx.x / LP);
  const int per_block = (int)(blockDim.x / LP);
  const int slot = blockIdx.x * per_block + owner;
  const bool writer = rep == 0;

  Req* sreq = (Req*)lds_raw;
  TermRec* sterms = (TermRec*)(lds_raw + staged_req_bytes());
  WTerm* swterms = (WTerm*)(lds_raw + staged_req_bytes());
  uint4* sjobs = nullptr;
  size_t base = 0;
  if (STAGED) {
    const uint4* src = (const uint4*)&L.reqs[0];
    uint4* dst = (uint4*)sreq;
    for (int i = threadIdx.x; i < (int)(sizeof(Req) / 16); i += blockDim.x)
      dst[i] = src[i];
    // the grid unit's terms end with a zero-coefficient term
    const int nt = L.reqs[0].nterms + (GRID ? 1 : 0);
    const uint4* tsrc = GRID ? (const uint4*)L.reqs[0].wterms
                             : (const uint4*)L.reqs[0].terms;
    uint4* tdst = (uint4*)sterms;
    for (int i = threadIdx.x; i < nt * 2; i += blockDim.x) tdst[i] = tsrc[i];
    base = staged_req_bytes() + (size_t)nt * sizeof(TermRec);
    if (GRID) {
      sjobs = (uint4*)(lds_raw + base);
      const int nj = job_records<G>(L.reqs[0]);
      const uint4* jsrc = L.reqs[0].jobs;
      for (int i = threadIdx.x; i < nj; i += blockDim.x) sjobs[i] = jsrc[i];
      base += (size_t)nj * sizeof(uint4);
    }
    __syncthreads();
  }
  // this path's LDS: its state (scalars, then vectors), then the unit's
  char* const pbase = lds_raw + base +
                      (size_t)owner * unit_lds_bytes<G, MODE>(L.expmax, L.vmax);
  PathScalars& S = *(PathScalars*)pbase;
  cx* const SV = (cx*)(pbase + ((sizeof(PathScalars) + 15) & ~size_t(15)));
  char* mine = pbase + state_lds_bytes(G);
  GroupLds<G> gl;
  SplitLds<G> sl;
  WaveLds<G> wl;
  if (GRID) {
    wl.bind(mine, L.expmax, L.vmax);
    // padding of the term values must read as zero (a staged launch has
    // one request, so one layout, and never writes the padding)
    for (int i = (int)(threadIdx.x % LP); i < L.vmax; i += LP)
      wl.V[i] = mk(0.0, 0.0);
    wave_lds_fence();
  } else {
    gl.bind(mine, L.expmax);
    if (SPLIT) sl.bind(mine + GroupLds<G>::bytes(L.expmax));
  }

  // the path's state block, saved by the previous launch
  constexpr int SQ = (int)(state_lds_bytes(G) / 16);
  if (L.resume) {
    const uint4* src = (const uint4*)(L.slots + (size_t)slot * state_lds_bytes(G));
    for (int i = (int)(threadIdx.x % LP); i < SQ; i += LP)
      ((uint4*)pbase)[i] = src[i];
  } else {
    S.pc = PC_IDLE;
  }
  wave_lds_fence();
  int units = 0;
  // passes of this wavefront's loop in which at least one of its paths ran
  // a unit: units / (passes * paths per wavefront) is the lane occupancy
  // left after path divergence
  int passes = 0;
  // Resumable launches end on a unit budget and, with max_ticks, on a time
  // budget: a slot whose units are cheap then keeps working until the
  // launch's end instead of idling after its unit budget while slots with
  // expensive units finish theirs. The clock is read every fourth unit;
  // every slot runs at least one unit per launch. (Results never depend
  // on where a launch ends: the state is saved and resumed exactly.)
  const unsigned long long t_launch = L.max_ticks ? wall_clock64() : 0ull;
  auto budget_spent = [&]() {
    if (L.max_units && units >= L.max_units) return true;
    return L.max_ticks && units > 0 && (units & 3) == 0 &&
           wall_clock64() - t_launch >= (unsigned long long)L.max_ticks;
  };
  // the last arc angle and its cos and sin (seg_t_cached): 4- and 8-lane
  // grids with the program in LDS, i.e. single-request solves, whose time
  // is their longest path's unit latency. The 16-lane kernels have no
  // registers to spare; in the engine's kernels six more registers cost
  // about what the saved sincos calls gain.
  double kth = __builtin_nan("");
  cx kcs = mk(0.0, 0.0);
  for (;;) {
    if (S.pc == PC_IDLE) {
      if (budget_spent()) break;
      // pop the next queue index; -1 when the queue is empty. One
      // fetch-and-add, never retried (a compare-and-swap loop serializes the
      // slots that pop together, at every launch's start, for milliseconds).
      // Pops that race past the end leave the head beyond the tail, by at
      // most one per slot and launch; those indices hold no item (the
      // engine appends its next items from the head).
      long long idx = -1;
      if (threadIdx.x % LP == 0) {
        const unsigned long long end =
            L.items ? __atomic_load_n(L.tail, __ATOMIC_RELAXED)
                    : (unsigned long long)L.total;
        if (__atomic_load_n(L.next, __ATOMIC_RELAXED) < end) {
          const unsigned long long h = atomicAdd(L.next, 1ull);
          if (h < end) idx = (long long)h;
        }
      }
      idx = __shfl(idx, 0, LP);
      if (idx < 0) break;
      int lo = 0;
      if (L.items) {
        const QItem it = L.items[(unsigned long long)idx % L.cap];
        lo = it.req;
        S.path = it.path;
      } else {
        if (!STAGED) {
          int hi = L.nreq - 1;
          while (lo < hi) {  // last request with base <= idx
            const int mid = (lo + hi + 1) >> 1;
            if (L.req_base[mid] <= idx)
              lo = mid;
            else
              hi = mid - 1;
          }
        }
        S.path = (int32_t)(idx - L.req_base[lo]);
      }
      S.req = lo;
      const Req& R = STAGED ? *sreq : L.reqs[lo];
      if (GRID && !STAGED) {
        // the previous path of this slot may have come from a request
        // whose values sat where this one's zero padding is
        for (int i = (int)(threadIdx.x % LP); i < L.vmax; i += LP)
          wl.V[i] = mk(0.0, 0.0);
      }
      Machine<G> M{R, S, SV, lane, writer};
      M.start_path((lane < R.N) ? R.start[(size_t)S.path * R.N + lane]
                                : mk(0.0, 0.0));
    } else {
      if (budget_spent()) break;
      const Req& R = STAGED ? *sreq : L.reqs[S.req];
      Machine<G> M{R, S, SV, lane, writer};
      // the unit's inputs, from the path's state in LDS
      auto inputs = [&](cx& t, cx& dt, cx& zl) {
        t = mk(0.0, 0.0);
        dt = mk(0.0, 0.0);
        if (S.pc != PC_BE) {
          const double sv = (S.pc == PC_TAN) ? S.se : S.sn;
          if constexpr (GRID && STAGED && G <= 8)
            t = seg_t_cached<LP == 64>(S, sv, kth, kcs);
          else
            t = seg_t<LP == 64>(S, sv);
          if (S.pc == PC_TAN)
            dt = (S.sk == 0) ? mk(S.sb - S.sa, 0.0)
                             : mk(0.0 * t.re - S.sdth * t.im,
                                  0.0 * t.im + S.sdth * t.re);
        }
        zl = (lane < R.N) ? SV[V_ZQ * G + lane] : mk(0.0, 0.0);
      };
      cx x;
      int ok;
      double be = 0.0;
      // Grid kernels run the fast unit (lu_one_slot, lu_slots_fast, fast
      // quotients) and redo it exactly when it says so.
      if (GRID) {
        {
          cx t, dt, zl;
          inputs(t, dt, zl);
          unit_wave<G, false>(R, table<STAGED>(swterms, R.wterms),
                              table<STAGED>(sjobs, R.jobs), lane, rep, S.pc, zl, t,
                              dt, wl, x, ok, be);
        }
        if (ok == 2) {
          // the fast LU's screen was inconclusive, or a quotient fell
          // outside the fast division's range: the same unit with the exact
          // pivot rule and divisions in every step (inputs read again from
          // LDS, so none of them stays in registers through the first unit)
          if (writer && lane == 0 && L.counters)
            atomicAdd(&L.counters[0], 1ull);
          __asm__ __volatile__("" ::: "memory");
          cx t, dt, zl;
          inputs(t, dt, zl);
          unit_wave<G, true>(R, table<STAGED>(swterms, R.wterms),
                             table<STAGED>(sjobs, R.jobs), lane, rep, S.pc, zl, t,
                             dt, wl, x, ok, be);
        }
      } else {
        cx t, dt, zl;
        inputs(t, dt, zl);
        unit<G, SPLIT>(R, STAGED ? sterms : R.terms, lane, rep, S.pc, zl, t,
                       dt, gl, sl, x, ok, be);
      }
      units++;
      // this pass of the wavefront ran a unit: its lowest active lane counts
      if ((int)__lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1)
        ++passes;
      if (S.pc == PC_TAN)
        M.on_tangent(x, ok);
      else if (S.pc == PC_NEWT)
        M.on_newton(x, ok);
      else
        M.on_backward_error(be);
    }
    const Req& R = STAGED ? *sreq : L.reqs[S.req];
    Machine<G> M{R, S, SV, lane, writer};
    M.run();
  }
  if (L.slots) {
    wave_lds_fence();
    char* dst = L.slots + (size_t)slot * state_lds_bytes(G);
    if (S.pc == PC_IDLE) {
      if (threadIdx.x % LP == 0) ((PathScalars*)dst)->pc = PC_IDLE;
    } else {
      for (int i = (int)(threadIdx.x % LP); i < SQ; i += LP)
        ((uint4*)dst)[i] = ((const uint4*)pbase)[i];
    }
  }
  if (writer && S.pc != PC_IDLE && lane == 0 && L.active) atomicAdd(L.active, 1u);
  if (writer && lane == 0 && L.counters) {
    atomicAdd(&L.counters[1], (unsigned long long)units);
    atomicMax(&L.counters[3], (unsigned long long)units);
  }
  if (L.counters) {
    for (int off = 32; off > 0; off >>= 1) passes += __shfl_xor(passes, off);
    if (threadIdx.x % 64 == 0)
      atomicAdd(&L.counters[2], (unsigned long long)passes * (64 / LP));
  }
}

}  // namespace phc
