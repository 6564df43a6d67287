// phc_runtime.hip -- host side of libphc_hip.so: the C ABI of include/phc.h.
//
// A call packs each solve request (system, random constants, options) into
// a device `Req` record, uploads start points, and runs the persistent
// tracking kernel over a device work queue of (request, path) items until
// every path is done. Requests are grouped by lanes-per-path class (G = 4,
// 8, 16 from N = n + 1); each class is one kernel instantiation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cstdarg>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/phc.h"
#include "phc_device.h"

using namespace phc;

struct phc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // streams of the second and later lanes classes of one run: the classes'
  // persistent launches overlap on the GPU and join the context stream
  std::vector<hipStream_t> side_streams;
  int num_cus = 0;
  int wg_threads = 0;  // 0: choose per launch
  // 0: by rule, 1: one path per G-lane group, 2: one path per wavefront
  // with split terms, 3: one path per G*G-lane grid
  int mapping = 0;
  int staging = -1;    // program in LDS: -1 when it fits, 0 never
  size_t lds_limit = 64 * 1024;  // largest LDS of one workgroup
  int wall_khz = 100000;         // rate of the kernels' wall clock (kHz)
  // autotuned launch choices, keyed by the shape of a lanes class (see
  // tune_key): mapping 1-3, workgroup threads, staging (phc_ctx_autotune,
  // phc_ctx_autotune_batch)
  std::map<std::array<int64_t, 4>, std::array<int, 3>> tuned;
  unsigned long long counters[4] = {0, 0, 0, 0};  // of the last run
  unsigned long long reruns = 0;  // units redone exactly in the last run
  // launch shape of every lanes class of the last run (phc_ctx_last_plan)
  std::vector<std::array<int32_t, 8>> last_plan;
  std::mutex mu;
  char err[512] = {0};
};

namespace {

// Message of the last failed call made by this thread (any context).
thread_local char t_err[512] = {0};

int fail(phc_ctx* ctx, int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof(t_err), fmt, ap);
  va_end(ap);
  if (ctx) memcpy(ctx->err, t_err, sizeof(t_err));
  return code;
}

#define HIPCHK(ctx, call)                                                  \
  do {                                                                     \
    hipError_t e_ = (call);                                                \
    if (e_ != hipSuccess)                                                  \
      return fail(ctx, PHC_ERR_DEVICE, "%s failed: %s", #call,             \
                  hipGetErrorString(e_));                                  \
  } while (0)

int lanes_for(const phc_ctx*, int N) {
  if (N <= 4) return 4;
  if (N <= 8) return 8;
  return 16;
}

// Device arena: one hipMalloc per call, bump-allocated.
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  ~Arena() {
    if (base) (void)hipFree(base);
  }
  static size_t al(size_t x) { return (x + 255) & ~size_t(255); }
};

struct HostReq {
  Req r;                         // device pointers filled at upload
  std::vector<TermRec> terms;
  std::vector<WTerm> wterms;     // grid unit tables (build_grid)
  std::vector<uint4> jobs;
  const double* start = nullptr; // host [npaths][N][2]
  phc_endpoints* out = nullptr;
  int64_t npaths = 0;
  int G = 8;
};

int validate_and_pack(phc_ctx* ctx, const phc_system* sys,
                      const phc_homotopy* hom, const phc_solve_options* o,
                      int64_t npaths, HostReq& h) {
  if (!sys || !hom || !o) return fail(ctx, PHC_ERR_INVALID_INPUT, "null input");
  const int n = sys->n, N = n + 1;
  if (n < 1) return fail(ctx, PHC_ERR_INVALID_INPUT, "n = %d < 1", n);
  // a valid system the kernels cannot represent: not an input error
  if (N > PHC_MAX_UNKNOWNS)
    return fail(ctx, PHC_ERR_CAPACITY, "n = %d: the kernels hold at most %d "
                "unknowns", n, PHC_MAX_UNKNOWNS - 1);
  if (sys->nterms < n || !sys->first || !sys->deg || !sys->hexps ||
      !sys->coefs)
    return fail(ctx, PHC_ERR_INVALID_INPUT, "incomplete system");
  if (npaths < 0) return fail(ctx, PHC_ERR_INVALID_INPUT, "npaths < 0");
  Req& r = h.r;
  memset(&r, 0, sizeof(r));
  r.n = n;
  r.N = N;
  r.nterms = sys->nterms;
  if (sys->first[0] != 0 || sys->first[n] != sys->nterms)
    return fail(ctx, PHC_ERR_INVALID_INPUT, "bad term offsets");
  int maxexp = 0;
  for (int i = 0; i <= n; ++i) r.first[i] = sys->first[i];
  for (int i = 0; i < n; ++i) {
    if (sys->first[i + 1] <= sys->first[i])
      return fail(ctx, PHC_ERR_INVALID_INPUT, "equation %d has no terms", i);
    if (sys->deg[i] < 1)
      return fail(ctx, PHC_ERR_INVALID_INPUT, "degree of equation %d < 1", i);
    r.deg[i] = sys->deg[i];
  }
  h.terms.resize(sys->nterms);
  for (int k = 0; k < sys->nterms; ++k) {
    TermRec& t = h.terms[k];
    memset(&t, 0, sizeof(t));
    t.re = sys->coefs[2 * k];
    t.im = sys->coefs[2 * k + 1];
    int tot = 0;
    for (int j = 0; j < N; ++j) {
      const int e = sys->hexps[(size_t)k * N + j];
      if (e < 0)
        return fail(ctx, PHC_ERR_INVALID_INPUT,
                    "exponent %d of term %d is negative", e, k);
      if (e > PHC_MAX_EXPONENT)
        return fail(ctx, PHC_ERR_CAPACITY,
                    "exponent %d of term %d: the kernels hold at most %d", e,
                    k, PHC_MAX_EXPONENT);
      t.e[j >> 2] |= (uint32_t)e << (8 * (j & 3));
      maxexp = std::max(maxexp, e);
      tot += e;
    }
    (void)tot;
  }
  r.maxexp = std::max(maxexp, 1);
  int dmax = 0;
  for (int i = 0; i < n; ++i) dmax = std::max(dmax, r.deg[i] - 1);
  while (dmax >> r.dbits) ++r.dbits;
  // Nonzero-exponent mask of each term in e[3] bits 16..31. Exponent bytes
  // of unknowns 0..7 live in e[0..1], so e[3] is free only when N <= 8; the
  // decode of wider groups ignores bytes at j >= N.
  if (N <= 8) {
    for (TermRec& t : h.terms) {
      uint32_t mask = 0;
      for (int j = 0; j < N; ++j)
        if ((t.e[j >> 2] >> (8 * (j & 3))) & 255u) mask |= 1u << j;
      t.e[3] |= mask << 16;
    }
  }
  r.maxterms = 0;
  for (int i = 0; i < n; ++i)
    r.maxterms = std::max(r.maxterms, sys->first[i + 1] - sys->first[i]);
  r.gamma = mk(hom->gamma[0], hom->gamma[1]);
  for (int i = 0; i < n; ++i) r.b[i] = mk(hom->b[2 * i], hom->b[2 * i + 1]);
  for (int j = 0; j < N; ++j)
    r.patch[j] = mk(hom->patch[2 * j], hom->patch[2 * j + 1]);
  const phc_tracker_options& t = o->tracker;
  const phc_endgame_options& e = o->endgame;
  r.step_init = t.step_init;
  r.step_max = t.step_max;
  r.step_min = t.step_min;
  r.ctol = t.corrector_tol;
  r.contraction = t.contraction;
  r.maxit = t.corrector_maxit;
  r.grow_after = t.grow_after;
  r.max_steps = t.max_steps;
  r.use_endgame = o->use_endgame != 0;
  r.t_start = e.t_start;
  r.rf = e.radius_factor;
  r.rmin = e.radius_min;
  r.loop_radius = e.loop_radius;
  r.agreement = e.winding_agreement;
  r.egtol = e.tol;
  r.rtol = e.residual_tol;
  r.closure = e.closure_factor;
  r.nsamples = e.nsamples;
  r.max_winding = e.max_winding;
  if (r.use_endgame && e.nsamples < 1)
    return fail(ctx, PHC_ERR_INVALID_INPUT, "nsamples < 1");
  if (r.use_endgame && e.nsamples > 4096)
    return fail(ctx, PHC_ERR_CAPACITY, "nsamples %d above 4096", e.nsamples);
  // dtheta = 2 pi / nsamples and log(1/radius_factor) exactly as the
  // reference computes them (src/endgame.f90:350, :225)
  r.dtheta = 2.0 * 3.14159265358979323846 / (double)(e.nsamples > 0 ? e.nsamples : 1);
  r.log_inv_rf = std::log(1.0 / e.radius_factor);
  r.npaths = npaths;
  h.npaths = npaths;
  h.G = lanes_for(ctx, N);
  return PHC_OK;
}

// One derivative job of the grid unit: term k, exponent e of its variable,
// the offset of pw(e-1, j) and the offsets of the term's other factors.
struct GridJob {
  int k, e, base, np, ns;
  int others[8];  // offsets of the other nonzero factors, ascending j
};

// Short-layout record of a job (the layout is described at WTerm).
uint4 short_record(const GridJob& jb) {
  uint4 rec = make_uint4((uint32_t)jb.k | ((uint32_t)jb.e << 20), 0u, 0u, 0u);
  rec.y = (uint32_t)jb.base;
  for (int b2 = 0; b2 < jb.np; ++b2)
    rec.y |= (uint32_t)jb.others[b2] << (8 * (1 + b2));
  // suffix in right-fold order: last factor first
  for (int b2 = 0; b2 < jb.ns; ++b2)
    rec.z |= (uint32_t)jb.others[jb.np + jb.ns - 1 - b2] << (8 * b2);
  return rec;
}

// Run layout of the 16-lane grid (short layout only). Each sum of the
// unit -- Jacobian entry (i, j) with its jobs in term order, or F_i with
// one step per four term values -- is a run of consecutive steps on one
// lane; the runs are dealt longest first, each to the lane with the fewest
// steps so far, so a lane holds one or more whole runs. Record w: bits
// 0..15 the V index of the step's four term values (the four zeros at n vs
// for derivative steps), bits 16..30 the run's sum index (i N + j, or n N + i
// for F_i) and bit 31 on the run's last step, where the lane stores the
// sum. Table [step][lane], the step count even (two steps per turn).
bool build_runs(HostReq& h, int G, std::vector<WTerm>& wt,
                const std::vector<std::vector<GridJob>>& lists, int maxterms,
                int maxnnz, int maxp, int maxs, int fl) {
  Req& r = h.r;
  const int N = r.N, n = r.n, nt = (int)h.terms.size();
  const int LP = std::min(G * G, 64), RG = LP / G, E = G / RG;
  const int vs = 4 * std::max(1, (maxterms + 3) / 4);
  const int zero_row = n * vs;
  if (zero_row + vs > 65535 || n * N + n > std::min(G * G, 1 << 15))
    return false;
  struct Run {
    int id, len, i, j;  // j == -1: F_i
  };
  std::vector<Run> runs;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < N; ++j) {
      const int len = (int)lists[((j % RG) * G + i) * E + j / RG].size();
      runs.push_back({i * N + j, std::max(len, 1), i, j});
    }
    const int nti = r.first[i + 1] - r.first[i];
    runs.push_back({n * N + i, std::max(1, (nti + 3) / 4), i, -1});
  }
  std::stable_sort(runs.begin(), runs.end(),
                   [](const Run& a, const Run& b) { return a.len > b.len; });
  std::vector<int> load(LP, 0);
  std::vector<std::vector<int>> dealt(LP);
  for (int x = 0; x < (int)runs.size(); ++x) {
    int best = 0;
    for (int q = 1; q < LP; ++q)
      if (load[q] < load[best]) best = q;
    dealt[best].push_back(x);
    load[best] += runs[x].len;
  }
  size_t M = (size_t)*std::max_element(load.begin(), load.end());
  M += M & 1;
  for (int i = 0; i < n; ++i)
    for (int k = r.first[i]; k < r.first[i + 1]; ++k)
      wt[k].slot = (uint32_t)(i * vs + (k - r.first[i]));
  wt[nt].slot = (uint32_t)zero_row;  // never written: not a real term
  // padding steps: the zero-coefficient term, factors pw(0, 0) = 1
  const uint4 pad = make_uint4((uint32_t)nt | (1u << 20), 0u, 0u,
                               (uint32_t)zero_row);
  h.jobs.assign(M * LP, pad);
  for (int q = 0; q < LP; ++q) {
    size_t m = 0;
    for (int x : dealt[q]) {
      const Run& ru = runs[x];
      for (int s = 0; s < ru.len; ++s, ++m) {
        uint4 rec = pad;
        if (ru.j < 0) {
          rec.w = (uint32_t)(ru.i * vs + 4 * s);
        } else {
          const auto& l = lists[((ru.j % RG) * G + ru.i) * E + ru.j / RG];
          if (s < (int)l.size()) rec = short_record(l[s]);
          rec.w = (uint32_t)zero_row;
        }
        if (s == ru.len - 1) rec.w |= ((uint32_t)ru.id << 16) | (1u << 31);
        h.jobs[m * LP + q] = rec;
      }
    }
  }
  h.wterms.swap(wt);
  r.njobs = (int32_t)M;
  r.vstride = vs;
  r.vtail = 4;
  r.gfold = maxnnz | (maxp << 8) | (maxs << 16) | (fl << 24) | (1 << 28);
  return true;
}

// Tables of the grid unit (phc_wave.h) for G*G lanes per path: compact
// terms (plus a zero-coefficient term at index nterms) with the LDS offsets
// e G + j of their nonzero factors and their V slots, and per lane (i, j)
// the job records of the terms of equation i with a nonzero exponent of
// z_j in term order, laid out [job][lane] with lane q = j G + i. False when
// the system does not fit the encoding (offsets above 255, 2^20 terms).
bool build_grid(HostReq& h, int G) {
  Req& r = h.r;
  const int N = r.N, n = r.n;
  const int LP = std::min(G * G, 64), RG = LP / G, E = G / RG;
  const int nt = (int)h.terms.size();
  if (N > G || nt + 1 >= (1 << 20)) return false;
  using Job = GridJob;
  std::vector<WTerm> wt(nt + 1);
  memset(wt.data(), 0, wt.size() * sizeof(WTerm));
  std::vector<std::vector<Job>> lists(LP * E);  // slot q E + e
  int maxnnz = 0, maxp = 0, maxs = 0, maxterms = 0;
  for (int i = 0; i < n; ++i) {
    maxterms = std::max(maxterms, r.first[i + 1] - r.first[i]);
    for (int k = r.first[i]; k < r.first[i + 1]; ++k) {
      const TermRec& t = h.terms[k];
      WTerm& w = wt[k];
      w.re = t.re;
      w.im = t.im;
      int var[16], ex[16], off[16], nz = 0;
      for (int j = 0; j < N; ++j) {
        const int e = (int)((t.e[j >> 2] >> (8 * (j & 3))) & 255u);
        if (e == 0) continue;
        const int o = e * G + j;
        if (o > 255) return false;
        var[nz] = j;
        ex[nz] = e;
        off[nz] = o;
        w.off |= (uint64_t)o << (8 * nz);
        ++nz;
      }
      w.nnz = (uint32_t)nz;
      maxnnz = std::max(maxnnz, nz);
      for (int a = 0; a < nz; ++a) {
        Job jb;
        memset(&jb, 0, sizeof(jb));
        jb.k = k;
        jb.e = ex[a];
        jb.base = ex[a] >= 2 ? (ex[a] - 1) * G + var[a] : 0;
        jb.np = a;
        jb.ns = nz - 1 - a;
        int c = 0;
        for (int b2 = 0; b2 < nz; ++b2)
          if (b2 != a) jb.others[c++] = off[b2];
        const int j = var[a];
        lists[((j % RG) * G + i) * E + j / RG].push_back(jb);
        maxp = std::max(maxp, jb.np);
        maxs = std::max(maxs, jb.ns);
      }
    }
  }
  // short layout: every term has <= 4 nonzero exponents
  const int fl = maxnnz <= 4 ? std::max(1, std::max(maxp, maxs)) : 0;
  if (fl && E > 1)
    return build_runs(h, G, wt, lists, maxterms, maxnnz, maxp, maxs, fl);
  size_t M = 0;
  for (const auto& l : lists) M = std::max(M, l.size());
  M = std::max(M, (size_t)((maxterms + 3) / 4));
  if (E == 1) M += M & 1;  // one slot per lane: two rows per step
  const int vstride = (int)(4 * M);
  for (int i = 0; i < n; ++i)
    for (int k = r.first[i]; k < r.first[i + 1]; ++k)
      wt[k].slot = (uint32_t)(i * vstride + (k - r.first[i]));
  wt[nt].slot = (uint32_t)(n * vstride);  // never written: not a real term
  // padding rows: the zero-coefficient term with no other factors (adds
  // +0), except for the general layout of the 16-lane grid, whose loop
  // skips WJOB_NONE rows
  // The general layout with one slot per lane keeps each term's prefix
  // products P_1..P_{m-1} and suffix products S_2..S_m (m nonzero factors)
  // in V after the four zeros and a one: term k's from ps_base[k] on,
  // prefixes first (WTerm::nnz bits 8 and up). A job record then names
  // them (z = V index of P_{q-1} | of S_{q+1} << 16; the one where a
  // product is empty) and the factor pw(e-1, j) (y).
  const bool ps_layout = !fl && E == 1;
  const int one_slot = n * vstride + 4;
  std::vector<int> ps_base(nt + 1, 0);
  int vtail = 4;
  if (ps_layout) {
    int ps = one_slot + 1;
    for (int k = 0; k < nt; ++k) {
      ps_base[k] = ps;
      const int nz = (int)wt[k].nnz;
      wt[k].nnz = (uint32_t)nz | ((uint32_t)ps << 8);
      ps += 2 * std::max(nz - 1, 0);
    }
    if (ps > 65535) return false;  // 16-bit indices in the job records
    vtail = ps - n * vstride;
  }
  const uint4 pad =
      ps_layout ? make_uint4((uint32_t)nt | (1u << 20), 0u,
                             (uint32_t)one_slot | ((uint32_t)one_slot << 16), 0u)
      : fl      ? make_uint4((uint32_t)nt | (1u << 20), 0u, 0u, 0u)
                : make_uint4(WJOB_NONE, 0u, 0u, 0u);
  h.jobs.assign(M * LP * E, pad);
  for (int qq = 0; qq < LP * E; ++qq) {
    for (size_t m = 0; m < lists[qq].size(); ++m) {
      const Job& jb = lists[qq][m];
      uint4 rec = make_uint4((uint32_t)jb.k | ((uint32_t)jb.e << 20), 0u, 0u, 0u);
      if (fl) {
        rec = short_record(jb);
      } else if (ps_layout) {
        // the job of the term's factor a = np (0-based) of nz = np + ns + 1
        const int a = jb.np, nz = jb.np + jb.ns + 1, b = ps_base[jb.k];
        const int pre = a >= 1 ? b + (a - 1) : one_slot;
        const int suf = a <= nz - 2 ? b + (nz - 1) + a : one_slot;
        rec.y = (uint32_t)jb.base;
        rec.z = (uint32_t)pre | ((uint32_t)suf << 16);
      } else {
        uint64_t offs = (uint64_t)jb.base;
        for (int b2 = 0; b2 < jb.np + jb.ns; ++b2)
          offs |= (uint64_t)jb.others[b2] << (8 * (1 + b2));
        rec.y = (uint32_t)(offs & 0xffffffffu);
        rec.z = (uint32_t)(offs >> 32);
        rec.w = (uint32_t)jb.np | ((uint32_t)jb.ns << 4);
      }
      h.jobs[m * LP * E + qq] = rec;
    }
  }
  h.wterms.swap(wt);
  r.njobs = (int32_t)M;
  r.vstride = vstride;
  r.vtail = vtail;
  r.gfold = maxnnz | (maxp << 8) | (maxs << 16) | (fl << 24);
  return true;
}

template <int G, int MODE>
hipError_t launch_track(const Launch& L, bool staged, int blocks,
                        int threads, size_t lds, hipStream_t s) {
  if (staged)
    hipLaunchKernelGGL((track_kernel<G, true, MODE>), dim3(blocks),
                       dim3(threads), lds, s, L);
  else
    hipLaunchKernelGGL((track_kernel<G, false, MODE>), dim3(blocks),
                       dim3(threads), lds, s, L);
  return hipGetLastError();
}

hipError_t launch_g(int G, const Launch& L, bool staged, int blocks,
                    int threads, size_t lds, hipStream_t s,
                    int mode = MODE_ROW) {
  switch (G) {
#ifndef PHC_ONLY_G8
    case 4:
      if (mode == MODE_GRID)
        return launch_track<4, MODE_GRID>(L, staged, blocks, threads, lds, s);
      if (mode == MODE_SPLIT)
        return launch_track<4, MODE_SPLIT>(L, staged, blocks, threads, lds, s);
      return launch_track<4, MODE_ROW>(L, staged, blocks, threads, lds, s);
    case 16:
      if (mode == MODE_GRID)
        return launch_track<16, MODE_GRID>(L, staged, blocks, threads, lds, s);
      return launch_track<16, MODE_ROW>(L, staged, blocks, threads, lds, s);
#endif
    default:
      if (mode == MODE_GRID)
        return launch_track<8, MODE_GRID>(L, staged, blocks, threads, lds, s);
      if (mode == MODE_SPLIT)
        return launch_track<8, MODE_SPLIT>(L, staged, blocks, threads, lds, s);
      return launch_track<8, MODE_ROW>(L, staged, blocks, threads, lds, s);
  }
}

// Largest LDS a staged workgroup may use (leaves room for two per CU).
constexpr size_t STAGED_LDS_LIMIT = 64 * 1024;

// LDS of one path (the device's unit_lds_bytes): its state, then the
// unit's power table and linear system (row and split mappings, plus the
// split mapping's exchange buffers) or the grid unit's tables.
size_t group_lds_bytes(int G, int expmax) {
  return state_lds_bytes(G) +
         ((size_t)(expmax + 1) * G + (size_t)G * G + G) * sizeof(cx) +
         (size_t)G * sizeof(int);
}
size_t split_lds_bytes(int G) {
  const size_t R = 64 / G;
  return (R * G * (G + 1) + (size_t)G * G + G) * sizeof(cx);
}
// Entries of a path's term-value array V for a request's grid tables: its
// n equations at vstride entries each, then the four zeros (and, for the
// one-slot general layout, the one and the prefix and suffix products).
int grid_v_entries(const Req& r) { return r.n * r.vstride + r.vtail; }

size_t grid_lds_bytes(int G, int expmax, int vmax) {
  return state_lds_bytes(G) + wave_lds_entries(G, expmax, vmax) * sizeof(cx);
}

// A prepared multi-request run for one lanes class, with device buffers.
struct ClassRun {
  int G = 8;
  std::vector<HostReq*> reqs;
  Req* d_reqs = nullptr;
  int64_t* d_base = nullptr;
  unsigned long long* d_next = nullptr;
  unsigned long long* d_counters = nullptr;
  unsigned int* d_active = nullptr;
  int64_t total = 0;
  int expmax = 1;
  int blocks = 0, threads = 0;
  size_t lds = 0;
  bool staged = false;
  int mode = MODE_ROW;
  int lp = 8;     // lanes per path
  int vmax = 0;   // grid unit: largest term count
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

}  // namespace

struct phc_staged {
  Arena arena;
  std::vector<std::unique_ptr<HostReq>> reqs;
  std::vector<ClassRun> classes;
  int32_t max_units = 0;  // units per path slot and launch (0: unbounded)
  ~phc_staged() {
    for (ClassRun& cr : classes) {
      if (cr.ev0) (void)hipEventDestroy(cr.ev0);
      if (cr.ev1) (void)hipEventDestroy(cr.ev1);
    }
  }
};

namespace {

void* arena_take(Arena& a, size_t bytes) {
  void* p = a.base + a.used;
  a.used += Arena::al(bytes);
  return p;
}

int64_t binade(int64_t v) {
  int64_t b = 0;
  while ((int64_t(1) << (b + 1)) <= std::max<int64_t>(v, 1)) ++b;
  return b;
}

// Key of the autotuner's cache. A class of one request: lanes class,
// unknowns, terms and the binade of the path count. A class of several
// requests (a batch): lanes class, minus the request count, the terms of
// all its requests and the binade of their path count.
std::array<int64_t, 4> tune_key(const HostReq& h) {
  return {h.G, h.r.N, h.r.nterms, binade(h.npaths)};
}
std::array<int64_t, 4> tune_key(const std::vector<HostReq*>& reqs) {
  if (reqs.size() == 1) return tune_key(*reqs[0]);
  int64_t paths = 0, terms = 0;
  for (const HostReq* h : reqs) {
    paths += h->npaths;
    terms += h->r.nterms;
  }
  return {reqs[0]->G, -(int64_t)reqs.size(), terms, binade(paths)};
}

// Lane mapping, launch shape and LDS of one lanes class. A class takes the
// autotuner's cached choice for its shape when there is one (tune_key:
// single requests and batch classes alike). Otherwise (mapping 0): the
// grid unit when every request fits its tables, else one path per
// wavefront with split terms when one path per group would leave most
// SIMDs idle, else one path per group. Mappings 1-3 force row, split or
// grid where the class allows it.
void plan_class(phc_ctx* ctx, ClassRun& cr) {
  const int G = cr.G;
  int mapping = ctx->mapping, wg_threads = ctx->wg_threads,
      staging = ctx->staging;
  if (mapping == 0) {
    auto it = ctx->tuned.find(tune_key(cr.reqs));
    if (it != ctx->tuned.end()) {
      mapping = it->second[0];
      wg_threads = it->second[1];
      staging = it->second[2];
    }
  }
  const int64_t cus = std::max(ctx->num_cus, 1);
  cr.total = 0;
  cr.expmax = 1;
  cr.vmax = 0;
  for (HostReq* h : cr.reqs) {
    cr.total += h->npaths;
    cr.expmax = std::max(cr.expmax, h->r.maxexp);
  }
  bool grid = mapping == 0 || mapping == 3;
  for (HostReq* h : cr.reqs)
    if (grid) grid = build_grid(*h, G);
  if (grid)
    for (HostReq* h : cr.reqs)
      cr.vmax = std::max(cr.vmax, grid_v_entries(h->r));
  bool can_split = G <= 8;
  for (HostReq* h : cr.reqs) can_split = can_split && h->r.N <= 64 / G;
  const int64_t row_waves = (std::max<int64_t>(cr.total, 1) * G + 63) / 64;
  const int64_t paths_needed = std::max<int64_t>(cr.total, 1);
  const size_t limit = std::min<size_t>(ctx->lds_limit, 160 * 1024);
  cr.blocks = 0;
  for (int attempt = 0; attempt < 3; ++attempt) {
    if (grid)
      cr.mode = MODE_GRID;
    else if (can_split && (mapping == 2 ||
                           (mapping == 0 && row_waves < 2 * 4 * cus)))
      cr.mode = MODE_SPLIT;
    else
      cr.mode = MODE_ROW;
    cr.lp = cr.mode == MODE_ROW ? G
                                : (cr.mode == MODE_SPLIT ? 64 : std::min(G * G, 64));
    const size_t per_path =
        cr.mode == MODE_GRID
            ? grid_lds_bytes(G, cr.expmax, cr.vmax)
            : group_lds_bytes(G, cr.expmax) +
                  (cr.mode == MODE_SPLIT ? split_lds_bytes(G) : 0);
    size_t prog = 0;
    if (cr.reqs.size() == 1) {
      prog = staged_req_bytes() + cr.reqs[0]->terms.size() * sizeof(TermRec);
      if (cr.mode == MODE_GRID)
        prog += sizeof(WTerm) + cr.reqs[0]->jobs.size() * sizeof(uint4);
    }
    int threads = wg_threads ? wg_threads : WG;
    threads = std::max(threads, cr.lp);
    if (!wg_threads) {
      while (threads > 64 && threads > cr.lp &&
             (paths_needed + threads / cr.lp - 1) / (threads / cr.lp) < 2 * cus)
        threads /= 2;
    }
    // every path keeps its power table (and, on the grid, its term values)
    // in LDS: shrink the workgroup until its paths fit; if one path does
    // not, give up the grid or split mapping for the row mapping
    while (threads > cr.lp && (size_t)(threads / cr.lp) * per_path > limit)
      threads /= 2;
    if ((size_t)(threads / cr.lp) * per_path > limit) {
      if (cr.mode == MODE_GRID) {
        grid = false;
        continue;
      }
      if (cr.mode == MODE_SPLIT) {
        can_split = false;
        continue;
      }
      cr.blocks = 0;  // no mapping holds one path: the caller reports it
      return;
    }
    const int per_wg = threads / cr.lp;
    int64_t blocks = (paths_needed + per_wg - 1) / per_wg;
    blocks = std::min<int64_t>(blocks, cus * 16 * (WG / threads));
    cr.blocks = (int)std::max<int64_t>(blocks, 1);
    cr.threads = threads;
    cr.lds = (size_t)per_wg * per_path;
    const size_t stage_limit =
        cr.mode == MODE_GRID ? limit : STAGED_LDS_LIMIT;
    cr.staged = staging != 0 && prog > 0 && cr.lds + prog <= stage_limit;
    if (cr.staged) cr.lds += prog;
    break;
  }
  if (cr.mode != MODE_GRID)
    for (HostReq* h : cr.reqs) {
      h->wterms.clear();
      h->jobs.clear();
      h->r.njobs = 0;
    }
}

// Allocates device memory for all requests, uploads inputs, prepares one
// ClassRun per lanes class.
int stage_requests(phc_ctx* ctx, phc_staged* st) {
  std::map<int, std::vector<HostReq*>> byG;
  for (auto& up : st->reqs) byG[up->G].push_back(up.get());
  std::vector<ClassRun> plans;
  size_t need = 4096;
  for (auto& kv : byG) {
    ClassRun cr;
    cr.G = kv.first;
    cr.reqs = kv.second;
    plan_class(ctx, cr);
    if (cr.blocks == 0)
      return fail(ctx, PHC_ERR_CAPACITY,
                  "one path's working set (power table of exponent %d, %d "
                  "lanes) exceeds the LDS of a workgroup", cr.expmax, cr.G);
    for (HostReq* h : cr.reqs) {
      const int N = h->r.N;
      need += Arena::al(h->terms.size() * sizeof(TermRec));
      need += Arena::al(h->wterms.size() * sizeof(WTerm));
      need += Arena::al(h->jobs.size() * sizeof(uint4));
      need += Arena::al((size_t)h->npaths * N * sizeof(cx)) * 2;  // start, z
      need += Arena::al((size_t)h->npaths * 4) * 8;
      need += Arena::al((size_t)h->npaths * 8) * 3;
      need += 256;
    }
    need += Arena::al(cr.reqs.size() * sizeof(Req));
    need += Arena::al((cr.reqs.size() + 1) * sizeof(int64_t));
    need += 256 * 3;
    plans.push_back(cr);
  }
  HIPCHK(ctx, hipMalloc(&st->arena.base, need));
  st->arena.size = need;

  for (ClassRun& cr : plans) {
    std::vector<Req> host_reqs;
    std::vector<int64_t> base(1, 0);
    for (HostReq* h : cr.reqs) {
      const int N = h->r.N;
      TermRec* dt = (TermRec*)arena_take(st->arena, h->terms.size() * sizeof(TermRec));
      HIPCHK(ctx, hipMemcpyAsync(dt, h->terms.data(),
                                 h->terms.size() * sizeof(TermRec),
                                 hipMemcpyHostToDevice, ctx->stream));
      h->r.terms = dt;
      h->r.wterms = nullptr;
      h->r.jobs = nullptr;
      if (cr.mode == MODE_GRID) {
        WTerm* dw = (WTerm*)arena_take(st->arena, h->wterms.size() * sizeof(WTerm));
        HIPCHK(ctx, hipMemcpyAsync(dw, h->wterms.data(),
                                   h->wterms.size() * sizeof(WTerm),
                                   hipMemcpyHostToDevice, ctx->stream));
        h->r.wterms = dw;
        if (!h->jobs.empty()) {
          uint4* dj = (uint4*)arena_take(st->arena, h->jobs.size() * sizeof(uint4));
          HIPCHK(ctx, hipMemcpyAsync(dj, h->jobs.data(),
                                     h->jobs.size() * sizeof(uint4),
                                     hipMemcpyHostToDevice, ctx->stream));
          h->r.jobs = dj;
        }
      }
      cx* ds = (cx*)arena_take(st->arena, (size_t)h->npaths * N * sizeof(cx));
      if (h->npaths)
        HIPCHK(ctx, hipMemcpyAsync(ds, h->start,
                                   (size_t)h->npaths * N * sizeof(cx),
                                   hipMemcpyHostToDevice, ctx->stream));
      h->r.start = ds;
      h->r.o_z = (cx*)arena_take(st->arena, (size_t)h->npaths * N * sizeof(cx));
      h->r.o_status = (int32_t*)arena_take(st->arena, (size_t)h->npaths * 4);
      h->r.o_wind = (int32_t*)arena_take(st->arena, (size_t)h->npaths * 4);
      h->r.o_egs = (int32_t*)arena_take(st->arena, (size_t)h->npaths * 4);
      h->r.o_acc = (int32_t*)arena_take(st->arena, (size_t)h->npaths * 4);
      h->r.o_rej = (int32_t*)arena_take(st->arena, (size_t)h->npaths * 4);
      h->r.o_nlin = (int32_t*)arena_take(st->arena, (size_t)h->npaths * 4);
      h->r.o_nbe = (int32_t*)arena_take(st->arena, (size_t)h->npaths * 4);
      h->r.o_done = (uint32_t*)arena_take(st->arena, 4);
      h->r.o_t = (double*)arena_take(st->arena, (size_t)h->npaths * 8);
      h->r.o_egerr = (double*)arena_take(st->arena, (size_t)h->npaths * 8);
      h->r.o_egres = (double*)arena_take(st->arena, (size_t)h->npaths * 8);
      HIPCHK(ctx, hipMemsetAsync(h->r.o_done, 0, 4, ctx->stream));
      host_reqs.push_back(h->r);
      base.push_back(base.back() + h->npaths);
    }
    cr.d_reqs = (Req*)arena_take(st->arena, host_reqs.size() * sizeof(Req));
    HIPCHK(ctx, hipMemcpyAsync(cr.d_reqs, host_reqs.data(),
                               host_reqs.size() * sizeof(Req),
                               hipMemcpyHostToDevice, ctx->stream));
    cr.d_base = (int64_t*)arena_take(st->arena, base.size() * sizeof(int64_t));
    HIPCHK(ctx, hipMemcpyAsync(cr.d_base, base.data(),
                               base.size() * sizeof(int64_t),
                               hipMemcpyHostToDevice, ctx->stream));
    cr.d_next = (unsigned long long*)arena_take(st->arena, 8);
    cr.d_counters = (unsigned long long*)arena_take(st->arena, 32);
    cr.d_active = (unsigned int*)arena_take(st->arena, 4);
    st->classes.push_back(cr);
  }
  return PHC_OK;
}

// Clears the queues on the context stream, then launches every lanes class
// at the same point of that stream: class 0 on the context stream, class i
// on side stream i-1, which the context stream waits for afterwards. A long
// path of one class then runs beside the other classes' work instead of
// after it. Each class's ev0 marks the common start, its ev1 its end.
int run_classes(phc_ctx* ctx, phc_staged* st) {
  const size_t nc = st->classes.size();
  while (ctx->side_streams.size() + 1 < nc) {
    hipStream_t s = nullptr;
    HIPCHK(ctx, hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    ctx->side_streams.push_back(s);
  }
  for (ClassRun& cr : st->classes) {
    HIPCHK(ctx, hipMemsetAsync(cr.d_next, 0, 8, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(cr.d_counters, 0, 32, ctx->stream));
    HIPCHK(ctx, hipMemsetAsync(cr.d_active, 0, 4, ctx->stream));
    for (HostReq* h : cr.reqs)
      HIPCHK(ctx, hipMemsetAsync(h->r.o_done, 0, 4, ctx->stream));
    if (!cr.ev0) HIPCHK(ctx, hipEventCreate(&cr.ev0));
    if (!cr.ev1) HIPCHK(ctx, hipEventCreate(&cr.ev1));
  }
  ctx->last_plan.clear();
  for (ClassRun& cr : st->classes)
    ctx->last_plan.push_back({cr.G, cr.blocks, cr.threads, (int32_t)cr.lds,
                              cr.staged ? 1 : 0, cr.mode, cr.lp,
                              (int32_t)cr.reqs.size()});
  for (ClassRun& cr : st->classes)
    HIPCHK(ctx, hipEventRecord(cr.ev0, ctx->stream));
  for (size_t i = 0; i < nc; ++i) {
    ClassRun& cr = st->classes[i];
    hipStream_t s = i == 0 ? ctx->stream : ctx->side_streams[i - 1];
    if (i) HIPCHK(ctx, hipStreamWaitEvent(s, cr.ev0, 0));
    Launch L;
    memset(&L, 0, sizeof(L));
    L.reqs = cr.d_reqs;
    L.req_base = cr.d_base;
    L.nreq = (int)cr.reqs.size();
    L.expmax = cr.expmax;
    L.vmax = cr.vmax;
    L.next = cr.d_next;
    L.total = cr.total;
    L.slots = nullptr;
    L.resume = 0;
    L.max_units = st->max_units;
    L.active = cr.d_active;
    L.counters = cr.d_counters;
    HIPCHK(ctx, launch_g(cr.G, L, cr.staged, cr.blocks, cr.threads, cr.lds,
                         s, cr.mode));
    HIPCHK(ctx, hipEventRecord(cr.ev1, s));
  }
  for (size_t i = 1; i < nc; ++i)
    HIPCHK(ctx, hipStreamWaitEvent(ctx->stream, st->classes[i].ev1, 0));
  return PHC_OK;
}

template <typename T>
int fetch_array(phc_ctx* ctx, T* host, const T* dev, size_t count) {
  if (!host || count == 0) return PHC_OK;
  HIPCHK(ctx, hipMemcpyAsync(host, dev, count * sizeof(T),
                             hipMemcpyDeviceToHost, ctx->stream));
  return PHC_OK;
}

int fetch_outputs(phc_ctx* ctx, phc_staged* st) {
  for (int i = 0; i < 4; ++i) ctx->counters[i] = 0;
  ctx->reruns = 0;
  for (ClassRun& cr : st->classes) {
    unsigned long long c[4];
    HIPCHK(ctx, hipMemcpyAsync(c, cr.d_counters, sizeof(c),
                               hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    ctx->reruns += c[0];       // lane 0 of a group adds
    ctx->counters[1] += c[1];  // lane 0 of a group adds
    ctx->counters[2] += c[2];  // lane 0 of a wavefront adds
    ctx->counters[3] = std::max<unsigned long long>(ctx->counters[3], c[3]);
    ctx->counters[0] += (unsigned long long)cr.blocks * (cr.threads / cr.lp);
  }
  for (auto& up : st->reqs) {
    HostReq& h = *up;
    phc_endpoints* o = h.out;
    if (!o) continue;
    const size_t np = (size_t)h.npaths;
    int rc = 0;
    rc |= fetch_array(ctx, o->status, h.r.o_status, np);
    rc |= fetch_array(ctx, o->t, h.r.o_t, np);
    rc |= fetch_array(ctx, o->z, (const double*)h.r.o_z, np * h.r.N * 2);
    rc |= fetch_array(ctx, o->winding, h.r.o_wind, np);
    rc |= fetch_array(ctx, o->eg_status, h.r.o_egs, np);
    rc |= fetch_array(ctx, o->eg_error, h.r.o_egerr, np);
    rc |= fetch_array(ctx, o->eg_residual, h.r.o_egres, np);
    rc |= fetch_array(ctx, o->naccepted, h.r.o_acc, np);
    rc |= fetch_array(ctx, o->nrejected, h.r.o_rej, np);
    rc |= fetch_array(ctx, o->nlin, h.r.o_nlin, np);
    rc |= fetch_array(ctx, o->nbe, h.r.o_nbe, np);
    if (rc) return PHC_ERR_DEVICE;
  }
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return PHC_OK;
}

// ------------------------------------------------- allocation-free scratch
// Device memory of phc_track_paths_noalloc beyond its caller's inputs and
// outputs: the request record, the term and grid tables, the work-queue
// head, the path-start offsets and the unit counters of one launch. Every
// GPU the library runs on holds one copy of this array until the process
// ends. Host code never passes `scratch` itself to a kernel or a copy (in
// host code the name is a host-side stand-in, not a device address): the
// device address comes from hipGetSymbolAddress, once per device, at the
// first context created on it.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)1 << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;                          // guards g_noalloc_ptr
void* g_noalloc_ptr[NOALLOC_MAX_DEVICES] = {nullptr};
std::mutex g_noalloc_run[NOALLOC_MAX_DEVICES];    // one call per device

// Looks up this device's copy of `scratch` (the device must be current).
int noalloc_setup(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return PHC_ERR_DEVICE;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return PHC_OK;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr)
    return fail(nullptr, PHC_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s",
                hipGetErrorString(e));
  g_noalloc_ptr[device] = ptr;
  return PHC_OK;
}

void* noalloc_ptr(int device) {
  if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  return g_noalloc_ptr[device];
}

// Clears `words` 32-bit words (a kernel, not hipMemsetAsync, so that the
// clearing is an ordinary kernel node if a caller captures the stream).
__global__ void zero_words(uint32_t* p, int words) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < words) p[i] = 0u;
}

}  // namespace

// ===================================================================== C ABI
extern "C" {

void phc_default_options(phc_solve_options* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->tracker.step_init = 0.05;
  o->tracker.step_max = 0.1;
  o->tracker.step_min = 1.0e-14;
  o->tracker.corrector_tol = 1.0e-10;
  o->tracker.contraction = 0.5;
  o->tracker.corrector_maxit = 3;
  o->tracker.grow_after = 5;
  o->tracker.max_steps = 20000;
  o->endgame.t_start = 0.1;
  o->endgame.radius_factor = 0.25;
  o->endgame.radius_min = 1.0e-13;
  o->endgame.loop_radius = 1.0e-5;
  o->endgame.winding_agreement = 0.2;
  o->endgame.tol = 1.0e-11;
  o->endgame.residual_tol = 1.0e-8;
  o->endgame.closure_factor = 0.1;
  o->endgame.nsamples = 8;
  o->endgame.max_winding = 16;
  o->infinity_tol = 1.0e-8;
  o->singular_cond = 1.0e10;
  o->real_tol = 1.0e-8;
  o->dedup_tol = 1.0e-8;
  o->singular_dedup_tol = 1.0e-6;
  o->use_endgame = 1;
  o->seed = 1;
  o->refine_maxit = 5;
}

int phc_ctx_create(int device, phc_ctx** out) {
  if (!out) return PHC_ERR_INVALID_INPUT;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return PHC_ERR_DEVICE;
  if (device < 0 || device >= ndev) return PHC_ERR_DEVICE;
  std::unique_ptr<phc_ctx> ctx(new phc_ctx);
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess) return PHC_ERR_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess)
    return PHC_ERR_DEVICE;
  ctx->num_cus = prop.multiProcessorCount;
  ctx->lds_limit = prop.sharedMemPerBlock;
  {
    int khz = 0;
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate,
                              device) == hipSuccess && khz > 0)
      ctx->wall_khz = khz;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) !=
      hipSuccess)
    return PHC_ERR_DEVICE;
  // this device's scratch of phc_track_paths_noalloc; a failure here only
  // disables that entry point (it reports the error when called)
  (void)noalloc_setup(device);
  *out = ctx.release();
  return PHC_OK;
}

int phc_ctx_destroy(phc_ctx* ctx) {
  if (!ctx) return PHC_OK;
  (void)hipSetDevice(ctx->device);
  for (hipStream_t s : ctx->side_streams) (void)hipStreamDestroy(s);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return PHC_OK;
}

int phc_ctx_lanes_per_path(const phc_ctx* ctx, int N) {
  return lanes_for(ctx, N);
}

const char* phc_ctx_last_error(const phc_ctx* ctx) {
  return ctx ? ctx->err : "no context";
}

const char* phc_thread_last_error(void) { return t_err; }

int phc_check_system(const phc_system* sys, const phc_solve_options* opts) {
  phc_homotopy hom;
  static const double zero[2 * PHC_MAX_UNKNOWNS] = {0.0};
  hom.gamma = zero;
  hom.b = zero;
  hom.patch = zero;
  HostReq h;
  return validate_and_pack(nullptr, sys, &hom, opts, 0, h);
}

int phc_ctx_shared(int device, phc_ctx** out) {
  if (!out) return PHC_ERR_INVALID_INPUT;
  *out = nullptr;
  if (device < 0 || device >= 64)
    return fail(nullptr, PHC_ERR_DEVICE, "device %d out of range", device);
  static std::mutex mu;
  static phc_ctx* shared[64] = {nullptr};
  std::lock_guard<std::mutex> lk(mu);
  if (!shared[device]) {
    const int rc = phc_ctx_create(device, &shared[device]);
    if (rc != PHC_OK) {
      shared[device] = nullptr;
      return fail(nullptr, rc, "no usable HIP device %d", device);
    }
  }
  *out = shared[device];
  return PHC_OK;
}

static std::atomic<int> g_default_device{0};
void phc_set_default_device(int device) { g_default_device.store(device); }
int phc_get_default_device(void) { return g_default_device.load(); }

int phc_ctx_counters(const phc_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return PHC_ERR_INVALID_INPUT;
  for (int i = 0; i < 4; ++i) out[i] = ctx->counters[i];
  return PHC_OK;
}

int phc_ctx_exact_reruns(const phc_ctx* ctx, uint64_t* out) {
  if (!ctx || !out) return PHC_ERR_INVALID_INPUT;
  *out = ctx->reruns;
  return PHC_OK;
}

int phc_ctx_set_mapping(phc_ctx* ctx, int mapping, int wg_threads) {
  if (!ctx || mapping < 0 || mapping > 3) return PHC_ERR_INVALID_INPUT;
  if (wg_threads != 0 && wg_threads != 64 && wg_threads != 128 &&
      wg_threads != 256)
    return PHC_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->mapping = mapping;
  ctx->wg_threads = wg_threads;
  return PHC_OK;
}

int phc_ctx_sync(phc_ctx* ctx) {
  if (!ctx) return PHC_ERR_INVALID_INPUT;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  return PHC_OK;
}

int phc_stage_create(phc_ctx* ctx, const phc_system* sys,
                     const phc_homotopy* hom, const phc_solve_options* opts,
                     int64_t npaths, const double* start, phc_staged** out) {
  if (!ctx || !out) return PHC_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<phc_staged> st(new phc_staged);
  std::unique_ptr<HostReq> h(new HostReq);
  int rc = validate_and_pack(ctx, sys, hom, opts, npaths, *h);
  if (rc) return rc;
  h->start = start;
  st->reqs.push_back(std::move(h));
  rc = stage_requests(ctx, st.get());
  if (rc) return rc;
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  *out = st.release();
  return PHC_OK;
}

int phc_stage_run(phc_ctx* ctx, phc_staged* st) {
  if (!ctx || !st) return PHC_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  return run_classes(ctx, st);
}

int phc_stage_fetch(phc_ctx* ctx, phc_staged* st, phc_endpoints* out) {
  if (!ctx || !st || st->reqs.empty()) return PHC_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  st->reqs[0]->out = out;
  return fetch_outputs(ctx, st);
}

int phc_stage_elapsed(phc_ctx* ctx, phc_staged* st, double* ms) {
  if (!ctx || !st || !ms) return PHC_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  // the classes start together (ev0), so the run takes the longest of them
  double span = 0.0;
  for (ClassRun& cr : st->classes) {
    if (!cr.ev0 || !cr.ev1) continue;
    HIPCHK(ctx, hipEventSynchronize(cr.ev1));
    float t = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&t, cr.ev0, cr.ev1));
    span = std::max(span, (double)t);
  }
  *ms = span;
  return PHC_OK;
}

int phc_stage_launch_info(phc_ctx* ctx, phc_staged* st, int32_t* info) {
  if (!ctx || !st || !info) return PHC_ERR_INVALID_INPUT;
  if (st->classes.empty()) return PHC_ERR_INVALID_INPUT;
  const ClassRun& cr = st->classes[0];
  info[0] = cr.G;
  info[1] = cr.blocks;
  info[2] = cr.threads;
  info[3] = (int32_t)cr.lds;
  info[4] = cr.staged ? 1 : 0;
  info[5] = cr.mode;
  info[6] = cr.lp;
  return PHC_OK;
}

int phc_stage_destroy(phc_ctx* ctx, phc_staged* st) {
  if (ctx) (void)hipSetDevice(ctx->device);
  delete st;
  return PHC_OK;
}

int phc_ctx_autotune(phc_ctx* ctx, const phc_system* sys,
                     const phc_homotopy* hom, const phc_solve_options* opts,
                     int64_t npaths, const double* start, int32_t* choice,
                     double* best_ms) {
  if (!ctx || !choice) return PHC_ERR_INVALID_INPUT;
  if (npaths <= 0 || !start)
    return fail(ctx, PHC_ERR_INVALID_INPUT, "autotune needs start points");
  HostReq probe;
  int rc = validate_and_pack(ctx, sys, hom, opts, npaths, probe);
  if (rc) return rc;
  const int G = probe.G, N = probe.r.N;
  const int64_t sample = std::min<int64_t>(npaths, 4096);
  int saved[3];
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    saved[0] = ctx->mapping;
    saved[1] = ctx->wg_threads;
    saved[2] = ctx->staging;
  }
  struct Cand {
    int mapping, wg, staging;
  };
  std::vector<Cand> cands;
  const int maps[3] = {3, 1, 2};
  for (int m : maps) {
    if (m == 2 && !(G <= 8 && N <= 64 / G)) continue;
    for (int wg : {64, 128, 256})
      for (int stg : {-1, 0}) cands.push_back({m, wg, stg});
  }
  std::vector<std::array<int, 4>> seen;  // actual mode, threads, staged
  double best = 1e300;
  std::array<int, 3> pick = {0, 0, -1};
  int measured = 0;
  for (const Cand& c : cands) {
    {
      std::lock_guard<std::mutex> lk(ctx->mu);
      ctx->mapping = c.mapping;
      ctx->wg_threads = c.wg;
      ctx->staging = c.staging;
    }
    phc_staged* st = nullptr;
    if (phc_stage_create(ctx, sys, hom, opts, sample, start, &st) != PHC_OK)
      continue;
    int32_t info[8] = {0};
    phc_stage_launch_info(ctx, st, info);
    const std::array<int, 4> id = {info[5], info[2], info[4], info[1]};
    bool dup = false;
    for (const auto& x : seen) dup = dup || x == id;
    double ms = 0.0;
    if (!dup) {
      seen.push_back(id);
      if (phc_stage_run(ctx, st) == PHC_OK &&          // code load, caches
          phc_stage_run(ctx, st) == PHC_OK &&
          phc_stage_elapsed(ctx, st, &ms) == PHC_OK) {
        ++measured;
        if (ms < best) {
          best = ms;
          const int mapping = info[5] == MODE_GRID ? 3 : (info[5] == MODE_SPLIT ? 2 : 1);
          pick = {mapping, info[2], info[4] ? -1 : 0};
        }
      }
    }
    phc_stage_destroy(ctx, st);
  }
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->mapping = saved[0];
  ctx->wg_threads = saved[1];
  ctx->staging = saved[2];
  if (!measured) return fail(ctx, PHC_ERR_DEVICE, "no launch shape could be timed");
  ctx->tuned[tune_key(probe)] = pick;
  choice[0] = pick[0];
  choice[1] = pick[1];
  choice[2] = pick[2] == 0 ? 0 : 1;
  if (best_ms) *best_ms = best;
  return PHC_OK;
}

int phc_ctx_autotune_batch(phc_ctx* ctx, int32_t nsys, const phc_system* sys,
                           const phc_homotopy* hom,
                           const phc_solve_options* opts,
                           const int64_t* npaths, const double* const* start,
                           int32_t* choices, int32_t* nclasses) {
  if (!ctx || !choices || !nclasses) return PHC_ERR_INVALID_INPUT;
  *nclasses = 0;
  if (nsys < 0 || (nsys > 0 && (!sys || !hom || !opts || !npaths || !start)))
    return fail(ctx, PHC_ERR_INVALID_INPUT, "null batch input");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  // the batch's lanes classes with their full path counts (the cache key)
  std::vector<std::unique_ptr<HostReq>> full;
  std::map<int, std::vector<int>> byG;  // G -> systems
  for (int s = 0; s < nsys; ++s) {
    if (npaths[s] == 0) continue;
    if (!start[s]) return fail(ctx, PHC_ERR_INVALID_INPUT, "autotune needs start points");
    std::unique_ptr<HostReq> h(new HostReq);
    int rc = validate_and_pack(ctx, &sys[s], &hom[s], &opts[s], npaths[s], *h);
    if (rc) return rc;
    byG[h->G].push_back(s);
    full.push_back(std::move(h));
  }
  struct Restore {  // the context's forced choices, whatever happens below
    phc_ctx* c;
    int m, w, st;
    ~Restore() {
      c->mapping = m;
      c->wg_threads = w;
      c->staging = st;
    }
  } restore{ctx, ctx->mapping, ctx->wg_threads, ctx->staging};
  int nc = 0;
  for (auto& kv : byG) {
    const int G = kv.first;
    std::vector<HostReq*> members;
    for (auto& up : full)
      if (up->G == G) members.push_back(up.get());
    const std::array<int64_t, 4> key = tune_key(members);
    // the sample: the first paths of every system of the class, about
    // 16 k paths in all
    const int64_t cap =
        std::max<int64_t>(16, 16384 / (int64_t)kv.second.size());
    bool can_split = G <= 8;
    for (HostReq* h : members) can_split = can_split && h->r.N <= 64 / G;
    std::vector<std::array<int, 4>> seen;  // mode, threads, staged, blocks
    double best = -1.0;
    std::array<int, 3> pick = {0, 0, -1};
    for (int m : {3, 1, 2}) {
      if (m == 2 && !can_split) continue;
      for (int wg : {64, 128, 256})
        for (int stg : {-1, 0}) {
          ctx->mapping = m;
          ctx->wg_threads = wg;
          ctx->staging = stg;
          std::unique_ptr<phc_staged> st(new phc_staged);
          bool packed = true;
          for (int s : kv.second) {
            std::unique_ptr<HostReq> h(new HostReq);
            if (validate_and_pack(ctx, &sys[s], &hom[s], &opts[s],
                                  std::min<int64_t>(npaths[s], cap), *h)) {
              packed = false;
              break;
            }
            h->start = start[s];
            st->reqs.push_back(std::move(h));
          }
          if (!packed || stage_requests(ctx, st.get()) != PHC_OK ||
              st->classes.size() != 1)
            continue;
          const ClassRun& cr = st->classes[0];
          const std::array<int, 4> id = {cr.mode, cr.threads, cr.staged ? 1 : 0,
                                         cr.blocks};
          bool dup = false;
          for (const auto& x : seen) dup = dup || x == id;
          if (dup) continue;
          seen.push_back(id);
          // bounded launches: every path slot runs at most this many units,
          // so the score is the class's unit throughput, not its longest path
          st->max_units = 64;
          unsigned long long c[4] = {0, 0, 0, 0};
          float ms = 0.f;
          if (run_classes(ctx, st.get()) != PHC_OK ||  // code load, caches
              hipStreamSynchronize(ctx->stream) != hipSuccess ||
              run_classes(ctx, st.get()) != PHC_OK ||
              hipMemcpyAsync(c, st->classes[0].d_counters, sizeof(c),
                             hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
              hipStreamSynchronize(ctx->stream) != hipSuccess ||
              hipEventElapsedTime(&ms, st->classes[0].ev0,
                                  st->classes[0].ev1) != hipSuccess)
            continue;
          const double score = (double)c[1] / std::max((double)ms, 1e-6);
          if (score > best) {
            best = score;
            const int mapping =
                cr.mode == MODE_GRID ? 3 : (cr.mode == MODE_SPLIT ? 2 : 1);
            pick = {mapping, cr.threads, cr.staged ? -1 : 0};
          }
        }
    }
    if (best < 0.0)
      return fail(ctx, PHC_ERR_DEVICE,
                  "no launch shape of lanes class %d could be timed", G);
    ctx->tuned[key] = pick;
    choices[4 * nc + 0] = G;
    choices[4 * nc + 1] = pick[0];
    choices[4 * nc + 2] = pick[1];
    choices[4 * nc + 3] = pick[2] == 0 ? 0 : 1;
    ++nc;
  }
  *nclasses = nc;
  return PHC_OK;
}

int phc_ctx_last_plan(const phc_ctx* ctx, int32_t max_classes, int32_t* info) {
  if (!ctx || max_classes < 0 || (max_classes > 0 && !info))
    return PHC_ERR_INVALID_INPUT;
  const int n = (int)ctx->last_plan.size();
  for (int c = 0; c < n && c < max_classes; ++c)
    for (int k = 0; k < 8; ++k) info[8 * c + k] = ctx->last_plan[c][k];
  return n;
}

int phc_ctx_clear_tuning(phc_ctx* ctx) {
  if (!ctx) return PHC_ERR_INVALID_INPUT;
  std::lock_guard<std::mutex> lk(ctx->mu);
  ctx->tuned.clear();
  return PHC_OK;
}

int phc_track_paths(phc_ctx* ctx, const phc_system* sys,
                    const phc_homotopy* hom, const phc_solve_options* opts,
                    int64_t npaths, const double* start, phc_endpoints* out) {
  if (!ctx) return PHC_ERR_DEVICE;
  if (npaths == 0) return PHC_OK;
  if (!start || !out) return fail(ctx, PHC_ERR_INVALID_INPUT, "null output");
  phc_staged* st = nullptr;
  int rc = phc_stage_create(ctx, sys, hom, opts, npaths, start, &st);
  if (rc) return rc;
  rc = phc_stage_run(ctx, st);
  if (!rc) rc = phc_stage_fetch(ctx, st, out);
  phc_stage_destroy(ctx, st);
  return rc;
}

int phc_track_batch(phc_ctx* ctx, int32_t nsys, const phc_system* sys,
                    const phc_homotopy* hom, const phc_solve_options* opts,
                    const int64_t* npaths, const double* const* start,
                    phc_endpoints* out) {
  if (!ctx) return PHC_ERR_DEVICE;
  if (nsys < 0 || (nsys > 0 && (!sys || !hom || !opts || !npaths || !start || !out)))
    return fail(ctx, PHC_ERR_INVALID_INPUT, "null batch input");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  std::unique_ptr<phc_staged> st(new phc_staged);
  for (int s = 0; s < nsys; ++s) {
    if (npaths[s] == 0) continue;
    std::unique_ptr<HostReq> h(new HostReq);
    int rc = validate_and_pack(ctx, &sys[s], &hom[s], &opts[s], npaths[s], *h);
    if (rc) return rc;
    h->start = start[s];
    h->out = &out[s];
    st->reqs.push_back(std::move(h));
  }
  if (st->reqs.empty()) return PHC_OK;
  int rc = stage_requests(ctx, st.get());
  if (rc) return rc;
  rc = run_classes(ctx, st.get());
  if (rc) return rc;
  return fetch_outputs(ctx, st.get());
}

int phc_track_paths_noalloc(phc_ctx* ctx, const phc_system* sys,
                            const phc_homotopy* hom,
                            const phc_solve_options* opts, int64_t npaths,
                            const double* d_start, phc_endpoints* d_out) {
  if (!ctx) return PHC_ERR_DEVICE;
  if (npaths == 0) return PHC_OK;
  if (!d_start || !d_out || !d_out->status || !d_out->t || !d_out->z ||
      !d_out->winding || !d_out->eg_status || !d_out->eg_error ||
      !d_out->naccepted || !d_out->nrejected)
    return fail(ctx, PHC_ERR_INVALID_INPUT, "null device input or output");
  std::lock_guard<std::mutex> lk(ctx->mu);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  char* const base = (char*)noalloc_ptr(ctx->device);
  if (!base)
    return fail(ctx, PHC_ERR_DEVICE, "no scratch address for device %d",
                ctx->device);
  std::lock_guard<std::mutex> run_lk(g_noalloc_run[ctx->device]);
  HostReq h;
  int rc = validate_and_pack(ctx, sys, hom, opts, npaths, h);
  if (rc) return rc;
  ClassRun cr;
  cr.G = h.G;
  cr.reqs.push_back(&h);
  plan_class(ctx, cr);
  if (cr.blocks == 0)
    return fail(ctx, PHC_ERR_CAPACITY,
                "one path's working set exceeds the LDS of a workgroup");
  // lay out the scratch: counters (cleared), queue head, offsets, request,
  // then the tables
  size_t used = 0;
  auto take = [&](size_t bytes) {
    char* p = base + used;
    used += Arena::al(bytes);
    return p;
  };
  unsigned long long* d_next = (unsigned long long*)take(8);
  unsigned long long* d_counters = (unsigned long long*)take(32);
  const size_t clear_bytes = used;
  int64_t* d_base = (int64_t*)take(2 * sizeof(int64_t));
  Req* d_req = (Req*)take(sizeof(Req));
  TermRec* d_terms = (TermRec*)take(h.terms.size() * sizeof(TermRec));
  WTerm* d_wterms = h.wterms.empty()
                        ? nullptr
                        : (WTerm*)take(h.wterms.size() * sizeof(WTerm));
  uint4* d_jobs =
      h.jobs.empty() ? nullptr : (uint4*)take(h.jobs.size() * sizeof(uint4));
  if (used > NOALLOC_SCRATCH_BYTES)
    return fail(ctx, PHC_ERR_CAPACITY,
                "the system's tables need %zu bytes of scratch; the "
                "allocation-free entry point holds %zu",
                used, NOALLOC_SCRATCH_BYTES);
  Req& r = h.r;
  r.terms = d_terms;
  r.wterms = cr.mode == MODE_GRID ? d_wterms : nullptr;
  r.jobs = cr.mode == MODE_GRID ? d_jobs : nullptr;
  r.start = (const cx*)d_start;
  r.o_status = d_out->status;
  r.o_t = d_out->t;
  r.o_z = (cx*)d_out->z;
  r.o_wind = d_out->winding;
  r.o_egs = d_out->eg_status;
  r.o_egerr = d_out->eg_error;
  r.o_egres = d_out->eg_residual;
  r.o_acc = d_out->naccepted;
  r.o_rej = d_out->nrejected;
  r.o_nlin = d_out->nlin;
  r.o_nbe = d_out->nbe;
  r.o_done = nullptr;
  const int64_t offs[2] = {0, npaths};
  hipStream_t s = ctx->stream;
  const int words = (int)(clear_bytes / 4);
  hipLaunchKernelGGL(zero_words, dim3((words + 255) / 256), dim3(256), 0, s,
                     (uint32_t*)base, words);
  HIPCHK(ctx, hipGetLastError());
  HIPCHK(ctx, hipMemcpyAsync(d_base, offs, sizeof(offs),
                             hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(d_req, &r, sizeof(Req), hipMemcpyHostToDevice, s));
  HIPCHK(ctx, hipMemcpyAsync(d_terms, h.terms.data(),
                             h.terms.size() * sizeof(TermRec),
                             hipMemcpyHostToDevice, s));
  if (r.wterms)
    HIPCHK(ctx, hipMemcpyAsync(d_wterms, h.wterms.data(),
                               h.wterms.size() * sizeof(WTerm),
                               hipMemcpyHostToDevice, s));
  if (r.jobs)
    HIPCHK(ctx, hipMemcpyAsync(d_jobs, h.jobs.data(),
                               h.jobs.size() * sizeof(uint4),
                               hipMemcpyHostToDevice, s));
  Launch L;
  memset(&L, 0, sizeof(L));
  L.reqs = d_req;
  L.req_base = d_base;
  L.nreq = 1;
  L.expmax = cr.expmax;
  L.vmax = cr.vmax;
  L.next = d_next;
  L.total = npaths;
  L.counters = d_counters;
  HIPCHK(ctx, launch_g(cr.G, L, cr.staged, cr.blocks, cr.threads, cr.lds, s,
                       cr.mode));
  unsigned long long c[4] = {0, 0, 0, 0};
  HIPCHK(ctx, hipMemcpyAsync(c, d_counters, sizeof(c), hipMemcpyDeviceToHost,
                             s));
  // the host tables above must outlive the copies: the call is synchronous
  HIPCHK(ctx, hipStreamSynchronize(s));
  ctx->counters[0] = (unsigned long long)cr.blocks * (cr.threads / cr.lp);
  ctx->counters[1] = c[1];
  ctx->counters[2] = c[2];
  ctx->counters[3] = c[3];
  ctx->reruns = c[0];
  return PHC_OK;
}

int64_t phc_noalloc_scratch_bytes(void) {
  return (int64_t)NOALLOC_SCRATCH_BYTES;
}

}  // extern "C"

#include "phc_server.inc"
