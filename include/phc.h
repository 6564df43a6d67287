/*
 * phc.h -- C ABI of the MI355X path-tracking library (libphc_hip.so).
 *
 * The library tracks the paths of polyhomocont's total-degree gamma-trick
 * homotopy on the GPU: for every start point it runs the predictor-corrector
 * tracker from t = 1 to the start of the end game and then the Cauchy end
 * game, and returns one endpoint record per path. Everything before the
 * tracking (random constants, start points) and after it (classification,
 * refinement, deduplication) stays on the host, in the Fortran layer that
 * keeps the reference API `solve(sys, res, status, options)`.
 *
 * Conventions
 *  - Complex numbers are two consecutive doubles (re, im), i.e. the layout
 *    of Fortran complex(c_double_complex) and C99 double _Complex.
 *  - All indices are 0-based. Arrays are caller-owned; the library copies
 *    what it needs into its own device memory during the call.
 *  - Return values: PHC_OK (0) or one of the PHC_ERR_* codes. Positive codes
 *    reuse the reference status codes (src/config.f90:17-23); negative codes
 *    are device/runtime errors. The library never falls back to a CPU path:
 *    without a usable GPU every tracking entry point returns PHC_ERR_DEVICE.
 *  - Every entry point is reentrant; state lives in the explicit context.
 *
 * Reference interfaces replaced (file:line in EVADEVacStab/polyhomocont):
 *  - phc_track_paths        <- the per-path loop of `solve`
 *                              (src/solve.f90:190-192: start_solution +
 *                              endpoint) and `endpoint` itself
 *                              (src/solve.f90:217-252) with everything it
 *                              calls: track_path/track_segment/predict_rk4/
 *                              tangent/newton_correct (src/tracker.f90:179-360),
 *                              cauchy_endgame/backward_error/better/loops
 *                              (src/endgame.f90:124-378), tdh_evaluate
 *                              (src/homotopy.f90:188-231), eval_terms via
 *                              sparse_evaluate_homogeneous
 *                              (src/sparse.f90:209-280), lu_factor/lu_solve/
 *                              norm_inf (src/linalg.f90:26-114,176-189).
 *  - phc_track_batch        <- the same per-path loop for many independent
 *                              `solve` calls at once (no reference
 *                              counterpart: the reference is serial).
 *  - phc_server_*           <- no reference counterpart: a continuous-
 *                              batching request loop around the same kernels.
 */
#ifndef PHC_H
#define PHC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
/* Library return codes; 1..3 are the reference codes (src/config.f90:17-23). */
#define PHC_OK 0
#define PHC_ERR_INVALID_INPUT 1
#define PHC_ERR_SINGULAR_MATRIX 2
#define PHC_ERR_NUMERICAL 3
#define PHC_ERR_DEVICE (-1)      /* no usable GPU, or a HIP call failed      */
#define PHC_ERR_CAPACITY (-2)    /* request exceeds a configured limit, or a
                                    valid system the kernels cannot hold
                                    (N > PHC_MAX_UNKNOWNS, an exponent above
                                    PHC_MAX_EXPONENT): the Fortran `solve`
                                    tracks such systems on the host         */
#define PHC_ERR_PENDING (-3)     /* server ticket not finished yet           */

/* Per-path status (src/tracker.f90:45-51, src/solve.f90:53). */
#define PHC_PATH_SUCCESS 0
#define PHC_PATH_FAILED_MIN_STEP 1
#define PHC_PATH_FAILED_MAX_STEPS 2
#define PHC_PATH_ENDGAME_FAILED 3

/* End game status (src/endgame.f90:64-71). */
#define PHC_ENDGAME_CONVERGED 0
#define PHC_ENDGAME_NOT_CONVERGED 1
#define PHC_ENDGAME_FAILED 2

/* Largest number of homotopy unknowns N = n + 1, and largest exponent of a
 * homogenized term, that the kernels support (term exponents are bytes). */
#define PHC_MAX_UNKNOWNS 16
#define PHC_MAX_EXPONENT 255

/* ---- options ----------------------------------------------------------- */
/* Field-for-field mirror of the reference option types; the defaults are the
 * reference defaults (phc_default_options). Fields are ordered doubles first
 * so that the C layout and the Fortran bind(C) layout carry no padding. */

/* tracker_options (src/tracker.f90:56-77) */
typedef struct phc_tracker_options {
  double step_init;      /* 0.05   */
  double step_max;       /* 0.1    */
  double step_min;       /* 1e-14  */
  double corrector_tol;  /* 1e-10  */
  double contraction;    /* 0.5    */
  int32_t corrector_maxit; /* 3    */
  int32_t grow_after;    /* 5      */
  int32_t max_steps;     /* 20000  */
  int32_t reserved;      /* 0      */
} phc_tracker_options;

/* endgame_options (src/endgame.f90:73-101) */
typedef struct phc_endgame_options {
  double t_start;           /* 0.1   */
  double radius_factor;     /* 0.25  */
  double radius_min;        /* 1e-13 */
  double loop_radius;       /* 1e-5  */
  double winding_agreement; /* 0.2   */
  double tol;               /* 1e-11 */
  double residual_tol;      /* 1e-8  */
  double closure_factor;    /* 0.1   */
  int32_t nsamples;         /* 8     */
  int32_t max_winding;      /* 16    */
} phc_endgame_options;

/* solve_options (src/solve.f90:57-84) */
typedef struct phc_solve_options {
  phc_tracker_options tracker;
  phc_endgame_options endgame;
  double infinity_tol;       /* 1e-8  */
  double singular_cond;      /* 1e10  */
  double real_tol;           /* 1e-8  */
  double dedup_tol;          /* 1e-8  */
  double singular_dedup_tol; /* 1e-6  */
  int32_t use_endgame;       /* 1 (logical .true.) */
  int32_t seed;              /* 1     */
  int32_t refine_maxit;      /* 5     */
  int32_t reserved;          /* 0     */
} phc_solve_options;

/* Fills *opts with the reference defaults. */
void phc_default_options(phc_solve_options *opts);

/* ---- inputs ------------------------------------------------------------ */
/* A sparse_system in homogenized form: the private components of the
 * reference type (src/sparse.f90:31-48) after add_equation has dropped zero
 * terms and appended the w exponent d_i - |a| (src/sparse.f90:119-138). */
typedef struct phc_system {
  int32_t n;              /* variables = equations; N = n + 1 unknowns   */
  int32_t nterms;         /* total number of terms                       */
  int32_t maxexp;         /* largest entry of hexps                      */
  int32_t reserved;
  const int32_t *first;   /* [n + 1] term offsets: equation i owns terms
                             first[i] .. first[i+1]-1 (0-based)          */
  const int32_t *deg;     /* [n] total degrees d_i >= 1                  */
  const int32_t *hexps;   /* [nterms][N] homogenized exponent vectors, w
                             last (Fortran hexps(N, nterms) column-major) */
  const double *coefs;    /* [nterms][2] complex coefficients            */
} phc_system;

/* Random constants of total_degree_homotopy (src/homotopy.f90:78-100),
 * drawn on the host (src/homotopy.f90:135-146). */
typedef struct phc_homotopy {
  const double *gamma;    /* [2]                                         */
  const double *b;        /* [n][2] start system constants, |b_i| = 1    */
  const double *patch;    /* [N][2] affine patch a, unit 2-norm          */
} phc_homotopy;

/* ---- outputs ----------------------------------------------------------- */
/* One record per path: what `endpoint` leaves in path_result
 * (src/solve.f90:217-252) plus the endpoint z and the end game's own
 * residual. Struct of arrays, each of length npaths; pointers marked
 * optional may be NULL. */
typedef struct phc_endpoints {
  int32_t *status;      /* PHC_PATH_*                                       */
  double *t;            /* real t reached (0 after a successful end game)   */
  double *z;            /* [npaths][N][2]: end game estimate on success,
                           else the last point tracked                      */
  int32_t *winding;     /* winding number (0 without end game)              */
  int32_t *eg_status;   /* PHC_ENDGAME_* (0 without end game)               */
  double *eg_error;     /* endgame_result%error (0 without end game)        */
  double *eg_residual;  /* optional: endgame_result%residual                */
  int32_t *naccepted;   /* accepted steps, main path + end game             */
  int32_t *nrejected;   /* rejected steps, main path + end game             */
  int32_t *nlin;        /* optional: linearizations (evaluate + LU + solve) */
  int32_t *nbe;         /* optional: backward-error evaluations             */
} phc_endpoints;

/* ---- context ----------------------------------------------------------- */
typedef struct phc_ctx phc_ctx;

/* Creates a context on HIP device `device` (stream, device arenas, kernel
 * configuration). Returns PHC_ERR_DEVICE if the device cannot be used. */
int phc_ctx_create(int device, phc_ctx **ctx);
int phc_ctx_destroy(phc_ctx *ctx);

/* Lanes per group (4, 8 or 16) for systems with N unknowns: the power of
 * two >= N that sizes the path's state. */
int phc_ctx_lanes_per_path(const phc_ctx *ctx, int N);

/* Device counters of the last tracking call on this context:
 * out[0] resident path slots of the launch (workgroups x paths each);
 * out[1] linearization units executed (all paths); out[2] unit slots
 * offered: over all wavefronts, the passes of the tracking loop in which at
 * least one of the wavefront's paths ran a unit, times the paths a
 * wavefront holds. out[1] / out[2] is the lane occupancy left after path
 * divergence (1 when a wavefront holds one path). out[3] the most units
 * one path slot ran: out[1] / (out[0] x out[3]) is the slot-time
 * utilisation (the share of the slots' time spent on units while the
 * busiest slot works, counting units as equal). */
int phc_ctx_counters(const phc_ctx *ctx, uint64_t *out);

/* Linearization units of the last plain or batched tracking call that the
 * grid kernels redid with the exact pivot rule and exact divisions: the
 * fast unit screens pivots by |a|^2 and divides through a reciprocal formed
 * once per divisor, and flags the rare unit where either does not decide
 * bit for bit (near ties, extreme magnitudes, NaN or infinite entries).
 * Diagnostics only: results never depend on it. No reference counterpart
 * (the reference divides directly). */
int phc_ctx_exact_reruns(const phc_ctx *ctx, uint64_t *out);

/* Lane mapping of later tracking calls. 0 chooses per launch: an
 * autotuned choice (phc_ctx_autotune, phc_ctx_autotune_batch) for lanes
 * classes of a tuned shape, else the min(G*G, 64)-lane grid (a path's linear system one
 * entry per lane, up to four per lane for G = 16, its evaluation spread
 * over all lanes) when the request fits the grid tables, else one path per
 * G-lane group. 1 forces one path per G-lane group (row per lane), 2 one
 * path per wavefront with the terms of each equation split across 64/G
 * replicas where N <= 64/G, 3 the grid where the tables fit. wg_threads: 0
 * (chosen per launch), 64, 128 or 256. Results do not depend on the
 * mapping. */
int phc_ctx_set_mapping(phc_ctx *ctx, int mapping, int wg_threads);

/* Last error message of the context (static storage, never NULL). With
 * several threads on one context it may be another thread's message;
 * phc_thread_last_error gives the calling thread's own. */
const char *phc_ctx_last_error(const phc_ctx *ctx);
/* Message of the last failed library call made by the calling thread
 * (thread-local storage, never NULL; "" if none failed). */
const char *phc_thread_last_error(void);

/* Whether the kernels can track this system with these options, without
 * touching a device: PHC_OK, PHC_ERR_INVALID_INPUT (the system or options
 * are malformed) or PHC_ERR_CAPACITY (valid, but beyond the kernels'
 * limits). The tracking entry points make the same checks. */
int phc_check_system(const phc_system *sys, const phc_solve_options *opts);

/* Process-wide context of a device, created on first use under a lock and
 * kept until the process ends (one per device index): the device-context
 * cache of callers that hold no context of their own, such as the Fortran
 * `solve`. Never pass it to phc_ctx_destroy. PHC_ERR_DEVICE if the device
 * cannot be used. */
int phc_ctx_shared(int device, phc_ctx **ctx);
/* Device index that callers without a context of their own use (default
 * 0; -1 asks them to track on the host CPU). Atomic. */
void phc_set_default_device(int device);
int phc_get_default_device(void);

/* ---- plain per-solve tracker ------------------------------------------ */
/* Tracks paths first_path .. first_path+npaths-1 of one system (path k uses
 * start point start[k - first_path]) and writes one endpoint record per
 * path. `start` is [npaths][N][2] (host memory, tdh_start_solution,
 * src/homotopy.f90:151-177). Synchronous. */
int phc_track_paths(phc_ctx *ctx, const phc_system *sys,
                    const phc_homotopy *hom, const phc_solve_options *opts,
                    int64_t npaths, const double *start, phc_endpoints *out);

/* Allocation-free variant of phc_track_paths, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * `d_start` ([npaths][N][2] doubles) and every array of `d_out` (npaths
 * entries; z [npaths][N][2]; eg_residual, nlin, nbe may be NULL) are
 * DEVICE memory the caller owns on the context's device, ready before the
 * call (synchronize the stream that wrote them). The library calls no
 * hipMalloc/hipFree: the request record, the term and grid tables and the
 * counters of the launch live in a static device array of
 * phc_noalloc_scratch_bytes() bytes (1 MiB) that each GPU holds until the
 * process ends; a system whose tables do not fit returns PHC_ERR_CAPACITY
 * (C2's 840 terms need about 100 KB). Calls on one GPU share that array:
 * the library serializes them per device within the process (the array
 * keeps the last call's contents), and a caller that ran two of them at
 * once from outside that serialization would race. Synchronous; results
 * are those of phc_track_paths, bit for bit. */
int phc_track_paths_noalloc(phc_ctx *ctx, const phc_system *sys,
                            const phc_homotopy *hom,
                            const phc_solve_options *opts, int64_t npaths,
                            const double *d_start, phc_endpoints *d_out);
int64_t phc_noalloc_scratch_bytes(void);

/* Same, with the system, constants, start points and outputs already in
 * device memory prepared by phc_stage_* below (used for timing with inputs
 * resident in HBM). Asynchronous on the context stream; phc_ctx_sync waits. */
typedef struct phc_staged phc_staged;
int phc_stage_create(phc_ctx *ctx, const phc_system *sys,
                     const phc_homotopy *hom, const phc_solve_options *opts,
                     int64_t npaths, const double *start, phc_staged **st);
int phc_stage_run(phc_ctx *ctx, phc_staged *st);
int phc_stage_fetch(phc_ctx *ctx, phc_staged *st, phc_endpoints *out);
int phc_stage_destroy(phc_ctx *ctx, phc_staged *st);
/* Device time (ms) of the tracking kernels of the last phc_stage_run,
 * measured with HIP events: from the point of the context stream where
 * every lanes class starts (the classes of one run are launched together
 * on their own streams) to the end of the last class; waits for the run
 * to finish. */
int phc_stage_elapsed(phc_ctx *ctx, phc_staged *st, double *ms);
/* Launch shape of the first lanes class (info has room for 8 values):
 * info[0] group size G (a power of two >= N), [1] workgroups, [2] threads
 * per workgroup, [3] LDS bytes per workgroup, [4] 1 if the system program
 * is staged in LDS, [5] lane mapping (0 one path per G-lane group, 1 one
 * path per wavefront with split evaluation, 2 one path per G*G-lane grid),
 * [6] lanes per path. */
int phc_stage_launch_info(phc_ctx *ctx, phc_staged *st, int32_t *info);
int phc_ctx_sync(phc_ctx *ctx);

/* Autotuner of the launch shape. Times, on the first min(npaths, 4096)
 * start points of this request, every lane mapping the request allows (grid,
 * row, split) with 64-, 128- and 256-thread workgroups, its program staged
 * in LDS or read from L2, and keeps the fastest for later single-request
 * launches of the same shape (lanes class, N, term count, log2 of the path
 * count) while the mapping is 0 (chosen per launch). choice[0] mapping
 * (1 row, 2 split, 3 grid), [1] threads per workgroup, [2] 1 if staged;
 * *best_ms (may be NULL) the winner's kernel time. Results never depend on
 * the choice. No counterpart in the reference (its tracker is serial). */
int phc_ctx_autotune(phc_ctx *ctx, const phc_system *sys,
                     const phc_homotopy *hom, const phc_solve_options *opts,
                     int64_t npaths, const double *start, int32_t *choice,
                     double *best_ms);
/* Autotuner of a batch (phc_track_batch): for every lanes class of the
 * batch, times each launch shape the class allows (lane mapping grid, row
 * or split; 64-, 128- or 256-thread workgroups; program staged in LDS when
 * the class holds one system) on a sample -- the first paths of every
 * system of the class, about 16 k paths in all -- with launches bounded to
 * 64 linearization units per path slot, and keeps the shape of highest
 * unit throughput for later batches of the same class shape (lanes class,
 * system count, total terms, log2 of the total path count) while the
 * mapping is 0. choices (room for 3 x 4 values) gets per class
 * {G, mapping (1 row, 2 split, 3 grid), threads, staged}; *nclasses the
 * class count. Results never depend on the choice. No counterpart in the
 * reference (its tracker is serial). */
int phc_ctx_autotune_batch(phc_ctx *ctx, int32_t nsys, const phc_system *sys,
                           const phc_homotopy *hom,
                           const phc_solve_options *opts,
                           const int64_t *npaths, const double *const *start,
                           int32_t *choices, int32_t *nclasses);
/* Forgets every autotuned choice. */
int phc_ctx_clear_tuning(phc_ctx *ctx);
/* Launch shapes of the lanes classes of the last phc_track_paths /
 * phc_track_batch / phc_stage_run call: for class c < max_classes,
 * info[8c .. 8c+7] = {G, workgroups, threads, LDS bytes, staged, mapping
 * (0 row, 1 split, 2 grid), lanes per path, systems}. Returns the class
 * count. No counterpart in the reference. */
int phc_ctx_last_plan(const phc_ctx *ctx, int32_t max_classes, int32_t *info);

/* ---- batched multi-system tracker ------------------------------------- */
/* Tracks the paths of nsys independent solves: one persistent launch per
 * lanes class (4, 8, 16), the classes running concurrently on their own
 * streams; system s has npaths[s] paths with start points start[s] and
 * writes out[s]. Equivalent, path by path, to nsys phc_track_paths calls.
 * Synchronous. */
int phc_track_batch(phc_ctx *ctx, int32_t nsys, const phc_system *sys,
                    const phc_homotopy *hom, const phc_solve_options *opts,
                    const int64_t *npaths, const double *const *start,
                    phc_endpoints *out);

/* ---- serving ----------------------------------------------------------- */
typedef struct phc_server phc_server;
typedef struct phc_server_config {
  int64_t max_resident_paths;  /* path slots kept resident in HBM          */
  int64_t hbm_budget_bytes;    /* admission budget for request state       */
  int32_t units_per_launch;    /* linearizations per path per kernel launch */
  int32_t launch_us;           /* time budget of a launch in microseconds:
                                  a path slot starts no unit after it (0:
                                  the unit budget alone bounds launches) */
} phc_server_config;

int phc_server_create(phc_ctx *ctx, const phc_server_config *cfg,
                      phc_server **srv);
/* Queues one solve request; the inputs are copied before returning. The
 * request is admitted to the device as soon as its state fits the budget. */
int phc_server_submit(phc_server *srv, const phc_system *sys,
                      const phc_homotopy *hom, const phc_solve_options *opts,
                      int64_t npaths, const double *start, phc_endpoints *out,
                      uint64_t *ticket);
/* One step of the engine. Every lanes class runs its own rounds of
 * launches; a call retires each class whose round has ended (records of
 * the requests it finished copied to their `out`), admits queued requests
 * that fit into the classes between rounds (FIFO within a class; a request
 * that waits holds back only later requests of its class), starts the
 * next round of every idle class with work and returns while the rounds
 * run, so the caller's host work until the next call overlaps the GPU.
 * When no round has ended and none could start, it first waits for one to
 * end (without holding the context). Returns the number of requests
 * completed by this call. */
int phc_server_progress(phc_server *srv);
/* phc_server_progress with the wait bounded: when no round has ended and
 * none could start, waits at most max_wait_us microseconds for one to end
 * (0: returns at once; negative: no bound, as phc_server_progress). For a
 * caller with host work of its own in hand, such as requests to submit.
 * No counterpart in the reference. */
int phc_server_progress_for(phc_server *srv, int64_t max_wait_us);
/* PHC_OK when the request's records are in its `out`, PHC_ERR_PENDING if
 * not yet; phc_server_wait blocks (driving progress) until done. */
int phc_server_poll(phc_server *srv, uint64_t ticket);
int phc_server_wait(phc_server *srv, uint64_t ticket);
/* Forgets a finished ticket (its records stay in the caller's `out`), or
 * cancels one still queued for admission. PHC_ERR_PENDING while its paths
 * are on the device; PHC_ERR_INVALID_INPUT for an unknown ticket. A
 * long-running server calls it once per request, after poll/wait. */
int phc_server_release(phc_server *srv, uint64_t ticket);
#define PHC_SERVER_MAX_CLASSES 16
/* Per lanes class of the engine (returns their number, at most
 * PHC_SERVER_MAX_CLASSES; out holds 8 per class): out[8c
 * + 0] lanes per group G, [1] mapping (0 row, 2 grid), [2] kernel launches
 * so far, [3] their kernel time in ms (HIP events around each round on the
 * class's stream), [4] linearization units run, [5] unit slots offered
 * (see phc_ctx_counters), [6] resident path slots, [7] term values a grid
 * path holds (0 for a row class). */
int phc_server_stats(phc_server *srv, double *out);
/* Lanes class (index into phc_server_stats) that serves a queued or
 * admitted ticket. */
int phc_server_ticket_class(phc_server *srv, uint64_t ticket);
/* Most paths one request may have (the smallest class's ring capacity). */
int64_t phc_server_max_request_paths(const phc_server *srv);
/* counts[0] queued, [1] admitted, [2] finished and not yet released. */
int phc_server_tickets(phc_server *srv, int64_t *counts);
int phc_server_destroy(phc_server *srv);

#ifdef __cplusplus
}
#endif

#endif /* PHC_H */
