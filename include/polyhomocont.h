/*
 * polyhomocont.h -- C-callable layer of the Fortran host library
 * (libpolyhomocont_amd.so).
 *
 * Fortran callers use the module `polyhomocont` directly:
 *   use polyhomocont
 *   call sys%init(n); call sys%add_equation(coefs, exps, status)
 *   call solve(sys, res, status, options)
 * exactly as with the reference (src/polyhomocont.f90:32-53,
 * src/solve.f90:154-163). The bind(C) functions below expose the same
 * `solve` to C and Python with flat arrays, for tests and benchmarks.
 *
 * A system is passed as the user passes it to `add_equation`
 * (src/sparse.f90:83-141): per equation a list of complex coefficients and
 * exponent vectors; zero terms are dropped and inputs validated by the
 * library's own builder.
 */
#ifndef POLYHOMOCONT_H
#define POLYHOMOCONT_H

#include <stdint.h>
#include "phc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct phc_user_system {
  int32_t n;               /* variables                                   */
  int32_t neq;             /* equations given (n for a square system)     */
  const int32_t *nterms;   /* [neq] terms of each equation, as given      */
  const double *coefs;     /* [sum nterms][2]                             */
  const int32_t *exps;     /* [sum nterms][n] exponent vector of each term */
} phc_user_system;

/* solve_result (src/solve.f90:121-150). Array capacities: npaths columns. */
typedef struct phc_solve_summary {
  int32_t n;
  int32_t npaths;
  int32_t nsolutions;
  int32_t nreal;
  int32_t nsingular;
  int32_t ninfinite;
  int32_t nfailed;
  int32_t nduplicates;
  double *solutions;          /* [npaths][n][2]  (first nsolutions used) */
  double *real_solutions;     /* [npaths][n]     (first nreal used)      */
  double *singular_solutions; /* [npaths][n][2]  (first nsingular used)  */
  int32_t *multiplicities;    /* [npaths]        (first nsingular used)  */
} phc_solve_summary;

/* path_result (src/solve.f90:86-119), struct of arrays of length npaths. */
typedef struct phc_path_results {
  int32_t *status;
  double *t;
  double *z;               /* [npaths][N][2]                              */
  double *x;               /* [npaths][n][2]; zeros where at infinity     */
  int32_t *at_infinity;
  int32_t *singular;
  int32_t *is_real;
  int32_t *winding;
  int32_t *eg_status;
  double *eg_error;
  double *cond;
  double *residual;
  int32_t *solution_index;
  int32_t *singular_index;
  int32_t *naccepted;
  int32_t *nrejected;
} phc_path_results;

/* Runs `solve` on the system. Returns the `status` argument of solve
 * (0 or err_invalid_input); a negative value means the GPU path could not
 * run (err_device: solve does not fall back to the host when the device
 * fails; phc_f_set_device(-1) asks for the host tracker explicitly).
 * Systems the kernels cannot hold (more than 15 variables, an exponent
 * above 255) are tracked on the host with the same results. */
int phc_f_solve(const phc_user_system *sys, const phc_solve_options *opts,
                phc_solve_summary *summary, phc_path_results *paths);

/* `solve` on the same system wrapped as a user evaluator (a poly_system
 * extension that supplies only the affine F(x) and its Jacobian, like the
 * user types of the reference's tests): the path that the Fortran `solve`
 * takes for every non-sparse system -- the host tracker, paths over the
 * OpenMP threads with one copy of the system per thread. */
int phc_f_solve_user(const phc_user_system *sys, const phc_solve_options *opts,
                     phc_solve_summary *summary, phc_path_results *paths);

/* Host threads of later solve calls from the calling thread (OpenMP's
 * omp_set_num_threads; nthreads <= 0 leaves it unchanged). Returns the
 * previous count. Results never depend on it. */
int32_t phc_f_set_threads(int32_t nthreads);

/* The random constants and start points `solve` uses for this system and
 * seed (tdh_init, tdh_start_solution). `nstart` receives the Bezout number.
 * Any output pointer may be NULL. start is [count][N][2]. */
int phc_f_homotopy(const phc_user_system *sys, int32_t seed, double *gamma,
                   double *b, double *patch, int64_t *nstart);
int phc_f_start_points(const phc_user_system *sys, int32_t seed,
                       int64_t first, int64_t count, double *start);

/* Packs the homogenized system as the device sees it (phc_system arrays).
 * Capacities: first[n+1], deg[n], hexps[nterms][N], coefs[nterms][2];
 * call with NULL arrays to get nterms and maxexp only. */
int phc_f_pack(const phc_user_system *sys, int32_t *nterms, int32_t *maxexp,
               int32_t *first, int32_t *deg, int32_t *hexps, double *coefs);

/* Host post-processing of `solve` on given endpoint records (classify,
 * add_solution/add_singular, collect_real; src/solve.f90:193-213). */
int phc_f_classify(const phc_user_system *sys, const phc_solve_options *opts,
                   int64_t npaths, const phc_endpoints *rec,
                   phc_solve_summary *summary, phc_path_results *paths);

/* Index of the HIP device used by phc_f_solve (default 0; -1: track
 * sparse systems on the host CPU). */
void phc_f_set_device(int32_t device);

#ifdef __cplusplus
}
#endif

#endif /* POLYHOMOCONT_H */
