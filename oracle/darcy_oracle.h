/*
 * CPU oracle for the darcy solve path -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch C++17/OpenMP restatement of the reference solve path
 * (include/darcy/level.hpp, multigrid.hpp, blas.hpp, solver.hpp and the batch
 * loop of src/bindings.cpp).  It exists to check the HIP implementation and to
 * serve as the CPU baseline of the benchmark.  Nothing in the product library
 * links, loads or calls it; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py do.
 *
 * Every function is C-ABI so the Python tests can call it through ctypes.
 * Status codes: 0 ok, 1 invalid input (offending flat index in *bad_index).
 */
#ifndef DARCY_ORACLE_H
#define DARCY_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Preconditioner ids (reference solver.hpp:31-35). */
enum { ORACLE_PC_NONE = 0, ORACLE_PC_JACOBI = 1, ORACLE_PC_MULTIGRID = 2 };

typedef struct oracle_opts {
    double tol;          /* relative residual target, solver.hpp:41 */
    int32_t max_iter;    /* solver.hpp:42 */
    int32_t precond;     /* ORACLE_PC_* */
    int32_t pre_sweeps;  /* multigrid.hpp:31 */
    int32_t post_sweeps; /* multigrid.hpp:33 */
    int32_t cycle_index; /* multigrid.hpp:42 (1 = V, 2 = W) */
    int32_t record_history;
    int32_t cycle_fp32;  /* 1: cycle in FP32 inside FP64 flexible CG (no reference counterpart) */
    int32_t threads;     /* OpenMP threads for batch loops; 0 = all */
    int32_t min_coarse_n; /* multigrid.hpp:43-44 (reference default 1) */
    int32_t max_levels;   /* multigrid.hpp:45-46 (reference default 32) */
} oracle_opts;

typedef struct oracle_stats {
    int32_t iterations;
    int32_t converged;
    int32_t levels;
    int32_t history_len;
    double residual;
} oracle_stats;

/* Single solve (solver.hpp:83-174).  u0 may be NULL (reference behaviour,
 * u = 0); history may be NULL or hold history_cap doubles. */
int oracle_solve(const double* a, const double* rhs, int64_t n, const oracle_opts* opts,
                 const double* u0, double* u_out, oracle_stats* stats, double* history,
                 int32_t history_cap, int64_t* bad_index);

/* Batch of `count` problems of one size n with a shared rhs, OpenMP dynamic
 * schedule across samples (bindings.cpp:140-197). */
int oracle_solve_batch(const double* a, const double* rhs, int64_t count, int64_t n,
                       const oracle_opts* opts, double* u_out, oracle_stats* stats,
                       int64_t* bad_index);

/* y = A x (level.hpp:105-119 via bindings.cpp:199-216). */
int oracle_apply_operator(const double* a, int64_t n, const double* x, double* y,
                          int64_t* bad_index);

/* z = M r, one cycle from zero (multigrid.hpp:88-93 via bindings.cpp:218-245).
 * cycle_fp32 selects the FP32 cycle. */
int oracle_apply_preconditioner(const double* a, int64_t n, const double* r, double* z,
                                int32_t pre_sweeps, int32_t post_sweeps, int32_t cycle_index,
                                int32_t cycle_fp32, int64_t* bad_index);

/* The same with the cycle options of `opts` (pre_sweeps, post_sweeps,
 * cycle_index, cycle_fp32, min_coarse_n, max_levels). */
int oracle_apply_preconditioner_opts(const double* a, int64_t n, const double* r, double* z,
                                     const oracle_opts* opts, int64_t* bad_index);

/* Number of levels for a fine size n (multigrid.hpp:55-71, level.hpp:164). */
int32_t oracle_num_levels(int64_t n);

/* The hierarchy's operators.  Writes, for each level l, the 6 arrays
 * e, w, n, s, diag, inv_diag (each n_l*n_l doubles) back to back into `out`,
 * level after level.  Returns the total number of doubles written, or the
 * number required when out == NULL. */
int64_t oracle_hierarchy(const double* a, int64_t n, double* out);

/* Maximum OpenMP threads of this build / set them (omp_set_num_threads of
 * the calling thread; 1 makes every reduction serial and deterministic). */
int32_t oracle_max_threads(void);
void oracle_set_num_threads(int32_t threads);

#ifdef __cplusplus
}
#endif
#endif
