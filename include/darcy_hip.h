/*
 * darcy_hip.h -- C ABI of the MI355X (gfx950) solve path.
 *
 * The drop-in boundary for the reference's solve path: plain pointers and
 * sizes, no exceptions, no C++ or torch types.  The C++17 API
 * (include/darcy/solver.hpp) and the pybind11 module darcy._core are built
 * on these entry points; other languages bind them directly (ctypes, cgo,
 * JNI).  Each entry point names the reference interface it replaces.
 *
 * Arrays are row-major n x n float64.  By default every array argument is a
 * host pointer (copied to and from device memory); with
 * darcy_opts.device_pointers = 1 they are device pointers already resident
 * in HBM and nothing crosses PCIe.
 *
 * Return codes: DARCY_OK, DARCY_INVALID_INPUT (message and offending index
 * via darcy_last_error / darcy_stats.bad_index), DARCY_HIP_ERROR,
 * DARCY_OUT_OF_MEMORY.
 *
 * A context owns a HIP stream and the device workspace; calls on one
 * context are serialised by an internal mutex.
 */
#ifndef DARCY_HIP_H
#define DARCY_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DARCY_OK 0
#define DARCY_INVALID_INPUT 1
#define DARCY_HIP_ERROR 2
#define DARCY_OUT_OF_MEMORY 3

/* Preconditioners (reference include/darcy/solver.hpp:31-35). */
#define DARCY_PC_NONE 0
#define DARCY_PC_JACOBI 1
#define DARCY_PC_MULTIGRID 2

/* Precision of the multigrid cycle. */
#define DARCY_FP64 0  /* everything FP64: the reference's arithmetic */
#define DARCY_MIXED 1 /* FP32 cycle inside an FP64 flexible CG */

/* Per-problem status (darcy_stats.status). */
#define DARCY_STATUS_CONVERGED 1
#define DARCY_STATUS_MAX_ITER 2
#define DARCY_STATUS_BREAKDOWN 3
#define DARCY_STATUS_INVALID 4

typedef struct darcy_ctx darcy_ctx;

/* Solver options: SolveOptions + MultigridOptions of the reference
 * (solver.hpp:40-46, multigrid.hpp:29-47) plus two additions. */
typedef struct darcy_opts {
    double tol;              /* relative residual target (default 1e-10) */
    int32_t max_iter;        /* default 10000 */
    int32_t preconditioner;  /* DARCY_PC_* (default multigrid) */
    int32_t pre_sweeps;      /* default 1 */
    int32_t post_sweeps;     /* default 1 */
    int32_t cycle_index;     /* 1 = V-cycle, 2 = W-cycle (default) */
    int32_t record_history;  /* keep the relative residual per iteration */
    int32_t precision;       /* DARCY_FP64 (default) or DARCY_MIXED */
    int32_t device_pointers; /* 1: array arguments are device pointers */
} darcy_opts;

/* Outcome of one solve: the scalar fields of SolveResult (solver.hpp:48-56). */
typedef struct darcy_stats {
    int32_t iterations;
    int32_t converged;
    int32_t levels;
    int32_t status;      /* DARCY_STATUS_* */
    double residual;
    int64_t bad_index;   /* first invalid permeability index, else -1 */
    int32_t history_len; /* entries written to `history` */
    int32_t reserved;
} darcy_stats;

/* Defaults of the reference (tol 1e-10, max_iter 10000, multigrid W-cycle,
 * one pre- and one post-sweep, FP64, host pointers). */
void darcy_default_opts(darcy_opts* opts);

/* Create a context on HIP device `device`.  No reference counterpart (the
 * reference has no device state). */
int darcy_ctx_create(int device, darcy_ctx** out);
void darcy_ctx_destroy(darcy_ctx* ctx);

/* Launch-shape tuning of a context (no reference counterpart; set by the
 * autotuner, or by hand):
 *   "grid_min_n"      single multigrid solves with n above this use the
 *                     multi-workgroup grid path (default 64)
 *   "grid_tail_side"  on the grid path, levels with side <= this run in one
 *                     workgroup (>= 32, default 64)
 * Neither changes any result: only the launch structure. */
int darcy_ctx_set_tuning(darcy_ctx* ctx, const char* key, int32_t value);
int darcy_ctx_get_tuning(const darcy_ctx* ctx, const char* key, int32_t* value);

/* Message of the last failed call on this context ("" if none). */
const char* darcy_last_error(const darcy_ctx* ctx);

/* Solve -div(a grad u) = f, u = 0 on the boundary, on an n x n grid.
 * Replaces darcy::solve (reference include/darcy/solver.hpp:83-174).
 * rhs: n x n forcing, or NULL for the constant f.  u0: initial guess, or
 * NULL for zero (the reference behaviour).  history: NULL, or room for
 * history_cap doubles (max_iter + 1 suffice). */
int darcy_solve(darcy_ctx* ctx, const double* a, const double* rhs, double f, int64_t n,
                const darcy_opts* opts, const double* u0, double* u_out, darcy_stats* stats,
                double* history, int32_t history_cap);

/* Solve `count` independent problems of possibly different sizes n[i] in
 * one persistent launch.  Replaces _core.solve_batch (reference
 * src/bindings.cpp:140-197), which needs one n for the whole batch.
 * rhs[i] may be NULL (then f[i], or 1.0 when f is NULL); u0 may be NULL or
 * hold NULL entries.  stats has `count` entries.  If any problem has an
 * invalid permeability the call returns DARCY_INVALID_INPUT naming the
 * lowest such problem; the other problems are still solved. */
int darcy_solve_batch(darcy_ctx* ctx, int64_t count, const int64_t* n, const double* const* a,
                      const double* const* rhs, const double* f, const darcy_opts* opts,
                      const double* const* u0, double* const* u_out, darcy_stats* stats);

/* y = A x for the finite-volume operator of a.  Replaces _core.apply_operator
 * (reference src/bindings.cpp:199-216, Level::apply level.hpp:105-119). */
int darcy_apply_operator(darcy_ctx* ctx, const double* a, int64_t n, const double* x, double* y,
                         int32_t device_pointers);

/* z = M r, one cycle from zero.  Replaces _core.apply_preconditioner
 * (reference src/bindings.cpp:218-245, Multigrid::apply multigrid.hpp:88-93).
 * precision selects the FP64 or the FP32 cycle. */
int darcy_apply_preconditioner(darcy_ctx* ctx, const double* a, int64_t n, const double* r, double* z,
                               int32_t pre_sweeps, int32_t post_sweeps, int32_t cycle_index,
                               int32_t precision, int32_t device_pointers);

/* Multigrid depth for side n (reference multigrid.hpp:55-71). */
int32_t darcy_num_levels(int64_t n);

/* Device memory helpers, for callers that keep inputs resident in HBM. */
int darcy_device_alloc(darcy_ctx* ctx, size_t bytes, void** out);
int darcy_device_free(darcy_ctx* ctx, void* ptr);
int darcy_copy_to_device(darcy_ctx* ctx, void* dst, const void* src, size_t bytes);
int darcy_copy_to_host(darcy_ctx* ctx, void* dst, const void* src, size_t bytes);

/* HIP-event time of the solve kernels of the last solve call on this
 * context, on the stream they ran on (milliseconds). */
int darcy_last_kernel_ms(const darcy_ctx* ctx, double* ms);

/* Achievable HBM bandwidth: a streaming copy of `bytes` repeated `repeats`
 * times, timed with HIP events; *gbps counts read + write bytes.  No
 * reference counterpart (measurement support). */
int darcy_bandwidth_probe(darcy_ctx* ctx, size_t bytes, int32_t repeats, double* gbps);

/* "gfx950, 256 CUs, HIP 7.2" style description of the context's device. */
int darcy_device_info(const darcy_ctx* ctx, char* buf, int32_t cap);

#ifdef __cplusplus
}
#endif
#endif
