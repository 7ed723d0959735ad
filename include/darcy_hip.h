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
 * (solver.hpp:40-46, multigrid.hpp:29-47) plus three additions.  Fill with
 * darcy_default_opts before setting fields. */
typedef struct darcy_opts {
    double tol;              /* relative residual target (default 1e-10) */
    int32_t max_iter;        /* default 10000 */
    int32_t preconditioner;  /* DARCY_PC_* (default multigrid) */
    int32_t pre_sweeps;      /* default 1 */
    int32_t post_sweeps;     /* default 1 */
    int32_t cycle_index;     /* 1 = V-cycle, 2 = W-cycle (default) */
    int32_t record_history;  /* keep the relative residual per iteration */
    int32_t precision;       /* addition: DARCY_FP64 (default) or DARCY_MIXED */
    int32_t device_pointers; /* addition: 1: array arguments are device pointers */
    int32_t min_coarse_n;    /* stop coarsening at a side <= this (default 1, multigrid.hpp:43-44) */
    int32_t max_levels;      /* hierarchy depth cap (default 32, multigrid.hpp:45-46) */
    int32_t breakdown_at;    /* addition, for tests of the breakdown exit: 0 (default) = off;
                                k > 0 makes <p, A p> of CG iteration k read as 0, so the solve
                                returns as solver.hpp:149-155 does, with k - 1 iterations */
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
 * one pre- and one post-sweep, coarsening to 1 x 1 within 32 levels, FP64,
 * host pointers, no fault injection). */
void darcy_default_opts(darcy_opts* opts);

/* Create a context on HIP device `device`.  No reference counterpart (the
 * reference has no device state). */
int darcy_ctx_create(int device, darcy_ctx** out);
void darcy_ctx_destroy(darcy_ctx* ctx);

/* Launch-shape tuning of a context (no reference counterpart; set by the
 * autotuner, or by hand):
 *   "grid_min_n"      single solves with n above this use the
 *                     multi-workgroup grid path (default 64)
 *   "grid_tail_side"  on the grid path, levels with side <= this run in one
 *                     workgroup (>= 32, default 64)
 *   "grid_graph"      1 (default): the grid path captures its cycle in a HIP
 *                     graph on first use and replays it each iteration;
 *                     0: launches the cycle's kernels one by one
 *   "grid_cg_per_cu"  workgroups per CU of the grid path's streaming CG
 *                     kernels (1..8, default 4; not autotuned)
 *   "batch_step_cells" cells per row-pipeline step on the levels the
 *                     persistent batch kernel keeps in global memory, in
 *                     the FP64 cycle (128..16384, default 1024)
 *   "batch_step_cells_fp32" the same in the FP32 cycle (default 4096)
 *   "batch_wide_side" a batch whose largest side is >= this runs on
 *                     1024-thread workgroups, one per CU, with 4x the
 *                     pipeline step (0: always 256-thread workgroups, four
 *                     per CU; 64..32768, default 384; not autotuned)
 * None of these changes any result: only the launch structure.  Two more
 * keys change the rounding of the grid path's cycle, not the method:
 *   "grid_fp64_dense_tail" 1 (default): the FP64 cycle applies its tail
 *                     levels (see grid_dense_side) as one dense FP64 matrix
 *                     per visit, formed from the FP64 recursive cycle on
 *                     unit vectors (the same linear map, rounded
 *                     differently: solutions agree to about 1e-14); 0: the
 *                     recursion itself, bit for bit the reference's cycle
 *   "grid_dense_side" 16 (default): the dense tail matrix covers the levels
 *                     from the first of side <= 16 (a 256 x 256 matrix at
 *                     most, read from L2 by the tail workgroup); 8: the
 *                     levels of side <= 8 (64 x 64, in LDS), which gives
 *                     the FP32 cycle of the persistent batch kernel bit for
 *                     bit. */
int darcy_ctx_set_tuning(darcy_ctx* ctx, const char* key, int32_t value);
int darcy_ctx_get_tuning(const darcy_ctx* ctx, const char* key, int32_t* value);

/* Message of the last failed call on this context ("" if none). */
const char* darcy_last_error(const darcy_ctx* ctx);

/* Solve -div(a grad u) = f, u = 0 on the boundary, on an n x n grid.
 * Replaces darcy::solve (reference include/darcy/solver.hpp:83-174).
 * rhs: n x n forcing, or NULL for the constant f.  u0: initial guess, or
 * NULL for zero (the reference behaviour).  history: NULL, or room for
 * history_cap doubles (max_iter + 1 suffice); entries beyond history_cap
 * are neither written nor counted in history_len. */
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

/* darcy_solve_batch without any device allocation or free.  No counterpart
 * in the reference; for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Every array is a device pointer (opts->device_pointers must
 * be 1); history is not recorded.  The solve tables and slots live in one
 * static device array of the library, 64 MiB per device
 * (darcy_scratch_bytes): a batch whose tables plus one solve slot of its
 * largest problem exceed it is rejected with DARCY_INVALID_INPUT (one
 * 1024^2 FP64 problem fits; the slots that fit run concurrently).  Calls on
 * one device share that array: calls that overlap in time on one device --
 * from two contexts, or two threads with their own contexts -- race. */
int darcy_solve_batch_noalloc(darcy_ctx* ctx, int64_t count, const int64_t* n, const double* const* a,
                              const double* const* rhs, const double* f, const darcy_opts* opts,
                              const double* const* u0, double* const* u_out, darcy_stats* stats);

/* Bytes of the static scratch array of the context's device (0 if its
 * address could not be looked up). */
int darcy_scratch_bytes(const darcy_ctx* ctx, size_t* bytes);

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

/* darcy_apply_preconditioner with every cycle option in `opts` (sweeps,
 * cycle_index, precision, device_pointers, min_coarse_n, max_levels; the
 * preconditioner field is ignored).  Replaces Multigrid(Level, MultigridOptions)
 * + Multigrid::apply (reference multigrid.hpp:55-71, :88-93) for hierarchies
 * other than the default. */
int darcy_apply_preconditioner_opts(darcy_ctx* ctx, const double* a, int64_t n, const double* r, double* z,
                                    const darcy_opts* opts);

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

/* Gaussian random fields and two-phase media for `count` samples on an
 * n x n grid.  Replaces gaussian_random_field + two_phase_permeability
 * (reference python/darcy/fields.py:31-83) for batches: xi holds the
 * count * n * n standard normals the reference draws (in its order, from
 * numpy.random.default_rng), and each sample is the orthonormal 2-D DCT-III
 * of xi scaled by the spectral amplitude
 * tau^(alpha-1) (pi^2 (k1^2 + k2^2) + tau^2)^(-alpha/2) (mean mode 0).
 * grf_out (may be NULL) receives the fields, a_out (may be NULL) the media
 * hi where the field is >= 0 and lo elsewhere.  Host pointers unless
 * device_pointers != 0.  The DCT is a direct FP64 sum, so values agree with
 * SciPy's FFT-based idctn to round-off (about 1e-16 of the field's scale). */
int darcy_grf_fields(darcy_ctx* ctx, int64_t count, int64_t n, double alpha, double tau, const double* xi,
                     double hi, double lo, double* grf_out, double* a_out, int32_t device_pointers);

/* ------------------------------------------------------------------------
 * Learned warm start.  No counterpart in the reference, whose solve always
 * starts from u = 0 (solver.hpp:99); its output is the u0 argument of
 * darcy_solve / darcy_solve_batch, and CG still runs to tol from it.
 *
 * The network maps (ln a, the forcing, a position channel) on an n x n grid
 * to u0 through a U-Net encoder-decoder with 3x3 kernels, GELU, bf16
 * weights and activations and FP32 accumulation.  Its finest level has side
 * m = n / sub: sub = 1 runs it on the grid itself (n a multiple of 64);
 * with DARCY_NET_HALF_RES, sub = 2 (n a multiple of 128), each network pixel
 * stands for a 2 x 2 block of cells: its input features carry the block's
 * four ln a values and its head gives a smooth channel, interpolated to the
 * cells, and a detail channel per cell of the block, a quarter of the work.
 * Level
 * l has side m / 2^l and channels 32, 64, 128, 256 (levels 0-3, the
 * encoder-decoder proper), then 256, 256, 256 (levels 4-6, a context stack
 * that gives every cell a receptive field spanning the grid):
 *   features: 8 channels per pixel: the sub^2 cells' (ln a - mu) / sd, 1,
 *        the mean position of the cells, zeros (mu, sd map ln 3, ln 12 to
 *        -1, +1)
 *   enc: D_0: conv 8->32 at m; D_l: stride-2 conv from level l-1 to level l
 *   dec: U_i (i = 5..0): conv at level i+1 to the channels of level i, x2
 *        nearest upsampling, + the encoder activation of level i, GELU
 *   head: conv 32->1 in FP32 (sub 1: y itself), or 32->5 (sub 2): y at
 *        cell (2I + di, 2J + dj) is channel 0 interpolated bilinearly to the
 *        cell's centre (weights 3/4, 1/4 per direction, the neighbour index
 *        clamped to the grid) plus channel 1 + 2 di + dj of pixel (I, J);
 *        u0 = f * pos * y / 64, pos = sin(pi (i + 1/2) / n) sin(pi (j + 1/2) / n).
 * Parameters are one FP32 vector: for each conv in the order D_0..D_6,
 * U_5..U_0, head: W[cout][3][3][cin] then bias[cout] (D_0 has 8 input
 * channels, sub^2 + 2 of them used).
 * ------------------------------------------------------------------------ */
typedef struct darcy_net darcy_net;

#define DARCY_NET_TRAIN 1    /* also allocate what training needs */
#define DARCY_NET_HALF_RES 2 /* sub = 2: the network at half the grid's resolution */

/* Length of the parameter vector of a network created with `flags`. */
int64_t darcy_net_param_count(int32_t flags);

/* A network for n x n problems, processing at most max_batch at a time
 * (max_batch * n * n at most 1.34e8); flags: DARCY_NET_TRAIN,
 * DARCY_NET_HALF_RES.  Parameters start at 0. */
int darcy_net_create(darcy_ctx* ctx, int64_t n, int64_t max_batch, int32_t flags, darcy_net** out);
void darcy_net_destroy(darcy_net* net);

/* Host parameter vector in / out; set also resets the optimizer state. */
int darcy_net_set_params(darcy_net* net, const float* params, int64_t count);
int darcy_net_get_params(darcy_net* net, float* params, int64_t count);

/* u0[i] (device, n*n doubles) from a[i] (device) and f[i] (host), for
 * `count` problems (any count; processed max_batch at a time). */
int darcy_net_predict(darcy_net* net, int64_t count, const double* const* a, const double* f, double* const* u0);

/* One Adam step (beta 0.9 / 0.999, eps 1e-8, each weight's step ratio bounded to
 * [-1, 1] times the rate) over a batch of
 * device-resident pairs (a, u) with f = 1, on the loss
 *   (sum_cells e^2 + 1000 e^T A e) / cells,  e = pos y - 64 u,
 * A the finite-volume operator of a (the energy norm CG converges in).
 * *loss (may be null) receives the loss before the step. */
int darcy_net_train_step(darcy_net* net, int64_t batch, const double* const* a, const double* const* u, float lr,
                         double* loss);

/* Loss and gradient (host vector of darcy_net_param_count floats) of the
 * same objective, without an update. */
int darcy_net_gradient(darcy_net* net, int64_t batch, const double* const* a, const double* const* u, double* loss,
                       float* grad);

/* Single-layer entry points on host arrays (bf16 as uint16 bit patterns,
 * NHWC activations, W[cout][3][3][cin]), for checking the conv kernels:
 * stride 1 or 2, cin and cout multiples of 8, h and w multiples of the
 * stride, b * (h/s) * (w/s) a multiple of 8.
 * z / act: [b][h/s][w/s][cout] (act may be null). */
int darcy_nn_conv_forward(darcy_ctx* ctx, int32_t b, int32_t h, int32_t w, int32_t cin, int32_t cout, int32_t stride,
                          const uint16_t* x, const uint16_t* wgt, const float* bias, uint16_t* z, uint16_t* act);
/* dx: [b][h][w][cin] from dz: [b][h/s][w/s][cout]. */
int darcy_nn_conv_backward_data(darcy_ctx* ctx, int32_t b, int32_t h, int32_t w, int32_t cin, int32_t cout,
                                int32_t stride, const uint16_t* dz, const uint16_t* wgt, uint16_t* dx);
/* dw: [cout][3][3][cin], db: [cout] (FP32), split over `splits` pixel ranges. */
int darcy_nn_conv_backward_weight(darcy_ctx* ctx, int32_t b, int32_t h, int32_t w, int32_t cin, int32_t cout,
                                  int32_t stride, int32_t splits, const uint16_t* x, const uint16_t* dz, float* dw,
                                  float* db);
/* C (16 x 16, FP32) = A (16 x 32) B (32 x 16), bf16 row-major, through one
 * matrix-core instruction: checks the operand layout the kernels assume. */
int darcy_nn_mfma_check(darcy_ctx* ctx, const uint16_t* a, const uint16_t* b, float* c);

#ifdef __cplusplus
}
#endif
#endif
