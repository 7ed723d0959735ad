// Argument blocks shared by the host launch code and the kernels.
#pragma once

#include <stdint.h>

#include "layout.hpp"

namespace darcy_hip {

// Solve status codes (match DARCY_STATUS_* of include/darcy_hip.h).
enum : int { kRunning = 0, kConverged = 1, kMaxIter = 2, kBreakdown = 3, kInvalid = 4 };

// What the persistent kernel does with each problem.
enum : int { kModeSolve = 0, kModeApplyOperator = 1, kModeApplyPreconditioner = 2 };
// Preconditioner ids (reference solver.hpp:31-35).
enum : int { kPcNone = 0, kPcJacobi = 1, kPcMultigrid = 2 };

// One problem of a batch; every pointer is device memory.
struct Problem {
    int32_t n;
    int32_t pad;
    const double* a;    // n x n permeability
    const double* rhs;  // n x n forcing, or nullptr for the constant `f`
    double f;
    const double* in;   // u0 for solves (nullable), x for apply_operator, r for apply_preconditioner
    double* out;        // u, A x, or M r
    double* history;    // max_iter + 1 doubles, or nullptr
};

// Per-problem outcome written by the kernel.
struct ProblemStats {
    int32_t iterations;
    int32_t status;  // kConverged, kMaxIter, kBreakdown, kInvalid
    int32_t levels;
    int32_t history_len;
    double residual;
    long long bad_index;
};

struct SolveParams {
    double tol;
    int32_t max_iter;
    int32_t precond;
    int32_t pre;
    int32_t post;
    int32_t gamma;
    int32_t mixed;  // FP32 cycle inside FP64 flexible CG
    int32_t mode;
    int32_t record_history;
    int32_t history_cap;  // entries the history buffer holds
    int32_t step_cells;   // cells per row-pipeline step of the global levels
    int32_t min_coarse;   // hierarchy options (multigrid.hpp:43-46)
    int32_t max_levels;
    int32_t breakdown_at;  // testing: iteration (1-based) whose <p, Ap> reads as 0; 0 = off
};

struct BatchArgs {
    const Problem* probs;
    ProblemStats* stats;
    const int32_t* order;  // processing order (largest first)
    uint32_t* next;        // work-queue counter, zero at launch
    char* slots;           // one slot per workgroup
    size_t slot_bytes;
    int32_t count;
    int32_t pad;
    SolveParams prm;
};

// CG scalars of one solve, resident in device memory.
struct GridState {
    double bnorm, rz, alpha, beta, res, pap;
    int it, status, first, pad;
};

struct GridCg {
    const Hier* hd;
    GridState* st;
    const double* a;
    const double* rhs;  // or nullptr: constant f
    double f;
    const double* u0;  // or nullptr
    double* u;         // solution (output buffer)
    double* history;   // or nullptr
    double tol;
    int max_iter;
    int mixed;
    int history_cap;   // entries `history` holds
    int breakdown_at;  // testing: iteration (1-based) whose <p, Ap> reads as 0; 0 = off
};

}  // namespace darcy_hip
