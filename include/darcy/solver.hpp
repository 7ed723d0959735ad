// darcy C++17 API on the MI355X solve path.
//
// Same types, names, defaults and error behaviour as the reference's public
// header (include/darcy/solver.hpp:31-199 and multigrid.hpp:29-47 of the
// reference); the work runs on the GPU through the C ABI of darcy_hip.h.
// Additions are marked; none changes a reference default.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

namespace darcy {

/// Signed 64-bit index, as the reference's config.hpp:51.
using index_t = std::ptrdiff_t;

enum class Preconditioner {
    None,       ///< plain CG
    Jacobi,     ///< diagonal scaling
    Multigrid,  ///< the default
};

/// Arithmetic of the multigrid cycle.  Addition: the reference is FP64 only.
enum class Precision {
    FP64,   ///< everything in FP64 (reference arithmetic)
    Mixed,  ///< FP32 cycle inside an FP64 flexible CG
};

struct MultigridOptions {
    int pre_sweeps = 1;
    int post_sweeps = 1;
    int cycle_index = 2;       ///< 1 = V-cycle, 2 = W-cycle
    index_t min_coarse_n = 1;  ///< only 1 is supported on the GPU path
    int max_levels = 32;       ///< only 32 is supported on the GPU path
};

struct SolveOptions {
    double tol = 1e-10;
    int max_iter = 10000;
    Preconditioner preconditioner = Preconditioner::Multigrid;
    MultigridOptions multigrid{};
    bool record_history = false;
    Precision precision = Precision::FP64;  ///< addition
};

struct SolveResult {
    std::vector<double> u;
    index_t n = 0;
    int iterations = 0;
    double residual = 0.0;
    bool converged = false;
    int levels = 1;
    std::vector<double> history;
};

/// Per-sample outcome of a batch (the lists of _core.solve_batch).
struct BatchResult {
    std::vector<int> iterations;
    std::vector<double> residuals;
    std::vector<bool> converged;
    std::vector<int> levels;
};

class InvalidInput : public std::invalid_argument {
  public:
    using std::invalid_argument::invalid_argument;
};

Preconditioner preconditioner_from_string(const std::string& name);
const char* preconditioner_name(Preconditioner p);

/// Throws InvalidInput at the first non-positive or non-finite entry
/// (host-side check, reference solver.hpp:68-77).
void validate_permeability(const double* a, index_t size);

/// Solve -div(a grad u) = f with u = 0 on the boundary (reference
/// solver.hpp:83-174).  Addition: optional initial guess u0.
SolveResult solve(const double* a, const double* rhs, index_t n, const SolveOptions& opts,
                  const double* u0 = nullptr);

/// Constant forcing (reference solver.hpp:177-180); the forcing is not
/// expanded to an n x n array.
SolveResult solve_constant_rhs(const double* a, index_t n, double f, const SolveOptions& opts);

/// `batch` problems of side n stored back to back in `a`, one shared forcing
/// (`rhs` n x n, or nullptr for constant `f`); u receives batch*n*n values
/// (reference src/bindings.cpp:140-197).
BatchResult solve_batch(const double* a, index_t batch, index_t n, const double* rhs, double f,
                        const SolveOptions& opts, double* u);

/// Problems of different sizes in one launch (addition).  rhs may be empty
/// (constant f for all) or hold one pointer per problem (nullptr = f).
BatchResult solve_batch_uneven(const std::vector<const double*>& a, const std::vector<index_t>& n,
                               const std::vector<const double*>& rhs, double f, const SolveOptions& opts,
                               const std::vector<double*>& u);

/// y = A x (reference bindings.cpp:199-216).
void apply_operator(const double* a, index_t n, const double* x, double* y);

/// z = M r from a zero start (reference bindings.cpp:218-245).
void apply_preconditioner(const double* a, index_t n, const double* r, double* z, const MultigridOptions& mg,
                          Precision precision = Precision::FP64);

/// Host threads used by the library's host loops (the reference's OpenMP
/// controls, config.hpp:54-62).
int max_threads() noexcept;
bool has_openmp() noexcept;
void set_num_threads(int threads) noexcept;

/// Device description and event time of the last solve's kernels (additions).
std::string device_info();
double last_kernel_ms();

/// Launch-shape tuning of the process-wide context (darcy_ctx_set_tuning
/// keys; no result depends on them).
void set_tuning(const std::string& key, int value);
int get_tuning(const std::string& key);

}  // namespace darcy
