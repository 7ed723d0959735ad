// The darcy C++17 API (include/darcy/solver.hpp) on top of the C ABI.
//
// Maps the reference's argument checks and messages (solver.hpp:85-87,
// :182-187) and its exceptions: InvalidInput for bad input, runtime_error for
// device failures.  One process-wide context serves all calls.
#include <algorithm>
#include <cmath>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "darcy/solver.hpp"
#include "darcy_hip.h"

namespace darcy {
namespace {

darcy_ctx* context() {
    static std::once_flag once;
    static darcy_ctx* ctx = nullptr;
    static int rc = DARCY_OK;
    std::call_once(once, [] { rc = darcy_ctx_create(0, &ctx); });
    if (!ctx) {
        throw std::runtime_error(rc == DARCY_HIP_ERROR ? "no HIP device available for the darcy GPU solver"
                                                       : "could not create the darcy GPU context");
    }
    return ctx;
}

void check(int rc, darcy_ctx* ctx) {
    if (rc == DARCY_OK) return;
    const std::string msg = darcy_last_error(ctx);
    if (rc == DARCY_INVALID_INPUT) throw InvalidInput(msg);
    throw std::runtime_error(msg);
}

// The hierarchy options, clamped to what 32-bit fields hold: a side limit
// beyond any grid side, or a depth beyond 32 levels, acts like the limit.
void set_hierarchy(darcy_opts& c, const MultigridOptions& m) {
    c.min_coarse_n = static_cast<int32_t>(std::min<index_t>(std::max<index_t>(m.min_coarse_n, 0), 1 << 30));
    c.max_levels = m.max_levels;
}

darcy_opts to_c(const SolveOptions& o) {
    darcy_opts c;
    darcy_default_opts(&c);
    c.tol = o.tol;
    c.max_iter = o.max_iter;
    c.preconditioner = o.preconditioner == Preconditioner::None     ? DARCY_PC_NONE
                       : o.preconditioner == Preconditioner::Jacobi ? DARCY_PC_JACOBI
                                                                    : DARCY_PC_MULTIGRID;
    c.pre_sweeps = o.multigrid.pre_sweeps;
    c.post_sweeps = o.multigrid.post_sweeps;
    c.cycle_index = o.multigrid.cycle_index;
    c.record_history = o.record_history ? 1 : 0;
    c.precision = o.precision == Precision::Mixed ? DARCY_MIXED : DARCY_FP64;
    set_hierarchy(c, o.multigrid);
    return c;
}

// solver.hpp:85-87
void check_scalars(index_t n, const SolveOptions& o) {
    if (n < 1) throw InvalidInput("grid size must be at least 1");
    if (o.tol <= 0.0) throw InvalidInput("tol must be positive");
    if (o.max_iter < 1) throw InvalidInput("max_iter must be at least 1");
}

// History buffer: one entry per iteration plus the initial residual, capped
// so that a huge max_iter does not allocate up front (a solve that runs
// longer keeps the first kHistoryCap entries).
constexpr std::size_t kHistoryCap = std::size_t{1} << 20;
std::size_t history_size(const SolveOptions& o) {
    return std::min<std::size_t>(static_cast<std::size_t>(o.max_iter) + 1, kHistoryCap);
}

int g_threads = 0;

}  // namespace

Preconditioner preconditioner_from_string(const std::string& name) {
    if (name == "multigrid" || name == "mg") return Preconditioner::Multigrid;
    if (name == "jacobi") return Preconditioner::Jacobi;
    if (name == "none") return Preconditioner::None;
    throw InvalidInput("unknown preconditioner '" + name + "'; expected one of: multigrid, jacobi, none");
}

const char* preconditioner_name(Preconditioner p) {
    switch (p) {
        case Preconditioner::Multigrid:
            return "multigrid";
        case Preconditioner::Jacobi:
            return "jacobi";
        case Preconditioner::None:
            return "none";
    }
    return "unknown";
}

void validate_permeability(const double* a, index_t size) {
    for (index_t i = 0; i < size; ++i) {
        if (!(a[i] > 0.0) || !std::isfinite(a[i])) {
            throw InvalidInput(
                "permeability must be finite and strictly positive everywhere; offending value at flat index " +
                std::to_string(static_cast<long long>(i)));
        }
    }
}

SolveResult solve(const double* a, const double* rhs, index_t n, const SolveOptions& opts, const double* u0) {
    check_scalars(n, opts);
    darcy_ctx* ctx = context();
    SolveResult out;
    out.n = n;
    out.u.assign(static_cast<std::size_t>(n * n), 0.0);
    darcy_opts c = to_c(opts);
    darcy_stats st{};
    std::vector<double> hist;
    if (opts.record_history) hist.resize(history_size(opts));
    check(darcy_solve(ctx, a, rhs, 1.0, n, &c, u0, out.u.data(), &st, hist.empty() ? nullptr : hist.data(),
                      static_cast<int32_t>(hist.size())),
          ctx);
    out.iterations = st.iterations;
    out.residual = st.residual;
    out.converged = st.converged != 0;
    out.levels = st.levels;
    if (opts.record_history) {
        hist.resize(static_cast<std::size_t>(st.history_len));
        out.history = std::move(hist);
    }
    return out;
}

SolveResult solve_constant_rhs(const double* a, index_t n, double f, const SolveOptions& opts) {
    check_scalars(n, opts);
    darcy_ctx* ctx = context();
    SolveResult out;
    out.n = n;
    out.u.assign(static_cast<std::size_t>(n * n), 0.0);
    darcy_opts c = to_c(opts);
    darcy_stats st{};
    std::vector<double> hist;
    if (opts.record_history) hist.resize(history_size(opts));
    check(darcy_solve(ctx, a, nullptr, f, n, &c, nullptr, out.u.data(), &st, hist.empty() ? nullptr : hist.data(),
                      static_cast<int32_t>(hist.size())),
          ctx);
    out.iterations = st.iterations;
    out.residual = st.residual;
    out.converged = st.converged != 0;
    out.levels = st.levels;
    if (opts.record_history) {
        hist.resize(static_cast<std::size_t>(st.history_len));
        out.history = std::move(hist);
    }
    return out;
}

BatchResult solve_batch_uneven(const std::vector<const double*>& a, const std::vector<index_t>& n,
                               const std::vector<const double*>& rhs, double f, const SolveOptions& opts,
                               const std::vector<double*>& u) {
    if (a.size() != n.size() || u.size() != n.size() || (!rhs.empty() && rhs.size() != n.size()))
        throw InvalidInput("a, n, rhs and u must describe the same number of problems");
    for (index_t s : n) check_scalars(s, opts);
    darcy_ctx* ctx = context();
    const auto count = static_cast<int64_t>(n.size());
    BatchResult r;
    r.iterations.assign(n.size(), 0);
    r.residuals.assign(n.size(), 0.0);
    r.converged.assign(n.size(), false);
    r.levels.assign(n.size(), 1);
    if (count == 0) return r;
    std::vector<int64_t> sizes(n.begin(), n.end());
    std::vector<double> fs(n.size(), f);
    std::vector<darcy_stats> st(n.size());
    darcy_opts c = to_c(opts);
    c.record_history = 0;
    check(darcy_solve_batch(ctx, count, sizes.data(), a.data(), rhs.empty() ? nullptr : rhs.data(), fs.data(), &c,
                            nullptr, u.data(), st.data()),
          ctx);
    for (std::size_t i = 0; i < n.size(); ++i) {
        r.iterations[i] = st[i].iterations;
        r.residuals[i] = st[i].residual;
        r.converged[i] = st[i].converged != 0;
        r.levels[i] = st[i].levels;
    }
    return r;
}

BatchResult solve_batch(const double* a, index_t batch, index_t n, const double* rhs, double f,
                        const SolveOptions& opts, double* u) {
    if (batch < 0 || n < 1) throw InvalidInput("a must have shape (batch, n, n)");
    const std::size_t b = static_cast<std::size_t>(batch), stride = static_cast<std::size_t>(n * n);
    std::vector<const double*> ap(b), rp;
    std::vector<double*> up(b);
    std::vector<index_t> ns(b, n);
    for (std::size_t s = 0; s < b; ++s) {
        ap[s] = a + s * stride;
        up[s] = u + s * stride;
    }
    if (rhs) rp.assign(b, rhs);
    return solve_batch_uneven(ap, ns, rp, f, opts, up);
}

void apply_operator(const double* a, index_t n, const double* x, double* y) {
    if (n < 1) throw InvalidInput("a must be a square 2-D array with at least one cell");
    darcy_ctx* ctx = context();
    check(darcy_apply_operator(ctx, a, n, x, y, 0), ctx);
}

void apply_preconditioner(const double* a, index_t n, const double* r, double* z, const MultigridOptions& mg,
                          Precision precision) {
    if (n < 1) throw InvalidInput("a must be a square 2-D array with at least one cell");
    if (mg.pre_sweeps < 1 || mg.post_sweeps < 1) throw InvalidInput("sweeps must be at least 1");
    if (mg.cycle_index < 1 || mg.cycle_index > 2) throw InvalidInput("cycle_index must be 1 or 2");
    darcy_ctx* ctx = context();
    if (mg.min_coarse_n != 1 || mg.max_levels != 32) {
        // Non-default hierarchies go through the options form of the call.
        darcy_opts c;
        darcy_default_opts(&c);
        c.pre_sweeps = mg.pre_sweeps;
        c.post_sweeps = mg.post_sweeps;
        c.cycle_index = mg.cycle_index;
        c.precision = precision == Precision::Mixed ? DARCY_MIXED : DARCY_FP64;
        set_hierarchy(c, mg);
        check(darcy_apply_preconditioner_opts(ctx, a, n, r, z, &c), ctx);
        return;
    }
    check(darcy_apply_preconditioner(ctx, a, n, r, z, mg.pre_sweeps, mg.post_sweeps, mg.cycle_index,
                                     precision == Precision::Mixed ? DARCY_MIXED : DARCY_FP64, 0),
          ctx);
}

int max_threads() noexcept {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

bool has_openmp() noexcept {
#ifdef _OPENMP
    return true;
#else
    return false;
#endif
}

void set_num_threads(int threads) noexcept {
#ifdef _OPENMP
    if (threads > 0) {
        g_threads = threads;
        omp_set_num_threads(threads);
    }
#else
    (void)threads;
#endif
}

std::string device_info() {
    darcy_ctx* ctx = context();
    char buf[256];
    darcy_device_info(ctx, buf, sizeof buf);
    return buf;
}

double last_kernel_ms() {
    darcy_ctx* ctx = context();
    double ms = 0.0;
    darcy_last_kernel_ms(ctx, &ms);
    return ms;
}

void set_tuning(const std::string& key, int value) {
    darcy_ctx* ctx = context();
    check(darcy_ctx_set_tuning(ctx, key.c_str(), value), ctx);
}

int get_tuning(const std::string& key) {
    darcy_ctx* ctx = context();
    int32_t v = 0;
    if (darcy_ctx_get_tuning(ctx, key.c_str(), &v) != DARCY_OK) throw InvalidInput("unknown tuning key: " + key);
    return v;
}

}  // namespace darcy
