// C-ABI shim around the reference's own header-only solver -- TEST
// INFRASTRUCTURE ONLY.
//
// Compiled by oracle/Makefile against the reference headers where they lie
// (include/darcy/solver.hpp, multigrid.hpp, level.hpp) into oracle/_ref/.
// It exposes the same entry points as the oracle (prefix ref_ instead of
// oracle_) so tests can check the oracle restatement against the reference
// itself, and so the benchmark can time the reference CPU path.  No reference
// source is copied: this file only includes the headers and calls them.
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include "darcy/solver.hpp"
#include "darcy_oracle.h"

namespace {

darcy::SolveOptions to_options(const oracle_opts& o) {
    darcy::SolveOptions s;
    s.tol = o.tol;
    s.max_iter = o.max_iter;
    s.preconditioner = o.precond == ORACLE_PC_NONE     ? darcy::Preconditioner::None
                       : o.precond == ORACLE_PC_JACOBI ? darcy::Preconditioner::Jacobi
                                                       : darcy::Preconditioner::Multigrid;
    s.multigrid.pre_sweeps = o.pre_sweeps;
    s.multigrid.post_sweeps = o.post_sweeps;
    s.multigrid.cycle_index = o.cycle_index;
    s.multigrid.min_coarse_n = o.min_coarse_n;
    s.multigrid.max_levels = o.max_levels;
    s.record_history = o.record_history != 0;
    return s;
}

void to_stats(const darcy::SolveResult& r, oracle_stats* st) {
    if (!st) return;
    st->iterations = r.iterations;
    st->converged = r.converged ? 1 : 0;
    st->levels = r.levels;
    st->history_len = static_cast<int32_t>(r.history.size());
    st->residual = r.residual;
}

}  // namespace

extern "C" int ref_solve(const double* a, const double* rhs, int64_t n, const oracle_opts* opts,
                         double* u_out, oracle_stats* stats, double* history, int32_t history_cap) {
    try {
        const darcy::SolveResult r = darcy::solve(a, rhs, static_cast<darcy::index_t>(n), to_options(*opts));
        std::memcpy(u_out, r.u.data(), r.u.size() * sizeof(double));
        to_stats(r, stats);
        if (history) {
            const std::size_t c = std::min<std::size_t>(r.history.size(), static_cast<std::size_t>(history_cap));
            for (std::size_t i = 0; i < c; ++i) history[i] = r.history[i];
        }
        return 0;
    } catch (const darcy::InvalidInput&) {
        return 1;
    } catch (...) {
        return 2;
    }
}

// Batch loop with the same scheduling as the reference binding: OpenMP
// dynamic over samples, each solve serial inside, first error reported.
extern "C" int ref_solve_batch(const double* a, const double* rhs, int64_t count, int64_t n,
                               const oracle_opts* opts, double* u_out, oracle_stats* stats) {
    const darcy::SolveOptions so = to_options(*opts);
    const int64_t stride = n * n;
    int rc_all = 0;
#if DARCY_HAS_OPENMP
    const int nt = opts->threads > 0 ? opts->threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic) num_threads(nt)
#endif
    for (int64_t s = 0; s < count; ++s) {
        int rc = 0;
        try {
            const darcy::SolveResult r = darcy::solve(a + s * stride, rhs, static_cast<darcy::index_t>(n), so);
            std::memcpy(u_out + s * stride, r.u.data(), r.u.size() * sizeof(double));
            to_stats(r, stats ? stats + s : nullptr);
        } catch (const darcy::InvalidInput&) {
            rc = 1;
        } catch (...) {
            rc = 2;
        }
        if (rc) {
#if DARCY_HAS_OPENMP
#pragma omp critical
#endif
            if (!rc_all) rc_all = rc;
        }
    }
    return rc_all;
}

extern "C" int ref_apply_operator(const double* a, int64_t n, const double* x, double* y) {
    try {
        darcy::validate_permeability(a, n * n);
        const darcy::Level level(a, static_cast<darcy::index_t>(n));
        level.apply(x, y);
        return 0;
    } catch (...) {
        return 1;
    }
}

extern "C" int ref_apply_preconditioner(const double* a, int64_t n, const double* r, double* z,
                                        int32_t pre_sweeps, int32_t post_sweeps, int32_t cycle_index) {
    try {
        darcy::validate_permeability(a, n * n);
        darcy::MultigridOptions mo;
        mo.pre_sweeps = pre_sweeps;
        mo.post_sweeps = post_sweeps;
        mo.cycle_index = cycle_index;
        const darcy::Multigrid mg(darcy::Level(a, static_cast<darcy::index_t>(n)), mo);
        mg.apply(r, z);
        return 0;
    } catch (...) {
        return 1;
    }
}

extern "C" int ref_apply_preconditioner_opts(const double* a, int64_t n, const double* r, double* z,
                                             const oracle_opts* opts) {
    try {
        darcy::validate_permeability(a, n * n);
        const darcy::Multigrid mg(darcy::Level(a, static_cast<darcy::index_t>(n)), to_options(*opts).multigrid);
        mg.apply(r, z);
        return 0;
    } catch (...) {
        return 1;
    }
}

// Per-level operator arrays, laid out like oracle_hierarchy (e, w, n, s,
// diag, inv_diag per level).  The reference Level exposes only diag and
// inv_diag publicly, so the faces are recovered by applying the level to unit
// vectors: A e_k gives -face at each neighbour.  Returns doubles written.
extern "C" int64_t ref_hierarchy_diag(const double* a, int64_t n, double* out) {
    const darcy::Multigrid mg(darcy::Level(a, static_cast<darcy::index_t>(n)), darcy::MultigridOptions{});
    int64_t w = 0;
    for (int l = 0; l < mg.num_levels(); ++l) {
        const darcy::Level& L = mg.level(l);
        const auto& d = L.diag();
        const auto& id = L.inv_diag();
        if (out) {
            std::memcpy(out + w, d.data(), d.size() * sizeof(double));
            std::memcpy(out + w + static_cast<int64_t>(d.size()), id.data(), id.size() * sizeof(double));
        }
        w += 2 * static_cast<int64_t>(d.size());
    }
    return w;
}

extern "C" int32_t ref_num_levels(int64_t n) {
    std::vector<double> a(static_cast<std::size_t>(n * n), 1.0);
    const darcy::Multigrid mg(darcy::Level(a.data(), static_cast<darcy::index_t>(n)), darcy::MultigridOptions{});
    return mg.num_levels();
}

extern "C" int32_t ref_max_threads(void) { return darcy::max_threads(); }

// The OpenMP thread count seen by this library's parallel regions (the
// reductions inside a solve), set through the runtime this library binds to.
extern "C" void ref_set_num_threads(int32_t threads) {
#if DARCY_HAS_OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}
