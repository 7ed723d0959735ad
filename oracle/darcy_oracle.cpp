// CPU oracle for the darcy solve path -- TEST INFRASTRUCTURE ONLY.
//
// A from-scratch C++17/OpenMP restatement of the reference algorithm, used as
// the checker for the HIP implementation and as the CPU baseline of the
// benchmark.  The product library never links, loads or calls this file.
//
// Floating-point operations are performed in the same order as the reference
// so that, built with the same flags, the results agree bit for bit:
//   fine operator ........ include/darcy/level.hpp:79-97
//   apply / residual ..... level.hpp:105-137
//   red-black GS sweep ... level.hpp:145-161
//   coarse size .......... level.hpp:164
//   Galerkin coarsening .. level.hpp:171-211
//   restrict / prolong ... level.hpp:214-237
//   hierarchy + cycle .... include/darcy/multigrid.hpp:55-146
//   BLAS-1 ............... include/darcy/blas.hpp:27-84
//   PCG driver ........... include/darcy/solver.hpp:68-174
//   batch loop ........... src/bindings.cpp:140-197
// Two additions have no reference counterpart: a warm start u0 and an FP32
// cycle inside an FP64 flexible CG (beta = -alpha <z_new, A p> / <z, r>,
// Notay's FCG(1) with r_new - r_old = -alpha A p).
#include "darcy_oracle.h"

#include <cmath>
#include <cstring>
#include <exception>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using idx = std::ptrdiff_t;

// Kernels run serially below this many cells (level.hpp:74, blas.hpp:225).
constexpr idx kSerialBelow = 4096;
// Symmetric sweeps of the coarsest solve when it is larger than 1x1
// (multigrid.hpp:146).
constexpr int kCoarsestSweeps = 8;
// Hierarchy depth cap (multigrid.hpp:46).
constexpr int kMaxLevels = 32;

#define ORACLE_PRAGMA(x) _Pragma(#x)
#ifdef _OPENMP
#define ORACLE_OMP(x) ORACLE_PRAGMA(omp x)
#else
#define ORACLE_OMP(x)
#endif

// One level of the hierarchy: six face-wise arrays, like level.hpp:253.
template <class T>
struct Op {
    idx n = 0;
    std::vector<T> e, w, no, so, dg, inv;
    bool par() const { return n * n >= kSerialBelow; }
    void resize(idx m) {
        n = m;
        const auto c = static_cast<std::size_t>(m * m);
        e.resize(c); w.resize(c); no.resize(c); so.resize(c); dg.resize(c); inv.resize(c);
    }
};

inline double hmean(double x, double y) { return 2.0 * x * y / (x + y); }

// level.hpp:79-97
Op<double> fine_operator(const double* a, idx n) {
    Op<double> L;
    L.resize(n);
    const bool par = L.par();
    ORACLE_OMP(parallel for schedule(static) if (par))
    for (idx i = 0; i < n; ++i) {
        for (idx j = 0; j < n; ++j) {
            const idx k = i * n + j;
            const double c = a[k];
            const double fe = (j + 1 < n) ? hmean(c, a[k + 1]) : c;
            const double fw = (j > 0) ? hmean(c, a[k - 1]) : c;
            const double fn = (i + 1 < n) ? hmean(c, a[k + n]) : c;
            const double fs = (i > 0) ? hmean(c, a[k - n]) : c;
            L.e[k] = fe;
            L.w[k] = fw;
            L.no[k] = fn;
            L.so[k] = fs;
            const double d = fe + fw + fn + fs;
            L.dg[k] = d;
            L.inv[k] = 1.0 / d;
        }
    }
    return L;
}

// level.hpp:164
inline idx half_up(idx n) { return (n + 1) / 2; }

// level.hpp:171-211: coarse faces are sums of the fine faces on the
// aggregate's edges; odd n leaves a width-one last aggregate.
Op<double> galerkin(const Op<double>& F) {
    const idx n = F.n, m = half_up(n);
    Op<double> C;
    C.resize(m);
    const bool par = F.par();
    ORACLE_OMP(parallel for schedule(static) if (par))
    for (idx I = 0; I < m; ++I) {
        const idx r0 = 2 * I, r1 = (2 * I + 1 < n) ? 2 * I + 1 : 2 * I;
        for (idx J = 0; J < m; ++J) {
            const idx c0 = 2 * J, c1 = (2 * J + 1 < n) ? 2 * J + 1 : 2 * J;
            double se = 0.0, sw = 0.0, sn = 0.0, ss = 0.0;
            for (idx i = r0; i <= r1; ++i) {
                se += F.e[i * n + c1];
                sw += F.w[i * n + c0];
            }
            for (idx j = c0; j <= c1; ++j) {
                sn += F.no[r1 * n + j];
                ss += F.so[r0 * n + j];
            }
            const idx k = I * m + J;
            C.e[k] = se;
            C.w[k] = sw;
            C.no[k] = sn;
            C.so[k] = ss;
            const double d = se + sw + sn + ss;
            C.dg[k] = d;
            C.inv[k] = 1.0 / d;
        }
    }
    return C;
}

template <class T>
Op<T> convert(const Op<double>& D) {
    Op<T> L;
    L.resize(D.n);
    for (std::size_t k = 0; k < D.e.size(); ++k) {
        L.e[k] = static_cast<T>(D.e[k]);
        L.w[k] = static_cast<T>(D.w[k]);
        L.no[k] = static_cast<T>(D.no[k]);
        L.so[k] = static_cast<T>(D.so[k]);
        // The cycle's diagonal is summed in T from the rounded faces, in the
        // order of level.hpp:92, and its reciprocal taken in T.
        L.dg[k] = ((L.e[k] + L.w[k]) + L.no[k]) + L.so[k];
        L.inv[k] = T(1) / L.dg[k];
    }
    return L;
}

// y = A x  (level.hpp:105-119); with b != nullptr, y = b - A x (level.hpp:122-137).
template <class T>
void stencil(const Op<T>& L, const T* x, const T* b, T* y) {
    const idx n = L.n;
    const bool par = L.par();
    ORACLE_OMP(parallel for schedule(static) if (par))
    for (idx i = 0; i < n; ++i) {
        for (idx j = 0; j < n; ++j) {
            const idx k = i * n + j;
            T v = L.dg[k] * x[k];
            if (j + 1 < n) v -= L.e[k] * x[k + 1];
            if (j > 0) v -= L.w[k] * x[k - 1];
            if (i + 1 < n) v -= L.no[k] * x[k + n];
            if (i > 0) v -= L.so[k] * x[k - n];
            y[k] = b ? b[k] - v : v;
        }
    }
}

// level.hpp:145-161: cells with (i + j) % 2 == colour.
template <class T>
void gs_colour(const Op<T>& L, T* x, const T* b, int colour) {
    const idx n = L.n;
    const bool par = L.par();
    ORACLE_OMP(parallel for schedule(static) if (par))
    for (idx i = 0; i < n; ++i) {
        for (idx j = (i + colour) & 1; j < n; j += 2) {
            const idx k = i * n + j;
            T v = b[k];
            if (j + 1 < n) v += L.e[k] * x[k + 1];
            if (j > 0) v += L.w[k] * x[k - 1];
            if (i + 1 < n) v += L.no[k] * x[k + n];
            if (i > 0) v += L.so[k] * x[k - n];
            x[k] = v * L.inv[k];
        }
    }
}

// level.hpp:214-227
template <class T>
void restrict_sum(idx n, const T* f, T* c) {
    const idx m = half_up(n);
    const bool par = n * n >= kSerialBelow;
    ORACLE_OMP(parallel for schedule(static) if (par))
    for (idx I = 0; I < m; ++I) {
        const idx r1 = (2 * I + 1 < n) ? 2 * I + 1 : 2 * I;
        for (idx J = 0; J < m; ++J) {
            const idx c1 = (2 * J + 1 < n) ? 2 * J + 1 : 2 * J;
            T s = 0;
            for (idx i = 2 * I; i <= r1; ++i)
                for (idx j = 2 * J; j <= c1; ++j) s += f[i * n + j];
            c[I * m + J] = s;
        }
    }
}

// level.hpp:230-237
template <class T>
void prolong_inject(idx n, const T* c, T* f) {
    const idx m = half_up(n);
    const bool par = n * n >= kSerialBelow;
    ORACLE_OMP(parallel for schedule(static) if (par))
    for (idx i = 0; i < n; ++i)
        for (idx j = 0; j < n; ++j) f[i * n + j] += c[(i / 2) * m + j / 2];
}

// --- BLAS-1 (blas.hpp) --------------------------------------------------
double vdot(const std::vector<double>& x, const std::vector<double>& y) {
    const double* px = x.data();
    const double* py = y.data();
    const idx n = static_cast<idx>(x.size());
    double s = 0.0;
    ORACLE_OMP(parallel for simd reduction(+ : s) schedule(static) if (n >= kSerialBelow))
    for (idx i = 0; i < n; ++i) s += px[i] * py[i];
    return s;
}

void vxpby(const std::vector<double>& x, double beta, std::vector<double>& y) {
    const double* px = x.data();
    double* py = y.data();
    const idx n = static_cast<idx>(x.size());
    ORACLE_OMP(parallel for simd schedule(static) if (n >= kSerialBelow))
    for (idx i = 0; i < n; ++i) py[i] = px[i] + beta * py[i];
}

double vcg_update(double alpha, const std::vector<double>& p, const std::vector<double>& ap,
                  std::vector<double>& u, std::vector<double>& r) {
    const double* pp = p.data();
    const double* pap = ap.data();
    double* pu = u.data();
    double* pr = r.data();
    const idx n = static_cast<idx>(u.size());
    double rr = 0.0;
    ORACLE_OMP(parallel for simd reduction(+ : rr) schedule(static) if (n >= kSerialBelow))
    for (idx i = 0; i < n; ++i) {
        pu[i] += alpha * pp[i];
        const double ri = pr[i] - alpha * pap[i];
        pr[i] = ri;
        rr += ri * ri;
    }
    return rr;
}

template <class T>
void vzero(std::vector<T>& x) {
    T* px = x.data();
    const idx n = static_cast<idx>(x.size());
    ORACLE_OMP(parallel for simd schedule(static) if (n >= kSerialBelow))
    for (idx i = 0; i < n; ++i) px[i] = T(0);
}

// --- hierarchy and cycle (multigrid.hpp) --------------------------------
// multigrid.hpp:55-63: coarsen while fewer than max_levels levels exist and
// the last one is wider than min_coarse_n and than 1.
std::vector<Op<double>> build_hierarchy(Op<double> fine, int64_t min_coarse_n = 1, int max_levels = kMaxLevels) {
    std::vector<Op<double>> lv;
    lv.push_back(std::move(fine));
    while (static_cast<int>(lv.size()) < max_levels) {
        if (lv.back().n <= min_coarse_n || lv.back().n <= 1) break;
        lv.push_back(galerkin(lv.back()));
    }
    return lv;
}

template <class T>
struct Cycle {
    std::vector<Op<T>> ops;
    std::vector<std::vector<T>> x, b, r;
    int pre = 1, post = 1, gamma = 2;

    void init(int pre_s, int post_s, int cyc) {
        pre = pre_s; post = post_s; gamma = cyc;
        for (const auto& L : ops) {
            const auto m = static_cast<std::size_t>(L.n * L.n);
            x.emplace_back(m);
            b.emplace_back(m);
            r.emplace_back(m);
        }
    }

    void coarsest(std::size_t l) {
        const Op<T>& L = ops[l];
        if (L.n == 1) {
            x[l][0] = b[l][0] * L.inv[0];
            return;
        }
        for (int s = 0; s < kCoarsestSweeps; ++s) {
            gs_colour(L, x[l].data(), b[l].data(), 0);
            gs_colour(L, x[l].data(), b[l].data(), 1);
        }
        for (int s = 0; s < kCoarsestSweeps; ++s) {
            gs_colour(L, x[l].data(), b[l].data(), 1);
            gs_colour(L, x[l].data(), b[l].data(), 0);
        }
    }

    // Dense coarse tail of the FP32 cycle (no reference counterpart).  From
    // level `tail` -- the first level below the finest with side <= 8 -- the
    // recursive cycle is a linear map of b (x fresh) and, on a repeated
    // W-cycle visit, x + M (b - A x).  Its matrix M is built once in FP64 by
    // running the reference's recursive cycle on unit vectors, rounded to T,
    // and applied with a sequential row sum.
    std::size_t tail = SIZE_MAX;
    std::size_t tdim = 0;
    std::vector<T> M;  // tdim x tdim, M[i * tdim + k]

    void build_tail(const std::vector<Op<double>>& hier) {
        for (std::size_t l = 1; l < hier.size(); ++l)
            if (hier[l].n <= 8) {
                tail = l;
                break;
            }
        if (tail == SIZE_MAX) return;
        Cycle<double> c;
        c.ops.assign(hier.begin() + static_cast<std::ptrdiff_t>(tail), hier.end());
        c.init(pre, post, gamma);
        tdim = static_cast<std::size_t>(hier[tail].n * hier[tail].n);
        M.assign(tdim * tdim, T(0));
        for (std::size_t k = 0; k < tdim; ++k) {
            for (std::size_t i = 0; i < tdim; ++i) c.b[0][i] = i == k ? 1.0 : 0.0;
            vzero(c.x[0]);
            c.visit(0, true);
            for (std::size_t i = 0; i < tdim; ++i) M[i * tdim + k] = static_cast<T>(c.x[0][i]);
        }
    }

    void dense_tail(std::size_t l, bool fresh) {
        const T* rhs = b[l].data();
        if (!fresh) {
            stencil(ops[l], x[l].data(), b[l].data(), r[l].data());
            rhs = r[l].data();
        }
        for (std::size_t i = 0; i < tdim; ++i) {
            T s = T(0);
            for (std::size_t k = 0; k < tdim; ++k) s += M[i * tdim + k] * rhs[k];
            x[l][i] = fresh ? s : x[l][i] + s;
        }
    }

    // `fresh`: x_l is zero on entry (the first of the gamma coarse visits).
    void visit(std::size_t l, bool fresh) {
        const Op<T>& L = ops[l];
        if (l == tail) {
            dense_tail(l, fresh);
            return;
        }
        if (l + 1 == ops.size()) {
            coarsest(l);
            return;
        }
        for (int s = 0; s < pre; ++s) {
            gs_colour(L, x[l].data(), b[l].data(), 0);
            gs_colour(L, x[l].data(), b[l].data(), 1);
        }
        stencil(L, x[l].data(), b[l].data(), r[l].data());
        restrict_sum(L.n, r[l].data(), b[l + 1].data());
        vzero(x[l + 1]);
        for (int g = 0; g < gamma; ++g) visit(l + 1, g == 0);
        prolong_inject(L.n, x[l + 1].data(), x[l].data());
        for (int s = 0; s < post; ++s) {
            gs_colour(L, x[l].data(), b[l].data(), 1);
            gs_colour(L, x[l].data(), b[l].data(), 0);
        }
    }

    // z = M r from x = 0 (multigrid.hpp:88-93); the FP32 variant converts at
    // the boundary of the cycle.
    void apply(const double* rin, double* z) {
        const std::size_t m = x[0].size();
        for (std::size_t k = 0; k < m; ++k) b[0][k] = static_cast<T>(rin[k]);
        vzero(x[0]);
        visit(0, true);
        for (std::size_t k = 0; k < m; ++k) z[k] = static_cast<double>(x[0][k]);
    }
};

// solver.hpp:68-77: the first offending index.
idx first_bad(const double* a, idx size) {
    for (idx i = 0; i < size; ++i)
        if (!(a[i] > 0.0) || !std::isfinite(a[i])) return i;
    return -1;
}

struct Outcome {
    int iterations = 0;
    bool converged = false;
    int levels = 1;
    double residual = 0.0;
    std::vector<double> history;
};

// solver.hpp:83-174 (+ warm start and flexible FP32-cycle variants).
int solve_one(const double* a, const double* rhs, idx n, const oracle_opts& o, const double* u0,
              double* u_out, Outcome& out, int64_t* bad) {
    if (n < 1 || !(o.tol > 0.0) || o.max_iter < 1) return 1;
    const idx size = n * n;
    const idx b_at = first_bad(a, size);
    if (b_at >= 0) {
        if (bad) *bad = b_at;
        return 1;
    }
    const double h = 1.0 / static_cast<double>(n + 1);
    const double h2 = h * h;
    Op<double> fine = fine_operator(a, n);

    const auto m = static_cast<std::size_t>(size);
    std::vector<double> u(m, 0.0), b(m);
    for (std::size_t k = 0; k < m; ++k) b[k] = rhs[k] * h2;
    const double b_norm = std::sqrt(vdot(b, b));
    if (b_norm == 0.0) {
        out.converged = true;
        std::memset(u_out, 0, m * sizeof(double));
        return 0;
    }

    const bool mg = o.precond == ORACLE_PC_MULTIGRID;
    const bool fp32 = mg && o.cycle_fp32;
    Cycle<double> c64;
    Cycle<float> c32;
    std::vector<Op<double>> hier;
    if (mg) {
        hier = build_hierarchy(std::move(fine), o.min_coarse_n, o.max_levels);
        out.levels = static_cast<int>(hier.size());
        if (fp32) {
            for (const auto& L : hier) c32.ops.push_back(convert<float>(L));
            c32.init(o.pre_sweeps, o.post_sweeps, o.cycle_index);
            c32.build_tail(hier);
        }
    } else {
        hier.push_back(std::move(fine));
    }
    const Op<double>& A = hier[0];
    if (mg && !fp32) {
        c64.ops = hier;
        c64.init(o.pre_sweeps, o.post_sweeps, o.cycle_index);
    }

    std::vector<double> r = b, z(m), p, ap(m);
    double res0 = 1.0;
    if (u0) {
        for (std::size_t k = 0; k < m; ++k) u[k] = u0[k];
        stencil(A, u.data(), b.data(), r.data());
        res0 = std::sqrt(vdot(r, r)) / b_norm;
    }
    auto precondition = [&](const std::vector<double>& in, std::vector<double>& to) {
        if (o.precond == ORACLE_PC_NONE) {
            to = in;
        } else if (o.precond == ORACLE_PC_JACOBI) {
            for (std::size_t k = 0; k < m; ++k) to[k] = in[k] * A.inv[k];
        } else if (fp32) {
            c32.apply(in.data(), to.data());
        } else {
            c64.apply(in.data(), to.data());
        }
    };

    out.residual = res0;
    if (o.record_history) out.history.push_back(res0);
    if (u0 && res0 < o.tol) {
        out.converged = true;
        std::memcpy(u_out, u.data(), m * sizeof(double));
        return 0;
    }
    precondition(r, z);
    p = z;
    double rz = vdot(r, z);

    for (int it = 0; it < o.max_iter; ++it) {
        stencil(A, p.data(), static_cast<const double*>(nullptr), ap.data());
        const double pap = vdot(p, ap);
        if (!(std::abs(pap) > 0.0) || !std::isfinite(pap)) {
            out.iterations = it;
            out.converged = false;
            std::memcpy(u_out, u.data(), m * sizeof(double));
            return 0;
        }
        const double alpha = rz / pap;
        const double rr = vcg_update(alpha, p, ap, u, r);
        out.residual = std::sqrt(rr) / b_norm;
        out.iterations = it + 1;
        if (o.record_history) out.history.push_back(out.residual);
        if (out.residual < o.tol) {
            out.converged = true;
            break;
        }
        precondition(r, z);
        const double rz_new = vdot(r, z);
        double beta;
        if (fp32) {
            beta = (-alpha * vdot(z, ap)) / rz;
        } else {
            beta = rz_new / rz;
        }
        rz = rz_new;
        vxpby(z, beta, p);
    }
    std::memcpy(u_out, u.data(), m * sizeof(double));
    return 0;
}

void fill_stats(const Outcome& oc, oracle_stats* st) {
    if (!st) return;
    st->iterations = oc.iterations;
    st->converged = oc.converged ? 1 : 0;
    st->levels = oc.levels;
    st->history_len = static_cast<int32_t>(oc.history.size());
    st->residual = oc.residual;
}

}  // namespace

extern "C" int oracle_solve(const double* a, const double* rhs, int64_t n, const oracle_opts* opts,
                            const double* u0, double* u_out, oracle_stats* stats, double* history,
                            int32_t history_cap, int64_t* bad_index) {
    Outcome oc;
    const int rc = solve_one(a, rhs, static_cast<idx>(n), *opts, u0, u_out, oc, bad_index);
    if (rc) return rc;
    fill_stats(oc, stats);
    if (history) {
        const std::size_t c = std::min<std::size_t>(oc.history.size(), static_cast<std::size_t>(history_cap));
        for (std::size_t i = 0; i < c; ++i) history[i] = oc.history[i];
    }
    return 0;
}

extern "C" int oracle_solve_batch(const double* a, const double* rhs, int64_t count, int64_t n,
                                  const oracle_opts* opts, double* u_out, oracle_stats* stats,
                                  int64_t* bad_index) {
    const idx stride = static_cast<idx>(n * n);
    int first_rc = 0;
    int64_t first_sample = -1;
    int64_t first_bad_at = -1;
#ifdef _OPENMP
    const int nt = opts->threads > 0 ? opts->threads : omp_get_max_threads();
#endif
    ORACLE_OMP(parallel for schedule(dynamic) num_threads(nt))
    for (int64_t s = 0; s < count; ++s) {
        Outcome oc;
        int64_t bad = -1;
        const int rc = solve_one(a + s * stride, rhs, static_cast<idx>(n), *opts, nullptr,
                                 u_out + s * stride, oc, &bad);
        if (rc) {
            ORACLE_OMP(critical)
            if (first_sample < 0 || s < first_sample) {
                first_sample = s;
                first_rc = rc;
                first_bad_at = bad;
            }
        } else {
            fill_stats(oc, stats ? stats + s : nullptr);
        }
    }
    if (first_rc) {
        if (bad_index) *bad_index = first_bad_at;
        return first_rc;
    }
    return 0;
}

extern "C" int oracle_apply_operator(const double* a, int64_t n, const double* x, double* y,
                                     int64_t* bad_index) {
    const idx b_at = first_bad(a, n * n);
    if (b_at >= 0) {
        if (bad_index) *bad_index = b_at;
        return 1;
    }
    const Op<double> L = fine_operator(a, static_cast<idx>(n));
    stencil(L, x, static_cast<const double*>(nullptr), y);
    return 0;
}

namespace {

int apply_cycle(const double* a, int64_t n, const double* r, double* z, int32_t pre_sweeps, int32_t post_sweeps,
                int32_t cycle_index, int32_t cycle_fp32, int64_t min_coarse_n, int32_t max_levels,
                int64_t* bad_index) {
    const idx b_at = first_bad(a, n * n);
    if (b_at >= 0) {
        if (bad_index) *bad_index = b_at;
        return 1;
    }
    if (pre_sweeps < 1 || post_sweeps < 1 || cycle_index < 1 || cycle_index > 2) return 1;
    auto hier = build_hierarchy(fine_operator(a, static_cast<idx>(n)), min_coarse_n, max_levels);
    if (cycle_fp32) {
        Cycle<float> c;
        for (const auto& L : hier) c.ops.push_back(convert<float>(L));
        c.init(pre_sweeps, post_sweeps, cycle_index);
        c.build_tail(hier);
        c.apply(r, z);
    } else {
        Cycle<double> c;
        c.ops = std::move(hier);
        c.init(pre_sweeps, post_sweeps, cycle_index);
        c.apply(r, z);
    }
    return 0;
}

}  // namespace

extern "C" int oracle_apply_preconditioner(const double* a, int64_t n, const double* r, double* z,
                                           int32_t pre_sweeps, int32_t post_sweeps, int32_t cycle_index,
                                           int32_t cycle_fp32, int64_t* bad_index) {
    return apply_cycle(a, n, r, z, pre_sweeps, post_sweeps, cycle_index, cycle_fp32, 1, kMaxLevels, bad_index);
}

extern "C" int oracle_apply_preconditioner_opts(const double* a, int64_t n, const double* r, double* z,
                                                const oracle_opts* o, int64_t* bad_index) {
    return apply_cycle(a, n, r, z, o->pre_sweeps, o->post_sweeps, o->cycle_index, o->cycle_fp32, o->min_coarse_n,
                       o->max_levels, bad_index);
}

extern "C" int32_t oracle_num_levels(int64_t n) {
    int32_t l = 1;
    while (n > 1 && l < kMaxLevels) {
        n = half_up(n);
        ++l;
    }
    return l;
}

extern "C" int64_t oracle_hierarchy(const double* a, int64_t n, double* out) {
    int64_t total = 0;
    {
        int64_t m = n;
        for (int32_t l = 0; l < oracle_num_levels(n); ++l) {
            total += 6 * m * m;
            m = half_up(m);
        }
    }
    if (!out) return total;
    const auto hier = build_hierarchy(fine_operator(a, static_cast<idx>(n)));
    double* o = out;
    for (const auto& L : hier) {
        for (const auto* v : {&L.e, &L.w, &L.no, &L.so, &L.dg, &L.inv}) {
            std::memcpy(o, v->data(), v->size() * sizeof(double));
            o += v->size();
        }
    }
    return total;
}

extern "C" int32_t oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

extern "C" void oracle_set_num_threads(int32_t threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}
