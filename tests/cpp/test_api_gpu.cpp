// The C++17 API (include/darcy/solver.hpp) linked against libdarcy_hip.so and
// run on the GPU: the checks of the reference's native suite
// (tests/cpp/test_core.cpp:78-231) that the public API can express, with
// the same dependency-free check / check_close harness style.
//
// Built by darcy-cpp_amd/build.py into tests/cpp/test_api_gpu; run by
// tests/test_gpu_cpp_api.py.  Exit status 0 when every check passes.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

#include "darcy/solver.hpp"

namespace {

int failures = 0;
int checks = 0;

void check(bool ok, const std::string& what) {
    ++checks;
    if (!ok) {
        ++failures;
        std::printf("  FAIL  %s\n", what.c_str());
    }
}

void check_close(double got, double want, double tol, const std::string& what) {
    const double scale = std::max({std::abs(got), std::abs(want), 1.0});
    check(std::abs(got - want) <= tol * scale, what + ": got " + std::to_string(got) + ", want " + std::to_string(want));
}

// Two phases, 3 and 12, drawn cell by cell.
std::vector<double> medium(darcy::index_t n, unsigned seed) {
    std::mt19937 gen(seed);
    std::bernoulli_distribution coin(0.5);
    std::vector<double> a(static_cast<std::size_t>(n * n));
    for (double& v : a) v = coin(gen) ? 12.0 : 3.0;
    return a;
}

std::vector<double> normals(darcy::index_t size, unsigned seed) {
    std::mt19937 gen(seed);
    std::normal_distribution<double> normal;
    std::vector<double> x(static_cast<std::size_t>(size));
    for (double& v : x) v = normal(gen);
    return x;
}

double dot(const std::vector<double>& x, const std::vector<double>& y) {
    double s = 0.0;
    for (std::size_t i = 0; i < x.size(); ++i) s += x[i] * y[i];
    return s;
}

std::vector<double> apply(const std::vector<double>& a, darcy::index_t n, const std::vector<double>& x) {
    std::vector<double> y(x.size());
    darcy::apply_operator(a.data(), n, x.data(), y.data());
    return y;
}

void operator_symmetric_positive_definite() {
    const darcy::index_t n = 12;
    const auto a = medium(n, 1);
    const auto x = normals(n * n, 2), y = normals(n * n, 3);
    check_close(dot(apply(a, n, x), y), dot(x, apply(a, n, y)), 1e-12, "<Ax, y> == <x, Ay>");
    for (unsigned s = 0; s < 10; ++s) {
        const auto v = normals(n * n, 100 + s);
        check(dot(v, apply(a, n, v)) > 0.0, "<Ax, x> > 0");
    }
}

void preconditioner_symmetric() {
    const darcy::index_t n = 24;
    const auto a = medium(n, 10);
    const auto r = normals(n * n, 11), s = normals(n * n, 12);
    for (int cycle : {1, 2}) {
        for (darcy::index_t coarse : {1, 4}) {
            darcy::MultigridOptions mo;
            mo.cycle_index = cycle;
            mo.min_coarse_n = coarse;
            std::vector<double> mr(r.size()), ms(s.size());
            darcy::apply_preconditioner(a.data(), n, r.data(), mr.data(), mo);
            darcy::apply_preconditioner(a.data(), n, s.data(), ms.data(), mo);
            check_close(dot(mr, s), dot(ms, r), 1e-10,
                        "M symmetric, cycle " + std::to_string(cycle) + ", min_coarse_n " + std::to_string(coarse));
        }
    }
}

void solver_converges_and_residual_is_honest() {
    for (darcy::index_t n : {48, 160}) {  // persistent kernel, grid path
        const auto a = medium(n, 13);
        darcy::SolveOptions opts;
        opts.tol = 1e-11;
        const darcy::SolveResult res = darcy::solve_constant_rhs(a.data(), n, 1.0, opts);
        check(res.converged, "converges at n=" + std::to_string(n));
        const double h2 = 1.0 / static_cast<double>((n + 1) * (n + 1));
        const auto au = apply(a, n, res.u);
        double rr = 0.0, bb = 0.0;
        for (std::size_t i = 0; i < au.size(); ++i) {
            rr += (h2 - au[i]) * (h2 - au[i]);
            bb += h2 * h2;
        }
        check(std::sqrt(rr / bb) < 10 * opts.tol, "recomputed residual at n=" + std::to_string(n));
        const int expected = static_cast<int>(std::ceil(std::log2(static_cast<double>(n)))) + 1;
        check(res.levels == expected, "levels = ceil(log2 n) + 1 at n=" + std::to_string(n));
        opts.precision = darcy::Precision::Mixed;
        const darcy::SolveResult mixed = darcy::solve_constant_rhs(a.data(), n, 1.0, opts);
        double du = 0.0, uu = 0.0;
        for (std::size_t i = 0; i < au.size(); ++i) {
            du += (mixed.u[i] - res.u[i]) * (mixed.u[i] - res.u[i]);
            uu += res.u[i] * res.u[i];
        }
        check(mixed.converged && std::sqrt(du / uu) < 1e-9, "FP32 cycle reaches the FP64 solution");
    }
}

void hierarchy_options() {
    const darcy::index_t n = 64;
    const auto a = medium(n, 20);
    darcy::SolveOptions opts;
    opts.multigrid.min_coarse_n = 8;  // 64 -> 32 -> 16 -> 8
    const darcy::SolveResult r = darcy::solve_constant_rhs(a.data(), n, 1.0, opts);
    check(r.converged && r.levels == 4, "min_coarse_n = 8 gives 4 levels");
    opts.multigrid.min_coarse_n = 1;
    opts.multigrid.max_levels = 2;
    const darcy::SolveResult s = darcy::solve_constant_rhs(a.data(), n, 1.0, opts);
    check(s.converged && s.levels == 2, "max_levels = 2 gives 2 levels");
}

void batch_equals_single_solves() {
    const darcy::index_t n = 32, batch = 3;
    std::vector<double> a;
    for (unsigned s = 0; s < batch; ++s) {
        const auto m = medium(n, 30 + s);
        a.insert(a.end(), m.begin(), m.end());
    }
    std::vector<double> u(a.size());
    const darcy::SolveOptions opts;
    const darcy::BatchResult br = darcy::solve_batch(a.data(), batch, n, nullptr, 1.0, opts, u.data());
    for (darcy::index_t s = 0; s < batch; ++s) {
        const darcy::SolveResult one = darcy::solve_constant_rhs(a.data() + s * n * n, n, 1.0, opts);
        check(br.iterations[static_cast<std::size_t>(s)] == one.iterations, "batch iterations = single");
        check(std::equal(one.u.begin(), one.u.end(), u.begin() + s * n * n), "batch u = single u, bit for bit");
    }
}

void invalid_input_throws() {
    const std::vector<double> bad{1.0, 2.0, -3.0, 4.0};
    bool threw = false;
    try {
        darcy::solve_constant_rhs(bad.data(), 2, 1.0, darcy::SolveOptions{});
    } catch (const darcy::InvalidInput& e) {
        threw = std::string(e.what()).find("flat index 2") != std::string::npos;
    }
    check(threw, "non-positive permeability throws InvalidInput naming index 2");
    threw = false;
    try {
        darcy::SolveOptions opts;
        opts.tol = 0.0;
        const std::vector<double> one{1.0};
        darcy::solve_constant_rhs(one.data(), 1, 1.0, opts);
    } catch (const darcy::InvalidInput&) {
        threw = true;
    }
    check(threw, "tol = 0 throws InvalidInput");
}

}  // namespace

int main() {
    operator_symmetric_positive_definite();
    preconditioner_symmetric();
    solver_converges_and_residual_is_honest();
    hierarchy_options();
    batch_equals_single_solves();
    invalid_input_throws();
    std::printf("%d checks, %d failures (%s)\n", checks, failures, darcy::device_info().c_str());
    return failures == 0 ? 0 : 1;
}
