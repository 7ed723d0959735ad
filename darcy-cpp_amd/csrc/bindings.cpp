// pybind11 module darcy._core: the reference's Python boundary
// (src/bindings.cpp:96-271 of the reference) with the same functions,
// keyword names, defaults, result fields and exceptions, backed by the GPU
// solve path.  Additions are keyword arguments with reference-preserving
// defaults (u0=None, precision="fp64") and the uneven batch entry point.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "darcy/solver.hpp"

namespace py = pybind11;
using namespace darcy;

#ifndef DARCY_VERSION
#define DARCY_VERSION "1.0.0"
#endif

namespace {

using Array = py::array_t<double, py::array::c_style | py::array::forcecast>;

Precision precision_from_string(const std::string& p) {
    if (p == "fp64") return Precision::FP64;
    if (p == "mixed") return Precision::Mixed;
    throw InvalidInput("unknown precision '" + p + "'; expected one of: fp64, mixed");
}

// bindings.cpp:33-48 of the reference: same checks, same order.
SolveOptions make_options(double tol, int max_iter, const std::string& preconditioner, int pre_sweeps,
                          int post_sweeps, int cycle_index, bool record_history, const std::string& precision) {
    SolveOptions opts;
    opts.tol = tol;
    opts.max_iter = max_iter;
    opts.preconditioner = preconditioner_from_string(preconditioner);
    opts.multigrid.pre_sweeps = pre_sweeps;
    opts.multigrid.post_sweeps = post_sweeps;
    opts.multigrid.cycle_index = cycle_index;
    opts.record_history = record_history;
    if (pre_sweeps < 1 || post_sweeps < 1) throw InvalidInput("pre_sweeps and post_sweeps must be at least 1");
    if (cycle_index < 1 || cycle_index > 2) throw InvalidInput("cycle_index must be 1 (V-cycle) or 2 (W-cycle)");
    opts.precision = precision_from_string(precision);
    return opts;
}

// The forcing: a scalar stays a scalar (no n x n expansion); an array must
// match a's shape (bindings.cpp:51-63).
struct Forcing {
    std::vector<double> values;
    double scalar = 1.0;
    bool is_array = false;
};

Forcing make_rhs(const py::object& f, index_t n) {
    Forcing r;
    if (py::isinstance<py::array>(f) || py::isinstance<py::list>(f) || py::isinstance<py::tuple>(f)) {
        Array arr = py::cast<Array>(f);
        auto buf = arr.request();
        if (buf.ndim != 2 || buf.shape[0] != n || buf.shape[1] != n)
            throw InvalidInput("f must be a scalar or an array with the same shape as a");
        const auto* p = static_cast<const double*>(buf.ptr);
        r.values.assign(p, p + n * n);
        r.is_array = true;
    } else {
        r.scalar = py::cast<double>(f);
    }
    return r;
}

// A new (n, n) array owning a copy of `data` (pybind11 copies a pointer it
// is given without a base object), so callers may write to it freely.
py::array_t<double> to_numpy(const std::vector<double>& data, index_t n) {
    const py::ssize_t side = static_cast<py::ssize_t>(n);
    return py::array_t<double>({side, side}, data.data());
}

// The compiler of this host module (the kernels are built by hipcc).
std::string host_compiler() {
    std::string v = __VERSION__;
    if (v.find("clang") == std::string::npos && v.find("Clang") == std::string::npos) v = "gcc " + v;
    return v + "; kernels: hipcc for gfx950";
}

index_t require_square_2d(const py::buffer_info& buf, const char* what) {
    if (buf.ndim != 2 || buf.shape[0] != buf.shape[1] || buf.shape[0] < 1)
        throw InvalidInput(std::string(what) + " must be a square 2-D array with at least one cell");
    return static_cast<index_t>(buf.shape[0]);
}

struct PyResult {
    py::array_t<double> u;
    int iterations;
    double residual;
    bool converged;
    int levels;
    std::vector<double> history;
};

std::string describe(const PyResult& r) {
    std::ostringstream os;
    os << "SolveResult(n=" << r.u.shape(0) << ", iterations=" << r.iterations << ", residual=" << std::scientific
       << std::setprecision(3) << r.residual << ", converged=" << (r.converged ? "True" : "False")
       << ", levels=" << r.levels << ")";
    return os.str();
}

}  // namespace

PYBIND11_MODULE(_core, m) {
    m.doc() =
        "darcy._core: the compiled boundary of the GPU solve path (C++17 API over the HIP kernels).\n"
        "The `darcy` package wraps these functions with keyword-only arguments and dataclass results.";

    py::register_exception<InvalidInput>(m, "InvalidInput", PyExc_ValueError);

    py::class_<PyResult>(m, "SolveResult", "Solution u and the convergence record of one solve.")
        .def_readonly("u", &PyResult::u)
        .def_readonly("iterations", &PyResult::iterations)
        .def_readonly("residual", &PyResult::residual)
        .def_readonly("converged", &PyResult::converged)
        .def_readonly("levels", &PyResult::levels)
        .def_readonly("history", &PyResult::history)
        .def("__repr__", &describe);

    m.def(
        "solve",
        [](Array a, py::object f, double tol, int max_iter, const std::string& preconditioner, int pre_sweeps,
           int post_sweeps, int cycle_index, bool record_history, py::object u0, const std::string& precision) {
            auto buf = a.request();
            const index_t n = require_square_2d(buf, "a");
            const SolveOptions opts = make_options(tol, max_iter, preconditioner, pre_sweeps, post_sweeps,
                                                   cycle_index, record_history, precision);
            const Forcing rhs = make_rhs(f, n);
            std::vector<double> guess;
            if (!u0.is_none()) {
                Array g = py::cast<Array>(u0);
                auto gb = g.request();
                if (gb.ndim != 2 || gb.shape[0] != n || gb.shape[1] != n)
                    throw InvalidInput("u0 must have the same shape as a");
                const auto* gp = static_cast<const double*>(gb.ptr);
                guess.assign(gp, gp + n * n);
            }
            const auto* a_ptr = static_cast<const double*>(buf.ptr);
            SolveResult r;
            {
                py::gil_scoped_release release;
                if (rhs.is_array)
                    r = solve(a_ptr, rhs.values.data(), n, opts, guess.empty() ? nullptr : guess.data());
                else if (guess.empty())
                    r = solve_constant_rhs(a_ptr, n, rhs.scalar, opts);
                else {
                    const std::vector<double> full(static_cast<std::size_t>(n * n), rhs.scalar);
                    r = solve(a_ptr, full.data(), n, opts, guess.data());
                }
            }
            return PyResult{to_numpy(r.u, n), r.iterations, r.residual, r.converged, r.levels, r.history};
        },
        py::arg("a"), py::arg("f") = 1.0, py::arg("tol") = 1e-10, py::arg("max_iter") = 10000,
        py::arg("preconditioner") = "multigrid", py::arg("pre_sweeps") = 1, py::arg("post_sweeps") = 1,
        py::arg("cycle_index") = 2, py::arg("record_history") = false, py::arg("u0") = py::none(),
        py::arg("precision") = "fp64",
        "u with -div(a grad u) = f in the unit square and u = 0 on its boundary, on the GPU.");

    m.def(
        "solve_batch",
        [](Array a, py::object f, double tol, int max_iter, const std::string& preconditioner, int pre_sweeps,
           int post_sweeps, int cycle_index, int threads, const std::string& precision) {
            (void)threads;  // the batch runs as one GPU launch; accepted for compatibility
            auto buf = a.request();
            if (buf.ndim != 3 || buf.shape[1] != buf.shape[2] || buf.shape[1] < 1)
                throw InvalidInput("a must have shape (batch, n, n)");
            const auto batch = static_cast<index_t>(buf.shape[0]);
            const auto n = static_cast<index_t>(buf.shape[1]);
            const SolveOptions opts =
                make_options(tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index, false, precision);
            const Forcing rhs = make_rhs(f, n);
            const auto* a_ptr = static_cast<const double*>(buf.ptr);
            py::array_t<double> out(std::vector<py::ssize_t>{static_cast<py::ssize_t>(batch),
                                                             static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(n)});
            auto* out_ptr = static_cast<double*>(out.request().ptr);
            BatchResult r;
            {
                py::gil_scoped_release release;
                r = solve_batch(a_ptr, batch, n, rhs.is_array ? rhs.values.data() : nullptr, rhs.scalar, opts,
                                out_ptr);
            }
            return py::make_tuple(out, r.iterations, r.residuals, r.converged);
        },
        py::arg("a"), py::arg("f") = 1.0, py::arg("tol") = 1e-10, py::arg("max_iter") = 10000,
        py::arg("preconditioner") = "multigrid", py::arg("pre_sweeps") = 1, py::arg("post_sweeps") = 1,
        py::arg("cycle_index") = 2, py::arg("threads") = 0, py::arg("precision") = "fp64",
        "(u, iterations, residuals, converged) of a (batch, n, n) stack, one persistent GPU launch.");

    m.def(
        "solve_batch_uneven",
        [](py::list arrays, double f, double tol, int max_iter, const std::string& preconditioner, int pre_sweeps,
           int post_sweeps, int cycle_index, const std::string& precision) {
            const SolveOptions opts =
                make_options(tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index, false, precision);
            std::vector<Array> keep;
            std::vector<const double*> ap;
            std::vector<index_t> ns;
            for (py::handle h : arrays) {
                keep.push_back(py::cast<Array>(h));
                auto b = keep.back().request();
                ns.push_back(require_square_2d(b, "a"));
                ap.push_back(static_cast<const double*>(b.ptr));
            }
            std::vector<py::array_t<double>> outs;
            std::vector<double*> up;
            for (index_t n : ns) {
                outs.emplace_back(std::vector<py::ssize_t>{static_cast<py::ssize_t>(n), static_cast<py::ssize_t>(n)});
                up.push_back(static_cast<double*>(outs.back().request().ptr));
            }
            BatchResult r;
            {
                py::gil_scoped_release release;
                r = solve_batch_uneven(ap, ns, {}, f, opts, up);
            }
            return py::make_tuple(outs, r.iterations, r.residuals, r.converged, r.levels);
        },
        py::arg("a"), py::arg("f") = 1.0, py::arg("tol") = 1e-10, py::arg("max_iter") = 10000,
        py::arg("preconditioner") = "multigrid", py::arg("pre_sweeps") = 1, py::arg("post_sweeps") = 1,
        py::arg("cycle_index") = 2, py::arg("precision") = "fp64",
        "A list of square media of any sides, solved together in one persistent GPU launch.");

    m.def(
        "apply_operator",
        [](Array a, Array x) {
            auto abuf = a.request();
            const index_t n = require_square_2d(abuf, "a");
            auto xbuf = x.request();
            if (require_square_2d(xbuf, "x") != n) throw InvalidInput("x must have the same shape as a");
            validate_permeability(static_cast<const double*>(abuf.ptr), n * n);
            std::vector<double> y(static_cast<std::size_t>(n * n));
            {
                py::gil_scoped_release release;
                apply_operator(static_cast<const double*>(abuf.ptr), n, static_cast<const double*>(xbuf.ptr),
                               y.data());
            }
            return to_numpy(y, n);
        },
        py::arg("a"), py::arg("x"), "y = A x for the harmonic-face finite-volume operator of a, computed on the GPU.");

    m.def(
        "apply_preconditioner",
        [](Array a, Array r, int pre_sweeps, int post_sweeps, int cycle_index, const std::string& precision) {
            auto abuf = a.request();
            const index_t n = require_square_2d(abuf, "a");
            auto rbuf = r.request();
            if (require_square_2d(rbuf, "r") != n) throw InvalidInput("r must have the same shape as a");
            validate_permeability(static_cast<const double*>(abuf.ptr), n * n);
            MultigridOptions mo;
            mo.pre_sweeps = pre_sweeps;
            mo.post_sweeps = post_sweeps;
            mo.cycle_index = cycle_index;
            if (pre_sweeps < 1 || post_sweeps < 1) throw InvalidInput("sweeps must be at least 1");
            if (cycle_index < 1 || cycle_index > 2) throw InvalidInput("cycle_index must be 1 or 2");
            const Precision prec = precision_from_string(precision);
            std::vector<double> z(static_cast<std::size_t>(n * n));
            {
                py::gil_scoped_release release;
                apply_preconditioner(static_cast<const double*>(abuf.ptr), n, static_cast<const double*>(rbuf.ptr),
                                     z.data(), mo, prec);
            }
            return to_numpy(z, n);
        },
        py::arg("a"), py::arg("r"), py::arg("pre_sweeps") = 1, py::arg("post_sweeps") = 1,
        py::arg("cycle_index") = 2, py::arg("precision") = "fp64",
        "z = M r: one multigrid cycle from z = 0 on the GPU (the CG preconditioner).");

    m.def("num_threads", &max_threads, "Host threads available to the library.");
    m.def("has_openmp", &has_openmp, "Whether the host layer was compiled with OpenMP.");
    m.def("set_num_threads", &set_num_threads, py::arg("threads"), "Set the host thread count.");
    m.def("device_info", &device_info, "Description of the GPU the solver runs on.");
    m.def("last_kernel_ms", &last_kernel_ms, "HIP-event time of the last solve's kernels (ms).");
    m.def("set_tuning", &set_tuning, py::arg("key"), py::arg("value"),
          "Set a launch-shape tuning value of the process context (results do not depend on it).");
    m.def("get_tuning", &get_tuning, py::arg("key"), "Current launch-shape tuning value.");
    m.def("num_levels", [](index_t n) {
        int l = 1;  // multigrid.hpp:55-71 of the reference
        while (n > 1 && l < 32) {
            n = (n + 1) / 2;
            ++l;
        }
        return l;
    });

    m.attr("__version__") = DARCY_VERSION;
    m.attr("__compiler__") = host_compiler();
}
