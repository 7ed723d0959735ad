"""ctypes access to the CPU oracle and to the reference shim.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py -- never by the product package.

``Oracle`` wraps oracle/lib/liboracle.so (the from-scratch restatement);
``Reference`` wraps oracle/_ref/libdarcy_ref.so (the reference's own headers
compiled by oracle/Makefile), which exists only where it was built.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "lib", "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libdarcy_ref.so")

PC = {"none": 0, "jacobi": 1, "multigrid": 2, "mg": 2}


class Opts(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("max_iter", C.c_int32),
        ("precond", C.c_int32),
        ("pre_sweeps", C.c_int32),
        ("post_sweeps", C.c_int32),
        ("cycle_index", C.c_int32),
        ("record_history", C.c_int32),
        ("cycle_fp32", C.c_int32),
        ("threads", C.c_int32),
        ("min_coarse_n", C.c_int32),
        ("max_levels", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("levels", C.c_int32),
        ("history_len", C.c_int32),
        ("residual", C.c_double),
    ]


@dataclass
class Result:
    u: np.ndarray
    iterations: int
    residual: float
    converged: bool
    levels: int
    history: np.ndarray = field(default_factory=lambda: np.empty(0))


def _dp(x: np.ndarray):
    return x.ctypes.data_as(C.POINTER(C.c_double))


def _opts(tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index, record_history, fp32, threads,
          min_coarse_n=1, max_levels=32):
    return Opts(tol, max_iter, PC[preconditioner], pre_sweeps, post_sweeps, cycle_index,
                int(record_history), int(fp32), threads, min_coarse_n, max_levels)


def _rhs(f, n):
    if np.isscalar(f):
        return np.full((n, n), float(f))
    f = np.ascontiguousarray(f, dtype=np.float64)
    if f.shape != (n, n):
        raise ValueError("f must be a scalar or an array with the same shape as a")
    return f


class Oracle:
    """The CPU restatement (oracle/darcy_oracle.cpp)."""

    prefix = "oracle_"

    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        self.lib = C.CDLL(path)
        L = self.lib
        i64, i32, dp = C.c_int64, C.c_int32, C.POINTER(C.c_double)
        L.oracle_solve.argtypes = [dp, dp, i64, C.POINTER(Opts), dp, dp, C.POINTER(Stats), dp, i32,
                                   C.POINTER(i64)]
        L.oracle_solve_batch.argtypes = [dp, dp, i64, i64, C.POINTER(Opts), dp, C.POINTER(Stats),
                                         C.POINTER(i64)]
        L.oracle_apply_operator.argtypes = [dp, i64, dp, dp, C.POINTER(i64)]
        L.oracle_apply_preconditioner.argtypes = [dp, i64, dp, dp, i32, i32, i32, i32, C.POINTER(i64)]
        L.oracle_apply_preconditioner_opts.argtypes = [dp, i64, dp, dp, C.POINTER(Opts), C.POINTER(i64)]
        L.oracle_num_levels.argtypes = [i64]
        L.oracle_num_levels.restype = i32
        L.oracle_hierarchy.argtypes = [dp, i64, dp]
        L.oracle_hierarchy.restype = i64
        L.oracle_max_threads.restype = i32
        L.oracle_set_num_threads.argtypes = [i32]

    def solve(self, a, f=1.0, *, tol=1e-10, max_iter=10000, preconditioner="multigrid", pre_sweeps=1,
              post_sweeps=1, cycle_index=2, record_history=False, fp32=False, u0=None, min_coarse_n=1,
              max_levels=32) -> Result:
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        rhs = _rhs(f, n)
        u = np.zeros((n, n))
        st = Stats()
        cap = max_iter + 2 if record_history else 0
        hist = np.zeros(max(cap, 1))
        bad = C.c_int64(-1)
        u0p = None
        if u0 is not None:
            u0 = np.ascontiguousarray(u0, dtype=np.float64)
            u0p = _dp(u0)
        o = _opts(tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index, record_history, fp32, 0,
                  min_coarse_n, max_levels)
        rc = self.lib.oracle_solve(_dp(a), _dp(rhs), n, C.byref(o), u0p, _dp(u), C.byref(st),
                                   _dp(hist) if record_history else None, cap, C.byref(bad))
        if rc:
            raise ValueError(f"oracle rejected input (rc={rc}, index={bad.value})")
        return Result(u, st.iterations, st.residual, bool(st.converged), st.levels,
                      hist[: st.history_len].copy() if record_history else np.empty(0))

    def solve_batch(self, a, f=1.0, *, tol=1e-10, max_iter=10000, preconditioner="multigrid", pre_sweeps=1,
                    post_sweeps=1, cycle_index=2, threads=0, fp32=False):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b, n = a.shape[0], a.shape[1]
        rhs = _rhs(f, n)
        u = np.zeros_like(a)
        st = (Stats * b)()
        bad = C.c_int64(-1)
        o = _opts(tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index, False, fp32, threads)
        rc = self.lib.oracle_solve_batch(_dp(a), _dp(rhs), b, n, C.byref(o), _dp(u), st, C.byref(bad))
        if rc:
            raise ValueError(f"oracle rejected input (rc={rc}, index={bad.value})")
        return (u, np.array([s.iterations for s in st]), np.array([s.residual for s in st]),
                np.array([bool(s.converged) for s in st]))

    def apply_operator(self, a, x):
        a = np.ascontiguousarray(a, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        bad = C.c_int64(-1)
        if self.lib.oracle_apply_operator(_dp(a), a.shape[0], _dp(x), _dp(y), C.byref(bad)):
            raise ValueError(f"invalid permeability at {bad.value}")
        return y

    def apply_preconditioner(self, a, r, *, pre_sweeps=1, post_sweeps=1, cycle_index=2, fp32=False,
                             min_coarse_n=1, max_levels=32):
        a = np.ascontiguousarray(a, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        bad = C.c_int64(-1)
        o = _opts(1e-10, 1, "multigrid", pre_sweeps, post_sweeps, cycle_index, False, fp32, 0, min_coarse_n,
                  max_levels)
        if self.lib.oracle_apply_preconditioner_opts(_dp(a), a.shape[0], _dp(r), _dp(z), C.byref(o), C.byref(bad)):
            raise ValueError("invalid input")
        return z

    def num_levels(self, n: int) -> int:
        return int(self.lib.oracle_num_levels(n))

    def hierarchy(self, a):
        """List of dicts (e, w, n, s, diag, inv_diag) per level."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        total = self.lib.oracle_hierarchy(_dp(a), n, None)
        buf = np.zeros(total)
        self.lib.oracle_hierarchy(_dp(a), n, _dp(buf))
        out, off, m = [], 0, n
        for _ in range(self.num_levels(n)):
            sz = m * m
            lv = {}
            for name in ("e", "w", "n", "s", "diag", "inv_diag"):
                lv[name] = buf[off: off + sz].reshape(m, m)
                off += sz
            out.append(lv)
            m = (m + 1) // 2
        return out

    def max_threads(self) -> int:
        return int(self.lib.oracle_max_threads())

    def set_num_threads(self, threads: int) -> None:
        """OpenMP threads of the calling thread (shared by every OpenMP
        library of the process, the reference build included)."""
        self.lib.oracle_set_num_threads(threads)


class Reference:
    """The reference's own headers behind oracle/ref_shim.cpp (oracle/_ref)."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        self.lib = C.CDLL(path)
        L = self.lib
        i64, i32, dp = C.c_int64, C.c_int32, C.POINTER(C.c_double)
        L.ref_solve.argtypes = [dp, dp, i64, C.POINTER(Opts), dp, C.POINTER(Stats), dp, i32]
        L.ref_solve_batch.argtypes = [dp, dp, i64, i64, C.POINTER(Opts), dp, C.POINTER(Stats)]
        L.ref_apply_operator.argtypes = [dp, i64, dp, dp]
        L.ref_apply_preconditioner.argtypes = [dp, i64, dp, dp, i32, i32, i32]
        L.ref_apply_preconditioner_opts.argtypes = [dp, i64, dp, dp, C.POINTER(Opts)]
        L.ref_hierarchy_diag.argtypes = [dp, i64, dp]
        L.ref_hierarchy_diag.restype = i64
        L.ref_num_levels.argtypes = [i64]
        L.ref_num_levels.restype = i32
        L.ref_max_threads.restype = i32
        L.ref_set_num_threads.argtypes = [i32]

    def max_threads(self) -> int:
        return int(self.lib.ref_max_threads())

    def set_num_threads(self, threads: int) -> None:
        """Threads of the reductions inside a solve (the library's own
        OpenMP runtime)."""
        self.lib.ref_set_num_threads(threads)

    def solve(self, a, f=1.0, *, tol=1e-10, max_iter=10000, preconditioner="multigrid", pre_sweeps=1,
              post_sweeps=1, cycle_index=2, record_history=False, min_coarse_n=1, max_levels=32) -> Result:
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        rhs = _rhs(f, n)
        u = np.zeros((n, n))
        st = Stats()
        cap = max_iter + 2 if record_history else 0
        hist = np.zeros(max(cap, 1))
        o = _opts(tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index, record_history, False, 0,
                  min_coarse_n, max_levels)
        rc = self.lib.ref_solve(_dp(a), _dp(rhs), n, C.byref(o), _dp(u), C.byref(st),
                                _dp(hist) if record_history else None, cap)
        if rc:
            raise ValueError(f"reference rejected input (rc={rc})")
        return Result(u, st.iterations, st.residual, bool(st.converged), st.levels,
                      hist[: st.history_len].copy() if record_history else np.empty(0))

    def solve_batch(self, a, f=1.0, *, tol=1e-10, max_iter=10000, preconditioner="multigrid", pre_sweeps=1,
                    post_sweeps=1, cycle_index=2, threads=0):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b, n = a.shape[0], a.shape[1]
        rhs = _rhs(f, n)
        u = np.zeros_like(a)
        st = (Stats * b)()
        o = _opts(tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index, False, False, threads)
        rc = self.lib.ref_solve_batch(_dp(a), _dp(rhs), b, n, C.byref(o), _dp(u), st)
        if rc:
            raise ValueError(f"reference rejected input (rc={rc})")
        return (u, np.array([s.iterations for s in st]), np.array([s.residual for s in st]),
                np.array([bool(s.converged) for s in st]))

    def apply_operator(self, a, x):
        a = np.ascontiguousarray(a, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.zeros_like(x)
        if self.lib.ref_apply_operator(_dp(a), a.shape[0], _dp(x), _dp(y)):
            raise ValueError("invalid input")
        return y

    def apply_preconditioner(self, a, r, *, pre_sweeps=1, post_sweeps=1, cycle_index=2, min_coarse_n=1,
                             max_levels=32):
        a = np.ascontiguousarray(a, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.zeros_like(r)
        o = _opts(1e-10, 1, "multigrid", pre_sweeps, post_sweeps, cycle_index, False, False, 0, min_coarse_n,
                  max_levels)
        if self.lib.ref_apply_preconditioner_opts(_dp(a), a.shape[0], _dp(r), _dp(z), C.byref(o)):
            raise ValueError("invalid input")
        return z

    def hierarchy_diag(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        total = self.lib.ref_hierarchy_diag(_dp(a), n, None)
        buf = np.zeros(total)
        self.lib.ref_hierarchy_diag(_dp(a), n, _dp(buf))
        out, off, m = [], 0, n
        while off < total:
            sz = m * m
            out.append((buf[off: off + sz].reshape(m, m), buf[off + sz: off + 2 * sz].reshape(m, m)))
            off += 2 * sz
            m = (m + 1) // 2
        return out

    def num_levels(self, n: int) -> int:
        return int(self.lib.ref_num_levels(n))


def reference_available() -> bool:
    return os.path.exists(REF_SO)
