"""ctypes binding of the C ABI (include/darcy_hip.h).

The pybind11 module covers the reference's API on host arrays.  This binding
exposes the rest of the boundary -- device-resident inputs, uneven batches
given as device pointer tables, kernel timing -- for callers that keep data
in HBM (the benchmark, data generation).
"""

from __future__ import annotations

import ctypes as C
import os
from typing import Sequence

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdarcy_hip.so")

OK, INVALID_INPUT, HIP_ERROR, OUT_OF_MEMORY = 0, 1, 2, 3
PC = {"none": 0, "jacobi": 1, "multigrid": 2, "mg": 2}
PRECISION = {"fp64": 0, "mixed": 1}


class Opts(C.Structure):
    _fields_ = [
        ("tol", C.c_double),
        ("max_iter", C.c_int32),
        ("preconditioner", C.c_int32),
        ("pre_sweeps", C.c_int32),
        ("post_sweeps", C.c_int32),
        ("cycle_index", C.c_int32),
        ("record_history", C.c_int32),
        ("precision", C.c_int32),
        ("device_pointers", C.c_int32),
        ("min_coarse_n", C.c_int32),
        ("max_levels", C.c_int32),
        ("breakdown_at", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("converged", C.c_int32),
        ("levels", C.c_int32),
        ("status", C.c_int32),
        ("residual", C.c_double),
        ("bad_index", C.c_int64),
        ("history_len", C.c_int32),
        ("reserved", C.c_int32),
    ]


STATS_DTYPE = np.dtype(Stats)  # the same layout as a numpy record
assert STATS_DTYPE.itemsize == C.sizeof(Stats)


def _load() -> C.CDLL:
    if not os.path.exists(_LIB_PATH):
        raise ImportError(f"{_LIB_PATH} is not built (python darcy-cpp_amd/build.py)")
    lib = C.CDLL(_LIB_PATH)
    vp, i64, i32, dp = C.c_void_p, C.c_int64, C.c_int32, C.POINTER(C.c_double)
    lib.darcy_default_opts.argtypes = [C.POINTER(Opts)]
    lib.darcy_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
    lib.darcy_ctx_destroy.argtypes = [vp]
    lib.darcy_last_error.argtypes = [vp]
    lib.darcy_last_error.restype = C.c_char_p
    lib.darcy_solve_batch.argtypes = [vp, i64, C.POINTER(i64), C.POINTER(vp), C.POINTER(vp), dp, C.POINTER(Opts),
                                      C.POINTER(vp), C.POINTER(vp), C.POINTER(Stats)]
    lib.darcy_device_alloc.argtypes = [vp, C.c_size_t, C.POINTER(vp)]
    lib.darcy_device_free.argtypes = [vp, vp]
    lib.darcy_copy_to_device.argtypes = [vp, vp, vp, C.c_size_t]
    lib.darcy_copy_to_host.argtypes = [vp, vp, vp, C.c_size_t]
    lib.darcy_last_kernel_ms.argtypes = [vp, dp]
    lib.darcy_device_info.argtypes = [vp, C.c_char_p, i32]
    lib.darcy_bandwidth_probe.argtypes = [vp, C.c_size_t, i32, dp]
    lib.darcy_solve.argtypes = [vp, vp, vp, C.c_double, i64, C.POINTER(Opts), vp, vp, C.POINTER(Stats), dp, i32]
    lib.darcy_apply_preconditioner_opts.argtypes = [vp, vp, i64, vp, vp, C.POINTER(Opts)]
    lib.darcy_solve_batch_noalloc.argtypes = lib.darcy_solve_batch.argtypes
    lib.darcy_scratch_bytes.argtypes = [vp, C.POINTER(C.c_size_t)]
    lib.darcy_grf_fields.argtypes = [vp, i64, i64, C.c_double, C.c_double, vp, C.c_double, C.c_double, vp, vp, i32]
    lib.darcy_ctx_set_tuning.argtypes = [vp, C.c_char_p, i32]
    lib.darcy_ctx_get_tuning.argtypes = [vp, C.c_char_p, C.POINTER(i32)]
    # Learned warm start.
    fp, u16p = C.POINTER(C.c_float), C.POINTER(C.c_uint16)
    lib.darcy_net_param_count.argtypes = [i32]
    lib.darcy_net_param_count.restype = i64
    lib.darcy_net_create.argtypes = [vp, i64, i64, i32, C.POINTER(vp)]
    lib.darcy_net_destroy.argtypes = [vp]
    lib.darcy_net_destroy.restype = None
    lib.darcy_net_set_params.argtypes = [vp, fp, i64]
    lib.darcy_net_get_params.argtypes = [vp, fp, i64]
    lib.darcy_net_predict.argtypes = [vp, i64, C.POINTER(vp), dp, C.POINTER(vp)]
    lib.darcy_net_train_step.argtypes = [vp, i64, C.POINTER(vp), C.POINTER(vp), C.c_float, dp]
    lib.darcy_net_gradient.argtypes = [vp, i64, C.POINTER(vp), C.POINTER(vp), dp, fp]
    lib.darcy_nn_conv_forward.argtypes = [vp, i32, i32, i32, i32, i32, i32, u16p, u16p, fp, u16p, u16p]
    lib.darcy_nn_conv_backward_data.argtypes = [vp, i32, i32, i32, i32, i32, i32, u16p, u16p, u16p]
    lib.darcy_nn_conv_backward_weight.argtypes = [vp, i32, i32, i32, i32, i32, i32, i32, u16p, u16p, fp, fp]
    lib.darcy_nn_mfma_check.argtypes = [vp, u16p, u16p, fp]
    return lib


def _ptr(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


def _ptr_table(ptrs: Sequence[int]) -> np.ndarray:
    """Device addresses as a contiguous uint64 array (a C pointer table)."""
    return np.ascontiguousarray(np.asarray(ptrs, dtype=np.uint64))


class Context:
    """A darcy_ctx: one device, one stream, its workspace."""

    def __init__(self, device: int = 0, apply_tuning: bool = True):
        self.lib = _load()
        self.ptr = C.c_void_p()
        rc = self.lib.darcy_ctx_create(device, C.byref(self.ptr))
        if rc != OK:
            raise RuntimeError(f"darcy_ctx_create failed (code {rc}): no usable HIP device")
        if apply_tuning:  # the autotuner's table for this device, if one was made
            from . import autotune

            autotune.apply(self)

    def close(self) -> None:
        if self.ptr:
            self.lib.darcy_ctx_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def _check(self, rc: int) -> None:
        if rc == OK:
            return
        msg = self.lib.darcy_last_error(self.ptr).decode()
        if rc == INVALID_INPUT:
            raise ValueError(msg)
        raise RuntimeError(msg)

    def info(self) -> str:
        buf = C.create_string_buffer(256)
        self.lib.darcy_device_info(self.ptr, buf, 256)
        return buf.value.decode()

    def bandwidth_probe(self, nbytes: int = 1 << 30, repeats: int = 10) -> float:
        """GB/s (read + write) of the library's streaming copy kernel."""
        g = C.c_double()
        self._check(self.lib.darcy_bandwidth_probe(self.ptr, nbytes, repeats, C.byref(g)))
        return g.value

    # ----------------------------------------------------------- memory
    def alloc(self, nbytes: int) -> int:
        p = C.c_void_p()
        self._check(self.lib.darcy_device_alloc(self.ptr, nbytes, C.byref(p)))
        return int(p.value)

    def free(self, dptr: int) -> None:
        self._check(self.lib.darcy_device_free(self.ptr, C.c_void_p(dptr)))

    def upload(self, dptr: int, host: np.ndarray) -> None:
        host = np.ascontiguousarray(host)
        self._check(self.lib.darcy_copy_to_device(self.ptr, C.c_void_p(dptr), host.ctypes.data, host.nbytes))

    def download(self, host: np.ndarray, dptr: int) -> None:
        assert host.flags.c_contiguous
        self._check(self.lib.darcy_copy_to_host(self.ptr, host.ctypes.data, C.c_void_p(dptr), host.nbytes))

    def last_kernel_ms(self) -> float:
        ms = C.c_double()
        self.lib.darcy_last_kernel_ms(self.ptr, C.byref(ms))
        return ms.value

    # ------------------------------------------------------------ solve
    def set_tuning(self, key: str, value: int) -> None:
        self._check(self.lib.darcy_ctx_set_tuning(self.ptr, key.encode(), value))

    def get_tuning(self, key: str) -> int:
        v = C.c_int32()
        self._check(self.lib.darcy_ctx_get_tuning(self.ptr, key.encode(), C.byref(v)))
        return v.value

    def _opts(self, **kw) -> Opts:
        o = Opts()
        self.lib.darcy_default_opts(C.byref(o))
        for key, value in kw.items():
            if key == "preconditioner":
                value = PC[value]
            elif key == "precision":
                value = PRECISION[value]
            setattr(o, key, int(value) if isinstance(value, bool) else value)
        return o

    def solve(self, a: np.ndarray, f: float = 1.0, *, rhs: np.ndarray | None = None, u0: np.ndarray | None = None,
              history_cap: int | None = None, **options) -> dict:
        """One solve of host arrays with any darcy_opts field as a keyword
        (tol, max_iter, preconditioner, pre_sweeps, post_sweeps, cycle_index,
        precision, record_history, min_coarse_n, max_levels, breakdown_at)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        o = self._opts(**options)
        u = np.empty((n, n))
        st = Stats()
        if rhs is not None:
            rhs = np.ascontiguousarray(rhs, dtype=np.float64)
        if u0 is not None:
            u0 = np.ascontiguousarray(u0, dtype=np.float64)
        cap = (o.max_iter + 1) if history_cap is None else history_cap
        hist = np.zeros(max(cap, 1)) if o.record_history else None
        self._check(self.lib.darcy_solve(
            self.ptr, a.ctypes.data, None if rhs is None else rhs.ctypes.data, C.c_double(f), C.c_int64(n),
            C.byref(o), None if u0 is None else u0.ctypes.data, u.ctypes.data, C.byref(st),
            None if hist is None else _ptr(hist, C.c_double), cap if hist is not None else 0))
        return {"u": u, "iterations": st.iterations, "converged": bool(st.converged), "levels": st.levels,
                "status": st.status, "residual": st.residual,
                "history": hist[: st.history_len].copy() if hist is not None else None}

    def apply_preconditioner(self, a: np.ndarray, r: np.ndarray, **options) -> np.ndarray:
        """z = M r with the cycle options of darcy_opts as keywords."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        r = np.ascontiguousarray(r, dtype=np.float64)
        z = np.empty_like(r)
        o = self._opts(**options)
        self._check(self.lib.darcy_apply_preconditioner_opts(self.ptr, a.ctypes.data, C.c_int64(a.shape[0]),
                                                             r.ctypes.data, z.ctypes.data, C.byref(o)))
        return z

    def grf(self, xi: np.ndarray, *, alpha: float = 2.0, tau: float = 3.0, hi: float = 12.0, lo: float = 3.0,
            field: bool = True, media: bool = True) -> tuple[np.ndarray | None, np.ndarray | None]:
        """Gaussian random fields and two-phase media of the normals xi
        (shape (count, n, n)), computed on the GPU: (fields, media)."""
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        if xi.ndim == 2:
            xi = xi[None]
        count, n = xi.shape[0], xi.shape[1]
        g = np.empty_like(xi) if field else None
        a = np.empty_like(xi) if media else None
        self._check(self.lib.darcy_grf_fields(self.ptr, count, n, C.c_double(alpha), C.c_double(tau), xi.ctypes.data,
                                              C.c_double(hi), C.c_double(lo), g.ctypes.data if field else None,
                                              a.ctypes.data if media else None, 0))
        return g, a

    def grf_device(self, count: int, n: int, xi_ptr: int, a_ptr: int, *, alpha: float = 2.0, tau: float = 3.0,
                   hi: float = 12.0, lo: float = 3.0, field_ptr: int | None = None) -> None:
        """Media (and optionally fields) of count * n * n normals already in
        device memory, written to device memory."""
        self._check(self.lib.darcy_grf_fields(self.ptr, count, n, C.c_double(alpha), C.c_double(tau),
                                              C.c_void_p(xi_ptr), C.c_double(hi), C.c_double(lo),
                                              C.c_void_p(field_ptr) if field_ptr else None, C.c_void_p(a_ptr), 1))

    def solve_device(self, n: int, a_ptr: int, u_ptr: int, *, u0_ptr: int | None = None, f: float = 1.0,
                     tol: float = 1e-10, max_iter: int = 10000, cycle_index: int = 2,
                     precision: str = "fp64") -> dict:
        """One solve whose a / u (and u0) live in device memory."""
        o = Opts()
        self.lib.darcy_default_opts(C.byref(o))
        o.tol, o.max_iter, o.cycle_index = tol, max_iter, cycle_index
        o.precision, o.device_pointers = PRECISION[precision], 1
        st = Stats()
        self._check(self.lib.darcy_solve(self.ptr, C.c_void_p(a_ptr), None, C.c_double(f), C.c_int64(n),
                                         C.byref(o), C.c_void_p(u0_ptr) if u0_ptr else None, C.c_void_p(u_ptr),
                                         C.byref(st), None, 0))
        return {"iterations": st.iterations, "converged": bool(st.converged), "residual": st.residual,
                "levels": st.levels}

    def scratch_bytes(self) -> int:
        b = C.c_size_t()
        self._check(self.lib.darcy_scratch_bytes(self.ptr, C.byref(b)))
        return b.value

    def solve_batch_device(
        self,
        sizes: Sequence[int],
        a_ptrs: Sequence[int],
        u_ptrs: Sequence[int],
        *,
        u0_ptrs: Sequence[int] | None = None,
        f: float = 1.0,
        tol: float = 1e-10,
        max_iter: int = 10000,
        preconditioner: str = "multigrid",
        pre_sweeps: int = 1,
        post_sweeps: int = 1,
        cycle_index: int = 2,
        precision: str = "fp64",
        no_alloc: bool = False,
    ) -> dict[str, np.ndarray]:
        """Solve problems whose a / u (and u0) live in device memory;
        no_alloc uses darcy_solve_batch_noalloc (the library's static
        scratch, no device allocation)."""
        count = len(sizes)
        o = Opts()
        self.lib.darcy_default_opts(C.byref(o))
        o.tol, o.max_iter, o.preconditioner = tol, max_iter, PC[preconditioner]
        o.pre_sweeps, o.post_sweeps, o.cycle_index = pre_sweeps, post_sweeps, cycle_index
        o.precision, o.device_pointers = PRECISION[precision], 1
        # The tables as numpy arrays: building them element by element in
        # ctypes costs about 2 ms per 1000 problems, a few percent of a solve.
        n = np.ascontiguousarray(sizes, dtype=np.int64)
        a, u = _ptr_table(a_ptrs), _ptr_table(u_ptrs)
        u0 = _ptr_table(u0_ptrs) if u0_ptrs is not None else None
        if n.shape != (count,) or a.shape != (count,) or u.shape != (count,) or (u0 is not None and u0.shape != (count,)):
            raise ValueError("one size, one a, one u (and one u0) pointer per problem")
        fs = np.full(count, float(f))
        st = np.zeros(count, dtype=STATS_DTYPE)
        entry = self.lib.darcy_solve_batch_noalloc if no_alloc else self.lib.darcy_solve_batch
        self._check(entry(self.ptr, count, _ptr(n, C.c_int64), _ptr(a, C.c_void_p), None, _ptr(fs, C.c_double),
                          C.byref(o), _ptr(u0, C.c_void_p) if u0 is not None else None, _ptr(u, C.c_void_p),
                          _ptr(st, Stats)))
        return {
            "iterations": st["iterations"].astype(np.int64),
            "converged": st["converged"] != 0,
            "residual": st["residual"].copy(),
            "levels": st["levels"].astype(np.int64),
        }

    # ------------------------------------------------- conv layer checks
    # bf16 tensors travel as uint16 bit patterns.
    def conv_forward(self, x: np.ndarray, w: np.ndarray, bias: np.ndarray, stride: int,
                     want_act: bool = True) -> tuple[np.ndarray, np.ndarray | None]:
        b, h, wd, cin = x.shape
        cout = w.shape[0]
        x = np.ascontiguousarray(x, dtype=np.uint16)
        w = np.ascontiguousarray(w, dtype=np.uint16)
        bias = np.ascontiguousarray(bias, dtype=np.float32)
        z = np.empty((b, h // stride, wd // stride, cout), dtype=np.uint16)
        act = np.empty_like(z) if want_act else None
        self._check(self.lib.darcy_nn_conv_forward(
            self.ptr, b, h, wd, cin, cout, stride, _ptr(x, C.c_uint16), _ptr(w, C.c_uint16),
            _ptr(bias, C.c_float), _ptr(z, C.c_uint16), _ptr(act, C.c_uint16) if want_act else None))
        return z, act

    def conv_backward_data(self, dz: np.ndarray, w: np.ndarray, in_shape: tuple[int, ...],
                           stride: int) -> np.ndarray:
        b, h, wd, cin = in_shape
        cout = w.shape[0]
        dz = np.ascontiguousarray(dz, dtype=np.uint16)
        w = np.ascontiguousarray(w, dtype=np.uint16)
        dx = np.empty((b, h, wd, cin), dtype=np.uint16)
        self._check(self.lib.darcy_nn_conv_backward_data(
            self.ptr, b, h, wd, cin, cout, stride, _ptr(dz, C.c_uint16), _ptr(w, C.c_uint16), _ptr(dx, C.c_uint16)))
        return dx

    def conv_backward_weight(self, x: np.ndarray, dz: np.ndarray, stride: int,
                             splits: int = 1) -> tuple[np.ndarray, np.ndarray]:
        b, h, wd, cin = x.shape
        cout = dz.shape[3]
        x = np.ascontiguousarray(x, dtype=np.uint16)
        dz = np.ascontiguousarray(dz, dtype=np.uint16)
        dw = np.empty((cout, 3, 3, cin), dtype=np.float32)
        db = np.empty(cout, dtype=np.float32)
        self._check(self.lib.darcy_nn_conv_backward_weight(
            self.ptr, b, h, wd, cin, cout, stride, splits, _ptr(x, C.c_uint16), _ptr(dz, C.c_uint16),
            _ptr(dw, C.c_float), _ptr(db, C.c_float)))
        return dw, db

    def mfma_check(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        a = np.ascontiguousarray(a, dtype=np.uint16)
        b = np.ascontiguousarray(b, dtype=np.uint16)
        assert a.shape == (16, 32) and b.shape == (32, 16)
        c = np.empty((16, 16), dtype=np.float32)
        self._check(self.lib.darcy_nn_mfma_check(self.ptr, _ptr(a, C.c_uint16), _ptr(b, C.c_uint16),
                                                 _ptr(c, C.c_float)))
        return c


NET_TRAIN, NET_HALF_RES = 1, 2  # darcy_net_create flags


def _net_flags(train: bool, sub: int) -> int:
    if sub not in (1, 2):
        raise ValueError("the network's sub-pixel factor must be 1 or 2")
    return (NET_TRAIN if train else 0) | (NET_HALF_RES if sub == 2 else 0)


def net_param_count(sub: int = 1) -> int:
    return int(_load().darcy_net_param_count(_net_flags(False, sub)))


class Net:
    """A darcy_net: the warm-start network on one context; `sub` = 2 runs it
    at half the grid's resolution (2 x 2 blocks of cells per network pixel)."""

    def __init__(self, ctx: Context, n: int, max_batch: int, train: bool = False, sub: int = 1):
        self.ctx, self.lib = ctx, ctx.lib
        self.n, self.max_batch, self.train, self.sub = n, max_batch, train, sub
        flags = _net_flags(train, sub)
        self.count = int(self.lib.darcy_net_param_count(flags))
        self.ptr = C.c_void_p()
        ctx._check(self.lib.darcy_net_create(ctx.ptr, n, max_batch, flags, C.byref(self.ptr)))

    def close(self) -> None:
        if self.ptr:
            self.lib.darcy_net_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params: np.ndarray) -> None:
        p = np.ascontiguousarray(params, dtype=np.float32)
        self.ctx._check(self.lib.darcy_net_set_params(self.ptr, _ptr(p, C.c_float), p.size))

    def get_params(self) -> np.ndarray:
        p = np.empty(self.count, dtype=np.float32)
        self.ctx._check(self.lib.darcy_net_get_params(self.ptr, _ptr(p, C.c_float), p.size))
        return p

    def predict_device(self, a_ptrs: Sequence[int], f: Sequence[float], u0_ptrs: Sequence[int]) -> None:
        count = len(a_ptrs)
        a, u = _ptr_table(a_ptrs), _ptr_table(u0_ptrs)
        fs = np.ascontiguousarray(f, dtype=np.float64)
        if fs.shape != (count,) or u.shape != (count,):
            raise ValueError("predict needs one f and one u0 pointer per problem")
        self.ctx._check(self.lib.darcy_net_predict(self.ptr, count, _ptr(a, C.c_void_p), _ptr(fs, C.c_double),
                                                   _ptr(u, C.c_void_p)))

    def train_step(self, a_ptrs: Sequence[int], u_ptrs: Sequence[int], lr: float, want_loss: bool = True) -> float:
        count = len(a_ptrs)
        if len(u_ptrs) != count:
            raise ValueError("one a and one u pointer per sample")
        a = (C.c_void_p * count)(*a_ptrs)
        u = (C.c_void_p * count)(*u_ptrs)
        loss = C.c_double(float("nan"))
        self.ctx._check(self.lib.darcy_net_train_step(self.ptr, count, a, u, lr,
                                                      C.byref(loss) if want_loss else None))
        return loss.value

    def gradient(self, a_ptrs: Sequence[int], u_ptrs: Sequence[int]) -> tuple[float, np.ndarray]:
        count = len(a_ptrs)
        if len(u_ptrs) != count:
            raise ValueError("one a and one u pointer per sample")
        a = (C.c_void_p * count)(*a_ptrs)
        u = (C.c_void_p * count)(*u_ptrs)
        loss = C.c_double()
        g = np.empty(self.count, dtype=np.float32)
        self.ctx._check(self.lib.darcy_net_gradient(self.ptr, count, a, u, C.byref(loss), _ptr(g, C.c_float)))
        return loss.value, g
