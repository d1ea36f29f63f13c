"""MI355X (gfx950) Sinkhorn-barycenter hot path of Wasserstein dictionary learning.

The product is the C-ABI shared library ``_lib/libwdl.so`` (HIP kernels,
declared in ``include/wdl_c.h``) and the header-only C++ KernelType in
``include/wdl/hip_kernel_type.hpp``.  This module is a thin ctypes layer over
the same C ABI for Python callers, tests and the benchmark.  It never falls
back to a CPU path: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os
import sys
import weakref

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_PKG, "_lib", "libwdl.so")

WDL_OK = 0
LOSS_TV, LOSS_QUADRATIC, LOSS_WASSERSTEIN, LOSS_KL = 1, 2, 3, 4
FLAG_LOG_DOMAIN, FLAG_WARM_RESTART, FLAG_DENSE, FLAG_NO_GRAPH, FLAG_WEIGHTS_ONLY = 1, 2, 4, 8, 16
FLAG_TILED = 32
FLAG_EXACT_LSE = 64
FLAG_BF16X3 = 128
FLAG_CHECKPOINT = 256
NOALLOC_MAX_VALUES = 2097152  # WDL_NOALLOC_MAX_VALUES
RUN_BATCHED, RUN_PLAIN = 0, 1
SERVER_POOLS, SERVER_MIXED = 0, 1
STAGE_FORWARD, STAGE_BACKWARD = 0, 1

_lib = None
_vp = ctypes.c_void_p
_d = ctypes.POINTER(ctypes.c_double)
_u8 = ctypes.POINTER(ctypes.c_ubyte)
_i = ctypes.c_int

# name -> (restype, argtypes); mirrors include/wdl_c.h
_SIGS = {
    "wdl_last_error": (ctypes.c_char_p, []),
    "wdl_version": (ctypes.c_char_p, []),
    "wdl_debug_fill": (_i, [_i]),
    "wdl_ctx_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "wdl_ctx_destroy": (_i, [_vp]),
    "wdl_ctx_stream": (_vp, [_vp]),
    "wdl_ctx_synchronize": (_i, [_vp]),
    "wdl_gauss2d_create": (_i, [_vp, _i, _i, ctypes.c_double, _i, ctypes.POINTER(_vp)]),
    "wdl_gauss2d_destroy": (_i, [_vp]),
    "wdl_gauss2d_table": (_i, [_vp, _d]),
    "wdl_gauss2d_convolve": (_i, [_vp, _d, _d, _i]),
    "wdl_gauss2d_log_convolve": (_i, [_vp, _d, _d, _i]),
    "wdl_gauss2d_log_convolve_signarray": (_i, [_vp, _d, _u8, _d, _i]),
    "wdl_gauss2d_convolve_device": (_i, [_vp, _vp, _vp, _i]),
    "wdl_gauss2d_log_convolve_noalloc": (_i, [_vp, _vp, _vp, _vp, _i]),
    "wdl_gauss1d_create": (_i, [_vp, _i, ctypes.c_double, ctypes.POINTER(_vp)]),
    "wdl_gauss1d_destroy": (_i, [_vp]),
    "wdl_gauss1d_convolve": (_i, [_vp, _d, _d, _i]),
    "wdl_gauss1d_convolve_device": (_i, [_vp, _vp, _vp, _i]),
    "wdl_gauss3d_create": (_i, [_vp, _i, _i, _i, ctypes.c_double, ctypes.POINTER(_vp)]),
    "wdl_gauss3d_destroy": (_i, [_vp]),
    "wdl_gauss3d_convolve": (_i, [_vp, _d, _d, _i]),
    "wdl_gauss3d_convolve_device": (_i, [_vp, _vp, _vp, _i]),
    "wdl_solver_create": (_i, [_vp, _i, _i, _i, _i, _i, ctypes.c_double, _i, _i,
                               ctypes.POINTER(_vp)]),
    "wdl_solver_destroy": (_i, [_vp]),
    "wdl_solver_footprint": (ctypes.c_longlong, [_i, _i, _i, _i, _i, _i]),
    "wdl_solver_set_observed": (_i, [_vp, _d]),
    "wdl_solver_set_observed_device": (_i, [_vp, _vp]),
    "wdl_solver_set_problem": (_i, [_vp, _d, _d]),
    "wdl_solver_set_problem_device": (_i, [_vp, _vp, _vp]),
    "wdl_solver_set_warm_state": (_i, [_vp, _d]),
    "wdl_solver_get_warm_state": (_i, [_vp, _d]),
    "wdl_solver_warm_commit": (_i, [_vp]),
    "wdl_solver_run": (_i, [_vp, _i]),
    "wdl_solver_run_stage": (_i, [_vp, _i, _i]),
    "wdl_solver_set_streams": (_i, [_vp, _i]),
    "wdl_solver_autotune": (_i, [_vp, _i, ctypes.POINTER(_i)]),
    "wdl_solver_set_tuning": (_i, [_vp, _i, _i, _i, _i, _i]),
    "wdl_solver_get_tuning": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i),
                                   ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
    "wdl_solver_set_iterations": (_i, [_vp, ctypes.POINTER(_i)]),
    "wdl_solver_get_results": (_i, [_vp, _d, _d, _d, _d]),
    "wdl_solver_get_grad_dic": (_i, [_vp, _d]),
    "wdl_solver_log_range_events": (_i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "wdl_gradient_both": (_i, [_vp, _d, _d, _d, _d, _d, _d]),
    "wdl_solver_set_scale": (_i, [_vp, ctypes.c_double, _i]),
    "wdl_evaluate_both": (ctypes.c_double, [_vp, _d, _d, _i, ctypes.c_double]),
    "wdl_evaluate_both_device": (_i, [_vp, _vp, _vp, ctypes.POINTER(ctypes.c_double)]),
    "wdl_export_histogram_csv": (_i, [ctypes.c_char_p, ctypes.c_char_p, _d, ctypes.c_longlong]),
    "wdl_export_histogram_raw": (_i, [ctypes.c_char_p, ctypes.c_char_p, _d, ctypes.c_longlong]),
    "wdl_export_histograms_csv": (_i, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_longlong), _i]),
    "wdl_export_histograms_raw": (_i, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p),
                                      ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_longlong), _i]),
    "wdl_load_csv_to_pdf": (ctypes.c_double, [ctypes.c_char_p, _d, ctypes.c_longlong,
                                             ctypes.POINTER(ctypes.c_longlong)]),
    "wdl_solver_set_slot": (_i, [_vp, _i, _d, _d, _i, ctypes.c_double]),
    "wdl_solver_projection_round": (_i, [_vp]),
    "wdl_solver_get_slot": (_i, [_vp, _i, _d, _d]),
    "wdl_ctx_query": (_i, [_vp]),
    "wdl_server_create": (_i, [_i, ctypes.c_longlong, _i, ctypes.POINTER(_vp)]),
    "wdl_server_destroy": (_i, [_vp]),
    "wdl_server_add_dictionary": (_i, [_vp, _i, _i, _i, ctypes.c_double, _i, _i, _i, _d,
                                       ctypes.POINTER(_i)]),
    "wdl_server_submit": (_i, [_vp, _i, _d, _i, _i, ctypes.c_double,
                               ctypes.POINTER(ctypes.c_longlong)]),
    "wdl_server_wait": (_i, [_vp, ctypes.c_longlong, ctypes.c_double, _d, _d, _d, _d]),
    "wdl_server_stats": (_i, [_vp, ctypes.POINTER(ctypes.c_longlong)]),
    "wdl_server_set_batching": (_i, [_vp, _i]),
    "wdl_group_create": (_i, [_vp, ctypes.POINTER(_vp), _i, ctypes.POINTER(_vp)]),
    "wdl_group_destroy": (_i, [_vp]),
    "wdl_group_run": (_i, [_vp]),
    "wdl_group_projection_round": (_i, [_vp]),
    "wdl_group_shape": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
}


class WdlError(RuntimeError):
    pass


def lib():
    """Loads libwdl.so; raises if it has not been built (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise WdlError("libwdl.so not built: run python -m wassersteindictionarylearning_amd._build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(status: int, what: str):
    if status != WDL_OK:
        raise WdlError("%s failed (%d): %s" % (what, status, lib().wdl_last_error().decode()))


def _dp(a):
    return None if a is None else a.ctypes.data_as(_d)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One device, one HIP stream (wdl_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        _check(lib().wdl_ctx_create(device, ctypes.byref(h)), "wdl_ctx_create")
        self.handle = h
        self.device = device
        self._children = weakref.WeakSet()

    @property
    def stream(self) -> int:
        return lib().wdl_ctx_stream(self.handle)

    def synchronize(self):
        _check(lib().wdl_ctx_synchronize(self.handle), "wdl_ctx_synchronize")

    def close(self):
        """Destroys the objects created on this context, then the context."""
        if self.handle:
            for child in list(self._children):
                child.close()
            lib().wdl_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass


class _LineKernel:
    """Shared surface of the 1-D and 3-D KernelTypes (convolve, its adjoint and
    the FP32 device form)."""

    _prefix = ""

    def _create(self, ctx, fn, *args):
        h = _vp()
        _check(getattr(lib(), fn)(ctx.handle, *args, ctypes.byref(h)), fn)
        self.handle, self.ctx = h, ctx
        ctx._children.add(self)

    def convolve(self, u, nvectors: int | None = None):
        u = _f64(u)
        nv = u.size // self.N if nvectors is None else nvectors
        out = np.zeros_like(u)
        fn = self._prefix + "_convolve"
        _check(getattr(lib(), fn)(self.handle, _dp(u), _dp(out), nv), fn)
        return out

    convolveAdjoint = convolve

    def convolve_device(self, d_in: int, d_out: int, nvectors: int):
        fn = self._prefix + "_convolve_device"
        _check(getattr(lib(), fn)(self.handle, d_in, d_out, nvectors), fn)

    def close(self):
        if self.handle:
            getattr(lib(), self._prefix + "_destroy")(self.handle)
            self.handle = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass


class HipGaussian1DKernel(_LineKernel):
    """Gaussian1DKernel (kernels.h:501-556): nvectors vectors of W values.
    Constructor argument order follows kernels.h:503 (gamma, W)."""

    _prefix = "wdl_gauss1d"

    def __init__(self, ctx: Context, gamma: float, W: int):
        self._create(ctx, "wdl_gauss1d_create", W, float(gamma))
        self.gamma, self.W, self.N = float(gamma), W, W


class HipGaussian3DKernel(_LineKernel):
    """Gaussian3DKernel (kernels.h:712-815): volumes of W x H x D, voxel
    (x, y, z) at z*W*H + y*W + x.  Constructor argument order follows
    kernels.h:714 (gamma, W, H, D)."""

    _prefix = "wdl_gauss3d"

    def __init__(self, ctx: Context, gamma: float, W: int, H: int, D: int):
        self._create(ctx, "wdl_gauss3d_create", W, H, D, float(gamma))
        self.gamma, self.W, self.H, self.D, self.N = float(gamma), W, H, D, W * H * D


class HipGaussian2DKernel:
    """The KernelType surface of kernels.h (Gaussian2DKernel / LogSignArrayGaussian2DKernel).

    Constructor argument order follows kernels.h:632 (gamma, W, H).
    """

    def __init__(self, ctx: Context, gamma: float, W: int, H: int, flags: int = 0):
        h = _vp()
        _check(lib().wdl_gauss2d_create(ctx.handle, W, H, float(gamma), flags, ctypes.byref(h)),
               "wdl_gauss2d_create")
        self.handle, self.ctx, self.gamma, self.W, self.H = h, ctx, float(gamma), W, H
        self.N = W * H
        self._groups = weakref.WeakSet()  # groups that hold this solver
        ctx._children.add(self)

    def kernel1d(self):
        out = np.zeros(max(self.W, self.H))
        _check(lib().wdl_gauss2d_table(self.handle, _dp(out)), "wdl_gauss2d_table")
        return out

    def convolve(self, u, nvectors: int | None = None):
        u = _f64(u)
        nv = u.size // self.N if nvectors is None else nvectors
        out = np.zeros_like(u)
        _check(lib().wdl_gauss2d_convolve(self.handle, _dp(u), _dp(out), nv), "wdl_gauss2d_convolve")
        return out

    convolveAdjoint = convolve

    def log_convolve(self, log_u, nvectors: int | None = None):
        u = _f64(log_u)
        nv = u.size // self.N if nvectors is None else nvectors
        out = np.zeros_like(u)
        _check(lib().wdl_gauss2d_log_convolve(self.handle, _dp(u), _dp(out), nv),
               "wdl_gauss2d_log_convolve")
        return out

    log_convolveAdjoint = log_convolve

    def log_convolve_signArray(self, log_u, signs, nvectors: int | None = None):
        """Returns (out, signs_out); the input sign array is not modified."""
        u = _f64(log_u)
        nv = u.size // self.N if nvectors is None else nvectors
        s = np.array(signs, dtype=np.uint8, copy=True)
        out = np.zeros_like(u)
        _check(lib().wdl_gauss2d_log_convolve_signarray(self.handle, _dp(u),
                                                        s.ctypes.data_as(_u8), _dp(out), nv),
               "wdl_gauss2d_log_convolve_signarray")
        return out, s

    log_convolve_signArrayAdjoint = log_convolve_signArray

    def convolve_device(self, d_in: int, d_out: int, nvectors: int):
        _check(lib().wdl_gauss2d_convolve_device(self.handle, d_in, d_out, nvectors),
               "wdl_gauss2d_convolve_device")

    def close(self):
        if self.handle:
            lib().wdl_gauss2d_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass


class SinkhornSolver:
    """P compute_gradient_both solves at once (wdl_solver), plus evaluate_both."""

    def __init__(self, ctx: Context, W, H, K, P, L, gamma, loss=LOSS_QUADRATIC, flags=0):
        h = _vp()
        _check(lib().wdl_solver_create(ctx.handle, W, H, K, P, L, float(gamma), loss, flags,
                                       ctypes.byref(h)), "wdl_solver_create")
        self.handle, self.ctx = h, ctx
        self.W, self.H, self.K, self.P, self.L = W, H, K, P, L
        self.N = W * H
        self._groups = weakref.WeakSet()  # groups that hold this solver
        ctx._children.add(self)

    @staticmethod
    def footprint(W, H, K, P, L, flags=0) -> int:
        return int(lib().wdl_solver_footprint(W, H, K, P, L, flags))

    def set_observed(self, observed):
        o = _f64(observed)
        assert o.size == self.P * self.N
        _check(lib().wdl_solver_set_observed(self.handle, _dp(o)), "wdl_solver_set_observed")

    def set_observed_device(self, d_obs: int):
        _check(lib().wdl_solver_set_observed_device(self.handle, d_obs),
               "wdl_solver_set_observed_device")

    def set_problem(self, atoms, lambdas):
        a, l = _f64(atoms), _f64(lambdas)
        assert a.size == self.K * self.N and l.size == self.P * self.K
        _check(lib().wdl_solver_set_problem(self.handle, _dp(a), _dp(l)), "wdl_solver_set_problem")

    def set_problem_device(self, d_atoms: int, d_lambdas: int):
        _check(lib().wdl_solver_set_problem_device(self.handle, d_atoms, d_lambdas),
               "wdl_solver_set_problem_device")

    def set_warm_state(self, b_storage):
        b = _f64(b_storage)
        _check(lib().wdl_solver_set_warm_state(self.handle, _dp(b)), "wdl_solver_set_warm_state")

    def get_warm_state(self):
        out = np.zeros((self.P, self.K, self.N))
        _check(lib().wdl_solver_get_warm_state(self.handle, _dp(out)), "wdl_solver_get_warm_state")
        return out

    def warm_commit(self):
        _check(lib().wdl_solver_warm_commit(self.handle), "wdl_solver_warm_commit")

    def run(self, mode: int = RUN_BATCHED):
        _check(lib().wdl_solver_run(self.handle, mode), "wdl_solver_run")

    def set_iterations(self, iterations=None):
        if iterations is None:
            _check(lib().wdl_solver_set_iterations(self.handle, None), "wdl_solver_set_iterations")
            return
        it = np.ascontiguousarray(iterations, dtype=np.int32)
        assert it.size == self.P
        _check(lib().wdl_solver_set_iterations(self.handle, it.ctypes.data_as(ctypes.POINTER(_i))),
               "wdl_solver_set_iterations")

    def set_streams(self, nstreams: int):
        _check(lib().wdl_solver_set_streams(self.handle, nstreams), "wdl_solver_set_streams")

    def autotune(self, reps: int = 3) -> int:
        """Picks the histogram-group (stream) count by timing; returns it."""
        n = ctypes.c_int()
        _check(lib().wdl_solver_autotune(self.handle, reps, ctypes.byref(n)), "wdl_solver_autotune")
        return n.value

    def set_tuning(self, nstreams=0, fwd_threads=0, bwd_threads=0, line_waves=0, line_group=0):
        _check(lib().wdl_solver_set_tuning(self.handle, nstreams, fwd_threads, bwd_threads,
                                           line_waves, line_group), "wdl_solver_set_tuning")

    def tuning(self) -> dict:
        v = [_i() for _ in range(5)]
        _check(lib().wdl_solver_get_tuning(self.handle, *[ctypes.byref(x) for x in v]),
               "wdl_solver_get_tuning")
        return dict(streams=v[0].value, fwd_threads=v[1].value, bwd_threads=v[2].value,
                    line_waves=v[3].value, line_group=v[4].value)

    def run_stage(self, stage: int, step: int):
        _check(lib().wdl_solver_run_stage(self.handle, stage, step), "wdl_solver_run_stage")

    def results(self, barycenters: bool = True):
        loss = np.zeros(self.P)
        gl = np.zeros((self.P, self.K))
        gd = np.zeros((self.K, self.N))
        q = np.zeros((self.P, self.N)) if barycenters else None
        _check(lib().wdl_solver_get_results(self.handle, _dp(loss), _dp(gl), _dp(gd), _dp(q)),
               "wdl_solver_get_results")
        return dict(loss=loss, grad_lambda=gl, grad_dic_sum=gd, bary=q)

    def log_range_events(self) -> int:
        c = ctypes.c_longlong()
        _check(lib().wdl_solver_log_range_events(self.handle, ctypes.byref(c)),
               "wdl_solver_log_range_events")
        return int(c.value)

    def grad_dic(self):
        out = np.zeros((self.P, self.K, self.N))
        _check(lib().wdl_solver_get_grad_dic(self.handle, _dp(out)), "wdl_solver_get_grad_dic")
        return out

    def set_scale(self, scale_dictionary: float, exp_weight: bool = True):
        _check(lib().wdl_solver_set_scale(self.handle, float(scale_dictionary), int(exp_weight)),
               "wdl_solver_set_scale")

    def evaluate_both(self, x):
        """The lbfgs_evaluate_t callback: returns (f, g)."""
        x = _f64(x)
        g = np.zeros_like(x)
        f = lib().wdl_evaluate_both(self.handle, _dp(x), _dp(g), x.size, 0.0)
        if np.isnan(f):
            err = lib().wdl_last_error().decode()
            if err:
                raise WdlError("wdl_evaluate_both: " + err)
        return f, g

    def evaluate_both_device(self, d_x: int, d_g: int) -> float:
        """evaluate_both with x and g in device memory (FP64 device pointers,
        P*K + K*N values each); returns f."""
        f = ctypes.c_double()
        _check(lib().wdl_evaluate_both_device(self.handle, d_x, d_g, ctypes.byref(f)),
               "wdl_evaluate_both_device")
        return f.value

    def set_slot(self, slot, observed=None, lambdas=None, iterations=0, step=0.0):
        o = None if observed is None else _f64(observed)
        lm = None if lambdas is None else _f64(lambdas)
        _check(lib().wdl_solver_set_slot(self.handle, slot, _dp(o), _dp(lm), int(iterations),
                                         float(step)), "wdl_solver_set_slot")

    def projection_round(self):
        _check(lib().wdl_solver_projection_round(self.handle), "wdl_solver_projection_round")

    def get_slot(self, slot):
        w = np.zeros(self.K)
        loss = ctypes.c_double()
        _check(lib().wdl_solver_get_slot(self.handle, slot, _dp(w), ctypes.byref(loss)),
               "wdl_solver_get_slot")
        return w, loss.value

    def close(self):
        if self.handle:
            for g in list(self._groups):
                g.close()
            lib().wdl_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass


GROUP_LAUNCH_KINDS = 12  # WDL_GROUP_LAUNCH_KINDS
GROUP_KIND_NAMES = ("xf1", "fwd_x", "fwd_y", "final", "dot", "xb1", "dot_fin", "bwd_x", "bwd_y",
                    "project", "warm_commit", "reduce_gd")


class SolverGroup:
    """Exact-shift log-domain solvers of uneven sizes run as one launch per
    stage kernel (wdl_group).  Results are read from the member solvers."""

    def __init__(self, ctx: Context, solvers):
        arr = (_vp * len(solvers))(*[s.handle for s in solvers])
        h = _vp()
        _check(lib().wdl_group_create(ctx.handle, arr, len(solvers), ctypes.byref(h)),
               "wdl_group_create")
        self.handle, self.ctx, self.members = h, ctx, list(solvers)
        for s in solvers:
            s._groups.add(self)
        ctx._children.add(self)

    def run(self):
        _check(lib().wdl_group_run(self.handle), "wdl_group_run")

    def projection_round(self):
        _check(lib().wdl_group_projection_round(self.handle), "wdl_group_projection_round")

    def shape(self) -> dict:
        grids = (_i * GROUP_LAUNCH_KINDS)()
        w, lds = _i(), _i()
        _check(lib().wdl_group_shape(self.handle, grids, ctypes.byref(w), ctypes.byref(lds)),
               "wdl_group_shape")
        return dict(grids=dict(zip(GROUP_KIND_NAMES, list(grids))), waves_per_line=w.value,
                    lds_bytes=lds.value)

    def close(self):
        if self.handle:
            lib().wdl_group_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass


class ProjectionServer:
    """Served projection loop (wdl_server): requests fit the weights of one
    histogram against a registered dictionary; a native scheduler thread
    batches them continuously on the device."""

    STATS = ("completed", "pools_created", "pools_released", "memory_waits", "bytes_in_use",
             "rounds")

    def __init__(self, device: int = 0, hbm_budget_bytes: int = 64 << 30,
                 max_slots_per_pool: int = 16, mixed: bool = False):
        """mixed: the pools of one iteration class run each round as one group
        launch (WDL_SERVER_MIXED) instead of on their own streams."""
        h = _vp()
        _check(lib().wdl_server_create(device, int(hbm_budget_bytes), int(max_slots_per_pool),
                                       ctypes.byref(h)), "wdl_server_create")
        self.handle = h
        self.K = {}
        if mixed:
            try:
                _check(lib().wdl_server_set_batching(h, SERVER_MIXED), "wdl_server_set_batching")
            except Exception:
                self.close()
                raise

    def add_dictionary(self, atoms, W, H, gamma, loss=LOSS_QUADRATIC, max_iterations=500,
                       flags=0) -> int:
        a = _f64(atoms)
        K = a.size // (W * H)
        d = ctypes.c_int()
        _check(lib().wdl_server_add_dictionary(self.handle, W, H, K, float(gamma), loss,
                                               int(max_iterations), flags, _dp(a),
                                               ctypes.byref(d)), "wdl_server_add_dictionary")
        self.K[d.value] = K
        return d.value

    def submit(self, dict_id, histogram, iterations, rounds, step=0.5) -> int:
        h = _f64(histogram)
        rid = ctypes.c_longlong()
        _check(lib().wdl_server_submit(self.handle, dict_id, _dp(h), int(iterations), int(rounds),
                                       float(step), ctypes.byref(rid)), "wdl_server_submit")
        return (rid.value, dict_id)

    def wait(self, request, timeout_s=60.0):
        """(weights, loss, latency_ms, queue_ms) of a submitted request."""
        rid, dict_id = request
        w = np.zeros(self.K[dict_id])
        loss, lat, qu = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        st = lib().wdl_server_wait(self.handle, rid, float(timeout_s), _dp(w), ctypes.byref(loss),
                                   ctypes.byref(lat), ctypes.byref(qu))
        if st == 1:
            raise TimeoutError("request %d not done after %g s" % (rid, timeout_s))
        _check(st, "wdl_server_wait")
        return w, loss.value, lat.value, qu.value

    def stats(self) -> dict:
        out = (ctypes.c_longlong * 6)()
        _check(lib().wdl_server_stats(self.handle, out), "wdl_server_stats")
        return dict(zip(self.STATS, [int(v) for v in out]))

    def close(self):
        if self.handle:
            lib().wdl_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        if not sys.is_finalizing():
            try:
                self.close()
            except Exception:
                pass


def debug_fill(on: bool) -> bool:
    """While on, solvers created fill the device buffers they do not initialise
    with NaN bytes (wdl_debug_fill): a read of unwritten memory shows in the
    results.  Returns the previous setting."""
    return bool(lib().wdl_debug_fill(1 if on else 0))


def export_histograms(folder, names, histograms, fmt="csv"):
    """ExportHistogramToCSV / ExportHistogramToRAW::exportHistogramsBatch
    (histogramIO.h:80-97, 222-238): one <folder>/<name>.csv|.raw per histogram.
    Returns 0, or -1 when any file could not be written (as the reference)."""
    if fmt not in ("csv", "raw"):
        raise ValueError("fmt must be 'csv' or 'raw'")
    if len(names) != len(histograms):
        raise ValueError("Histogram vector and filename vector are of different sizes")
    hs = [_f64(h).ravel() for h in histograms]
    n = len(hs)
    cn = (ctypes.c_char_p * max(n, 1))(*[str(x).encode() for x in names])
    cv = (_d * max(n, 1))(*[_dp(h) for h in hs])
    cl = (ctypes.c_longlong * max(n, 1))(*[h.size for h in hs])
    f = lib().wdl_export_histograms_csv if fmt == "csv" else lib().wdl_export_histograms_raw
    st = f(str(folder).encode(), cn, cv, cl, n)
    if st not in (0, -1):
        _check(st, "wdl_export_histograms_" + fmt)
    return st


def export_histogram(folder, name, histogram, fmt="csv"):
    """ExportHistogramTo{CSV,RAW}::exportHistogram (histogramIO.h:62-78, 208-220)."""
    return export_histograms(folder, [name], [histogram], fmt)


def load_csv_to_pdf(filename):
    """load_csv_to_pdf (histogramIO.h:314-337): (sum, values / sum); sum is -1
    and values empty when the file cannot be opened."""
    n = ctypes.c_longlong(0)
    fn = str(filename).encode()
    s = lib().wdl_load_csv_to_pdf(fn, None, 0, ctypes.byref(n))
    if s == -1.0 and n.value == 0:
        return s, np.zeros(0)
    out = np.empty(n.value)
    s = lib().wdl_load_csv_to_pdf(fn, _dp(out), out.size, ctypes.byref(n))
    return s, out
