"""ctypes access to oracle/_ref/*.so -- the reference's own templates compiled
from its sources by oracle/Makefile.

TEST INFRASTRUCTURE ONLY (same rule as oracle.py): used by tests/ to pin the
numpy oracle, by tests/golden/make_golden.py to write fixtures, and by
bench.py's CPU-baseline leg.  ``available()`` is False when the objects were
not built; callers skip then.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(_HERE, "_ref")
_libs: dict = {}

_d = ctypes.POINTER(ctypes.c_double)
_u8 = ctypes.POINTER(ctypes.c_ubyte)
_i, _f = ctypes.c_int, ctypes.c_double


def _lib(name: str):
    if name not in _libs:
        path = os.path.join(REF_DIR, name)
        _libs[name] = ctypes.CDLL(path) if os.path.exists(path) else None
        lib = _libs[name]
        if lib is not None and name == "libwdlref.so":
            for fn in ("ref_convolve", "ref_sse_convolve", "ref_log_convolve"):
                getattr(lib, fn).argtypes = [_i, _i, _f, _d, _d, _i]
            lib.ref_kernel_table.argtypes = [_i, _i, _f, _d]
            lib.ref_convolve1d.argtypes = [_i, _f, _d, _d, _i]
            lib.ref_convolve3d.argtypes = [_i, _i, _i, _f, _d, _d, _i]
            lib.ref_log_convolve_signarray.argtypes = [_i, _i, _f, _d, _u8, _d, _i]
            lib.ref_abs_log.argtypes = [_d, _u8, _d, _i]
            lib.ref_abs_exp.argtypes = [_d, _u8, _d, _i]
            for fn in ("ref_gradient_both", "ref_gradient_both_sse", "ref_gradient_both_log"):
                getattr(lib, fn).argtypes = [_i, _i, _f, _i, _i, _i, _d, _d, _d, _d, _d, _d, _d, _d]
            lib.ref_evaluate_both.argtypes = [_i, _i, _i, _f, _i, _i, _i, _i, _f, _i, _d, _d, _d, _d,
                                              _d, _d]
            lib.ref_evaluate_both.restype = _f
            lib.ref_num_threads.restype = _i
            lib.ref_set_threads.argtypes = [_i]
        if lib is not None and name == "libwdlref_log.so":
            lib.ref_evaluate_both_log.argtypes = [_i, _i, _f, _i, _i, _i, _i, _f, _i, _d, _d, _d, _d,
                                                  _d, _d]
            lib.ref_evaluate_both_log.restype = _f
    return _libs[name]


def available() -> bool:
    return _lib("libwdlref.so") is not None and _lib("libwdlref_log.so") is not None


def _p(a):
    return None if a is None else a.ctypes.data_as(_d)


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def kernel_table(W, H, gamma):
    out = np.zeros(max(W, H))
    _lib("libwdlref.so").ref_kernel_table(W, H, gamma, _p(out))
    return out


def _conv(fn, u, W, H, gamma):
    u = _c(u)
    out = np.zeros_like(u)
    getattr(_lib("libwdlref.so"), fn)(W, H, gamma, _p(u), _p(out), u.size // (W * H))
    return out


def convolve(u, W, H, gamma):
    return _conv("ref_convolve", u, W, H, gamma)


def convolve1d(u, W, gamma):
    u = _c(u)
    out = np.zeros_like(u)
    _lib("libwdlref.so").ref_convolve1d(W, gamma, _p(u), _p(out), u.size // W)
    return out


def convolve3d(u, W, H, D, gamma):
    u = _c(u)
    out = np.zeros_like(u)
    _lib("libwdlref.so").ref_convolve3d(W, H, D, gamma, _p(u), _p(out), u.size // (W * H * D))
    return out


def sse_convolve(u, W, H, gamma):
    return _conv("ref_sse_convolve", u, W, H, gamma)


def log_convolve(u, W, H, gamma):
    return _conv("ref_log_convolve", u, W, H, gamma)


def log_convolve_signarray(u, signs, W, H, gamma):
    u = _c(u)
    s = np.array(signs, dtype=np.uint8, copy=True)
    out = np.zeros_like(u)
    _lib("libwdlref.so").ref_log_convolve_signarray(W, H, gamma, _p(u), s.ctypes.data_as(_u8),
                                                    _p(out), u.size // (W * H))
    return out, s


def abs_log(v, signs):
    v = _c(v)
    s = np.array(signs, dtype=np.uint8, copy=True)
    out = np.zeros_like(v)
    _lib("libwdlref.so").ref_abs_log(_p(v), s.ctypes.data_as(_u8), _p(out), v.size)
    return out, s


def abs_exp(v, signs):
    v = _c(v)
    s = np.array(signs, dtype=np.uint8, copy=True)
    out = np.zeros_like(v)
    _lib("libwdlref.so").ref_abs_exp(_p(v), s.ctypes.data_as(_u8), _p(out), v.size)
    return out


def gradient_both(atoms, lam, obs, W, H, gamma, L, loss=2, b0=None, variant="plain"):
    """variant: 'plain' (Gaussian2DKernel), 'sse' (GaussianSSE2DKernel), 'log' (log signArray)."""
    atoms, lam, obs = _c(atoms), _c(lam), _c(obs)
    K = lam.size
    N = W * H
    b0 = None if b0 is None else _c(b0)
    gl, gd, bary, bl = np.zeros(K), np.zeros(K * N), np.zeros(N), np.zeros(K * N)
    fn = {"plain": "ref_gradient_both", "sse": "ref_gradient_both_sse",
          "log": "ref_gradient_both_log"}[variant]
    getattr(_lib("libwdlref.so"), fn)(W, H, gamma, K, L, loss, _p(atoms), _p(lam), _p(obs),
                                       _p(b0), _p(gl), _p(gd), _p(bary), _p(bl))
    return dict(grad_lambda=gl, grad_dic=gd.reshape(K, N), bary=bary, b_last=bl.reshape(K, N))


def evaluate_both(x, observed, W, H, gamma, K, P, L, loss=2, scale_dic=1.0, exp_weight=True,
                  b_storage=None, log_domain=False, kernel="plain"):
    x, observed = _c(x), _c(observed)
    N = W * H
    g = np.zeros_like(x)
    atoms_init = np.full(K * N, 1.0 / N)
    bs = None if b_storage is None else _c(b_storage).copy()
    bt = None if b_storage is None else np.zeros(P * K * N)
    if log_domain:
        f = _lib("libwdlref_log.so").ref_evaluate_both_log(
            W, H, gamma, K, P, L, loss, scale_dic, int(exp_weight), _p(atoms_init), _p(observed),
            _p(bs), _p(bt), _p(x), _p(g))
    else:
        f = _lib("libwdlref.so").ref_evaluate_both(
            1 if kernel == "sse" else 0, W, H, gamma, K, P, L, loss, scale_dic, int(exp_weight),
            _p(atoms_init), _p(observed), _p(bs), _p(bt), _p(x), _p(g))
    return f, g, (None if bt is None else bt.reshape(P, K, N))


def num_threads() -> int:
    return int(_lib("libwdlref.so").ref_num_threads())


def set_threads(n: int) -> None:
    _lib("libwdlref.so").ref_set_threads(int(n))
