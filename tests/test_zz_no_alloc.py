"""The allocation-free log-domain convolution (wdl_gauss2d_log_convolve_noalloc)
against the allocating entry points it varies (wdl_gauss2d_log_convolve and
wdl_gauss2d_log_convolve_signarray, kernels.h:1073-1162), through the C ABI:
bit-equal outputs and sign arrays, two calls in a row with the second smaller,
for the reference-order FP64 passes and the exact-shift passes."""
import ctypes

import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from wassersteindictionarylearning_amd import _hip

pytestmark = pytest.mark.gpu


def _noalloc(k, log_u, signs=None):
    """One call of the allocation-free entry point on device copies; returns
    (out, signs_out)."""
    L = wdl.lib()
    u = np.ascontiguousarray(log_u, np.float64)
    nv = u.size // k.N
    din = _hip.DeviceArray.from_host(u, np.float64)
    dout = _hip.DeviceArray(u.shape, np.float64)
    dsg = None
    if signs is not None:
        dsg = _hip.DeviceArray.from_host(np.asarray(signs, np.uint8), np.uint8)
    st = L.wdl_gauss2d_log_convolve_noalloc(k.handle, din.ptr, dsg.ptr if dsg else None,
                                            dout.ptr, nv)
    assert st == 0, L.wdl_last_error()
    k.ctx.synchronize()
    out = dout.download()
    sg = dsg.download() if dsg else None
    for d in (din, dout, dsg):
        if d is not None:
            d.free()
    return out, sg


def _inputs(rng, nv, W, H, spread):
    u = rng.uniform(-spread, 0.0, (nv, H, W))
    u[0, 0, :3] = -np.inf  # zero entries of the linear values
    nbits = (nv * W * H + 7) // 8
    return u, rng.integers(0, 256, nbits).astype(np.uint8)


@pytest.mark.parametrize("flags", [0, wdl.FLAG_EXACT_LSE])
def test_noalloc_matches_allocating_entry(flags):
    rng = np.random.default_rng(5)
    ctx = wdl.Context(0)
    k_big = wdl.HipGaussian2DKernel(ctx, 1.0, 72, 56, flags)
    k_small = wdl.HipGaussian2DKernel(ctx, 2.0, 24, 40, flags)
    # two calls in a row, the second smaller (fewer images, then a smaller grid)
    for k, nv in ((k_big, 5), (k_big, 2), (k_small, 1)):
        u, bits = _inputs(rng, nv, k.W, k.H, 30.0)
        want = k.log_convolve(u)
        got, _ = _noalloc(k, u)
        assert np.array_equal(got, want, equal_nan=True), (flags, k.W, nv)
        want_s, want_bits = k.log_convolve_signArray(u, bits)
        got_s, got_bits = _noalloc(k, u, bits)
        assert np.array_equal(got_s, want_s, equal_nan=True), (flags, k.W, nv)
        assert np.array_equal(got_bits, want_bits), (flags, k.W, nv)
    ctx.close()


def test_noalloc_rejects_sizes_beyond_its_scratch():
    ctx = wdl.Context(0)
    k = wdl.HipGaussian2DKernel(ctx, 1.0, 256, 256)
    L = wdl.lib()
    nv = wdl.NOALLOC_MAX_VALUES // (256 * 256) + 1
    st = L.wdl_gauss2d_log_convolve_noalloc(k.handle, ctypes.c_void_p(16), None,
                                            ctypes.c_void_p(16), nv)
    assert st == -1 and b"WDL_NOALLOC_MAX_VALUES" in L.wdl_last_error()
    # the largest size that fits is accepted
    nv = wdl.NOALLOC_MAX_VALUES // (256 * 256)
    u = np.random.default_rng(1).uniform(-5, 0, (nv, 256, 256))
    got, _ = _noalloc(k, u)
    assert np.array_equal(got, k.log_convolve(u))
    ctx.close()
