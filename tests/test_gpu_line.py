"""Parity of the 1-D and 3-D KernelTypes (wdl_gauss1d / wdl_gauss3d: the
drop-ins for Gaussian1DKernel and Gaussian3DKernel, kernels.h:501-815) with
the FP64 oracle and the compiled reference's golden vectors.

Tolerance: FP32 MFMA passes, as the 2-D operator -- max |out - ref| / max |ref|
<= 2e-6 (three passes for 3-D: 3e-6)."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import fixtures, maxrel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("W,g,nv", [(37, 2.0, 3), (128, 25.0, 5), (129, 1.0, 2), (300, 9.0, 4),
                                    (2048, 4.0, 2), (5, 1000.0, 70), (64, 0.3, 200)])
def test_convolve1d_matches_oracle(ctx, W, g, nv):
    print("convolve1d W=%d gamma=%g nv=%d" % (W, g, nv), flush=True)
    u = np.random.default_rng(W + nv).uniform(0, 1, nv * W)
    k = wdl.HipGaussian1DKernel(ctx, g, W)
    assert maxrel(k.convolve(u), O.convolve1d(u, W, g)) < 2e-6
    k.close()


@pytest.mark.parametrize("name,d", fixtures("conv1d"))
def test_convolve1d_golden(ctx, name, d):
    W, g = int(d["W"]), float(d["gamma"])
    print("convolve1d golden", name, flush=True)
    k = wdl.HipGaussian1DKernel(ctx, g, W)
    assert maxrel(k.convolve(d["u"]), d["out"]) < 2e-6
    k.close()


@pytest.mark.parametrize("W,H,D,g,nv", [(8, 6, 5, 2.0, 2), (20, 16, 12, 1.0, 3),
                                        (64, 64, 32, 4.0, 1), (6, 5, 136, 4.0, 2),
                                        (130, 70, 4, 1.0, 1), (16, 16, 1, 2.0, 3)])
def test_convolve3d_matches_oracle(ctx, W, H, D, g, nv):
    print("convolve3d W=%d H=%d D=%d gamma=%g nv=%d" % (W, H, D, g, nv), flush=True)
    u = np.random.default_rng(W * H + D).uniform(0, 1, nv * W * H * D)
    k = wdl.HipGaussian3DKernel(ctx, g, W, H, D)
    assert maxrel(k.convolve(u), O.convolve3d(u, W, H, D, g)) < 3e-6
    k.close()


@pytest.mark.parametrize("name,d", fixtures("conv3d"))
def test_convolve3d_golden(ctx, name, d):
    W, H, D, g = int(d["W"]), int(d["H"]), int(d["D"]), float(d["gamma"])
    print("convolve3d golden", name, flush=True)
    k = wdl.HipGaussian3DKernel(ctx, g, W, H, D)
    assert maxrel(k.convolve(d["u"]), d["out"]) < 3e-6
    k.close()


def test_device_forms(ctx):
    # FP32 device pointers, ordered on the context stream; the FP64 drop-in
    # computes in FP32 too, so the two agree to the bit.
    from wassersteindictionarylearning_amd import _hip
    print("convolve1d / convolve3d device forms", flush=True)
    rng = np.random.default_rng(3)
    for k, n, nv in ((wdl.HipGaussian1DKernel(ctx, 2.0, 200), 200, 6),
                     (wdl.HipGaussian3DKernel(ctx, 1.5, 24, 20, 18), 24 * 20 * 18, 2)):
        u = rng.uniform(0, 1, nv * n).astype(np.float32)
        du, dout = _hip.DeviceArray.from_host(u), _hip.DeviceArray(u.shape)
        k.convolve_device(du.ptr, dout.ptr, nv)
        ctx.synchronize()
        np.testing.assert_array_equal(dout.download(), k.convolve(u.astype(np.float64)))
        du.free()
        dout.free()
        k.close()


def test_errors(ctx):
    with pytest.raises(wdl.WdlError):
        wdl.HipGaussian1DKernel(ctx, 1.0, 4096)  # longer than 2048
    with pytest.raises(wdl.WdlError):
        wdl.HipGaussian1DKernel(ctx, 5000.0, 300)  # band beyond the table for a long line
    with pytest.raises(wdl.WdlError):
        wdl.HipGaussian3DKernel(ctx, -1.0, 8, 8, 8)
