"""Parity of the HIP separable Gaussian convolution with the FP64 oracle and the
reference's golden vectors.  The HIP path computes in FP32 (MFMA), so results
are compared at FP32 tolerance: max |gpu - ref| / max |ref| <= 2e-6."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import fixtures, maxrel

pytestmark = pytest.mark.gpu
TOL = 2e-6


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("W,H,g,nv", [(8, 8, 1.0, 1), (12, 9, 2.0, 3), (16, 16, 0.5, 2),
                                      (33, 17, 1.0, 2), (64, 64, 2.0, 4), (128, 128, 1.0, 5),
                                      (100, 77, 25.0, 2), (128, 64, 0.5, 3), (1, 5, 1.0, 2),
                                      (127, 128, 4.0, 1)])
def test_convolve_matches_oracle(ctx, W, H, g, nv):
    print("convolve W=%d H=%d gamma=%g nv=%d" % (W, H, g, nv), flush=True)
    u = np.random.default_rng(W * 1000 + H).uniform(0, 1, nv * W * H)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    out = k.convolve(u)
    assert maxrel(out, O.convolve(u, W, H, g)) < TOL
    np.testing.assert_allclose(k.kernel1d(), O.kernel_table(W, H, g), rtol=1e-15)
    k.close()


@pytest.mark.parametrize("name,d", fixtures("conv"))
def test_convolve_golden(ctx, name, d):
    W, H, g = int(d["W"]), int(d["H"]), float(d["gamma"])
    print("convolve golden", name, flush=True)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    assert maxrel(k.convolve(d["u"]), d["out"]) < TOL
    assert maxrel(k.convolveAdjoint(d["u"]), d["out"]) < TOL


@pytest.mark.parametrize("W,H,g", [(128, 128, 1.0), (64, 48, 0.5), (96, 96, 2.0)])
def test_band_skip_is_exact(ctx, W, H, g):
    # Chunks skipped under the band hold only exact FP32 zeros of the factor,
    # so the banded and the dense product agree bit for bit.
    print("band vs dense W=%d H=%d" % (W, H), flush=True)
    u = np.random.default_rng(7).uniform(0, 1, 3 * W * H)
    kb = wdl.HipGaussian2DKernel(ctx, g, W, H)
    kd = wdl.HipGaussian2DKernel(ctx, g, W, H, flags=wdl.FLAG_DENSE)
    np.testing.assert_array_equal(kb.convolve(u), kd.convolve(u))


def test_convolve_device(ctx):
    from wassersteindictionarylearning_amd import _hip
    W, H, g, nv = 128, 128, 1.0, 40
    print("convolve_device nv=%d" % nv, flush=True)
    u = np.random.default_rng(9).uniform(0, 1, (nv, H, W))
    du = _hip.DeviceArray.from_host(u)
    dout = _hip.DeviceArray(u.shape)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    k.convolve_device(du.ptr, dout.ptr, nv)
    ctx.synchronize()
    assert maxrel(dout.download(), O.convolve(u, W, H, g)) < TOL
    du.free()
    dout.free()


def test_errors_are_reported(ctx):
    with pytest.raises(wdl.WdlError):
        wdl.HipGaussian2DKernel(ctx, -1.0, 8, 8)
    with pytest.raises(wdl.WdlError):  # tiled size, band half-width 71 > the 32-pixel halo
        wdl.HipGaussian2DKernel(ctx, 50.0, 200, 8)
