"""Parity of the log-domain KernelType entry points (log_convolve and
log_convolve_signArray, FP64 on the GPU) with the reference's golden vectors
and the FP64 oracle.  Both sum in the reference's order, so finite outputs
agree to ~1e-12 absolute, -inf outputs and every sign bit exactly."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import fixtures

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def same_log(a, b, tol=1e-11):
    fin = np.isfinite(b)
    np.testing.assert_array_equal(np.isfinite(a), fin)
    np.testing.assert_array_equal(a[~fin], b[~fin])
    assert np.max(np.abs(a[fin] - b[fin]), initial=0) < tol


@pytest.mark.parametrize("name,d", fixtures("lconv"))
def test_log_convolve_golden(ctx, name, d):
    W, H, g = int(d["W"]), int(d["H"]), float(d["gamma"])
    print("log_convolve golden", name, flush=True)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    same_log(k.log_convolve(d["u"]), d["out"])
    same_log(k.log_convolveAdjoint(d["u"]), d["out"])


@pytest.mark.parametrize("name,d", fixtures("sconv"))
def test_log_convolve_signarray_golden(ctx, name, d):
    W, H, g = int(d["W"]), int(d["H"]), float(d["gamma"])
    print("log_convolve_signArray golden", name, flush=True)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    out, s = k.log_convolve_signArray(d["u"], d["signs_in"])
    np.testing.assert_array_equal(s, d["signs_out"])
    same_log(out, d["out"], 1e-10)


@pytest.mark.parametrize("W,H,g,nv", [(128, 128, 0.5, 2), (37, 64, 2.0, 3), (64, 100, 1.0, 1)])
def test_log_convolve_matches_oracle(ctx, W, H, g, nv):
    print("log_convolve W=%d H=%d" % (W, H), flush=True)
    rng = np.random.default_rng(W + H)
    x = rng.uniform(-40, 40, nv * W * H)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    same_log(k.log_convolve(x), O.log_convolve(x, W, H, g), 1e-10)
    signs = rng.integers(0, 256, (x.size + 7) // 8 + 1).astype(np.uint8)  # one spare byte
    out, s = k.log_convolve_signArray(x, signs)
    o2, s2 = O.log_convolve_signarray(x, signs, W, H, g)
    np.testing.assert_array_equal(s, s2)
    same_log(out, o2, 1e-9)
