"""Parity of the tiled path (images with a side above 128: 64 x 64 output tiles
with a halo, one workgroup each) with the FP64 oracle, and with the
LDS-resident path on sizes both can run (WDL_FLAG_TILED forces the tiles).

Tolerances: the drop-in convolution max |gpu - ref| / max |ref| <= 2e-6 (FP32
MFMA); solver outputs relative L2 <= 1e-5 against the oracle, as for the
resident path.  Tiled and resident kernels share their arithmetic per pixel
(same factor chunks in the same order, the halo adds only exact zeros), so
the standard-domain barycenters agree bit for bit and the gradients to 1e-6
(the weight-gradient dots are summed per tile first; log-domain shifts are
per tile region, which moves dL/dlambda by up to a few 1e-6)."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import maxrel, rel
from wassersteindictionarylearning_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("W,H,g,nv", [(256, 256, 1.0, 2), (200, 130, 0.5, 3), (129, 300, 2.0, 1),
                                      (130, 70, 1.0, 2), (512, 512, 2.0, 1), (1000, 8, 1.0, 1)])
def test_tiled_convolve_matches_oracle(ctx, W, H, g, nv):
    print("tiled convolve W=%d H=%d gamma=%g nv=%d" % (W, H, g, nv), flush=True)
    u = np.random.default_rng(W * 1000 + H).uniform(0, 1, nv * W * H)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    assert maxrel(k.convolve(u), O.convolve(u, W, H, g)) < 2e-6
    k.close()


@pytest.mark.parametrize("W,H", [(256, 192), (130, 140)])
def test_tiled_convolve_device(ctx, W, H):
    # FP32 device form: float4 groups when W % 4 == 0, scalar otherwise.
    from wassersteindictionarylearning_amd import _hip
    g, nv = 1.0, 7
    print("tiled convolve_device W=%d H=%d nv=%d" % (W, H, nv), flush=True)
    u = np.random.default_rng(11).uniform(0, 1, (nv, H, W))
    du = _hip.DeviceArray.from_host(u)
    dout = _hip.DeviceArray(u.shape)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H)
    k.convolve_device(du.ptr, dout.ptr, nv)
    ctx.synchronize()
    assert maxrel(dout.download(), O.convolve(u, W, H, g)) < 2e-6
    du.free()
    dout.free()
    k.close()


def run(ctx, atoms, lam, obs, W, H, g, L, loss, flags=0, mode=wdl.RUN_BATCHED, iters=None,
        streams=None):
    P, K = lam.shape
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, loss, flags)
    if streams:
        s.set_streams(streams)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    if iters is not None:
        s.set_iterations(iters)
    s.run(mode)
    r = s.results()
    if not flags & wdl.FLAG_WEIGHTS_ONLY:
        r["grad_dic"] = s.grad_dic()
    if flags & wdl.FLAG_LOG_DOMAIN:
        r["events"] = s.log_range_events()
    if flags & wdl.FLAG_WARM_RESTART:
        r["b_temp"] = s.get_warm_state()
    s.close()
    return r


def problem(W, H, K, P, seed, curves=False):
    rng = np.random.default_rng(seed)
    if curves and W == H:
        obs = synthetic.curve_histograms(rng, W, P)
    elif W == H:
        obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, W / 6))
    else:
        obs = rng.uniform(0.2, 1, (P, W * H))
        obs /= obs.sum(1, keepdims=True)
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    return atoms, lam, obs


@pytest.mark.parametrize("W,H,K,P,L,g,logd", [
    (128, 128, 3, 9, 30, 1.0, 0), (96, 80, 2, 3, 12, 0.5, 0), (70, 66, 2, 2, 10, 2.0, 0),
    (128, 128, 3, 9, 30, 1.0, 1), (96, 80, 2, 3, 12, 0.5, 1)])
def test_tiled_equals_resident(ctx, W, H, K, P, L, g, logd):
    print("tiled vs resident W=%d H=%d K=%d P=%d L=%d log=%d" % (W, H, K, P, L, logd), flush=True)
    atoms, lam, obs = problem(W, H, K, P, 100 + W + L)
    loss = 4 if logd else 2
    f = wdl.FLAG_LOG_DOMAIN if logd else 0
    a = run(ctx, atoms, lam, obs, W, H, g, L, loss, f)
    b = run(ctx, atoms, lam, obs, W, H, g, L, loss, f | wdl.FLAG_TILED)
    for k in ("bary", "grad_dic", "grad_dic_sum", "loss"):
        assert rel(b[k], a[k]) < 1e-6, (k, rel(b[k], a[k]))
    # dL/dlambda: each within ~3e-6 of the FP64 oracle on these cases
    assert rel(b["grad_lambda"], a["grad_lambda"]) < (5e-6 if logd else 1e-6)
    if logd:  # the tile's shifts come from its region rows, so roundings differ
        assert a["events"] == 0 and b["events"] == 0
    else:
        np.testing.assert_array_equal(b["bary"], a["bary"])


@pytest.mark.parametrize("W,H,K,P,L,g,loss", [
    (192, 192, 3, 2, 10, 1.0, 2), (160, 136, 2, 3, 8, 0.5, 4), (130, 130, 2, 2, 6, 2.0, 1)])
def test_tiled_gradient_matches_oracle(ctx, W, H, K, P, L, g, loss):
    print("tiled gradient W=%d H=%d K=%d P=%d L=%d gamma=%g loss=%d" % (W, H, K, P, L, g, loss),
          flush=True)
    atoms, lam, obs = problem(W, H, K, P, W + 3 * H + L)
    r = run(ctx, atoms, lam, obs, W, H, g, L, loss)
    for p in range(P):
        o = O.gradient_both(atoms, lam[p], obs[p], W, H, g, L, loss)
        assert rel(r["bary"][p], o["bary"]) < TOL, p
        assert rel(r["grad_lambda"][p], o["grad_lambda"]) < TOL, p
        assert rel(r["grad_dic"][p], o["grad_dic"]) < TOL, p
        assert abs(r["loss"][p] - O.loss_value(loss, o["bary"], obs[p])) < TOL * abs(r["loss"][p])


@pytest.mark.parametrize("W,H,K,P,L,g,loss,curves", [
    (136, 136, 2, 2, 6, 0.5, 4, True), (144, 132, 2, 1, 5, 1.0, 2, False)])
def test_tiled_log_gradient_matches_oracle(ctx, W, H, K, P, L, g, loss, curves):
    print("tiled log gradient W=%d H=%d K=%d P=%d L=%d gamma=%g" % (W, H, K, P, L, g), flush=True)
    atoms, lam, obs = problem(W, H, K, P, 7 * W + L, curves)
    r = run(ctx, atoms, lam, obs, W, H, g, L, loss, wdl.FLAG_LOG_DOMAIN)
    assert r["events"] == 0
    for p in range(P):
        o = O.gradient_both_log(atoms, lam[p], obs[p], W, H, g, L, loss)
        assert rel(r["bary"][p], o["bary"]) < TOL, p
        assert rel(r["grad_lambda"][p], o["grad_lambda"]) < TOL, p
        assert rel(r["grad_dic"][p], o["grad_dic"]) < TOL, p


@pytest.mark.parametrize("logd", [0, 1])
def test_tiled_plain_equals_batched(ctx, logd):
    # One solve per launch sequence vs all P together on 2 streams: same bits.
    W = H = 160
    K, P, L, g = 2, 9, 6, 1.0
    print("tiled plain vs batched log=%d" % logd, flush=True)
    atoms, lam, obs = problem(W, H, K, P, 31)
    f = wdl.FLAG_LOG_DOMAIN if logd else 0
    a = run(ctx, atoms, lam, obs, W, H, g, L, 2, f, mode=wdl.RUN_PLAIN)
    b = run(ctx, atoms, lam, obs, W, H, g, L, 2, f, streams=2)
    for k in ("grad_lambda", "grad_dic", "bary", "loss"):
        np.testing.assert_array_equal(b[k], a[k], err_msg=k)
    one = run(ctx, atoms, lam[4:5], obs[4:5], W, H, g, L, 2, f)
    np.testing.assert_array_equal(one["grad_lambda"][0], a["grad_lambda"][4])
    np.testing.assert_array_equal(one["grad_dic"][0], a["grad_dic"][4])


@pytest.mark.parametrize("logd", [0, 1])
def test_tiled_uneven_iterations_weights_only_warm(ctx, logd):
    W = H = 40
    K, P, L, g = 3, 5, 9, 1.0
    iters = np.array([5, 9, 7, 6, 9], np.int32)
    print("tiled uneven iterations log=%d" % logd, flush=True)
    atoms, lam, obs = problem(W, H, K, P, 77)
    f = wdl.FLAG_TILED | wdl.FLAG_WARM_RESTART | (wdl.FLAG_LOG_DOMAIN if logd else 0)
    full = run(ctx, atoms, lam, obs, W, H, g, L, 2, f, iters=iters)
    wo = run(ctx, atoms, lam, obs, W, H, g, L, 2, f | wdl.FLAG_WEIGHTS_ONLY, iters=iters)
    np.testing.assert_array_equal(wo["grad_lambda"], full["grad_lambda"])
    np.testing.assert_array_equal(wo["bary"], full["bary"])
    solve = O.gradient_both_log if logd else O.gradient_both
    ones = np.ones((K, W * H))
    for p in range(P):
        o = solve(atoms, lam[p], obs[p], W, H, g, int(iters[p]), 2, b0=ones)
        assert rel(full["grad_lambda"][p], o["grad_lambda"]) < TOL, p
        assert rel(full["grad_dic"][p], o["grad_dic"]) < TOL, p
        assert rel(full["bary"][p], o["bary"]) < TOL, p
        assert rel(full["b_temp"][p], o["b_last"]) < TOL, p


def test_tiled_errors(ctx):
    with pytest.raises(wdl.WdlError):  # the dense flag is for the resident path only
        wdl.SinkhornSolver(ctx, 256, 256, 2, 1, 3, 1.0, 2, wdl.FLAG_DENSE)
    for flags in (0, wdl.FLAG_LOG_DOMAIN):  # gamma 12: factors reach 35 pixels, halo above 32
        with pytest.raises(wdl.WdlError):
            wdl.SinkhornSolver(ctx, 256, 256, 2, 1, 3, 12.0, 2, flags)


def test_log_config_full_size_against_reference(ctx):
    # The small-epsilon log-domain configuration at its full spatial size (the
    # benchmark's own inputs: 256x256 curve histograms, K=6, gamma=0.5, KL loss,
    # tiled log-domain kernels), with 8 Sinkhorn iterations instead of 200 so the
    # compiled reference's FP64 compute_gradient_both_log_signArray (oracle/_ref)
    # finishes in seconds; one histogram of two.
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    c = synthetic.make_problem("log_small_eps", P=2)
    n, K, g, loss = c["n"], c["K"], c["gamma"], c["loss"]
    L = 8
    print("log configuration %dx%d K=%d L=%d" % (n, n, K, L), flush=True)
    r = run(ctx, c["atoms"], c["weights"], c["obs"], n, n, g, L, loss, wdl.FLAG_LOG_DOMAIN)
    assert r["events"] == 0
    o = ref.gradient_both(c["atoms"], c["weights"][0], c["obs"][0], n, n, g, L, loss,
                          variant="log")
    for k in ("bary", "grad_lambda", "grad_dic"):
        e = rel(r[k][0], o[k])
        print("%s: %.2e" % (k, e), flush=True)
        assert e < TOL, k
