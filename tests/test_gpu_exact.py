"""Parity of the exact-shift log-domain path (WDL_FLAG_EXACT_LSE: the
reference's exact per-output maximum, FP32 sums over windows outside which every term is below e^-41) with the
FP64 oracle, in the drop-in convolution and in the batched solver, including
inputs whose log values spread far beyond the shifted MFMA form's range, lines
of one and two positions, lines of 1024, and lines that are -inf throughout.

Tolerances: unsigned log-convolution outputs within 2e-6 absolute (relative
in linear terms); signed outputs within 1e-6 of the magnitude scale, i.e.
|s exp(out - A) - s' exp(ref - A)| <= 1e-6 with A the log-convolution of |u|;
solver outputs relative L2 <= 1e-5."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import fixtures, rel
from wassersteindictionarylearning_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-5
EXACT = wdl.FLAG_LOG_DOMAIN | wdl.FLAG_EXACT_LSE


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def close_log(a, b, tol=2e-6):
    fin = np.isfinite(b)
    np.testing.assert_array_equal(np.isfinite(a), fin)
    np.testing.assert_array_equal(a[~fin], b[~fin])
    assert np.max(np.abs(a[fin] - b[fin]), initial=0) < tol


def close_signed(out, s, ref, sref, scale, n, tol=1e-6):
    pos = O.unpack_signs(s, n)
    pref = O.unpack_signs(sref, n)
    with np.errstate(invalid="ignore", over="ignore"):
        a = np.where(pos, 1.0, -1.0) * np.exp(out - scale)
        b = np.where(pref, 1.0, -1.0) * np.exp(ref - scale)
    fin = np.isfinite(scale)
    assert np.max(np.abs(a[fin] - b[fin]), initial=0) < tol


def spread_input(rng, W, H, nv, slope):
    # log values along steep ramps (|x| up to slope * side) with noise and some -inf
    y, x = np.mgrid[0:H, 0:W]
    out = []
    for v in range(nv):
        sx, sy = rng.uniform(-slope, slope, 2)
        f = sx * x + sy * y + rng.normal(0, 3, (H, W))
        f[rng.uniform(0, 1, (H, W)) < 0.02] = -np.inf
        out.append(f.ravel())
    return np.concatenate(out)


@pytest.mark.parametrize("W,H,g,nv,slope", [(64, 64, 1.0, 2, 0.5), (37, 50, 0.5, 3, 30.0),
                                            (128, 96, 2.0, 1, 80.0), (300, 20, 1.0, 1, 10.0),
                                            (1, 7, 1.0, 2, 5.0), (2, 3, 0.5, 2, 5.0),
                                            (1024, 3, 1.0, 1, 2.0), (5, 1024, 4.0, 1, 0.2)])
def test_exact_log_convolve_matches_oracle(ctx, W, H, g, nv, slope):
    print("exact log_convolve W=%d H=%d slope=%g" % (W, H, slope), flush=True)
    rng = np.random.default_rng(W * 7 + H)
    x = spread_input(rng, W, H, nv, slope)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H, wdl.FLAG_EXACT_LSE)
    ref = O.log_convolve(x, W, H, g)
    close_log(k.log_convolve(x), ref)
    signs = rng.integers(0, 256, (x.size + 7) // 8).astype(np.uint8)
    out, s = k.log_convolve_signArray(x, signs)
    o2, s2 = O.log_convolve_signarray(x, signs, W, H, g)
    close_signed(out, s, o2, s2, ref, x.size)
    k.close()


def test_exact_log_convolve_infinite_lines(ctx):
    # whole rows and columns of -inf (zero histogram bins) next to finite ones
    print("exact log_convolve -inf lines", flush=True)
    rng = np.random.default_rng(5)
    W, H, g = 70, 40, 1.0
    x = rng.normal(0, 20, (H, W))
    x[3] = -np.inf
    x[:, 10:13] = -np.inf
    x[20:, 40:] = -np.inf
    x = x.ravel()
    k = wdl.HipGaussian2DKernel(ctx, g, W, H, wdl.FLAG_EXACT_LSE)
    ref = O.log_convolve(x, W, H, g)
    close_log(k.log_convolve(x), ref)
    signs = rng.integers(0, 256, (x.size + 7) // 8).astype(np.uint8)
    out, s = k.log_convolve_signArray(x, signs)
    o2, s2 = O.log_convolve_signarray(x, signs, W, H, g)
    close_signed(out, s, o2, s2, ref, x.size)
    xi = np.full(W * H, -np.inf)
    np.testing.assert_array_equal(k.log_convolve(xi), xi)
    k.close()


@pytest.mark.parametrize("name,d", fixtures("lconv"))
def test_exact_log_convolve_golden(ctx, name, d):
    W, H, g = int(d["W"]), int(d["H"]), float(d["gamma"])
    print("exact log_convolve golden", name, flush=True)
    k = wdl.HipGaussian2DKernel(ctx, g, W, H, wdl.FLAG_EXACT_LSE)
    close_log(k.log_convolve(d["u"]), d["out"])


def run(ctx, atoms, lam, obs, W, H, g, L, loss, flags, iters=None):
    P, K = lam.shape
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, loss, flags)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    if iters is not None:
        s.set_iterations(iters)
    s.run()
    r = s.results()
    if not flags & wdl.FLAG_WEIGHTS_ONLY:
        r["grad_dic"] = s.grad_dic()
    r["events"] = s.log_range_events()
    if flags & wdl.FLAG_WARM_RESTART:
        r["b_temp"] = s.get_warm_state()
    s.close()
    return r


@pytest.mark.parametrize("W,H,K,P,L,g,loss", [(24, 24, 3, 3, 10, 1.0, 4), (40, 36, 2, 2, 8, 0.5, 2),
                                              (136, 136, 2, 1, 5, 1.0, 4)])
def test_exact_solver_matches_oracle(ctx, W, H, K, P, L, g, loss):
    print("exact solver W=%d H=%d K=%d P=%d L=%d" % (W, H, K, P, L), flush=True)
    rng = np.random.default_rng(W + L)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, W / 6)) if W == H else \
        rng.uniform(0.2, 1, (P, W * H))
    obs = obs / obs.sum(1, keepdims=True)
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    r = run(ctx, atoms, lam, obs, W, H, g, L, loss, EXACT)
    for p in range(P):
        o = O.gradient_both_log(atoms, lam[p], obs[p], W, H, g, L, loss)
        assert rel(r["bary"][p], o["bary"]) < TOL, p
        assert rel(r["grad_lambda"][p], o["grad_lambda"]) < TOL, p
        assert rel(r["grad_dic"][p], o["grad_dic"]) < TOL, p


def test_exact_solver_far_transport(ctx):
    # Sharp blob atoms at gamma = 1: the scalings diverge (the standard domain
    # overflows, log values along a line spread over hundreds of nats); the
    # shifted MFMA form counts range events, the exact-shift path matches FP64.
    W = H = 32
    K, P, L, g = 3, 2, 30, 1.0
    print("exact solver, far transport", flush=True)
    rng = np.random.default_rng(5)
    atoms = synthetic.blob_histograms(rng, W, K, (1, 2), (1.0, 2.0))
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (1.0, 3.0))
    lam = np.full((P, K), 1.0 / K)
    mf = run(ctx, atoms, lam, obs, W, H, g, L, 2, wdl.FLAG_LOG_DOMAIN)
    ex = run(ctx, atoms, lam, obs, W, H, g, L, 2, EXACT)
    sd = run(ctx, atoms, lam, obs, W, H, g, L, 2, 0)
    print("range events: shifted MFMA %d, standard domain %d, exact %d"
          % (mf["events"], sd["events"], ex["events"]), flush=True)
    assert mf["events"] > 0  # log b spreads over ~200 nats along a line here
    assert sd["events"] > 0  # b ~ e^200 is beyond FP32
    assert ex["events"] == 0
    # dL/dlambda is ill-conditioned here: the dot products sum terms up to ~9e2
    # times larger than a component (0.0011 from terms summing to 0.74), so the
    # FP32 rounding of g (~1e-7) shows at ~1e-5; its bound here is 3e-5
    for p in range(P):
        o = O.gradient_both_log(atoms, lam[p], obs[p], W, H, g, L, 2)
        assert rel(ex["bary"][p], o["bary"]) < TOL, p
        assert rel(ex["grad_lambda"][p], o["grad_lambda"]) < 3e-5, p
        assert rel(ex["grad_dic"][p], o["grad_dic"]) < TOL, p


@pytest.mark.parametrize("W,H", [(20, 20), (48, 48)])
def test_exact_uneven_weights_only_warm(ctx, W, H):
    # the weights-only backward (one fused kernel per step) against the full
    # one (separate dot and input kernels) bit for bit, at images of fewer and
    # of more pixels than a workgroup has threads
    K, P, L, g = 3, 4, 9, 1.0
    iters = np.array([5, 9, 7, 6], np.int32)
    print("exact uneven iterations W=%d H=%d" % (W, H), flush=True)
    rng = np.random.default_rng(77)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, 5))
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    f = EXACT | wdl.FLAG_WARM_RESTART
    full = run(ctx, atoms, lam, obs, W, H, g, L, 2, f, iters)
    wo = run(ctx, atoms, lam, obs, W, H, g, L, 2, f | wdl.FLAG_WEIGHTS_ONLY, iters)
    np.testing.assert_array_equal(wo["grad_lambda"], full["grad_lambda"])
    ones = np.ones((K, W * H))
    for p in range(P):
        o = O.gradient_both_log(atoms, lam[p], obs[p], W, H, g, int(iters[p]), 2, b0=ones)
        assert rel(full["grad_lambda"][p], o["grad_lambda"]) < TOL, p
        assert rel(full["grad_dic"][p], o["grad_dic"]) < TOL, p
        assert rel(full["b_temp"][p], o["b_last"]) < TOL, p


def test_exact_projection_config_against_reference(ctx):
    # The projection configuration at full size: 128x128, a K=8 blob dictionary,
    # gamma=1, weights gradient only, warm-restart store at its initial ones,
    # exact shifts, each request with its own L_p.  The shortest and the longest
    # of 8 requests against the compiled reference's FP64
    # compute_gradient_both_log_signArray (oracle/_ref), whose weight gradient the
    # weights-only chain computes.
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    c = synthetic.make_problem("projection", P=8)
    n, K, L, g, loss = c["n"], c["K"], c["L"], c["gamma"], c["loss"]
    atoms = synthetic.blob_histograms(np.random.default_rng(7), n, K, c["nblobs"], c["sigma"])
    iters = c["iterations"]
    print("exact projection configuration %dx%d K=%d L_p=%s" % (n, n, K, list(iters)), flush=True)
    f = EXACT | wdl.FLAG_WEIGHTS_ONLY | wdl.FLAG_WARM_RESTART
    r = run(ctx, atoms, c["weights"], c["obs"], n, n, g, L, loss, f, iters)
    assert r["events"] == 0
    for p in (int(np.argmin(iters)), int(np.argmax(iters))):
        o = ref.gradient_both(atoms, c["weights"][p], c["obs"][p], n, n, g, int(iters[p]), loss,
                              variant="log")
        e = rel(r["grad_lambda"][p], o["grad_lambda"])
        print("request %d, L_p=%d: weight-gradient error %.2e" % (p, iters[p], e), flush=True)
        assert e < TOL, p
        assert rel(r["b_temp"][p], o["b_last"]) < TOL, p


@pytest.mark.parametrize("weights_only", [False, True])
def test_exact_many_groups_and_plain_form(ctx, weights_only):
    # P=36 histograms in 1 and 3 stream groups and in the plain form (one solve
    # per launch sequence): the group offsets of the scratch images, iteration
    # counts and dot partial sums give the same bits in every split, and one
    # histogram of each of the 3 groups matches the oracle
    W = H = 16
    K, P, L, g = 3, 36, 8, 1.0
    rng = np.random.default_rng(36)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, 4))
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    iters = rng.integers(5, L + 1, P).astype(np.int32)
    f = EXACT | wdl.FLAG_WARM_RESTART | (wdl.FLAG_WEIGHTS_ONLY if weights_only else 0)
    print("exact, P=%d in 1 and 3 groups and plain, weights only %s" % (P, weights_only), flush=True)
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, 2, f)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    s.set_iterations(iters)
    out = {}
    for name, groups, mode in (("one", 1, wdl.RUN_BATCHED), ("three", 3, wdl.RUN_BATCHED),
                               ("plain", 3, wdl.RUN_PLAIN)):
        s.set_streams(groups)
        s.run(mode)
        r = s.results()
        r["b_temp"] = s.get_warm_state()
        if not weights_only:
            r["grad_dic"] = s.grad_dic()
        out[name] = r
    s.close()
    for name in ("three", "plain"):
        for k in out["one"]:
            np.testing.assert_array_equal(out[name][k], out["one"][k], err_msg="%s %s" % (name, k))
    ones = np.ones((K, W * H))
    for p in (0, 12 + 5, P - 1):  # one histogram of each of the 3 groups
        o = O.gradient_both_log(atoms, lam[p], obs[p], W, H, g, int(iters[p]), 2, b0=ones)
        assert rel(out["one"]["grad_lambda"][p], o["grad_lambda"]) < TOL, p
        assert rel(out["one"]["b_temp"][p], o["b_last"]) < TOL, p
        if not weights_only:
            assert rel(out["one"]["grad_dic"][p], o["grad_dic"]) < TOL, p
            assert rel(out["one"]["bary"][p], o["bary"]) < TOL, p
