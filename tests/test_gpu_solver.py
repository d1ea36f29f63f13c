"""Parity of the batched HIP Sinkhorn barycenter + unrolled gradient
(compute_gradient_both for P histograms) and of evaluate_both with the FP64
oracle and the reference's golden vectors.

Tolerance (FP32 pipeline, FP64 gradient dot products): relative L2 error
<= 1e-5 for the weight gradient, the dictionary gradient, the barycenters and
the loss, as stated for the hot path."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import fixtures, rel
from wassersteindictionarylearning_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-5


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def solve_gpu(ctx, atoms, lam, obs, W, H, g, L, loss, flags=0, mode=wdl.RUN_BATCHED, b0=None):
    P, K = lam.shape
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, loss, flags)
    s.set_observed(obs)
    if b0 is not None:
        s.set_warm_state(b0)
    s.set_problem(atoms, lam)
    s.run(mode)
    r = s.results()
    r["grad_dic"] = s.grad_dic()
    r["events"] = s.log_range_events()
    if b0 is not None:
        r["b_temp"] = s.get_warm_state()
    s.close()
    return r


def oracle_all(atoms, lam, obs, W, H, g, L, loss, b0=None):
    rs = [O.gradient_both(atoms, lam[p], obs[p], W, H, g, L, loss,
                          b0=None if b0 is None else b0[p]) for p in range(lam.shape[0])]
    return dict(grad_lambda=np.stack([r["grad_lambda"] for r in rs]),
                grad_dic=np.stack([r["grad_dic"] for r in rs]),
                bary=np.stack([r["bary"] for r in rs]),
                b_last=np.stack([r["b_last"] for r in rs]),
                loss=np.array([O.loss_value(loss, rs[p]["bary"], obs[p], W, H, g)
                               for p in range(len(rs))]))


def check(r, o):
    assert rel(r["bary"], o["bary"]) < TOL
    assert rel(r["loss"], o["loss"]) < TOL
    for p in range(o["grad_lambda"].shape[0]):
        assert rel(r["grad_lambda"][p], o["grad_lambda"][p]) < TOL, p
        assert rel(r["grad_dic"][p], o["grad_dic"][p]) < TOL, p
    assert rel(r["grad_dic_sum"], o["grad_dic"].sum(0)) < TOL


@pytest.mark.parametrize("W,H,K,P,L,g,loss", [
    (12, 12, 3, 2, 6, 2.0, 2), (16, 12, 2, 3, 10, 1.0, 4), (20, 20, 3, 9, 8, 0.5, 1),
    (64, 64, 4, 4, 25, 2.0, 2), (128, 128, 5, 2, 100, 1.0, 2), (32, 32, 3, 3, 10, 2.0, 3),
    (64, 64, 4, 2, 25, 2.0, 3)])
def test_gradient_both_matches_oracle(ctx, W, H, K, P, L, g, loss):
    print("gradient_both W=%d H=%d K=%d P=%d L=%d gamma=%g loss=%d" % (W, H, K, P, L, g, loss),
          flush=True)
    rng = np.random.default_rng(W + 7 * P + L)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (1.5, W / 6)) if W == H else \
        rng.uniform(0.2, 1, (P, W * H))
    obs = obs / obs.sum(1, keepdims=True)
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    r = solve_gpu(ctx, atoms, lam, obs, W, H, g, L, loss)
    check(r, oracle_all(atoms, lam, obs, W, H, g, L, loss))
    assert r["events"] == 0  # every scaling stayed in FP32 range


@pytest.mark.parametrize("name,d", fixtures("grad"))
def test_gradient_both_golden(ctx, name, d):
    W, H, L, g, loss = int(d["W"]), int(d["H"]), int(d["L"]), float(d["gamma"]), int(d["loss"])
    print("gradient_both golden", name, flush=True)
    K = d["lam"].size
    warm = d["b0"].size > 0
    flags = wdl.FLAG_WARM_RESTART if warm else 0
    r = solve_gpu(ctx, d["atoms"], d["lam"][None], d["obs"][None], W, H, g, L, loss, flags,
                  b0=d["b0"][None] if warm else None)
    assert rel(r["grad_lambda"][0], d["grad_lambda"]) < TOL
    assert rel(r["grad_dic"][0], d["grad_dic"]) < TOL
    assert rel(r["bary"][0], d["bary"]) < TOL
    if warm:
        assert rel(r["b_temp"][0], d["b_last"]) < TOL


@pytest.mark.parametrize("nstreams", [1, 3])
def test_plain_equals_batched_and_graph_replays(ctx, nstreams):
    # Same kernels, one solve per launch sequence vs all at once (on one or
    # several streams): identical bits.
    W = H = 32
    K, P, L, g = 3, 10, 12, 1.0
    print("plain vs batched P=%d streams=%d" % (P, nstreams), flush=True)
    rng = np.random.default_rng(3)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, 6))
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g)
    s.set_streams(nstreams)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    s.run(wdl.RUN_PLAIN)
    plain = s.results()
    outs = []
    for rep in range(3):  # graph replays, with eager work and syncs in between
        s.run(wdl.RUN_BATCHED)
        outs.append(s.results())
        s.run(wdl.RUN_PLAIN)
        s.results()
    s.close()
    for o in outs:
        for k in ("grad_lambda", "grad_dic_sum", "bary", "loss"):
            np.testing.assert_array_equal(o[k], plain[k], err_msg=k)
    ne = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, flags=wdl.FLAG_NO_GRAPH)
    ne.set_streams(nstreams)
    ne.set_observed(obs)
    ne.set_problem(atoms, lam)
    ne.run()
    eager = ne.results()
    ne.close()
    for k in ("grad_lambda", "grad_dic_sum", "bary", "loss"):
        np.testing.assert_array_equal(eager[k], plain[k], err_msg=k)


def test_each_histogram_same_alone_and_in_batch(ctx):
    W = H = 24
    K, P, L, g = 3, 8, 10, 1.0
    print("alone vs batch P=%d" % P, flush=True)
    rng = np.random.default_rng(5)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, 5))
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    full = solve_gpu(ctx, atoms, lam, obs, W, H, g, L, 2)
    for p in (0, 5):
        one = solve_gpu(ctx, atoms, lam[p:p + 1], obs[p:p + 1], W, H, g, L, 2)
        np.testing.assert_array_equal(one["grad_lambda"][0], full["grad_lambda"][p])
        np.testing.assert_array_equal(one["bary"][0], full["bary"][p])
        np.testing.assert_array_equal(one["grad_dic"][0], full["grad_dic"][p])


@pytest.mark.parametrize("name,d", [f for f in fixtures("eval") if not f[1]["log_domain"]])
def test_evaluate_both_golden(ctx, name, d):
    W, H, K, P, L = (int(d[k]) for k in ("W", "H", "K", "P", "L"))
    print("evaluate_both golden", name, flush=True)
    warm = d["b_storage"].size > 0
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, float(d["gamma"]), int(d["loss"]),
                           wdl.FLAG_WARM_RESTART if warm else 0)
    s.set_observed(d["obs"])
    s.set_scale(float(d["scale"]), bool(d["exp_weight"]))
    if warm:
        s.set_warm_state(d["b_storage"])
    f, g = s.evaluate_both(d["x"])
    assert abs(f - float(d["f"])) <= TOL * abs(float(d["f"]))
    assert rel(g, d["g"]) < TOL
    if warm:
        assert rel(s.get_warm_state(), d["b_temp"]) < TOL
        s.warm_commit()
        f2, _ = s.evaluate_both(d["x"])  # now seeded from b_temp
        o2, _, _ = O.evaluate_both(d["x"], d["obs"], W, H, float(d["gamma"]), K, P, L,
                                   int(d["loss"]), float(d["scale"]), bool(d["exp_weight"]),
                                   b_storage=d["b_temp"])
        assert abs(f2 - o2) <= TOL * abs(o2)
    s.close()


def test_evaluate_both_config_shape(ctx):
    # The plain small configuration (64x64, K=4, L=25, gamma=2), 4 of its 16 histograms.
    c = synthetic.make_problem("plain_small", P=4)
    n, P, K, L, g = c["n"], c["P"], c["K"], c["L"], c["gamma"]
    N = n * n
    print("evaluate_both plain_small P=%d" % P, flush=True)
    scale = (N / P) / c["scale_factor"]
    x = O.encode_variables(c["weights"], c["atoms"], scale)
    s = wdl.SinkhornSolver(ctx, n, n, K, P, L, g, c["loss"])
    s.set_observed(c["obs"])
    s.set_scale(scale, True)
    f, gr = s.evaluate_both(x)
    s.close()
    fo, go, _ = O.evaluate_both(x, c["obs"], n, n, g, K, P, L, c["loss"], scale)
    assert abs(f - fo) <= TOL * abs(fo)
    assert rel(gr[:P * K], go[:P * K]) < TOL
    assert rel(gr[P * K:], go[P * K:]) < TOL


@pytest.mark.parametrize("flags,expw", [(0, True), (0, False), (wdl.FLAG_LOG_DOMAIN, True),
                                        (wdl.FLAG_WEIGHTS_ONLY, True)])
def test_evaluate_both_device_entry(ctx, flags, expw):
    # x and g in device memory (wdl_evaluate_both_device).  The host callback
    # runs the same device path plus the copies of x and g, so the two agree
    # to the bit; both against the FP64 oracle's evaluate_both within TOL.
    from wassersteindictionarylearning_amd import _hip
    c = synthetic.make_problem("plain_small", P=3)
    n, P, K, L, g = c["n"], c["P"], c["K"], c["L"], c["gamma"]
    N = n * n
    scale = (N / P) / c["scale_factor"]
    print("evaluate_both_device flags=%d exp_weight=%d" % (flags, expw), flush=True)
    x = O.encode_variables(c["weights"], c["atoms"], scale, expw)
    s = wdl.SinkhornSolver(ctx, n, n, K, P, L, g, c["loss"], flags)
    s.set_observed(c["obs"])
    s.set_scale(scale, expw)
    f_host, g_host = s.evaluate_both(x)
    dx = _hip.DeviceArray.from_host(x, np.float64)
    dg = _hip.DeviceArray(x.shape, np.float64)
    f_dev = s.evaluate_both_device(dx.ptr, dg.ptr)
    g_dev = dg.download()
    dx.free()
    dg.free()
    s.close()
    assert f_dev == f_host
    np.testing.assert_array_equal(g_dev, g_host)
    fo, go, _ = O.evaluate_both(x, c["obs"], n, n, g, K, P, L, c["loss"], scale, expw,
                                log_domain=bool(flags & wdl.FLAG_LOG_DOMAIN))
    assert abs(f_dev - fo) <= TOL * abs(fo)
    assert rel(g_dev[:P * K], go[:P * K]) < TOL
    if flags & wdl.FLAG_WEIGHTS_ONLY:
        assert not np.any(g_dev[P * K:])  # no atom gradients in a weights-only solve
    else:
        assert rel(g_dev[P * K:], go[P * K:]) < TOL


@pytest.mark.parametrize("flags", [0, wdl.FLAG_LOG_DOMAIN])
def test_uneven_iterations_and_weights_only(ctx, flags):
    # Each histogram runs its own number of Sinkhorn iterations (and as many
    # backward steps); the weights-only mode gives the same weight gradients.
    W = H = 20
    K, P, L, g = 3, 7, 10, 1.0
    # (with very few iterations dL/dlambda is a handful of dot products with
    # heavy cancellation -- one iteration: 9e-5 in FP32, four in the log
    # domain: 1.2e-5 -- beyond what FP32 inputs resolve to 1e-5; counts here
    # start at 5)
    iters = np.array([5, 7, 10, 6, 5, 10, 8], np.int32)
    print("uneven iterations flags=%d" % flags, flush=True)
    rng = np.random.default_rng(21)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, 5))
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    out = {}
    for extra in (0, wdl.FLAG_WEIGHTS_ONLY):
        s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, 2, flags | extra | wdl.FLAG_WARM_RESTART)
        s.set_observed(obs)
        s.set_problem(atoms, lam)
        s.set_iterations(iters)
        s.run()
        out[extra] = s.results()
        out[extra]["b_temp"] = s.get_warm_state()
        if not extra:
            out[extra]["grad_dic"] = s.grad_dic()
        s.close()
    full, wo = out[0], out[wdl.FLAG_WEIGHTS_ONLY]
    np.testing.assert_array_equal(wo["grad_lambda"], full["grad_lambda"])
    np.testing.assert_array_equal(wo["bary"], full["bary"])
    assert not wo["grad_dic_sum"].any()
    solve = O.gradient_both_log if flags else O.gradient_both
    ones = np.ones((K, W * H))  # b_storage starts at 1 in both domains (:380-384)
    for p in range(P):
        o = solve(atoms, lam[p], obs[p], W, H, g, int(iters[p]), 2, b0=ones)
        assert rel(full["grad_lambda"][p], o["grad_lambda"]) < TOL, p
        assert rel(full["grad_dic"][p], o["grad_dic"]) < TOL, p
        assert rel(full["bary"][p], o["bary"]) < TOL, p
        b_last = o["b_last"]
        assert rel(full["b_temp"][p], b_last) < TOL, p


def test_autotune_keeps_results(ctx):
    # The tuner only regroups histograms over streams: results stay bit-identical.
    W = H = 32
    K, P, L, g = 3, 20, 8, 1.0
    print("autotune P=%d" % P, flush=True)
    rng = np.random.default_rng(9)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, 6))
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    s.set_streams(1)
    s.run()
    ref = s.results()
    n = s.autotune(2)
    assert n in (1, 2, 4, 8)
    s.run()
    r = s.results()
    s.close()
    for k in ("grad_lambda", "grad_dic_sum", "bary", "loss"):
        np.testing.assert_array_equal(r[k], ref[k], err_msg=k)


def test_full_batched_config_against_reference(ctx):
    # The batched configuration at full size (128x128, P=64, K=5, L=100,
    # gamma=1, quadratic loss; the benchmark's own inputs): all 64 histograms in
    # one batched run, three of them against the compiled reference's FP64
    # compute_gradient_both (oracle/_ref, GaussianSSE2DKernel), and the plain
    # form (one solve per launch sequence) equal to the batched one to the bit.
    from oracle import ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    c = synthetic.make_problem("batched_std")
    n, P, K, L, g, loss = c["n"], c["P"], c["K"], c["L"], c["gamma"], c["loss"]
    print("full batched configuration %dx%d P=%d K=%d L=%d" % (n, n, P, K, L), flush=True)
    r = solve_gpu(ctx, c["atoms"], c["weights"], c["obs"], n, n, g, L, loss)
    plain = solve_gpu(ctx, c["atoms"], c["weights"], c["obs"], n, n, g, L, loss,
                      mode=wdl.RUN_PLAIN)
    for k in ("grad_lambda", "grad_dic", "bary", "loss"):
        np.testing.assert_array_equal(plain[k], r[k], err_msg=k)
    for p in (0, 31, 63):
        o = ref.gradient_both(c["atoms"], c["weights"][p], c["obs"][p], n, n, g, L, loss,
                              variant="sse")
        for k in ("grad_lambda", "grad_dic", "bary"):
            assert rel(r[k][p], o[k]) < TOL, (p, k)
