"""Full-size parity against the compiled reference's FP64 output.

The fixtures (tests/golden/large_*.npz) hold what the reference's own
compute_gradient_both_log_signArray (inverseWasserstein.h:678-814, built from
its sources into oracle/_ref) gives on the benchmark workloads at their real
sizes and lengths; tests/golden/make_large.py wrote them and regenerates the
inputs from seeds.  Each test first checks that it regenerated the same inputs
(sums stored with the fixture), then runs the HIP solver on them.

Tolerance: relative L2 error <= 1e-5 for dL/dlambda, barycenters and
dL/datoms (FP32 arithmetic, FP64 log shifts and histories); the weights after
the served loop's projection rounds <= 5e-5 (each exponentiated-gradient step
divides the projected gradient by its largest component, which amplifies the
gradient's error; see tests/test_gpu_serving.py)."""
import os

import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from tests._util import GOLDEN, rel
from tests.golden import make_large as ML
from wassersteindictionarylearning_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-5
EXACT = wdl.FLAG_LOG_DOMAIN | wdl.FLAG_EXACT_LSE


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def fixture(name):
    path = os.path.join(GOLDEN, name + ".npz")
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def same_inputs(z, *arrays):
    np.testing.assert_allclose(ML.checksums(*arrays), z["check"], rtol=1e-12, atol=0)


def test_log_config_full_length_against_reference(ctx):
    # 256x256 curve histograms, K=6, gamma=0.5, KL loss, tiled log-domain
    # kernels, all L=200 Sinkhorn iterations: histogram 0 of the benchmark's 32
    z = fixture("large_c3")
    c = synthetic.make_problem("log_small_eps")
    same_inputs(z, c["obs"][0], c["atoms"], c["weights"][0])
    n, K, g, L, loss = c["n"], c["K"], c["gamma"], int(z["L"]), c["loss"]
    print("log configuration %dx%d K=%d L=%d against the reference" % (n, n, K, L), flush=True)
    zm = fixture("large_c3_more")  # histograms 1-3: dL/dlambda and the barycenter
    P = 1 + len(zm["requests"])
    for j, p in enumerate(zm["requests"]):
        np.testing.assert_allclose(ML.checksums(c["obs"][p], c["atoms"], c["weights"][p]),
                                   zm["check"][j], rtol=1e-12, atol=0)
    out = {}
    for name, f in (("full", wdl.FLAG_LOG_DOMAIN),
                    ("checkpointed", wdl.FLAG_LOG_DOMAIN | wdl.FLAG_CHECKPOINT)):
        s = wdl.SinkhornSolver(ctx, n, n, K, P, L, g, loss, f)
        s.set_observed(c["obs"][:P])
        s.set_problem(c["atoms"], c["weights"][:P])
        s.run()
        out[name] = s.results()
        out[name]["grad_dic"] = s.grad_dic()
        out[name]["events"] = s.log_range_events()
        s.close()
    # the checkpointed histories (17 of 200 iterations kept) give the same bits
    for k in out["full"]:
        np.testing.assert_array_equal(out["checkpointed"][k], out["full"][k], err_msg=k)
    r, gd = out["full"], out["full"]["grad_dic"]
    assert r["events"] == 0
    for k, got in (("bary", r["bary"][0]), ("grad_lambda", r["grad_lambda"][0]),
                   ("grad_dic", gd[0])):
        e = rel(got, z[k])
        print("%s: %.2e" % (k, e), flush=True)
        assert e < TOL, k
    for j, p in enumerate(zm["requests"]):
        for k in ("bary", "grad_lambda"):
            e = rel(r[k][p], zm[k][j])
            print("histogram %d %s: %.2e" % (p, k, e), flush=True)
            assert e < TOL, (p, k)


@pytest.mark.parametrize("n", [128, 256, 512])
def test_served_sizes_warm_weights_only_against_reference(ctx, n):
    # one request per served grid size, weights-only exact-shift solve seeded
    # from a warm-restart store that is not the initial one: 512x512 at L=20,
    # 256x256 at L=300, 128x128 at L=400
    z = fixture("large_served_%d" % n)
    c = ML.served_case(n)
    same_inputs(z, c["atoms"], c["hist"], c["lam"], c["lb0"])
    K, g, L = c["K"], c["gamma"], c["L"]
    print("served size %dx%d K=%d gamma=%g L=%d" % (n, n, K, g, L), flush=True)
    s = wdl.SinkhornSolver(ctx, n, n, K, 1, L, g, 2,
                           EXACT | wdl.FLAG_WEIGHTS_ONLY | wdl.FLAG_WARM_RESTART)
    s.set_observed(c["hist"][None])
    s.set_problem(c["atoms"], c["lam"][None])
    s.set_warm_state(c["lb0"][None])
    s.run()
    r = s.results(barycenters=False)
    s.close()
    e = rel(r["grad_lambda"][0], z["grad_lambda"])
    print("weight-gradient error %.2e" % e, flush=True)
    assert e < TOL


def test_projection_sample_against_reference(ctx):
    # four requests spread over the 4096 of the projection workload (128x128,
    # K=8 blob dictionary, gamma=1), each at its own L_p, in one batched solve
    z = fixture("large_projection")
    c = synthetic.make_problem("projection")
    atoms = ML.projection_dictionary(c)
    idx = z["requests"]
    same_inputs(z, atoms, c["obs"][idx], c["weights"][idx])
    n, K, g, L, loss = c["n"], c["K"], c["gamma"], c["L"], c["loss"]
    print("projection requests %s, L_p %s" % (list(idx), list(z["iterations"])), flush=True)
    s = wdl.SinkhornSolver(ctx, n, n, K, len(idx), L, g, loss,
                           EXACT | wdl.FLAG_WEIGHTS_ONLY | wdl.FLAG_WARM_RESTART)
    s.set_observed(c["obs"][idx])
    s.set_problem(atoms, c["weights"][idx])
    s.set_iterations(c["iterations"][idx])
    s.run()
    r = s.results(barycenters=False)
    s.close()
    for j, p in enumerate(idx):
        e = rel(r["grad_lambda"][j], z["grad_lambda"][j])
        print("request %d: weight-gradient error %.2e" % (p, e), flush=True)
        assert e < TOL, p


def test_served_rounds_against_reference():
    # two requests of the served benchmark's Poisson workload through the
    # server: three projection rounds each, against the same rounds driven by
    # the reference's gradients
    z = fixture("large_served_rounds")
    wl, arrivals, _ = synthetic.served_workload()
    sv = wdl.ProjectionServer(0, 16 << 30, 4)
    ids, handles = {}, []
    for j, k in enumerate(z["requests"]):
        _, n, L, h = arrivals[int(k)]
        K, g, atoms = wl[n]
        np.testing.assert_allclose(ML.checksums(atoms, h), z["check"][j], rtol=1e-12, atol=0)
        if n not in ids:
            ids[n] = sv.add_dictionary(atoms, n, n, g, wdl.LOSS_QUADRATIC, 500, EXACT)
        print("served request %d: %dx%d K=%d L=%d" % (k, n, n, K, L), flush=True)
        handles.append(sv.submit(ids[n], h, L, int(z["rounds"]), float(z["step"])))
    res = [sv.wait(hd, 300.0) for hd in handles]
    sv.close()
    for j, (w, loss, _, _) in enumerate(res):
        ew, el = rel(w, z["weights"][j]), abs(loss - z["loss"][j]) / abs(z["loss"][j])
        print("request %d: weights %.2e, loss %.2e" % (z["requests"][j], ew, el), flush=True)
        assert ew < 5e-5 and el < TOL, j
