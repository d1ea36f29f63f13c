"""The served projection loop: per-histogram slots of a batched solver
(wdl_solver_set_slot / projection_round) and the continuous-batching server
(wdl_server_*), against the oracle's projection (project_weights) and against
themselves (each request alone vs in a mix).

Tolerance: the loss of the last round relative <= 1e-5 against the FP64
oracle, the weights after the rounds <= 1e-5 (slots) and <= 5e-5 (server
requests, whose steps amplify the gradient error: see the test).  A request's result does not
depend on the other requests in its pool (every slot's arithmetic is its own),
so alone and in the mix it is the same to the bit."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import rel
from wassersteindictionarylearning_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-5
POOL_FLAGS = wdl.FLAG_WEIGHTS_ONLY | wdl.FLAG_WARM_RESTART


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def dictionary(W, H, K, seed, blobs=False):
    # uniform random atoms keep the standard-domain scalings in FP32 range;
    # sharp blob atoms (blobs=True) need the log domain
    rng = np.random.default_rng(seed)
    if blobs:
        return synthetic.blob_histograms(rng, W, K, (1, 2), (1.5, W / 5))
    a = rng.uniform(0.1, 1, (K, W * H))
    return a / a.sum(1, keepdims=True)


def observed(W, H, count, seed):
    rng = np.random.default_rng(seed)
    if W == H:
        return synthetic.blob_histograms(rng, W, count, (1, 3), (1.5, W / 5))
    o = rng.uniform(0.1, 1, (count, W * H))
    return o / o.sum(1, keepdims=True)


@pytest.mark.parametrize("logd", [0, 1])
def test_projection_rounds_match_oracle(ctx, logd):
    W = H = 24
    K, P, L, g, rounds, step = 3, 4, 12, 1.0, 3, 0.4
    iters = [12, 8, 10]
    print("projection rounds log=%d" % logd, flush=True)
    atoms = dictionary(W, H, K, 1)
    obs = observed(W, H, 3, 2)
    f = POOL_FLAGS | (wdl.FLAG_LOG_DOMAIN if logd else 0)
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, 2, f)
    s.set_problem(atoms, np.full((P, K), 1.0 / K))
    s.set_iterations(np.zeros(P, np.int32))
    for p in range(3):
        s.set_slot(p, obs[p], np.ones(K), iters[p], step)
    for _ in range(rounds):
        s.projection_round()
    for p in range(3):
        w, loss = s.get_slot(p)
        ow, ol = O.project_weights(atoms, obs[p], W, H, g, iters[p], rounds, step, 2,
                                   log_domain=bool(logd))
        assert rel(w, ow) < TOL, (p, w, ow)
        assert abs(loss - ol) < TOL * abs(ol), (p, loss, ol)
    w3, loss3 = s.get_slot(3)  # idle slot: untouched
    np.testing.assert_allclose(w3, np.full(K, 1.0 / K), rtol=1e-7)
    assert loss3 == 0.0
    s.close()


def test_projection_round_without_steps_renormalises_only(ctx):
    """Slots whose mirror-descent step was never set (set_iterations only) step
    0: a projection round only renormalises their weights, on every solver --
    also one created where an earlier solver's steps lay in device memory."""
    W = H = 24
    K, P, L, g = 3, 4, 12, 1.0
    print("projection round without steps", flush=True)
    atoms = dictionary(W, H, K, 1)
    obs = observed(W, H, P, 2)
    lam0 = np.random.default_rng(5).uniform(0.2, 1, (P, K))
    iters = np.array([12, 8, 10, 6], np.int32)
    flags = POOL_FLAGS | wdl.FLAG_LOG_DOMAIN
    out = []
    for trial in range(3):
        s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, 2, flags)
        s.set_observed(obs)
        s.set_problem(atoms, lam0)
        if trial == 0:  # leaves steps of 0.5 in the memory the next solvers get
            for p in range(P):
                s.set_slot(p, None, None, int(iters[p]), 0.5)
        else:
            s.set_iterations(iters)
        s.projection_round()
        out.append(np.array([s.get_slot(p)[0] for p in range(P)]))
        s.close()
    assert not np.allclose(out[0], lam0 / lam0.sum(1, keepdims=True), rtol=1e-3)
    np.testing.assert_array_equal(out[1], out[2])
    np.testing.assert_allclose(out[1], lam0 / lam0.sum(1, keepdims=True), rtol=1e-6)


def serve(requests, dicts, budget, max_slots=8, max_iterations=64, mixed=False):
    """Runs the requests [(dict index, histogram, iterations, rounds, step)] through a
    fresh server; returns [(weights, loss)] and its stats.  A dictionary is
    (atoms, W, H, gamma[, flags])."""
    sv = wdl.ProjectionServer(0, budget, max_slots, mixed=mixed)
    ids = [sv.add_dictionary(d[0], d[1], d[2], d[3], 2, max_iterations,
                             d[4] if len(d) > 4 else 0) for d in dicts]
    handles = [sv.submit(ids[d], h, it, r, st) for (d, h, it, r, st) in requests]
    out = [sv.wait(hd, 120.0) for hd in handles]
    stats = sv.stats()
    sv.close()
    return [(w, loss) for (w, loss, _, _) in out], stats


def mixed_requests():
    dicts = [(dictionary(32, 32, 4, 10), 32, 32, 1.0),
             (dictionary(48, 40, 3, 11), 48, 40, 2.0),
             (dictionary(136, 136, 3, 12), 136, 136, 0.5),  # tiled kernels
             (dictionary(40, 40, 3, 14, blobs=True), 40, 40, 1.0,  # exact-shift log domain
              wdl.FLAG_LOG_DOMAIN | wdl.FLAG_EXACT_LSE)]
    rng = np.random.default_rng(13)
    reqs = []
    for k in range(16):
        d = [0, 1, 2, 0, 1, 3, 2, 1, 0, 3, 1, 2, 0, 1, 3, 0][k]
        W, H = dicts[d][1], dicts[d][2]
        h = observed(W, H, 1, 100 + k)[0]
        reqs.append((d, h, int(rng.integers(6, 41)), int(rng.integers(2, 4)), 0.4))
    return dicts, reqs


def test_server_same_alone_and_in_mix():
    dicts, reqs = mixed_requests()
    print("server: %d requests in a mix" % len(reqs), flush=True)
    mix, stats = serve(reqs, dicts, 8 << 30)
    assert stats["completed"] == len(reqs)
    for k in (0, 2, 5, 6, 13):
        print("server: request %d alone" % k, flush=True)
        alone, _ = serve([reqs[k]], dicts, 8 << 30)
        np.testing.assert_array_equal(alone[0][0], mix[k][0])
        assert alone[0][1] == mix[k][1]


def test_server_matches_oracle():
    dicts, reqs = mixed_requests()
    print("server vs oracle", flush=True)
    sel = [1, 6, 9]  # 48x40, 136x136 (tiled) and 40x40 exact-shift log-domain requests
    res, _ = serve([reqs[k] for k in sel], dicts, 8 << 30)
    for (w, loss), k in zip(res, sel):
        d, h, it, r, st = reqs[k]
        a, W, H, g = dicts[d][:4]
        ow, ol = O.project_weights(a, h, W, H, g, it, r, st, 2, log_domain=len(dicts[d]) > 4)
        # each step divides the projected gradient by its largest component, up
        # to 13x smaller than max|dL/dlambda| on these requests (FP64 oracle), so
        # the gradients' 1e-5 shows on the weights as up to step * 13 * 1e-5
        assert rel(w, ow) < 5e-5, (k, w, ow)
        assert abs(loss - ol) < TOL * abs(ol), (k, loss, ol)


def test_server_admission_waits_for_memory():
    # A budget that holds one pool of either dictionary but not both: requests
    # of the second wait until the first pool is idle and released.
    dicts = [(dictionary(32, 32, 4, 20), 32, 32, 1.0), (dictionary(40, 40, 3, 21), 40, 40, 1.0)]
    # one-slot pools (a quarter of the budget holds less than one request)
    f0 = wdl.SinkhornSolver.footprint(32, 32, 4, 1, 32, POOL_FLAGS)
    f1 = wdl.SinkhornSolver.footprint(40, 40, 3, 1, 32, POOL_FLAGS)
    budget = int(max(f0, f1) * 1.5)
    assert f0 + f1 > budget
    reqs = [(0, observed(32, 32, 1, 30 + k)[0], 20, 3, 0.4) for k in range(3)] + \
           [(1, observed(40, 40, 1, 40 + k)[0], 16, 2, 0.4) for k in range(3)]
    print("server admission, budget %d bytes" % budget, flush=True)
    res, stats = serve(reqs, dicts, budget, max_slots=2, max_iterations=32)
    assert stats["completed"] == 6
    assert stats["memory_waits"] >= 1 and stats["pools_released"] >= 1, stats
    for k in (0, 5):
        alone, _ = serve([reqs[k]], dicts, budget, max_slots=2, max_iterations=32)
        np.testing.assert_array_equal(alone[0][0], res[k][0])


def test_server_rejects_bad_requests():
    sv = wdl.ProjectionServer(0, 1 << 30, 4)
    d = sv.add_dictionary(dictionary(16, 16, 2, 5), 16, 16, 1.0, 2, 20)
    with pytest.raises(wdl.WdlError):
        sv.submit(d, np.ones(256) / 256, 21, 1)  # more iterations than the dictionary allows
    with pytest.raises(wdl.WdlError):
        sv.submit(d + 1, np.ones(256) / 256, 5, 1)  # unknown dictionary
    sv.close()


def test_server_mixed_batching_same_as_pools():
    # Mixed batching: the exact-shift pools of one iteration class (here 32x32,
    # 48x40, 136x136 and 64x64 dictionaries, K 2..4, gamma 0.5..2) run each
    # round as one group launch; a standard-domain dictionary keeps its own
    # stream.  Every request's weights and loss equal the pooled mode's to the
    # bit, and one request matches the oracle's projection.
    X = wdl.FLAG_LOG_DOMAIN | wdl.FLAG_EXACT_LSE
    dicts = [(dictionary(32, 32, 4, 50, blobs=True), 32, 32, 1.0, X),
             (dictionary(48, 40, 3, 51), 48, 40, 2.0, X),
             (dictionary(136, 136, 2, 52, blobs=True), 136, 136, 0.5, X),
             (dictionary(64, 64, 3, 53, blobs=True), 64, 64, 1.0, X),
             (dictionary(24, 24, 3, 54), 24, 24, 1.0)]
    rng = np.random.default_rng(55)
    reqs = []
    for k in range(14):
        d = [0, 1, 2, 3, 0, 4, 1, 3, 2, 0, 4, 1, 3, 0][k]
        W, H = dicts[d][1], dicts[d][2]
        h = observed(W, H, 1, 200 + k)[0]
        reqs.append((d, h, int(rng.integers(5, 60)), int(rng.integers(1, 4)), 0.4))
    print("server: %d requests, pooled and mixed batching" % len(reqs), flush=True)
    pooled, _ = serve(reqs, dicts, 8 << 30)
    mixed, stats = serve(reqs, dicts, 8 << 30, mixed=True)
    assert stats["completed"] == len(reqs)
    for k in range(len(reqs)):
        np.testing.assert_array_equal(mixed[k][0], pooled[k][0], err_msg="request %d" % k)
        assert mixed[k][1] == pooled[k][1], k
    d, h, it, r, st = reqs[1]
    a, W, H, g = dicts[d][:4]
    ow, ol = O.project_weights(a, h, W, H, g, it, r, st, 2, log_domain=True)
    assert rel(mixed[1][0], ow) < 5e-5
    assert abs(mixed[1][1] - ol) < TOL * abs(ol)
