"""Solver groups (wdl_group): exact-shift log-domain solvers of uneven grid
sizes, atom counts, histogram counts, iteration counts, gamma and losses run
as one launch per stage kernel.

Each member's results after a group run must equal, bit for bit, what the
same member leaves when it runs alone (the same kernel bodies on the same
inputs; only the grid a workgroup belongs to differs), and one histogram of
each member is checked against the FP64 oracle (relative L2 <= 1e-5, the
bar of the other exact-shift solver tests)."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from oracle import oracle as O
from tests._util import rel
from wassersteindictionarylearning_amd import synthetic

pytestmark = pytest.mark.gpu
TOL = 1e-5
EXACT = wdl.FLAG_LOG_DOMAIN | wdl.FLAG_EXACT_LSE

# (W, H, K, P, L, gamma, loss, per-histogram iterations or None)
MEMBERS = [
    (24, 24, 3, 3, 10, 1.0, wdl.LOSS_KL, [10, 6, 0]),
    (40, 36, 2, 2, 8, 0.5, wdl.LOSS_QUADRATIC, None),
    (136, 136, 2, 1, 5, 1.0, wdl.LOSS_KL, None),
    (300, 20, 2, 2, 6, 2.0, wdl.LOSS_TV, [6, 4]),
]


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def problem(m, seed):
    W, H, K, P, L, g, loss, it = m
    rng = np.random.default_rng(seed)
    obs = rng.uniform(0.2, 1, (P, W * H)) if W != H else \
        synthetic.blob_histograms(rng, W, P, (1, 3), (2, max(3, W / 6)))
    obs = obs / obs.sum(1, keepdims=True)
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    return obs, atoms, lam


def make(ctx, m, flags, seed):
    W, H, K, P, L, g, loss, it = m
    obs, atoms, lam = problem(m, seed)
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, loss, flags)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    if it is not None:
        s.set_iterations(np.array(it, np.int32))
    s.set_streams(1)
    return s


def outputs(s, full=True):
    r = s.results()
    if full:
        r["grad_dic"] = s.grad_dic()
    return r


@pytest.mark.parametrize("weights_only", [False, True])
def test_group_equals_members_alone(ctx, weights_only):
    flags = EXACT | (wdl.FLAG_WEIGHTS_ONLY if weights_only else 0)
    print("group of %d members, weights only %s" % (len(MEMBERS), weights_only), flush=True)
    alone = []
    for k, m in enumerate(MEMBERS):
        s = make(ctx, m, flags, 100 + k)
        s.run()
        alone.append(outputs(s, not weights_only))
        s.close()
    ms = [make(ctx, m, flags, 100 + k) for k, m in enumerate(MEMBERS)]
    g = wdl.SolverGroup(ctx, ms)
    sh = g.shape()
    print("group shape", sh, flush=True)
    assert sh["waves_per_line"] == 2  # lines of 136 and 300 need two waves
    for rep in range(2):  # the first run captures the graph, the second replays it
        g.run()
        for k, s in enumerate(ms):
            got = outputs(s, not weights_only)
            for key, v in alone[k].items():
                if v is None or (weights_only and key in ("grad_dic_sum",)):
                    continue
                np.testing.assert_array_equal(got[key], v, err_msg="member %d %s rep %d"
                                              % (k, key, rep))
    g.close()
    # one histogram of each member against the oracle (its own iteration count)
    for k, m in enumerate(MEMBERS):
        W, H, K, P, L, gam, loss, it = m
        obs, atoms, lam = problem(m, 100 + k)
        p = 0
        lp = L if it is None else it[p]
        o = O.gradient_both_log(atoms, lam[p], obs[p], W, H, gam, lp, loss)
        e = rel(alone[k]["grad_lambda"][p], o["grad_lambda"])
        print("member %d (%dx%d, L_p=%d): grad_lambda error %.2e" % (k, W, H, lp, e), flush=True)
        assert e < TOL, k
        if not weights_only:
            assert rel(alone[k]["grad_dic"][p], o["grad_dic"]) < TOL, k
            assert rel(alone[k]["bary"][p], o["bary"]) < TOL, k
    for s in ms:
        s.close()


def test_group_projection_rounds_equal_members_alone(ctx):
    # served-style members: weights only, warm-restart store, slots set one by
    # one with their own iterations (one idle), two projection rounds
    flags = EXACT | wdl.FLAG_WEIGHTS_ONLY | wdl.FLAG_WARM_RESTART
    shapes = [(32, 32, 4, 3, 12, 1.0), (48, 40, 3, 2, 9, 0.5), (136, 136, 2, 2, 6, 2.0)]
    iters = [[12, 7, 0], [9, 5], [6, 3]]
    print("group projection rounds", flush=True)

    def build():
        ss = []
        for k, (W, H, K, P, L, g) in enumerate(shapes):
            rng = np.random.default_rng(300 + k)
            atoms = synthetic.blob_histograms(rng, W, K, (1, 2), (2, 5)) if W == H else \
                rng.uniform(0.1, 1, (K, W * H))
            s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, wdl.LOSS_KL, flags)
            s.set_problem(atoms, np.full((P, K), 1.0 / K))
            s.set_iterations(np.zeros(P, np.int32))
            s.set_streams(1)
            for p in range(P):
                h = rng.uniform(0.2, 1, W * H)
                s.set_slot(p, h / h.sum(), np.full(K, 1.0 / K), iters[k][p], 0.5)
            ss.append(s)
        return ss

    def state(ss):
        out = []
        for s in ss:
            out.append(dict(slots=[s.get_slot(p) for p in range(s.P)],
                            warm=s.get_warm_state(), gl=s.results()["grad_lambda"]))
        return out

    solo = build()
    for _ in range(2):
        for s in solo:
            s.projection_round()
    ref = state(solo)
    for s in solo:
        s.close()
    ms = build()
    g = wdl.SolverGroup(ctx, ms)
    for _ in range(2):
        g.projection_round()
    got = state(ms)
    for k in range(len(shapes)):
        np.testing.assert_array_equal(got[k]["gl"], ref[k]["gl"], err_msg="member %d" % k)
        np.testing.assert_array_equal(got[k]["warm"], ref[k]["warm"], err_msg="member %d" % k)
        for p in range(shapes[k][3]):
            np.testing.assert_array_equal(got[k]["slots"][p][0], ref[k]["slots"][p][0])
            assert got[k]["slots"][p][1] == ref[k]["slots"][p][1]
    g.close()
    for s in ms:
        s.close()


def test_group_rejects_unsupported_members(ctx):
    a = wdl.SinkhornSolver(ctx, 16, 16, 2, 1, 4, 1.0, wdl.LOSS_KL, EXACT)
    b = wdl.SinkhornSolver(ctx, 16, 16, 2, 1, 4, 1.0, wdl.LOSS_KL, wdl.FLAG_LOG_DOMAIN)
    c = wdl.SinkhornSolver(ctx, 16, 16, 2, 1, 4, 1.0, wdl.LOSS_KL, EXACT | wdl.FLAG_WEIGHTS_ONLY)
    d = wdl.SinkhornSolver(ctx, 16, 16, 2, 1, 100, 1.0, wdl.LOSS_KL,
                          EXACT | wdl.FLAG_CHECKPOINT)  # long enough for the ring to pay
    for bad in ([a, b], [a, c], [a, a], [a, d]):
        with pytest.raises(wdl.WdlError):
            wdl.SolverGroup(ctx, bad)
    g = wdl.SolverGroup(ctx, [a])
    with pytest.raises(wdl.WdlError):  # the projection round needs weights-only warm members
        g.projection_round()
    g.close()
    for s in (a, b, c, d):
        s.close()
