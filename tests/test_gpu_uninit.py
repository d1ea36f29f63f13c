"""Reads of device memory no launch has written: every solver form run twice,
once on buffers as allocated and once with wdl_debug_fill on (every buffer the
solver does not initialise itself filled with NaN bytes at creation).  The
results must be the same to the bit -- a read before the first write would
carry NaN or the memory's old contents into them (the per-slot steps of the
projection round were such a read until they were zeroed at creation)."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from wassersteindictionarylearning_amd import _hip, synthetic

pytestmark = pytest.mark.gpu

LOG, EXACT, CKPT = wdl.FLAG_LOG_DOMAIN, wdl.FLAG_EXACT_LSE, wdl.FLAG_CHECKPOINT
WARM, WONLY = wdl.FLAG_WARM_RESTART, wdl.FLAG_WEIGHTS_ONLY

# name: W, H, K, P, L, gamma, loss, flags, histogram groups, per-histogram L (None: L)
FORMS = {
    "standard": (48, 48, 3, 8, 20, 1.0, wdl.LOSS_QUADRATIC, 0, 2, None),
    "standard_checkpointed": (48, 48, 3, 8, 40, 1.0, wdl.LOSS_QUADRATIC, CKPT, 2, None),
    "standard_warm": (48, 48, 3, 4, 10, 1.0, wdl.LOSS_QUADRATIC, WARM, 1, None),
    "log": (64, 64, 3, 4, 10, 1.0, wdl.LOSS_KL, LOG, 2, None),
    "tiled_log": (136, 136, 2, 2, 6, 0.5, wdl.LOSS_KL, LOG, 1, None),
    "tiled_standard": (136, 136, 2, 2, 6, 2.0, wdl.LOSS_QUADRATIC, 0, 1, None),
    "exact": (40, 36, 3, 4, 8, 1.0, wdl.LOSS_KL, LOG | EXACT, 2, [8, 4, 7, 3]),
    "exact_checkpointed": (40, 36, 3, 4, 30, 1.0, wdl.LOSS_KL, LOG | EXACT | CKPT, 1,
                           [30, 15, 29, 3]),
    "exact_weights_only": (40, 36, 3, 4, 8, 1.0, wdl.LOSS_QUADRATIC, LOG | EXACT | WONLY, 1,
                           [8, 4, 7, 3]),
}


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def problem(W, H, K, P, seed, blobs):
    """Blob histograms for the exact-shift forms (any spread of log values);
    uniform random ones, whose scalings stay in range, for the others."""
    rng = np.random.default_rng(seed)
    if blobs and W == H:
        obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, W / 6))
        atoms = synthetic.blob_histograms(rng, W, K, (1, 3), (2, W / 6))
    else:
        obs = rng.uniform(0.1, 1, (P, W * H))
        atoms = rng.uniform(0.1, 1, (K, W * H))
        obs /= obs.sum(1, keepdims=True)
        atoms /= atoms.sum(1, keepdims=True)
    return obs, atoms, rng.uniform(0.2, 1, (P, K))


def solve(ctx, form, fill):
    W, H, K, P, L, g, loss, flags, ng, lp = FORMS[form]
    obs, atoms, lam = problem(W, H, K, P, 3, bool(flags & EXACT))
    prev = wdl.debug_fill(fill)
    try:
        s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, loss, flags)
    finally:
        wdl.debug_fill(prev)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    if lp is not None:
        s.set_iterations(np.array(lp, np.int32))
    s.set_streams(ng)
    s.run()
    if flags & WARM:
        s.run()  # the second run starts from the first one's scalings
    r = s.results()
    out = {k: np.asarray(r[k]) for k in ("grad_lambda", "bary", "loss") if k in r}
    if not flags & WONLY:
        out["grad_dic"] = s.grad_dic()
    if form == "standard":  # evaluate_both on the device as well
        x = np.concatenate([np.log(lam / lam.sum(1, keepdims=True)).ravel(),
                            np.log(atoms / atoms.sum(1, keepdims=True)).ravel()])
        dx = _hip.DeviceArray.from_host(x, np.float64)
        dg = _hip.DeviceArray(x.shape, np.float64)
        out["evaluate_f"] = np.array([s.evaluate_both_device(dx.ptr, dg.ptr)])
        out["evaluate_g"] = dg.download()
        dx.free()
        dg.free()
    s.close()
    return out


@pytest.mark.parametrize("form", sorted(FORMS))
def test_results_do_not_read_unwritten_memory(ctx, form):
    print("form %s" % form, flush=True)
    a = solve(ctx, form, False)
    b = solve(ctx, form, True)
    for k in a:
        assert np.all(np.isfinite(a[k])), (form, k)
        np.testing.assert_array_equal(a[k], b[k], err_msg="%s: %s" % (form, k))


def test_projection_round_does_not_read_unwritten_memory(ctx):
    """Slots given iterations but no step (set_iterations), one projection
    round, the weights and the weight gradient."""
    W = H = 32
    K, P, L = 3, 6, 12
    obs, atoms, lam = problem(W, H, K, P, 7, True)
    outs = []
    for fill in (False, True):
        prev = wdl.debug_fill(fill)
        try:
            s = wdl.SinkhornSolver(ctx, W, H, K, P, L, 1.0, wdl.LOSS_QUADRATIC,
                                   LOG | EXACT | WONLY | WARM)
        finally:
            wdl.debug_fill(prev)
        s.set_observed(obs)
        s.set_problem(atoms, lam)
        s.set_iterations(np.array([12, 5, 0, 9, 1, 12], np.int32))
        s.projection_round()
        outs.append((np.array([s.get_slot(p)[0] for p in range(P)]),
                     np.asarray(s.results()["grad_lambda"])))
        s.close()
    assert np.all(np.isfinite(outs[0][0]))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    np.testing.assert_array_equal(outs[0][1], outs[1][1])
