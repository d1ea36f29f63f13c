"""Checkpointed histories (WDL_FLAG_CHECKPOINT) against full histories.

The reference keeps every iteration's b, conv_b and phi per solve
(inverseWasserstein.h:551-558).  With the flag the solver keeps a ring of
ceil(sqrt(L)) + 2 of them plus phi at every segment boundary and recomputes each
segment's forward inside the backward, with the same kernels on the same
inputs, so every output must equal the full-history run bit for bit: losses,
weight and atom gradients, barycenters, the warm-restart store, range events.
Cases cover every solver form (resident, tiled, log domain, exact shifts,
weights only, warm restart, several stream groups and the plain form) and
iteration counts on, before and after segment boundaries, 1 and 0 (idle)."""
import numpy as np
import pytest

import wassersteindictionarylearning_amd as wdl
from wassersteindictionarylearning_amd import synthetic

pytestmark = pytest.mark.gpu
LOG = wdl.FLAG_LOG_DOMAIN
EXACT = wdl.FLAG_LOG_DOMAIN | wdl.FLAG_EXACT_LSE
WO = wdl.FLAG_WEIGHTS_ONLY
WARM = wdl.FLAG_WARM_RESTART


@pytest.fixture(scope="module")
def ctx():
    c = wdl.Context(0)
    yield c
    c.close()


def solve(ctx, W, H, K, P, L, g, loss, flags, iters, obs, atoms, lam, groups, mode):
    s = wdl.SinkhornSolver(ctx, W, H, K, P, L, g, loss, flags)
    s.set_observed(obs)
    s.set_problem(atoms, lam)
    if iters is not None:
        s.set_iterations(iters)
    s.set_streams(groups)
    s.run(mode)
    out = s.results()
    if not flags & WO:
        out["grad_dic"] = s.grad_dic()
    if flags & WARM:
        out["b_temp"] = s.get_warm_state()
    out["events"] = np.array([s.log_range_events()])
    s.close()
    return out


# (W, H, K, P, L, gamma, loss, flags, uneven, groups, mode)
CASES = [
    (32, 32, 3, 20, 23, 1.0, 2, 0, True, 2, wdl.RUN_BATCHED),      # resident, 2 groups
    (32, 32, 3, 5, 23, 1.0, 2, WARM, True, 1, wdl.RUN_PLAIN),      # plain form, warm store
    (32, 32, 3, 6, 17, 1.0, 4, LOG, True, 2, wdl.RUN_BATCHED),     # resident log domain
    (136, 136, 2, 2, 11, 1.0, 2, 0, False, 1, wdl.RUN_BATCHED),    # tiled
    (136, 136, 2, 3, 10, 0.5, 4, LOG | WARM, True, 1, wdl.RUN_BATCHED),  # tiled log domain
    (24, 24, 3, 6, 14, 1.0, 2, EXACT, True, 2, wdl.RUN_BATCHED),   # exact shifts
    (24, 24, 3, 6, 14, 1.0, 2, EXACT | WO | WARM, True, 2, wdl.RUN_BATCHED),  # projection form
]


@pytest.mark.parametrize("case", CASES, ids=[str(i) for i in range(len(CASES))])
def test_checkpointed_equals_full_histories(ctx, case):
    W, H, K, P, L, g, loss, flags, uneven, groups, mode = case
    print("checkpoint case %s" % (case,), flush=True)
    rng = np.random.default_rng(W + 7 * L + flags)
    obs = synthetic.blob_histograms(rng, W, P, (1, 3), (2, W / 6))
    atoms, lam = synthetic.init_atoms_weights(rng, K, P, W * H)
    iters = None
    if uneven:
        seg = int(np.ceil(np.sqrt(L)))
        # on, before and after segment boundaries, the full length, one, idle
        pick = [L, seg, seg + 1, 2 * seg, 2 * seg - 1, 1, 0, L - 1]
        iters = np.array([pick[p % len(pick)] for p in range(P)], np.int32)
    full = solve(ctx, W, H, K, P, L, g, loss, flags, iters, obs, atoms, lam, groups, mode)
    ck = solve(ctx, W, H, K, P, L, g, loss, flags | wdl.FLAG_CHECKPOINT, iters, obs, atoms, lam,
               groups, mode)
    for k in full:
        np.testing.assert_array_equal(ck[k], full[k], err_msg=k)
    assert np.isfinite(full["grad_lambda"]).all()

