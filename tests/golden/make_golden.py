"""Writes tests/golden/*.npz: input/output vectors of the reference's own
functions, produced by running the reference templates compiled from
/root/reference/cpp (oracle/Makefile -> oracle/_ref/*.so).

Run here (the container that holds the reference): python tests/golden/make_golden.py
Fixtures are data only (inputs and the reference's outputs), kept small.
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
from oracle import ref  # noqa: E402


def blob_hist(rng, W, H, nb=(1, 3), sig=(1.5, 3.0), floor=1e-6):
    y, x = np.mgrid[0:H, 0:W].astype(float)
    h = np.zeros((H, W))
    for _ in range(rng.integers(nb[0], nb[1] + 1)):
        cx, cy = rng.uniform(0, W), rng.uniform(0, H)
        s = rng.uniform(*sig)
        h += rng.uniform(0.5, 1.0) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
    h += floor
    return (h / h.sum()).ravel()


def main():
    assert ref.available(), "build oracle/_ref first: make -C oracle"
    rng = np.random.default_rng(20260001)
    out = {}

    # a1: factor tables
    for (W, H, g) in [(12, 9, 2.0), (64, 64, 0.5), (7, 30, 25.0)]:
        out["table_%d_%d_%g" % (W, H, g)] = dict(W=W, H=H, gamma=g, out=ref.kernel_table(W, H, g))

    # a2: convolve (plain and AVX kernels)
    for j, (W, H, g, nv) in enumerate([(12, 9, 2.0, 3), (16, 16, 25.0, 2), (20, 7, 0.5, 1),
                                       (33, 17, 1.0, 2)]):
        u = rng.uniform(0, 1, nv * W * H)
        d = dict(W=W, H=H, gamma=g, u=u, out=ref.convolve(u, W, H, g))
        if W % 4 == 0:  # the AVX kernel needs 32-byte aligned rows
            d["out_sse"] = ref.sse_convolve(u, W, H, g)
        out["conv_%d" % j] = d

    # a3: log_convolve
    for j, (W, H, g, nv, lo, hi) in enumerate([(10, 7, 0.5, 2, -30, 30), (9, 12, 2.0, 1, -5, 5),
                                               (16, 16, 1.0, 2, -60, 10)]):
        u = rng.uniform(lo, hi, nv * W * H)
        if j == 1:
            u[::7] = -np.inf          # log of exact zeros
        out["lconv_%d" % j] = dict(W=W, H=H, gamma=g, u=u, out=ref.log_convolve(u, W, H, g))

    # a4: log_convolve_signArray (sign bits in and out)
    for j, (W, H, g, nv) in enumerate([(10, 7, 0.5, 2), (12, 12, 2.0, 3), (5, 3, 1.0, 1)]):
        u = rng.uniform(-20, 20, nv * W * H)
        if j == 1:
            u[::5] = -np.inf
        signs = rng.integers(0, 256, (nv * W * H + 7) // 8).astype(np.uint8)
        o, s2 = ref.log_convolve_signarray(u, signs, W, H, g)
        out["sconv_%d" % j] = dict(W=W, H=H, gamma=g, u=u, signs_in=signs, out=o, signs_out=s2)

    # a5: myAbsLog / myAbsExp on the values of signArray.h:67
    vals = np.array([1, -2, -4, 8, -16, 32, 64, 128, -255, -127, 23, 56, -84, 19, -26, -154.0])
    lv, s = ref.abs_log(vals, np.zeros(2, np.uint8))
    out["signs_0"] = dict(v=vals, log=lv, signs=s, back=ref.abs_exp(lv, s))

    # a6-a8, a10: compute_gradient_both (plain kernel) with each loss; warm b0
    cases = [(12, 12, 3, 6, 2.0, 2, False), (12, 12, 3, 6, 2.0, 4, False),
             (12, 12, 3, 6, 2.0, 1, False), (10, 8, 2, 4, 1.0, 2, True),
             (16, 16, 4, 10, 0.5, 2, False)]
    for j, (W, H, K, L, g, loss, warm) in enumerate(cases):
        N = W * H
        atoms = rng.uniform(0, 1, (K, N))
        lam = rng.uniform(0, 1, K)
        obs = blob_hist(rng, W, H)
        b0 = rng.uniform(0.5, 2.0, (K, N)) if warm else None
        r = ref.gradient_both(atoms, lam, obs, W, H, g, L, loss, b0=b0)
        out["grad_%d" % j] = dict(W=W, H=H, K=K, L=L, gamma=g, loss=loss, atoms=atoms, lam=lam,
                                  obs=obs, b0=(b0 if warm else np.zeros(0)), **r)

    # a9: compute_gradient_both_log_signArray
    for j, (W, H, K, L, g, loss) in enumerate([(10, 10, 3, 5, 0.5, 4), (12, 9, 2, 6, 1.0, 2)]):
        N = W * H
        atoms = rng.uniform(0, 1, (K, N))
        lam = rng.uniform(0, 1, K)
        obs = blob_hist(rng, W, H, floor=1e-8)
        r = ref.gradient_both(atoms, lam, obs, W, H, g, L, loss, variant="log")
        out["glog_%d" % j] = dict(W=W, H=H, K=K, L=L, gamma=g, loss=loss, atoms=atoms, lam=lam,
                                  obs=obs, **r)

    # a12 / a13: evaluate_both, incl. exp_weight off and warm restart; log build
    W, H, K, P, L, g = 8, 8, 2, 3, 4, 1.0
    N = W * H
    obs = np.stack([blob_hist(rng, W, H) for _ in range(P)])
    x = np.concatenate([np.log(rng.uniform(0.1, 1, P * K)) * 2.0, np.log(rng.uniform(0.1, 1, K * N))])
    for j, (ew, warm, logd, loss) in enumerate([(True, False, False, 2), (False, False, False, 2),
                                                (True, True, False, 2), (True, False, True, 4),
                                                (True, False, False, 4)]):
        xx = x.copy()
        if not ew:
            xx[:P * K] = rng.uniform(0.1, 1, P * K) * 2.0
        bs = rng.uniform(0.5, 2.0, (P, K, N)) if warm else None
        f, gr, bt = ref.evaluate_both(xx, obs, W, H, g, K, P, L, loss, 2.0, ew, b_storage=bs,
                                      log_domain=logd)
        out["eval_%d" % j] = dict(W=W, H=H, K=K, P=P, L=L, gamma=g, loss=loss, scale=2.0,
                                  exp_weight=ew, log_domain=logd, x=xx, obs=obs,
                                  b_storage=(bs if warm else np.zeros(0)), f=f, g=gr,
                                  b_temp=(bt if warm else np.zeros(0)))

    for name, d in out.items():
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
    wasserstein_fixtures()
    line_fixtures()
    print("wrote %d fixtures to %s" % (len(out), HERE))


def wasserstein_fixtures():
    # the Wasserstein loss (loss.h:92-154), with a generator of its own so the
    # fixtures above keep their values
    rw = np.random.default_rng(303)
    for j, (W, H, K, L, g, logd) in enumerate([(12, 12, 3, 6, 2.0, False),
                                               (10, 12, 2, 5, 1.0, True)]):
        N = W * H
        atoms = rw.uniform(0, 1, (K, N))
        lam = rw.uniform(0, 1, K)
        obs = blob_hist(rw, W, H)
        r = ref.gradient_both(atoms, lam, obs, W, H, g, L, 3, variant="log" if logd else "plain")
        d = dict(W=W, H=H, K=K, L=L, gamma=g, loss=3, atoms=atoms, lam=lam, obs=obs, **r)
        if not logd:
            d["b0"] = np.zeros(0)
        name = "glog_2" if logd else "grad_5"
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
    W, H, K, P, L, g = 8, 8, 2, 3, 4, 1.0
    N = W * H
    obs = np.stack([blob_hist(rw, W, H) for _ in range(P)])
    xx = np.concatenate([np.log(rw.uniform(0.1, 1, P * K)) * 2.0,
                         np.log(rw.uniform(0.1, 1, K * N))])
    f, gr, _ = ref.evaluate_both(xx, obs, W, H, g, K, P, L, 3, 2.0, True)
    np.savez_compressed(os.path.join(HERE, "eval_5.npz"),
                        W=W, H=H, K=K, P=P, L=L, gamma=g, loss=3, scale=2.0, exp_weight=True,
                        log_domain=False, x=xx, obs=obs, b_storage=np.zeros(0), f=f, g=gr,
                        b_temp=np.zeros(0))


def line_fixtures():
    # Gaussian1DKernel / Gaussian3DKernel::convolve (kernels.h:533-549,
    # :762-807), with a generator of their own
    rl = np.random.default_rng(404)
    for j, (W, g, nv) in enumerate([(37, 2.0, 3), (200, 9.0, 2), (64, 500.0, 2)]):
        u = rl.uniform(0, 1, nv * W)
        np.savez_compressed(os.path.join(HERE, "conv1d_%d.npz" % j), W=W, gamma=g, u=u,
                            out=ref.convolve1d(u, W, g))
    for j, (W, H, D, g, nv) in enumerate([(8, 6, 5, 2.0, 2), (20, 16, 12, 1.0, 1),
                                          (6, 5, 136, 4.0, 1)]):
        u = rl.uniform(0, 1, nv * W * H * D)
        np.savez_compressed(os.path.join(HERE, "conv3d_%d.npz" % j), W=W, H=H, D=D, gamma=g, u=u,
                            out=ref.convolve3d(u, W, H, D, g))


if __name__ == "__main__":
    main()
