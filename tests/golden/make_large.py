"""Writes tests/golden/large_*.npz: outputs of the reference's own
compute_gradient_both_log_signArray (inverseWasserstein.h:678-814, built from
/root/reference/cpp into oracle/_ref by oracle/Makefile) on the benchmark's
full-size workloads, where the numpy oracle would take hours:

  large_c3          the 256x256 curve workload (32 histograms, K=6, gamma=0.5,
                    KL loss) at its full L=200, histogram 0: dL/dlambda, the
                    barycenter and dL/datoms.
  large_c3_more     histograms 1-3 of the same workload at L=200: dL/dlambda and
                    the barycenter.
  large_served_<n>  one served-size request at 512x512 (L=20), 256x256 (L=300)
                    and 128x128 (L=400) against a blob dictionary, warm-started
                    from a seeded log b^0: dL/dlambda.
  large_projection  four requests of the 4096-request projection workload at
                    their own L_p (store at its initial value): dL/dlambda.
  large_served_rounds  two requests of the served loop's Poisson workload, their
                    three projection rounds driven by the reference's gradients
                    (each round warm-started from the previous b^L, then the
                    exponentiated-gradient step of the serving loop): the
                    weights after the rounds and the last round's loss.

Inputs are not stored: they are regenerated from seeds by the package's
workload generator (synthetic.py) and by served_case() below; each fixture
keeps sums of its inputs so a test can see that it regenerated the same ones.
The large outputs (barycenter, dL/datoms) are stored rounded to FP32: that
rounding (6e-8 relative) is far below the 1e-5 tolerance they are checked at.

Run here (the container that holds the reference), about 30 minutes on 8 cores:
    python tests/golden/make_large.py [name ...]
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
from wassersteindictionarylearning_amd import synthetic  # noqa: E402

# served-size requests: (grid side, atoms, gamma, Sinkhorn iterations, seed)
SERVED_CASES = {512: (512, 5, 1.0, 20, 512), 256: (256, 4, 2.0, 300, 256),
                128: (128, 8, 1.0, 400, 128)}
PROJECTION_SAMPLE = (0, 1365, 2730, 4095)   # requests of the projection workload
C3_MORE = (1, 2, 3)                          # further histograms of the 256x256 workload
SERVED_SAMPLE = (16, 23)                     # arrivals of the served workload (both 64x64, K=4)
SERVED_ROUNDS, SERVED_STEP = 3, 0.5          # as the served benchmark leg runs them


def served_case(n):
    """Inputs of one served-size request: a blob dictionary (1-5 blobs, sigma
    3-20), a blob histogram, random weights and a smooth warm-start log b^0 per
    atom (an offset plus a cosine ridge of a few nats)."""
    n, K, gamma, L, seed = SERVED_CASES[n]
    rng = np.random.default_rng(seed)
    atoms = synthetic.blob_histograms(rng, n, K, (1, 5), (3, 20))
    hist = synthetic.blob_histograms(rng, n, 1, (1, 5), (3, 20))[0]
    lam = rng.uniform(0.2, 1.0, K)
    lam /= lam.sum()
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    lb0 = np.empty((K, n * n))
    for i in range(K):
        c, a = rng.uniform(-3, 3), rng.uniform(0.5, 2.0)
        fx, fy, ph = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0), rng.uniform(0, 2 * np.pi)
        lb0[i] = (c + a * np.cos(2 * np.pi * (fx * x + fy * y) / n + ph)).ravel()
    return dict(n=n, K=K, gamma=gamma, L=L, atoms=atoms, hist=hist, lam=lam, lb0=lb0)


def projection_dictionary(c):
    """The projection workload's fixed dictionary (K blob atoms, seed 7)."""
    return synthetic.blob_histograms(np.random.default_rng(7), c["n"], c["K"], c["nblobs"],
                                     c["sigma"])


def checksums(*arrays):
    """Order-sensitive sums of the inputs (a regenerated input that differs shows)."""
    out = []
    for a in arrays:
        a = np.asarray(a, np.float64).ravel()
        out += [a.sum(), (a * np.arange(1, a.size + 1) / a.size).sum()]
    return np.array(out)


def _save(name, **kw):
    path = os.path.join(HERE, name + ".npz")
    np.savez_compressed(path, **kw)
    print("wrote %s (%d bytes)" % (path, os.path.getsize(path)), flush=True)


def make_c3(ref):
    c = synthetic.make_problem("log_small_eps")
    n, K, g, L, loss = c["n"], c["K"], c["gamma"], c["L"], c["loss"]
    o = ref.gradient_both(c["atoms"], c["weights"][0], c["obs"][0], n, n, g, L, loss, variant="log")
    _save("large_c3", L=L, grad_lambda=o["grad_lambda"], bary=o["bary"].astype(np.float32),
          grad_dic=o["grad_dic"].astype(np.float32),
          check=checksums(c["obs"][0], c["atoms"], c["weights"][0]))


def make_c3_more(ref):
    c = synthetic.make_problem("log_small_eps")
    n, K, g, L, loss = c["n"], c["K"], c["gamma"], c["L"], c["loss"]
    gl, bary, chk = [], [], []
    for p in C3_MORE:
        o = ref.gradient_both(c["atoms"], c["weights"][p], c["obs"][p], n, n, g, L, loss,
                              variant="log")
        gl.append(o["grad_lambda"])
        bary.append(o["bary"].astype(np.float32))
        chk.append(checksums(c["obs"][p], c["atoms"], c["weights"][p]))
    _save("large_c3_more", L=L, requests=np.array(C3_MORE), grad_lambda=np.array(gl),
          bary=np.array(bary), check=np.array(chk))


def make_served(ref, n):
    s = served_case(n)
    o = ref.gradient_both(s["atoms"], s["lam"], s["hist"], n, n, s["gamma"], s["L"], 2,
                          b0=s["lb0"], variant="log")
    _save("large_served_%d" % n, grad_lambda=o["grad_lambda"],
          check=checksums(s["atoms"], s["hist"], s["lam"], s["lb0"]))


def make_projection(ref):
    c = synthetic.make_problem("projection")
    atoms = projection_dictionary(c)
    n, g, loss = c["n"], c["gamma"], c["loss"]
    gl = []
    for p in PROJECTION_SAMPLE:
        o = ref.gradient_both(atoms, c["weights"][p], c["obs"][p], n, n, g, int(c["iterations"][p]),
                              loss, variant="log")
        gl.append(o["grad_lambda"])
    idx = np.array(PROJECTION_SAMPLE)
    _save("large_projection", requests=idx, iterations=c["iterations"][idx],
          grad_lambda=np.array(gl),
          check=checksums(atoms, c["obs"][idx], c["weights"][idx]))


def reference_projection(ref, atoms, hist, n, gamma, L, rounds, step):
    """The served loop's projection rounds with the reference's log-domain
    gradient: each round warm-started from the previous b^L (the store starts
    at 1, inverseWasserstein.h:380-384), then
        lambda_i <- lambda_i exp(-step g_i / max_j |g_j|) / Z,
        g = resultLa - <lambda, resultLa>     (the projection of :1095-1099).
    Returns the weights after the last step and the last round's quadratic loss
    (loss.h:25-32)."""
    K = atoms.shape[0]
    lam = np.full(K, 1.0 / K)
    b = np.ones((K, n * n))
    f = None
    for _ in range(rounds):
        o = ref.gradient_both(atoms, lam, hist, n, n, gamma, L, 2, b0=b, variant="log")
        b = o["b_last"]
        f = float(((o["bary"] - hist) ** 2).sum())
        G = o["grad_lambda"]
        gp = G - lam @ G
        m = np.abs(gp).max()
        if m > 0 and np.isfinite(m):
            lam = lam * np.exp(-step * gp / m)
            lam = lam / lam.sum()
    return lam, f


def make_served_rounds(ref):
    wl, arrivals, _ = synthetic.served_workload()
    ws, fs, chk = [], [], []
    for k in SERVED_SAMPLE:
        _, n, L, h = arrivals[k]
        K, g, atoms = wl[n]
        lam, f = reference_projection(ref, atoms, h, n, g, L, SERVED_ROUNDS, SERVED_STEP)
        ws.append(lam)
        fs.append(f)
        chk.append(checksums(atoms, h))
    _save("large_served_rounds", requests=np.array(SERVED_SAMPLE), weights=np.array(ws),
          loss=np.array(fs),
          rounds=SERVED_ROUNDS, step=SERVED_STEP, check=np.array(chk))


def main(names):
    from oracle import ref
    assert ref.available(), "build oracle/_ref first: make -C oracle"
    jobs = {"c3": lambda: make_c3(ref), "c3_more": lambda: make_c3_more(ref),
            "projection": lambda: make_projection(ref),
            "served_rounds": lambda: make_served_rounds(ref)}
    for n in SERVED_CASES:
        jobs["served_%d" % n] = (lambda n=n: make_served(ref, n))
    for name in names or list(jobs):
        t0 = time.perf_counter()
        jobs[name]()
        print("%s: %.0f s" % (name, time.perf_counter() - t0), flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
