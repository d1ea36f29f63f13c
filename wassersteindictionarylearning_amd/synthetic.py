"""Synthetic inputs for the benchmark workloads (histograms, atoms, weights).

Blob histograms: each is a sum of isotropic Gaussian blobs (centres uniform
in the grid, sigma and amplitude uniform in stated ranges) plus a floor,
normalised to sum 1, on a pixel-unit grid.  Curve histograms: polylines of
random segments rasterised with a Gaussian profile.  Atoms and weights are
uniform(0,1), normalised -- the initialisation regress_both receives
(main_dictionary_learning.cpp:142-159 with random atoms).  All draws come
from one numpy PCG64 stream per seed: histograms, then atoms, then weights.
"""
from __future__ import annotations

import numpy as np


def blob_histograms(rng, n, count, nblobs=(1, 3), sigma=(3.0, 8.0), floor=1e-6):
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    out = np.empty((count, n * n))
    for c in range(count):
        h = np.zeros((n, n))
        for _ in range(int(rng.integers(nblobs[0], nblobs[1] + 1))):
            cx, cy = rng.uniform(0, n, 2)
            s = rng.uniform(*sigma)
            a = rng.uniform(0.5, 1.0)
            h += a * np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * s * s))
        h += floor
        out[c] = (h / h.sum()).ravel()
    return out


def curve_histograms(rng, n, count, segments=(3, 6), width=1.5, floor=1e-8):
    y, x = np.mgrid[0:n, 0:n].astype(np.float64)
    out = np.empty((count, n * n))
    for c in range(count):
        nseg = int(rng.integers(segments[0], segments[1] + 1))
        pts = rng.uniform(0, n, (nseg + 1, 2))
        h = np.zeros((n, n))
        for a, b in zip(pts[:-1], pts[1:]):
            d = b - a
            t = np.clip(((x - a[0]) * d[0] + (y - a[1]) * d[1]) / max(float(d @ d), 1e-12), 0, 1)
            dist2 = (x - a[0] - t * d[0]) ** 2 + (y - a[1] - t * d[1]) ** 2
            h = np.maximum(h, np.exp(-dist2 / (2 * width * width)))
        h += floor
        out[c] = (h / h.sum()).ravel()
    return out


def init_atoms_weights(rng, K, P, N):
    atoms = rng.uniform(0, 1, (K, N))
    atoms /= atoms.sum(axis=1, keepdims=True)
    w = rng.uniform(0, 1, (P, K))
    w /= w.sum(axis=1, keepdims=True)
    return atoms, w


# Workloads (sizes of the benchmark configurations).
WORKLOADS = {
    "plain_small": dict(n=64, P=16, K=4, gamma=2.0, L=25, loss=2, nblobs=(1, 3), sigma=(3, 8),
                        scale_factor=100.0),
    "batched_std": dict(n=128, P=64, K=5, gamma=1.0, L=100, loss=2, nblobs=(1, 4),
                        sigma=(4, 16), scale_factor=100.0),
    "log_small_eps": dict(n=256, P=32, K=6, gamma=0.5, L=200, loss=4, curves=True,
                          scale_factor=100.0),
    "projection": dict(n=128, P=4096, K=8, gamma=1.0, L=100, Lmin=20, loss=2, nblobs=(1, 5),
                       sigma=(3, 20), scale_factor=100.0),
}


def served_workload(seed=11, rate=2.0, duration_s=20.0, sizes=(64, 128, 256, 512),
                    probs=(0.4, 0.3, 0.2, 0.1), max_iterations=500):
    """The served projection loop's inputs: one random-blob dictionary per grid
    size (K uniform 4..10, gamma from {0.5, 1, 2}, blobs 1-5 with sigma 3-20),
    then Poisson arrivals at `rate` per second for `duration_s` seconds, each a
    blob histogram with a grid size drawn with `probs` and a Sinkhorn iteration
    count uniform in 20..max_iterations.  Returns (dictionaries {n: (K, gamma,
    atoms)}, arrivals [(t, n, L, histogram)], the generator for later draws)."""
    rng = np.random.default_rng(seed)
    dicts = {}
    for n in sizes:
        K = int(rng.integers(4, 11))
        g = float(rng.choice([0.5, 1.0, 2.0]))
        dicts[n] = (K, g, blob_histograms(rng, n, K, (1, 5), (3, 20)))
    arrivals, t = [], 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration_s:
            break
        n = int(rng.choice(sizes, p=list(probs)))
        L = int(rng.integers(20, max_iterations + 1))
        arrivals.append((t, n, L, blob_histograms(rng, n, 1, (1, 5), (3, 20))[0]))
    return dicts, arrivals, rng


def make_problem(name: str, seed: int | None = None, P: int | None = None):
    """Inputs of one workload.  The projection workload's atoms stand for a
    learned dictionary (random blobs, as the configuration states)."""
    c = dict(WORKLOADS[name])
    if P is not None:
        c["P"] = P
    rng = np.random.default_rng(20260000 + list(WORKLOADS).index(name) if seed is None else seed)
    n, Pn, K = c["n"], c["P"], c["K"]
    if c.get("curves"):
        obs = curve_histograms(rng, n, Pn)
    else:
        obs = blob_histograms(rng, n, Pn, c["nblobs"], c["sigma"])
    atoms, w = init_atoms_weights(rng, K, Pn, n * n)
    c.update(obs=obs, atoms=atoms, weights=w)
    if "Lmin" in c:  # per-request iteration counts, uniform in Lmin..L
        c["iterations"] = rng.integers(c["Lmin"], c["L"] + 1, Pn).astype(np.int32)
    return c
