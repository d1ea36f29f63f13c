"""FP64 CPU oracle for the Sinkhorn-barycenter hot path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker: the product (``wassersteindictionarylearning_amd`` and its HIP
library) never calls it.

Each function restates, in vectorised numpy float64, one function of the
reference (cited file:line into matthieuheitz/WassersteinDictionaryLearning,
``cpp/``).  Summation order differs from the reference's loops (matrix
products instead of running sums), so agreement with the reference is to
FP64 rounding (~1e-14 relative), not bit-exact.

How each function is pinned (tests/test_oracle.py):

* ``kernel_table``, ``convolve``: the compiled reference (oracle/_ref),
  golden fixtures, a dense brute-force 4-D operator at tiny sizes, the delta
  response closed form, and the adjoint identity.
* ``log_convolve``: compiled reference, fixtures, and log(convolve(exp x))
  where exp does not under/overflow.
* ``log_convolve_signarray`` and the sign helpers: compiled reference,
  fixtures, and the signed closed form log|K(s*exp x)|.
* ``gradient_both`` / ``gradient_both_log``: compiled reference, fixtures,
  the K=1 closed form of the forward, central finite differences of the
  L-step map, and agreement between the two domains.
* ``evaluate_both``: compiled reference, fixtures (incl. warm restart
  state), and finite differences of f(x).
* ``loss_*``: textbook formulas checked against the compiled reference
  through gradient_both with each loss.
* ``project_weights`` (the serving loop's projection; no counterpart in the
  reference): built from the pinned ``gradient_both`` / ``gradient_both_log``;
  its update is pinned by step 0 (weights unchanged), K = 1 (weight 1), the
  descent property (the loss falls over rounds; a sign error raises it) and
  the simplex (weights stay positive and sum to 1).
"""
from __future__ import annotations

import numpy as np

EPSILON = 1e-250  # kernels.h:30

LOSS_TV, LOSS_QUADRATIC, LOSS_WASSERSTEIN, LOSS_KL = 1, 2, 3, 4  # main_dictionary_learning.cpp:189-197


# ----------------------------------------------------------------------------
# a1 / a2: separable Gaussian convolution (standard domain)
# ----------------------------------------------------------------------------
def kernel_table(W: int, H: int, gamma: float) -> np.ndarray:
    """kernel1d[d] = max(1e-250, exp(-d^2/gamma)), d < max(W,H).  kernels.h:632-641."""
    d = np.arange(max(W, H), dtype=np.float64)
    return np.maximum(EPSILON, np.exp(-(d * d) / gamma))


def toeplitz(n: int, table: np.ndarray) -> np.ndarray:
    """Dense n x n factor M[j,k] = table[|j-k|] (the 1-D pass of kernels.h:676-696)."""
    idx = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :])
    return table[idx]


def convolve(u: np.ndarray, W: int, H: int, gamma: float) -> np.ndarray:
    """Gaussian2DKernel::convolve, kernels.h:669-699.

    u holds nvectors row-major H x W images (x fastest).  Pass x along rows,
    then pass y along columns; the operator is symmetric so the adjoint
    (kernels.h:700-702) is the same call.
    """
    t = kernel_table(W, H, gamma)
    KW, KH = toeplitz(W, t), toeplitz(H, t)
    U = np.asarray(u, np.float64).reshape(-1, H, W)
    R = KH @ (U @ KW)          # KW symmetric: U @ KW == sum_k u[i,k] kd[|j-k|]
    return R.reshape(np.shape(u))


convolve_adjoint = convolve


def convolve1d(u: np.ndarray, W: int, gamma: float) -> np.ndarray:
    """Gaussian1DKernel::convolve, kernels.h:533-549 (table :508-510):
    out[i*W + j] = sum_k kernel1d[|j-k|] u[i*W + k] for each of the nvectors
    vectors; the adjoint (:550-552) is the same call."""
    t = kernel_table(W, 1, gamma)
    U = np.asarray(u, np.float64).reshape(-1, W)
    return (U @ toeplitz(W, t)).reshape(np.shape(u))


def convolve3d(u: np.ndarray, W: int, H: int, D: int, gamma: float) -> np.ndarray:
    """Gaussian3DKernel::convolve, kernels.h:762-807 (table :719-723 over
    max(W,H,D)): voxel (x, y, z) at z*W*H + y*W + x; pass x along rows, pass y
    along columns of every z-slice (:768-789), then pass z (:791-803)."""
    n = max(W, H, D)
    d = np.arange(n, dtype=np.float64)
    t = np.maximum(EPSILON, np.exp(-(d * d) / gamma))
    KW, KH, KD = toeplitz(W, t), toeplitz(H, t), toeplitz(D, t)
    U = np.asarray(u, np.float64).reshape(-1, D, H, W)
    R = np.einsum("zk,nkhw->nzhw", KD, KH @ (U @ KW))
    return R.reshape(np.shape(u))


# ----------------------------------------------------------------------------
# a3 / a4 / a5: log-domain convolution and the sign array
# ----------------------------------------------------------------------------
def _lse_pass(x: np.ndarray, gamma: float, sign: np.ndarray | None = None):
    """One exact 1-D log-sum-exp pass along the last axis.

    out_j = m_j + log sum_k s_k exp(-(j-k)^2/gamma + x_k - m_j),
    m_j = max(-1e99, max_k(-(j-k)^2/gamma + x_k))      kernels.h:1082-1091, 1125-1135
    Returns (out, sign_out) where sign_out is True for conv >= 0 (signArray.h:11).
    """
    n = x.shape[-1]
    d = np.arange(n, dtype=np.float64)
    C = -((d[:, None] - d[None, :]) ** 2) / gamma          # [j, k]
    out = np.empty_like(x)
    sgn_out = np.empty(x.shape, dtype=bool)
    xs = x.reshape(-1, n)
    ss = None if sign is None else sign.reshape(-1, n)
    o = out.reshape(-1, n)
    so = sgn_out.reshape(-1, n)
    step = max(1, (1 << 22) // (n * n))
    for r0 in range(0, xs.shape[0], step):
        e = C[None, :, :] + xs[r0:r0 + step, None, :]      # [row, j, k]
        m = np.maximum(-1e99, e.max(axis=2))
        with np.errstate(invalid="ignore", over="ignore"):
            w = np.exp(e - m[:, :, None])
        if ss is not None:
            w = np.where(ss[r0:r0 + step, None, :], w, -w)
        conv = w.sum(axis=2)
        with np.errstate(divide="ignore"):
            o[r0:r0 + step] = np.log(np.abs(conv)) + m
        so[r0:r0 + step] = conv >= 0
    return out, sgn_out


def log_convolve(log_u: np.ndarray, W: int, H: int, gamma: float) -> np.ndarray:
    """LogSignArrayGaussian2DKernel::log_convolve, kernels.h:1073-1112 (no epsilon floor)."""
    X = np.asarray(log_u, np.float64).reshape(-1, H, W)
    T, _ = _lse_pass(X, gamma)                                   # along x (rows)
    R, _ = _lse_pass(np.ascontiguousarray(T.transpose(0, 2, 1)), gamma)   # along y
    return R.transpose(0, 2, 1).reshape(np.shape(log_u))


log_convolve_adjoint = log_convolve


def unpack_signs(signs: np.ndarray, n: int) -> np.ndarray:
    """Bit i of the array is byte i/8, bit i%8 (LSB first); 1 means >= 0.  signArray.h:8-22."""
    return np.unpackbits(np.asarray(signs, np.uint8), bitorder="little")[:n].astype(bool)


def pack_signs(pos: np.ndarray, nbytes: int | None = None, into: np.ndarray | None = None) -> np.ndarray:
    """Inverse of unpack_signs; bits past len(pos) are kept from ``into`` (setSign touches only bit i)."""
    pos = np.asarray(pos, bool).ravel()
    nb = (pos.size + 7) // 8 if nbytes is None else nbytes
    out = np.zeros(nb, np.uint8) if into is None else np.array(into, np.uint8, copy=True)
    bits = np.unpackbits(out, bitorder="little")
    bits[:pos.size] = pos
    return np.packbits(bits, bitorder="little")[:nb]


def abs_log(v: np.ndarray):
    """myAbsLog, signArray.h:24-29: returns (log|v|, v >= 0)."""
    v = np.asarray(v, np.float64)
    with np.errstate(divide="ignore"):
        return np.log(np.abs(v)), v >= 0


def abs_exp(log_v: np.ndarray, pos: np.ndarray):
    """myAbsExp, signArray.h:31-35: sign * exp(log_v)."""
    return np.where(pos, 1.0, -1.0) * np.exp(log_v)


def log_convolve_signarray(log_u: np.ndarray, signs: np.ndarray, W: int, H: int, gamma: float):
    """LogSignArrayGaussian2DKernel::log_convolve_signArray, kernels.h:1117-1159.

    Inputs: log|u| (nvec*H*W) and the packed sign array (bit v*N+idx).  Returns
    (log|K u|, new packed sign array) -- the reference overwrites the same
    sign array with the output signs (kernels.h:1153).
    """
    X = np.asarray(log_u, np.float64).reshape(-1, H, W)
    nvec = X.shape[0]
    pos = unpack_signs(signs, nvec * H * W).reshape(nvec, H, W)
    T, tpos = _lse_pass(X, gamma, pos)
    R, rpos = _lse_pass(np.ascontiguousarray(T.transpose(0, 2, 1)), gamma,
                        np.ascontiguousarray(tpos.transpose(0, 2, 1)))
    out = R.transpose(0, 2, 1).reshape(np.shape(log_u))
    new_signs = pack_signs(rpos.transpose(0, 2, 1).ravel(), len(signs), into=signs)
    return out, new_signs


# ----------------------------------------------------------------------------
# a10: losses (loss.h)
# ----------------------------------------------------------------------------
def safe_log(x):
    """loss.h:3-5."""
    return np.log(np.maximum(EPSILON, x))


def wasserstein_scalings(q, o, W, H, gamma, n_iter=50):
    """WassersteinLoss's Bregman projections, loss.h:103-117 / :132-146: a, b from
    1, n_iter times c = K b, a = q / c, c = K a, b = o / c."""
    a = np.ones_like(q)
    b = np.ones_like(q)
    for _ in range(n_iter):
        a = q / convolve(b, W, H, gamma)
        b = o / convolve(a, W, H, gamma)
    return a, b


def loss_value(kind: int, q: np.ndarray, o: np.ndarray, W=None, H=None, gamma=None) -> float:
    if kind == LOSS_WASSERSTEIN:         # loss.h:100-128
        a, b = wasserstein_scalings(q, o, W, H, gamma)
        c = convolve(b, W, H, gamma)
        return float(gamma * np.sum(safe_log(a) * q + safe_log(b) * o - a * c))
    if kind == LOSS_QUADRATIC:           # loss.h:25-32
        return float(np.sum((q - o) ** 2))
    if kind == LOSS_TV:                  # loss.h:49-56
        return float(0.5 * np.sum(np.abs(q - o)))
    if kind == LOSS_KL:                  # loss.h:74-81
        return float(np.sum(q * safe_log(q / o) - q + o))
    raise ValueError("loss kind %d not in the oracle" % kind)


def loss_gradient(kind: int, q: np.ndarray, o: np.ndarray, W=None, H=None, gamma=None) -> np.ndarray:
    if kind == LOSS_WASSERSTEIN:         # loss.h:130-152
        a, _ = wasserstein_scalings(q, o, W, H, gamma)
        return gamma * safe_log(a)
    if kind == LOSS_QUADRATIC:           # loss.h:34-40
        return 2.0 * (q - o)
    if kind == LOSS_TV:                  # loss.h:58-64
        return np.where(q - o > 0, 0.5, -0.5)
    if kind == LOSS_KL:                  # loss.h:83-89
        return safe_log(q / o)
    raise ValueError("loss kind %d not in the oracle" % kind)


# ----------------------------------------------------------------------------
# a11 / a6-a8: normalise + forward Sinkhorn + unrolled backward
# ----------------------------------------------------------------------------
def normalize_values(atoms: np.ndarray, lam: np.ndarray):
    """Problem::normalize_values, inverseWasserstein.h:53-74."""
    lam = lam / lam.sum()
    atoms = atoms / atoms.sum(axis=1, keepdims=True)
    return atoms, lam


def gradient_both(atoms, lam, obs, W, H, gamma, L, loss=LOSS_QUADRATIC, b0=None, history=False):
    """compute_gradient_both, inverseWasserstein.h:544-674.

    atoms (K,N), lam (K,), obs (N,).  Returns dict with grad_lambda (K,),
    grad_dic (K,N), bary (N,), b_last (K,N) and, if history, the b / conv_b /
    phi histories.
    """
    atoms = np.asarray(atoms, np.float64)
    lam = np.asarray(lam, np.float64)
    K, N = atoms.shape
    atoms, lam = normalize_values(atoms, lam)
    conv = lambda u: convolve(u, W, H, gamma)
    b = [np.ones((K, N)) if b0 is None else np.asarray(b0, np.float64).reshape(K, N).copy()]
    cb, phi = [], [None]
    q = None
    for s in range(1, L + 1):                      # :565-593
        c = conv(b[s - 1]); cb.append(c)           # :568
        a = atoms / c                              # :569-574
        f = conv(a); phi.append(f)                 # :575
        q = np.exp((lam[:, None] * safe_log(f)).sum(axis=0))   # :577-585
        b.append(q[None, :] / f)                   # :587-592
    g = loss_gradient(loss, q, obs, W, H, gamma)                # :597
    # atom gradient :600-629
    n = g.copy()
    v = np.zeros((K, N))
    GD = np.zeros((K, N))
    for s in range(L, 0, -1):
        t = (lam[:, None] * n[None, :] - v) * b[s]
        c = conv(t)
        GD += c / cb[s - 1]
        t = -atoms * c / cb[s - 1] ** 2
        if s == 1:
            break
        v = conv(t) / phi[s - 1]
        n = v.sum(axis=0)
    # weight gradient :635-673
    g = g * q
    GL = np.zeros(K)
    r = np.zeros((K, N))
    for s in range(L, 0, -1):
        GL += (np.log(phi[s]) * g[None, :]).sum(axis=1)     # plain log, :649
        u = (lam[:, None] * g[None, :] - r) / phi[s]
        if s != 1:
            cu = conv(u) * (-atoms / cb[s - 1] ** 2)
            r = conv(cu) * b[s - 1]
            g = r.sum(axis=0)
    out = dict(grad_lambda=GL, grad_dic=GD, bary=q, b_last=b[L])
    if history:
        out.update(b=np.stack(b), conv_b=np.stack(cb), phi=np.stack(phi[1:]))
    return out


def gradient_both_log(atoms, lam, obs, W, H, gamma, L, loss=LOSS_QUADRATIC, b0=None):
    """compute_gradient_both_log_signArray, inverseWasserstein.h:678-814.

    Same outputs as gradient_both; b is carried as log b (b0 defaults to 0).
    """
    atoms = np.asarray(atoms, np.float64)
    lam = np.asarray(lam, np.float64)
    K, N = atoms.shape
    atoms, lam = normalize_values(atoms, lam)
    lconv = lambda u: log_convolve(u, W, H, gamma)
    sconv = lambda u, p: log_convolve_signarray(u, pack_signs(p), W, H, gamma)
    log_atoms = np.log(atoms)
    b = [np.zeros((K, N)) if b0 is None else np.asarray(b0, np.float64).reshape(K, N).copy()]
    cb, phi = [], [None]
    lq = None
    for s in range(1, L + 1):                      # :696-725
        c = lconv(b[s - 1]); cb.append(c)
        a = log_atoms - c
        f = lconv(a); phi.append(f)
        lq = (lam[:, None] * f).sum(axis=0)
        b.append(lq[None, :] - f)
    q = np.exp(lq)                                 # :727-728
    g = loss_gradient(loss, q, obs, W, H, gamma)
    # atom gradient :736-770 ; v holds -v of the standard form
    n = g.copy()
    v = np.zeros((K, N))
    GD = np.zeros((K, N))
    for s in range(L, 0, -1):
        lt, pos = abs_log(lam[:, None] * n[None, :] + v)
        lt = lt + b[s]
        lc, bits = sconv(lt, pos)
        pos = unpack_signs(bits, K * N).reshape(K, N)
        GD += abs_exp(lc - cb[s - 1], pos)
        lt = log_atoms + lc - 2.0 * cb[s - 1]
        if s == 1:
            break
        lv, bits = sconv(lt, pos)
        pos = unpack_signs(bits, K * N).reshape(K, N)
        v = abs_exp(lv - phi[s - 1], pos)
        n = -v.sum(axis=0)
    # weight gradient :774-811 ; r holds -r of the standard form
    g = g * q
    GL = np.zeros(K)
    r = np.zeros((K, N))
    for s in range(L, 0, -1):
        GL += (phi[s] * g[None, :]).sum(axis=1)
        lu, pos = abs_log(lam[:, None] * g[None, :] + r)
        lu = lu - phi[s]
        if s != 1:
            lc, bits = sconv(lu, pos)
            pos = unpack_signs(bits, K * N).reshape(K, N)
            lc = lc + log_atoms - 2.0 * cb[s - 1]
            lr, bits = sconv(lc, pos)
            pos = unpack_signs(bits, K * N).reshape(K, N)
            r = abs_exp(lr + b[s - 1], pos)
            g = -r.sum(axis=0)
    return dict(grad_lambda=GL, grad_dic=GD, bary=q, b_last=b[L])


# ----------------------------------------------------------------------------
# a12 / a13: the L-BFGS callback and warm-restart state
# ----------------------------------------------------------------------------
def decode_variables(x, K, P, N, scale_dic, exp_weight=True):
    """Variable transform of evaluate_both, inverseWasserstein.h:998-1031."""
    x = np.asarray(x, np.float64)
    w = np.exp(x[:P * K] / scale_dic) if exp_weight else x[:P * K] / scale_dic
    w = w.reshape(P, K)
    w = w / w.sum(axis=1, keepdims=True)
    a = np.exp(x[P * K:]).reshape(K, N)
    a = a / a.sum(axis=1, keepdims=True)
    return w, a


def evaluate_both(x, observed, W, H, gamma, K, P, L, loss=LOSS_QUADRATIC, scale_dic=1.0,
                  exp_weight=True, b_storage=None, log_domain=False):
    """WassersteinRegression::evaluate_both, inverseWasserstein.h:986-1145.

    Returns (f, g, b_temp).  With b_storage (P,K,N) the warm-restart path is
    taken: each solve starts from b_storage[p] (:1057-1061) and b^L is saved
    into b_temp[p] (:1081-1085).
    """
    N = W * H
    observed = np.asarray(observed, np.float64).reshape(P, N)
    w, a = decode_variables(x, K, P, N, scale_dic, exp_weight)
    solve = gradient_both_log if log_domain else gradient_both
    g = np.zeros(P * K + K * N)
    f = 0.0
    b_temp = None if b_storage is None else np.zeros((P, K, N))
    for p in range(P):                                        # :1052-1113
        b0 = None if b_storage is None else np.asarray(b_storage).reshape(P, K, N)[p]
        res = solve(a, w[p], observed[p], W, H, gamma, L, loss, b0=b0)
        if b_temp is not None:
            b_temp[p] = res["b_last"]
        f += loss_value(loss, res["bary"], observed[p], W, H, gamma)  # :1089-1091
        GL = res["grad_lambda"]
        g[p * K:(p + 1) * K] = GL - np.dot(w[p], GL)          # :1095-1099
        GD = res["grad_dic"]
        GD = GD - (a * GD).sum(axis=1, keepdims=True)         # :1103-1108
        g[P * K:] += GD.ravel()                               # :1110-1112
    v2 = np.concatenate([w.ravel(), a.ravel()])
    if exp_weight:                                            # :1125-1133
        g *= v2
    else:
        g[P * K:] *= v2[P * K:]
    g[:P * K] /= scale_dic                                    # :1135-1136
    return f, g, b_temp


def encode_variables(weights, atoms, scale_dic, exp_weight=True):
    """Inverse transform done by regress_both before L-BFGS, inverseWasserstein.h:389-401."""
    w = np.asarray(weights, np.float64).ravel()
    xw = np.log(w) * scale_dic if exp_weight else w * scale_dic
    return np.concatenate([xw, np.log(np.asarray(atoms, np.float64).ravel())])


# ----------------------------------------------------------------------------
# The serving loop's projection (no counterpart in the reference)
# ----------------------------------------------------------------------------
def project_weights(atoms, obs, W, H, gamma, L, rounds, step, loss=LOSS_QUADRATIC,
                    log_domain=False):
    """Fit the weights of one histogram against a fixed dictionary, as the
    projection rounds of the served loop do: `rounds` weight-gradient
    evaluations (compute_gradient_both, inverseWasserstein.h:544-674, or its log
    form), each warm-started from the previous b^L (:1057-1061, :1081-1085; the
    store starts at 1, :380-384), each followed by the exponentiated-gradient step
        lambda_i <- lambda_i exp(-step g_i / max_j |g_j|) / Z,
        g = resultLa - <lambda, resultLa>            (the projection of :1095-1099).
    Returns (weights after the last step, loss of the last evaluation)."""
    atoms = np.asarray(atoms, np.float64)
    K, N = atoms.shape
    lam = np.full(K, 1.0 / K)
    b = np.ones((K, N))
    solve = gradient_both_log if log_domain else gradient_both
    f = None
    for _ in range(rounds):
        r = solve(atoms, lam, obs, W, H, gamma, L, loss, b0=b)
        b = r["b_last"]
        f = loss_value(loss, r["bary"], obs, W, H, gamma)
        G = r["grad_lambda"]
        g = G - lam @ G
        m = np.abs(g).max()
        if m > 0 and np.isfinite(m):
            lam = lam * np.exp(-step * g / m)
            lam = lam / lam.sum()
    return lam, f
