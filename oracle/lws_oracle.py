"""numpy restatement of the learned-warm-start network (test infrastructure
only): features, forward, the Rayleigh-quotient loss and the backward pass,
with optional BF16 rounding at the points where the device kernels round
(layer activations, aggregates, weights, gradients fed to the matrix units).
Its backward pass is pinned by central finite differences in
tests/test_lws.py.

Parameter layout (floats): W0 [64 x 8] | b0 [64] | 4 x (W_l [64 x 128] |
b_l [64]) | w_o [64] | b_o.
"""
from __future__ import annotations

import numpy as np

H, NIN, NL = 64, 8, 4


def offsets():
    o = {"W0": 0, "b0": H * NIN}
    p = H * NIN + H
    for l in range(NL):
        o[f"W{l + 1}"] = p
        o[f"b{l + 1}"] = p + H * 2 * H
        p += H * 2 * H + H
    o["wo"] = p
    o["bo"] = p + H
    o["n"] = p + H + 1
    return o


OFF = offsets()
NPARAM = OFF["n"]


def init_params(seed=0, scale=1.0):
    rng = np.random.default_rng(seed)
    P = np.zeros(NPARAM, dtype=np.float64)
    P[OFF["W0"]:OFF["b0"]] = rng.standard_normal(H * NIN) * np.sqrt(2.0 / NIN) * scale
    for l in range(NL):
        a = OFF[f"W{l + 1}"]
        P[a:a + H * 2 * H] = rng.standard_normal(H * 2 * H) * np.sqrt(1.0 / (2 * H)) * scale
    P[OFF["wo"]:OFF["bo"]] = rng.standard_normal(H) * np.sqrt(1.0 / H) * scale
    return P


def unpack(P):
    W0 = P[OFF["W0"]:OFF["b0"]].reshape(H, NIN)
    b0 = P[OFF["b0"]:OFF["b0"] + H]
    Ws = [P[OFF[f"W{l + 1}"]:OFF[f"b{l + 1}"]].reshape(H, 2 * H) for l in range(NL)]
    bs = [P[OFF[f"b{l + 1}"]:OFF[f"b{l + 1}"] + H] for l in range(NL)]
    wo = P[OFF["wo"]:OFF["bo"]]
    bo = P[OFF["bo"]]
    return W0, b0, Ws, bs, wo, bo


def bf16(x):
    """Round to BF16 (nearest even), returned as float32."""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def features(coord, rowlen, level):
    """coord: M x 3 (segment), rowlen: neighbours per row, level: cut index."""
    lo = coord.min(0)
    ln = coord.max(0) - lo
    order = sorted(range(3), key=lambda d: (-ln[d], d))
    L = ln[order[0]] if ln[order[0]] > 0 else 1.0
    M = coord.shape[0]
    x = np.zeros((M, NIN))
    for k, d in enumerate(order):
        x[:, k] = (coord[:, d] - lo[d] - 0.5 * ln[d]) / L
    x[:, 3] = rowlen / 26.0
    x[:, 4] = level / 16.0
    x[:, 5] = 1.0
    return x


def forward(P, x, A, emulate_bf16=False):
    """A: dense M x M shared-vertex weight matrix of the segment graph."""
    r = bf16 if emulate_bf16 else (lambda v: v)
    W0, b0, Ws, bs, wo, bo = unpack(P)
    d = A.sum(1)
    dinv = np.where(d > 0, 1.0 / np.where(d > 0, d, 1), 0.0)
    h = [r(np.maximum(x @ W0.T + b0, 0.0))]
    a = []
    for l in range(NL):
        al = r((A @ h[l]) * dinv[:, None])
        a.append(al)
        z = np.concatenate([h[l], al], 1) @ r(Ws[l]).T + bs[l]
        h.append(r(np.maximum(z, 0.0)))
    y = h[NL] @ wo + bo
    return y, dict(h=h, a=a, dinv=dinv, x=x)


def laplacian(A):
    return np.diag(A.sum(1)) - A


def loss(y, A, lam):
    """R(y) / lam - 1, R the Rayleigh quotient of the graph Laplacian of A
    on the centred prediction."""
    yc = y - y.mean()
    return (yc @ laplacian(A) @ yc) / (yc @ yc) / lam - 1.0


def dloss(y, A, lam):
    yc = y - y.mean()
    Ly = laplacian(A) @ y
    Q = yc @ yc
    R = (y @ Ly) / Q
    return 2.0 * (Ly - R * yc) / (Q * lam)


def backward(P, cache, A, dy, emulate_bf16=False):
    r = bf16 if emulate_bf16 else (lambda v: v)
    W0, b0, Ws, bs, wo, bo = unpack(P)
    h, a, dinv, x = cache["h"], cache["a"], cache["dinv"], cache["x"]
    G = np.zeros_like(P)
    G[OFF["wo"]:OFF["bo"]] = dy @ h[NL]
    G[OFF["bo"]] = dy.sum()
    g = r(np.outer(dy, wo))
    for l in range(NL - 1, -1, -1):
        dz = r(g * (h[l + 1] > 0))
        G[OFF[f"b{l + 1}"]:OFF[f"b{l + 1}"] + H] = dz.sum(0)
        X = np.concatenate([h[l], a[l]], 1)
        G[OFF[f"W{l + 1}"]:OFF[f"b{l + 1}"]] = (dz.T @ X).ravel()
        dha = r(dz @ r(Ws[l]))
        g = r(dha[:, :H] + A.T @ (dha[:, H:] * dinv[:, None]))
    dz0 = r(g * (h[0] > 0))
    G[OFF["b0"]:OFF["b0"] + H] = dz0.sum(0)
    G[OFF["W0"]:OFF["b0"]] = (dz0.T @ x).ravel()
    return G
