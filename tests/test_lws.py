"""Learned warm start: the numpy oracle's backward pass is pinned by finite
differences (CPU); the device network (BF16 MFMA forward and backward) is
checked against the oracle with BF16 rounding emulated at the same points
(GPU)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import lws_oracle as W
from parrsb_amd import meshgen as M


def segment_graph(dims=(5, 4, 3), seed=0):
    vtx, xyz, nv = M.box_hex(*dims, jitter=0.1, seed=seed)
    nel = vtx.shape[0]
    _, vid = np.unique(vtx.ravel(), return_inverse=True)
    B = sp.csr_matrix((np.ones(vtx.size), (np.repeat(np.arange(nel), nv), vid)))
    A = (B @ B.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    A.sort_indices()
    coord = xyz.mean(1)
    return A, coord


def test_oracle_backward_matches_finite_differences():
    A, coord = segment_graph()
    Ad = A.toarray()
    rowlen = np.diff(A.indptr)
    x = W.features(coord, rowlen, 3)
    rng = np.random.default_rng(1)
    t = np.cos(np.pi * x[:, 0]) + 0.1 * rng.standard_normal(x.shape[0])
    P = W.init_params(2)
    y, cache = W.forward(P, x, Ad)
    G = W.backward(P, cache, Ad, W.dloss(y, t))
    idx = np.concatenate([rng.choice(np.arange(W.OFF["W0"], W.OFF["b0"]), 8),
                          rng.choice(np.arange(W.OFF["W1"], W.OFF["b1"]), 8),
                          rng.choice(np.arange(W.OFF["W4"], W.OFF["b4"]), 8),
                          [W.OFF["b2"] + 3, W.OFF["wo"] + 5, W.OFF["bo"]]])
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (W.loss(W.forward(Pp, x, Ad)[0], t) - W.loss(W.forward(Pm, x, Ad)[0], t)) / (2 * eps)
        assert abs(fd - G[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, G[i])


def test_loss_is_sign_and_scale_invariant():
    rng = np.random.default_rng(3)
    t = rng.standard_normal(50)
    y = rng.standard_normal(50)
    assert abs(W.loss(y, t) - W.loss(-3.0 * y + 7.0, t)) < 1e-12
    assert abs(W.loss(2.0 * t + 1.0, t)) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(5, 4, 3), (9, 8, 7)])
def test_device_network_matches_oracle(dims):
    import parrsb_amd as pa
    A, coord = segment_graph(dims)
    Ad = A.toarray()
    rowlen = np.diff(A.indptr)
    level = 5.0
    rng = np.random.default_rng(4)
    t = np.cos(np.pi * W.features(coord, rowlen, level)[:, 0]) + 0.05 * rng.standard_normal(coord.shape[0])
    t = (t - t.mean()) / np.linalg.norm(t - t.mean())
    P = W.init_params(5).astype(np.float32).astype(np.float64)
    y, grad, L = pa.lws_test(coord.T.copy(), level, A.indptr, A.indices, A.data.astype(np.int32),
                             t.astype(np.float32), P)
    x = W.features(coord, rowlen, level)
    yo, cache = W.forward(P, x, Ad, emulate_bf16=True)
    # forward: BF16 activations, FP32 accumulation in a different order
    assert np.max(np.abs(y - yo)) <= 2e-2 * np.max(np.abs(yo)) + 1e-3
    Lo = W.loss(yo, t)
    assert abs(L - Lo) <= 2e-2 * max(Lo, 1e-3)
    Go = W.backward(P, cache, Ad, W.dloss(yo, t), emulate_bf16=True)
    cos = grad @ Go / (np.linalg.norm(grad) * np.linalg.norm(Go))
    assert cos > 0.995, cos
    assert np.linalg.norm(grad - Go) <= 0.05 * np.linalg.norm(Go)
