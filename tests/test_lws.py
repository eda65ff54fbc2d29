"""Learned warm start: the numpy oracle's loss is pinned by closed forms and
its backward pass by finite differences (CPU); the device network (BF16 MFMA
forward and backward) is checked against the oracle with BF16 rounding
emulated at the same points (GPU)."""
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
    lam = np.linalg.eigvalsh(W.laplacian(Ad))[1]
    rng = np.random.default_rng(1)
    P = W.init_params(2)
    y, cache = W.forward(P, x, Ad)
    G = W.backward(P, cache, Ad, W.dloss(y, Ad, lam))
    idx = np.concatenate([rng.choice(np.arange(W.OFF["W0"], W.OFF["b0"]), 8),
                          rng.choice(np.arange(W.OFF["W1"], W.OFF["b1"]), 8),
                          rng.choice(np.arange(W.OFF["W4"], W.OFF["b4"]), 8),
                          [W.OFF["b2"] + 3, W.OFF["wo"] + 5, W.OFF["bo"]]])
    eps = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = (W.loss(W.forward(Pp, x, Ad)[0], Ad, lam) -
              W.loss(W.forward(Pm, x, Ad)[0], Ad, lam)) / (2 * eps)
        assert abs(fd - G[i]) <= 1e-6 * (1 + abs(G).max()) + 1e-4 * abs(fd), (i, fd, G[i])


def test_loss_closed_forms():
    """Zero on the Fiedler vector, lambda_k / lambda_2 - 1 on the k-th
    eigenvector, invariant under y -> a y + b, and the gradient vanishes on
    eigenvectors (stationary points of the Rayleigh quotient)."""
    A, _ = segment_graph((4, 3, 3))
    Ad = A.toarray()
    w, V = np.linalg.eigh(W.laplacian(Ad))
    lam = w[1]
    assert abs(W.loss(V[:, 1], Ad, lam)) < 1e-10
    assert abs(W.loss(V[:, 5], Ad, lam) - (w[5] / w[1] - 1)) < 1e-9
    rng = np.random.default_rng(3)
    y = rng.standard_normal(Ad.shape[0])
    assert abs(W.loss(y, Ad, lam) - W.loss(-3.0 * y + 7.0, Ad, lam)) < 1e-9
    assert np.abs(W.dloss(V[:, 3], Ad, lam)).max() < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("dims", [(5, 4, 3), (9, 8, 7)])
def test_device_network_matches_oracle(dims):
    import parrsb_amd as pa
    A, coord = segment_graph(dims)
    Ad = A.toarray()
    rowlen = np.diff(A.indptr)
    level = 5.0
    lam = float(np.linalg.eigvalsh(W.laplacian(Ad))[1])
    P = W.init_params(5).astype(np.float32).astype(np.float64)
    y, grad, L = pa.lws_test(coord.T.copy(), level, A.indptr, A.indices, A.data.astype(np.int32),
                             lam, P)
    x = W.features(coord, rowlen, level)
    yo, cache = W.forward(P, x, Ad, emulate_bf16=True)
    # forward: BF16 activations, FP32 accumulation in a different order
    assert np.max(np.abs(y - yo)) <= 2e-2 * np.max(np.abs(yo)) + 1e-3
    Lo = W.loss(yo, Ad, lam)
    assert abs(L - Lo) <= 2e-2 * max(Lo, 1e-3)
    Go = W.backward(P, cache, Ad, W.dloss(yo, Ad, lam), emulate_bf16=True)
    cos = grad @ Go / (np.linalg.norm(grad) * np.linalg.norm(Go))
    assert cos > 0.995, cos
    assert np.linalg.norm(grad - Go) <= 0.05 * np.linalg.norm(Go)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["zero", "nan"])
def test_failed_warm_start_falls_back_to_reference_start(kind):
    """A network whose prediction is constant (zero weights) or not finite
    (NaN weights) sends every segment to the reference start: the partition,
    the iteration counts and the Fiedler values equal the plain solver's."""
    import parrsb_amd as pa
    vtx, xyz, nv = M.box_hex(12, 10, 8, jitter=0.1, seed=3)
    P = np.zeros(pa.warm_start_nparams(), dtype=np.float32)
    if kind == "nan":
        P[:] = np.nan
    try:
        pa.set_warm_start(None)
        part0, _ = pa.part_mesh(vtx, xyz, nv, 16)
        r0 = pa.last_report()
        pa.set_warm_start(P, 1, 0)
        part1, _ = pa.part_mesh(vtx, xyz, nv, 16)
        r1 = pa.last_report()
    finally:
        pa.set_warm_start(None)
    assert r1["warm_start"] == 1 and r0["warm_start"] == 0
    assert r1["nfallback"] == r1["nsegs"]
    assert np.array_equal(part0, part1)
    assert r1["niter_max"] == r0["niter_max"] and r1["niter_min"] == r0["niter_min"]


@pytest.mark.gpu
def test_warm_start_levels_report_fallback_iterations():
    """Device API: with a random (untrained) network and one learned pass,
    segments either meet tol from the prediction (fallback 0) or fall back
    after exactly one pass of max_iter iterations and then repeat the plain
    solve exactly."""
    import parrsb_amd as pa
    vtx, xyz, nv = M.box_hex(16, 12, 10, jitter=0.1, seed=4)
    nel = vtx.shape[0]
    with pa.Device() as dev:
        dP = dev.upload(W.init_params(7).astype(np.float32))
        dev.set_mesh(vtx, xyz, nv)
        dev.set_segments([0], [nel], [4], [0])
        dev.prepartition(1)
        dev.level_laplacian()
        plain = dev.fiedler(1, 64, 50, 50, 1e-5)
        dev.lws_set_params(dP, 0.0, 1)
        warm = dev.fiedler(1, 64, 50, 50, 1e-5, init_mode=2)
        # above the size limit: the reference start, nothing else
        dev.lws_set_params(dP, 0.0, 1, 100)
        gated = dev.fiedler(1, 64, 50, 50, 1e-5, init_mode=2)
        dev.free(dP)
    assert gated["fallback"][0] == -1
    assert gated["niter"][0] == plain["niter"][0] and gated["lambda"][0] == plain["lambda"][0]
    fb = int(warm["fallback"][0])
    if fb:
        # the fallback is the plain solve, bit for bit
        assert fb - 1 == 50  # one learned pass of max_iter iterations
        assert warm["niter"][0] == plain["niter"][0] + 50
        assert warm["lambda"][0] == plain["lambda"][0]
        assert warm["resid"][0] == plain["resid"][0]
    else:
        # the stop rule is the reference's (CG residual reduced by tol), which
        # bounds the eigen-residual only loosely; a smooth prediction meets
        # it after few iterations with a larger, still small, eigen-residual
        assert warm["niter"][0] < 50
        assert warm["resid"][0] < 2e-2


@pytest.mark.gpu
def test_trained_start_gains_on_small_segments():
    """The committed network (parrsb_amd/lws_weights.bin, trained on the GPU
    by parrsb_amd.lws_train) on a held-out warped pipe: at the levels whose
    segments are within the size gate (<= 2048 elements) the learned start
    needs fewer Lanczos iterations than the reference start; the levels
    above the gate are solved exactly as without it; the edge cut stays
    within 2%."""
    import os
    import parrsb_amd as pa
    from parrsb_amd import lws_train
    path = os.path.join(os.path.dirname(pa.__file__), "lws_weights.bin")
    P = lws_train.load_params(path)
    vtx, xyz, nv = M.pipe_hex(8, 64, 128, seed=4321, amp=0.12)
    nparts = 64
    pa.set_precision(32)
    try:
        pa.set_warm_start(None)
        part0, _ = pa.part_mesh(vtx, xyz, nv, nparts)
        r0 = pa.last_report()
        pa.set_warm_start(P, 1, 2048)
        part1, _ = pa.part_mesh(vtx, xyz, nv, nparts)
        r1 = pa.last_report()
    finally:
        pa.set_warm_start(None)
        pa.set_precision(64)
    n = vtx.shape[0]
    gated = [l for l in range(r0["nlevels"]) if n / 2 ** l <= 2048]
    above = [l for l in range(r0["nlevels"]) if l not in gated]
    assert gated, "no level within the gate"
    for l in above:
        assert r1["niter_mean"][l] == r0["niter_mean"][l]
    plain = sum(r0["niter_mean"][l] for l in gated)
    learned = sum(r1["niter_mean"][l] for l in gated)
    print("gated levels", gated, "niter_mean plain", plain, "learned", learned)
    assert learned <= 0.95 * plain
    q0, q1 = pa.quality(part0, vtx, nv, nparts), pa.quality(part1, vtx, nv, nparts)
    assert q1["edge_cut"] <= 1.02 * q0["edge_cut"]

