"""MG-preconditioned inverse iteration (rsb_algo = 1; fiedler.c:94-321,
multigrid.c:184-317, mat.c:17-61 and 704-764).

The oracle's pieces are pinned by what the mathematics fixes:
  * every coarse matrix is P^T L P for the aggregation P (row i -> i /
    factor) with its diagonal reset to the row sum -- a zero-row-sum
    Laplacian, so equal to P^T L P exactly (dense numpy, integer values);
  * the projection solves L x = b for b orthogonal to the constant vector
    (against the pseudo-inverse) to its stopping tolerance;
  * the inverse iteration converges to the Fiedler vector: a strip of quads
    is a weighted path graph, lambda_2 = 2 (2 - 2 cos(pi / n)) and
    v_2(i) ~ cos(pi (i + 1/2) / n) in closed form, and on small boxes v_2
    and lambda_2 match dense eigh;
  * its V-cycle is linear and one preconditioned step lowers the energy
    error of a Laplacian solve.
The device path must build the same coarse matrices (bit-exact) and reach
the same Fiedler vectors, iteration counts and partitions as the oracle."""
import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M


def flush(msg):
    print(msg, flush=True)


def dense_laplacian(oracle, vtx, nv, elems):
    rp, col, w = oracle.dual_graph(vtx[elems], nv)
    n = len(elems)
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = w[rp[i]:rp[i + 1]]
    return np.diag(A.sum(1)) - A


def strip(n):
    """1 x n strip of quads: the dual graph is a path with weights 2."""
    vtx, xyz, nv = M.quad_grid(n, 1)
    return vtx, xyz, nv


@pytest.mark.parametrize("factor", [2, 3, 4])
def test_oracle_coarse_matrices_are_galerkin_products(oracle, factor):
    vtx, xyz, nv = M.box_hex(6, 5, 4, jitter=0.1, seed=2)
    n = vtx.shape[0]
    el = np.random.default_rng(factor).permutation(n).astype(np.int32)
    L = dense_laplacian(oracle, vtx, nv, el)
    sizes, _ = oracle.mg_levels(vtx, nv, el, factor)
    assert sizes[0] == n and sizes[-1] == 1
    Lf = L
    for lev in range(1, len(sizes)):
        nf = Lf.shape[0]
        nc = (nf + factor - 1) // factor
        assert sizes[lev] == nc
        P = np.zeros((nf, nc))
        P[np.arange(nf), np.arange(nf) // factor] = 1.0
        Lc = P.T @ Lf @ P
        rp, col, w, d = oracle.mg_level_matrix(vtx, nv, el, lev, factor)
        got = np.zeros((nc, nc))
        for i in range(nc):
            got[i, col[rp[i]:rp[i + 1]]] = -w[rp[i]:rp[i + 1]]
        got[np.arange(nc), np.arange(nc)] = d
        assert np.array_equal(got, Lc), lev
        Lf = Lc


def test_oracle_projection_solves_the_laplacian(oracle):
    vtx, xyz, nv = M.box_hex(7, 6, 5, jitter=0.1, seed=3)
    n = vtx.shape[0]
    el = np.arange(n, dtype=np.int32)
    L = dense_laplacian(oracle, vtx, nv, el)
    b = np.random.default_rng(0).standard_normal(n)
    b -= b.mean()
    b /= np.linalg.norm(b)
    tol = 1e-8
    x, it = oracle.mg_project(vtx, nv, el, b, miter=50, tol=tol)
    assert it < 50
    assert np.linalg.norm(L @ x - b) <= 1.01 * tol * np.linalg.norm(b)
    xs = np.linalg.pinv(L) @ b
    assert np.linalg.norm((x - x.mean()) - xs) <= 1e-6 * np.linalg.norm(xs)


def test_oracle_vcycle_is_linear_and_reduces_the_error(oracle):
    vtx, xyz, nv = M.box_hex(8, 6, 5, jitter=0.1, seed=4)
    n = vtx.shape[0]
    el = np.arange(n, dtype=np.int32)
    L = dense_laplacian(oracle, vtx, nv, el)
    rng = np.random.default_rng(1)
    r1, r2 = rng.standard_normal(n), rng.standard_normal(n)
    z1 = oracle.mg_vcycle(vtx, nv, el, r1)
    z2 = oracle.mg_vcycle(vtx, nv, el, r2)
    z12 = oracle.mg_vcycle(vtx, nv, el, 2.0 * r1 - 3.0 * r2)
    assert np.allclose(z12, 2.0 * z1 - 3.0 * z2, rtol=0, atol=1e-12 * np.abs(z12).max())
    # one preconditioned Richardson step lowers the energy error of L x = b
    xs = rng.standard_normal(n)
    xs -= xs.mean()
    b = L @ xs
    z = oracle.mg_vcycle(vtx, nv, el, b)
    z -= z.mean()
    e0, e1 = xs, xs - z
    assert e1 @ L @ e1 < 0.9 * (e0 @ L @ e0)


@pytest.mark.parametrize("n", [16, 37, 64])
def test_oracle_fiedler_on_a_strip_has_the_closed_form(oracle, n):
    vtx, xyz, nv = strip(n)
    el = np.arange(n, dtype=np.int32)
    f, it, npass = oracle.fiedler_mg(vtx, nv, el)
    v2 = np.cos(np.pi * (np.arange(n) + 0.5) / n)
    v2 /= np.linalg.norm(v2)
    assert abs(abs(f @ v2) - 1) < 1e-9
    lam = 2.0 * (2.0 - 2.0 * np.cos(np.pi / n))
    L = dense_laplacian(oracle, vtx, nv, el)
    assert abs(f @ L @ f - lam) < 1e-8 * lam
    assert it >= npass >= 1


def test_oracle_fiedler_matches_dense_eigh(oracle):
    vtx, xyz, nv = M.pipe_hex(3, 8, 12, seed=3, amp=0.1)
    n = vtx.shape[0]
    el = np.arange(n, dtype=np.int32)
    L = dense_laplacian(oracle, vtx, nv, el)
    ev, V = np.linalg.eigh(L)
    assert ev[2] - ev[1] > 0.1 * ev[1]  # a simple lambda_2
    f, it, npass = oracle.fiedler_mg(vtx, nv, el)
    assert abs(abs(f @ V[:, 1]) - 1) < 1e-8
    assert abs(f @ L @ f - ev[1]) < 1e-8 * ev[1]


@pytest.mark.parametrize("n", [1, 2, 3, 7, 20, 50])
def test_oracle_jacobi_smallest_eigenpair_matches_lapack(oracle, n):
    """The Gram refinement's eigen-solve (cyclic Jacobi) against LAPACK
    (numpy eigh) on random symmetric matrices, a repeated smallest
    eigenvalue's leading digits included."""
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + np.diag(rng.uniform(0.0, 1.0, n))
    lam, v = oracle.sym_smallest(A)
    ev, V = np.linalg.eigh(A)
    assert abs(lam - ev[0]) <= 1e-11 * max(1.0, abs(ev).max())
    assert abs(np.linalg.norm(v) - 1) < 1e-12
    assert np.linalg.norm(A @ v - lam * v) <= 1e-10 * max(1.0, abs(ev).max())


@pytest.mark.parametrize("mesh", ["pipe", "box"])
def test_oracle_gram_refinement_converges_to_the_fiedler_vector(oracle, mesh):
    """rsb_mg_grammian = 1 (fiedler.c:263-307): each pass's z is replaced by
    the Rayleigh-Ritz vector of the Laplacian over the orthonormalised
    iterates; the result is still v_2 (dense eigh) and, since the refined
    start is at least as good as the last iterate, it takes no more passes
    than without the refinement."""
    if mesh == "pipe":
        vtx, xyz, nv = M.pipe_hex(3, 8, 12, seed=3, amp=0.1)
    else:
        vtx, xyz, nv = M.box_hex(14, 9, 6, jitter=0.1, seed=2)
    n = vtx.shape[0]
    el = np.arange(n, dtype=np.int32)
    L = dense_laplacian(oracle, vtx, nv, el)
    ev, V = np.linalg.eigh(L)
    f0, it0, np0 = oracle.fiedler_mg(vtx, nv, el, tol=1e-8)
    f1, it1, np1 = oracle.fiedler_mg(vtx, nv, el, tol=1e-8, grammian=1)
    assert abs(abs(f1 @ V[:, 1]) - 1) < 1e-8
    assert abs(f1 @ L @ f1 - ev[1]) < 1e-8 * ev[1]
    assert np1 <= np0 and np1 >= 1


@pytest.mark.parametrize("grammian", [0, 1])
def test_oracle_partition_with_mg_is_balanced_and_close_to_lanczos(oracle, grammian):
    vtx, xyz, nv = M.box_hex(12, 10, 8, jitter=0.1, seed=9)
    P = 8
    a, _ = oracle.part_mesh(vtx, xyz, nv, P, rsb_algo=1, mg_factor=2, mg_grammian=grammian)
    b, _ = oracle.part_mesh(vtx, xyz, nv, P)
    assert np.array_equal(np.bincount(a, minlength=P), np.bincount(b, minlength=P))
    qa, qb = oracle.quality(a, vtx, nv, P), oracle.quality(b, vtx, nv, P)
    assert qa["edge_cut"] <= 1.1 * qb["edge_cut"]


# ---------------------------------------------------------------------------
# device

def _one_segment(vtx, xyz, nv, p=2):
    d = pa.Device()
    d.set_mesh(vtx, xyz, nv)
    d.set_segments([0], [vtx.shape[0]], [p], [0])
    d.prepartition(1)
    d.level_laplacian()
    return d


@pytest.mark.gpu
@pytest.mark.parametrize("factor", [2, 4])
def test_device_mg_levels_equal_oracle(oracle, factor):
    flush(f"mg levels factor {factor}")
    vtx, xyz, nv = M.box_hex(9, 7, 6, jitter=0.1, seed=5)
    with _one_segment(vtx, xyz, nv) as d:
        order = d.order()
        d.fiedler_mg(1, factor=factor)
        sizes, _ = oracle.mg_levels(vtx, nv, order, factor)
        for lev in range(1, len(sizes)):
            got = d.mg_level_csr(lev, 0)
            ref = oracle.mg_level_matrix(vtx, nv, order, lev, factor)
            for g_, r_ in zip(got, ref):
                assert np.array_equal(g_, r_), lev


@pytest.mark.gpu
def test_device_mg_fiedler_matches_oracle(oracle):
    flush("mg fiedler one segment")
    vtx, xyz, nv = M.pipe_hex(3, 8, 12, seed=3, amp=0.1)
    with _one_segment(vtx, xyz, nv) as d:
        order = d.order()
        out = d.fiedler_mg(1)
        f = d.fiedler_values()
    fo, it, npass = oracle.fiedler_mg(vtx, nv, order)
    assert abs(abs(f @ fo) - 1) < 1e-9
    assert f @ fo > 0  # same start, same sign
    assert out["npass"][0] == npass
    assert abs(out["niter"][0] - it) <= 1


@pytest.mark.gpu
def test_device_mg_batched_segments_match_oracle(oracle):
    flush("mg batched segments")
    vtx, xyz, nv = M.box_hex(14, 11, 9, seed=6)
    n = vtx.shape[0]
    sizes = [300, 1, 700, n - 1001]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(starts, sizes, [2, 2, 3, 5], [0, 2, 4, 7])
        d.prepartition(1)
        order = d.order()
        d.level_laplacian()
        out = d.fiedler_mg(4)
        f = d.fiedler_values()
    for s, (a, sz) in enumerate(zip(starts, sizes)):
        if sz < 2:
            assert f[a] == 0.0
            continue
        fo, it, npass = oracle.fiedler_mg(vtx, nv, order[a:a + sz])
        assert abs(abs(f[a:a + sz] @ fo) - 1) < 1e-8, s
        assert out["npass"][s] == npass, s


@pytest.mark.gpu
def test_device_mg_partition_matches_oracle(oracle):
    flush("mg partition 16^3 P=8")
    vtx, xyz, nv = M.box_hex(16, 16, 16, jitter=0.1, seed=1)
    P = 8
    part, _ = pa.part_mesh(vtx, xyz, nv, P, rsb_algo=1)
    ref, rep = oracle.part_mesh(vtx, xyz, nv, P, rsb_algo=1, mg_factor=2)
    assert np.array_equal(np.bincount(part, minlength=P), np.bincount(ref, minlength=P))
    q, qo = pa.quality(part, vtx, nv, P), oracle.quality(ref, vtx, nv, P)
    assert abs(q["edge_cut"] - qo["edge_cut"]) <= 0.02 * qo["edge_cut"]
    r = pa.last_report()
    assert r["niter_max"][0] > 0


@pytest.mark.gpu
def test_device_mg_option_errors():
    vtx, xyz, nv = M.box_hex(4, 4, 4, seed=1)
    for kw in (dict(rsb_algo=1, rsb_mg_factor=1),
               dict(rsb_algo=1, rsb_mg_grammian=1, rsb_max_passes=65), dict(rsb_algo=2)):
        with pytest.raises(pa.ParrsbError) as e:
            pa.part_mesh(vtx, xyz, nv, 2, **kw)
        assert e.value.rc == 3


@pytest.mark.gpu
def test_device_mg_graph_replays_equal_eager_launches(monkeypatch):
    """Each CG iteration after the first runs as a replay of one captured
    graph; with the graph off (PARRSB_AMD_MG_GRAPH=0) the same kernels run
    eagerly.  Every replay feeds the next, so equal results over a whole
    partition (eager solves in between) check all of them."""
    flush("mg graph vs eager")
    vtx, xyz, nv = M.box_hex(12, 10, 9, jitter=0.1, seed=7)
    a, _ = pa.part_mesh(vtx, xyz, nv, 6, rsb_algo=1)
    ra = pa.last_report()
    monkeypatch.setenv("PARRSB_AMD_MG_GRAPH", "0")
    b, _ = pa.part_mesh(vtx, xyz, nv, 6, rsb_algo=1)
    rb = pa.last_report()
    monkeypatch.delenv("PARRSB_AMD_MG_GRAPH")
    c, _ = pa.part_mesh(vtx, xyz, nv, 6, rsb_algo=1)
    assert np.array_equal(a, b) and np.array_equal(a, c)
    assert ra["niter_max"] == rb["niter_max"] and ra["lambda_min"] == rb["lambda_min"]


@pytest.mark.gpu
def test_device_mg_lds_tiles_equal_global_form(monkeypatch):
    """The level-0 smoother and SpMV read the 16-bit LDS tiles; with the
    tiles off (PARRSB_AMD_NO_TILES=1) they read the 32-bit global form.  The
    entries and their order are the same, so the solves are bit-identical
    (a mesh big enough for several tiles per segment and halos)."""
    flush("mg lds tiles vs global form")
    vtx, xyz, nv = M.box_hex(24, 20, 18, jitter=0.1, seed=8)
    a, _ = pa.part_mesh(vtx, xyz, nv, 6, rsb_algo=1)
    ra = pa.last_report()
    monkeypatch.setenv("PARRSB_AMD_NO_TILES", "1")
    b, _ = pa.part_mesh(vtx, xyz, nv, 6, rsb_algo=1)
    rb = pa.last_report()
    monkeypatch.delenv("PARRSB_AMD_NO_TILES")
    assert np.array_equal(a, b)
    assert ra["niter_max"] == rb["niter_max"] and ra["lambda_min"] == rb["lambda_min"]


# ---------------------------------------------------------------------------
# Gram-matrix refinement on the device (rsb_mg_grammian = 1)

@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 2, 3, 7, 20, 50, 63])
def test_device_jacobi_equals_oracle_bit_for_bit(oracle, n):
    """The Ritz step's eigen-solve: one wave splits the rows of each
    rotation, every entry taking the serial sweep's operations in order
    (no contraction), so eigenvalue and vector equal the oracle's exactly;
    the oracle's is pinned against LAPACK."""
    flush(f"device jacobi n={n}")
    rng = np.random.default_rng(100 + n)
    B = rng.standard_normal((n, n))
    A = B @ B.T + np.diag(rng.uniform(0.0, 1.0, n))
    lam, v = pa.test_sym_smallest(A)
    lo, vo = oracle.sym_smallest(A)
    assert lam == lo
    assert np.array_equal(v, vo)


@pytest.mark.gpu
@pytest.mark.parametrize("mesh", ["pipe", "box"])
def test_device_mg_gram_fiedler_matches_oracle(oracle, mesh):
    flush(f"mg gram fiedler {mesh}")
    if mesh == "pipe":
        vtx, xyz, nv = M.pipe_hex(3, 8, 12, seed=3, amp=0.1)
    else:
        vtx, xyz, nv = M.box_hex(14, 9, 6, jitter=0.1, seed=2)
    for tol in (1e-5, 1e-8):
        with _one_segment(vtx, xyz, nv) as d:
            order = d.order()
            out = d.fiedler_mg(1, tol=tol, grammian=1)
            f = d.fiedler_values()
        fo, it, npass = oracle.fiedler_mg(vtx, nv, order, tol=tol, grammian=1)
        assert abs(abs(f @ fo) - 1) < 1e-9, tol
        assert f @ fo > 0, tol
        assert out["npass"][0] == npass, tol
        assert abs(out["niter"][0] - it) <= npass, tol


@pytest.mark.gpu
def test_device_mg_gram_batched_segments_match_oracle(oracle):
    """Segments grow their bases independently (some stop early, one is too
    small to bisect): each equals its own oracle solve."""
    flush("mg gram batched segments")
    vtx, xyz, nv = M.box_hex(14, 11, 9, seed=6)
    n = vtx.shape[0]
    sizes = [300, 1, 700, n - 1001]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(starts, sizes, [2, 2, 3, 5], [0, 2, 4, 7])
        d.prepartition(1)
        order = d.order()
        d.level_laplacian()
        out = d.fiedler_mg(4, tol=1e-7, grammian=1)
        f = d.fiedler_values()
    for s, (a, sz) in enumerate(zip(starts, sizes)):
        if sz < 2:
            assert f[a] == 0.0
            continue
        fo, it, npass = oracle.fiedler_mg(vtx, nv, order[a:a + sz], tol=1e-7, grammian=1)
        assert abs(abs(f[a:a + sz] @ fo) - 1) < 1e-8, s
        assert out["npass"][s] == npass, s


@pytest.mark.gpu
def test_device_mg_gram_partition_matches_oracle(oracle):
    flush("mg gram partition 16^3 P=8")
    vtx, xyz, nv = M.box_hex(16, 16, 16, jitter=0.1, seed=1)
    P = 8
    part, _ = pa.part_mesh(vtx, xyz, nv, P, rsb_algo=1, rsb_mg_grammian=1)
    ref, rep = oracle.part_mesh(vtx, xyz, nv, P, rsb_algo=1, mg_factor=2, mg_grammian=1)
    assert np.array_equal(np.bincount(part, minlength=P), np.bincount(ref, minlength=P))
    q, qo = pa.quality(part, vtx, nv, P), oracle.quality(ref, vtx, nv, P)
    assert abs(q["edge_cut"] - qo["edge_cut"]) <= 0.02 * qo["edge_cut"]
    assert pa.last_report()["niter_max"][0] > 0
