"""Pins the CPU oracle against mathematics that fixes each of its functions:
closed forms, brute force, library routines (scipy) and invariants.  The
reference has no golden vectors or known-answer tests of its own and cannot be
built here (it needs gslib), so these checks are what pins the oracle."""
import itertools

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
import scipy.sparse.csgraph as csg

from parrsb_amd import meshgen as M


# ---------------------------------------------------------------- helpers
def incidence_adjacency(vtx, nv):
    """Shared-vertex counts A = B B^T from the element/vertex incidence B."""
    vtx = np.asarray(vtx).reshape(-1, nv)
    nel = vtx.shape[0]
    _, vid = np.unique(vtx.ravel(), return_inverse=True)
    B = sp.csr_matrix((np.ones(vtx.size), (np.repeat(np.arange(nel), nv), vid)),
                      shape=(nel, vid.max() + 1))
    A = (B @ B.T).tocsr()
    A.setdiag(0)
    A.eliminate_zeros()
    return A


def dense_laplacian(vtx, nv, elems=None):
    A = incidence_adjacency(vtx, nv)
    if elems is not None:
        A = A[elems][:, elems]
    A = A.toarray()
    return np.diag(A.sum(1)) - A


def closed_form_box_counts(dims):
    """Off-diagonal count and total weight of a structured box's dual graph:
    w(offset) = prod_d s(|delta_d|), s(0) = 2, s(1) = 1."""
    nnz = 0
    wsum = 0
    for off in itertools.product((-1, 0, 1), repeat=len(dims)):
        if not any(off):
            continue
        cnt = 1
        w = 1
        for n, d in zip(dims, off):
            cnt *= n - abs(d)
            w *= 2 if d == 0 else 1
        nnz += cnt
        wsum += cnt * w
    return nnz, wsum


# ---------------------------------------------------------------- ranks
def test_rank_counts_exhaustive(oracle):
    """sort.c:48-70 assigns ranks per global index; count them and compare
    with the oracle's closed forms."""
    for n in range(0, 70):
        for p in range(1, 20):
            counts = np.zeros(p, dtype=int)
            nelt, nrem = n // p, n % p
            for i in range(n):
                if nrem == 0:
                    counts[i // nelt] += 1
                else:
                    s = p - nrem
                    t1 = nelt * s
                    counts[i // nelt if i < t1 else s + (i - t1) // (nelt + 1)] += 1
            for r in range(p):
                assert oracle.rank_count(n, p, r) == counts[r]
            assert oracle.lower_count(n, p) == counts[: (p + 1) // 2].sum()
            e0, e1 = oracle.balance_expected(n, p, 0), oracle.balance_expected(n, p, 1)
            assert e0 + e1 == n
            # each half ends up with its ranks at q or q+1 elements
            q = n // p
            assert q * ((p + 1) // 2) <= e0 <= (q + 1) * ((p + 1) // 2)
            assert q * (p // 2) <= e1 <= (q + 1) * (p // 2)


# ---------------------------------------------------------------- centroids
def test_centroids_exact_on_dyadic_lattice(oracle):
    vtx, xyz, nv = M.box_hex(4, 3, 2, jitter=0.0, h=1.0 / 16)
    c = oracle.centroids(xyz.reshape(-1, nv * 3), nv)
    k, j, i = np.meshgrid(np.arange(2), np.arange(3), np.arange(4), indexing="ij")
    ref = np.stack([i.ravel() + 0.5, j.ravel() + 0.5, k.ravel() + 0.5], -1) / 16
    assert np.array_equal(c, ref)


def test_centroids_sum_in_corner_order(oracle):
    vtx, xyz, nv = M.box_hex(5, 4, 3, jitter=0.1, seed=7)
    c = oracle.centroids(xyz, nv)
    acc = np.zeros((xyz.shape[0], 3))
    for v in range(nv):
        acc = acc + xyz[:, v, :]
    assert np.array_equal(c, acc / nv)
    vtx, xyz, nv = M.quad_grid(6, 5)
    c = oracle.centroids(xyz, nv)
    acc = np.zeros((xyz.shape[0], 2))
    for v in range(nv):
        acc = acc + xyz[:, v, :]
    assert np.array_equal(c[:, :2], acc / nv) and not c[:, 2].any()


# ---------------------------------------------------------------- dual graph
@pytest.mark.parametrize("mesh", ["hex", "quad", "holed"])
def test_dual_graph_matches_incidence_product(oracle, mesh):
    if mesh == "hex":
        vtx, xyz, nv = M.box_hex(5, 4, 3)
    elif mesh == "quad":
        vtx, xyz, nv = M.quad_grid(7, 5)
    else:
        vtx, xyz, nv = M.holed_quads(40, 32, seed=11, frac=0.1, rmin=2, rmax=5)
    rowptr, col, w = oracle.dual_graph(vtx, nv)
    A = incidence_adjacency(vtx, nv)
    A.sort_indices()
    assert np.array_equal(rowptr, A.indptr)
    assert np.array_equal(col, A.indices)
    assert np.array_equal(w, A.data.astype(np.int32))


@pytest.mark.parametrize("dims", [(16, 16, 16), (7, 5, 3), (9, 4)])
def test_dual_graph_closed_form_counts(oracle, dims):
    if len(dims) == 3:
        vtx, _, nv = M.box_hex(*dims)
    else:
        vtx, _, nv = M.quad_grid(*dims)
    rowptr, col, w = oracle.dual_graph(vtx, nv)
    nnz, wsum = closed_form_box_counts(dims)
    assert rowptr[-1] == nnz and w.sum() == wsum
    if dims == (16, 16, 16):
        assert (nnz, wsum) == (93240, 205560)


# ---------------------------------------------------------------- Laplacian
def test_laplacian_apply_matches_dense(oracle):
    rng = np.random.default_rng(0)
    vtx, xyz, nv = M.box_hex(6, 5, 4)
    n = vtx.shape[0]
    elems = np.sort(rng.choice(n, size=80, replace=False)).astype(np.int32)
    for sub in (np.arange(n, dtype=np.int32), elems, elems[::-1].copy()):
        L = dense_laplacian(vtx, nv, sub)
        u = rng.standard_normal(sub.size)
        v = oracle.laplacian_apply(vtx, nv, sub, u)
        assert np.allclose(v, L @ u, rtol=0, atol=1e-12 * np.abs(L).sum(1).max())
        assert abs(v.sum()) < 1e-10 * np.linalg.norm(u) * 56


# ---------------------------------------------------------------- TQLI
@pytest.mark.parametrize("n", [1, 2, 3, 10, 50])
def test_tqli_matches_scipy(oracle, n):
    rng = np.random.default_rng(n)
    d = rng.uniform(0.5, 5, n)
    e = rng.uniform(-1, 1, max(n - 1, 0))
    if n >= 10:
        e[3] = 0.0  # forces a deflation in the middle
    rc, ev, z = oracle.tqli(d, e)
    assert rc == 0
    ref = sla.eigh_tridiagonal(d, e, eigvals_only=True)
    assert np.allclose(np.sort(ev), ref, rtol=1e-12, atol=1e-12)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    for k in range(n):
        assert abs(np.linalg.norm(z[k]) - 1) < 1e-12
        assert np.linalg.norm(T @ z[k] - ev[k] * z[k]) < 1e-10 * np.abs(T).max()


# ---------------------------------------------------------------- CG-Lanczos
def _start_vector(n):
    v = np.arange(1, n + 1, dtype=np.float64)
    v[: n // 2] += 1000.0 * n
    v -= v.sum() / n
    return v / np.sqrt(v @ v)


def test_cg_lanczos_tridiagonal_is_projection(oracle):
    """Saad, Iterative Methods for Sparse Linear Systems, 6.7.3: T built from
    the CG coefficients equals Q L Q^T for the normalised residual basis Q
    (checked before orthogonality is lost)."""
    vtx, xyz, nv = M.box_hex(6, 6, 5, seed=3)
    n = vtx.shape[0]
    elems = np.arange(n, dtype=np.int32)
    k, d, e, Q = oracle.lanczos_pass(vtx, nv, elems, _start_vector(n), 8, tol=1e-12)
    assert k == 8
    L = dense_laplacian(vtx, nv)
    Qk = Q[:k]
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    assert np.allclose(Qk @ Qk.T, np.eye(k), atol=1e-9)
    assert np.allclose(Qk @ L @ Qk.T, T, atol=1e-9 * np.abs(T).max())


def test_lanczos_stops_after_one_step_on_an_eigenvector(oracle):
    vtx, xyz, nv = M.box_hex(7, 5, 4, jitter=0.0)
    L = dense_laplacian(vtx, nv)
    lam, V = np.linalg.eigh(L)
    assert lam[2] - lam[1] > 1e-3  # lambda_2 of a 7x5x4 box is simple
    n = L.shape[0]
    f, it, lam_r, rn = oracle.fiedler(vtx, nv, np.arange(n, dtype=np.int32), init=V[:, 1])
    assert it == 1
    assert abs(lam_r - lam[1]) < 1e-10
    assert abs(abs(f @ V[:, 1]) - 1) < 1e-12


def test_fiedler_matches_dense_eigensolver(oracle):
    """The start vector is built from positions in the pre-partition order
    (fiedler.c:572-578), so the elements are passed in RCB order, as rsb.c
    does (rsb.c:288-301)."""
    vtx, xyz, nv = M.box_hex(9, 6, 5, seed=4)
    n = vtx.shape[0]
    order = oracle.rcb_order(oracle.centroids(xyz, nv), 3, np.arange(n, dtype=np.int32), 2)
    L = dense_laplacian(vtx, nv, order)
    lam, V = np.linalg.eigh(L)
    f, it, lam_r, rn = oracle.fiedler(vtx, nv, order)
    assert abs(lam_r - lam[1]) / lam[1] < 1e-6
    assert abs(abs(f @ V[:, 1]) - 1) < 1e-6
    assert abs(f.sum()) < 1e-10 and abs(np.linalg.norm(f) - 1) < 1e-12
    assert rn[2] < rn[1]  # final residual below target


def test_fiedler_on_degenerate_cube_lies_in_eigenspace(oracle):
    """16^3 box: lambda_2 has multiplicity 3, so only the eigenspace is fixed."""
    vtx, xyz, nv = M.box_hex(16, 16, 16, seed=1)
    A = incidence_adjacency(vtx, nv)
    L = sp.diags(np.asarray(A.sum(1)).ravel()) - A
    lam, V = np.linalg.eigh(L.toarray())
    assert abs(lam[3] - lam[1]) < 1e-10 * lam[1] and lam[4] - lam[3] > 0.1
    n = L.shape[0]
    f, it, lam_r, rn = oracle.fiedler(vtx, nv, np.arange(n, dtype=np.int32))
    assert abs(lam_r - lam[1]) / lam[1] < 1e-6
    proj = V[:, 1:4].T @ f
    assert abs(np.linalg.norm(proj) - 1) < 1e-5


# ---------------------------------------------------------------- components
def test_components_match_scipy(oracle):
    rng = np.random.default_rng(5)
    vtx, xyz, nv = M.holed_quads(60, 50, seed=9, frac=0.2, rmin=2, rmax=8)
    A = incidence_adjacency(vtx, nv)
    n = vtx.shape[0]
    for trial in range(6):
        elems = np.arange(n, dtype=np.int32) if trial == 0 else \
            np.sort(rng.choice(n, size=n // (trial + 1), replace=False)).astype(np.int32)
        ref = csg.connected_components(A[elems][:, elems], directed=False)[0]
        assert oracle.components(vtx, nv, elems) == ref


# ---------------------------------------------------------------- RCB
def _check_rcb(coord, ndim, order, p, oracle):
    """Walk the RCB tree of a result: every node must be split at its
    virtual-rank boundary along its axis of largest extent (ties -> lower
    axis), every key left of the split <= every key right of it."""
    n = order.size
    if n <= 1:
        return
    c = coord[order]
    ext = c[:, :ndim].max(0) - c[:, :ndim].min(0)
    axis = 0
    for d in range(1, ndim):
        if ext[d] > ext[axis]:
            axis = d
    nl = oracle.lower_count(n, p) if p > 1 else n // 2
    assert c[:nl, axis].max(initial=-np.inf) <= c[nl:, axis].min(initial=np.inf)
    _check_rcb(coord, ndim, order[:nl], (p + 1) // 2 if p > 1 else 1, oracle)
    _check_rcb(coord, ndim, order[nl:], p // 2 if p > 1 else 1, oracle)


@pytest.mark.parametrize("p", [1, 2, 3, 5, 8, 13])
def test_rcb_split_structure(oracle, p):
    rng = np.random.default_rng(p)
    for ndim in (2, 3):
        coord = np.zeros((150, 3))
        coord[:, :ndim] = rng.uniform(0, 1, (150, ndim)) * [3, 1, 2][:ndim]
        start = rng.permutation(150).astype(np.int32)
        out = oracle.rcb_order(coord, ndim, start, p)
        assert sorted(out.tolist()) == list(range(150))
        _check_rcb(coord, ndim, out, p, oracle)


def test_rcb_ties_keep_current_order(oracle):
    coord = np.zeros((4, 3))
    coord[:, 0] = [1, 0, 1, 0]
    out = oracle.rcb_order(coord, 3, np.arange(4, dtype=np.int32), 1)
    # x-extent 1 -> stable sort by x: [1,3,0,2]; halves have zero extent and
    # keep their order.
    assert out.tolist() == [1, 3, 0, 2]
    out = oracle.rcb_order(coord, 3, np.array([2, 3, 0, 1], dtype=np.int32), 1)
    assert out.tolist() == [3, 1, 2, 0]


# ---------------------------------------------------------------- RIB
def test_sym_dominant_matches_eigh(oracle):
    """Jacobi dominant eigenvector vs LAPACK (numpy.linalg.eigh), with the
    sign fixed by the largest component; 2 x 2 and 3 x 3, random and
    nearly diagonal matrices."""
    rng = np.random.default_rng(11)
    for trial in range(200):
        n = 2 + trial % 2
        B = rng.standard_normal((n, n)) * (1e-6 if trial % 7 == 0 else 1.0)
        A = B @ B.T + (np.diag(rng.uniform(0, 3, n)) if trial % 5 == 0 else 0)
        w, V = np.linalg.eigh(A)
        if w[-1] - w[-2] < 1e-6 * abs(w[-1]):
            continue
        v = V[:, -1] * np.sign(V[np.argmax(np.abs(V[:, -1])), -1])
        ev = oracle.sym_dominant(A, n)
        assert np.allclose(ev, v, atol=1e-12), (A, ev, v)


def test_sym_dominant_degenerate_picks_lowest_axis(oracle):
    assert oracle.sym_dominant(np.eye(3) * 2.0, 3).tolist() == [1.0, 0.0, 0.0]
    assert oracle.sym_dominant(np.diag([1.0, 3.0, 3.0]), 3).tolist() == [0.0, 1.0, 0.0]


def test_rib_axis_is_principal_axis(oracle):
    """Points stretched along a known direction: the axis is that direction
    (sign: largest component positive) and the mean is the sample mean."""
    rng = np.random.default_rng(3)
    d = np.array([0.48, -0.6, 0.64])
    d /= np.linalg.norm(d)
    pts = rng.standard_normal((400, 1)) * 5.0 * d + rng.standard_normal((400, 3)) * 0.2 + 1.5
    m, ev = oracle.rib_axis(pts, 3, np.arange(400, dtype=np.int32))
    assert np.allclose(m, pts.mean(0), rtol=1e-13)
    C = (pts - pts.mean(0)).T @ (pts - pts.mean(0))
    w, V = np.linalg.eigh(C)
    v = V[:, -1] * np.sign(V[np.argmax(np.abs(V[:, -1])), -1])
    assert np.allclose(ev, v, atol=1e-12)
    assert abs(abs(ev @ d) - 1) < 1e-3


def _check_rib(coord, ndim, order, p, oracle):
    """Every node is split at its virtual-rank boundary along the principal
    axis of its own centroids (numpy eigh): keys left <= keys right."""
    n = order.size
    if n <= 1:
        return
    c = coord[order][:, :ndim]
    x = c - c.mean(0)
    w, V = np.linalg.eigh(x.T @ x)
    key = x @ V[:, -1]
    nl = oracle.lower_count(n, p) if p > 1 else n // 2
    lo, hi = key[:nl], key[nl:]
    # either orientation of the axis; sign is the oracle's convention
    tol = 1e-9 * (np.abs(key).max() + 1)
    assert lo.max(initial=-np.inf) <= hi.min(initial=np.inf) + tol or \
        lo.min(initial=np.inf) >= hi.max(initial=-np.inf) - tol
    _check_rib(coord, ndim, order[:nl], (p + 1) // 2 if p > 1 else 1, oracle)
    _check_rib(coord, ndim, order[nl:], p // 2 if p > 1 else 1, oracle)


@pytest.mark.parametrize("p", [1, 2, 3, 5, 8, 13])
def test_rib_split_structure(oracle, p):
    rng = np.random.default_rng(100 + p)
    for ndim in (2, 3):
        coord = np.zeros((150, 3))
        coord[:, :ndim] = rng.uniform(0, 1, (150, ndim)) @ rng.standard_normal((ndim, ndim))
        start = rng.permutation(150).astype(np.int32)
        out = oracle.rib_order(coord, ndim, start, p)
        assert sorted(out.tolist()) == list(range(150))
        _check_rib(coord, ndim, out, p, oracle)


def test_rib_on_axis_aligned_box_is_rcb_like(oracle):
    """A box elongated along y: the first RIB split is a y split, as RCB's."""
    vtx, xyz, nv = M.box_hex(6, 14, 5, jitter=0.1, seed=2)
    c = oracle.centroids(xyz, nv)
    n = c.shape[0]
    o = oracle.rib_order(c, 3, np.arange(n, dtype=np.int32), 2)
    nl = oracle.lower_count(n, 2)
    assert c[o[:nl], 1].max() < c[o[nl:], 1].min()


# ---------------------------------------------------------------- full path
def test_part_mesh_16cube_is_octant_split(oracle):
    vtx, xyz, nv = M.box_hex(16, 16, 16, jitter=0.1, seed=1)
    part, rep = oracle.part_mesh(vtx, xyz, nv, 8)
    q = oracle.quality(part, vtx, nv, 8)
    # three plane cuts: one 16x16 section, two 16x8, four 8x8; a section a x b
    # cuts 4ab + 4(a-1)b + 4a(b-1) + 4(a-1)(b-1) of shared-vertex weight
    sec = lambda a, b: 4 * a * b + 4 * (a - 1) * b + 4 * a * (b - 1) + 4 * (a - 1) * (b - 1)
    assert q["edge_cut"] == sec(16, 16) + 2 * sec(16, 8) + 4 * sec(8, 8) == 11164
    assert q["max_part"] == q["min_part"] == 512
    assert q["components"] == 8 and q["max_components"] == 1
    # neighbouring parts after each cut: halves touch 1, quarters (sharing
    # the centre line) 3, octants (sharing the centre vertex) 7
    assert rep["nbrs_min"] == rep["nbrs_max"] == [1, 3, 7]
    assert rep["nbrs_mean"] == [1.0, 3.0, 7.0]
    assert rep["ncomp_max"] == [1, 1, 1]
    c = oracle.centroids(xyz, nv)
    for r in range(8):
        box = c[part == r]
        # an octant: each axis range lies in one half of the cube
        for d in range(3):
            assert box[:, d].max() < 0.5 + 0.01 or box[:, d].min() > 0.5 - 0.01


def test_part_mesh_uneven_parts_balance(oracle):
    vtx, xyz, nv = M.box_hex(16, 16, 16, jitter=0.1, seed=1)
    part, rep = oracle.part_mesh(vtx, xyz, nv, 3)
    sizes = np.bincount(part, minlength=3)
    # cut 1: lower ranks {0,1} hold 2730 after the sort but expect 2731, so
    # the upper side sends one interface element; cut 2 on the lower half
    # (2731 over 2 ranks) again moves one element.
    assert sizes.tolist() == [1366, 1365, 1365]
    assert rep["nbalanced"] == [1, 1]


def brute_balance(vtx, elems, fied, nl, sender_lower, scount):
    """The balance step of rsb.c:118-229 by sets: the sender's interface
    elements are those sharing a vertex id with any element of the other
    half; if there are at least scount of them, the scount nearest the cut
    change side (upper side: smallest Fiedler values, lower side: largest;
    equal values in position order), and each half keeps the positions of
    its members in order."""
    n = len(elems)
    V = [set(int(v) for v in vtx[e]) for e in elems]
    snd = range(0, nl) if sender_lower else range(nl, n)
    rcv = range(nl, n) if sender_lower else range(0, nl)
    touched = set()
    for i in rcv:
        touched |= V[i]
    iface = [i for i in snd if V[i] & touched]
    if scount <= 0 or len(iface) < scount:
        return np.asarray(elems), np.asarray(fied), nl, iface
    key = (lambda i: (-fied[i], i)) if sender_lower else (lambda i: (fied[i], i))
    move = set(sorted(iface, key=key)[:scount])
    low = [(i < nl) != (i in move) for i in range(n)]
    new = [i for i in range(n) if low[i]] + [i for i in range(n) if not low[i]]
    return np.asarray(elems)[new], np.asarray(fied)[new], sum(low), iface


def _sorted_segment(oracle, vtx, xyz, nv, n, p_rcb=4):
    """A connected-ish segment: the first n elements of the mesh's RCB order,
    in ascending order of their Fiedler values (the state rsb.c:308 leaves)."""
    ndim = 3 if nv == 8 else 2
    c = oracle.centroids(xyz, nv)
    order = oracle.rcb_order(c, ndim, np.arange(vtx.shape[0], dtype=np.int32), p_rcb)[:n]
    order = np.sort(order).astype(np.int32)
    order = oracle.rcb_order(c, ndim, order, 2)
    f, _, _, _ = oracle.fiedler(vtx, nv, order)
    s = np.argsort(f, kind="stable")
    return order[s], f[s]


@pytest.mark.parametrize("case", ["holed_p64", "hex_p13", "holed_lower_sends", "ties"])
def test_balance_step_matches_brute_force(oracle, case):
    """oracle_balance_step against the set-based restatement above, on
    segments where many elements must move (n = 64q + 32 over 64 ranks: the
    upper half is 16 over; 13 ranks with remainder 6: 3 over), on a lower
    sender (which the rank arithmetic never produces, but the step takes),
    and on Fiedler values with many exact ties (position order decides)."""
    if case.startswith("holed"):
        vtx, xyz, nv = M.holed_quads(72, 60, seed=9, frac=0.15, rmin=2, rmax=6)
    else:
        vtx, xyz, nv = M.box_hex(14, 12, 10, jitter=0.1, seed=3)
    nel = vtx.shape[0]
    if case == "holed_p64":
        p = 64
        n = (nel - 32) // 64 * 64 + 32
    elif case == "hex_p13":
        p = 13
        n = (nel - 6) // 13 * 13 + 6
    else:
        p = 2
        n = nel - nel % 7
    el, f = _sorted_segment(oracle, vtx, xyz, nv, n)
    if case == "ties":
        f = np.round(f * 8) / 8  # a handful of distinct values, runs of ties
        s = np.argsort(f, kind="stable")
        el, f = el[s], f[s]
    nl = oracle.lower_count(n, p)
    e0, e1 = oracle.balance_expected(n, p, 0), oracle.balance_expected(n, p, 1)
    if case in ("holed_p64", "hex_p13"):
        assert nl < e0 and n - nl > e1  # the upper half sends
        snd, sc = 0, (n - nl) - e1
        assert sc == {"holed_p64": 16, "hex_p13": 3}[case]
    elif case == "holed_lower_sends":
        nl, snd, sc = n // 2 + 40, 1, 40
    else:
        nl, snd, sc = n // 2 - 25, 0, 25
    got_el, got_f, got_nl = oracle.balance_step(vtx, nv, el, f, nl, snd, sc)
    ref_el, ref_f, ref_nl, iface = brute_balance(vtx, el, f, nl, snd, sc)
    assert len(iface) >= sc  # something moves
    assert got_nl == ref_nl == (nl - sc if snd else nl + sc)
    assert np.array_equal(got_el, ref_el) and np.array_equal(got_f, ref_f)
    # the movers are interface elements of the sender and nothing else moved
    moved = set(got_el[:got_nl]) ^ set(el[:nl])
    assert len(moved) == sc and moved <= set(el[iface])
    # too few interface elements: nothing moves (rsb.c:194, 221-224)
    big = len(iface) + 1
    o2, f2, nl2 = oracle.balance_step(vtx, nv, el, f, nl, snd, big)
    assert nl2 == nl and np.array_equal(o2, el) and np.array_equal(f2, f)


def test_part_mesh_moves_many_elements(oracle):
    """A whole partition whose first cut must move 16 elements (n = 64q + 32
    into 64 parts): every cut ends at the balanced sizes and the first level
    reports a move."""
    vtx, xyz, nv = M.holed_quads(72, 60, seed=9, frac=0.15, rmin=2, rmax=6)
    n = (vtx.shape[0] - 32) // 64 * 64 + 32
    vtx, xyz = np.ascontiguousarray(vtx[:n]), np.ascontiguousarray(xyz[:n])
    part, rep = oracle.part_mesh(vtx, xyz, nv, 64)
    assert rep["nbalanced"][0] == 1

    def rec(n, p, out):
        if p == 1:
            out.append(n)
            return
        nl = oracle.balance_expected(n, p, 0)
        rec(nl, (p + 1) // 2, out)
        rec(n - nl, p // 2, out)
    sizes = []
    rec(n, 64, sizes)
    assert np.bincount(part, minlength=64).tolist() == sizes


def test_quality_matches_incidence(oracle):
    rng = np.random.default_rng(2)
    vtx, xyz, nv = M.box_hex(6, 5, 4)
    part = rng.integers(0, 5, vtx.shape[0]).astype(np.int32)
    q = oracle.quality(part, vtx, nv, 5)
    A = sp.triu(incidence_adjacency(vtx, nv)).tocoo()
    cut = part[A.row] != part[A.col]
    assert q["edge_cut"] == A.data[cut].sum() and q["cut_edges"] == cut.sum()
    vol = 0
    for v in np.unique(vtx):
        k = np.unique(part[(vtx == v).any(1)]).size
        vol += k * (k - 1)
    assert q["comm_volume"] == vol
    Af = incidence_adjacency(vtx, nv)
    tot = sum(csg.connected_components(Af[part == r][:, part == r])[0] for r in range(5))
    assert q["components"] == tot
