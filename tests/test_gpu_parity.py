"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer and index work is compared bit for bit: centroids, dual graph, RCB
orders, level Laplacian, tridiagonal eigen-solve, component counts, balance
sizes and quality metrics.  Floating-point Lanczos results are compared with
tolerances written in each test (the summation order of the device
reductions is not the oracle's, so trajectories agree to rounding)."""
import numpy as np
import pytest

import parrsb_amd as pa
from parrsb_amd import meshgen as M

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


MESHES = {
    "hex": lambda: M.box_hex(9, 7, 5, jitter=0.1, seed=3),
    "quad": lambda: M.quad_grid(23, 17),
    "holed": lambda: M.holed_quads(48, 40, seed=11, frac=0.12, rmin=2, rmax=6),
}


@pytest.mark.parametrize("name", sorted(MESHES))
def test_centroids_and_dual_graph_bit_exact(oracle, name):
    flush(f"graph {name}")
    vtx, xyz, nv = MESHES[name]()
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        assert np.array_equal(d.centroids(), oracle.centroids(xyz, nv))
        rp, col, w = d.graph()
    orp, ocol, ow = oracle.dual_graph(vtx, nv)
    assert np.array_equal(rp, orp) and np.array_equal(col, ocol) and np.array_equal(w, ow)


def test_dual_graph_closed_form_16cube():
    vtx, xyz, nv = M.box_hex(16, 16, 16)
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        rp, col, w = d.graph()
    assert rp[-1] == 93240 and w.sum() == 205560


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("p", [1, 2, 3, 8, 13])
def test_rcb_order_bit_exact(oracle, name, p):
    flush(f"rcb {name} p={p}")
    vtx, xyz, nv = MESHES[name]()
    n = vtx.shape[0]
    ndim = 3 if nv == 8 else 2
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments([0], [n], [p], [0])
        d.prepartition(1)
        got = d.order()
    if p == 1:
        # segments with one rank are final: the pre-partition leaves them alone
        ref = np.arange(n, dtype=np.int32)
    else:
        ref = oracle.rcb_order(oracle.centroids(xyz, nv), ndim, np.arange(n, dtype=np.int32), p)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("rankmax", ["0", "64", "1024"])
def test_rcb_tail_kernel_equals_global_rounds(oracle, monkeypatch, rankmax):
    """Nodes of <= 1024 elements are finished by one LDS kernel; the global
    rounds alone (PARRSB_AMD_RCB_TAIL=0) and the oracle give the same orders,
    also next to a segment with more ranks than elements (which keeps the
    global rounds).  rankmax: the sub-node size up to which the kernel ranks
    by counting instead of sorting (0: always the bitonic sort, 1024: always
    counting)."""
    flush(f"rcb tail, rankmax {rankmax}")
    monkeypatch.setenv("PARRSB_AMD_RCB_RANKMAX", rankmax)
    vtx, xyz, nv = M.box_hex(36, 30, 24, jitter=0.1, seed=12)
    n = vtx.shape[0]
    c = oracle.centroids(xyz, nv)
    for sizes, ps in (([n - 3000, 3000], [7, 1]), ([n - 10, 10], [13, 16])):
        starts = [0, sizes[0]]
        orders = []
        for env in ("1", "0"):
            monkeypatch.setenv("PARRSB_AMD_RCB_TAIL", env)
            with pa.Device() as d:
                d.set_mesh(vtx, xyz, nv)
                d.set_segments(starts, sizes, ps, [0, ps[0]])
                d.prepartition(1)
                orders.append(d.order())
            monkeypatch.delenv("PARRSB_AMD_RCB_TAIL", raising=False)
        assert np.array_equal(orders[0], orders[1])
        for s0, sz, p in zip(starts, sizes, ps):
            seg = np.arange(s0, s0 + sz, dtype=np.int32)
            ref = oracle.rcb_order(c, 3, seg, p) if p > 1 else seg
            assert np.array_equal(orders[0][s0:s0 + sz], ref)


def test_rcb_many_segments_bit_exact(oracle):
    vtx, xyz, nv = M.quad_grid(40, 31)  # lattice centroids: ties everywhere
    n = vtx.shape[0]
    c = oracle.centroids(xyz, nv)
    sizes = [311, 1, 0, 400, 528]
    ps = [3, 4, 1, 1, 6]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(starts, sizes, ps, np.arange(5))
        d.prepartition(1)
        got = d.order()
    for s0, sz, p in zip(starts, sizes, ps):
        seg = np.arange(s0, s0 + sz, dtype=np.int32)
        ref = oracle.rcb_order(c, 2, seg, p) if p > 1 else seg
        assert np.array_equal(got[s0:s0 + sz], ref)


def test_level_laplacian_is_induced_subgraph(oracle):
    vtx, xyz, nv = M.box_hex(10, 8, 6, seed=5)
    n = vtx.shape[0]
    sizes = [150, 180, 150]
    starts = [0, 150, 330]
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(starts, sizes, [2, 3, 2], [0, 2, 5])
        d.prepartition(1)
        order = d.order()
        d.level_laplacian()
        rp, col, w, row_elem = d.level_csr()
    segof = np.zeros(n, dtype=int)
    for s, (a, b) in enumerate(zip(starts, sizes)):
        segof[order[a:a + b]] = s
    orp, ocol, ow = oracle.dual_graph(vtx, nv)
    real = row_elem >= 0
    assert real.sum() == n
    for r in np.nonzero(real)[0]:
        e = row_elem[r]
        mine = sorted((row_elem[c], ww) for c, ww in zip(col[rp[r]:rp[r + 1]], w[rp[r]:rp[r + 1]]))
        ref = sorted((c, ww) for c, ww in zip(ocol[orp[e]:orp[e + 1]], ow[orp[e]:orp[e + 1]])
                     if segof[c] == segof[e])
        assert mine == ref


def test_tridiagonal_solver_bit_exact(oracle):
    """The QL solve against the oracle's, bit for bit: up to 128 with the
    eigenvector matrix in LDS, above (passes of rsb_max_iter > 128) with it
    in global memory."""
    rng = np.random.default_rng(0)
    for n in [1, 2, 3, 7, 20, 50, 64, 100, 128, 129, 200, 300]:
        d = rng.uniform(0.1, 4, n)
        e = -rng.uniform(0.01, 1.5, max(n - 1, 0))
        if n > 10:
            e[5] = 0.0
        lam, y = pa.test_tqli(d, e)
        rc, ev, z = oracle.tqli(d, e)
        imin = int(np.argmin(np.abs(ev)))  # first index of the minimum
        assert lam == ev[imin]
        assert np.array_equal(y, z[imin]), n


def test_fast_tridiagonal_solver_matches_lapack():
    """The FP32 path's eigen-solve (multisection + inverse iteration) against
    LAPACK (scipy eigh_tridiagonal): smallest-|lambda| eigenpair to rounding,
    sign y[0] >= 0; includes a zero off-diagonal (split matrix), a
    Lanczos-like spectrum spread over 5 decades and a clustered pair."""
    import scipy.linalg as sla
    rng = np.random.default_rng(1)
    cases = []
    for n in [1, 2, 3, 7, 20, 50, 64, 100, 128]:
        d = rng.uniform(0.1, 4, n)
        e = -rng.uniform(0.01, 1.5, max(n - 1, 0))
        if n > 10:
            e[5] = 0.0
        cases.append((d, e))
    # T = Q^T diag(lam) Q tridiagonalised by Lanczos on a diagonal matrix
    lam_true = np.concatenate([[1e-4, 1.3e-4], np.geomspace(1e-3, 10.0, 48)])
    q = rng.standard_normal(50)
    q /= np.linalg.norm(q)
    Q = [q]
    al, be = [], []
    for j in range(50):
        v = lam_true * Q[j]
        al.append(Q[j] @ v)
        v -= al[j] * Q[j] + (be[j - 1] * Q[j - 1] if j else 0)
        for u in Q:
            v -= (u @ v) * u
        if j < 49:
            be.append(np.linalg.norm(v))
            Q.append(v / be[j])
    cases.append((np.array(al), np.array(be)))
    for d, e in cases:
        n = d.size
        lam, y = pa.test_tqli(d, e, fast=True)
        w, V = sla.eigh_tridiagonal(d, e) if n > 1 else (d.copy(), np.ones((1, 1)))
        i = int(np.argmin(np.abs(w)))
        scale = np.abs(w).max()
        assert abs(lam - w[i]) <= 1e-13 * scale, (n, lam, w[i])
        v = V[:, i] * (1 if V[0, i] >= 0 else -1)
        gap = np.min(np.abs(np.delete(w, i) - w[i])) if n > 1 else scale
        assert np.abs(y - v).max() <= 1e-14 * scale / gap + 1e-12, n
        assert y[0] >= 0


def _segments_after_rcb(vtx, xyz, nv, p):
    d = pa.Device()
    d.set_mesh(vtx, xyz, nv)
    n = vtx.shape[0]
    d.set_segments([0], [n], [p], [0])
    d.prepartition(1)
    return d


@pytest.mark.parametrize("tiles", ["lds", "global"])
@pytest.mark.parametrize("dims", [(9, 6, 5), (12, 10, 7), (40, 30, 20), (70, 52)])
def test_lanczos_fp64_matches_oracle(oracle, dims, tiles, monkeypatch):
    """Both forms of the FP64 Laplacian apply: the 16-bit neighbour entries
    of the LDS tiles, and the global 32-bit path (PARRSB_AMD_NO_TILES=1);
    hex boxes and a quad grid."""
    flush(f"lanczos fp64 {dims} {tiles}")
    if tiles == "global":
        monkeypatch.setenv("PARRSB_AMD_NO_TILES", "1")
    if len(dims) == 3:
        vtx, xyz, nv = M.box_hex(*dims, seed=4)
    else:
        vtx, xyz, nv = M.quad_grid(*dims)
    with _segments_after_rcb(vtx, xyz, nv, 2) as d:
        order = d.order()
        d.level_laplacian()
        wide, _ = d.level_form()
        assert (wide > 0) == (tiles == "global")
        out = d.fiedler(1, precision=64)
        f = d.fiedler_values()
    fo, it, lam, rn = oracle.fiedler(vtx, nv, order)
    # same algorithm and start; reductions in another order: 1e-9 relative on
    # lambda, the Fiedler vectors agree to 1e-7 (lambda_2 is simple here)
    assert abs(out["lambda"][0] - lam) <= 1e-9 * abs(lam)
    assert abs(out["niter"][0] - it) <= 1
    assert abs(abs(f @ fo) - 1) < 1e-7
    assert out["resid"][0] < 1e-3


@pytest.mark.parametrize("name", ["hex", "warped", "quad"])
def test_lanczos_fp32_vertex_form(oracle, name, monkeypatch):
    """The FP32 apply in the vertex (gather-scatter) form of the LDS tiles
    (PARRSB_AMD_VTILES=1) against the neighbour-entry form and the FP64
    oracle: the same operator, so lambda agrees with the oracle to 1e-4 and
    with the entry form to 1e-5, vectors to 1e-3 in angle, iteration counts
    within 2; most tiles of the level have the form."""
    flush(f"lanczos fp32 vertex form {name}")
    vtx, xyz, nv = {"hex": lambda: M.box_hex(40, 30, 20, jitter=0.1, seed=4),
                    "warped": lambda: M.warped_box_hex(40, seed=3),
                    "quad": lambda: M.quad_grid(140, 104)}[name]()
    # both forms with the r-update as its own pass (the vertex form has no
    # folded variant)
    monkeypatch.setenv("PARRSB_AMD_LZ_FOLD_TILES", "0")
    res = {}
    for form in ("1", "0"):
        monkeypatch.setenv("PARRSB_AMD_VTILES", form)
        with _segments_after_rcb(vtx, xyz, nv, 2) as d:
            order = d.order()
            d.level_laplacian()
            wide, vtiles = d.level_form()
            res[form] = (d.fiedler(1, precision=32), d.fiedler_values(), vtiles)
            ntiles = (vtx.shape[0] + 1023) // 1024
    (ov, fv, nvt), (oe, fe, nvt0) = res["1"], res["0"]
    assert nvt0 == 0 and nvt >= 0.9 * ntiles
    fo, it, lam, rn = oracle.fiedler(vtx, nv, order)
    assert abs(ov["lambda"][0] - lam) <= 1e-4 * abs(lam)
    assert abs(ov["lambda"][0] - oe["lambda"][0]) <= 1e-5 * abs(lam)
    assert abs(abs(fv @ fo) - 1) < 1e-3 and abs(abs(fv @ fe) - 1) < 1e-3
    assert abs(ov["niter"][0] - oe["niter"][0]) <= 2


@pytest.mark.parametrize("name", ["hex", "warped", "quad"])
def test_lanczos_fp32_folded_update(oracle, name, monkeypatch):
    """FP32 passes on few running tiles fold the r-update into the next
    apply and take the new residual's norm from r.r - 2 alpha r.w +
    alpha^2 w.w over the stored vectors.  The iterates are the same
    vectors rounded the same way (one fused multiply-add), so the solve
    agrees with the separate r-update pass to FP32 rounding: lambda to 1e-5,
    vectors to 1e-4 in angle, and with the FP64 oracle as the FP32 path does.
    Rounding moves the iteration at which a slowly falling residual crosses
    tol: the counts agree within 1% (at least 3)."""
    flush(f"lanczos fp32 folded update {name}")
    vtx, xyz, nv = {"hex": lambda: M.box_hex(40, 30, 20, jitter=0.1, seed=4),
                    "warped": lambda: M.warped_box_hex(40, seed=3),
                    "quad": lambda: M.quad_grid(140, 104)}[name]()
    res = {}
    for tiles in ("100000", "0"):
        monkeypatch.setenv("PARRSB_AMD_LZ_FOLD_TILES", tiles)
        with _segments_after_rcb(vtx, xyz, nv, 2) as d:
            order = d.order()
            d.level_laplacian()
            res[tiles] = (d.fiedler(1, precision=32), d.fiedler_values(), d.lanczos_traffic())
    (of, ff, tf), (ou, fu, tu) = res["100000"], res["0"]
    fo, it, lam, rn = oracle.fiedler(vtx, nv, order)
    assert abs(of["lambda"][0] - lam) <= 1e-4 * abs(lam)
    assert abs(of["lambda"][0] - ou["lambda"][0]) <= 1e-5 * abs(lam)
    assert abs(abs(ff @ fu) - 1) < 1e-4 and abs(abs(ff @ fo) - 1) < 1e-3
    assert abs(of["niter"][0] - ou["niter"][0]) <= max(3, 0.01 * ou["niter"][0])
    # the byte model follows the launches: 7 vector streams per folded
    # iteration against 9 (apply 5 + update 4)
    assert tf[0] < tu[0]


@pytest.mark.parametrize("name", ["hex", "warped", "quad"])
def test_lanczos_12bit_entries_equal_16bit_form(oracle, name, monkeypatch):
    """The tiles' 12-bit form (indices ordered by weight class, the class
    counts per row) is the same operator as the 16-bit entries summed in
    another order: FP64 lambda agrees to 1e-11, the vectors to 1e-9 and the
    iteration counts within one; FP32 as the FP32 path does.  The FP64 solve
    reads the 12-bit form only with PARRSB_AMD_E12=2."""
    flush(f"lanczos 12-bit entries {name}")
    vtx, xyz, nv = {"hex": lambda: M.box_hex(40, 30, 20, jitter=0.1, seed=4),
                    "warped": lambda: M.warped_box_hex(40, seed=3),
                    "quad": lambda: M.quad_grid(140, 104)}[name]()
    for prec in (64, 32):
        res = {}
        for form in ("2", "1", "0"):
            monkeypatch.setenv("PARRSB_AMD_E12", form)
            with _segments_after_rcb(vtx, xyz, nv, 2) as d:
                d.level_laplacian()
                assert d.level_e12() == (form != "0")
                res[form] = (d.fiedler(1, precision=prec), d.fiedler_values())
        # the default ("1"): FP32 reads the 12-bit form, FP64 the 16-bit one
        same = res["0"] if prec == 64 else res["2"]
        assert np.array_equal(res["1"][1], same[1]), prec
        (o12, f12), (o16, f16) = res["2"], res["0"]
        lam = o16["lambda"][0]
        assert abs(o12["lambda"][0] - lam) <= (1e-11 if prec == 64 else 1e-5) * abs(lam), prec
        assert abs(abs(f12 @ f16) - 1) < (1e-9 if prec == 64 else 1e-4), prec
        assert abs(o12["niter"][0] - o16["niter"][0]) <= max(1 if prec == 64 else 3,
                                                             0.01 * o16["niter"][0]), prec


def test_lanczos_12bit_form_falls_back_on_other_weights(oracle, monkeypatch):
    """A level with a weight other than 1, 2 or 4 (face neighbours of an
    element with one corner renumbered share three vertices) keeps the 16-bit
    entries: level_e12() is 0, the FP32 solve equals the one with the 12-bit
    form switched off bit for bit, and the FP64 solve agrees with the
    oracle."""
    flush("lanczos 12-bit fallback")
    vtx, xyz, nv = M.box_hex(16, 12, 10, jitter=0.1, seed=4)
    vtx = vtx.copy()
    picks = np.arange(7, vtx.shape[0], 97)
    vtx[picks, 0] = vtx.max() + 1 + np.arange(picks.size)
    rp, col, w = oracle.dual_graph(vtx, nv)
    assert (w == 3).any()
    res = {}
    for form in ("1", "0"):
        monkeypatch.setenv("PARRSB_AMD_E12", form)
        with _segments_after_rcb(vtx, xyz, nv, 2) as d:
            order = d.order()
            d.level_laplacian()
            assert d.level_e12() == 0
            res[form] = (d.fiedler(1, precision=32), d.fiedler_values(),
                         d.fiedler(1, precision=64), d.fiedler_values())
    assert np.array_equal(res["1"][1], res["0"][1])
    out, f = res["1"][2], res["1"][3]
    fo, it, lam, rn = oracle.fiedler(vtx, nv, order)
    assert abs(out["lambda"][0] - lam) <= 1e-9 * abs(lam)
    assert abs(abs(f @ fo) - 1) < 1e-7


def test_lanczos_batched_segments_match_oracle(oracle):
    """Four segments of different sizes solved in one batch: each equals the
    oracle's solve of that segment alone."""
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
        out = d.fiedler(4, precision=64)
        f = d.fiedler_values()
    for s, (a, sz) in enumerate(zip(starts, sizes)):
        if sz < 2:
            assert f[a] == 0.0
            continue
        fo, it, lam, rn = oracle.fiedler(vtx, nv, order[a:a + sz])
        assert abs(out["lambda"][s] - lam) <= 1e-9 * abs(lam), s
        assert abs(abs(f[a:a + sz] @ fo) - 1) < 1e-6, s


def test_lanczos_fp32_close_to_fp64(oracle):
    vtx, xyz, nv = M.box_hex(12, 10, 7, seed=4)
    with _segments_after_rcb(vtx, xyz, nv, 2) as d:
        order = d.order()
        d.level_laplacian()
        out = d.fiedler(1, precision=32)
        f = d.fiedler_values()
    fo, it, lam, rn = oracle.fiedler(vtx, nv, order)
    # FP32 vectors: lambda to 1e-4 relative, vector to 1e-3 in angle
    assert abs(out["lambda"][0] - lam) <= 1e-4 * abs(lam)
    assert abs(abs(f @ fo) - 1) < 1e-3


def _upper_sends(oracle, n, p):
    nl = oracle.lower_count(n, p)
    e0, e1 = oracle.balance_expected(n, p, 0), oracle.balance_expected(n, p, 1)
    assert nl < e0 and n - nl > e1
    return nl, (n - nl) - e1


def test_bisect_components_and_balance(oracle):
    """Holed quads cut over 64 ranks with n = 64q + 32, so the upper half is
    16 elements over and must send: the components of each half before
    balancing equal the oracle's count on the same halves, and the balanced
    order equals the oracle's balance step applied to the device's own
    Fiedler order, element for element."""
    vtx, xyz, nv = M.holed_quads(72, 60, seed=9, frac=0.15, rmin=2, rmax=6)
    n = (vtx.shape[0] - 32) // 64 * 64 + 32
    vtx, xyz = np.ascontiguousarray(vtx[:n]), np.ascontiguousarray(xyz[:n])
    p = 64
    with _segments_after_rcb(vtx, xyz, nv, p) as d:
        d.level_laplacian()
        d.fiedler(1, precision=64)
        d.fiedler_sort()
        order, f = d.order(), d.fiedler_values()
        nl, sc = _upper_sends(oracle, n, p)
        assert sc == 16
        # rsbgpu_bisect takes the sending half's index (1 = upper)
        nl2, ncomp, moved = d.bisect([nl], [1], [sc], 1)
        got = d.order()
    assert np.all(np.diff(f) >= 0)
    assert ncomp[0] == oracle.components(vtx, nv, order[:nl])
    assert ncomp[1] == oracle.components(vtx, nv, order[nl:])
    ref, _, ref_nl = oracle.balance_step(vtx, nv, order, f, nl, 0, sc)
    assert moved[0] == 1 and nl2[0] == ref_nl == nl + sc
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("case", ["ties_upper", "ties_lower", "too_few"])
def test_balance_pick_with_chosen_values(oracle, case):
    """The balance step on three segments at once with Fiedler values set
    from the host: runs of exactly equal values (position order decides
    among them), the lower half as sender (largest values first), and a
    request for more elements than the sender's interface holds (nothing
    moves).  The device order after rsbgpu_bisect equals the oracle's
    balance step on each segment."""
    vtx, xyz, nv = M.holed_quads(60, 54, seed=5, frac=0.1, rmin=2, rmax=5)
    n = vtx.shape[0]
    sizes = [n // 3, n // 3, n - 2 * (n // 3)]
    starts = [0, sizes[0], sizes[0] + sizes[1]]
    rng = np.random.default_rng(7)
    f = np.empty(n)
    for a, m in zip(starts, sizes):
        # a smooth field along the segment plus coarse quantisation: ties
        x = np.sort(rng.random(m))
        f[a:a + m] = np.round(x * 6) / 6
    snd = {"ties_upper": [0, 0, 0], "ties_lower": [1, 1, 1], "too_few": [0, 1, 0]}[case]
    nl = [m // 2 - 7 if s == 0 else m // 2 + 7 for m, s in zip(sizes, snd)]
    sc = [9, 21, 5] if case != "too_few" else [10**6, 10**6, 10**6]
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(starts, sizes, [2, 2, 2], [0, 2, 4])
        d.level_laplacian()
        d.set_fiedler_values(f)
        # the oracle's flag is "the lower half sends"; rsbgpu_bisect takes the
        # sending half's index (0 = lower, 1 = upper)
        nl2, _, moved = d.bisect(nl, [1 - x for x in snd], sc, 0)
        got = d.order()
    for s in range(3):
        a, m = starts[s], sizes[s]
        el = np.arange(a, a + m, dtype=np.int32)
        ref, _, ref_nl = oracle.balance_step(vtx, nv, el, f[a:a + m], nl[s], snd[s], sc[s])
        assert nl2[s] == ref_nl, s
        assert np.array_equal(got[a:a + m], ref), s
        assert moved[s] == (case != "too_few")


def test_components_of_many_blocks_match_oracle(oracle):
    """Union-find inside tiles (LDS) and across them: six segments of a grid
    of 8 x 8 quad blocks separated by removed rows and columns (no shared
    vertices across a gap), so each half holds many components, some split
    between tiles; every count equals the oracle's on the same half."""
    flush("components of many blocks")
    vtx, xyz, nv = M.quad_grid(134, 107)
    j, i = np.divmod(np.arange(vtx.shape[0]), 134)
    keep = (i % 9 != 8) & (j % 9 != 8)
    vtx, xyz = np.ascontiguousarray(vtx[keep]), np.ascontiguousarray(xyz[keep])
    n = vtx.shape[0]
    sizes = [n // 6] * 5 + [n - 5 * (n // 6)]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int32)
    ps = [2, 3, 2, 5, 2, 4]
    r0 = np.concatenate([[0], np.cumsum(ps)[:-1]]).astype(np.int32)
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(starts, sizes, ps, r0)
        d.prepartition(1)
        d.level_laplacian()
        d.fiedler(6, precision=32)
        d.fiedler_sort()
        order = d.order()
        nl = [oracle.lower_count(sz, p) for sz, p in zip(sizes, ps)]
        _, ncomp, _ = d.bisect(nl, [-1] * 6, [0] * 6, 1)
    total = 0
    for s in range(6):
        a, b, c = starts[s], starts[s] + nl[s], starts[s] + sizes[s]
        want = (oracle.components(vtx, nv, order[a:b]), oracle.components(vtx, nv, order[b:c]))
        assert (ncomp[2 * s], ncomp[2 * s + 1]) == want, s
        total += sum(want)
    assert total >= 100


def test_full_partition_16cube_matches_oracle(oracle):
    flush("full 16^3 P=8")
    vtx, xyz, nv = M.box_hex(16, 16, 16, jitter=0.1, seed=1)
    part, _ = pa.part_mesh(vtx, xyz, nv, 8)
    ref, rep = oracle.part_mesh(vtx, xyz, nv, 8)
    q = pa.quality(part, vtx, nv, 8)
    qo = oracle.quality(ref, vtx, nv, 8)
    assert q["edge_cut"] == qo["edge_cut"] == 11164
    assert q == qo
    r = pa.last_report()
    assert r["ncomp_max"] == [1, 1, 1]
    assert [abs(a - b) <= 1 for a, b in zip(r["niter_max"], rep["niter_max"])] == [True] * 3


@pytest.mark.parametrize("precision", [64, 32])
def test_passes_longer_than_128_iterations(oracle, precision):
    """rsb_max_iter = 150 (fiedler.c:513-561 takes any value): the segments'
    tridiagonal matrices exceed the LDS solve and go through the one with
    its eigenvector matrix in global memory (both precisions).  FP64 gives
    the oracle's partition quality and iteration counts; FP32 its part sizes
    and an edge cut within 2%."""
    flush(f"rsb_max_iter 150, fp{precision}")
    vtx, xyz, nv = M.quad_grid(300, 3)  # a long strip: the first cut takes 459 iterations
    pa.set_precision(precision)
    try:
        part, opts = pa.part_mesh(vtx, xyz, nv, 4, rsb_max_iter=150)
    finally:
        pa.set_precision(64)
    r = pa.last_report()
    ref, rep = oracle.part_mesh(vtx, xyz, nv, 4, max_iter=150)
    assert opts.rsb_max_iter == 150
    assert max(rep["niter_max"]) > 150  # full passes longer than the LDS solve takes
    q, qo = pa.quality(part, vtx, nv, 4), oracle.quality(ref, vtx, nv, 4)
    if precision == 64:
        assert q == qo
        assert all(abs(a - b) <= 1 for a, b in zip(r["niter_max"], rep["niter_max"]))
    else:
        assert q["max_part"] == qo["max_part"] and q["min_part"] == qo["min_part"]
        assert abs(q["edge_cut"] - qo["edge_cut"]) <= 0.02 * qo["edge_cut"]


@pytest.mark.parametrize("P", [3, 5, 7, 12])
def test_uneven_part_counts_match_oracle_sizes(oracle, P):
    flush(f"uneven P={P}")
    vtx, xyz, nv = M.box_hex(16, 16, 16, jitter=0.1, seed=1)
    part, _ = pa.part_mesh(vtx, xyz, nv, P)
    ref, rep = oracle.part_mesh(vtx, xyz, nv, P)
    assert np.array_equal(np.bincount(part, minlength=P), np.bincount(ref, minlength=P))
    assert rep["nbalanced"] == pa.last_report()["nbalanced"]
    # FP64 Lanczos, the same pre-partition, sort and balance: the same parts
    assert np.array_equal(part, ref)


def test_holed_quads_partition_close_to_oracle(oracle):
    flush("holed quads P=16")
    vtx, xyz, nv = M.holed_quads(128, 96, seed=3, frac=0.08, rmin=2, rmax=10)
    P = 16
    part, _ = pa.part_mesh(vtx, xyz, nv, P)
    ref, rep = oracle.part_mesh(vtx, xyz, nv, P)
    q, qo = pa.quality(part, vtx, nv, P), oracle.quality(ref, vtx, nv, P)
    assert q["max_part"] == qo["max_part"] and q["min_part"] == qo["min_part"]
    assert abs(q["edge_cut"] - qo["edge_cut"]) <= 0.02 * qo["edge_cut"]
    assert abs(q["components"] - qo["components"]) <= max(1, 0.02 * qo["components"])


def test_quality_bit_exact(oracle):
    rng = np.random.default_rng(1)
    vtx, xyz, nv = M.holed_quads(50, 40, seed=2, frac=0.1, rmin=2, rmax=5)
    part = rng.integers(0, 7, vtx.shape[0]).astype(np.int32)
    assert pa.quality(part, vtx, nv, 7) == oracle.quality(part, vtx, nv, 7)


def test_fp32_partition_quality(oracle):
    vtx, xyz, nv = M.box_hex(24, 20, 16, jitter=0.1, seed=2)
    P = 16
    ref, _ = oracle.part_mesh(vtx, xyz, nv, P)
    qo = oracle.quality(ref, vtx, nv, P)
    pa.set_precision(32)
    try:
        part, _ = pa.part_mesh(vtx, xyz, nv, P)
    finally:
        pa.set_precision(64)
    q = pa.quality(part, vtx, nv, P)
    assert q["max_part"] == qo["max_part"] and q["min_part"] == qo["min_part"]
    assert abs(q["edge_cut"] - qo["edge_cut"]) <= 0.02 * qo["edge_cut"]


def test_lanczos_plateau_stop(monkeypatch):
    """The opt-in plateau stop (PARRSB_AMD_LZ_PLATEAU, off by default) on the
    64^3 warped box into 4 parts, whose second level has a nearly repeated
    lambda_2 (the reference rule runs all 50 passes there): it ends those
    passes early, only where the Ritz residual from the Lanczos relation is at
    most the ceiling, so the device's explicit residual of the returned vector
    stays under it (this pins the estimate), and the parts stay balanced and
    connected.  (Where lambda_2 is nearly repeated the stopped vector is
    another mixture of the near-equal modes than the one 50 passes reach, so
    the cut can differ; the bench measures the effect on whole workloads.)"""
    flush("lanczos plateau stop")
    vtx, xyz, nv = M.warped_box_hex(64, seed=5)
    reps = {}
    for ratio in ("0", "0.9"):
        monkeypatch.setenv("PARRSB_AMD_LZ_PLATEAU", ratio)
        monkeypatch.setenv("PARRSB_AMD_LZ_PLATEAU_MAX", "1e-2")
        part, _ = pa.part_mesh(vtx, xyz, nv, 4)
        reps[ratio] = (pa.last_report(), pa.quality(part, vtx, nv, 4))
    (r0, q0), (r1, q1) = reps["0"], reps["0.9"]
    assert r0["npass_max"][1] == 50
    assert 4 <= r1["npass_max"][1] < 50
    assert r1["resid_max"][1] <= 1e-2 * (1 + 1e-6)
    assert q1["max_part"] - q1["min_part"] <= 1 and q1["max_components"] == 1


def test_deterministic():
    vtx, xyz, nv = M.box_hex(20, 18, 14, jitter=0.1, seed=8)
    a, _ = pa.part_mesh(vtx, xyz, nv, 24)
    b, _ = pa.part_mesh(vtx, xyz, nv, 24)
    assert np.array_equal(a, b)


def test_rcb_partitioner_matches_oracle_ranks(oracle):
    vtx, xyz, nv = M.box_hex(10, 9, 8, seed=2)
    n = vtx.shape[0]
    P = 6
    part, _ = pa.part_mesh(vtx, xyz, nv, P, partitioner=1)
    order = oracle.rcb_order(oracle.centroids(xyz, nv), 3, np.arange(n, dtype=np.int32), P)
    ref = np.zeros(n, dtype=np.int32)
    # ranks own consecutive ranges of the RCB order, split by the virtual-rank
    # arithmetic
    def ranges(start, length, p, r0):
        if p == 1:
            ref[order[start:start + length]] = r0
            return
        l = oracle.lower_count(length, p)
        ranges(start, l, (p + 1) // 2, r0)
        ranges(start + l, length - l, p // 2, r0 + (p + 1) // 2)
    ranges(0, n, P, 0)
    assert np.array_equal(part, ref)


def test_neighbors_match_brute_force(oracle):
    """RSB_NEIGHBORS on arbitrary segments: distinct other segments sharing a
    vertex, against a numpy count over the incidence structure."""
    vtx, xyz, nv = M.holed_quads(48, 40, seed=11, frac=0.12, rmin=2, rmax=6)
    n = vtx.shape[0]
    rng = np.random.default_rng(3)
    cuts = np.sort(rng.choice(np.arange(1, n), size=11, replace=False))
    cs = np.concatenate([[0], cuts]).astype(np.int32)
    ln = np.diff(np.concatenate([cs, [n]])).astype(np.int32)
    with pa.Device() as d:
        d.set_mesh(vtx, xyz, nv)
        d.set_segments(cs, ln, np.full(cs.size, 2), np.arange(cs.size) * 2)
        order = d.order()
        got = d.neighbors(cs.size)
    seg = np.empty(n, dtype=np.int64)
    for s, (a, l) in enumerate(zip(cs, ln)):
        seg[order[a:a + l]] = s
    pairs = set()
    by_v = {}
    for e in range(n):
        for v in vtx[e]:
            by_v.setdefault(int(v), set()).add(int(seg[e]))
    for ss in by_v.values():
        for a in ss:
            for b in ss:
                if a != b:
                    pairs.add((a, b))
    ref = np.zeros(cs.size, dtype=np.int32)
    for a, _ in pairs:
        ref[a] += 1
    assert np.array_equal(got, ref)


def test_neighbors_report_16cube():
    vtx, xyz, nv = M.box_hex(16, 16, 16, jitter=0.1, seed=1)
    pa.part_mesh(vtx, xyz, nv, 8)
    r = pa.last_report()
    assert r["nbrs_min"] == r["nbrs_max"] == [1, 3, 7]
    assert r["nbrs_mean"] == [1.0, 3.0, 7.0]


def test_part_mesh_seq_is_order_within_part(oracle):
    """parrsb_part_mesh_seq: seq numbers each part's elements 0..size-1; with
    the RCB partitioner it is the position inside the rank's range of the
    RCB order (oracle), and RSB gives the same part[] as parrsb_part_mesh."""
    vtx, xyz, nv = M.box_hex(10, 9, 8, seed=2)
    n = vtx.shape[0]
    P = 6
    part, seq, _ = pa.part_mesh_seq(vtx, xyz, nv, P, partitioner=1)
    order = oracle.rcb_order(oracle.centroids(xyz, nv), 3, np.arange(n, dtype=np.int32), P)
    ref = np.zeros(n, dtype=np.int32)

    def ranges(start, length, p):
        if p == 1:
            ref[order[start:start + length]] = np.arange(length)
            return
        l = oracle.lower_count(length, p)
        ranges(start, l, (p + 1) // 2)
        ranges(start + l, length - l, p // 2)
    ranges(0, n, P)
    assert np.array_equal(seq, ref)
    part2, seq2, _ = pa.part_mesh_seq(vtx, xyz, nv, P)
    part3, _ = pa.part_mesh(vtx, xyz, nv, P)
    assert np.array_equal(part2, part3)
    for q in range(P):
        assert np.array_equal(np.sort(seq2[part2 == q]), np.arange(np.count_nonzero(part2 == q)))


def test_lanczos_graph_chunks_equal_eager_launches(monkeypatch):
    """PARRSB_AMD_LZ_GRAPH=1 replays eight CG-Lanczos iterations as one
    captured graph (captured again when the running lists change); every
    replay feeds the next and eager launches run between them (restarts,
    Ritz, compaction), so equal partitions and iteration counts over a whole
    call check all of them."""
    vtx, xyz, nv = M.box_hex(24, 20, 16, jitter=0.1, seed=5)
    pa.set_precision(32)
    try:
        monkeypatch.setenv("PARRSB_AMD_LZ_GRAPH", "1")
        a, _ = pa.part_mesh(vtx, xyz, nv, 16)
        ra = pa.last_report()
        monkeypatch.setenv("PARRSB_AMD_LZ_GRAPH", "0")
        b, _ = pa.part_mesh(vtx, xyz, nv, 16)
        rb = pa.last_report()
    finally:
        pa.set_precision(64)
    assert np.array_equal(a, b)
    assert ra["niter_max"] == rb["niter_max"] and ra["lambda_min"] == rb["lambda_min"]
