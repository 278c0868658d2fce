"""Parity of the HIP P2M and L2P kernels (particles.hip) with the oracle's
harmonics R and dR and dot product -- needs an MI355X.

P2M of a cell is sum_i q_i R(P, p_i - c) (the M of tests/test_m2l.cpp:70
summed over charges); L2P is dot(L, R(P, p - c)) and dot(L, dR(P, p - c))
(tests/test_m2l.cpp:112-116).  All calls go through the C ABI.
"""
import math

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")


@pytest.fixture(scope="module")
def sfmm():
    import solidfmm_amd
    solidfmm_amd.load()
    return solidfmm_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def make_cells(sfmm, rng, counts, orders, spread=0.5):
    """Cells with centres in [-2, 2]^3, points within `spread` of the centre,
    blocks packed in order with a gap of 2 reals between them."""
    n = int(sum(counts))
    cells = np.zeros(len(counts), dtype=sfmm.CELL_DTYPE)
    pts = np.zeros((n, 4))
    first = 0
    off = 0
    for c, (k, P) in enumerate(zip(counts, orders)):
        ctr = rng.uniform(-2, 2, 3)
        cells[c] = (ctr[0], ctr[1], ctr[2], off, first, k, P, 0)
        pts[first:first + k, :3] = ctr + rng.uniform(-spread, spread, (k, 3))
        pts[first:first + k, 3] = rng.uniform(-1, 1, k)
        first += k
        off += P * (P + 1) + 2
    return cells, pts, off


def p2m_oracle(orc, cells, pts, pool):
    for c in cells:
        P = int(c["order"])
        o = int(c["offset"])
        for i in range(int(c["first"]), int(c["first"] + c["count"])):
            x, y, z, q = pts[i]
            pool[o:o + P * (P + 1)] += q * orc.harmonic("R", P, x - c["x"], y - c["y"], z - c["z"])
    return pool


def l2p_oracle(orc, cells, pts, pool, phi, grad):
    for c in cells:
        P = int(c["order"])
        o = int(c["offset"])
        L = pool[o:o + P * (P + 1)]
        for i in range(int(c["first"]), int(c["first"] + c["count"])):
            x, y, z, _ = pts[i]
            d = (x - c["x"], y - c["y"], z - c["z"])
            phi[i] += orc.dot(L, P, orc.harmonic("R", P, *d), P)
            g = orc.harmonic("dR", P, *d).reshape(3, -1)
            for k in range(3):
                grad[i, k] += orc.dot(L, P, np.ascontiguousarray(g[k]), P)
    return phi, grad


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("P", [1, 2, 5, 10, 16, 24])
def test_p2m_matches_oracle(sfmm, P):
    orc = oracle.Oracle()
    rng = np.random.default_rng(100 + P)
    counts = [0, 1, 7, 64, 65, 150]
    cells, pts, plen = make_cells(sfmm, rng, counts, [P] * len(counts))
    pool0 = rng.standard_normal(plen)
    want = p2m_oracle(orc, cells, pts, pool0.copy())
    ctx = sfmm.Context(P)
    cs = sfmm.CellSet(ctx, cells, len(pts))
    pool = dev(pool0)
    cs.p2m(dev(pts), pool)
    torch.cuda.synchronize()
    got = pool.cpu().numpy()
    assert rel(got, want) < 1e-13
    # the 2-real gaps between blocks are untouched
    mask = np.ones(plen, bool)
    for c in cells:
        mask[int(c["offset"]):int(c["offset"]) + P * (P + 1)] = False
    assert np.array_equal(got[mask], pool0[mask])


@pytest.mark.parametrize("P", [1, 2, 5, 10, 16, 24])
def test_l2p_matches_oracle(sfmm, P):
    orc = oracle.Oracle()
    rng = np.random.default_rng(200 + P)
    counts = [0, 1, 7, 64, 65, 150]
    cells, pts, plen = make_cells(sfmm, rng, counts, [P] * len(counts))
    pool = rng.standard_normal(plen)
    phi0 = rng.standard_normal(len(pts))
    grad0 = rng.standard_normal((len(pts), 3))
    phi_w, grad_w = l2p_oracle(orc, cells, pts, pool, phi0.copy(), grad0.copy())
    ctx = sfmm.Context(P)
    cs = sfmm.CellSet(ctx, cells, len(pts))
    phi, grad = dev(phi0), dev(grad0)
    cs.l2p(dev(pts), dev(pool), phi, grad)
    phi_only = dev(phi0)
    cs.l2p(dev(pts), dev(pool), phi_only)
    torch.cuda.synchronize()
    assert rel(phi.cpu().numpy() - phi0, phi_w - phi0) < 1e-12
    assert rel(grad.cpu().numpy() - grad0, grad_w - grad0) < 1e-12
    assert np.array_equal(phi_only.cpu().numpy(), phi.cpu().numpy())


def test_mixed_orders_and_cell_order(sfmm):
    """Cells of orders 1..20 in a shuffled list against one P=20 context."""
    orc = oracle.Oracle()
    rng = np.random.default_rng(7)
    orders = list(range(1, 21))
    counts = [int(k) for k in rng.integers(0, 90, len(orders))]
    cells, pts, plen = make_cells(sfmm, rng, counts, orders)
    cells = cells[rng.permutation(len(cells))]
    pool0 = rng.standard_normal(plen)
    want = p2m_oracle(orc, cells, pts, pool0.copy())
    ctx = sfmm.Context(20)
    cs = sfmm.CellSet(ctx, cells, len(pts))
    pool = dev(pool0)
    P_pts = dev(pts)
    cs.p2m(P_pts, pool)
    phi = torch.zeros(len(pts), dtype=torch.float64, device="cuda")
    grad = torch.zeros((len(pts), 3), dtype=torch.float64, device="cuda")
    cs.l2p(P_pts, pool, phi, grad)
    torch.cuda.synchronize()
    assert rel(pool.cpu().numpy(), want) < 1e-13
    phi_w, grad_w = l2p_oracle(orc, cells, pts, want, np.zeros(len(pts)), np.zeros((len(pts), 3)))
    assert rel(phi.cpu().numpy(), phi_w) < 1e-12
    assert rel(grad.cpu().numpy(), grad_w) < 1e-12


def test_p2m_m2l_l2p_physics(sfmm):
    """Charges in one cell, targets in a well-separated cell: P2M -> M2L -> L2P
    reproduces the direct potential sum q/|x - y| and its gradient
    -q (x - y)/|x - y|^3 to the truncation error of order 16 (cell radii
    0.87, centres 9.2 apart: ((0.87 + 0.87)/9.2)^16 ~ 3e-12)."""
    P = 16
    rng = np.random.default_rng(11)
    ns, nt = 200, 150
    src = np.zeros((ns, 4))
    src[:, :3] = rng.uniform(-0.5, 0.5, (ns, 3))
    src[:, 3] = rng.uniform(0.1, 1.0, ns)
    ctr_t = np.array([8.0, 2.0, -4.0])
    tgt = np.zeros((nt, 4))
    tgt[:, :3] = ctr_t + rng.uniform(-0.5, 0.5, (nt, 3))
    pts = np.vstack([src, tgt])
    n = P * (P + 1)
    cells_m = np.zeros(1, dtype=sfmm.CELL_DTYPE)
    cells_m[0] = (0, 0, 0, 0, 0, ns, P, 0)
    cells_l = np.zeros(1, dtype=sfmm.CELL_DTYPE)
    cells_l[0] = (ctr_t[0], ctr_t[1], ctr_t[2], 0, ns, nt, P, 0)
    ctx = sfmm.Context(P)
    d_pts = dev(pts)
    M = torch.zeros(n, dtype=torch.float64, device="cuda")
    L = torch.zeros(n, dtype=torch.float64, device="cuda")
    sfmm.CellSet(ctx, cells_m, len(pts)).p2m(d_pts, M)
    items = np.zeros(1, dtype=sfmm.ITEM_DTYPE)
    # M2L shift: target centre minus source centre
    items[0] = (0, 0, P, P, ctr_t[0], ctr_t[1], ctr_t[2])
    sfmm.Plan(ctx, sfmm.M2L, items).execute(M, L)
    phi = torch.zeros(len(pts), dtype=torch.float64, device="cuda")
    grad = torch.zeros((len(pts), 3), dtype=torch.float64, device="cuda")
    sfmm.CellSet(ctx, cells_l, len(pts)).l2p(d_pts, L, phi, grad)
    torch.cuda.synchronize()
    d = tgt[:, None, :3] - src[None, :, :3]
    r = np.sqrt((d ** 2).sum(-1))
    phi_w = (src[None, :, 3] / r).sum(1)
    grad_w = -(src[None, :, 3, None] * d / r[..., None] ** 3).sum(1)
    assert rel(phi.cpu().numpy()[ns:], phi_w) < 1e-8
    assert rel(grad.cpu().numpy()[ns:], grad_w) < 1e-7
    assert np.all(phi.cpu().numpy()[:ns] == 0.0)


def test_cellset_validation(sfmm):
    ctx = sfmm.Context(8)
    c = np.zeros(1, dtype=sfmm.CELL_DTYPE)
    c[0] = (0, 0, 0, 0, 0, 4, 9, 0)
    with pytest.raises(sfmm.OutOfRangeError):
        sfmm.CellSet(ctx, c, 4)
    c[0] = (0, 0, 0, 1, 0, 4, 8, 0)
    with pytest.raises(sfmm.LogicError):
        sfmm.CellSet(ctx, c, 4)
    c[0] = (0, 0, 0, 0, 2, 4, 8, 0)
    with pytest.raises(sfmm.LogicError):
        sfmm.CellSet(ctx, c, 5)
    c[0] = (0, 0, 0, 0, 0, 4, 8, 0)
    cs = sfmm.CellSet(ctx, c, 4)
    with pytest.raises(sfmm.LogicError):
        cs.p2m(dev(np.zeros((4, 4))), torch.zeros(71, dtype=torch.float64, device="cuda"))
    with pytest.raises(sfmm.LogicError):
        cs.p2m(dev(np.zeros((3, 4))), torch.zeros(72, dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("D", [2, 3, 4])
def test_vector_valued_p2m_l2p(sfmm, D):
    """D charges per point and D blocks per cell (solid<real>(P, D)): every
    component equals the scalar operators applied to that component."""
    orc = oracle.Oracle()
    P = 9
    rng = np.random.default_rng(300 + D)
    counts = [0, 1, 33, 64, 90]
    n = sum(counts)
    cells = np.zeros(len(counts), dtype=sfmm.CELL_DTYPE)
    pts = np.zeros((n, 3 + D))
    first = off = 0
    for c, k in enumerate(counts):
        ctr = rng.uniform(-2, 2, 3)
        cells[c] = (ctr[0], ctr[1], ctr[2], off, first, k, P, D)
        pts[first:first + k, :3] = ctr + rng.uniform(-0.5, 0.5, (k, 3))
        pts[first:first + k, 3:] = rng.uniform(-1, 1, (k, D))
        first += k
        off += D * P * (P + 1) + 2
    pool0 = rng.standard_normal(off)
    ctx = sfmm.Context(P)
    cs = sfmm.CellSet(ctx, cells, n)
    pool = dev(pool0)
    cs.p2m(dev(pts), pool)
    phi = torch.zeros(n * D, dtype=torch.float64, device="cuda")
    grad = torch.zeros(n * D * 3, dtype=torch.float64, device="cuda")
    cs.l2p(dev(pts), pool, phi, grad)
    torch.cuda.synchronize()
    got_pool = pool.cpu().numpy()
    got_phi = phi.cpu().numpy().reshape(n, D)
    got_grad = grad.cpu().numpy().reshape(n, D, 3)
    blk = P * (P + 1)
    for d in range(D):
        # the scalar oracle on component d, with the cells' d-th blocks
        sc = cells.copy()
        sc["offset"] = cells["offset"] + d * blk
        want_pool = p2m_oracle(orc, sc, np.c_[pts[:, :3], pts[:, 3 + d]], pool0.copy())
        for c in sc:
            o = int(c["offset"])
            assert rel(got_pool[o:o + blk], want_pool[o:o + blk]) < 1e-13
        phi_w, grad_w = l2p_oracle(orc, sc, np.c_[pts[:, :3], pts[:, 3 + d]], want_pool, np.zeros(n),
                                   np.zeros((n, 3)))
        assert rel(got_phi[:, d], phi_w) < 1e-12
        assert rel(got_grad[:, d], grad_w) < 1e-12
    # the gaps between cells are untouched
    mask = np.ones(off, bool)
    for c in cells:
        mask[int(c["offset"]):int(c["offset"]) + D * blk] = False
    assert np.array_equal(got_pool[mask], pool0[mask])


def test_cellset_rejects_mixed_dims(sfmm):
    ctx = sfmm.Context(6)
    c = np.zeros(2, dtype=sfmm.CELL_DTYPE)
    c[0] = (0, 0, 0, 0, 0, 1, 6, 3)
    c[1] = (0, 0, 0, 200, 1, 1, 6, 2)
    with pytest.raises(sfmm.LogicError):
        sfmm.CellSet(ctx, c, 2)
    c[1]["dim"] = 5
    c[0]["dim"] = 5
    with pytest.raises(sfmm.LogicError):
        sfmm.CellSet(ctx, c, 2)
