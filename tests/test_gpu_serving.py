"""Continuous-batching request loop on the GPU: every request's result is
the same alone and in the mix -- needs an MI355X.

The FMM passes of a small Plummer octree (orders 20 at the root to 6 at
the leaves) are cut into requests.  "Alone": each request runs by itself,
synchronised, in list order.  "Mix": the request loop admits them into 4
concurrent streams under a small item budget, in arrival order, with the
hazard rules of solidfmm_amd.serving.  Both must leave identical pools (the
same kernels see the same updates per cell in the same order), and the
alone run must match the oracle translating the same lists.
"""
import numpy as np
import pytest

from solidfmm_amd import octree as O
from solidfmm_amd import serving as S
from solidfmm_amd import workload as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sfmm():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    import solidfmm_amd
    solidfmm_amd.load()
    return solidfmm_amd


def workload(npts, seed, lo, hi, lists="vlist"):
    pts = O.plummer(npts, seed)
    tree = O.Octree(pts, leaf_size=64)
    if lists == "mac":
        from solidfmm_amd import fmm as F
        P, off, pool_len, passes = F.mac_translation_lists(tree, pts, 0.6, 20, 6)
    else:
        P, off, pool_len, passes = O.translation_lists(tree, 20, 6)
    m0 = W.uniform(seed, 7, pool_len, -1.0, 1.0)
    for c in range(len(tree)):  # Im C_n^0 = 0
        p = int(P[c])
        m0[int(off[c]) + (np.arange(p) + 1) ** 2] = 0.0
    reqs = S.make_requests(passes, seed, lo=lo, hi=hi)
    return tree, pool_len, m0, reqs


def pools_from(m0, pool_len):
    return (torch.from_numpy(m0.copy()).cuda(), torch.zeros(pool_len, dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("precision,kind,lists", [(0, "python", "vlist"), (2, "python", "vlist"),
                                                   (0, "native", "vlist"), (2, "native", "vlist"),
                                                   (0, "native", "mac")])
def test_alone_equals_mix(sfmm, precision, kind, lists):
    tree, pool_len, m0, reqs = workload(20000, 5, 300, 3000, lists)
    ctx = sfmm.Context(20, precision=precision)
    alone = pools_from(m0, pool_len)
    S.run_sequential(ctx, alone, reqs)
    torch.cuda.synchronize()
    mix = pools_from(m0, pool_len)
    budget = 8 << 20  # a few requests at a time (M2L requests carry scratch)
    cls = S.Server if kind == "python" else S.NativeServer
    server = cls(ctx, mix, len(tree), slots=4, budget_bytes=budget)
    stats = server.run(reqs, realtime=False)
    print(f"precision={precision} requests={len(reqs)} stats={stats}", flush=True)
    for a, b in zip(alone, mix):
        assert torch.equal(a, b)
    assert stats["max_inflight"] >= 2
    assert stats["arena_peak_bytes"] <= budget
    assert all(r.done is not None for r in reqs)


def test_alone_matches_oracle(sfmm, orc):
    tree, pool_len, m0, reqs = workload(4000, 6, 200, 1000)
    ctx = sfmm.Context(20)
    pools = pools_from(m0, pool_len)
    S.run_sequential(ctx, pools, reqs)
    got_m, got_l = (p.cpu().numpy() for p in pools)
    m, l = m0.copy(), np.zeros(pool_len)
    for r in reqs:
        src = m if r.op in (W.M2M, W.M2L) else l
        dst = m if r.op == W.M2M else l
        orc.translate(r.op, r.items, src.copy() if src is dst else src, dst)
    P = tree.orders(20, 6)
    off = np.concatenate([[0], np.cumsum(P * (P + 1))[:-1]])
    import oracle
    for got, want in ((got_m, m), (got_l, l)):
        worst = 0.0
        for c in range(len(tree)):
            p = int(P[c])
            blk = slice(int(off[c]), int(off[c]) + p * (p + 1))
            if not np.any(want[blk]):
                continue
            e = oracle.degree_errors(got[blk][None], want[blk][None], p).max()
            worst = max(worst, e)
        # sums over many translations of orders up to 20: FP64 tolerance of
        # order 20 (tests/helpers.tol_fp64) with the 10x max headroom
        assert worst <= 2e-8, worst


@pytest.mark.parametrize("kind", ["python", "native"])
def test_poisson_run_reports(sfmm, kind):
    tree, pool_len, m0, reqs = workload(20000, 8, 300, 3000)
    ctx = sfmm.Context(20)
    S.poisson_arrivals(reqs, 2000.0, 3)
    cls = S.Server if kind == "python" else S.NativeServer
    server = cls(ctx, pools_from(m0, pool_len), len(tree), slots=4)
    st = server.run(reqs, realtime=True)
    assert st["translations"] == sum(len(r.items) for r in reqs)
    assert 0 < st["latency_ms"]["p50"] <= st["latency_ms"]["p99"] <= st["latency_ms"]["max"]
    assert st["translations_per_s"] > 0


def test_native_server_rejects_and_continues(sfmm):
    """A request with orders above the context's, and one that can never fit
    in the arena, finish with their error codes; the others still run and
    match running alone."""
    import ctypes
    tree, pool_len, m0, reqs = workload(8000, 9, 300, 2000)
    ctx = sfmm.Context(20)
    alone = pools_from(m0, pool_len)
    S.run_sequential(ctx, alone, reqs)
    mix = pools_from(m0, pool_len)
    server = S.NativeServer(ctx, mix, len(tree), slots=2, budget_bytes=4 << 20)
    bad = S.Request(10 ** 6, W.M2L, reqs[0].items.copy(), reqs[0].src_cells, reqs[0].dst_cells)
    bad.items["p_in"] = 21
    rid_bad = server.submit(bad)
    big = [r for r in reqs if r.op == W.M2L][0]
    huge = S.Request(10 ** 6 + 1, W.M2L, np.tile(big.items, 400), big.src_cells, big.dst_cells)
    rid_huge = server.submit(huge)
    ok = [(r, server.submit(r)) for r in reqs]
    lib = server._lib
    assert lib.sfmm_server_drain(server._h) == 0
    times = np.zeros(5)
    rc = ctypes.c_int()
    assert lib.sfmm_server_request_info(server._h, rid_bad, times.ctypes.data, ctypes.byref(rc)) == 0
    assert rc.value == sfmm.SFMM_ERR_RANGE if hasattr(sfmm, "SFMM_ERR_RANGE") else rc.value == -2
    assert lib.sfmm_server_request_info(server._h, rid_huge, times.ctypes.data, ctypes.byref(rc)) == 0
    assert rc.value == -4
    for r, rid in ok:
        assert lib.sfmm_server_request_info(server._h, rid, times.ctypes.data, ctypes.byref(rc)) == 0
        assert rc.value == 0 and times[2] >= 0 and times[4] >= times[1]
    server.close()
    for a, b in zip(alone, mix):
        assert torch.equal(a, b)
