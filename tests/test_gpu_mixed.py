"""Mixed-order lists (items of different p_in, p_out in one plan), HIP
graphs and the autotuner -- needs an MI355X.  Calls go through the C ABI.

The oracle translates each item at its own orders; the GPU runs each order
bucket max(p_in, p_out) in the per-item-order variant of that bucket's
kernel.  Tolerances: each item against tol_fp64 of its own orders.
"""
import numpy as np
import pytest

from solidfmm_amd import workload as W
from tests.helpers import assert_mixed_parity

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

AUTO, LANE, MFMA, GENERIC = 0, 1, 2, 3


@pytest.fixture(scope="module")
def sfmm():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    import solidfmm_amd
    solidfmm_amd.load()
    return solidfmm_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run(sfmm, op, items, src, dst0, P, family=AUTO):
    ctx = sfmm.Context(P)
    ctx.set_kernels(family)
    plan = sfmm.Plan(ctx, op, items)
    s, d = dev(src), dev(dst0)
    plan.execute(s, d)
    torch.cuda.synchronize()
    return d.cpu().numpy(), plan


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("family,pmin,pmax", [(LANE, 1, 12), (MFMA, 1, 24), (GENERIC, 1, 24), (AUTO, 4, 20),
                                              (AUTO, 1, 30)])
def test_mixed_orders(sfmm, orc, op, family, pmin, pmax):
    items, src, dst = W.mixed_batch(700, 0x31337 + 97 * op + pmax + family, pmin, pmax, op=op)
    dst0 = W.uniform(0x777, op, len(dst), -1.0, 1.0)
    print(f"op={op} family={family} orders {pmin}..{pmax}", flush=True)
    got, plan = run(sfmm, op, items, src, dst0, pmax, family)
    info = plan.launch_info()
    assert sum(c for *_, c in info) == len(items)
    if family in (LANE, MFMA):
        assert all(f == family for f, *_ in info)
        assert any(rt for _, _, rt, _ in info)
    want = orc.translate(op, items, src, dst0.copy())
    assert_mixed_parity(got, want, items, label=f"op={op} family={family}", case=(op, src, dst0))


def test_zero_shift_mixed(sfmm, orc):
    for op in (W.M2M, W.L2L):
        items, src, dst = W.mixed_batch(300, 1234 + op, 1, 16, op=op)
        items["x"][::4] = 0.0
        items["y"][::4] = 0.0
        items["z"][::4] = 0.0
        dst0 = W.uniform(0x778, op, len(dst), -1.0, 1.0)
        for family in (LANE, MFMA, GENERIC):
            pmax = 12 if family == LANE else 16
            sel = np.maximum(items["p_in"], items["p_out"]) <= pmax
            got, _ = run(sfmm, op, items[sel], src, dst0, pmax, family)
            want = orc.translate(op, items[sel], src, dst0.copy())
            assert_mixed_parity(got, want, items[sel], label=f"op={op} family={family}", case=(op, src, dst0))


def test_graph_replay_matches_execute(sfmm):
    items, src, dst = W.mixed_batch(5000, 4242, 4, 20)
    ctx = sfmm.Context(20)
    plan = sfmm.Plan(ctx, W.M2L, items)
    s = dev(src)
    de = dev(dst)
    eager = []
    for _ in range(3):
        plan.execute(s, de)
        torch.cuda.synchronize()
        eager.append(de.cpu().numpy())
    dg = dev(dst)
    plan.capture(s, dg)
    torch.cuda.synchronize()
    assert float(dg.abs().max()) == 0.0  # capturing runs nothing
    for k in range(3):
        plan.replay()
        torch.cuda.synchronize()
        assert np.array_equal(dg.cpu().numpy(), eager[k]), f"replay {k}"
        x = torch.randn(1 << 16, device="cuda")  # unrelated eager work between replays
        float(x.sum())
        torch.cuda.synchronize()


def test_autotune_keeps_results(sfmm, orc):
    items, src, dst = W.mixed_batch(3000, 777, 4, 20)
    ctx = sfmm.Context(20)
    ctx.autotune(W.M2L, items)
    tuning = ctx.tuning(W.M2L)
    for pb in range(4, 21):
        assert tuning[pb] in (LANE, MFMA, GENERIC)
        if tuning[pb] == LANE:
            assert pb <= 12
        if tuning[pb] == MFMA:
            assert pb <= 24
    plan = sfmm.Plan(ctx, W.M2L, items)
    for fam, order, rt, count in plan.launch_info():
        if fam != GENERIC:
            assert fam == tuning[order]
    s, d = dev(src), dev(dst)
    plan.execute(s, d)
    torch.cuda.synchronize()
    want = orc.translate(W.M2L, items, src, dst.copy())
    assert_mixed_parity(d.cpu().numpy(), want, items, label="autotuned", case=(W.M2L, src, dst))
    ctx.set_tuning(W.M2L, 10, GENERIC)
    assert ctx.tuning(W.M2L)[10] == GENERIC
    with pytest.raises(sfmm.UnsupportedError):
        ctx.set_tuning(W.M2L, 16, LANE)


def test_mixed_full_size_sample(sfmm, orc):
    """The 500,000-item mixed M2L workload (orders 4..20): a sample of 1,500
    items against the oracle; one launch per order bucket."""
    n = 500_000
    items, src, dst = W.mixed_batch(n, 0x5011DF02, 4, 20)
    ctx = sfmm.Context(20)
    plan = sfmm.Plan(ctx, W.M2L, items)
    s = dev(src)
    d = dev(dst)
    plan.execute(s, d)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    idx = np.linspace(0, n - 1, 1500).astype(np.int64)
    sub = items[idx]
    want = orc.translate(W.M2L, sub, src, np.zeros_like(dst))
    # compare only the sampled blocks
    assert_mixed_parity(got, want, sub, label="sample", case=(W.M2L, src, None))
    assert np.isfinite(got).all()
    assert plan.launches <= 17


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
def test_heavily_repeated_targets(sfmm, orc, op):
    """Targets hit hundreds of times (as in FMM interaction lists): per-item
    scratch blocks and a segmented reduction in list order."""
    import oracle
    n, nt, P = 3000, 9, 14
    items, src, _ = W.mixed_batch(n, 0xAB + op, 4, P, op=op)
    items["p_out"] = P  # one order per target solid
    blk = P * (P + 1)
    tgt = (W.uniform01(0x55, op, n) * nt).astype(np.int64)
    items["dst"] = (tgt * blk).astype(np.uint64)
    dst0 = W.uniform(0x66, op, nt * blk, -1.0, 1.0)
    got, plan = run(sfmm, op, items, src, dst0, P)
    want = orc.translate(op, items, src, dst0.copy())
    # each target sums ~330 contributions, added in another order than the
    # oracle's one-by-one accumulation
    e = oracle.degree_errors(got, want, P).max(axis=1)
    assert e.max() <= 1e-11, e.max()
    again, _ = run(sfmm, op, items, src, dst0, P)
    assert np.array_equal(again, got)  # bitwise reproducible


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
def test_autotune_on_repeated_targets(sfmm, orc, op):
    """The autotuner times the plan each family would get, with the list's
    repeated targets (scratch blocks, runs, reduction), on scratch pools
    sized per bucket: one bucket of 3,000 items into 7 targets of order 12
    beside 24 independent translations of order 20.  The tuned plan stays
    within tolerance of the oracle and uses the tuned families."""
    import oracle
    from tests.helpers import assert_parity
    n, nt, P = 3000, 7, 12
    a, src_a, _ = W.mixed_batch(n, 0xAC + op, 4, P, op=op)
    a["p_out"] = P
    blk = P * (P + 1)
    a["dst"] = ((W.uniform01(0x57, op, n) * nt).astype(np.int64) * blk).astype(np.uint64)
    shift = "m2l" if op == W.M2L else "octant"
    b, src_b, _ = W.uniform_batch(op, 20, 20, 24, 0xAD + op, shift=shift)
    b["src"] += np.uint64(len(src_a))
    b["dst"] += np.uint64(nt * blk)
    items = np.concatenate([a, b])
    src = np.concatenate([src_a, src_b])
    dst0 = W.uniform(0x67, op, nt * blk + 24 * 420, -1.0, 1.0)
    ctx = sfmm.Context(20)
    ctx.autotune(op, items)
    tuning = ctx.tuning(op)
    plan = sfmm.Plan(ctx, op, items)
    for fam, order, rt, count in plan.launch_info():
        if fam != GENERIC:  # generic launches mix buckets
            assert fam == tuning[order]
    d = dev(dst0)
    plan.execute(dev(src), d)
    torch.cuda.synchronize()
    got = d.cpu().numpy()
    want = orc.translate(op, items, src, dst0.copy())
    e = oracle.degree_errors(got[:nt * blk], want[:nt * blk], P).max(axis=1)
    assert e.max() <= 1e-11, e.max()
    assert_parity(got[nt * blk:], want[nt * blk:], 20, 20, label=f"op={op} order 20")


def test_prepared_plan_matches_created(sfmm):
    items, src, _ = W.mixed_batch(4000, 0x99, 4, 16)
    items["dst"] = (np.arange(4000) % 50 * 272).astype(np.uint64)  # 80 items per target
    ctx = sfmm.Context(16)
    s = dev(src)
    a = torch.zeros(50 * 272, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a)
    sfmm.Plan(ctx, W.M2L, items).execute(s, a)
    p = sfmm.Plan(ctx, W.M2L, items, prepared=True)
    assert p.device_bytes > 4000 * 48  # reduction tables and scratch included
    buf = torch.empty(p.device_bytes, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream()
    p.bind(buf.data_ptr(), p.device_bytes, stream)
    p.execute(s, b, stream)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
