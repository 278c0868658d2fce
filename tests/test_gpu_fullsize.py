"""The benchmark workloads at their full sizes -- needs an MI355X.

  - M2M and L2L, FP64, P = 16, 1,000,000 translations each (octant shifts);
  - M2L at P = 10, 16 and 24, 200,000 translations each, in FP64, FP32 and
    double-word FP32;
  - the serving workload: the MAC interaction lists of a 1,000,000-point
    Plummer octree cut into requests, through the native request loop.

Whole batches are checked through size-independent properties (linearity
over the batch, every value finite, alone equal to the mix bit for bit) and
sampled items against the binary128 truth (oracle/truth.c) or the FP64
oracle.
"""
import numpy as np
import pytest

import oracle
from solidfmm_amd import workload as W
from tests.helpers import FLOOR

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def sfmm():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    import solidfmm_amd
    solidfmm_amd.load()
    return solidfmm_amd


def sample_truth(op, items, src, got, P, k=2000):
    """Errors of `got` on k evenly spaced items against the truth: per item
    (max over degrees) relative error and error in rounding units."""
    n = len(items)
    idx = np.linspace(0, n - 1, k).astype(np.int64)
    sub = items[idx].copy()
    blk = P * (P + 1)
    sub["dst"] = np.arange(k, dtype=np.uint64) * np.uint64(blk)
    hi, lo, mag = oracle.Truth().translate(op, sub, src, k * blk, magnitude=True)
    g = got.reshape(n, -1)[idx].reshape(-1)
    ed = oracle.truth_degree_errors(g, hi, lo, P)
    s = oracle.error_scale(mag, hi, P)
    units = np.where(s > 0, ed / np.where(s > 0, s, 1.0), 0.0).max(axis=1)
    return ed.max(axis=1), units


def q(e):
    return f"p50 {np.percentile(e, 50):.2e} p99 {np.percentile(e, 99):.2e} max {e.max():.2e}"


_P16 = {}


def p16_batch():
    """1,000,000 P=16 sources with octant shifts (the same for M2M and L2L:
    the generator's streams do not depend on the operator)."""
    if not _P16:
        items, src, _ = W.uniform_batch(W.M2M, 16, 16, 1_000_000, 0x5011DF00, shift="octant")
        _P16["v"] = (items, src)
    return _P16["v"]


@pytest.mark.parametrize("op", [W.M2M, W.L2L])
def test_m2m_l2l_p16_full_size(sfmm, op):
    """1,000,000 translations at P=16: linearity T(2a - b) = 2 T(a) - T(b)
    over the whole batch (b: the sources shifted by one block), and 2,000
    sampled items against the truth at the FP64 floor of order 16."""
    n, P = 1_000_000, 16
    items, src = p16_batch()
    blk = P * (P + 1)
    print(f"op={op} n={n} P={P}", flush=True)
    ctx = sfmm.Context(P)
    plan = sfmm.Plan(ctx, op, items)
    assert plan.launches == 1
    src_b = np.roll(src, blk)
    outs = []
    for s in (src, src_b, 2 * src - src_b):
        d = torch.zeros(n * blk, dtype=torch.float64, device="cuda")
        plan.execute(torch.from_numpy(s).cuda(), d)
        torch.cuda.synchronize()
        outs.append(d.cpu().numpy())
        del d
    a, b, c = outs
    assert np.isfinite(a).all() and np.isfinite(c).all()

    def norms(v):
        return oracle.degree_errors(v, np.zeros_like(v), P)
    lin = (norms(c - (2 * a - b)) / (2 * norms(a) + norms(b))).max(axis=1)
    print(f"  linearity {q(lin)}", flush=True)
    assert np.percentile(lin, 99) <= 2 * FLOOR[P] and lin.max() <= 10 * FLOOR[P]
    e, units = sample_truth(op, items, src, a, P)
    print(f"  against truth {q(e)} units max {units.max():.1f}", flush=True)
    # reference on 2,000 items of this workload: p99 6.0e-13 (M2M) and
    # 7.9e-13 (L2L), at most 3.9 rounding units (tests/golden/truth_errors.npz)
    assert np.percentile(e, 99) <= FLOOR[P] and units.max() <= 8.0


@pytest.mark.parametrize("P", [10, 16, 24])
def test_precision_study_full_size(sfmm, P):
    """200,000 M2L translations per order in the three arithmetics.  On
    2,000 sampled items against the truth: FP64 at its floor; double-word
    FP32 (about 48 bits) within 64x the FP64 error (twice the ratio of the
    unit roundoffs, 2^-48 / 2^-53); FP32 within 2^30 x (likewise)."""
    n = 200_000
    items, src, _ = W.uniform_batch(W.M2L, P, P, n, 0x5011DF03 + P)
    blk = P * (P + 1)
    s = torch.from_numpy(src).cuda()
    got = {}
    for name, prec in (("fp64", 0), ("fp32", 1), ("split32", 2)):
        print(f"P={P} precision={name}", flush=True)
        ctx = sfmm.Context(P, precision=prec)
        plan = sfmm.Plan(ctx, W.M2L, items)
        d = torch.zeros(n * blk, dtype=torch.float64, device="cuda")
        plan.execute(s, d)
        torch.cuda.synchronize()
        got[name] = d.cpu().numpy()
        assert np.isfinite(got[name]).all()
    p99 = {}
    for name, v in got.items():
        e, units = sample_truth(W.M2L, items, src, v, P)
        p99[name] = float(np.percentile(e, 99))
        print(f"  {name}: {q(e)} units max {units.max():.1f}", flush=True)
    assert p99["fp64"] <= FLOOR[P]
    assert p99["split32"] <= 64 * p99["fp64"]
    assert p99["fp32"] <= 2.0 ** 30 * p99["fp64"]


def test_serving_full_size(sfmm, orc):
    """The serving workload: MAC lists of a 1,000,000-point Plummer octree
    (orders 20 at the root to 6 at the leaves, 11M translations) cut into
    requests of 1k-50k translations.  The native loop (6 slots, 1 GiB arena)
    leaves the pools equal bit for bit to running every request alone in
    arrival order; the M2L results of 16 sampled target cells match the
    oracle on the same lists."""
    from solidfmm_amd import fmm as F
    from solidfmm_amd import octree as O
    from solidfmm_amd import serving as S
    pts = O.plummer(1_000_000, 5)
    tree = O.Octree(pts, leaf_size=64)
    P, off, pool_len, passes = F.mac_translation_lists(tree, pts, 0.6, 20, 6)
    m0 = W.uniform(5, 7, pool_len, -1.0, 1.0)
    for c in range(len(tree)):
        p = int(P[c])
        m0[int(off[c]) + (np.arange(p) + 1) ** 2] = 0.0
    reqs = S.make_requests(passes, 5, lo=1000, hi=50000)
    n_tr = sum(len(r.items) for r in reqs)
    print(f"cells={len(tree)} requests={len(reqs)} translations={n_tr}", flush=True)
    assert n_tr > 10_000_000
    ctx = sfmm.Context(20)

    def pools():
        return (torch.from_numpy(m0.copy()).cuda(), torch.zeros(pool_len, dtype=torch.float64, device="cuda"))
    alone = pools()
    S.run_sequential(ctx, alone, reqs)
    torch.cuda.synchronize()
    mix = pools()
    server = S.NativeServer(ctx, mix, len(tree), slots=6, budget_bytes=1 << 30)
    stats = server.run(reqs, realtime=False)
    print(f"  mix: max in flight {stats['max_inflight']}, {stats['translations_per_s']:.3g} translations/s", flush=True)
    assert stats["max_inflight"] >= 2
    for a, b in zip(alone, mix):
        assert torch.equal(a, b)
    # The M2L pass alone on the final multipoles (nothing writes them after
    # the upward pass) into zeroed locals: 16 sampled target cells, each the
    # sum of its whole interaction list, against the oracle on the same
    # items, bounded in rounding units of the summed translations.
    from tests.helpers import scale_limit, scaled_differences
    m_final = alone[0]
    m2l_reqs = [r for r in reqs if r.op == W.M2L]
    only = (m_final.clone(), torch.zeros(pool_len, dtype=torch.float64, device="cuda"))
    S.run_sequential(ctx, only, m2l_reqs)
    torch.cuda.synchronize()
    got_l = only[1].cpu().numpy()
    m_host = m_final.cpu().numpy()
    items = np.concatenate([r.items for r in m2l_reqs])
    rng = np.random.default_rng(3)
    targets = rng.choice(np.unique(items["dst"]), 16, replace=False)
    sel = items[np.isin(items["dst"], targets)]
    want = orc.translate(W.M2L, sel, m_host, np.zeros(pool_len))
    r = scaled_differences(got_l, want, W.M2L, sel, m_host)
    print(f"  M2L of {len(sel)} items into 16 cells: at most {r.max():.1f} rounding units", flush=True)
    assert r.max() <= scale_limit(20, 20)
