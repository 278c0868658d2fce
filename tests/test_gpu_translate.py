"""Parity of the HIP translation kernels with the oracle and the golden
vectors of the compiled reference -- needs an MI355X.

Every call goes through the C ABI (libsolidfmm_gpu.so).  Tolerances are the
per-degree normwise relative bounds of tests/helpers.tol_fp64.
"""
import numpy as np
import pytest

import oracle
from solidfmm_amd import workload as W
from tests.helpers import assert_parity, load_golden, orders, tol_fp64

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _has_gpu():
    return torch.cuda.is_available()


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not _has_gpu():
        pytest.fail("GPU tests need a visible GPU")


@pytest.fixture(scope="module")
def sfmm():
    import solidfmm_amd
    solidfmm_amd.load()
    return solidfmm_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run_gpu(sfmm, op, items, src, dst0, P=None, family=0):
    P = P or max(int(items["p_in"].max(initial=1)), int(items["p_out"].max(initial=1)))
    ctx = sfmm.Context(P)
    ctx.set_kernels(family)
    plan = sfmm.Plan(ctx, op, items)
    s, d = dev(src), dev(dst0)
    plan.execute(s, d)
    torch.cuda.synchronize()
    return d.cpu().numpy()


GOLDEN = load_golden()
LANE_MAX = 12   # one lane per translation
MFMA_MAX = 24   # 16 translations per wave, swaps on the matrix cores
AUTO, LANE, MFMA, GENERIC = 0, 1, 2, 3


def _kernel_available(items):
    pi, po = orders(items)
    return max(pi, po) <= 86


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_golden(sfmm, case):
    op, items, src, dst0, dst = GOLDEN[case]
    if not _kernel_available(items):
        pytest.skip("orders without a kernel in this library")
    print(f"case={case}", flush=True)
    got = run_gpu(sfmm, op, items, src, dst0)
    pi, po = orders(items)
    assert_parity(got, dst, pi, po, label=case, case=(op, items, src, dst0))


SQUARE_CASES = ([(LANE, P) for P in range(1, LANE_MAX + 1)] + [(MFMA, P) for P in range(1, MFMA_MAX + 1)]
                + [(GENERIC, P) for P in (1, 2, 3, 5, 10, 16, 24, 30)])


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_golden_generic(sfmm, case):
    """Every golden case through the generic kernel as well."""
    op, items, src, dst0, dst = GOLDEN[case]
    got = run_gpu(sfmm, op, items, src, dst0, family=GENERIC)
    pi, po = orders(items)
    assert_parity(got, dst, pi, po, label=case, case=(op, items, src, dst0))


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("pi,po", [(4, 20), (20, 4), (13, 7), (7, 13), (2, 17), (17, 16), (30, 25)])
def test_generic_mixed_orders(sfmm, orc, op, pi, po):
    """Order pairs the specialised kernels do not cover (random inputs).
    L2L keeps p_in <= p_out + 1 (the reference's own range, see the oracle)."""
    if op == W.L2L and pi > po + 1:
        pytest.skip("l2l truncation covered by the oracle against the naive formula")
    shift = "m2l" if op == W.M2L else "octant"
    items, src, dst = W.uniform_batch(op, pi, po, 40, 0xBEEF + 7 * pi + po + op, shift=shift)
    dst0 = W.random_solids(9, 4, 40, po).reshape(-1)
    got = run_gpu(sfmm, op, items, src, dst0, family=AUTO)
    want = orc.translate(op, items, src, dst0.copy())
    assert_parity(got, want, pi, po, label=f"{op} {pi}->{po}", case=(op, items, src, dst0))


def physical_batch(orc, op, pi, po, n, seed):
    """Expansions of point charges (decaying coefficients): multipoles of
    charges within |s| < 0.87 for M2M / M2L, local expansions of a charge
    about 4 away for L2L; M2L shifts |x| in [3,4], M2M / L2L octant shifts."""
    rng = np.random.default_rng(seed)
    src = np.zeros(n * pi * (pi + 1))
    for i in range(n):
        p = rng.uniform(-0.5, 0.5, 3)
        blk = (orc.harmonic("S", pi, *(np.array([3.0, 1.0, -2.0]) + p)) if op == W.L2L
               else orc.harmonic("R", pi, *p))
        src[i * pi * (pi + 1):(i + 1) * pi * (pi + 1)] = blk
    sh = W.random_shifts(seed, 10, n, 3.0, 4.0) if op == W.M2L else W.octant_shifts(seed, 20, n)
    items = W.make_items(np.arange(n, dtype=np.uint64) * np.uint64(pi * (pi + 1)),
                         np.arange(n, dtype=np.uint64) * np.uint64(po * (po + 1)), pi, po, sh)
    return items, src, np.zeros(n * po * (po + 1))


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("pi,po", [(40, 40), (60, 30), (86, 86)])
def test_generic_high_orders(sfmm, orc, op, pi, po):
    """Orders up to the FP64 limit 86 on physical expansions.  At these
    orders the highest degrees carry few correct digits in any FP64
    evaluation, so the bound is 20x the oracle's own distance from the naive
    O(P^4) formulas on the same inputs (or tol_fp64, whichever is larger)."""
    if op == W.L2L and pi > po + 1:
        pytest.skip("l2l truncation covered by the oracle against the naive formula")
    items, src, dst = physical_batch(orc, op, pi, po, 6, 31 + pi + op)
    got = run_gpu(sfmm, op, items, src, dst, family=AUTO)
    want = orc.translate(op, items, src, dst.copy())
    naive = orc.translate(op, items, src, dst.copy(), naive=True)
    spread = float(oracle.degree_errors(want, naive, po).max())
    assert_parity(got, want, pi, po, tol=max(tol_fp64(pi, po), 20 * spread), label=f"{op} {pi}->{po}")


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("family,P", SQUARE_CASES)
def test_oracle_parity_square(sfmm, orc, op, family, P):
    n = 200  # several waves and a partial one
    shift = "m2l" if op == W.M2L else "octant"
    items, src, dst = W.uniform_batch(op, P, P, n, 0xA5A5 + 17 * P + op, shift=shift)
    dst0 = W.random_solids(5, 3, n, P).reshape(-1)
    print(f"op={op} family={family} P={P}", flush=True)
    got = run_gpu(sfmm, op, items, src, dst0, family=family)
    want = orc.translate(op, items, src, dst0.copy())
    assert_parity(got, want, P, P, label=f"op={op} family={family} P={P}", case=(op, items, src, dst0))


def test_repeated_targets_apply_in_order(sfmm, orc):
    # 40 items over 3 targets, shuffled, plus zero-order items that do nothing.
    items, src, _ = W.uniform_batch(W.M2L, 10, 10, 40, 11, distinct_targets=False, n_targets=3)
    perm = np.argsort(W.uniform01(11, 5, 40))
    items["dst"] = items["dst"][perm]
    dst0 = W.random_solids(11, 6, 3, 10).reshape(-1)
    got = run_gpu(sfmm, W.M2L, items, src, dst0)
    want = orc.translate(W.M2L, items, src, dst0.copy())
    # Each target sums ~13 contributions whose cancellation is counted in
    # the per-target relative error; a lost or doubled update is O(1).
    assert_parity(got, want, 10, 10, tol=1e-12, case=(W.M2L, items, src, dst0))


@pytest.mark.parametrize("family", [LANE, MFMA])
def test_zero_shift_and_mixed_lanes(sfmm, orc, family):
    for op in (W.M2M, W.L2L):
        items, src, dst = W.uniform_batch(op, 8, 8, 130, 12, shift="octant")
        items["x"][::3] = 0.0
        items["y"][::3] = 0.0
        items["z"][::3] = 0.0
        dst0 = W.random_solids(12, 7, 130, 8).reshape(-1)
        got = run_gpu(sfmm, op, items, src, dst0, family=family)
        want = orc.translate(op, items, src, dst0.copy())
        assert_parity(got, want, 8, 8, case=(op, items, src, dst0))


def test_validation_writes_nothing(sfmm):
    items, src, dst = W.uniform_batch(W.M2L, 6, 6, 10, 13)
    ctx = sfmm.Context(6)
    bad = items.copy()
    bad["p_in"][7] = 7
    with pytest.raises(sfmm.OutOfRangeError):
        sfmm.Plan(ctx, W.M2L, bad)
    odd = items.copy()
    odd["src"][3] += 1
    with pytest.raises(sfmm.LogicError):
        sfmm.Plan(ctx, W.M2L, odd)
    plan = sfmm.Plan(ctx, W.M2L, items)
    d = dev(np.zeros(len(dst) - 2))
    with pytest.raises(sfmm.LogicError):
        plan.execute(dev(src), d)
    torch.cuda.synchronize()
    assert float(d.abs().max()) == 0.0


def test_context_errors(sfmm):
    with pytest.raises(sfmm.LogicError):
        sfmm.Context(0)
    with pytest.raises(sfmm.OrderOverflowError):
        sfmm.Context(10 ** 6)


def test_empty_and_zero_order_lists(sfmm):
    ctx = sfmm.Context(4)
    empty = np.zeros(0, dtype=W.ITEM_DTYPE)
    plan = sfmm.Plan(ctx, W.M2L, empty)
    plan.execute(dev(np.zeros(2)), dev(np.zeros(2)))
    items, src, dst = W.uniform_batch(W.M2L, 4, 4, 5, 14)
    items["p_in"] = 0
    plan = sfmm.Plan(ctx, W.M2L, items)
    assert plan.launches == 0


def test_deterministic(sfmm):
    items, src, dst = W.uniform_batch(W.M2L, 10, 10, 5000, 15)
    a = run_gpu(sfmm, W.M2L, items, src, dst)
    b = run_gpu(sfmm, W.M2L, items, src, dst)
    assert np.array_equal(a, b)


def test_full_size_m2l_p10(sfmm, orc):
    """The 100,000-translation P=10 workload: a 2,000-item sample against the
    oracle, plus linearity T(2a - b) = 2 T(a) - T(b) over the whole batch."""
    n = 100_000
    items, src, dst = W.uniform_batch(W.M2L, 10, 10, n, 0x5011DF00)
    src2 = W.random_solids(0x5011DF00, 50, n, 10).reshape(-1)
    ctx = sfmm.Context(10)
    plan = sfmm.Plan(ctx, W.M2L, items)
    assert plan.launches == 1
    outs = []
    for s in (src, src2, 2 * src - src2):
        d = dev(np.zeros_like(dst))
        plan.execute(dev(s), d)
        outs.append(d.cpu().numpy())
    a, b, c = outs
    # Scale of the operands, not of their difference, which can cancel.
    def norms(v):  # per-degree 2-norms (degree_errors against zero)
        return oracle.degree_errors(v, np.zeros_like(v), 10)
    lin = (norms(c - (2 * a - b)) / (2 * norms(a) + norms(b))).max(axis=1)
    # internal cancellation inside single translations leaves outliers
    assert np.percentile(lin, 99) <= 1e-13 and lin.max() <= 1e-12, (np.percentile(lin, 99), lin.max())
    idx = np.linspace(0, n - 1, 2000).astype(np.int64)
    sub = items[idx].copy()
    sub["dst"] = np.arange(len(idx), dtype=np.uint64) * np.uint64(110)
    want = orc.translate(W.M2L, sub, src, np.zeros(len(idx) * 110))
    got = a.reshape(n, 110)[idx]
    assert_parity(got, want, 10, 10, label="sample", case=(W.M2L, sub, src, None))


GOLDEN32 = load_golden("golden_f32.npz")


@pytest.mark.parametrize("case", sorted(GOLDEN32))
def test_golden_float_path(sfmm, orc, case):
    """The reference's float outputs (tests/golden/golden_f32.npz).  Float
    solids take the drop-in's float path: widened to FP64, translated on the
    GPU, rounded to float once.  Against the FP64 result on the same float
    inputs, every value is within one float rounding plus FP64 noise, and
    the per-degree errors stay within the reference float path's own."""
    op, items, src32, dst0_32, dst32 = GOLDEN32[case]
    pi, po = orders(items)
    src, dst0 = src32.astype(np.float64), dst0_32.astype(np.float64)
    exact = orc.translate(op, items, src, dst0.copy())
    got = run_gpu(sfmm, op, items, src, dst0).astype(np.float32).astype(np.float64)
    scale = np.zeros_like(exact)
    P = po
    blk = P * (P + 1)
    for b in range(len(exact) // blk):
        for n in range(P):
            sl = slice(b * blk + n * (n + 1), b * blk + (n + 1) * (n + 2))
            scale[sl] = np.abs(exact[sl]).max()
    bound = 2.0 ** -24 * np.abs(exact) + 1e-12 * scale
    assert np.all(np.abs(got - exact) <= bound), float((np.abs(got - exact) / bound).max())
    e = oracle.degree_errors(got, exact, P).max(axis=1)
    e_ref = oracle.degree_errors(dst32.astype(np.float64), exact, P).max(axis=1)
    print(f"  {case}: float path max {e.max():.2e}, reference float max {e_ref.max():.2e}", flush=True)
    assert np.median(e) <= np.median(e_ref) and e.max() <= e_ref.max()
