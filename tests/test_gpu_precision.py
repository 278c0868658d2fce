"""Reduced-precision paths (FP32 and split FP32 = double-word FP32) against
the FP64 oracle -- needs an MI355X.  Calls go through the C ABI.

Error measure: per translation, the largest per-degree normwise relative
error against the oracle.  Random U[-1,1] coefficients make some single
translations badly conditioned (their FP64 error already reaches 1e-7 at
P=16), so the bounds are relative to the FP64 GPU kernel's error on the
same items: the unit roundoff of double-word FP32 (2^-48) is 32x that of
FP64 and the one of FP32 (2^-24) is 2^29 times it, and the tests allow
twice those factors on the 50th and 99th percentiles.  Absolute targets:
split FP32 p99 <= 1e-11 up to P = 12.
"""
import numpy as np
import pytest

import oracle
from solidfmm_amd import workload as W
from tests.helpers import assert_mixed_parity

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FP64, FP32, SPLIT = 0, 1, 2
RATIO = {FP32: 2.0 ** 30, SPLIT: 64.0}


@pytest.fixture(scope="module")
def sfmm():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")
    import solidfmm_amd
    solidfmm_amd.load()
    return solidfmm_amd


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda")


def run(sfmm, op, items, src, dst0, P, precision, family=0):
    ctx = sfmm.Context(P, precision=precision)
    ctx.set_kernels(family)
    plan = sfmm.Plan(ctx, op, items)
    s, d = dev(src), dev(dst0)
    plan.execute(s, d)
    torch.cuda.synchronize()
    return d.cpu().numpy(), plan


def errors(got, want, P):
    return oracle.degree_errors(got, want, P).max(axis=1)


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("P", [3, 8, 10, 12, 16, 24])
def test_precision_against_fp64(sfmm, orc, op, P):
    n = 1500
    shift = "m2l" if op == W.M2L else "octant"
    items, src, dst = W.uniform_batch(op, P, P, n, 0xFACE + 31 * P + op, shift=shift)
    want = orc.translate(op, items, src, dst.copy())
    ref, _ = run(sfmm, op, items, src, dst, P, FP64)
    e64 = errors(ref, want, P)
    for precision in (FP32, SPLIT):
        print(f"op={op} P={P} precision={precision}", flush=True)
        got, plan = run(sfmm, op, items, src, dst, P, precision)
        e = errors(got, want, P)
        assert np.isfinite(got).all()
        for q in (50, 99):
            bound = RATIO[precision] * max(np.percentile(e64, q), 1e-16)
            assert np.percentile(e, q) <= bound, (precision, q, np.percentile(e, q), bound)
        if precision == SPLIT and P <= 12:
            assert np.percentile(e, 99) <= 1e-11


@pytest.mark.parametrize("precision,P", [(FP32, 8), (FP32, 13), (FP32, 16), (SPLIT, 8), (SPLIT, 10)])
def test_lane_and_generic_precision_kernels_agree(sfmm, orc, precision, P):
    """The one-lane-per-translation kernels (FP32 to order 16, double-word
    to 10) against the one-workgroup kernel and the FP64 error."""
    items, src, dst = W.uniform_batch(W.M2L, P, P, 800, 0xBEAD + P)
    a, pa = run(sfmm, W.M2L, items, src, dst, P, precision, family=1)
    b, pb = run(sfmm, W.M2L, items, src, dst, P, precision, family=3)
    assert pa.launch_info()[0][0] == 1 and pb.launch_info()[0][0] == 3
    want = orc.translate(W.M2L, items, src, dst.copy())
    ref, _ = run(sfmm, W.M2L, items, src, dst, P, FP64)
    e64 = np.percentile(errors(ref, want, P), 99)
    for got in (a, b):
        assert np.percentile(errors(got, want, P), 99) <= RATIO[precision] * e64


@pytest.mark.parametrize("P", [1, 5, 8, 13, 16, 17, 20, 24])
def test_fp32_mfma_kernel(sfmm, orc, P):
    """The FP32 M2L kernel on the matrix cores (orders 1..24) against the
    FP64 oracle, at twice the ratio of unit roundoffs times the FP64
    kernel's error on the same items, and against the one-workgroup FP32
    kernel; a second list sends six items to every target (more than the
    plans apply in rounds), which runs the kernel in scratch mode (scratch
    blocks overwritten, then reduced)."""
    n = 1000
    items, src, dst = W.uniform_batch(W.M2L, P, P, n, 0xF32 + P)
    blk = P * (P + 1)
    rep = items.copy()
    rep["dst"] = (np.arange(n, dtype=np.uint64) // 6) * np.uint64(blk)
    for case, its in (("distinct", items), ("repeated", rep)):
        print(f"P={P} targets={case}", flush=True)
        got, plan = run(sfmm, W.M2L, its, src, dst, P, FP32, family=2)
        assert all(f == 2 for f, *_ in plan.launch_info())
        gen, _ = run(sfmm, W.M2L, its, src, dst, P, FP32, family=3)
        want = orc.translate(W.M2L, its, src, dst.copy())
        ref, _ = run(sfmm, W.M2L, its, src, dst, P, FP64)
        assert np.isfinite(got).all()
        e64, e, eg = (errors(v, want, P) for v in (ref, got, gen))
        for q in (50, 99):
            bound = RATIO[FP32] * max(np.percentile(e64, q), 1e-16)
            assert np.percentile(e, q) <= bound, (case, q, np.percentile(e, q), bound)
        # the same arithmetic as the one-workgroup kernel up to the order of
        # the sums: the two stay within a few FP32 roundings of each other
        assert np.percentile(e, 99) <= 4.0 * max(np.percentile(eg, 99), 2.0 ** -24)


@pytest.mark.parametrize("precision", [FP32, SPLIT])
def test_precision_mixed_orders(sfmm, orc, precision):
    items, src, dst = W.mixed_batch(600, 0xD00D + precision, 1, 24)
    dst0 = W.uniform(0x779, precision, len(dst), -1.0, 1.0)
    got, _ = run(sfmm, W.M2L, items, src, dst0, 24, precision)
    want = orc.translate(W.M2L, items, src, dst0.copy())
    assert np.isfinite(got).all()
    if precision == SPLIT:
        # double-word: the FP64 tolerances, 100x
        assert_mixed_parity(got, want, items, scale=100.0, label="split mixed")
    # per item against the FP64 kernel's error on the same items, at twice
    # the ratio of unit roundoffs (50th and 99th percentiles)
    ref, _ = run(sfmm, W.M2L, items, src, dst0, 24, FP64)
    e64 = oracle.mixed_errors(ref, want, items)
    e = oracle.mixed_errors(got, want, items)
    for q in (50, 99):
        bound = RATIO[precision] * max(np.percentile(e64, q), 1e-16)
        assert np.percentile(e, q) <= bound, (precision, q, np.percentile(e, q), bound)


def test_precision_context_limits(sfmm):
    with pytest.raises(sfmm.OrderOverflowError):
        sfmm.Context(25, precision=FP32)
    ctx = sfmm.Context(16, precision=SPLIT)
    ctx.set_kernels(2)  # MFMA kernels: FP32 square M2L only
    items, _, _ = W.uniform_batch(W.M2L, 16, 16, 4, 1)
    with pytest.raises(sfmm.UnsupportedError):
        sfmm.Plan(ctx, W.M2L, items)
    ctx32 = sfmm.Context(16, precision=FP32)
    ctx32.set_kernels(2)
    assert sfmm.Plan(ctx32, W.M2L, items).launch_info()[0][0] == 2
    m2m, _, _ = W.uniform_batch(W.M2M, 16, 16, 4, 1, shift="octant")
    with pytest.raises(sfmm.UnsupportedError):
        sfmm.Plan(ctx32, W.M2M, m2m)
    with pytest.raises(sfmm.LogicError):
        sfmm.Context(4, precision=7)
