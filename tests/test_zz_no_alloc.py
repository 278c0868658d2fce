"""The allocation-free entry point sfmm_translate_noalloc, called through the
C ABI, against sfmm_plan_create + sfmm_plan_execute on the same lists --
needs an MI355X.

Both run the same plan and kernels, so the results must be equal bit for
bit.  Two calls in a row, the second with a smaller list, exercise the
shared scratch array holding the first call's contents; an oversized list
is refused before anything is enqueued.
"""
import ctypes

import numpy as np
import pytest

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


def noalloc(sfmm, ctx, op, items, s, d, stream):
    items = np.ascontiguousarray(items, dtype=W.ITEM_DTYPE)
    return sfmm.load().sfmm_translate_noalloc(ctx.handle, op, items.ctypes.data, len(items), s.data_ptr(),
                                              s.numel(), d.data_ptr(), d.numel(), ctypes.c_void_p(stream))


def with_plan(sfmm, ctx, op, items, src, dst0):
    plan = sfmm.Plan(ctx, op, items)
    s = torch.from_numpy(src).cuda()
    d = torch.from_numpy(dst0.copy()).cuda()
    plan.execute(s, d)
    torch.cuda.synchronize()
    return d.cpu().numpy()


def cases():
    # 1: M2L P=10, 20,000 items into 1,000 targets (each hit 20 times, so
    #    per-item scratch blocks and the reduction run in the array)
    items1, src1, _ = W.uniform_batch(W.M2L, 10, 10, 20000, 0xA110C, distinct_targets=False, n_targets=1000)
    dst1 = W.random_solids(0xA110C, 7, 1000, 10).reshape(-1)
    # 2: smaller: mixed-order M2M (orders 2..8), 3,000 items, distinct targets
    items2, src2, dst2 = W.mixed_batch(3000, 0xA110D, 2, 8, op=W.M2M)
    dst2 = W.uniform(0xA110D, 9, len(dst2), -1.0, 1.0)
    return [(W.M2L, 10, items1, src1, dst1), (W.M2M, 8, items2, src2, dst2)]


def test_noalloc_matches_plan_execute(sfmm):
    stream = torch.cuda.current_stream().cuda_stream
    for k, (op, P, items, src, dst0) in enumerate(cases()):
        print(f"call {k + 1}: op={op} P={P} items={len(items)}", flush=True)
        ctx = sfmm.Context(P)
        s = torch.from_numpy(src).cuda()
        d = torch.from_numpy(dst0.copy()).cuda()
        rc = noalloc(sfmm, ctx, op, items, s, d, stream)
        assert rc == 0, sfmm.load().sfmm_last_error(ctx.handle)
        torch.cuda.synchronize()
        got = d.cpu().numpy()
        want = with_plan(sfmm, ctx, op, items, src, dst0)
        assert np.array_equal(got, want), float(np.abs(got - want).max())
        assert not np.array_equal(got, dst0)


def test_noalloc_refuses_oversized_list(sfmm):
    """100,000 M2L P=10 items into 100 targets need 94 MB of scratch blocks:
    refused with SFMM_ERR_LOGIC, targets untouched."""
    items, src, _ = W.uniform_batch(W.M2L, 10, 10, 100000, 0xA110E, distinct_targets=False, n_targets=100)
    ctx = sfmm.Context(10)
    s = torch.from_numpy(src).cuda()
    d = torch.zeros(100 * 110, dtype=torch.float64, device="cuda")
    rc = noalloc(sfmm, ctx, W.M2L, items, s, d, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == -1
    assert b"SFMM_NOALLOC_SCRATCH_BYTES" in sfmm.load().sfmm_last_error(ctx.handle)
    assert float(d.abs().max()) == 0.0
