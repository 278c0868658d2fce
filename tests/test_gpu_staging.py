"""Host-pool staging (sfmm_staging_*, sfmm_translate_host) and the binds
of prepared plans through pinned memory (sfmm_plan_bind_staged,
sfmm_plan_copy_tables + sfmm_plan_bind_pinned) -- needs an MI355X.  The
staged paths must give bitwise the results of the same plan on device
pools, grow their buffers across calls of different sizes, and reject items
outside the reserved pools and buffers that are missing or too small.  All
calls go through the C ABI."""
import numpy as np
import pytest

from solidfmm_amd import workload as W

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")


def _device_result(sf, ctx, op, items, src, dst):
    d_src = torch.from_numpy(src).cuda()
    d_dst = torch.from_numpy(dst.copy()).cuda()
    sf.Plan(ctx, op, items).execute(d_src, d_dst)
    torch.cuda.synchronize()
    return d_dst.cpu().numpy()


@pytest.mark.parametrize("op,P,n", [(W.M2L, 10, 20000), (W.M2M, 16, 3000), (W.L2L, 7, 5000), (W.M2L, 20, 500)])
def test_staging_equals_device_pools(op, P, n):
    import solidfmm_amd as sf
    items, src, dst = W.uniform_batch(op, P, P, n, 0x5011DF20 + P)
    dst = np.random.default_rng(P).uniform(-1, 1, len(dst))
    ctx = sf.Context(P)
    want = _device_result(sf, ctx, op, items, src, dst)
    st = sf.Staging(ctx)
    for rep in range(2):  # the second call reuses the grown buffers
        hs, hd = st.reserve(len(src), len(dst))
        hs[:] = src
        hd[:] = dst
        st.translate(op, items, len(src), len(dst))
        assert np.array_equal(hd, want)
    # translate_host over pageable pools
    got = dst.copy()
    ctx.translate_host(op, items, src, got)
    assert np.array_equal(got, want)


def test_staging_repeated_targets_and_sizes():
    """Targets repeated 50 times (scratch-mode plans) and calls that grow
    and shrink the pools."""
    import solidfmm_amd as sf
    P = 8
    ctx = sf.Context(P)
    st = sf.Staging(ctx)
    for n, reps in [(400, 50), (20000, 3), (37, 1)]:
        items, src, _ = W.uniform_batch(W.M2L, P, P, n, 0x5011DF21 + n)
        items["dst"] = (np.arange(n) // reps).astype(np.uint64) * np.uint64(P * (P + 1))
        dst = np.zeros(((n + reps - 1) // reps) * P * (P + 1))
        want = _device_result(sf, ctx, W.M2L, items, src, dst)
        hs, hd = st.reserve(len(src), len(dst))
        hs[:] = src
        hd[:] = dst
        st.translate(W.M2L, items, len(src), len(dst))
        assert np.array_equal(hd, want)


def test_staging_rejects_items_outside_pools():
    import solidfmm_amd as sf
    P = 6
    ctx = sf.Context(P)
    st = sf.Staging(ctx)
    items, src, dst = W.uniform_batch(W.M2L, P, P, 10, 1)
    hs, hd = st.reserve(len(src), len(dst))
    with pytest.raises(sf.LogicError):
        st.translate(W.M2L, items, len(src), len(dst) - 1)
    with pytest.raises(sf.LogicError):
        st.translate(W.M2L, items, len(src) + 10 ** 9, len(dst))


@pytest.mark.parametrize("route", ["bind", "staged", "pinned"])
@pytest.mark.parametrize("reps", [1, 20])
def test_prepared_plan_binds_equal_owning_plan(route, reps):
    """A prepared plan bound into caller memory by each route gives bitwise
    the owning plan's result (distinct targets: rounds; repeated targets:
    scratch tables and reduction)."""
    import solidfmm_amd as sf
    P, n = 9, 6000
    ctx = sf.Context(P)
    items, src, _ = W.uniform_batch(W.M2L, P, P, n, 0x5011DF22 + reps)
    items["dst"] = (np.arange(n) // reps).astype(np.uint64) * np.uint64(P * (P + 1))
    dst = np.random.default_rng(reps).uniform(-1, 1, ((n + reps - 1) // reps) * P * (P + 1))
    want = _device_result(sf, ctx, W.M2L, items, src, dst)
    plan = sf.Plan(ctx, W.M2L, items, prepared=True)
    dev = torch.empty(plan.device_bytes + 256, dtype=torch.uint8, device="cuda")
    ptr = (dev.data_ptr() + 255) // 256 * 256
    host = torch.empty(plan.host_bytes, dtype=torch.uint8, pin_memory=True)
    assert plan.host_bytes > 0
    if route == "bind":
        plan.bind(ptr, plan.device_bytes)
    elif route == "staged":
        plan.bind_staged(ptr, plan.device_bytes, host.data_ptr(), plan.host_bytes)
    else:
        with pytest.raises(sf.LogicError):
            plan.copy_tables(host.data_ptr(), plan.host_bytes - 1)
        plan.copy_tables(host.data_ptr(), plan.host_bytes)
        with pytest.raises(sf.LogicError):
            plan.bind_pinned(ptr, plan.device_bytes, None)
        plan.bind_pinned(ptr, plan.device_bytes, host.data_ptr())
    d_src = torch.from_numpy(src).cuda()
    d_dst = torch.from_numpy(dst.copy()).cuda()
    plan.execute(d_src, d_dst)
    torch.cuda.synchronize()
    assert np.array_equal(d_dst.cpu().numpy(), want)
    with pytest.raises(sf.LogicError):  # too small a device buffer
        plan.bind(ptr, plan.device_bytes - 1)
