"""Host-pool staging (sfmm_staging_*, sfmm_translate_host) -- needs an
MI355X.  The staged path must give bitwise the results of the same plan on
device pools, grow its buffers across calls of different sizes, and reject
items outside the reserved pools.  All calls go through the C ABI."""
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
