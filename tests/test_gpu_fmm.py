"""A whole far-field FMM evaluation on the GPU (solidfmm_amd/fmm.py) against
direct summation -- needs an MI355X.

P2M, M2M, M2L over the acceptance-criterion lists, L2L and L2P run through
the C ABI; the near field of sampled targets is added on the host from the
near pairs; the sum must equal the direct potential sum_j q_j/|x_i - x_j|
and its gradient to the truncation error, which must fall with the order.
"""
import numpy as np
import pytest

from solidfmm_amd import octree as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a visible GPU")


def _system(n, seed):
    rng = np.random.default_rng(seed)
    pts = np.c_[O.plummer(n, seed), rng.uniform(0.5, 1.5, n) / n]
    return pts


def _errors(fmm, pts, sample):
    phi, grad = fmm.evaluate()
    torch.cuda.synchronize()
    phi = phi.cpu().numpy()
    grad = grad.cpu().numpy()
    tree = fmm.tree
    srt = pts[tree.order]
    rank = np.argsort(tree.order)  # sorted position of each original point
    leaf_of = np.empty(len(pts), dtype=np.int64)
    for c in np.nonzero(tree.leaf)[0]:
        leaf_of[tree.first[c]:tree.first[c] + tree.count[c]] = c
    ns, nt = fmm.near_pairs()
    ephi, egrad = [], []
    for i in sample:
        x = pts[i, :3]
        d = x - pts[:, :3]
        r = np.linalg.norm(d, axis=1)
        r[i] = np.inf
        want_phi = (pts[:, 3] / r).sum()
        want_grad = -(pts[:, 3, None] * d / r[:, None] ** 3).sum(0)
        near = ns[nt == leaf_of[rank[i]]]
        idx = np.concatenate([np.arange(tree.first[s], tree.first[s] + tree.count[s]) for s in near])
        dn = x - srt[idx, :3]
        rn = np.linalg.norm(dn, axis=1)
        rn[rn == 0] = np.inf  # the point itself (and exact duplicates of it)
        got_phi = phi[i] + (srt[idx, 3] / rn).sum()
        got_grad = grad[i] - (srt[idx, 3, None] * dn / rn[:, None] ** 3).sum(0)
        ephi.append(abs(got_phi - want_phi) / abs(want_phi))
        egrad.append(np.linalg.norm(got_grad - want_grad) / np.linalg.norm(want_grad))
    return np.max(ephi), np.max(egrad)


@pytest.mark.parametrize("theta", [0.5, 0.7])
def test_fmm_matches_direct_sum_and_converges(theta):
    import solidfmm_amd as sf
    from solidfmm_amd.fmm import FMM
    pts = _system(20000, 0x5011DF10)
    rng = np.random.default_rng(1)
    sample = rng.choice(len(pts), 150, replace=False)
    errs = {}
    for P in (6, 10, 16):
        ctx = sf.Context(P)
        fmm = FMM(ctx, pts, leaf_size=32, theta=theta)
        errs[P] = _errors(fmm, pts, sample)
        print(f"theta {theta} P {P}: potential {errs[P][0]:.2e} gradient {errs[P][1]:.2e} "
              f"({fmm.translations} translations)", flush=True)
    # the truncation error falls geometrically with the order, like theta^P
    assert errs[16][0] < 5 * theta ** 16 and errs[16][1] < 50 * theta ** 16
    assert errs[10][0] < 5 * theta ** 10
    assert errs[16][0] < errs[10][0] < errs[6][0]
    assert errs[16][1] < errs[10][1] < errs[6][1]


def test_fmm_mixed_orders_and_repeat():
    """Orders falling from 16 at the root to 8 at the leaves (mixed-order
    M2L items); a second evaluate gives bitwise the same result."""
    import solidfmm_amd as sf
    from solidfmm_amd.fmm import FMM
    pts = _system(12000, 0x5011DF11)
    ctx = sf.Context(16)
    fmm = FMM(ctx, pts, leaf_size=24, theta=0.5, p_root=16, p_leaf=8)
    assert len(np.unique(fmm.orders)) > 3
    e = _errors(fmm, pts, np.random.default_rng(2).choice(len(pts), 100, replace=False))
    assert e[0] < 5 * 0.5 ** 8 and e[1] < 50 * 0.5 ** 8
    a = [t.clone() for t in fmm.evaluate()]
    b = fmm.evaluate()
    torch.cuda.synchronize()
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_fmm_vector_charges():
    """Three charges per point (a vortex method's vector potential): every
    component of the potential and of its gradient matches direct sums."""
    import solidfmm_amd as sf
    from solidfmm_amd.fmm import FMM
    n = 12000
    rng = np.random.default_rng(5)
    pts = np.c_[O.plummer(n, 0x5011DF12), rng.uniform(-1, 1, (n, 3)) / n]
    ctx = sf.Context(14)
    fmm = FMM(ctx, pts, leaf_size=32, theta=0.5)
    assert fmm.dim == 3
    phi, grad = fmm.evaluate()
    torch.cuda.synchronize()
    phi = phi.cpu().numpy()
    grad = grad.cpu().numpy()
    assert phi.shape == (n, 3) and grad.shape == (n, 3, 3)
    ep, eg = [], []
    for i in rng.choice(n, 60, replace=False):
        d = pts[i, :3] - pts[:, :3]
        r = np.linalg.norm(d, axis=1)
        r[i] = np.inf
        want_p = (pts[:, 3:] / r[:, None]).sum(0)
        want_g = -(pts[:, 3:, None] * (d / r[:, None] ** 3)[:, None, :]).sum(0)
        near_p, near_g = fmm.near_field_at(pts, i)
        ep.append(np.abs(phi[i] + near_p - want_p).max() / np.abs(want_p).max())
        eg.append(np.abs(grad[i] + near_g - want_g).max() / np.abs(want_g).max())
    print(f"vector FMM: potential {max(ep):.2e} gradient {max(eg):.2e}", flush=True)
    assert max(ep) < 5e-5 and max(eg) < 5e-3
