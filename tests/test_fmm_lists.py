"""The FMM list builder (solidfmm_amd/fmm.py) -- host code, CPU.

The dual tree walk must cover every ordered pair of points exactly once:
either by one M2L between cells satisfying the acceptance criterion or by
one near pair of leaves (a leaf with itself included).  The passes must be
ordered so every expansion is complete before it is read.
"""
import numpy as np
import pytest

from solidfmm_amd import fmm as F
from solidfmm_amd import octree as O
from solidfmm_amd.workload import L2L, M2L, M2M


@pytest.mark.parametrize("dist,theta,leaf", [("plummer", 0.5, 16), ("plummer", 0.8, 8), ("uniform", 0.6, 32),
                                             ("clustered", 0.5, 4)])
def test_every_pair_covered_once(dist, theta, leaf):
    rng = np.random.default_rng(3)
    n = 2500
    if dist == "plummer":
        pts = O.plummer(n, 0x5011DF0C)
    elif dist == "uniform":
        pts = rng.uniform(-1, 1, (n, 3))
    else:  # a few tight clusters, with exact duplicates
        pts = np.repeat(rng.standard_normal((50, 3)), n // 50, axis=0) + 1e-3 * rng.standard_normal((n, 3))
        pts[:20] = pts[0]
    tree = O.Octree(pts, leaf_size=leaf)
    balls = F.cell_balls(tree, pts)
    ms, mt, ns, nt = F.mac_lists(tree, balls, theta)
    # acceptance criterion on every M2L pair; near pairs are leaves
    d = np.linalg.norm(balls[ms, :3] - balls[mt, :3], axis=1)
    assert np.all(balls[ms, 3] + balls[mt, 3] < theta * d)
    assert tree.leaf[ns].all() and tree.leaf[nt].all()
    cover = np.zeros((n, n), dtype=np.int32)
    for s, t in zip(np.concatenate([ms, ns]), np.concatenate([mt, nt])):
        cover[tree.first[t]:tree.first[t] + tree.count[t], tree.first[s]:tree.first[s] + tree.count[s]] += 1
    assert cover.min() == 1 and cover.max() == 1
    # every leaf is near itself exactly once
    selfp = ns[ns == nt]
    assert np.array_equal(np.sort(selfp), np.nonzero(tree.leaf)[0])


def test_pass_order_and_centres():
    pts = O.plummer(4000, 0x5011DF0D)
    tree = O.Octree(pts, leaf_size=32)
    P, off, pool_len, passes = F.mac_translation_lists(tree, pts, theta=0.5, p_root=12, p_leaf=6)
    ops = [p[0] for p in passes]
    first_m2l = ops.index(M2L)
    assert all(o == M2M for o in ops[:first_m2l]) and all(o == L2L for o in ops[first_m2l + 1:])
    levels = [p[4] for p in passes]
    assert levels[:first_m2l] == sorted(levels[:first_m2l], reverse=True)
    assert levels[first_m2l + 1:] == sorted(levels[first_m2l + 1:])
    # shifts are between the cells' smallest balls; blocks do not overlap
    balls = F.cell_balls(tree, pts)
    for op, it, s, d, _ in passes:
        sh = np.c_[it["x"], it["y"], it["z"]]
        assert np.allclose(sh, balls[d, :3] - balls[s, :3], rtol=0, atol=1e-12)
        assert np.array_equal(it["src"], off[s]) and np.array_equal(it["dst"], off[d])
    assert np.all(np.diff(off) == P[:-1] * (P[:-1] + 1)) and pool_len == off[-1] + P[-1] * (P[-1] + 1)
    # every cell's ball encloses its points
    srt = pts[tree.order]
    for c in range(0, len(tree), 7):
        p = srt[tree.first[c]:tree.first[c] + tree.count[c]]
        assert np.linalg.norm(p - balls[c, :3], axis=1).max() <= balls[c, 3] * (1 + 1e-12) + 1e-300
