"""Smallest enclosing balls (solidfmm_amd/csrc/sphere.cpp) -- host code, CPU.

Pinned by: the compiled reference's bounding_sphere (sphere.cpp:546-627) on
random and degenerate point and ball sets; closed forms (two points, the
equilateral triangle, the regular tetrahedron, the cube, two balls, nested
balls); brute force over all support sets of at most four inputs on small
sets; enclosure of every input on large sets.
"""
import itertools
import math

import numpy as np
import pytest

import solidfmm_amd as sf


def brute_force_ball(pts):
    """Smallest ball of (n <= 8) points: the smallest circumball of a subset of
    at most four points (centre in the subset's affine hull) that encloses
    all points."""
    best = None
    for k in range(1, 5):
        for sub in itertools.combinations(range(len(pts)), k):
            q = pts[list(sub)]
            d = q[1:] - q[0]
            if k == 1:
                c = q[0]
            else:
                G = 2 * d @ d.T
                if abs(np.linalg.det(G)) < 1e-12:
                    continue
                lam = np.linalg.solve(G, (d * d).sum(1))
                c = q[0] + lam @ d
            R = np.linalg.norm(q - c, axis=1).max()
            if np.linalg.norm(pts - c, axis=1).max() <= R * (1 + 1e-12) + 1e-15:
                if best is None or R < best[3]:
                    best = np.array([*c, R])
    return best


def check_encloses(ball, pts, radii=None):
    r = np.zeros(len(pts)) if radii is None else radii
    reach = np.linalg.norm(pts[:, :3] - ball[:3], axis=1) + r
    assert reach.max() <= ball[3] * (1 + 1e-13) + 1e-300


def test_closed_forms():
    b = sf.bounding_sphere(np.array([[0.0, 0, 0], [2.0, 0, 0]]))
    assert np.allclose(b, [1, 0, 0, 1], atol=1e-15)
    s3 = math.sqrt(3.0)
    tri = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.5, s3 / 2, 0]])
    b = sf.bounding_sphere(tri)
    assert np.allclose(b[:3], [0.5, s3 / 6, 0], atol=1e-15) and b[3] == pytest.approx(1 / s3, rel=1e-14)
    tet = np.array([[1.0, 1, 1], [1.0, -1, -1], [-1.0, 1, -1], [-1.0, -1, 1]])
    b = sf.bounding_sphere(tet)
    assert np.allclose(b[:3], 0, atol=1e-15) and b[3] == pytest.approx(math.sqrt(3.0), rel=1e-14)
    cube = np.array(list(itertools.product([0.0, 2.0], repeat=3)))
    b = sf.bounding_sphere(cube)
    assert np.allclose(b[:3], 1, atol=1e-15) and b[3] == pytest.approx(math.sqrt(3.0), rel=1e-14)
    # an interior point changes nothing; a single point is a ball of radius 0
    assert np.allclose(sf.bounding_sphere(np.vstack([cube, [[1.0, 1.2, 0.7]]])), b, atol=1e-15)
    assert np.array_equal(sf.bounding_sphere(np.array([[3.0, -1, 2]])), [3, -1, 2, 0])
    # balls: two apart -> (d + r1 + r2)/2 on the line; nested -> the outer one
    b = sf.bounding_sphere(spheres=np.array([[0.0, 0, 0, 1.0], [4.0, 0, 0, 0.5]]))
    assert b[3] == pytest.approx((4 + 1 + 0.5) / 2, rel=1e-14) and b[0] == pytest.approx(-1 + b[3], rel=1e-14)
    b = sf.bounding_sphere(spheres=np.array([[0.0, 0, 0, 3.0], [0.5, 0.2, -0.1, 1.0], [-1.0, 0.0, 0.0, 0.5]]))
    assert np.allclose(b, [0, 0, 0, 3], atol=1e-15)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 8])
def test_brute_force(n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        pts = rng.standard_normal((n, 3))
        b = sf.bounding_sphere(pts)
        w = brute_force_ball(pts)
        assert np.abs(b[:3] - w[:3]).max() <= 1e-12 * w[3]
        assert b[3] == pytest.approx(w[3], rel=1e-12)


@pytest.mark.parametrize("shape", ["gauss", "cube", "sphere_surface", "plane", "line", "duplicates", "far"])
def test_large_sets_enclose(shape):
    rng = np.random.default_rng(5)
    n = 5000
    if shape == "gauss":
        pts = rng.standard_normal((n, 3))
    elif shape == "cube":
        pts = rng.uniform(-1, 1, (n, 3))
    elif shape == "sphere_surface":
        v = rng.standard_normal((n, 3))
        pts = v / np.linalg.norm(v, axis=1, keepdims=True)
    elif shape == "plane":
        pts = np.c_[rng.uniform(-1, 1, (n, 2)), np.zeros(n)]
    elif shape == "line":
        pts = np.outer(rng.uniform(-1, 1, n), [0.3, -0.5, 0.8])
    elif shape == "duplicates":
        pts = np.repeat(rng.standard_normal((7, 3)), 50, axis=0)
    else:
        pts = 1e6 + rng.uniform(-1e-3, 1e-3, (n, 3))
    b = sf.bounding_sphere(pts)
    check_encloses(b, pts)
    if shape == "sphere_surface":
        assert b[3] <= 1 + 1e-12 and b[3] >= 0.999
    if shape == "line":
        t = pts @ np.array([0.3, -0.5, 0.8]) / np.dot([0.3, -0.5, 0.8], [0.3, -0.5, 0.8])
        assert b[3] == pytest.approx((t.max() - t.min()) / 2 * np.linalg.norm([0.3, -0.5, 0.8]), rel=1e-12)
    # at least two inputs touch the boundary (a smallest ball is tight)
    d = np.linalg.norm(pts - b[:3], axis=1)
    slack = 1e-9 * b[3] + 1e-14 * np.abs(b[:3]).max()
    assert np.sum(d >= b[3] - slack) >= 2


def test_batched_equals_single():
    rng = np.random.default_rng(9)
    pts = np.c_[rng.standard_normal((8000, 3)), rng.uniform(0, 1, 8000)]
    counts = rng.integers(1, 200, 40)
    first = np.concatenate([[0], np.cumsum(counts)[:-1]])
    out = sf.bounding_spheres(pts[:counts.sum()], first, counts, threads=3)
    for s in range(len(counts)):
        assert np.array_equal(out[s], sf.bounding_sphere(pts[first[s]:first[s] + counts[s]]))


def test_errors():
    with pytest.raises(sf.LogicError):
        sf.bounding_sphere(np.zeros((0, 3)))
    with pytest.raises(sf.LogicError):
        sf.bounding_sphere(spheres=np.array([[0.0, 0, 0, -1.0]]))
    with pytest.raises(sf.LogicError):
        sf.bounding_spheres(np.zeros((4, 3)), [2], [3])
    with pytest.raises(sf.LogicError):
        sf.bounding_spheres(np.zeros((4, 3)), [0], [0])


@pytest.mark.reference
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 17, 100, 2000])
def test_points_equal_reference(ref, n):
    rng = np.random.default_rng(100 + n)
    for pts in (rng.standard_normal((n, 3)), rng.uniform(0, 1, (n, 3)), np.c_[rng.uniform(-1, 1, (n, 2)), np.zeros(n)]):
        a = sf.bounding_sphere(pts)
        b = ref.bounding_sphere(points=pts)
        assert np.abs(a[:3] - b[:3]).max() <= 1e-10 * max(b[3], 1e-300)
        assert a[3] == pytest.approx(b[3], rel=1e-10, abs=1e-300)


@pytest.mark.reference
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 9, 64, 500])
def test_balls_equal_reference(ref, n):
    rng = np.random.default_rng(200 + n)
    for _ in range(3):
        b4 = np.c_[rng.standard_normal((n, 3)), rng.uniform(0, 0.8, n)]
        a = sf.bounding_sphere(spheres=b4)
        b = ref.bounding_sphere(spheres=b4)
        check_encloses(a, b4[:, :3], b4[:, 3])
        assert np.abs(a[:3] - b[:3]).max() <= 1e-10 * b[3]
        assert a[3] == pytest.approx(b[3], rel=1e-10)


def _run_check(name):
    import os
    import subprocess
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (oracle/Makefile target dropin)")
    r = subprocess.run([path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout.splitlines()


def test_dropin_cxx_api_matches_reference_program():
    """The same C++ program over the public sphere API (oracle/sphere_check.cpp),
    built against the reference and against the drop-in: double results agree
    to rounding, float results to float precision, the empty set throws."""
    ref_lines = _run_check("ref_sphere_check")
    mine = _run_check("dropin_sphere_check")
    assert len(ref_lines) == len(mine)
    for a, b in zip(ref_lines, mine):
        if a.startswith("empty"):
            assert a == b == "empty: logic_error"
            continue
        ka, kb = a.split()[:3], b.split()[:3]
        assert ka == kb
        va = np.array([float(t) for t in a.split()[3:]])
        vb = np.array([float(t) for t in b.split()[3:]])
        tol = 1e-12 if ka[2] == "8" else 2e-6
        scale = max(va[3], 1e-300)
        assert np.abs(va - vb).max() <= tol * max(scale, 1.0), (a, b)
