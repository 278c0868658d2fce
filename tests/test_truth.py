"""The binary128 truth (oracle/truth.c) and the FP64 error floors -- CPU only.

What pins the truth:
  - its two routes, the naive O(P^4) sums over solid harmonics and the
    rotation algorithm, share nothing but the inputs and agree to ~1e-30;
  - exact closed forms on the z axis, compared with rational arithmetic:
    S_j^0(0,0,r) = j!/r^(j+1), R_j^0(0,0,r) = r^j/j!, and the operators
    M2L  L_n^m = (-1)^(n+m) (n+k)!/r^(n+k+1) M_k^m,
    M2M  M_n^m = sum_k M_k^m (-r)^(n-k)/(n-k)!,
    L2L  L_n^m = sum_k L_k^m r^(k-n)/(k-n)!
    (doc/solidfmm.tex:1404-1433 with R, S on the z axis);
  - the compiled reference's golden vectors sit within its measured
    rounding scale of the truth.
Then the FP64 oracle's error against the truth is measured per order and
checked against the floor table of tests/helpers.py.
"""
import math
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle
from solidfmm_amd import workload as W
from tests.golden.gen_truth import SAMPLES, errors, reference_output, sample
from tests.helpers import FLOOR, load_golden, orders, scale_limit

T = oracle.Truth()


@pytest.fixture(scope="module")
def orc():
    return oracle.Oracle()


def _exact(hi, lo):
    return Fraction(float(hi)) + Fraction(float(lo))


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("pi,po", [(1, 1), (2, 2), (5, 5), (10, 10), (16, 16), (24, 24), (12, 3), (3, 12),
                                   (20, 7), (24, 2), (9, 8)])
def test_routes_agree(op, pi, po):
    """Naive sums and rotations in binary128, with nonzero initial targets,
    shifts along +-z (rho = 0) and, for M2M / L2L, zero shifts."""
    shift = "m2l" if op == W.M2L else "octant"
    n = 24
    items, src, dst = W.uniform_batch(op, pi, po, n, 0x7A0 + 31 * pi + po + op, shift=shift)
    items["x"][1] = items["y"][1] = 0.0
    items["x"][2] = items["y"][2] = 0.0
    items["z"][2] = -items["z"][2]
    if op != W.M2L:
        items["x"][3] = items["y"][3] = items["z"][3] = 0.0
    dst0 = W.random_solids(3, 4, n, po).reshape(-1)
    hn, ln = T.translate(op, items, src, len(dst), dst0=dst0, method=T.NAIVE)
    hr, lr, mag = T.translate(op, items, src, len(dst), dst0=dst0, method=T.ROTATION, magnitude=True)
    d = np.abs((hr - hn) + (lr - ln))
    # relative to the magnitude bound of the terms: the routes round at 1e-34
    # (entries the algorithm never touches, Im C_n^0, hold rounding residue)
    assert np.all(d <= 1e-29 * mag + 1e-32 * mag.max()), float((d / (mag + 1e-3 * mag.max())).max())


@pytest.mark.parametrize("j", [0, 1, 2, 5, 10, 23, 40])
def test_harmonics_z_axis_exact(j):
    r = 1.5
    P = j + 1
    i = j * (j + 1)  # Re C_j^0
    hi, lo = T.harmonic("S", P, 0.0, 0.0, r)
    want = Fraction(math.factorial(j)) / Fraction(r) ** (j + 1)
    assert abs(_exact(hi[i], lo[i]) - want) <= Fraction(1, 10 ** 30) * want
    hi, lo = T.harmonic("R", P, 0.0, 0.0, r)
    want = Fraction(r) ** j / math.factorial(j)
    assert abs(_exact(hi[i], lo[i]) - want) <= Fraction(1, 10 ** 30) * want


def _z_axis_expected(op, P_in, P_out, k, m, c, r):
    """Exact target coefficients (complex Fractions) of a source whose only
    nonzero coefficient is C_k^m = c, shifted by (0, 0, r)."""
    out = {}
    for n in range(P_out):
        if m > n:
            continue
        if op == W.M2L:
            v = Fraction((-1) ** (n + m) * math.factorial(n + k)) / Fraction(r) ** (n + k + 1)
        elif op == W.M2M:
            if k > n:
                continue
            v = Fraction(-r) ** (n - k) / math.factorial(n - k)
        else:
            if k < n:
                continue
            v = Fraction(r) ** (k - n) / math.factorial(k - n)
        out[(n, m)] = (v * c[0], v * c[1])
    return out


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("k,m", [(0, 0), (3, 0), (3, 2), (7, 7), (9, 4)])
def test_operators_z_axis_exact(method, op, k, m):
    P_in, P_out = 10, 10
    r = 1.25 if op != W.M2L else 2.5
    c = (Fraction(3, 4), Fraction(0) if m == 0 else Fraction(-5, 8))
    src = np.zeros(P_in * (P_in + 1))
    src[k * (k + 1) + m] = float(c[0])
    src[(k + 1) * (k + 1) + m] = float(c[1])
    items = W.make_items(np.array([0], dtype=np.uint64), np.array([0], dtype=np.uint64), P_in, P_out,
                         np.array([[0.0, 0.0, r]]))
    hi, lo = T.translate(op, items, src, P_out * (P_out + 1), method=method)
    want = _z_axis_expected(op, P_in, P_out, k, m, c, r)
    scale = max([abs(v) for pair in want.values() for v in pair] + [Fraction(1)])
    for n in range(P_out):
        for mm in range(n + 1):
            got = (_exact(hi[n * (n + 1) + mm], lo[n * (n + 1) + mm]),
                   _exact(hi[(n + 1) * (n + 1) + mm], lo[(n + 1) * (n + 1) + mm]))
            w = want.get((n, mm), (Fraction(0), Fraction(0)))
            for g, e in zip(got, w):
                assert abs(g - e) <= Fraction(1, 10 ** 29) * scale, (op, method, n, mm, float(g), float(e))


@pytest.mark.parametrize("case", sorted(load_golden()))
def test_reference_golden_within_rounding_scale(case):
    """The compiled reference's outputs lie within its rounding scale of the
    truth: a sign, index or term mistake in the truth would put them
    thousands of units away."""
    op, items, src, dst0, dst = load_golden()[case]
    pi, po = orders(items)
    if op == W.L2L and pi > po + 1:
        pytest.skip("the reference's L2L truncation defect (tests/test_oracle.py)")
    hi, lo, mag = T.translate(op, items, src, len(dst), dst0=dst0, magnitude=True)
    d = np.abs((dst - hi) - lo)
    units = d / (2.0 ** -53 * np.maximum(mag, 1e-300))
    assert units.max() <= scale_limit(pi, po) / 2, float(units.max())


@pytest.mark.parametrize("P", [2, 5, 8, 10, 12, 16, 20, 24])
def test_oracle_floor_matches_table(orc, P):
    """The FP64 oracle's 99th-percentile error against the truth on 2,000
    random translations per operator is at most the floor table's entry,
    and not far below it (the table is a measured floor, not a loose
    bound)."""
    worst = 0.0
    for op, shift in ((W.M2L, "m2l"), (W.M2M, "octant"), (W.L2L, "octant")):
        items, src, dst = W.uniform_batch(op, P, P, 2000, 0x7B00 + P + 100 * op, shift=shift)
        hi, lo = T.translate(op, items, src, len(dst))
        got = orc.translate(op, items, src, np.zeros_like(dst))
        e = oracle.truth_degree_errors(got, hi, lo, P).max(axis=1)
        worst = max(worst, float(np.percentile(e, 99)))
    assert FLOOR[P] / 4 <= worst <= FLOOR[P], (worst, FLOOR[P])


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("P", [4, 10, 16, 24])
def test_oracle_error_within_rounding_scale(orc, op, P):
    """Every oracle item's error against the truth is within half the
    scale limit (the limit bounds a difference of two evaluations)."""
    shift = "m2l" if op == W.M2L else "octant"
    items, src, dst = W.uniform_batch(op, P, P, 500, 0x7C00 + P + 100 * op, shift=shift)
    hi, lo, mag = T.translate(op, items, src, len(dst), magnitude=True)
    got = orc.translate(op, items, src, np.zeros_like(dst))
    e = oracle.truth_degree_errors(got, hi, lo, P)
    s = oracle.error_scale(mag, hi, P)
    units = np.where(s > 0, e / np.where(s > 0, s, 1.0), 0.0).max(axis=1)
    assert units.max() <= scale_limit(P, P) / 2, float(units.max())


def test_truncating_l2l_truth_matches_oracle(orc):
    """L2L with p_in >= p_out + 2: the oracle (tables as far as they are
    read) agrees with both binary128 routes."""
    items, src, dst = W.uniform_batch(W.L2L, 16, 8, 200, 0x7D00, shift="octant")
    hi, lo = T.translate(W.L2L, items, src, len(dst), method=T.NAIVE)
    got = orc.translate(W.L2L, items, src, np.zeros_like(dst))
    e = oracle.truth_degree_errors(got, hi, lo, 8).max(axis=1)
    assert e.max() <= 1e-13


@pytest.mark.parametrize("name", list(SAMPLES))
def test_truth_fixture_matches_rotation_route(name):
    """tests/golden/truth_errors.npz holds the reference's and the oracle's
    errors against the naive route; against the rotation route they are the
    same (the GPU tests use the rotation route)."""
    if not oracle.Reference.available():
        pytest.skip("compiled reference not present")
    fx = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "truth_errors.npz"))
    op, items, src, dst = sample(name)
    hi, lo, mag = T.translate(op, items, src, len(dst), magnitude=True)
    for who, v in (("ref", reference_output(oracle.Reference(), op, items, src, dst)),
                   ("orc", oracle.Oracle().translate(op, items, src, np.zeros_like(dst)))):
        e, u = errors(v, hi, lo, mag, items)
        assert np.allclose(e, fx[f"{name}/e_{who}"], rtol=1e-6, atol=1e-20)
        assert np.allclose(u, fx[f"{name}/u_{who}"], rtol=1e-6, atol=1e-9)
