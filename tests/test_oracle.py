"""Pins of the CPU oracle (oracle/oracle.c) -- CPU only.

Each oracle function is checked against something it does not restate:
  swap matrices     Dehnen's B_4 table (the reference's one stored golden
                    vector, tests/test_swap_matrix.cpp:24-35); the compiled
                    reference's swap_matrix; involution identities
  folded blocks     SR^2 = I, SI^2 = I on m,l >= 1, transposed blocks equal
                    D^-1 SR^T D with D = diag(1,2,..,2)
  euler             cos/sin of the angles from atan2 (closed form)
  translations      golden vectors of the compiled reference; the naive
                    O(P^4) formulas (doc/solidfmm.tex:1404-1433); physical
                    convergence to 1/r (tests/test_m2l.cpp pattern); L2L
                    exactness; M2M and L2L composition
  harmonics R, S    the compiled reference's harmonics; |x|^-1 and 1 at
                    degree 0; z-axis closed forms j!/r^(j+1) and r^j/j!
  gradient dR       central finite differences of R; the compiled
                    reference's harmonics::dR; the closed form on the z axis
"""
import math

import numpy as np
import pytest

import oracle
from solidfmm_amd import workload as W
from tests.helpers import assert_parity, load_golden, orders, tol_fp64

# Dehnen (2014) eq. (64): 16 * B_4^{m,l}, rows m = -4..4, columns l = -4..4.
B4x16 = np.array([
    [1, -8, 28, -56, 70, -56, 28, -8, 1],
    [-1, 6, -14, 14, 0, -14, 14, -6, 1],
    [1, -4, 4, 4, -10, 4, 4, -4, 1],
    [-1, 2, 2, -6, 0, 6, -2, -2, 1],
    [1, 0, -4, 0, 6, 0, -4, 0, 1],
    [-1, -2, 2, 6, 0, -6, -2, 2, 1],
    [1, 4, 4, -4, -10, -4, 4, 4, 1],
    [-1, -6, -14, -14, 0, 14, 14, 6, 1],
    [1, 8, 28, 56, 70, 56, 28, 8, 1],
], dtype=float)

GOLDEN = load_golden()


def test_swap_B4_known_answer(orc):
    got = np.array([[orc.swap_entry(4, m, l) for l in range(-4, 5)] for m in range(-4, 5)])
    assert np.array_equal(got * 16, B4x16)


@pytest.mark.reference
def test_swap_matrices_equal_reference(orc, ref):
    for n in range(0, 16):
        for m in range(-n, n + 1):
            for l in range(-n, n + 1):
                assert orc.swap_entry(n, m, l) == ref.swap_entry(n, m, l), (n, m, l)


@pytest.mark.parametrize("n", range(1, 20))
def test_folded_blocks_identities(orc, n):
    SR, SI, TR, TI = (orc.folded(k, n) for k in range(4))
    I = np.eye(n + 1)
    assert np.abs(SR @ SR - I).max() == 0.0
    assert np.abs((SI @ SI - I)[1:, 1:]).max() == 0.0
    assert np.all(SI[:, 0] == 0) and np.all(SI[0, :] == 0)
    D = np.diag([1.0] + [2.0] * n)
    assert np.abs(TR - np.linalg.inv(D) @ SR.T @ D).max() == 0.0
    assert np.abs(TI - np.linalg.inv(D) @ SI.T @ D).max() == 0.0


def test_euler_closed_form(orc):
    for (x, y, z) in [(1.0, 2.0, 3.0), (-0.3, 0.7, -2.0), (0.0, 0.0, 1.5), (0.0, 0.0, -2.0), (2.0, 0.0, 0.0)]:
        k = 12
        e = orc.euler(x, y, z, k)
        r = math.sqrt(x * x + y * y + z * z)
        rho = math.hypot(x, y)
        alpha = math.atan2(x, y) if rho > 0 else 0.0
        beta = math.atan2(-rho, z)
        for i in range(k):
            assert e[0, i] == pytest.approx(r ** i, rel=1e-14)
            assert e[1, i] == pytest.approx(r ** -i, rel=1e-14)
            assert e[2, i] == pytest.approx(math.cos(i * alpha), abs=1e-14)
            assert e[3, i] == pytest.approx(math.sin(i * alpha), abs=1e-14)
            assert e[4, i] == pytest.approx(math.cos(i * beta), abs=1e-14)
            assert e[5, i] == pytest.approx(math.sin(i * beta), abs=1e-14)


@pytest.mark.parametrize("case", sorted(GOLDEN))
def test_oracle_matches_golden(orc, case):
    op, items, src, dst0, dst = GOLDEN[case]
    got = orc.translate(op, items, src, dst0.copy())
    pi, po = orders(items)
    assert_parity(got, dst, pi, po, label=case, case=(op, items, src, dst0))


@pytest.mark.parametrize("op", [W.M2M, W.M2L, W.L2L])
@pytest.mark.parametrize("pi,po", [(1, 1), (3, 3), (6, 6), (10, 10), (13, 13), (5, 9), (9, 5), (10, 3), (2, 11)])
def test_oracle_matches_naive(orc, op, pi, po):
    shift = "m2l" if op == W.M2L else "octant"
    items, src, dst = W.uniform_batch(op, pi, po, 12, 4242 + 31 * pi + po, shift=shift)
    a = orc.translate(op, items, src, dst.copy())
    b = orc.translate(op, items, src, dst.copy(), naive=True)
    assert_parity(a, b, pi, po, label="naive", case=(op, items, src, dst))


def test_harmonics_closed_forms(orc):
    P = 12
    r = 1.7
    S = orc.harmonic("S", P, 0.0, 0.0, r)
    R = orc.harmonic("R", P, 0.0, 0.0, r)
    for j in range(P):
        assert S[j * (j + 1)] == pytest.approx(math.factorial(j) / r ** (j + 1), rel=1e-14)
        assert R[j * (j + 1)] == pytest.approx(r ** j / math.factorial(j), rel=1e-14)
    x, y, z = 0.3, -1.1, 0.8
    assert orc.harmonic("S", 1, x, y, z)[0] == pytest.approx(1 / math.sqrt(x * x + y * y + z * z))
    assert orc.harmonic("R", 1, x, y, z)[0] == 1.0


@pytest.mark.reference
def test_harmonics_equal_reference(orc, ref):
    for (x, y, z) in [(0.3, -1.1, 0.8), (1.0, 2.0, -0.5)]:
        for kind in ("R", "S"):
            a = orc.harmonic(kind, 15, x, y, z)
            b = ref.harmonic(kind, 15, x, y, z)
            assert np.abs(a - b).max() <= 1e-15 * np.abs(b).max()


def test_harmonic_gradient_finite_differences(orc):
    P = 12
    x, y, z = 0.31, -0.72, 0.55
    g = orc.harmonic("dR", P, x, y, z).reshape(3, -1)
    h = 1e-5
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (orc.harmonic("R", P, x + e[0], y + e[1], z + e[2])
              - orc.harmonic("R", P, x - e[0], y - e[1], z - e[2])) / (2 * h)
        assert np.abs(g[d] - fd).max() <= 1e-9 * max(1.0, np.abs(fd).max())
    # on the z axis only the m = 0 (d/dz: r^(j-1)/(j-1)!) and m = 1 (d/dx, d/dy)
    # entries survive
    r = 1.3
    g = orc.harmonic("dR", P, 0.0, 0.0, r).reshape(3, -1)
    for j in range(1, P):
        assert g[2][j * (j + 1)] == pytest.approx(r ** (j - 1) / math.factorial(j - 1), rel=1e-14)


@pytest.mark.reference
def test_harmonic_gradient_equals_reference(orc, ref):
    for P in (1, 2, 5, 12):
        for (x, y, z) in [(0.3, -1.1, 0.8), (1.0, 2.0, -0.5)]:
            a = orc.harmonic("dR", P, x, y, z)
            b = ref.harmonic("dR", P, x, y, z)
            assert np.abs(a - b).max() <= 1e-15 * max(1.0, np.abs(b).max())


def _potential_after(orc, op, P):
    """Point charge expansion through one translation, evaluated; the
    geometry of the reference's tests/test_m2l.cpp:57-60."""
    xs, ys, zs = 1.1, 0.2, 0.9
    xa, ya, za = 1.0, 0.0, 1.0
    xb, yb, zb = 0.3, 0.1, 2.0
    x, y, z = 0.1, 0.2, 2.5
    rr = math.sqrt((x - xs) ** 2 + (y - ys) ** 2 + (z - zs) ** 2)
    M = orc.harmonic("R", P, xs - xa, ys - ya, zs - za)
    L = np.zeros(P * (P + 1))
    items = W.make_items(np.array([0], np.uint64), np.array([0], np.uint64), P, P,
                         np.array([[xb - xa, yb - ya, zb - za]]))
    orc.translate(op, items, M, L)
    val = orc.dot(L, P, orc.harmonic("R", P, x - xb, y - yb, z - zb), P)
    return abs(val - 1 / rr) * rr


def test_physics_m2l_converges(orc):
    errs = [_potential_after(orc, W.M2L, P) for P in (4, 8, 12, 16, 20, 30)]
    assert all(b < a for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-9


def test_l2l_exact_and_round_trip(orc):
    P = 14
    L = W.random_solids(7, 1, 1, P).reshape(-1)
    sh = np.array([[0.3, -0.4, 0.25]])
    it = W.make_items(np.array([0], np.uint64), np.array([0], np.uint64), P, P, sh)
    back = W.make_items(np.array([0], np.uint64), np.array([0], np.uint64), P, P, -sh)
    A = orc.translate(W.L2L, it, L, np.zeros_like(L))
    B = orc.translate(W.L2L, back, A, np.zeros_like(L))
    assert oracle.degree_errors(B, L, P).max() < 1e-12
    # exactness: evaluation at a point agrees before and after
    pt = np.array([0.2, 0.1, -0.3])
    v0 = orc.dot(L, P, orc.harmonic("R", P, *pt), P)
    v1 = orc.dot(A, P, orc.harmonic("R", P, *(pt - sh[0])), P)
    assert v1 == pytest.approx(v0, rel=1e-12)


@pytest.mark.parametrize("op", [W.M2M, W.L2L])
def test_composition(orc, op):
    P = 12
    C = W.random_solids(8, 1, 1, P).reshape(-1)
    r1 = np.array([[0.3, -0.2, 0.4]])
    r2 = np.array([[-0.1, 0.5, 0.2]])
    z = np.array([0], np.uint64)
    one = lambda s: W.make_items(z, z, P, P, s)  # noqa: E731
    A = orc.translate(op, one(r1), C, np.zeros_like(C))
    B = orc.translate(op, one(r2), A, np.zeros_like(C))
    D = orc.translate(op, one(r1 + r2), C, np.zeros_like(C))
    assert oracle.degree_errors(B, D, P).max() < 1e-12


def test_zero_shift_is_copy_add(orc):
    for op in (W.M2M, W.L2L):
        src = W.random_solids(3, 1, 1, 6).reshape(-1)
        dst = W.random_solids(3, 2, 1, 8).reshape(-1)
        it = W.make_items(np.array([0], np.uint64), np.array([0], np.uint64), 6, 8, np.zeros((1, 3)))
        got = orc.translate(op, it, src, dst.copy())
        want = dst.copy()
        want[:42] += src
        assert np.array_equal(got, want)


@pytest.mark.reference
def test_reference_avx512_lane_defect_documented(orc, ref):
    """microkernel_avx512f.cpp:2182 stores r20 for r21: in 16-lane batches
    lanes 12-15 see Im C_n^m (m = 2 mod 4, n >= 3) of lane l-4.  Calls of
    <= 12 items avoid it; whole-list calls show it on AVX-512 hosts."""
    items, src, dst = W.uniform_batch(W.M2L, 10, 10, 16, 99)
    want = orc.translate(W.M2L, items, src, dst.copy())
    safe = dst.copy()
    ref.translate(W.M2L, 10, items, src, safe, chunk=12)
    assert oracle.degree_errors(safe, want, 10).max() < 1e-13
    whole = dst.copy()
    ref.translate(W.M2L, 10, items, src, whole, chunk=0)
    per_item = oracle.degree_errors(whole, want, 10).max(axis=1)
    assert per_item[:12].max() < 1e-13
    with open("/proc/cpuinfo") as f:
        avx512 = "avx512f" in f.read()
    if avx512:
        assert per_item[12:].min() > 1e-3


@pytest.mark.reference
def test_reference_l2l_truncation_defect_documented(orc, ref):
    """l2l with p_in >= p_out + 2 reads Euler tables past index
    min(p_in,p_out) (translations.cpp:414-417 vs :198); the oracle computes
    them, and agrees with the naive formula."""
    items, src, dst = W.uniform_batch(W.L2L, 10, 3, 4, 5, shift="octant")
    naive = orc.translate(W.L2L, items, src, dst.copy(), naive=True)
    mine = orc.translate(W.L2L, items, src, dst.copy())
    theirs = dst.copy()
    ref.translate(W.L2L, 10, items, src, theirs)
    assert oracle.degree_errors(mine, naive, 3).max() < 1e-13
    assert oracle.degree_errors(theirs, naive, 3).max() > 1e-6


@pytest.mark.reference
def test_reference_error_types(ref):
    # 0 ok, 1 logic_error, 2 out_of_range, 3 overflow_error
    assert ref.check_errors(1, 10, 10, 1, 10, 1) == 0
    assert ref.check_errors(1, 10, 11, 1, 10, 1) == 2
    assert ref.check_errors(1, 10, 10, 2, 10, 1) == 1
    assert ref.check_errors(1, 0, 0, 1, 0, 1) == 1
    assert ref.check_errors(1, 87, 10, 1, 10, 1) == 3


@pytest.mark.reference
def test_reference_pool_protocol(ref):
    """The CPU baseline's run of the reference (oracle/ref_shim.cpp,
    ref_pool_translate_f64): workers pinned one per core, a warm-up call,
    timed calls on zeroed targets.  With at most 8 items per call it stays
    clear of the AVX-512 lane defect, so its targets equal the reference's
    single-thread result; the host description names the microkernel."""
    items, src, dst = W.uniform_batch(W.M2L, 10, 10, 8, 77)
    want = dst.copy()
    ref.translate(W.M2L, 10, items, src, want)
    got = dst.copy()
    secs, used, pinned = ref.pool_translate(W.M2L, 10, items, src, got, threads=2, warmup=4, repeats=3)
    assert used == 2 and pinned == 2
    assert len(secs) == 3 and np.all(secs > 0)
    assert np.array_equal(got, want)
    info = ref.host_info()
    assert info["physical_cores"] >= 1 and info["nproc"] >= info["physical_cores"]
    assert info["microkernel"].startswith("microkernel_double_")
