"""Shared pieces of the test-suite: tolerances and golden vectors."""
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# 99th percentile of the per-degree normwise relative error of an FP64
# evaluation against the binary128 truth (oracle/truth.c), per order
# P = max(p_in, p_out): the largest over M2L (|x| in [2,4]), M2M and L2L
# (octant shifts), over the compiled reference and the FP64 oracle, and over
# three seeds of 2,000 random-input translations each, rounded up to a
# monotone envelope.  tests/test_truth.py re-measures the oracle against it.
# At P = 1 the entry covers accumulation into random targets, which cancels
# (the translation itself is exact to 2 roundings there).
FLOOR = {1: 3e-15, 2: 6e-15, 3: 1e-14, 4: 1.6e-14, 5: 1.6e-14, 6: 2e-14, 7: 3.5e-14, 8: 4e-14, 9: 5e-14,
         10: 5.5e-14, 11: 1.2e-13, 12: 1.6e-13, 13: 2.7e-13, 14: 5e-13, 15: 7.5e-13, 16: 1.6e-12, 17: 3.5e-12,
         18: 5.5e-12, 19: 8e-12, 20: 1.9e-11, 21: 4e-11, 22: 5.6e-11, 23: 1.4e-10, 24: 2e-10}
U = 2.0 ** -53


def tol_fp64(p_in, p_out):
    """99th-percentile bound on the per-degree normwise relative difference
    between two FP64 evaluations (kernel against oracle or reference).

    Up to order 24 it is 2 x FLOOR[max(p_in, p_out)]: the difference of
    two evaluations whose own errors sit at the measured floor (1.1e-13 at
    P = 10).  Unequal orders follow the larger
    one (M2L 12 -> 3 and 24 -> 2 measure at the floors of 12 and 24).
    Above 24 the inputs are physical (decaying) expansions and the oracle
    and the naive formulas already differ by 7e-9 at P=30 and 2e-6 at P=40.
    """
    P = max(p_in, p_out)
    if P <= 24:
        return 2.0 * FLOOR[max(P, 1)]
    return 5e-8 if P <= 30 else 1e-4


def scale_limit(p_in, p_out):
    """Bound on a difference of two FP64 evaluations in units of the
    algorithm's first-order rounding scale u ||M_n|| (oracle.error_scale):
    10 max(P, 4).  A single evaluation measures up to 34 units at M2L P=10,
    53 at 16 and 103 at 24 (cancellation in the low output degrees); M2M and
    L2L stay below 6."""
    return 10.0 * max(p_in, p_out, 4)


def scaled_differences(got, want, op, items, src, dst0=None):
    """Per item: max over degrees of ||got_n - want_n|| / (u ||M_n||), with M
    the binary128 magnitude propagation of the rotation algorithm over the
    same list (oracle.Truth); degrees with M_n = 0 must match exactly."""
    import oracle
    dst_len = len(want)
    _, _, mag = oracle.Truth().translate(op, items, src, dst_len, dst0=dst0, magnitude=True)
    out = np.zeros(len(items))
    for po in np.unique(items["p_out"]):
        po = int(po)
        idx = np.nonzero(items["p_out"] == po)[0]
        cols = items["dst"][idx].astype(np.int64)[:, None] + np.arange(po * (po + 1))
        g, w, m = got[cols], want[cols], mag[cols]
        r = np.zeros((len(idx), po))
        for n in range(po):
            sl = slice(n * (n + 1), (n + 1) * (n + 2))
            d = np.linalg.norm(g[:, sl] - w[:, sl], axis=1)
            s = U * np.linalg.norm(m[:, sl], axis=1)
            r[:, n] = np.where(s > 0, d / np.where(s > 0, s, 1.0), np.where(d > 0, np.inf, 0.0))
        out[idx] = r.max(axis=1)
    return out


def assert_parity(got, want, p_in, p_out, tol=None, label="", case=None):
    """Per-translation error e = max over degrees of the per-degree normwise
    relative error.  Requires p99(e) <= tol.

    With the list given (case = (op, items, src, dst0)) every item is also
    bounded in units of the rounding scale, and the p99 bar applies from
    1,000 items on, the 90th percentile below that (the floors are 99th
    percentiles of 2,000 items; in 200 the 99th is the second worst): a
    random input can make a degree's result small by cancellation, which
    inflates its relative error for every FP64 implementation (the reference
    and the oracle reach 3.5e-12 and 4.3e-12 against the truth on one of
    4096 P=10 M2L translations, where the 99th percentile is 5e-14), but not
    its error in units of u ||M_n||.  Without the list the largest item may
    reach 30 tol.
    """
    import oracle
    tol = tol_fp64(p_in, p_out) if tol is None else tol
    e = oracle.degree_errors(got, want, p_out).max(axis=1)
    p99 = float(np.percentile(e, 99)) if len(e) else 0.0
    mx = float(e.max()) if len(e) else 0.0
    if case is None or len(e) >= 1000:
        assert p99 <= tol, f"{label}: p99 {p99:.2e} max {mx:.2e} tol {tol:.1e}"
    elif len(e):
        p90 = float(np.percentile(e, 90))
        assert p90 <= tol, f"{label}: p90 {p90:.2e} max {mx:.2e} tol {tol:.1e}"
    if case is not None and len(e):
        op, items, src, dst0 = case
        r = scaled_differences(np.asarray(got).reshape(-1), np.asarray(want).reshape(-1), op, items, src, dst0)
        lim = scale_limit(p_in, p_out)
        assert r.max() <= lim, f"{label}: difference {r.max():.1f} rounding units (limit {lim:.0f})"
    else:
        assert mx <= 30 * tol, f"{label}: p99 {p99:.2e} max {mx:.2e} tol {tol:.1e}"
    return p99, mx


def load_golden(name="golden_f64.npz"):
    """{case: (op, items, src, dst0, dst)} from a golden file."""
    g = np.load(os.path.join(GOLDEN_DIR, name), allow_pickle=False)
    out = {}
    for k in g.files:
        case, field = k.split("/")
        if case == "meta":
            continue
        out.setdefault(case, {})[field] = g[k]
    return {c: (int(v["op"]), v["items"], v["src"], v["dst0"], v["dst"]) for c, v in out.items()}


def orders(items):
    return int(items["p_in"].max()), int(items["p_out"].max())


def assert_mixed_parity(got, want, items, scale=1.0, label="", case=None):
    """assert_parity for mixed orders: each item against tol_fp64 of its own
    orders (times `scale`); p99 of the ratios <= 1.  The largest ratio is
    bounded by 30, or, with case = (op, src, dst0), every item's difference
    by scale x scale_limit rounding units."""
    import oracle
    e = oracle.mixed_errors(got, want, items)
    tol = np.array([tol_fp64(int(a), int(b)) for a, b in zip(items["p_in"], items["p_out"])]) * scale
    ratio = e / tol
    p99 = float(np.percentile(ratio, 99)) if len(ratio) else 0.0
    mx = float(ratio.max()) if len(ratio) else 0.0
    if case is None or len(e) >= 1000:
        assert p99 <= 1.0, f"{label}: error/tol p99 {p99:.2f} max {mx:.2f}"
    elif len(e):
        p90 = float(np.percentile(ratio, 90))
        assert p90 <= 1.0, f"{label}: error/tol p90 {p90:.2f} max {mx:.2f}"
    if case is not None and len(e):
        op, src, dst0 = case
        r = scaled_differences(np.asarray(got), np.asarray(want), op, items, src, dst0)
        lim = np.array([scale_limit(int(a), int(b)) for a, b in zip(items["p_in"], items["p_out"])]) * scale
        assert (r / lim).max() <= 1.0, f"{label}: difference up to {(r / lim).max():.2f} x the rounding-unit limit"
    else:
        assert mx <= 30.0, f"{label}: error/tol p99 {p99:.2f} max {mx:.2f}"
    return p99, mx
