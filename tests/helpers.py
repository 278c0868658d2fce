"""Shared pieces of the test-suite: tolerances and golden vectors."""
import os

import numpy as np

GOLDEN_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def tol_fp64(p_in, p_out):
    """Per-degree normwise relative tolerance of FP64 results.

    1e-13 up to order 10, which covers the P=10 headline.  Higher orders are
    limited by the conditioning of the translations themselves, not by the
    implementation: on the golden cases the reference, the oracle and the
    naive O(P^4) formulas already disagree by up to 8e-13 (M2L 12 -> 3,
    where terms of size ~1e3 cancel to a degree-0 result of size ~5),
    2e-12 (P=16), 1.5e-11 (P=20), 1e-10 (P=24) and 4e-10 (M2L from
    P_in=20).  The tolerance is a small multiple of that floor; when the
    source has more degrees than the target (p_in > p_out) the low output
    degrees cancel harder, and the tolerance is ten times larger.
    """
    P = max(p_in, p_out)
    if P <= 10:
        t = 1e-13
    elif P <= 12:
        t = 2e-12
    elif P <= 16:
        t = 1e-11
    elif P <= 24:
        t = 2e-9
    else:
        # random O(1) coefficients at every degree: the oracle and the naive
        # formulas already differ by 7e-9 at P=30 and 2e-6 at P=40
        t = 5e-8 if P <= 30 else 1e-4
    return 10 * t if p_in > p_out else t


def assert_parity(got, want, p_in, p_out, tol=None, label=""):
    """Per-translation error e = max over degrees of the per-degree normwise
    relative error.  Requires p99(e) <= tol and max(e) <= 10 tol.

    The max gets headroom because a random input can make a degree's result
    small by cancellation, which inflates its relative error for any FP64
    implementation: over 200 random P=10 M2L translations the compiled
    reference and the oracle (same algorithm, both FP64) differ by up to
    4.4e-13 at the worst item while the 99th percentile is 5e-14.
    """
    import oracle
    tol = tol_fp64(p_in, p_out) if tol is None else tol
    e = oracle.degree_errors(got, want, p_out).max(axis=1)
    p99 = float(np.percentile(e, 99)) if len(e) else 0.0
    mx = float(e.max()) if len(e) else 0.0
    assert p99 <= tol and mx <= 10 * tol, f"{label}: p99 {p99:.2e} max {mx:.2e} tol {tol:.1e}"
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


def assert_mixed_parity(got, want, items, scale=1.0, label=""):
    """assert_parity for mixed orders: each item against tol_fp64 of its own
    orders (times `scale`); p99 of the ratios <= 1 and max <= 10."""
    import oracle
    e = oracle.mixed_errors(got, want, items)
    tol = np.array([tol_fp64(int(a), int(b)) for a, b in zip(items["p_in"], items["p_out"])]) * scale
    ratio = e / tol
    p99 = float(np.percentile(ratio, 99)) if len(ratio) else 0.0
    mx = float(ratio.max()) if len(ratio) else 0.0
    assert p99 <= 1.0 and mx <= 10.0, f"{label}: error/tol p99 {p99:.2f} max {mx:.2f}"
    return p99, mx
