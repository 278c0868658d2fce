"""The reference's own physics tests, compiled unchanged against the drop-in
C++ API (include/solidfmm/*.hpp + libsolidfmm_gpu.so), print the same
convergence tables as when compiled against the reference library.

oracle/Makefile (target `dropin`) builds both programs from the reference's
tests/test_{m2l,m2m,l2l}.cpp in the build container; each prints, per order
P (float 0..18, double 0..40), the relative error of an expansion of a unit
point charge against 1/r after translating it with 32 identical items into
one target (list accumulation, lane padding).  Needs a GPU for the drop-in
side; skipped where the programs were not built.
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
LINE = re.compile(r"P =\s*(\d+)\.(.*)$")
NUM = re.compile(r"([A-Za-z][A-Za-z0-9 \-]*?):\s*([-+0-9.eE]+|-?nan|-?inf)")


def run(path):
    r = subprocess.run([path], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    tables, section = {}, None
    for line in r.stdout.splitlines():
        if "single precision" in line:
            section = "float"
        elif "double precision" in line:
            section = "double"
        m = LINE.search(line)
        if m and section:
            P = int(m.group(1))
            for name, val in NUM.findall(m.group(2)):
                tables[(section, P, name.strip())] = float(val.rstrip("."))
    return tables


@pytest.mark.parametrize("name", ["test_m2l", "test_m2m", "test_l2l"])
def test_reference_program_through_dropin(name):
    ref = os.path.join(REF_DIR, "ref_" + name)
    mine = os.path.join(REF_DIR, "dropin_" + name)
    if not (os.path.exists(ref) and os.path.exists(mine)):
        pytest.skip("reference test programs not built (oracle/Makefile target dropin)")
    print(f"running {name}", flush=True)
    import math
    a, b = run(ref), run(mine)
    assert a.keys() == b.keys() and len(a) > 100
    # float noise floor of each column: the reference's smallest float value
    floor = {}
    for (section, P, name), v in a.items():
        if section == "float":
            floor[name] = min(floor.get(name, 1.0), v)
    for key, va in a.items():
        vb = b[key]
        section, P, name = key
        # Each table prints r |approx - 1/r| (the potential scaled to 1) with
        # 5 significant digits, so two printed values of the same error
        # differ by at most one unit of the last printed digit.  Beyond that
        # the approximations themselves differ by rounding: FP64 (double), or
        # the reference's float arithmetic, a few float ulps of the unit
        # potential, at most 16 or twice the column's float floor (float:
        # the drop-in translates in FP64 and rounds to float once).
        quantum = 10.0 ** (math.floor(math.log10(abs(va))) - 4) if va > 0 else 0.0
        noise = 1e-13 if section == "double" else max(16 * 2.0 ** -24, 2 * floor[name])
        assert abs(va - vb) <= quantum + noise, (key, va, vb)


# ---------------------------------------------------------------------------
# The whole public translation API: oracle/dropin_check.cpp built against
# both libraries (translation_info, *_unchecked, classic, and the drop-in's
# multithreaded_translator), double and float, dim 1 and 3, repeated targets.
# ---------------------------------------------------------------------------

FORMS = ("ti", "unchecked", "classic", "mt", "mt_unchecked")


def read_records(path):
    import numpy as np
    out = {}
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        n = int.from_bytes(data[pos:pos + 4], "little")
        name = data[pos + 4:pos + 4 + n].decode()
        pos += 4 + n
        cnt = int.from_bytes(data[pos:pos + 8], "little")
        pos += 8
        out[name] = np.frombuffer(data[pos:pos + 8 * cnt], dtype="<f8").copy()
        pos += 8 * cnt
    return out


@pytest.fixture(scope="module")
def api_runs(tmp_path_factory):
    ref = os.path.join(REF_DIR, "ref_dropin_check")
    mine = os.path.join(REF_DIR, "dropin_dropin_check")
    if not (os.path.exists(ref) and os.path.exists(mine)):
        pytest.skip("API check programs not built (oracle/Makefile target dropin)")
    out = {}
    for who, path in (("ref", ref), ("mine", mine)):
        f = str(tmp_path_factory.mktemp(who) / "records.bin")
        print(f"running {os.path.basename(path)}", flush=True)
        r = subprocess.run([path, f], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        errs = {}
        for line in r.stdout.splitlines():
            if line.startswith("err "):
                _, name, code, unchanged = line.split()
                errs[name] = (int(code), int(unchanged))
        out[who] = (read_records(f), errs)
    return out


def block_degree_errors(got, want, orders, dim):
    """Per (target, dimension, degree): ||got_n - want_n|| / ||want_n||."""
    import numpy as np
    e, pos = [], 0
    for P in orders.astype(int):
        for _ in range(dim):
            blk = slice(pos, pos + P * (P + 1))
            g, w = got[blk], want[blk]
            for n in range(P):
                sl = slice(n * (n + 1), (n + 1) * (n + 2))
                r = np.linalg.norm(w[sl])
                e.append(np.linalg.norm(g[sl] - w[sl]) / (r if r > 0 else 1.0))
            pos += P * (P + 1)
    assert pos == len(got)
    return np.array(e)


@pytest.mark.parametrize("dim", [1, 3])
@pytest.mark.parametrize("op", ["m2m", "m2l", "l2l"])
def test_api_forms_double(api_runs, op, dim):
    """Every double form of the drop-in against the reference's
    translation_info result on the same list (48 items into 12 targets of
    orders 1..12)."""
    import numpy as np
    from tests.helpers import tol_fp64
    (ref, _), (mine, _) = api_runs["ref"], api_runs["mine"]
    base = f"f64_{op}_d{dim}"
    assert np.array_equal(ref[base + "_initial"], mine[base + "_initial"])
    orders = ref[base + "_orders"]
    want = ref[base + "_ti"]
    assert not np.array_equal(want, ref[base + "_initial"])
    for form in FORMS:
        e = block_degree_errors(mine[f"{base}_{form}"], want, orders, dim)
        print(f"  {base} {form}: p50 {np.percentile(e, 50):.2e} max {e.max():.2e}", flush=True)
        # targets sum four translations of mixed orders up to 12
        assert np.percentile(e, 50) <= tol_fp64(12, 12) and e.max() <= 30 * tol_fp64(12, 12), (form, e.max())


@pytest.mark.parametrize("dim", [1, 3])
@pytest.mark.parametrize("op", ["m2m", "m2l", "l2l"])
def test_api_forms_float(api_runs, op, dim):
    """Float solids: the drop-in translates in FP64 and rounds each target
    value to float once, so against the double result on the same (float)
    inputs every value is within one float rounding plus FP64 noise, and the
    per-degree errors stay within the reference float path's own."""
    import numpy as np
    (ref, _), (mine, _) = api_runs["ref"], api_runs["mine"]
    base = f"f32_{op}_d{dim}"
    exact = ref[f"f64p10_{op}_d{dim}_ti"]
    orders = ref[base + "_orders"]
    e_ref = block_degree_errors(ref[base + "_ti"], exact, orders, dim)
    for form in FORMS:
        got = mine[f"{base}_{form}"]
        # per value: half a float ulp of the value plus FP64-level noise on
        # the scale of its degree block
        pos, scale = 0, np.zeros_like(exact)
        for P in orders.astype(int):
            for _ in range(dim):
                for n in range(P):
                    sl = slice(pos + n * (n + 1), pos + (n + 1) * (n + 2))
                    scale[sl] = np.abs(exact[sl]).max()
                pos += P * (P + 1)
        bound = 2.0 ** -24 * np.abs(exact) + 1e-12 * scale
        assert np.all(np.abs(got - exact) <= bound), (form, float((np.abs(got - exact) / bound).max()))
        e = block_degree_errors(got, exact, orders, dim)
        print(f"  {base} {form}: p50 {np.percentile(e, 50):.2e} max {e.max():.2e} | reference float "
              f"p50 {np.percentile(e_ref, 50):.2e} max {e_ref.max():.2e}", flush=True)
        assert np.percentile(e, 50) <= np.percentile(e_ref, 50) and e.max() <= e_ref.max()


def test_api_errors(api_runs):
    """The checked forms throw what the reference throws, before writing
    anything: order above the operator's (out_of_range), dimension mismatch
    and P = 0 (logic_error), P above 86 / 18 (overflow_error)."""
    (_, ref), (_, mine) = api_runs["ref"], api_runs["mine"]
    assert ref.keys() == mine.keys() and len(ref) >= 20
    for name, (code, unchanged) in ref.items():
        assert mine[name][0] == code, (name, mine[name], code)
        if code:
            assert mine[name][1] == 1 and unchanged == 1, name
    assert ref["f64_handle_above_max"][0] == 3 and ref["f32_handle_above_max"][0] == 3
    assert ref["f64_dimension_mismatch"][0] == 1 and ref["f64_source_order_above_P"][0] == 2
