"""ctypes bindings of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

``Oracle`` wraps liboracle.so, the FP64 restatement of the reference
algorithm (oracle.c).  ``Truth`` wraps libtruth.so, the same operators in
binary128 by two independent routes (truth.c), against which FP64 results'
errors are measured.  ``Reference`` wraps _ref/libsolidfmm_ref.so, the
reference library compiled from its own sources (Makefile, ref_shim.cpp);
it exists only where it was built.

Only tests/, the smoke check and the CPU-baseline leg of bench.py may use
this package.  The product path never imports it.

Functions whose result no independent check pins: none.  The pins are
listed per function in tests/test_oracle.py.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
TRUTH_LIB = os.path.join(HERE, "libtruth.so")
REF_LIB = os.path.join(HERE, "_ref", "libsolidfmm_ref.so")
REF_TREE = "/root/reference"

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


def build(reference=True):
    """Compile liboracle.so, and, when the reference tree is present, the
    reference library and the programs that run the reference's public API
    through both libraries (Makefile target dropin; needs the GPU library
    built first)."""
    targets = ["oracle"]
    if reference and os.path.isdir(os.path.join(REF_TREE, "solidfmm")):
        targets.append("ref")
        if os.path.exists(os.path.join(os.path.dirname(HERE), "solidfmm_amd", "lib", "libsolidfmm_gpu.so")):
            targets.append("dropin")
    subprocess.run(["make", "-s", "-C", HERE, "-j8"] + targets, check=True)


def _ptr(a, dtype=np.float64):
    assert a.dtype == dtype and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp if dtype == np.float64 else _fp)


class Oracle:
    def __init__(self, path=LIB):
        if not os.path.exists(path):
            build(reference=False)
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.orc_translate.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, _dp, _dp]
        L.orc_translate_naive.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, _dp, _dp]
        L.orc_harmonic.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _dp]
        L.orc_swap_entry.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_folded_matrix.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.orc_euler.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_dot.argtypes = [_dp, ctypes.c_int, _dp, ctypes.c_int]
        L.orc_dot.restype = ctypes.c_double

    def translate(self, op, items, src, dst, naive=False):
        """dst += T(src) for every item, in list order (dst is updated in place)."""
        items = np.ascontiguousarray(items)
        assert src.dtype == np.float64 and dst.dtype == np.float64
        f = self.lib.orc_translate_naive if naive else self.lib.orc_translate
        rc = f(op, items.ctypes.data, len(items), _ptr(src), _ptr(dst))
        if rc != 0:
            raise RuntimeError(f"oracle translate failed ({rc})")
        return dst

    def harmonic(self, kind, P, x, y, z):
        """R or S (P(P+1) reals), or dR (3 blocks of P(P+1): d/dx, d/dy, d/dz)."""
        out = np.zeros((3 if kind == "dR" else 1) * P * (P + 1))
        if self.lib.orc_harmonic({"R": 0, "S": 1, "dR": 2}[kind], P, x, y, z, _ptr(out)) != 0:
            raise ValueError("bad harmonic request")
        return out

    def swap_entry(self, n, m, l):
        v = ctypes.c_double()
        if self.lib.orc_swap_entry(n, m, l, ctypes.byref(v)) != 0:
            raise ValueError("bad swap entry")
        return v.value

    def folded(self, kind, n):
        out = np.zeros((n + 1) * (n + 1))
        if self.lib.orc_folded_matrix(kind, n, _ptr(out)) != 0:
            raise ValueError("bad folded matrix")
        return out.reshape(n + 1, n + 1)

    def euler(self, x, y, z, k):
        out = np.zeros(6 * k)
        self.lib.orc_euler(x, y, z, k, _ptr(out))
        return out.reshape(6, k)

    def dot(self, a, PA, b, PB):
        return self.lib.orc_dot(_ptr(a), PA, _ptr(b), PB)


class Truth:
    """Binary128 evaluation of the operators (truth.c), results as hi + lo."""

    NAIVE, ROTATION = 0, 1

    def __init__(self, path=TRUTH_LIB):
        if not os.path.exists(path):
            build(reference=False)
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.truth_translate.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, _dp, ctypes.c_void_p,
                                      ctypes.c_size_t, _dp, _dp, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        L.truth_harmonic.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_double, _dp, _dp]

    def translate(self, op, items, src, dst_len, dst0=None, method=ROTATION, threads=0, magnitude=False):
        """(hi, lo) of dst0 + T(src) over the list, in list order per target.

        magnitude=True also returns M, the rotation route run on magnitudes
        (|A| |v| for every linear step): u * M is the first-order scale of
        the rounding error of any FP64 evaluation of the algorithm."""
        items = np.ascontiguousarray(items)
        src = np.ascontiguousarray(src, dtype=np.float64)
        hi = np.zeros(dst_len)
        lo = np.zeros(dst_len)
        d0 = None
        if dst0 is not None:
            d0 = np.ascontiguousarray(dst0, dtype=np.float64)
            assert len(d0) >= dst_len
        mag = np.zeros(dst_len) if magnitude else None
        threads = threads or min(16, len(os.sched_getaffinity(0)))
        rc = self.lib.truth_translate(op, items.ctypes.data, len(items), _ptr(src),
                                      None if d0 is None else d0.ctypes.data, dst_len, _ptr(hi), _ptr(lo),
                                      None if mag is None else mag.ctypes.data, method, threads)
        if rc != 0:
            raise RuntimeError(f"truth translate failed ({rc})")
        return (hi, lo, mag) if magnitude else (hi, lo)

    def harmonic(self, kind, P, x, y, z):
        hi = np.zeros(P * (P + 1))
        lo = np.zeros(P * (P + 1))
        if self.lib.truth_harmonic({"R": 0, "S": 1}[kind], P, x, y, z, _ptr(hi), _ptr(lo)) != 0:
            raise ValueError("bad harmonic request")
        return hi, lo


class Reference:
    """The compiled reference (single-thread entry points over host solids)."""

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        self.lib = ctypes.CDLL(path)
        L = self.lib
        L.ref_translate_f64.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                        _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, _dp]
        L.ref_translate_f32.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                        _fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_size_t, _dp]
        L.ref_swap_entry.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.ref_harmonic.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_double, ctypes.c_double,
                                   ctypes.c_double, _dp]
        L.ref_bounding_sphere.argtypes = [ctypes.c_int, ctypes.c_size_t, _dp, _dp]
        L.ref_check_errors.argtypes = [ctypes.c_int] + [ctypes.c_size_t] * 5
        L.ref_pool_translate_f64.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                             _dp, _dp, ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _dp,
                                             ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
        L.ref_host_info.argtypes = [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                                    ctypes.c_size_t, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]

    @staticmethod
    def available(path=REF_LIB):
        return os.path.exists(path)

    # Calls of at most this many items keep every item in SIMD lanes 0-11,
    # away from the AVX-512 solid2buf defect (microkernel_avx512f.cpp:2182).
    SAFE_CHUNK = 12

    def translate(self, op, P_op, items, src, dst, threads=1, repeats=1, chunk=SAFE_CHUNK):
        """dst += T(src) through the reference; returns the best wall time (s).

        ``chunk`` = 0 issues the whole list in one call (full 16-lane
        batches, as the reference's own callers do).
        """
        items = np.ascontiguousarray(items)
        secs = ctypes.c_double()
        if src.dtype == np.float64:
            rc = self.lib.ref_translate_f64(op, P_op, items.ctypes.data, len(items), _ptr(src), _ptr(dst),
                                            threads, repeats, chunk, ctypes.byref(secs))
        else:
            rc = self.lib.ref_translate_f32(op, P_op, items.ctypes.data, len(items), _ptr(src, np.float32),
                                            _ptr(dst, np.float32), threads, repeats, chunk, ctypes.byref(secs))
        if rc != 0:
            raise RuntimeError(f"reference translate failed ({rc})")
        return secs.value

    def pool_translate(self, op, P_op, items, src, dst, threads=0, warmup=1024, repeats=5):
        """The reference's batch protocol on a pinned worker pool (one thread
        per physical core of the affinity mask, or `threads`): warm-up on the
        first `warmup` items, then `repeats` timed calls over the list.
        Returns (seconds per call, workers, workers pinned)."""
        items = np.ascontiguousarray(items)
        secs = np.zeros(repeats)
        used, pinned = ctypes.c_int(), ctypes.c_int()
        rc = self.lib.ref_pool_translate_f64(op, P_op, items.ctypes.data, len(items), _ptr(src), _ptr(dst),
                                             threads, warmup, repeats, _ptr(secs), ctypes.byref(used),
                                             ctypes.byref(pinned))
        if rc != 0:
            raise RuntimeError(f"reference pool translate failed ({rc})")
        return secs, used.value, pinned.value

    def host_info(self):
        """cpu_model, nproc, cpus in the affinity mask, physical cores in it,
        and the microkernel the reference selects here (with its tile)."""
        model = ctypes.create_string_buffer(256)
        kern = ctypes.create_string_buffer(64)
        ints = [ctypes.c_int() for _ in range(5)]
        self.lib.ref_host_info(model, 256, ctypes.byref(ints[0]), ctypes.byref(ints[1]), ctypes.byref(ints[2]),
                               kern, 64, ctypes.byref(ints[3]), ctypes.byref(ints[4]))
        return {"cpu_model": model.value.decode(), "nproc": ints[0].value, "affinity_cpus": ints[1].value,
                "physical_cores": ints[2].value, "microkernel": kern.value.decode(),
                "microkernel_tile": [ints[3].value, ints[4].value]}

    def swap_entry(self, n, m, l):
        v = ctypes.c_double()
        self.lib.ref_swap_entry(n, m, l, ctypes.byref(v))
        return v.value

    def harmonic(self, kind, P, x, y, z):
        out = np.zeros((3 if kind == "dR" else 1) * P * (P + 1))
        self.lib.ref_harmonic({"R": 0, "S": 1, "dR": 2}[kind], P, x, y, z, _ptr(out))
        return out

    def bounding_sphere(self, points=None, spheres=None):
        """The reference's smallest enclosing ball (x, y, z, r) of points
        (n, 3) or of balls (n, 4)."""
        out = np.zeros(4)
        if spheres is not None:
            data = np.ascontiguousarray(spheres, dtype=np.float64).reshape(-1)
            kind, n = 1, len(data) // 4
        else:
            p = np.asarray(points, dtype=np.float64)
            data = np.ascontiguousarray(p[:, :3].T).reshape(-1)
            kind, n = 0, len(p)
        if self.lib.ref_bounding_sphere(kind, n, _ptr(data), _ptr(out)) != 0:
            raise ValueError("reference bounding_sphere raised")
        return out

    def check_errors(self, op, P_op, P_in, dim_in, P_out, dim_out):
        return self.lib.ref_check_errors(op, P_op, P_in, dim_in, P_out, dim_out)


def degree_errors(got, ref, P):
    """Per-degree normwise relative error of solids (rows) of order P.

    e_n = ||got_n - ref_n||_2 / ||ref_n||_2 over Re and Im of m = 0..n
    (the 2(n+1) reals of degree n).  Returns an array (count, P); degrees
    whose reference norm is zero report the absolute error instead.
    """
    got = np.asarray(got).reshape(-1, P * (P + 1))
    ref = np.asarray(ref).reshape(-1, P * (P + 1))
    out = np.zeros((got.shape[0], P))
    for n in range(P):
        idx = list(range(n * (n + 1), (n + 1) * (n + 2)))
        d = np.linalg.norm(got[:, idx] - ref[:, idx], axis=1)
        r = np.linalg.norm(ref[:, idx], axis=1)
        out[:, n] = np.where(r > 0, d / np.where(r > 0, r, 1.0), d)
    return out


def truth_degree_errors(got, hi, lo, P):
    """degree_errors against a binary128 truth given as hi + lo: the
    difference is formed as (got - hi) - lo, so the truth's own rounding to
    FP64 does not enter the error."""
    got = np.asarray(got).reshape(-1, P * (P + 1))
    hi = np.asarray(hi).reshape(-1, P * (P + 1))
    lo = np.asarray(lo).reshape(-1, P * (P + 1))
    out = np.zeros((got.shape[0], P))
    for n in range(P):
        idx = slice(n * (n + 1), (n + 1) * (n + 2))
        d = np.linalg.norm((got[:, idx] - hi[:, idx]) - lo[:, idx], axis=1)
        r = np.linalg.norm(hi[:, idx], axis=1)
        out[:, n] = np.where(r > 0, d / np.where(r > 0, r, 1.0), d)
    return out


def error_scale(mag, hi, P):
    """Per-item, per-degree first-order rounding scale of the rotation
    algorithm: u ||M_n||_2 / ||truth_n||_2 with M the magnitude propagation
    (Truth.translate(magnitude=True)) and u = 2^-53."""
    mag = np.asarray(mag).reshape(-1, P * (P + 1))
    hi = np.asarray(hi).reshape(-1, P * (P + 1))
    out = np.zeros((mag.shape[0], P))
    for n in range(P):
        idx = slice(n * (n + 1), (n + 1) * (n + 2))
        m = np.linalg.norm(mag[:, idx], axis=1)
        r = np.linalg.norm(hi[:, idx], axis=1)
        out[:, n] = 2.0 ** -53 * np.where(r > 0, m / np.where(r > 0, r, 1.0), m)
    return out


def mixed_truth_errors(got, hi, lo, items):
    """Per-item error against the truth (max over degrees) for mixed orders."""
    e = np.zeros(len(items))
    for po in np.unique(items["p_out"]):
        idx = np.nonzero(items["p_out"] == po)[0]
        cols = items["dst"][idx].astype(np.int64)[:, None] + np.arange(int(po) * (int(po) + 1))
        e[idx] = truth_degree_errors(got[cols], hi[cols], lo[cols], int(po)).max(axis=1)
    return e


def mixed_errors(got, want, items):
    """Per-item error (max over degrees of the per-degree normwise relative
    error) for lists of mixed orders; target blocks sit at items["dst"]."""
    e = np.zeros(len(items))
    for po in np.unique(items["p_out"]):
        idx = np.nonzero(items["p_out"] == po)[0]
        cols = items["dst"][idx].astype(np.int64)[:, None] + np.arange(int(po) * (int(po) + 1))
        e[idx] = degree_errors(got[cols], want[cols], int(po)).max(axis=1)
    return e
