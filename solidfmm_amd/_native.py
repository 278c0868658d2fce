"""ctypes binding of libsolidfmm_gpu.so (the C ABI of include/solidfmm_gpu.h).

The library is built in-tree by solidfmm_amd/build.py.  There is no CPU
fallback: if the library is missing or no GPU is visible, calls raise.
"""
import ctypes
import os

import numpy as np

from .workload import ITEM_DTYPE

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libsolidfmm_gpu.so")

SFMM_OK = 0
SFMM_ERR_LOGIC = -1
SFMM_ERR_RANGE = -2
SFMM_ERR_OVERFLOW = -3
SFMM_ERR_ALLOC = -4
SFMM_ERR_HIP = -5
SFMM_ERR_UNSUPPORTED = -6

SFMM_FP64 = 0
SFMM_FP32 = 1
SFMM_SPLIT32 = 2
PRECISIONS = {"fp64": SFMM_FP64, "fp32": SFMM_FP32, "split32": SFMM_SPLIT32}

KERNELS_AUTO, KERNELS_LANE, KERNELS_MFMA, KERNELS_GENERIC = 0, 1, 2, 3


def swap_block(n, part):
    """The folded swap block of degree n (part 0 real, 1 imaginary) as the
    kernels use it; host only."""
    out = np.zeros((n + 1) * (n + 1))
    rc = load().sfmm_swap_block(n, part, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    _check(rc, None, "sfmm_swap_block")
    return out.reshape(n + 1, n + 1)

# Every function include/solidfmm_gpu.h declares, with its ctypes signature.
_c = ctypes
_vp = _c.c_void_p
SIGNATURES = {
    "sfmm_max_order": (_c.c_size_t, []),
    "sfmm_create": (_c.c_int, [_c.POINTER(_vp), _c.c_int, _c.c_size_t, _c.c_int]),
    "sfmm_destroy": (None, [_vp]),
    "sfmm_order": (_c.c_size_t, [_vp]),
    "sfmm_last_error": (_c.c_char_p, [_vp]),
    "sfmm_set_kernels": (_c.c_int, [_vp, _c.c_int]),
    "sfmm_plan_create": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _c.POINTER(_vp)]),
    "sfmm_plan_prepare": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _c.POINTER(_vp)]),
    "sfmm_plan_bind": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp]),
    "sfmm_plan_device_bytes": (_c.c_size_t, [_vp]),
    "sfmm_plan_host_bytes": (_c.c_size_t, [_vp]),
    "sfmm_plan_bind_staged": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp]),
    "sfmm_plan_copy_tables": (_c.c_int, [_vp, _vp, _c.c_size_t]),
    "sfmm_plan_bind_pinned": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp, _vp]),
    "sfmm_device": (_c.c_int, [_vp]),
    "sfmm_server_create": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp, _c.c_size_t, _c.c_size_t, _c.c_int, _c.c_size_t,
                                      _c.c_int, _c.POINTER(_vp)]),
    "sfmm_server_submit": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp, _c.c_size_t,
                                      _c.POINTER(_c.c_uint64)]),
    "sfmm_server_submit_at": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp, _c.c_size_t,
                                         _c.c_double, _c.POINTER(_c.c_uint64)]),
    "sfmm_server_drain": (_c.c_int, [_vp]),
    "sfmm_server_request_info": (_c.c_int, [_vp, _c.c_uint64, _vp, _c.POINTER(_c.c_int)]),
    "sfmm_server_counters": (_c.c_int, [_vp, _vp]),
    "sfmm_server_clock_ms": (_c.c_double, [_vp]),
    "sfmm_server_destroy": (None, [_vp]),
    "sfmm_staging_create": (_c.c_int, [_vp, _c.POINTER(_vp)]),
    "sfmm_staging_destroy": (None, [_vp]),
    "sfmm_staging_reserve": (_c.c_int, [_vp, _c.c_size_t, _c.c_size_t, _c.POINTER(_vp), _c.POINTER(_vp)]),
    "sfmm_staging_translate": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _c.c_size_t, _c.c_size_t]),
    "sfmm_bounding_sphere": (_c.c_int, [_c.c_size_t, _vp, _vp, _vp, _vp]),
    "sfmm_bounding_sphere_of_spheres": (_c.c_int, [_c.c_size_t, _vp, _vp]),
    "sfmm_bounding_spheres_of_ranges": (_c.c_int, [_vp, _c.c_size_t, _c.c_size_t, _vp, _vp, _c.c_size_t, _vp,
                                                   _c.c_int]),
    "sfmm_cellset_create": (_c.c_int, [_vp, _vp, _c.c_size_t, _c.c_size_t, _c.POINTER(_vp)]),
    "sfmm_cellset_destroy": (None, [_vp]),
    "sfmm_p2m": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp]),
    "sfmm_l2p": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp, _vp, _vp]),
    "sfmm_plan_destroy": (None, [_vp]),
    "sfmm_plan_launches": (_c.c_size_t, [_vp]),
    "sfmm_plan_extents": (None, [_vp, _c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)]),
    "sfmm_plan_execute": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp]),
    "sfmm_launch": (_c.c_int, [_vp, _c.c_int, _c.c_uint32, _c.c_uint32, _vp, _c.c_size_t, _vp, _vp, _vp]),
    "sfmm_translate_host": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp, _c.c_size_t]),
    "sfmm_translate_noalloc": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _vp, _c.c_size_t, _vp, _c.c_size_t,
                                          _vp]),
    "sfmm_swap_block": (_c.c_int, [_c.c_int, _c.c_int, _c.POINTER(_c.c_double)]),
    "sfmm_probe_arith": (_c.c_int, [_c.c_int, _c.c_int, _c.POINTER(_c.c_double)]),
    "sfmm_synchronize": (_c.c_int, [_vp, _vp]),
    "sfmm_plan_launch_info": (_c.c_int, [_vp, _c.c_size_t, _c.POINTER(_c.c_int), _c.POINTER(_c.c_uint32),
                                         _c.POINTER(_c.c_int), _c.POINTER(_c.c_size_t)]),
    "sfmm_plan_capture": (_c.c_int, [_vp, _vp, _c.c_size_t, _vp, _c.c_size_t]),
    "sfmm_plan_replay": (_c.c_int, [_vp, _vp]),
    "sfmm_autotune": (_c.c_int, [_vp, _c.c_int, _vp, _c.c_size_t, _vp]),
    "sfmm_get_tuning": (_c.c_int, [_vp, _c.c_int, _c.POINTER(_c.c_int), _c.c_size_t]),
    "sfmm_set_tuning": (_c.c_int, [_vp, _c.c_int, _c.c_uint32, _c.c_int]),
}


class SolidfmmError(RuntimeError):
    """Base class; subclasses mirror the reference's exception types."""


class LogicError(SolidfmmError, ValueError):
    """std::logic_error in the reference."""


class OutOfRangeError(SolidfmmError, IndexError):
    """std::out_of_range in the reference."""


class OrderOverflowError(SolidfmmError, OverflowError):
    """std::overflow_error in the reference."""


class DeviceAllocError(SolidfmmError, MemoryError):
    """std::bad_alloc in the reference."""


class HipError(SolidfmmError):
    """A HIP runtime call failed."""


class UnsupportedError(SolidfmmError, NotImplementedError):
    """A valid request this library has no kernel for."""


_ERRORS = {
    SFMM_ERR_LOGIC: LogicError,
    SFMM_ERR_RANGE: OutOfRangeError,
    SFMM_ERR_OVERFLOW: OrderOverflowError,
    SFMM_ERR_ALLOC: DeviceAllocError,
    SFMM_ERR_HIP: HipError,
    SFMM_ERR_UNSUPPORTED: UnsupportedError,
}

_lib = None


def load(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} is missing: run `python -m solidfmm_amd.build` first")
        lib = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def _check(rc, ctx=None, what=""):
    if rc == SFMM_OK:
        return
    msg = ""
    if ctx is not None:
        raw = load().sfmm_last_error(ctx)
        msg = raw.decode() if raw else ""
    raise _ERRORS.get(rc, SolidfmmError)(f"{what}: {msg} (code {rc})")


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)  # torch.cuda.Stream


def _dev_ptr(t, dtype="float64"):
    """(pointer, length) of a contiguous device tensor of reals."""
    import torch
    assert isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()
    assert t.dtype == getattr(torch, dtype)
    return t.data_ptr(), t.numel()


class Context:
    """Device context for orders up to P (operator_handle in the reference).

    ``precision``: SFMM_FP64 (default), SFMM_FP32 or SFMM_SPLIT32 (bulk
    arithmetic in FP32 or double-word FP32; FP64 pools, geometry and
    accumulation; orders up to 24)."""

    def __init__(self, P, device=0, precision=SFMM_FP64):
        lib = load()
        h = ctypes.c_void_p()
        rc = lib.sfmm_create(ctypes.byref(h), device, P, precision)
        _check(rc, None, f"sfmm_create(P={P})")
        self._h = h
        self.P = P
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            load().sfmm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def plan(self, op, items):
        return Plan(self, op, items)

    def set_kernels(self, family):
        """Kernel family for plans made afterwards: KERNELS_AUTO, _LANE, _MFMA
        or _GENERIC."""
        _check(load().sfmm_set_kernels(self._h, family), self._h, "sfmm_set_kernels")

    def autotune(self, op, items, stream=None):
        """Time every kernel family on each order bucket of ``items`` and keep
        the fastest for plans made afterwards (blocking)."""
        items = np.ascontiguousarray(items, dtype=ITEM_DTYPE)
        rc = load().sfmm_autotune(self._h, op, items.ctypes.data, len(items), _stream_ptr(stream))
        _check(rc, self._h, "sfmm_autotune")

    def tuning(self, op, upto=None):
        """Family used per order bucket (index = max(p_in, p_out))."""
        n = (upto or self.P) + 1
        out = (ctypes.c_int * n)()
        _check(load().sfmm_get_tuning(self._h, op, out, n), self._h, "sfmm_get_tuning")
        return list(out)

    def set_tuning(self, op, order, family):
        _check(load().sfmm_set_tuning(self._h, op, order, family), self._h, "sfmm_set_tuning")

    def launch(self, op, p_in, p_out, dev_items, n, src, dst, stream=None):
        """Unchecked launch over device-resident items (``dev_items`` is a uint8
        device tensor holding ``n`` items)."""
        sp, _ = _dev_ptr(src)
        dp, _ = _dev_ptr(dst)
        rc = load().sfmm_launch(self._h, op, p_in, p_out, dev_items.data_ptr(), n, sp, dp, _stream_ptr(stream))
        _check(rc, self._h, "sfmm_launch")

    def translate_host(self, op, items, src, dst):
        """dst += T(src) with host (numpy) pools; blocking."""
        items = np.ascontiguousarray(items, dtype=ITEM_DTYPE)
        assert src.dtype == np.float64 and dst.dtype == np.float64 and dst.flags["C_CONTIGUOUS"]
        src = np.ascontiguousarray(src)
        rc = load().sfmm_translate_host(self._h, op, items.ctypes.data, len(items), src.ctypes.data, src.size,
                                        dst.ctypes.data, dst.size)
        _check(rc, self._h, "sfmm_translate_host")
        return dst

    def synchronize(self, stream=None):
        _check(load().sfmm_synchronize(self._h, _stream_ptr(stream)), self._h, "sfmm_synchronize")


class Plan:
    """A validated translation list resident on the device.

    With ``prepared=True`` the plan is only validated and grouped on the host
    (sfmm_plan_prepare); ``bind(dev_ptr, nbytes, stream)`` then copies its
    tables into caller device memory of ``device_bytes`` bytes
    asynchronously (sfmm_plan_bind)."""

    def __init__(self, ctx, op, items, prepared=False):
        items = np.ascontiguousarray(items, dtype=ITEM_DTYPE)
        self.ctx = ctx
        self.op = op
        self.n = len(items)
        h = ctypes.c_void_p()
        f = load().sfmm_plan_prepare if prepared else load().sfmm_plan_create
        rc = f(ctx.handle, op, items.ctypes.data, len(items), ctypes.byref(h))
        _check(rc, ctx.handle, "sfmm_plan_prepare" if prepared else "sfmm_plan_create")
        self._h = h

    @property
    def device_bytes(self):
        """Device memory the plan needs (items, and reduction tables plus
        scratch when targets repeat)."""
        return load().sfmm_plan_device_bytes(self._h)

    def bind(self, ptr, nbytes, stream=None):
        _check(load().sfmm_plan_bind(self._h, ptr, nbytes, _stream_ptr(stream)), self.ctx.handle, "sfmm_plan_bind")

    @property
    def host_bytes(self):
        """Bytes of the tables a bind copies (sfmm_plan_host_bytes)."""
        return load().sfmm_plan_host_bytes(self._h)

    def bind_staged(self, ptr, nbytes, host_ptr, host_nbytes, stream=None):
        """bind() through caller pinned host memory (sfmm_plan_bind_staged)."""
        _check(load().sfmm_plan_bind_staged(self._h, ptr, nbytes, host_ptr, host_nbytes, _stream_ptr(stream)),
               self.ctx.handle, "sfmm_plan_bind_staged")

    def copy_tables(self, host_ptr, host_nbytes):
        """Host half of bind_staged (sfmm_plan_copy_tables)."""
        _check(load().sfmm_plan_copy_tables(self._h, host_ptr, host_nbytes), self.ctx.handle,
               "sfmm_plan_copy_tables")

    def bind_pinned(self, ptr, nbytes, host_ptr, stream=None):
        """Device half of bind_staged (sfmm_plan_bind_pinned)."""
        _check(load().sfmm_plan_bind_pinned(self._h, ptr, nbytes, host_ptr, _stream_ptr(stream)), self.ctx.handle,
               "sfmm_plan_bind_pinned")

    @property
    def launches(self):
        return load().sfmm_plan_launches(self._h)

    @property
    def extents(self):
        a, b = ctypes.c_size_t(), ctypes.c_size_t()
        load().sfmm_plan_extents(self._h, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def execute(self, src, dst, stream=None):
        """dst += T(src); src, dst are contiguous float64 device tensors."""
        sp, sn = _dev_ptr(src)
        dp, dn = _dev_ptr(dst)
        rc = load().sfmm_plan_execute(self._h, sp, sn, dp, dn, _stream_ptr(stream))
        _check(rc, self.ctx.handle, "sfmm_plan_execute")

    def launch_info(self):
        """[(family, order bound, per-item orders, items)] per launch."""
        out = []
        for i in range(self.launches):
            f, o, r, c = ctypes.c_int(), ctypes.c_uint32(), ctypes.c_int(), ctypes.c_size_t()
            rc = load().sfmm_plan_launch_info(self._h, i, ctypes.byref(f), ctypes.byref(o), ctypes.byref(r),
                                              ctypes.byref(c))
            _check(rc, self.ctx.handle, "sfmm_plan_launch_info")
            out.append((f.value, o.value, bool(r.value), c.value))
        return out

    def capture(self, src, dst):
        """Capture the plan on these pools into a HIP graph (see replay)."""
        sp, sn = _dev_ptr(src)
        dp, dn = _dev_ptr(dst)
        _check(load().sfmm_plan_capture(self._h, sp, sn, dp, dn), self.ctx.handle, "sfmm_plan_capture")
        self._graph_pools = (src, dst)  # keep the captured pools alive

    def replay(self, stream=None):
        """dst += T(src) on the captured pools, as one graph launch."""
        _check(load().sfmm_plan_replay(self._h, _stream_ptr(stream)), self.ctx.handle, "sfmm_plan_replay")

    def close(self):
        if getattr(self, "_h", None):
            load().sfmm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


CELL_DTYPE = np.dtype([("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("offset", "<u8"), ("first", "<u4"),
                       ("count", "<u4"), ("order", "<u4"), ("dim", "<u4")])
assert CELL_DTYPE.itemsize == 48


class CellSet:
    """Validated device-resident cells for P2M / L2P (sfmm_cellset)."""

    def __init__(self, ctx, cells, n_points):
        cells = np.ascontiguousarray(cells, dtype=CELL_DTYPE)
        self.ctx = ctx
        self.dim = int(max(cells["dim"][0], 1)) if len(cells) else 1
        h = ctypes.c_void_p()
        rc = load().sfmm_cellset_create(ctx.handle, cells.ctypes.data, len(cells), n_points, ctypes.byref(h))
        _check(rc, ctx.handle, "sfmm_cellset_create")
        self._h = h

    def p2m(self, points, pool, stream=None):
        """pool += P2M of points ((n, 3 + dim) float64 device tensor: x, y, z,
        then the dim charges)."""
        pp, pn = _dev_ptr(points)
        mp, mn = _dev_ptr(pool)
        _check(load().sfmm_p2m(self._h, pp, pn // (3 + self.dim), mp, mn, _stream_ptr(stream)), self.ctx.handle,
               "sfmm_p2m")

    def l2p(self, points, pool, phi, grad=None, stream=None):
        """phi += potential, grad += gradient of the local expansions at the points."""
        pp, pn = _dev_ptr(points)
        lp, ln = _dev_ptr(pool)
        fp, fn = _dev_ptr(phi)
        gp, gn = _dev_ptr(grad) if grad is not None else (None, 3 * fn)
        npts = pn // (3 + self.dim)
        if fn < self.dim * npts or gn < 3 * self.dim * npts:
            raise LogicError("sfmm_l2p: phi needs dim reals and grad 3 dim reals per point")
        _check(load().sfmm_l2p(self._h, pp, npts, lp, ln, fp, gp, _stream_ptr(stream)), self.ctx.handle, "sfmm_l2p")

    def close(self):
        if getattr(self, "_h", None):
            load().sfmm_cellset_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def bounding_sphere(points=None, spheres=None):
    """Smallest enclosing ball (x, y, z, r) of an (n, >=3) array of points or
    of an (n, 4) array of balls (solidfmm::bounding_sphere)."""
    out = np.zeros(4)
    if spheres is not None:
        b = np.ascontiguousarray(spheres, dtype=np.float64).reshape(-1, 4)
        _check(load().sfmm_bounding_sphere_of_spheres(len(b), b.ctypes.data, out.ctypes.data), None,
               "sfmm_bounding_sphere_of_spheres")
        return out
    p = np.asarray(points, dtype=np.float64)
    x, y, z = (np.ascontiguousarray(p[:, d]) for d in range(3))
    _check(load().sfmm_bounding_sphere(len(p), x.ctypes.data, y.ctypes.data, z.ctypes.data, out.ctypes.data), None,
           "sfmm_bounding_sphere")
    return out


def bounding_spheres(points, first, count, threads=0):
    """Smallest balls of points[first[s]:first[s]+count[s]] for every s, as an
    (n_sets, 4) array; points is an (n, >=3) float64 array (rows of any
    width, the first three columns are x, y, z)."""
    p = np.ascontiguousarray(points, dtype=np.float64)
    f = np.ascontiguousarray(first, dtype=np.uint64)
    c = np.ascontiguousarray(count, dtype=np.uint64)
    out = np.zeros((len(f), 4))
    _check(load().sfmm_bounding_spheres_of_ranges(p.ctypes.data, p.shape[0], p.shape[1], f.ctypes.data, c.ctypes.data,
                                                  len(f), out.ctypes.data, threads), None,
           "sfmm_bounding_spheres_of_ranges")
    return out


class Staging:
    """Pinned host pools + device pools + stream for host-memory callers
    (sfmm_staging).  reserve() returns numpy views of the pinned source and
    target pools; translate() runs dst += T(src) over them and returns with
    the result in the target view."""

    def __init__(self, ctx):
        self.ctx = ctx
        h = ctypes.c_void_p()
        _check(load().sfmm_staging_create(ctx.handle, ctypes.byref(h)), ctx.handle, "sfmm_staging_create")
        self._h = h

    def reserve(self, src_len, dst_len):
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        _check(load().sfmm_staging_reserve(self._h, src_len, dst_len, ctypes.byref(a), ctypes.byref(b)),
               self.ctx.handle, "sfmm_staging_reserve")
        src = np.ctypeslib.as_array(ctypes.cast(a, ctypes.POINTER(ctypes.c_double)), shape=(max(src_len, 1),))
        dst = np.ctypeslib.as_array(ctypes.cast(b, ctypes.POINTER(ctypes.c_double)), shape=(max(dst_len, 1),))
        return src[:src_len], dst[:dst_len]

    def translate(self, op, items, src_len, dst_len):
        items = np.ascontiguousarray(items, dtype=ITEM_DTYPE)
        _check(load().sfmm_staging_translate(self._h, op, items.ctypes.data, len(items), src_len, dst_len),
               self.ctx.handle, "sfmm_staging_translate")

    def close(self):
        if getattr(self, "_h", None):
            load().sfmm_staging_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
