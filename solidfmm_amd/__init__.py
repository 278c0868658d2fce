"""MI355X-native translation operators (M2M, M2L, L2L) on solid harmonics.

The compute path is hand-written HIP for gfx950 behind a C ABI
(include/solidfmm_gpu.h, built into solidfmm_amd/lib/libsolidfmm_gpu.so).
This package is its Python face: ``Context`` (operator_handle), ``Plan``
(a validated, device-resident translation list) and ``Translator``, which
mirrors the reference's multithreaded_translator batch interface
(translations.hpp:47-87) over device-resident pools.
"""
from . import workload
from .workload import ITEM_DTYPE, M2M, M2L, L2L, solid_size
from ._native import (SFMM_FP64, SFMM_FP32, SFMM_SPLIT32, PRECISIONS)
from ._native import (Context, Plan, CellSet, CELL_DTYPE, Staging, bounding_sphere, bounding_spheres, load, LIB_PATH, SolidfmmError, LogicError, OutOfRangeError,
                      OrderOverflowError, DeviceAllocError, HipError, UnsupportedError)


class Translator:
    """Batch interface of multithreaded_translator<double> over device pools.

    ``m2m/m2l/l2l(items, src_pool, dst_pool)`` validate every item before
    writing anything (translations.cpp:634-741) and apply items that share a
    target in list order; ``*_unchecked`` launch directly and require
    uniform orders and distinct targets.  Pools are contiguous float64
    device tensors; items are a numpy array of ``ITEM_DTYPE``.
    """

    def __init__(self, max_expansion_order, device=0):
        self.ctx = Context(max_expansion_order, device)

    def _run(self, op, items, src, dst, stream=None):
        plan = Plan(self.ctx, op, items)
        plan.execute(src, dst, stream)
        return plan

    def m2m(self, items, src, dst, stream=None):
        return self._run(M2M, items, src, dst, stream)

    def m2l(self, items, src, dst, stream=None):
        return self._run(M2L, items, src, dst, stream)

    def l2l(self, items, src, dst, stream=None):
        return self._run(L2L, items, src, dst, stream)


__all__ = ["SFMM_FP64", "SFMM_FP32", "SFMM_SPLIT32", "PRECISIONS", "workload", "ITEM_DTYPE", "M2M", "M2L", "L2L",
           "solid_size", "Context", "Plan", "CellSet", "CELL_DTYPE", "Staging", "bounding_sphere", "bounding_spheres", "Translator", "load", "LIB_PATH", "SolidfmmError", "LogicError", "OutOfRangeError", "OrderOverflowError",
           "DeviceAllocError", "HipError", "UnsupportedError"]
