"""Continuous-batching request loop for translation requests on one GPU.

A request is one list of translations of one operator (items with pool
offsets) plus the cells it reads and writes.  The loop keeps the coefficient
pools resident on the device and runs requests concurrently on a fixed set
of HIP streams ("slots"):

* Arrival.  Requests arrive at given times (e.g. a Poisson process) and
  wait in arrival order.
* Admission.  The oldest waiting request is admitted when a slot is free,
  its device memory (items, and reduction tables plus per-item scratch
  blocks when its targets repeat) fits in the arena (the HBM budget for
  in-flight requests), and it has no hazard with a request in flight: it writes no
  cell an in-flight request reads or writes, and reads no cell an in-flight
  request writes.  Admission is in arrival order (a blocked head blocks the
  queue), so every cell sees its updates in arrival order and the pools end
  exactly as if the requests had run one by one.
* Preparation.  On arrival a request's plan is validated and grouped on
  the host by a pool of worker threads (sfmm_plan_prepare; the C call
  releases the interpreter lock), while the GPU runs earlier requests.
* Execution.  On admission the prepared tables are copied into the arena
  and the kernels launched on the slot's stream (sfmm_plan_bind: no device
  allocation, nothing waits for the device); an event marks completion.
* Completion.  Events are polled; a completed request releases its slot,
  arena space and hazards, and its latency (completion on the GPU clock
  minus arrival) is recorded.

Requests that arrive while others are on the GPU are thus admitted as slots
and memory free up, and their kernels overlap on the device.  The reference
has no counterpart (its multithreaded_translator, translations.hpp:47-87,
serialises whole batches with a mutex).
"""
import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from .workload import ITEM_DTYPE, M2L, M2M, L2L, uniform01

# pools read / written per operator: 0 = multipole pool, 1 = local pool
POOLS = {M2M: (0, 0), M2L: (0, 1), L2L: (1, 1)}
ITEM_BYTES = ITEM_DTYPE.itemsize
ALIGN = 256


class Request:
    """One translation request: operator, items, cells read / written."""

    __slots__ = ("rid", "op", "items", "src_cells", "dst_cells", "arrival", "admitted", "done", "gpu_ms",
                 "noticed", "slot", "offset", "plan", "future", "ev_start", "ev_end")

    def __init__(self, rid, op, items, src_cells, dst_cells, arrival=0.0):
        self.rid = rid
        self.op = op
        self.items = np.ascontiguousarray(items, dtype=ITEM_DTYPE)
        self.src_cells = np.unique(src_cells)
        self.dst_cells = np.unique(dst_cells)
        self.arrival = float(arrival)  # seconds after the start of the run
        self.admitted = None
        self.done = None      # ms after the start, GPU clock
        self.gpu_ms = None    # kernel time of the request's own launches
        self.noticed = None   # ms after the start when the loop saw completion (host clock)
        self.slot = None
        self.offset = None
        self.plan = None
        self.future = None
        self.ev_start = None
        self.ev_end = None

    @property
    def nbytes(self):
        return len(self.items) * ITEM_BYTES


class Arena:
    """First-fit allocator of byte ranges of a device buffer (the HBM budget
    for in-flight item arrays)."""

    def __init__(self, nbytes):
        self.size = int(nbytes)
        self.free = [(0, self.size)]  # sorted, disjoint (offset, length)
        self.used = 0
        self.peak = 0

    def alloc(self, nbytes):
        need = max(ALIGN, (int(nbytes) + ALIGN - 1) // ALIGN * ALIGN)
        for i, (off, ln) in enumerate(self.free):
            if ln >= need:
                if ln == need:
                    self.free.pop(i)
                else:
                    self.free[i] = (off + need, ln - need)
                self.used += need
                self.peak = max(self.peak, self.used)
                return off, need
        return None

    def release(self, off, need):
        self.used -= need
        self.free.append((off, need))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged


class Hazards:
    """Per-cell counts of in-flight readers and writers, per pool."""

    def __init__(self, ncells):
        self.readers = [np.zeros(ncells, dtype=np.int32) for _ in range(2)]
        self.writers = [np.zeros(ncells, dtype=np.int32) for _ in range(2)]

    def conflict(self, r):
        sp, dp = POOLS[r.op]
        if self.writers[dp][r.dst_cells].any() or self.readers[dp][r.dst_cells].any():
            return True
        return bool(self.writers[sp][r.src_cells].any())

    def acquire(self, r, k=1):
        sp, dp = POOLS[r.op]
        self.readers[sp][r.src_cells] += k
        self.writers[dp][r.dst_cells] += k

    def release(self, r):
        self.acquire(r, -1)


class Server:
    """Continuous batching over `slots` HIP streams with a device arena of
    `budget_bytes` for the requests in flight.  `pools` = (multipole pool, local pool),
    contiguous float64 device tensors."""

    def __init__(self, ctx, pools, ncells, slots=4, budget_bytes=256 << 20, workers=8):
        import torch

        from ._native import Plan
        self._torch = torch
        self._Plan = Plan
        self.ctx = ctx
        self.pools = pools
        self.slots = slots
        self.streams = [torch.cuda.Stream() for _ in range(slots)]
        self.arena_buf = torch.empty(int(budget_bytes), dtype=torch.uint8, device="cuda")
        self.arena = Arena(budget_bytes)
        self.hazards = Hazards(ncells)
        self.workers = workers

    def run(self, requests, realtime=True):
        """Serve `requests` (arrival times in seconds after the start); returns
        statistics.  With realtime=False every request is available at once."""
        torch = self._torch
        pending = deque(sorted(requests, key=lambda r: (r.arrival, r.rid)))
        waiting = deque()
        inflight = []
        free_slots = list(range(self.slots))
        ref = torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        ref.record(torch.cuda.current_stream())
        ref.synchronize()
        t0 = time.perf_counter()
        max_inflight = 0
        blocked = {"prepare": 0, "slot": 0, "arena": 0, "hazard": 0}
        pool = ThreadPoolExecutor(max_workers=self.workers)
        prepare = lambda r: self._Plan(self.ctx, r.op, r.items, prepared=True)  # noqa: E731
        while pending or waiting or inflight:
            now = time.perf_counter() - t0
            while pending and (not realtime or pending[0].arrival <= now):
                r = pending.popleft()
                r.future = pool.submit(prepare, r)
                waiting.append(r)
            progressed = False
            # completions
            still = []
            for r in inflight:
                if r.ev_end.query():
                    self._finish(r, ref)
                    free_slots.append(r.slot)
                    progressed = True
                else:
                    still.append(r)
            inflight = still
            # admission, oldest first
            while waiting:
                r = waiting[0]
                if not r.future.done():
                    blocked["prepare"] += 1
                    break
                if not free_slots:
                    blocked["slot"] += 1
                    break
                if self.hazards.conflict(r):
                    blocked["hazard"] += 1
                    break
                need = r.future.result().device_bytes
                got = self.arena.alloc(need)
                if got is None:
                    if not inflight and self.arena.used == 0:
                        raise MemoryError(f"request {r.rid} ({need} B) exceeds the device budget "
                                          f"({self.arena.size} B)")
                    blocked["arena"] += 1
                    break
                waiting.popleft()
                r.slot = free_slots.pop()
                r.offset = got
                r.admitted = time.perf_counter() - t0
                self._launch(r)
                inflight.append(r)
                progressed = True
            max_inflight = max(max_inflight, len(inflight))
            if not progressed:
                if inflight:
                    time.sleep(20e-6)
                elif pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 1e-3)))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        pool.shutdown()
        return self._stats(requests, wall, max_inflight, blocked)

    def _launch(self, r):
        torch = self._torch
        s = self.streams[r.slot]
        off, need = r.offset
        ptr = self.arena_buf.data_ptr() + off
        sp, dp = POOLS[r.op]
        r.ev_start = torch.cuda.Event(enable_timing=True)
        r.ev_end = torch.cuda.Event(enable_timing=True)
        r.plan = r.future.result()
        r.future = None
        r.ev_start.record(s)
        r.plan.bind(ptr, need, s)
        self.hazards.acquire(r)
        r.plan.execute(self.pools[sp], self.pools[dp], s)
        r.ev_end.record(s)

    def _finish(self, r, ref):
        r.done = ref.elapsed_time(r.ev_end)
        r.gpu_ms = r.ev_start.elapsed_time(r.ev_end)
        self.hazards.release(r)
        self.arena.release(*r.offset)
        r.plan.close()
        r.plan = None

    def _stats(self, requests, wall, max_inflight, blocked):
        lat = np.array([r.done - 1e3 * r.arrival for r in requests])
        n = sum(len(r.items) for r in requests)
        busy = sum(r.gpu_ms for r in requests)
        first = min(1e3 * r.arrival for r in requests)
        last = max(r.done for r in requests)
        span = last - first
        return {"requests": len(requests), "translations": int(n), "wall_s": wall,
                "span_ms": span, "translations_per_s": n / (span * 1e-3) if span > 0 else 0.0,
                "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                               "max": float(lat.max())},
                "max_inflight": max_inflight, "arena_peak_bytes": self.arena.peak,
                "arena_bytes": self.arena.size, "sum_request_gpu_ms": busy,
                "overlap": busy / span if span > 0 else 0.0, "blocked_polls": blocked}


class NativeServer:
    """The same continuous-batching policy run by the library's native loop
    (sfmm_server_*): preparation threads, admission, launches and completion
    polling are C++; Python only submits requests at their arrival times.
    `pools` = (multipole pool, local pool), contiguous float64 device
    tensors."""

    def __init__(self, ctx, pools, ncells, slots=4, budget_bytes=256 << 20, workers=None):
        import ctypes
        import os

        from . import _native as N
        self._c = ctypes
        self._lib = N.load()
        self._check = N._check
        self.ctx = ctx
        self.pools = pools
        h = ctypes.c_void_p()
        rc = self._lib.sfmm_server_create(ctx.handle, pools[0].data_ptr(), pools[0].numel(), pools[1].data_ptr(),
                                          pools[1].numel(), int(ncells), int(slots), int(budget_bytes),
                                          int(workers or max(1, min(16, os.cpu_count() or 1) - 2)),
                                          ctypes.byref(h))
        self._check(rc, ctx.handle, "sfmm_server_create")
        self._h = h

    def submit(self, r, arrival_ms=-1.0):
        """Submit request r now, or held until `arrival_ms` on the server's
        clock; returns its id."""
        c = self._c
        src = np.ascontiguousarray(r.src_cells, dtype=np.uint32)
        dst = np.ascontiguousarray(r.dst_cells, dtype=np.uint32)
        rid = c.c_uint64()
        rc = self._lib.sfmm_server_submit_at(self._h, r.op, r.items.ctypes.data, len(r.items), src.ctypes.data,
                                             len(src), dst.ctypes.data, len(dst), float(arrival_ms), c.byref(rid))
        self._check(rc, self.ctx.handle, "sfmm_server_submit_at")
        return rid.value

    def run(self, requests, realtime=True, lead_ms=None):
        """Replay `requests`: every request is handed to the server up front,
        held until base + its arrival time (or base for all, realtime=False),
        where base lies `lead_ms` ahead so that the submissions finish first;
        wait for all and return the statistics of Server.run.  Latency is
        completion (GPU clock) minus the scheduled arrival."""
        c = self._c
        order = sorted(requests, key=lambda r: (r.arrival, r.rid))
        if lead_ms is None:
            lead_ms = 50.0 + 0.5 * len(order)
        base = self._lib.sfmm_server_clock_ms(self._h) + lead_ms
        rids = [(r, self.submit(r, base + (1e3 * r.arrival if realtime else 0.0))) for r in order]
        t0 = time.perf_counter()
        self._check(self._lib.sfmm_server_drain(self._h), self.ctx.handle, "sfmm_server_drain")
        wall = time.perf_counter() - t0
        times = np.zeros(5)
        status = c.c_int()
        for r, rid in rids:
            self._check(self._lib.sfmm_server_request_info(self._h, rid, times.ctypes.data, c.byref(status)),
                        self.ctx.handle, "sfmm_server_request_info")
            self._check(status.value, self.ctx.handle, f"request {r.rid}")
            r.done = times[2] - base
            r.gpu_ms = times[3]
            r.admitted = (times[1] - base) * 1e-3
            r.noticed = times[4] - base
        cnt = np.zeros(7, dtype=np.uint64)
        self._lib.sfmm_server_counters(self._h, cnt.ctypes.data)
        arr = {id(r): (1e3 * r.arrival if realtime else 0.0) for r in requests}
        lat = np.array([r.done - arr[id(r)] for r in requests])
        n = sum(len(r.items) for r in requests)
        busy = sum(r.gpu_ms for r in requests)
        span = max(r.done for r in requests) - min(arr.values())
        # host clock minus GPU clock (aligned at the server's creation)
        seen = np.array([r.noticed - r.done for r in requests])
        return {"requests": len(requests), "translations": int(n), "wall_s": wall, "span_ms": span,
                "translations_per_s": n / (span * 1e-3) if span > 0 else 0.0,
                "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                               "max": float(lat.max())},
                "max_inflight": int(cnt[0]), "arena_peak_bytes": int(cnt[1]), "arena_bytes": int(cnt[2]),
                "sum_request_gpu_ms": busy, "overlap": busy / span if span > 0 else 0.0,
                "blocked_polls": {"prepare": int(cnt[3]), "slot": int(cnt[4]), "arena": int(cnt[5]),
                                  "hazard": int(cnt[6])},
                "notice_delay_ms": {"p50": float(np.percentile(seen, 50)), "p99": float(np.percentile(seen, 99)),
                                    "max": float(seen.max())}}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sfmm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_requests(passes, seed, lo=1000, hi=50000):
    """Cut each pass of translation_lists() into requests of U[lo, hi]
    translations, cut at target boundaries (a target's items stay in one
    request); a pass smaller than lo is one request."""
    reqs = []
    k = 0
    for op, items, src, dst, _level in passes:
        n = len(items)
        start = 0
        while start < n:
            size = lo + int(uniform01(seed, 200, 1, offset=k)[0] * (hi - lo + 1))
            end = min(n, start + size)
            while end < n and dst[end] == dst[end - 1]:
                end += 1
            reqs.append(Request(len(reqs), op, items[start:end], src[start:end], dst[start:end]))
            start = end
            k += 1
    return reqs


def poisson_arrivals(requests, rate, seed):
    """Arrival times of a Poisson process of `rate` requests per second,
    assigned in list order."""
    u = uniform01(seed, 300, len(requests))
    gaps = -np.log(np.maximum(1.0 - u, 1e-300)) / rate
    t = np.cumsum(gaps) - gaps[0]
    for r, a in zip(requests, t):
        r.arrival = float(a)
    return requests


def run_sequential(ctx, pools, requests):
    """Reference execution: each request alone, one after another in list
    order, synchronising after each.  Returns per-request GPU ms."""
    import torch

    from ._native import Plan
    out = []
    s = torch.cuda.current_stream()
    for r in requests:
        sp, dp = POOLS[r.op]
        plan = Plan(ctx, r.op, r.items)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        plan.execute(pools[sp], pools[dp], s)
        b.record(s)
        b.synchronize()
        out.append(a.elapsed_time(b))
        plan.close()
    return out
