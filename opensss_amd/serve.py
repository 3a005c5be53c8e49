"""Served scatter estimation: resident subjects, continuous admission.

A subject (scanner, mu-map, scatter points: one ``phantoms.Workload``-like
instance) is created and traced once and stays resident in HBM; a request
names a subject and brings an activity image.  Serving a request uploads the
image, refreshes the emission terms, runs the scatter sum and downloads the
compact sinogram into pinned memory, all queued on the subject's context
stream.

``ScatterServer`` admits waiting requests whenever a slot and device memory
free up (continuous batching, not a queue drained batch by batch):

- at most ``max_inflight`` requests are on the GPU at once, one per subject
  (a subject's activity and output buffers are single, so a second request
  for a busy subject waits while requests for other subjects go ahead);
- subjects are spread over ``n_contexts`` contexts (one HIP stream each), so
  the work of requests for different subjects overlaps on the device;
- a subject is made resident only when its device-memory footprint fits the
  budget; idle subjects are evicted least recently used first to make room,
  and a request whose subject cannot fit yet waits;
- with ``autotune`` (or a ``tuner``), a subject takes the tile shape tuned
  for its shape key (``autotune.TileTuner``); the first subject of a key
  times the candidates on its resident tables.

Every request's result is the one-instance path's result for the same
subject and activity, bit for bit: the same kernels in the same order on
the subject's own buffers, whatever else is in flight.
"""
from __future__ import annotations

import ctypes
import time
from collections import OrderedDict, deque
from dataclasses import dataclass
from typing import Any, Callable, Optional

import numpy as np

from . import _native
from .autotune import TileTuner
from .sss import ScatterSubject

SSS_ERR_NOMEM = -3


@dataclass
class Request:
    key: Any                           # subject id
    activity: np.ndarray               # (X, Y, Z) float64
    arrival: float = 0.0               # seconds from the start of the stream
    admitted: Optional[float] = None
    done: Optional[float] = None
    tag: Any = None
    output: Optional[np.ndarray] = None


def footprint_estimate(inst) -> int:
    """Device bytes of a traced subject, before it exists: volumes, the
    per-half-path table and the sample table at half the grid's axis sum
    per walk (about what the phantoms' walks average)."""
    X, Y, Z = (int(v) for v in inst.attenuation.shape)
    n_vox = X * Y * Z
    nru, ndu = int(inst.n_rings_used), int(inst.n_det_used)
    n_hp = (0 if inst.scatter_points is None else len(inst.scatter_points)) * nru * ndu
    walk = (X + Y + Z) // 2
    out = int(inst.n_bins) * nru * nru * ndu * (ndu + 1) // 2 * 4
    return 4 * n_vox * 8 + n_hp * 120 + n_hp * walk * 48 + 3 * out


class ScatterServer:
    def __init__(self, instances: dict, attenuation_table, n_contexts: int = 2,
                 max_inflight: Optional[int] = None, hbm_budget: Optional[int] = None,
                 precision: str = 'fp32', contexts=None, autotune: bool = False,
                 tuner: Optional[TileTuner] = None):
        self.instances = dict(instances)
        # tile shapes by timing (autotune), per subject shape
        self.tuner = tuner if tuner is not None else (TileTuner() if autotune else None)
        self.table = attenuation_table
        self.precision = precision
        self._own_ctx = contexts is None
        self.ctxs = list(contexts) if contexts else [_native.Context(0) for _ in range(max(1, n_contexts))]
        self.max_inflight = max_inflight or 2 * len(self.ctxs)
        if hbm_budget is None:
            free, _ = self.ctxs[0].memory()
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.queue: deque = deque()
        self.inflight: dict = {}                  # key -> (request, pinned address)
        self.resident: OrderedDict = OrderedDict()  # key -> (subject, bytes), LRU first
        self._ctx_of: dict = {}
        self.stats = dict(loaded=0, evicted=0, admitted=0, completed=0, peak_inflight=0)

    # -- subjects ------------------------------------------------------------
    def resident_bytes(self) -> int:
        return sum(b for _, b in self.resident.values())

    def _context(self, key):
        if key not in self._ctx_of:
            self._ctx_of[key] = len(self._ctx_of) % len(self.ctxs)
        return self.ctxs[self._ctx_of[key]]

    def _evict_one(self) -> bool:
        for key in self.resident:
            if key not in self.inflight:
                s, _ = self.resident.pop(key)
                s.close()
                self.stats['evicted'] += 1
                return True
        return False

    def load(self, key) -> Optional[ScatterSubject]:
        """The resident subject of key, created and traced if it fits the
        budget (evicting idle subjects), else None."""
        if key in self.resident:
            self.resident.move_to_end(key)
            return self.resident[key][0]
        inst = self.instances[key]
        need = footprint_estimate(inst)
        while self.resident and self.resident_bytes() + need > self.hbm_budget:
            if not self._evict_one():
                return None
        if need > self.hbm_budget:
            raise MemoryError(f'subject {key!r} needs about {need} bytes, budget {self.hbm_budget}')
        while True:
            s = None
            try:
                s = ScatterSubject(inst.attenuation, inst.image_size, inst.geometry, inst.normals, inst.det_size,
                                   self.table, inst.energy_resolution, inst.energy_threshold, inst.n_rings_used,
                                   inst.n_det_used, inst.tof_resolution, inst.tof_range, inst.n_bins,
                                   scatter_points=inst.scatter_points, precision=self.precision,
                                   context=self._context(key))
                s.trace()
                if self.tuner is not None:
                    self.tuner.apply(s, inst, self.precision, activity=inst.activity)
                break
            except _native.SssError as e:
                if s is not None:
                    s.close()
                # device memory ran out below the estimate: drop the
                # contexts' cached blocks and an idle subject, then retry
                if e.code != SSS_ERR_NOMEM:
                    raise
                for c in self.ctxs:
                    c.release_cache()
                if not self._evict_one():
                    raise MemoryError(f'subject {key!r} does not fit in device memory') from e
        self.resident[key] = (s, int(s.info['device_bytes']))
        self.stats['loaded'] += 1
        return s

    # -- requests --------------------------------------------------------------
    def submit(self, req: Request) -> None:
        self.queue.append(req)

    def admit(self, now: float) -> int:
        """Start every waiting request that a free slot, an idle subject and
        the memory budget allow, oldest first; returns how many started."""
        started = 0
        for _ in range(len(self.queue)):
            if len(self.inflight) >= self.max_inflight:
                break
            req = self.queue.popleft()
            try:
                s = None if req.key in self.inflight else self.load(req.key)
            except BaseException:
                # the request goes back to the front of the queue, for the
                # caller to retry or drop
                self.queue.appendleft(req)
                raise
            if s is None:
                self.queue.append(req)
                continue
            s.set_activity(req.activity)
            s.scatter()
            addr = s.ctx.host_alloc(s.output_nbytes)
            s.output_compact_async(addr)
            req.admitted = now
            self.inflight[req.key] = (req, addr)
            self.stats['admitted'] += 1
            self.stats['peak_inflight'] = max(self.stats['peak_inflight'], len(self.inflight))
            started += 1
        return started

    def poll(self, now: float, on_result: Callable) -> int:
        """Hand each finished request's output to on_result(request, view);
        the view (pinned memory) is valid during the call only."""
        finished = [k for k, _ in self.inflight.items() if self.resident[k][0].done()]
        for key in finished:
            req, addr = self.inflight.pop(key)
            s = self.resident[key][0]
            n = int(np.prod(s.output_shape))
            ct = ctypes.c_double if s.output_dtype == np.float64 else ctypes.c_float
            view = np.ctypeslib.as_array((ct * n).from_address(addr)).reshape(s.output_shape)
            req.done = now
            try:
                on_result(req, view)
            finally:
                s.ctx.host_free(addr)
            self.stats['completed'] += 1
        return len(finished)

    def idle(self) -> bool:
        return not self.queue and not self.inflight

    def close(self) -> None:
        for key in list(self.inflight):
            self.resident[key][0].wait()
            _, addr = self.inflight.pop(key)
            self.resident[key][0].ctx.host_free(addr)
        for s, _ in self.resident.values():
            s.close()
        self.resident.clear()
        if self._own_ctx:
            for c in self.ctxs:
                c.close()


def copy_result(req: Request, view: np.ndarray) -> None:
    req.output = view.copy()


def serve(server: ScatterServer, requests, on_result: Callable = copy_result, clock=time.perf_counter):
    """Replays requests (sorted by arrival, seconds from the start) against
    the clock: each is submitted when it arrives, admitted when it fits and
    completed when its download is done.  Returns the start time."""
    t0 = clock()
    nxt = 0
    n = len(requests)
    while nxt < n or not server.idle():
        now = clock() - t0
        while nxt < n and requests[nxt].arrival <= now:
            server.submit(requests[nxt])
            nxt += 1
        server.poll(clock() - t0, on_result)
        started = server.admit(clock() - t0)
        if not started and not server.inflight and nxt < n:
            time.sleep(max(0.0, min(requests[nxt].arrival - (clock() - t0), 0.002)))
        elif not started:
            time.sleep(0.0001)
    return t0
