"""Many scatter estimates on one GPU: the batched paths.

An instance is anything with the attributes of `phantoms.Workload` (the
reference's SingleScatterSimulationTOF arguments plus an explicit scatter
point list).

`GraphBatch` holds traced subjects resident on one context and estimates
them together: one persistent scatter launch per kernel variant over every
member's work (per-XCD work queues), replayed as one HIP graph with a new
activity image per member and launch (the repeated estimates inside a
reconstruction).  `run_graph_batch` runs instances through it in resident
groups that fit a device-memory budget.

`run_batch` keeps several instances in flight on several contexts (one HIP
stream each) and overlaps their phases:

    create + activity upload + walk counts        (host work, queued GPU work)
    table sizing + walks + scatter + reduce +
    download into pinned memory                   (queued GPU work)
    wait + copy out + destroy                     (waits for that instance only)

so one instance's host-side preparation and waits hide behind the others'
kernels.  Each context keeps the device blocks of the instances it has
finished for the next ones (no hipMalloc/hipFree per instance).  Results are
those of the one-instance path, bit for bit (same kernels, same order).
"""
import ctypes
from collections import deque

import numpy as np

from . import _native
from .sss import ScatterSubject


def _subject(inst, attenuation_table, ctx, precision):
    return ScatterSubject(inst.attenuation, inst.image_size, inst.geometry, inst.normals, inst.det_size,
                          attenuation_table, inst.energy_resolution, inst.energy_threshold,
                          inst.n_rings_used, inst.n_det_used, inst.tof_resolution, inst.tof_range,
                          inst.n_bins, scatter_points=inst.scatter_points, precision=precision,
                          context=ctx)


def _collect(s, addr):
    s.wait()
    n = int(np.prod(s.output_shape))
    ct = ctypes.c_double if s.output_dtype == np.float64 else ctypes.c_float
    out = np.ctypeslib.as_array((ct * n).from_address(addr)).reshape(s.output_shape).copy()
    s.ctx.host_free(addr)
    s.close()
    return out


SSS_ERR_NOMEM = -3


def _out_of_memory(e):
    return isinstance(e, _native.SssError) and e.code == SSS_ERR_NOMEM


def run_batch(instances, attenuation_table, n_streams=4, precision='fp32', contexts=None,
              memory_budget=None):
    """Compact outputs (B, NRU^2, NDU(NDU+1)/2) of every instance, in order.

    n_streams contexts take the instances round-robin; up to 2 x n_streams + 1
    instances hold device memory at once, fewer when their estimated
    footprints (``serve.footprint_estimate``) would pass `memory_budget`
    bytes (default: 60% of the free device memory at the start); an
    allocation that still fails releases the contexts' cached blocks and is
    retried.  An
    instance's download is queued right behind its scatter into pinned
    memory, so collecting it waits for that instance only.
    """
    from .serve import footprint_estimate
    own = contexts is None
    ctxs = contexts or [_native.Context(0) for _ in range(max(1, int(n_streams)))]
    lag = len(ctxs)
    queued, running = deque(), deque()
    outs = [None] * len(instances)
    held = {}   # instance -> estimated device bytes while it is queued or running
    if memory_budget is None:
        memory_budget = int(0.6 * ctxs[0].memory()[0])

    def retire_one():
        if not running:
            advance()
        k, s, addr = running.popleft()
        outs[k] = _collect(s, addr)
        held.pop(k, None)

    def drain():
        while running:
            retire_one()

    def make_room(need):
        """Collect the oldest instances while the held estimates and this one
        pass the budget."""
        while sum(held.values()) + need > memory_budget and (queued or running):
            retire_one()

    def with_memory(fn):
        """fn(); when device memory runs out, the contexts' cached blocks are
        released and fn retried, then once more after every running
        instance has been collected."""
        for attempt in range(3):
            try:
                return fn()
            except _native.SssError as e:
                if not _out_of_memory(e) or attempt == 2:
                    raise
            if attempt == 1:
                drain()
            for c in ctxs:
                c.release_cache()

    def advance():
        k, s = queued.popleft()
        try:
            with_memory(s.trace_end)
            s.scatter()
            addr = s.ctx.host_alloc(s.output_nbytes)
        except BaseException:
            s.close()
            raise
        try:
            s.output_compact_async(addr)
        except BaseException:
            s.ctx.host_free(addr)
            s.close()
            raise
        running.append((k, s, addr))

    try:
        for i, inst in enumerate(instances):
            need = footprint_estimate(inst)
            make_room(need)
            s = with_memory(lambda: _subject(inst, attenuation_table, ctxs[i % lag], precision))
            held[i] = need
            s.set_activity(inst.activity)
            s.trace_begin()
            queued.append((i, s))
            if len(queued) > lag:
                advance()
            if len(running) > lag:
                retire_one()
        while queued:
            advance()
        while running:
            retire_one()
    finally:
        # on an error, instances still queued or running are released: a
        # running one's download targets its pinned block, so it is waited
        # for before the block goes back to the context
        for _, s in queued:
            s.close()
        for _, s, addr in running:
            try:
                s.wait()
            finally:
                s.ctx.host_free(addr)
                s.close()
        if own:
            for c in ctxs:
                c.close()
    return outs


def run_plain(instances, attenuation_table, precision='fp32', context=None):
    """The one-instance path: create, trace, scatter, download, destroy, one
    instance at a time (what one call of the reference's function does)."""
    outs = []
    for inst in instances:
        s = _subject(inst, attenuation_table, context, precision)
        try:
            s.run(inst.activity)
            outs.append(s.output_compact())
        finally:
            s.close()
    return outs


class GraphBatch:
    """Traced subjects of one context estimated as one unit (``sss_batch``).

    A launch queues every member's staged activity upload and refresh, one
    persistent scatter launch per kernel variant present over all members'
    work, the chunk reduces and the output downloads into pinned memory;
    ``launch(graph=True)`` replays that sequence as one HIP graph (captured
    on the first launch, and again after a member's walks were redone),
    ``launch(graph=False)`` queues it directly.  Each member's output equals
    its own ``ScatterSubject.scatter`` output bit for bit.  Members need a
    host activity image (``set_activity``) before the batch is made and must
    stay open while it lives (``close`` the batch first).
    """

    def __init__(self, subjects, context=None):
        self.subjects = list(subjects)
        if not self.subjects:
            raise ValueError('GraphBatch needs at least one subject')
        self.ctx = context or self.subjects[0].ctx
        self._lib = _native.load_library()
        arr = (ctypes.c_void_p * len(self.subjects))(*[s.handle for s in self.subjects])
        h = ctypes.c_void_p()
        _native.check(self._lib.sss_batch_create(self.ctx.handle, arr, len(self.subjects), ctypes.byref(h)),
                      'sss_batch_create')
        self.handle = h

    def __len__(self):
        return len(self.subjects)

    def set_activity(self, i: int, ActivityMap):
        """Stage member i's next activity image (waits for the last launch)."""
        s = self.subjects[i]
        lam = np.asarray(ActivityMap)
        if tuple(lam.shape) != s.grid:
            raise ValueError(f'ActivityMap shape {lam.shape} != {s.grid}')
        flat = s._volume(lam)
        _native.check(self._lib.sss_batch_set_activity(self.handle, int(i), _native.ptr(flat)),
                      'sss_batch_set_activity')

    def launch(self, graph: bool = True):
        fn = self._lib.sss_batch_graph_launch if graph else self._lib.sss_batch_submit
        _native.check(fn(self.handle), 'sss_batch_graph_launch' if graph else 'sss_batch_submit')

    def wait(self):
        _native.check(self._lib.sss_batch_wait(self.handle), 'sss_batch_wait')

    def output(self, i: int) -> np.ndarray:
        """Member i's compact output of the last launch (waits for it)."""
        s = self.subjects[i]
        out = np.empty(s.output_shape, dtype=s.output_dtype)
        _native.check(self._lib.sss_batch_output(self.handle, int(i), _native.ptr(out)), 'sss_batch_output')
        return out

    def elapsed_ms(self) -> float:
        """Duration of the last launch (HIP events around it)."""
        ms = ctypes.c_float()
        _native.check(self._lib.sss_batch_elapsed_ms(self.handle, ctypes.byref(ms)), 'sss_batch_elapsed_ms')
        return float(ms.value)

    @property
    def info(self) -> dict:
        n = ctypes.c_int32()
        g = ctypes.c_int32()
        _native.check(self._lib.sss_batch_info(self.handle, ctypes.byref(n), ctypes.byref(g)), 'sss_batch_info')
        return {'scatter_launches': int(n.value), 'graph_ready': bool(g.value)}

    def close(self):
        if getattr(self, 'handle', None):
            self._lib.sss_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def resident_groups(instances, memory_budget):
    """Consecutive runs of instances whose estimated footprints
    (``serve.footprint_estimate``) fit `memory_budget` bytes together (an
    instance above the budget alone makes a group of one)."""
    from .serve import footprint_estimate
    groups, cur, used = [], [], 0
    for k, inst in enumerate(instances):
        need = footprint_estimate(inst)
        if cur and used + need > memory_budget:
            groups.append(cur)
            cur, used = [], 0
        cur.append(k)
        used += need
    if cur:
        groups.append(cur)
    return groups


def run_graph_batch(instances, attenuation_table, precision='fp32', context=None, memory_budget=None,
                    graph=True, iterations=1, activity=None, timings=None, tuner=None):
    """Compact outputs of every instance through `GraphBatch`.

    Instances go in resident groups (``resident_groups``; the budget
    defaults to 60% of the free device memory): a group's subjects are
    created, given their activity and traced, then estimated by one batch
    launch (one persistent scatter launch per kernel variant).  With
    ``iterations`` > 1 the group is launched that many times, member k's
    image for launch j being ``activity(k, j)`` (default: the instance's own
    image), as the repeated estimates of a reconstruction do; the outputs
    are those of the last launch.  ``timings``, a list, receives one dict per
    group: {'members', 'setup_s', 'launch_ms': [per launch]}.  With a
    ``tuner`` (``autotune.TileTuner``) each subject takes the tile shape
    tuned for its shape key (timed on the first subject of a new key).
    """
    import time
    own = context is None
    ctx = context or _native.Context(0)
    outs = [None] * len(instances)
    try:
        if memory_budget is None:
            memory_budget = int(0.6 * ctx.memory()[0])
        for group in resident_groups(instances, memory_budget):
            t0 = time.perf_counter()
            subs = []
            batch = None
            try:
                for k in group:
                    inst = instances[k]
                    s = _subject(inst, attenuation_table, ctx, precision)
                    subs.append(s)
                    s.set_activity(inst.activity if activity is None else activity(k, 0))
                    s.trace()
                    if tuner is not None:
                        tuner.apply(s, inst, precision)
                batch = GraphBatch(subs, ctx)
                setup = time.perf_counter() - t0
                ms = []
                for j in range(max(1, int(iterations))):
                    if j > 0 or activity is not None:
                        for m, k in enumerate(group):
                            batch.set_activity(m, instances[k].activity if activity is None else activity(k, j))
                    batch.launch(graph=graph)
                    ms.append(batch.elapsed_ms())
                for m, k in enumerate(group):
                    outs[k] = batch.output(m)
                if timings is not None:
                    timings.append({'members': list(group), 'setup_s': setup, 'launch_ms': ms})
            finally:
                if batch is not None:
                    batch.close()
                for s in subs:
                    s.close()
    finally:
        if own:
            ctx.close()
    return outs
