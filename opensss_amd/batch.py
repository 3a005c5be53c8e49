"""Many scatter estimates on one GPU: the batched path.

An instance is anything with the attributes of `phantoms.Workload` (the
reference's SingleScatterSimulationTOF arguments plus an explicit scatter
point list).  `run_batch` keeps several instances in flight on several
contexts (one HIP stream each) and overlaps their phases:

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
