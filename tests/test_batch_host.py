"""Host logic of the batched path (opensss_amd.batch.run_batch), on the CPU:
outputs come back in instance order whatever the pipeline's interleaving,
instances go round-robin to the contexts, at most 2 x contexts + 1
instances hold device memory at once, and every subject is closed.  The GPU test of
the same module checks the real outputs (tests/test_gpu_batch.py)."""
import ctypes
from types import SimpleNamespace

import numpy as np

from opensss_amd import batch


class Ctx:
    def __init__(self):
        self.blocks = {}

    def memory(self):
        return 1 << 40, 1 << 40

    def release_cache(self):
        pass

    def host_alloc(self, nbytes):
        buf = ctypes.create_string_buffer(max(1, nbytes))
        self.blocks[ctypes.addressof(buf)] = buf
        return ctypes.addressof(buf)

    def host_free(self, addr):
        del self.blocks[addr]


class Subject:
    alive = 0
    peak = 0
    contexts = []

    def __init__(self, attenuation, *args, context=None, **kw):
        self.ctx = context
        self.key = float(attenuation.flat[0])
        self.output_shape = (2, 2)
        self.output_dtype = np.float32
        self.output_nbytes = 16
        self.stage = 'created'
        Subject.alive += 1
        Subject.peak = max(Subject.peak, Subject.alive)
        Subject.contexts.append(context)

    def set_activity(self, lam):
        self.lam = float(np.asarray(lam).sum())

    def trace_begin(self):
        assert self.stage == 'created'
        self.stage = 'counting'

    def trace_end(self):
        assert self.stage == 'counting'
        self.stage = 'traced'

    def scatter(self):
        assert self.stage == 'traced'
        self.stage = 'scattered'

    def output_compact_async(self, addr):
        assert self.stage == 'scattered'
        self.addr = addr

    def wait(self):
        np.ctypeslib.as_array((ctypes.c_float * 4).from_address(self.addr))[:] = self.key * self.lam

    def close(self):
        if self.stage != 'closed':
            Subject.alive -= 1
        self.stage = 'closed'


def test_run_batch_order_contexts_and_residency(monkeypatch):
    monkeypatch.setattr(batch, 'ScatterSubject', Subject)
    Subject.alive = Subject.peak = 0
    Subject.contexts = []
    insts = [SimpleNamespace(attenuation=np.full((2, 2, 2), float(i + 1)), activity=np.full((2, 2, 2), 0.5 * i),
                             image_size=None, geometry=None, normals=None, det_size=None,
                             energy_resolution=0.1, energy_threshold=435.0, n_rings_used=1, n_det_used=2,
                             tof_resolution=214.0, tof_range=3335.64, n_bins=1, scatter_points=None)
             for i in range(11)]
    ctxs = [Ctx(), Ctx(), Ctx()]
    outs = batch.run_batch(insts, None, contexts=ctxs)
    for i, o in enumerate(outs):
        assert o.shape == (2, 2) and np.all(o == np.float32((i + 1) * 0.5 * i * 8)), i
    assert Subject.contexts == [ctxs[i % 3] for i in range(11)]
    assert Subject.alive == 0 and Subject.peak <= 2 * len(ctxs) + 1, Subject.peak
    assert all(not c.blocks for c in ctxs)


class FailingSubject(Subject):
    """Fails in scatter for the instance whose mu value is 5."""

    def scatter(self):
        if self.key == 5.0:
            raise RuntimeError('planted failure')
        super().scatter()


def test_run_batch_error_releases_pinned_blocks(monkeypatch):
    """An error inside the pipeline closes every subject still queued or
    running and hands their pinned output blocks back to their contexts."""
    import pytest
    monkeypatch.setattr(batch, 'ScatterSubject', FailingSubject)
    Subject.alive = Subject.peak = 0
    Subject.contexts = []
    insts = [SimpleNamespace(attenuation=np.full((2, 2, 2), float(i + 1)), activity=np.full((2, 2, 2), 1.0),
                             image_size=None, geometry=None, normals=None, det_size=None,
                             energy_resolution=0.1, energy_threshold=435.0, n_rings_used=1, n_det_used=2,
                             tof_resolution=214.0, tof_range=3335.64, n_bins=1, scatter_points=None)
             for i in range(9)]
    ctxs = [Ctx(), Ctx()]
    with pytest.raises(RuntimeError, match='planted'):
        batch.run_batch(insts, None, contexts=ctxs)
    assert Subject.alive == 0
    assert all(not c.blocks for c in ctxs)


class ShortOfMemory(Subject):
    """trace_end runs out of device memory once per subject of mu value 3."""
    failed = set()

    def trace_end(self):
        from opensss_amd import _native
        if self.key == 3.0 and self.key not in ShortOfMemory.failed:
            ShortOfMemory.failed.add(self.key)
            raise _native.SssError(-3, 'sss_subject_trace_end', 'hipMalloc failed: out of memory')
        super().trace_end()


class CacheCtx(Ctx):
    def __init__(self):
        super().__init__()
        self.released = 0

    def release_cache(self):
        self.released += 1


def test_run_batch_releases_caches_and_retries_when_memory_runs_out(monkeypatch):
    monkeypatch.setattr(batch, 'ScatterSubject', ShortOfMemory)
    Subject.alive = Subject.peak = 0
    Subject.contexts = []
    ShortOfMemory.failed = set()
    insts = [SimpleNamespace(attenuation=np.full((2, 2, 2), float(i + 1)), activity=np.full((2, 2, 2), 1.0),
                             image_size=None, geometry=None, normals=None, det_size=None,
                             energy_resolution=0.1, energy_threshold=435.0, n_rings_used=1, n_det_used=2,
                             tof_resolution=214.0, tof_range=3335.64, n_bins=1, scatter_points=None)
             for i in range(6)]
    ctxs = [CacheCtx(), CacheCtx()]
    outs = batch.run_batch(insts, None, contexts=ctxs)
    assert all(np.all(o == np.float32((i + 1) * 8.0)) for i, o in enumerate(outs))
    assert all(c.released == 1 for c in ctxs)
    assert Subject.alive == 0
