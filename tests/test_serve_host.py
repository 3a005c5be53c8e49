"""Host logic of the served path (opensss_amd.serve), on the CPU.

The subjects are stand-ins with the resident subject's interface: a request
"computes" a known function of its activity, and a download finishes a few
polls after it is queued, so the admission order, the slot and
one-request-per-subject limits, the memory budget and the eviction rules are
checked without a GPU.  The GPU tests of the same module check the real
outputs (tests/test_gpu_serve.py).
"""
import ctypes
from types import SimpleNamespace

import numpy as np
import pytest

from opensss_amd import serve


class FakeContext:
    def __init__(self):
        self.blocks = {}

    def memory(self):
        return 1 << 40, 1 << 40

    def host_alloc(self, nbytes):
        buf = ctypes.create_string_buffer(max(1, nbytes))
        addr = ctypes.addressof(buf)
        self.blocks[addr] = buf
        return addr

    def host_free(self, addr):
        del self.blocks[addr]

    def close(self):
        pass


class FakeSubject:
    """Output of a request: (key + 1) * sum(activity) in every cell; the
    download completes on the third poll after it is queued."""
    live = {}
    log = []
    timed = []   # keys of the subjects whose scatter was timed
    sizes = {}   # device bytes per subject key (default 1000)

    def __init__(self, attenuation, *args, context=None, scatter_points=None, precision='fp32',
                 tile_rings=0, **kw):
        self.ctx = context
        self.key = int(attenuation.flat[0])
        self.tile_rings = tile_rings
        self.output_shape = (1, 2, 3)
        self.output_dtype = np.float32
        self.output_nbytes = 24
        self.info = {'device_bytes': FakeSubject.sizes.get(self.key, 1000)}
        self.lam = None
        self.pending = None
        self.polls = 0
        self.busy = False
        assert self.key not in FakeSubject.live, 'subject loaded twice at once'
        FakeSubject.live[self.key] = self
        FakeSubject.log.append(('load', self.key))

    def trace(self):
        pass

    def run(self, lam):
        self.set_activity(lam)

    def scatter_kernel_ms(self):
        # the stand-in's fastest tile shape: 2 rings for even keys, 4 for odd
        best = 2 if self.key % 2 == 0 else 4
        FakeSubject.timed.append(self.key)
        return 1.0 if self.tile_rings == best else 2.0

    def set_tiling(self, tile_rings=0, n_chunks=0):
        self.tile_rings = tile_rings

    def set_activity(self, lam):
        assert not self.busy, 'activity changed while a request is in flight'
        self.lam = np.asarray(lam)

    def scatter(self):
        self.value = np.float32((self.key + 1) * float(self.lam.sum()))

    def output_compact_async(self, addr):
        self.busy = True
        self.pending = addr
        self.polls = 0

    def done(self):
        self.polls += 1
        if self.polls >= 3 and self.pending is not None:
            arr = np.ctypeslib.as_array((ctypes.c_float * 6).from_address(self.pending))
            arr[:] = self.value
            self.pending = None
            self.busy = False
        return self.pending is None

    def wait(self):
        while not self.done():
            pass

    def close(self):
        assert not self.busy, 'evicted while a request is in flight'
        FakeSubject.live.pop(self.key, None)
        FakeSubject.log.append(('close', self.key))


def _instance(key):
    mu = np.full((4, 4, 2), float(key))
    return SimpleNamespace(attenuation=mu, image_size=None, geometry=None, normals=None, det_size=None,
                           energy_resolution=0.1, energy_threshold=435.0, n_rings_used=2, n_det_used=4,
                           tof_resolution=214.0, tof_range=3335.64, n_bins=1,
                           scatter_points=np.zeros((3, 3), np.int32))


@pytest.fixture
def fake(monkeypatch):
    monkeypatch.setattr(serve, 'ScatterSubject', FakeSubject)
    FakeSubject.live = {}
    FakeSubject.log = []
    FakeSubject.timed = []
    FakeSubject.sizes = {}
    return [FakeContext(), FakeContext()]


def _requests(keys, seed=0):
    rng = np.random.default_rng(seed)
    return [serve.Request(key=k, activity=rng.uniform(0.5, 1.5, size=(4, 4, 2)), arrival=0.0, tag=i)
            for i, k in enumerate(keys)]


def test_every_request_gets_its_own_result(fake):
    subs = {k: _instance(k) for k in range(3)}
    reqs = _requests([0, 1, 2, 0, 0, 1, 2, 2, 1, 0])
    server = serve.ScatterServer(subs, None, contexts=fake, max_inflight=2)
    serve.serve(server, reqs)
    assert server.stats['completed'] == len(reqs) and server.stats['loaded'] == 3
    assert server.stats['peak_inflight'] == 2
    for r in reqs:
        want = np.float32((r.key + 1) * float(r.activity.sum()))
        assert r.output.shape == (1, 2, 3) and np.all(r.output == want), r.tag
    # requests of one subject keep their order
    for k in subs:
        mine = [r for r in reqs if r.key == k]
        assert all(a.admitted <= b.admitted for a, b in zip(mine, mine[1:]))
    server.close()
    assert not FakeSubject.live


def test_busy_subject_does_not_block_others(fake):
    subs = {k: _instance(k) for k in range(2)}
    reqs = _requests([0, 0, 1])
    server = serve.ScatterServer(subs, None, contexts=fake, max_inflight=2)
    for r in reqs:
        server.submit(r)
    assert server.admit(0.0) == 2          # request 0 and request 2; request 1 waits for subject 0
    assert set(server.inflight) == {0, 1}
    assert reqs[1].admitted is None and reqs[2].admitted == 0.0
    while not server.idle():
        server.poll(1.0, serve.copy_result)
        server.admit(1.0)
    assert all(r.output is not None for r in reqs)
    server.close()


def test_memory_budget_evicts_idle_subjects_only(fake):
    subs = {k: _instance(k) for k in range(3)}
    FakeSubject.sizes = {k: serve.footprint_estimate(w) for k, w in subs.items()}
    budget = int(1.5 * serve.footprint_estimate(subs[0]))     # one subject at a time
    reqs = _requests([0, 1, 2, 1, 0, 2, 2, 0])
    server = serve.ScatterServer(subs, None, contexts=fake, max_inflight=2, hbm_budget=budget)
    serve.serve(server, reqs)
    assert server.stats['completed'] == len(reqs)
    assert server.stats['evicted'] > 0 and server.stats['loaded'] > 3
    assert server.stats['peak_inflight'] == 1    # a second subject never fits beside the first
    for r in reqs:
        assert np.all(r.output == np.float32((r.key + 1) * float(r.activity.sum())))
    server.close()


def test_subject_larger_than_budget_is_an_error(fake):
    subs = {0: _instance(0)}
    server = serve.ScatterServer(subs, None, contexts=fake, hbm_budget=10)
    with pytest.raises(MemoryError):
        server.load(0)
    server.close()


def test_subjects_spread_over_contexts(fake):
    subs = {k: _instance(k) for k in range(4)}
    server = serve.ScatterServer(subs, None, contexts=fake)
    ctxs = [server.load(k).ctx for k in range(4)]
    assert ctxs == [fake[0], fake[1], fake[0], fake[1]]
    server.close()


def test_autotune_keeps_the_fastest_tile_shape(fake):
    """Subjects of one shape key share the tuned shape: subject 2 has subject
    0's key (8 rings x 16 detectors) and is not timed; subject 1 (24
    detectors) has its own key."""
    from opensss_amd.autotune import shape_key
    subs = {k: _instance(k) for k in range(3)}
    for k, w in subs.items():
        w.activity = np.ones((4, 4, 2))
        w.n_rings_used, w.n_det_used = 8, (24 if k % 2 else 16)
    reqs = _requests([0, 1, 2, 0])
    server = serve.ScatterServer(subs, None, contexts=fake, autotune=True)
    serve.serve(server, reqs)
    assert server.tuner.best == {shape_key(subs[0]): (2, 0), shape_key(subs[1]): (4, 0)}
    assert [server.resident[k][0].tile_rings for k in range(3)] == [2, 4, 2]
    assert 2 not in FakeSubject.timed and server.tuner.timed_subjects == 2
    for r in reqs:
        assert np.all(r.output == np.float32((r.key + 1) * float(r.activity.sum())))
    server.close()


def test_request_stays_queued_when_its_subject_cannot_load(fake):
    """A load that raises (the subject alone exceeds the budget) leaves the
    request at the front of the queue instead of dropping it."""
    subs = {0: _instance(0), 1: _instance(1)}
    FakeSubject.sizes = {}
    server = serve.ScatterServer(subs, None, contexts=fake, hbm_budget=10)
    reqs = _requests([1, 0])
    for r in reqs:
        server.submit(r)
    with pytest.raises(MemoryError):
        server.admit(0.0)
    assert [r.tag for r in server.queue] == [0, 1]
    server.queue.clear()
    server.close()
