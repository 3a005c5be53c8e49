"""Host logic of the tile-shape tuner (opensss_amd.autotune), on the CPU:
shape keys, the candidate list, the choice of the fastest shape, reuse per
key and the JSON file.  tests/test_gpu_serve.py times real subjects."""
from types import SimpleNamespace

import numpy as np

from opensss_amd.autotune import TileTuner, shape_key


def _inst(shape=(64, 64, 32), n_bins=15, ndu=64, nru=16, n_sp=1000):
    return SimpleNamespace(attenuation=np.zeros(shape), n_bins=n_bins, n_det_used=ndu, n_rings_used=nru,
                           scatter_points=np.zeros((n_sp, 3)), activity=np.ones(shape))


class Timed:
    """Stand-in subject: scatter_kernel_ms is a function of the tiling."""

    def __init__(self, cost):
        self.cost = cost
        self.shape = (0, 0)
        self.calls = []
        self.lam = None

    def set_tiling(self, tile_rings=0, n_chunks=0):
        self.shape = (tile_rings, n_chunks)

    def set_activity(self, lam):
        self.lam = lam

    def scatter(self):
        self.calls.append(self.shape)

    def scatter_kernel_ms(self):
        return self.cost(*self.shape)


def test_shape_keys():
    a = _inst()
    assert shape_key(a) == shape_key(_inst(shape=(60, 70, 30)))        # same walk bucket
    assert shape_key(a) != shape_key(_inst(shape=(128, 128, 64)))
    assert shape_key(a) != shape_key(_inst(n_bins=21))
    assert shape_key(a) != shape_key(_inst(ndu=48))
    assert shape_key(a) != shape_key(_inst(nru=8))
    assert shape_key(a) != shape_key(a, 'fp64')


def test_candidates_respect_the_scanner():
    t = TileTuner(ring_candidates=(1, 2, 4, 8), chunk_points=(0, 32))
    assert t.candidates(_inst(nru=4, ndu=64, n_sp=100)) == [(1, 0), (1, 4), (2, 0), (2, 4), (4, 0), (4, 4)]
    # 8 detectors per block need 8 sampled detectors
    assert [r for r, _ in t.candidates(_inst(nru=16, ndu=4, n_sp=10))] == [2, 2, 4, 4, 8, 8]
    assert t.candidates(_inst(nru=1, ndu=1, n_sp=10)) == [(0, 0), (0, 1)]


def test_tune_picks_the_fastest_and_reuses_it(tmp_path):
    path = str(tmp_path / 'tiles.json')
    t = TileTuner(path, repeats=3)
    inst = _inst(n_sp=640)
    s = Timed(lambda r, c: 0.5 if (r, c) == (4, 0) else 3.0 + r + (0.5 if c else 0.0))
    best = t.tune(s, inst, activity=inst.activity)
    assert best == (4, 0) and s.shape == best and s.lam is inst.activity
    assert len(s.calls) == 3 * len(t.candidates(inst))
    # a second subject of the key is set without timing
    s2 = Timed(lambda r, c: 1.0)
    assert t.apply(s2, _inst(shape=(62, 66, 31), n_sp=640)) == best and s2.calls == [] and s2.shape == best
    # a new key is timed
    s3 = Timed(lambda r, c: float(r))
    assert t.apply(s3, _inst(n_bins=3)) == (1, 0) and s3.calls
    again = TileTuner(path)
    assert again.best == t.best and again.times == t.times
