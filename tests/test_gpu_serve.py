"""GPU: the served path (resident subjects, continuous admission).

Each served result equals the one-instance path's result for the same
subject and activity, bit for bit, alone and in the mix: requests for three
subjects of different sizes and bin counts arrive together and overlap on
two contexts.  A device-memory budget that holds one subject at a time
forces evictions and reloads without changing any result.
"""
import dataclasses

import numpy as np
import pytest

from opensss_amd import phantoms
from tests.sss_cases import WATER

pytestmark = pytest.mark.gpu


def _subjects():
    return {
        'a': phantoms.small_tof(seed=11, n_sp=30, n_bins=5),
        'b': phantoms.small_tof(seed=12, n_sp=24, n_bins=1, shape=(20, 20, 10)),
        'c': phantoms.small_tof(seed=13, n_sp=36, n_bins=9, n_det=32, ndu=16),
    }


def _requests(subs, n, seed):
    from opensss_amd.serve import Request
    rng = np.random.default_rng(seed)
    keys = list(subs)
    out = []
    for i in range(n):
        k = keys[i % len(keys)] if i < len(keys) else keys[int(rng.integers(len(keys)))]
        lam = subs[k].activity * (1.0 + 0.05 * rng.standard_normal(subs[k].activity.shape))
        out.append(Request(key=k, activity=lam, arrival=0.0, tag=i))
    return out


def _alone(subs, req):
    from opensss_amd.batch import run_plain
    return run_plain([dataclasses.replace(subs[req.key], activity=req.activity)], WATER)[0]


def test_served_results_match_alone():
    from opensss_amd.serve import ScatterServer, serve
    subs = _subjects()
    reqs = _requests(subs, 12, seed=1)
    server = ScatterServer(subs, WATER, n_contexts=2, max_inflight=3)
    try:
        serve(server, reqs)
        stats = dict(server.stats)
    finally:
        server.close()
    assert stats['completed'] == len(reqs) and stats['loaded'] == 3 and stats['evicted'] == 0
    assert stats['peak_inflight'] >= 2   # requests overlapped on the device
    for r in reqs:
        assert r.done is not None and r.admitted is not None and r.done >= r.admitted
        want = _alone(subs, r)
        assert r.output.shape == want.shape
        assert np.array_equal(r.output, want), f'request {r.tag} (subject {r.key}) differs from its run alone'


def test_admission_by_footprint_evicts_and_reloads():
    from opensss_amd.serve import ScatterServer, footprint_estimate, serve
    subs = _subjects()
    reqs = _requests(subs, 9, seed=2)
    budget = int(1.2 * max(footprint_estimate(w) for w in subs.values()))
    server = ScatterServer(subs, WATER, n_contexts=2, max_inflight=2, hbm_budget=budget)
    try:
        serve(server, reqs)
        stats = dict(server.stats)
        assert len(server.resident) >= 1
    finally:
        server.close()
    assert stats['completed'] == len(reqs)
    assert stats['evicted'] > 0 and stats['loaded'] > 3
    for r in reqs:
        assert np.array_equal(r.output, _alone(subs, r)), f'request {r.tag} (subject {r.key})'


def test_tile_shapes_give_identical_results_and_autotune_serves_them(tmp_path):
    """Every tile shape gives the same bits (each cell keeps its order of
    sums), whether set at create or by re-tiling a traced subject; the tuner
    times the candidates on a resident subject and keeps the fastest per
    shape key (saved and reloaded through its JSON file); the server with
    autotune serves the one-instance path's results with the tuned shape."""
    from opensss_amd.autotune import TileTuner, shape_key
    from opensss_amd.serve import ScatterServer, serve
    from opensss_amd.sss import ScatterSubject
    w = phantoms.small_tof(seed=21, n_sp=30, n_bins=7, n_rings=8, nru=8, ndu=16)

    def subject(**kw):
        return ScatterSubject(w.attenuation, w.image_size, w.geometry, w.normals, w.det_size, WATER,
                              w.energy_resolution, w.energy_threshold, w.n_rings_used, w.n_det_used,
                              w.tof_resolution, w.tof_range, w.n_bins, scatter_points=w.scatter_points, **kw)

    outs = []
    for rb in (1, 2, 4, 8):
        s = subject(tile_rings=rb)
        try:
            s.run(w.activity)
            outs.append(s.output_compact())
        finally:
            s.close()
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    s = subject()
    try:
        s.run(w.activity)
        for rb in (8, 1, 2, 4, 0):
            s.set_tiling(rb, 0)
            s.scatter()
            assert np.array_equal(s.output_compact(), outs[0]), rb
        s.set_tiling(0, 10)   # shorter chunks: the slice sums regroup (FP32 rounding)
        s.scatter()
        assert s.info['n_chunks'] >= 10
        np.testing.assert_allclose(s.output_compact(), outs[0], rtol=1e-5, atol=1e-6 * outs[0].max())
        path = str(tmp_path / 'tiles.json')
        tuner = TileTuner(path)
        best = tuner.tune(s, w)
        key = shape_key(w)
        assert tuner.best[key] == best and len(tuner.times[key]) == len(tuner.candidates(w)) == 8
        assert all(t > 0 for t in tuner.times[key].values())
        assert TileTuner(path).best == {key: best}
    finally:
        s.close()
    subs = {'a': w}
    reqs = _requests(subs, 3, seed=5)
    server = ScatterServer(subs, WATER, n_contexts=1, tuner=TileTuner(path))
    try:
        serve(server, reqs)
        assert server.tuner.timed_subjects == 0   # the key was tuned above
    finally:
        server.close()
    for r in reqs:
        want = _alone(subs, r)
        if best[1] == 0:
            assert np.array_equal(r.output, want)
        else:
            np.testing.assert_allclose(r.output, want, rtol=1e-5, atol=1e-6 * want.max())


def test_short_random_stream_served_equals_alone():
    """A short Poisson stream over three subjects of the served distribution
    (full-size random instances, activity x (1 + 0.05 N(0,1)) per request):
    every served output equals its run alone."""
    from opensss_amd.serve import Request, ScatterServer, serve
    subs = {f'r{s}': phantoms.random_instance(s) for s in (11, 14, 3)}
    rng = np.random.default_rng(5)
    keys = list(subs)
    t, reqs = 0.0, []
    for i in range(9):
        t += float(rng.exponential(1 / 50.0))
        k = keys[int(rng.integers(len(keys)))]
        lam = subs[k].activity * (1.0 + 0.05 * rng.standard_normal(subs[k].activity.shape))
        reqs.append(Request(key=k, activity=lam, arrival=t, tag=i))
    server = ScatterServer(subs, WATER, n_contexts=2, max_inflight=3)
    try:
        serve(server, reqs)
    finally:
        server.close()
    for r in reqs:
        assert r.output is not None
        assert np.array_equal(r.output, _alone(subs, r)), f'request {r.tag} ({r.key})'
