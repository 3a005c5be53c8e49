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
