"""GPU: the batched path (several instances in flight on several contexts,
device blocks reused through each context's cache) gives the one-instance
path's outputs bit for bit, and a resident subject reused across activity
images after other subjects came and went keeps its results."""
import numpy as np
import pytest

from opensss_amd import phantoms
from opensss_amd.batch import run_batch, run_plain
from tests.sss_cases import WATER, gpu_subject

pytestmark = pytest.mark.gpu


def _instances():
    # small uneven instances: different grids, scanners, bin counts and sizes
    out = []
    for seed, n_bins, n_sp in [(1, 7, 30), (2, 15, 50), (3, 1, 20), (4, 21, 40), (5, 14, 25), (6, 3, 35)]:
        w = phantoms.small_tof(seed=seed, n_bins=n_bins, n_sp=n_sp)
        out.append(w)
    return out


def test_batch_matches_one_instance_path():
    ws = _instances()
    plain = run_plain(ws, WATER)
    for n_streams in (1, 3):
        got = run_batch(ws, WATER, n_streams=n_streams)
        for a, b in zip(got, plain):
            assert a.shape == b.shape and np.array_equal(a, b)


def test_context_cache_reuses_blocks():
    from opensss_amd import _native
    ctx = _native.Context(0)
    try:
        ws = _instances()[:3]
        first = run_batch(ws, WATER, contexts=[ctx])
        cached = ctx.cached_bytes
        assert cached > 0
        again = run_batch(ws, WATER, contexts=[ctx])
        assert ctx.cached_bytes == cached          # the second pass allocated nothing new
        for a, b in zip(first, again):
            assert np.array_equal(a, b)
        ctx.release_cache()
        assert ctx.cached_bytes == 0
    finally:
        ctx.close()


def test_resident_subject_between_others():
    w = phantoms.small_tof(seed=9, n_sp=30)
    s = gpu_subject(w, 'fp32')
    try:
        s.run(w.activity)
        ref = s.output_compact()
        run_batch(_instances()[:2], WATER, contexts=[s.ctx])   # other subjects on the same context
        s.set_activity(w.activity * 2.0)
        s.scatter()
        twice = s.output_compact()
        s.set_activity(w.activity)
        s.scatter()
        assert np.array_equal(s.output_compact(), ref)
        # linear in the activity (exactly, but for cells far below float's
        # normal range, where the terms are subnormal)
        assert np.allclose(twice, 2.0 * ref, rtol=1e-6, atol=1e-12 * np.abs(ref).max())
    finally:
        s.close()


@pytest.mark.parametrize('seeds', [(3, 11, 14, 1)])
def test_full_size_random_instances_batched_equal_one_instance(seeds):
    """Instances of the uneven batch's distribution at their full sizes
    (grids 32^3/2 .. 128^2 x 64, 20-32 detectors x 8-20 rings, 5-20 bins,
    430-2480 scatter points): the batched path over two contexts returns the
    one-instance outputs bit for bit."""
    ws = [phantoms.random_instance(s) for s in seeds]
    plain = run_plain(ws, WATER)
    got = run_batch(ws, WATER, n_streams=2)
    for w, a, b in zip(ws, got, plain):
        assert a.shape == b.shape and np.array_equal(a, b), w.name
        assert np.abs(a).max() > 0


def test_random_instance_subset_vs_oracle(oracle):
    """One full-size instance of the batch distribution (96x96x48 grid, 48
    detectors x 12 rings, 21 bins) on a scatter-point subset against the
    oracle, to the FP32 bar."""
    from tests.sss_cases import gpu_compact, oracle_compact, rel_err_above
    w = phantoms.random_instance(8)
    pts = w.scatter_points[::40]
    got = gpu_compact(w, 'fp32', points=pts)
    want = oracle_compact(oracle, w, points=pts)
    rel, absn = rel_err_above(got, want)
    print(f'{w.name} subset {len(pts)} sp: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= 1e-4 and absn <= 1e-6
