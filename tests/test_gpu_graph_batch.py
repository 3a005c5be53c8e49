"""GPU: batches (``sss_batch``).  One persistent scatter launch per kernel
variant over several resident subjects, queued directly or replayed as a HIP
graph, gives every subject its own one-instance output bit for bit; graph
replays with new activity images (with other work on the same context
between replays) equal fresh subjects; a member whose walks are redone makes
the batch rebuild its lists and graph."""
import dataclasses

import numpy as np
import pytest

from opensss_amd import _native, phantoms
from opensss_amd.autotune import TileTuner
from opensss_amd.batch import GraphBatch, resident_groups, run_graph_batch, run_plain
from tests.sss_cases import WATER, gpu_compact, gpu_subject

pytestmark = pytest.mark.gpu


def _members():
    """Small subjects covering the kernel variants: sliding window (15 bins),
    relative window h = 2 (21 bins) and its wide-bin form (6 bins), one TOF
    bin (zero table columns at large offsets), two subjects of one bin
    without TOF (the one-bin pair loop; their activity refreshes share one
    launch), an FP64 subject, and other tile shapes (8 rings x 24
    detectors)."""
    out = []
    for n_bins, precision, n_sp, seed, kw in [(15, 'fp32', 30, 1, {}), (21, 'fp32', 25, 2, {}),
                                             (6, 'fp32', 20, 3, {}), (1, 'fp32', 20, 4, {}),
                                             (1, 'fp32', 24, 8, {}), (1, 'fp32', 17, 9, {'nru': 8, 'ndu': 24}),
                                             (7, 'fp64', 15, 5, {}), (15, 'fp32', 18, 6, {'nru': 8, 'ndu': 24})]:
        w = phantoms.small_tof(seed=seed, n_bins=n_bins, n_sp=n_sp, **kw)
        if seed in (8, 9):
            w = dataclasses.replace(w, tof_resolution=1e9, tof_range=1000.0)
        out.append((w, precision))
    return out


def _open(ws):
    subs = []
    for w, precision in ws:
        s = gpu_subject(w, precision)
        s.run(w.activity)
        subs.append(s)
    return subs


def _close(batch, subs):
    if batch is not None:
        batch.close()
    for s in subs:
        s.close()


def test_batch_launch_equals_each_subject():
    ws = _members()
    subs = _open(ws)
    batch = None
    try:
        want = [s.output_compact() for s in subs]
        batch = GraphBatch(subs)
        # the two 15-bin sliding-window members share one launch, and so do
        # the two one-bin members
        variants = {(p, s.info['tof_fold'], w.n_bins, w.tof_resolution) for (w, p), s in zip(ws, subs)}
        assert batch.info['scatter_launches'] == len(variants) == len(subs) - 2
        for graph in (False, True, True, False):
            batch.launch(graph=graph)
            for i in range(len(subs)):
                got = batch.output(i)
                assert got.dtype == want[i].dtype and np.array_equal(got, want[i]), (graph, i)
        assert batch.info['graph_ready']
        assert batch.elapsed_ms() > 0
    finally:
        _close(batch, subs)


def test_graph_replays_with_new_images():
    """Each replay stages new images: multiplicative noise (same zero
    voxels: the refresh keeps the fold table), and in one replay a zeroed
    block in half the members (the refresh inside the graph rebuilds their
    tables).  Between replays, another subject runs on the same context."""
    ws = _members()
    subs = _open(ws)
    other_w = phantoms.small_tof(seed=20, n_bins=9, n_sp=12)
    other = gpu_subject(other_w, 'fp32')
    batch = None
    rng = np.random.default_rng(7)
    try:
        batch = GraphBatch(subs)
        for j in range(4):
            images = []
            for i, (w, _) in enumerate(ws):
                lam = w.activity * np.clip(1.0 + 0.05 * rng.standard_normal(w.activity.shape), 0.5, 1.5)
                if j == 2 and i % 2 == 0:
                    X, Y, Z = lam.shape
                    lam[X // 3:X // 2, Y // 3:Y // 2, :] = 0.0
                images.append(lam)
                batch.set_activity(i, lam)
            batch.launch(graph=True)
            other.run(other_w.activity * (1.0 + 0.1 * j))
            other.output_compact()
            for i, (w, precision) in enumerate(ws):
                want = gpu_compact(w, precision, lam=images[i])
                assert np.array_equal(batch.output(i), want), (j, i)
    finally:
        other.close()
        _close(batch, subs)


def test_redone_walks_rebuild_the_batch():
    ws = _members()[:3]
    subs = _open(ws)
    batch = None
    try:
        want = [s.output_compact() for s in subs]
        batch = GraphBatch(subs)
        batch.launch(graph=True)
        assert batch.info['graph_ready']
        subs[1].trace()   # new walks: the captured graph no longer fits
        assert not batch.info['graph_ready']
        batch.launch(graph=True)
        assert batch.info['graph_ready']
        for i in range(len(subs)):
            assert np.array_equal(batch.output(i), want[i])
    finally:
        _close(batch, subs)


def test_member_of_a_live_batch_is_not_destroyed():
    ws = _members()[:2]
    subs = _open(ws)
    batch = GraphBatch(subs)
    try:
        with pytest.raises(_native.SssError):
            subs[0].close()
        assert subs[0].handle is not None
    finally:
        _close(batch, subs)
    assert all(s.handle is None for s in subs)


def test_batch_rejects_bad_members():
    w, _ = _members()[0]
    s = gpu_subject(w, 'fp32')
    try:
        with pytest.raises(_native.SssError):   # not traced, no activity
            GraphBatch([s])
        s.run(w.activity)
        with pytest.raises(_native.SssError):   # listed twice
            GraphBatch([s, s])
    finally:
        s.close()


def test_run_graph_batch_matches_one_instance_path():
    """Small uneven instances in resident groups (a budget that forces
    several groups), two launches per group with the instances' own images;
    and full-size random instances from the batched workload's distribution."""
    ws = []
    for seed, n_bins, n_sp in [(1, 7, 30), (2, 15, 50), (3, 1, 20), (4, 21, 40), (5, 14, 25), (6, 3, 35)]:
        ws.append(phantoms.small_tof(seed=seed, n_bins=n_bins, n_sp=n_sp))
    plain = run_plain(ws, WATER)
    from opensss_amd.serve import footprint_estimate
    budget = 2 * max(footprint_estimate(w) for w in ws)
    assert len(resident_groups(ws, budget)) >= 2
    timings = []
    # tile shapes from the tuner (whole-slice chunks: the same sums in the same order)
    tuner = TileTuner(chunk_points=(0,))
    got = run_graph_batch(ws, WATER, memory_budget=budget, iterations=2, timings=timings, tuner=tuner)
    assert sum(len(t['members']) for t in timings) == len(ws) and tuner.timed_subjects >= 1
    for a, b in zip(got, plain):
        assert np.array_equal(a, b)
    big = [phantoms.random_instance(s) for s in (11, 34, 49)]
    got = run_graph_batch(big, WATER)
    for a, b in zip(got, run_plain(big, WATER)):
        assert np.array_equal(a, b)
