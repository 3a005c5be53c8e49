"""Pins the oracle's float64 voxel tracer (restating RayTracing3DTOF.py:29-180).

Checks that a plausible mistake in the walk would fail: exact-rational brute
force on small grids (voxel sequence, lengths), the analytic segment/box
length, the TOF sample recurrence s_k = s_{k-1} + L_k/2, and hand-derived
tie and axis-parallel cases.
"""
from fractions import Fraction

import numpy as np
import pytest

from tests.exact_tracer import exact_walk


def _rand_segments(rng, n, bounds, scale=1.6):
    lo = np.array(bounds[:3])
    hi = np.array(bounds[3:])
    c = (lo + hi) / 2
    h = (hi - lo) / 2
    a = c + (rng.uniform(-1, 1, size=(n, 3)) * h * scale)
    b = c + (rng.uniform(-1, 1, size=(n, 3)) * h * scale)
    return np.concatenate([a, b], axis=1)


@pytest.mark.parametrize('grid,bounds', [
    ((4, 4, 3), (-2.0, -2.0, -1.5, 2.0, 2.0, 1.5)),
    ((6, 6, 5), (-3.1, -3.1, -2.2, 3.1, 3.1, 2.2)),
    ((5, 5, 7), (0.0, 0.0, 0.0, 1.0, 1.0, 1.4)),
])
def test_trace_matches_exact_rational_walk(oracle, grid, bounds):
    rng = np.random.default_rng(11)
    segs = _rand_segments(rng, 300, bounds)
    checked = 0
    for seg in segs:
        L, idx, s = oracle.trace(grid, bounds, seg)
        eidx, elen = exact_walk(grid, bounds, seg)
        if len(eidx) == 0:
            assert len(idx) == 0
            continue
        assert list(idx) == eidx
        el = np.array([float(v) for v in elen])
        np.testing.assert_allclose(L, el, rtol=1e-12, atol=1e-13)
        checked += 1
    assert checked > 150


def test_total_length_is_segment_inside_box(oracle):
    grid, bounds = (10, 10, 8), (-5.0, -5.0, -4.0, 5.0, 5.0, 4.0)
    rng = np.random.default_rng(3)
    for seg in _rand_segments(rng, 200, bounds, scale=2.0):
        L, idx, s = oracle.trace(grid, bounds, seg)
        p0, p1 = seg[:3], seg[3:]
        # analytic clip of the segment with the box (slab method, exact fractions)
        t0, t1 = Fraction(0), Fraction(1)
        inside = True
        for a in range(3):
            d = Fraction(p1[a]) - Fraction(p0[a])
            lo, hi = Fraction(bounds[a]), Fraction(bounds[a + 3])
            if d == 0:
                inside &= lo <= Fraction(p0[a]) <= hi
                continue
            u, v = (lo - Fraction(p0[a])) / d, (hi - Fraction(p0[a])) / d
            t0, t1 = max(t0, min(u, v)), min(t1, max(u, v))
        expect = float(max(t1 - t0, 0)) * float(np.linalg.norm(p1 - p0)) if inside else 0.0
        assert abs(L.sum() - expect) <= 1e-11 * max(1.0, expect)


def test_sample_recurrence_is_half_cumulative(oracle):
    """s_k = s_{k-1} + L_k/2 (the reference's quirk, not the voxel midpoint)."""
    grid, bounds = (8, 8, 6), (-4.0, -4.0, -3.0, 4.0, 4.0, 3.0)
    L, idx, s = oracle.trace(grid, bounds, [-3.7, -2.2, -2.5, 3.1, 3.3, 2.1])
    expect = np.cumsum(L / 2)
    np.testing.assert_array_equal(s, expect)
    midpoint = np.cumsum(L) - L / 2
    assert not np.allclose(s[1:], midpoint[1:])


def test_diagonal_tie_steps_y_before_x(oracle):
    """x and y planes crossed at the same t: y advances first (zero-length voxel)."""
    grid, bounds = (4, 4, 4), (0.0, 0.0, 0.0, 4.0, 4.0, 4.0)
    L, idx, s = oracle.trace(grid, bounds, [0.5, 0.5, 0.5, 3.5, 3.5, 0.5])
    xyz = [(1, 1, 1), (1, 2, 1), (2, 2, 1), (2, 3, 1), (3, 3, 1), (3, 4, 1), (4, 4, 1)]
    expect_idx = [x + (y - 1) * 4 + (z - 1) * 16 for x, y, z in xyz]
    assert list(idx) == expect_idx
    d = np.sqrt(2.0)
    np.testing.assert_allclose(L, [0.5 * d, 0.0, d, 0.0, d, 0.0, 0.5 * d], atol=1e-15)


def test_triple_tie_steps_z_first(oracle):
    grid, bounds = (3, 3, 3), (0.0, 0.0, 0.0, 3.0, 3.0, 3.0)
    L, idx, s = oracle.trace(grid, bounds, [0.5, 0.5, 0.5, 2.5, 2.5, 2.5])
    xyz = [(1, 1, 1), (1, 1, 2), (1, 2, 2), (2, 2, 2), (2, 2, 3), (2, 3, 3), (3, 3, 3)]
    assert list(idx) == [x + (y - 1) * 3 + (z - 1) * 9 for x, y, z in xyz]


def test_axis_parallel_and_miss(oracle):
    grid, bounds = (5, 5, 2), (0.0, 0.0, 0.0, 5.0, 5.0, 2.0)
    L, idx, s = oracle.trace(grid, bounds, [-1.0, 2.5, 0.5, 7.0, 2.5, 0.5])
    assert list(idx) == [x + 2 * 5 for x in range(1, 6)]
    np.testing.assert_allclose(L, np.ones(5), atol=1e-14)
    L, idx, s = oracle.trace(grid, bounds, [-1.0, 7.5, 0.5, 7.0, 7.5, 0.5])
    assert len(idx) == 0


def test_start_on_face_gives_zero_first_length(oracle):
    """A scatter point on a voxel face (x, y on the upper edge) starts with L = 0."""
    grid, bounds = (4, 4, 2), (-2.0, -2.0, -1.0, 2.0, 2.0, 1.0)
    L, idx, s = oracle.trace(grid, bounds, [0.0, 0.0, 0.5, 5.0, 1.0, 0.5])
    assert L[0] == 0.0
    assert idx[0] == 2 + (2 - 1) * 4 + (2 - 1) * 16


def test_segment_ending_inside(oracle):
    grid, bounds = (6, 6, 6), (0.0, 0.0, 0.0, 6.0, 6.0, 6.0)
    seg = [0.3, 0.2, 0.1, 2.7, 4.1, 3.3]
    L, idx, s = oracle.trace(grid, bounds, seg)
    eidx, elen = exact_walk(grid, bounds, seg)
    assert list(idx) == eidx
    assert abs(L.sum() - np.linalg.norm(np.array(seg[3:]) - seg[:3])) < 1e-12


def test_walk_inside_voxel_face_is_truncated_at_expected_count(oracle):
    """Segment starting on the x = 0 face with a tiny x drift: the float64 end
    voxel (x) disagrees with the first step, the reference loop would not end.
    The walk stops after |dX|+|dY|+|dZ|+1 voxels and keeps the total length."""
    grid, bounds = (8, 8, 4), (-4.0, -4.0, -2.0, 4.0, 4.0, 2.0)
    seg = [0.0, 1.0, -0.5, 6.123233995736766e-16, 10.0, -1.0]
    L, idx, s = oracle.trace(grid, bounds, seg)
    clip_t = (4.0 - 1.0) / 9.0
    total = clip_t * np.linalg.norm(np.array(seg[3:]) - seg[:3])
    assert len(idx) == 4           # start y=5 .. end y=8, no x step counted
    assert abs(L.sum() - total) < 1e-12
