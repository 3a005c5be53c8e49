"""Exact-rational voxel walk used to pin the oracle's tracer.

Independent formulation of the reference's traversal
(``Python/openSSS/RayTracing3DTOF.py:29-180``): every plane crossing of the
clipped segment is listed with its exact parameter t (``fractions.Fraction``),
crossings are ordered by t and, at equal t, z before y before x (what the
reference's strict comparisons x<y, x<z, y<z produce), and the visited voxels
and lengths follow from the sorted crossings.  No float64 arithmetic.
"""
from fractions import Fraction as F
import math


def _clip(p0, p1, lo, hi):
    tmin, tmax = F(0), F(1)
    for a in range(3):
        d = p1[a] - p0[a]
        if d == 0:
            if p0[a] < lo[a] or p0[a] > hi[a]:
                return None
            continue
        t0 = (lo[a] - p0[a]) / d
        t1 = (hi[a] - p0[a]) / d
        if t0 > t1:
            t0, t1 = t1, t0
        tmin = max(tmin, t0)
        tmax = min(tmax, t1)
    if tmin >= tmax:
        return None
    return tmin, tmax


def exact_walk(grid, bounds, seg):
    """Returns (list of 1-based linear indices, list of exact lengths as Fraction)."""
    n = [int(v) for v in grid]
    lo = [F(bounds[i]) for i in range(3)]
    hi = [F(bounds[i + 3]) for i in range(3)]
    p0 = [F(seg[i]) for i in range(3)]
    p1 = [F(seg[i + 3]) for i in range(3)]
    clip = _clip(p0, p1, lo, hi)
    if clip is None:
        return [], []
    tmin, tmax = clip
    vox = [(hi[a] - lo[a]) / n[a] for a in range(3)]
    start = [p0[a] + (p1[a] - p0[a]) * tmin for a in range(3)]
    cur = [min(max(1, math.ceil((start[a] - lo[a]) / vox[a])), n[a]) for a in range(3)]
    events = []
    for a in range(3):
        d = p1[a] - p0[a]
        if d == 0:
            continue
        step = 1 if d > 0 else -1
        k = cur[a] if d > 0 else cur[a] - 1   # next plane index in the direction of travel
        while True:
            t = (lo[a] + k * vox[a] - p0[a]) / d
            if t >= tmax:
                break
            events.append((t, 2 - a, a, step))
            k += step
    events.sort()
    seq = [list(cur)]
    ts = [tmin]
    for t, _, a, step in events:
        cur[a] += step
        seq.append(list(cur))
        ts.append(t)
    ts.append(tmax)
    length = math.sqrt(float(sum((p1[a] - p0[a]) ** 2 for a in range(3))))
    idx = [x + (y - 1) * n[1] + (z - 1) * n[1] * n[0] for x, y, z in seq]
    lens = [(ts[i + 1] - ts[i]) * F(length) for i in range(len(seq))]
    return idx, lens
