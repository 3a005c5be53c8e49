"""Span (axial mashing) helpers, host side.

Drop-ins for ``Python/openSSS/DataReduction.py:102-172`` (``SinogramLUT``) and
``:174-218`` (``SumLORsCounts``).  ``ComputeLORCoordinates`` lives in
:mod:`opensss_amd.sinogram` and is re-exported here.
"""
from __future__ import annotations

import numpy as np

from .sinogram import ComputeLORCoordinates  # noqa: F401  (same module surface)


def _span_plane_count(n_rings: int, span: int) -> int:
    """Plane count formula of ``DataReduction.py:123``."""
    k = np.arange(1, np.round(n_rings / span) + 1)
    return int(n_rings * 2 - 1 + np.sum(4 * (n_rings - (span + 1) / 2 - span * (k - 1)) - 2))


def _span_plane_of(ring1: int, ring2: int, n_rings: int, span: int) -> int:
    """1-based span plane of the (1-based) ring pair, ``DataReduction.py:135-152``.

    Planes of segment 0 come first (ring1+ring2-1); every further segment adds
    two blocks of 2*(NR - (span+1)/2 - span*(i-1)) - 1 planes, the negative
    ring-difference block before the positive one.
    """
    half = (span + 1) / 2
    diff = ring2 - ring1
    if abs(diff) < half:
        return int(ring1 + ring2 - 1)
    seg = int(np.floor((abs(diff) - half) / span)) + 1
    pos = 2 * n_rings - 1
    last = seg + 1 if diff > 0 else seg
    for i in range(1, last):
        jump = 1 if (i == last - 1 and diff > 0) else 2
        pos = pos + ((n_rings - half - span * (i - 1)) * 2 - 1) * jump
    return int((ring1 + ring2 - 1) - seg * span + (span - 1) / 2 + pos)


def SinogramLUT(NrRings: int, AxialMashing: int, MaxRingDifference: int | None = None):
    """Span-mashed plane table with the reference's outputs and dtypes.

    Returns ``(SinogramIndex f32 (P, 2), RingLUT bool (2NR-1, 2NR-1),
    LookUpTable int16 (NR, NR), SinogramCounts f32 (P_all,))`` where
    ``SinogramIndex[p] = (mean ring2, mean ring1)`` over the contributing pairs,
    computed in float32 exactly as ``DataReduction.py:154-164``.
    """
    NR = int(NrRings)
    span = int(AxialMashing)
    n_planes = _span_plane_count(NR, span)
    ring_sum = np.zeros((n_planes, 2), dtype=np.float32)
    counts = np.zeros(n_planes, dtype=np.float32)
    lut = np.full((NR, NR), -1, dtype=np.int16)
    for r1 in range(1, NR + 1):
        for r2 in range(1, NR + 1):
            if MaxRingDifference is not None and abs(r1 - r2) > MaxRingDifference:
                continue
            p = _span_plane_of(r1, r2, NR, span) - 1
            ring_sum[p, 0] += r2 - 1
            ring_sum[p, 1] += r1 - 1
            counts[p] += 1
            lut[r1 - 1, r2 - 1] = p
    ring_sum[:, 0] /= counts + np.float32(1e-8)
    ring_sum[:, 1] /= counts + np.float32(1e-8)
    plane_rings = ring_sum[counts != 0, :]
    ring_lut = np.zeros((NR * 2 - 1, NR * 2 - 1), dtype=bool)
    for p in range(plane_rings.shape[0]):
        ring_lut[int(plane_rings[p, 0] * 2), int(plane_rings[p, 1] * 2)] = True
    return plane_rings, ring_lut, lut, counts


def SumLORsCounts(Geometry: np.ndarray, LookUpTable: np.ndarray, SinogramIndex: np.ndarray,
                  Mash: int, Span: int, NrSinograms: int) -> np.ndarray:
    """Per-cell LOR multiplicity, ``DataReduction.py:174-218``.

    Every ring pair contributes Mash**2 to each cell of its (span) plane.
    """
    n_det = Geometry.shape[1]
    n_ang = n_det // 2
    NR = Geometry.shape[0]
    planes = LookUpTable if Span > 1 else SinogramIndex
    n_out = NrSinograms if Span > 1 else NR ** 2
    per_plane = np.zeros(n_out, dtype=np.float64)
    np.add.at(per_plane, np.asarray(planes[:NR, :NR], dtype=np.int64).ravel(), 1.0)
    out = np.empty((n_det + 1, n_ang, n_out), dtype=np.float32)
    out[...] = (per_plane * float(Mash ** 2)).astype(np.float32)[None, None, :]
    return out
