"""Tile-shape autotuner of the scatter launch.

A scatter launch's work unit is one wave's tile of 8 x 8 detector-pair cells
for one chunk of scatter points, across all TOF bins.  Its shape has two
knobs that change no physics:

- ``tile_rings``: rings per block of a tile's rows and columns (1, 2, 4 or
  8; the block holds 8 / tile_rings detectors).  Lanes of a tile row walk
  half paths of neighbouring rings or detectors, so this sets how evenly the
  lanes of a wave stay busy.
- ``n_chunks``: scatter points per wave work item (whole normalisation
  slices by default, or fewer points per item).  Each chunk costs a wave's
  set-up and a partial array to reduce; shorter chunks keep more waves
  resident on fewer half paths.

Which shape is fastest depends on the subject's size.  `TileTuner` times the
candidates on a traced subject.  ``ScatterSubject.set_tiling`` re-tiles it
without redoing its walks, and HIP events bracket the scatter kernel.  The
tuner keeps the best shape per shape key: mean walk length (from the voxel
grid), TOF bins, sampled detectors and rings, and precision.  Keys live in
memory and optionally in a JSON file, so later subjects of a key take the
tuned shape without timing.  The tile shape never changes an output value;
``n_chunks`` changes only the grouping in which a slice's chunk sums are
added (FP32 rounding).
"""
from __future__ import annotations

import json
import math
import os
from typing import Dict, Optional, Sequence, Tuple

Shape = Tuple[int, int]   # (tile_rings, n_chunks); 0 = the library default


def shape_key(inst, precision: str = 'fp32') -> str:
    """Tuning key of an instance: walk-length bucket (half the voxel grid's
    axis sum, to the nearest half power of two), bins, sampled detectors and
    rings, precision."""
    X, Y, Z = (int(v) for v in inst.attenuation.shape)
    walk = max(1.0, (X + Y + Z) / 2.0)
    bucket = round(2.0 * math.log2(walk)) / 2.0
    return f'{precision}:walk{bucket:g}:b{int(inst.n_bins)}:ndu{int(inst.n_det_used)}:nru{int(inst.n_rings_used)}'


class TileTuner:
    """Best tile shape per shape key, found by timing.

    ring_candidates: tile_rings values to try (those above the sampled rings
    or with more detectors per block than sampled detectors are skipped);
    chunk_points: scatter points per chunk to try (0: the default, whole
    slices of up to 64 points); repeats: timings per candidate (the fastest
    counts).  path: JSON file that keeps the results between runs.
    """

    def __init__(self, path: Optional[str] = None, ring_candidates: Sequence[int] = (1, 2, 4, 8),
                 chunk_points: Sequence[int] = (0, 32), repeats: int = 2):
        self.path = path
        self.ring_candidates = tuple(int(r) for r in ring_candidates)
        self.chunk_points = tuple(int(c) for c in chunk_points)
        self.repeats = max(1, int(repeats))
        self.best: Dict[str, Shape] = {}
        self.times: Dict[str, Dict[str, float]] = {}
        self.timed_subjects = 0
        if path and os.path.exists(path):
            with open(path) as f:
                data = json.load(f)
            self.best = {k: (int(v[0]), int(v[1])) for k, v in data.get('best', {}).items()}
            self.times = data.get('times', {})

    def candidates(self, inst) -> list:
        nru, ndu = int(inst.n_rings_used), int(inst.n_det_used)
        n_sp = 0 if inst.scatter_points is None else len(inst.scatter_points)
        rings = [r for r in self.ring_candidates if r <= nru and 8 // r <= max(1, ndu)] or [0]
        chunks = []
        for p in self.chunk_points:
            n = 0 if p <= 0 else max(1, -(-n_sp // p))
            if n not in chunks:
                chunks.append(n)
        return [(r, c) for r in rings for c in chunks]

    def tune(self, subject, inst, precision: str = 'fp32', activity=None) -> Shape:
        """Times every candidate shape on `subject` (traced; given `activity`,
        default the instance's image, if it has none yet), leaves it on the
        fastest and records that shape under the instance's key."""
        key = shape_key(inst, precision)
        if activity is not None or not getattr(subject, 'has_activity', True):
            subject.set_activity(inst.activity if activity is None else activity)
        times = {}
        for shape in self.candidates(inst):
            subject.set_tiling(*shape)
            ms = []
            for _ in range(self.repeats):
                subject.scatter()
                ms.append(subject.scatter_kernel_ms())
            times[shape] = min(ms)
        best = min(times, key=times.get)
        subject.set_tiling(*best)
        self.best[key] = best
        self.times[key] = {f'{r},{c}': t for (r, c), t in times.items()}
        self.timed_subjects += 1
        self.save()
        return best

    def apply(self, subject, inst, precision: str = 'fp32', activity=None) -> Shape:
        """The key's tuned shape on `subject`, timing the candidates first
        when the key is new."""
        key = shape_key(inst, precision)
        if key not in self.best:
            return self.tune(subject, inst, precision, activity)
        shape = self.best[key]
        if shape != (0, 0):
            subject.set_tiling(*shape)
        return shape

    def save(self) -> None:
        if not self.path:
            return
        tmp = self.path + '.tmp'
        with open(tmp, 'w') as f:
            json.dump({'best': {k: list(v) for k, v in self.best.items()}, 'times': self.times}, f, indent=1)
        os.replace(tmp, self.path)
