"""Sinogram-coordinate helpers (host side, numpy).

Drop-in for ``Python/openSSS/Sinogram.py:27-141`` (``SinogramCoordinates``) and
``Python/openSSS/DataReduction.py:25-100`` (``ComputeLORCoordinates``).  The
reference evaluates these with numba double loops; here the same integer
arithmetic is evaluated on whole detector-pair grids at once.  Geometry-only
tables: built once per scanner and uploaded to HBM by the scatter subject.
"""
from __future__ import annotations

import numpy as np


def _pair_coordinates(n_crystals: int, shift: int, min_crystal_difference: int,
                      radial_size: int) -> np.ndarray:
    """(angular, radial) of every ordered detector pair, -1 where skipped.

    Integer rules follow ``Sinogram.py:66-122``: crystal ids are shifted by
    ``shift`` (wrapping below zero), ordered into IdA <= IdB, and the radial /
    angular bins come from floor divisions on IdA+IdB and IdB-IdA.
    """
    N = int(n_crystals)
    ids = np.arange(N, dtype=np.int64) % N - shift
    ids = np.where(ids < 0, ids + N, ids)
    c1 = ids[:, None]
    c2 = ids[None, :]
    ida = np.minimum(c1, c2)
    idb = np.maximum(c1, c2)
    ssum = ida + idb
    diff = idb - ida
    outer = (ssum >= (3 * N) // 2) | (ssum < N // 2)
    # floor-division semantics of Python ints (numba keeps them)
    rad_outer = np.where(diff == 0, (-N) // 2,
                         np.floor_divide(diff - 1, 2) - np.floor_divide(N - (diff + 1), 2))
    rad_inner = np.where(diff == 0, N // 2,
                         np.floor_divide(N - (diff + 1), 2) - np.floor_divide(diff - 1, 2))
    radial = np.where(outer, rad_outer, rad_inner)
    angular = np.where(
        ssum < N // 2, np.floor_divide(2 * ida + N + radial, 2),
        np.where(ssum >= (3 * N) // 2, np.floor_divide(2 * ida - N + radial, 2),
                 np.floor_divide(2 * ida - radial, 2)))
    out = np.full((N, N, 2), -1, dtype=np.int64)
    keep = diff >= min_crystal_difference
    out[..., 0] = np.where(keep, angular, -1)
    out[..., 1] = np.where(keep, radial + radial_size // 2, -1)
    return out


def sinogram_index_table(n_rings: int) -> np.ndarray:
    """Ring pair -> sinogram slice, closed form of ``Sinogram.py:124-139``."""
    NR = int(n_rings)
    r1 = np.arange(NR)[:, None]
    r2 = np.arange(NR)[None, :]
    d = np.abs(r1 - r2)
    base = (2 * d - 1) * NR - d * (d - 1)
    idx = np.where(r2 < r1, base + r2, base + (NR - d) + r1)
    idx = np.where(d == 0, np.broadcast_to(r1, (NR, NR)), idx)
    return idx.astype(np.int64)


def SinogramCoordinates(NrSectorsTrans: int, NrSectorsAxial: int, NrModulesAxial: int,
                        NrModulesTrans: int, NrCrystalsTrans: int, NrCrystalsAxial: int,
                        MinSectorDifference: int = 0, ID_Shidt: bool = True):
    """Same signature and outputs as ``Sinogram.py:28-141``.

    Returns ``(LORCoordinates int16 (N, N, 2), SinogramIndex int16 (NR, NR),
    DetectorShift int)``; ``LORCoordinates[d1, d2] = (angular, radial)``.
    """
    n_rings = NrSectorsAxial * NrModulesAxial * NrCrystalsAxial
    n_crystals = NrSectorsTrans * NrModulesTrans * NrCrystalsTrans
    min_cd = MinSectorDifference * NrModulesTrans * NrCrystalsTrans
    radial_size = n_crystals - 2 * (min_cd - 1) - 1
    shift = (NrModulesTrans * NrCrystalsTrans) // 2 if ID_Shidt else 0
    lor = _pair_coordinates(n_crystals, shift, min_cd, radial_size)
    sino = sinogram_index_table(n_rings)
    return lor.astype(np.int16), sino.astype(np.int16), int(shift)


def ComputeLORCoordinates(NrSectorsTrans: int, NrModulesTrans: int, NrCrystalsTrans: int,
                          MinSectorDifference: int = 0) -> np.ndarray:
    """``DataReduction.py:25-100``: the shifted pair table, skipped pairs left at 0."""
    n_crystals = NrSectorsTrans * NrModulesTrans * NrCrystalsTrans
    min_cd = MinSectorDifference * NrModulesTrans * NrCrystalsTrans
    radial_size = n_crystals - 2 * (min_cd - 1) - 1
    shift = (NrModulesTrans * NrCrystalsTrans) // 2
    lor = _pair_coordinates(n_crystals, shift, min_cd, radial_size)
    lor[lor < 0] = 0
    return lor.astype(np.int16)
