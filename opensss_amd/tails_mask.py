"""Tail mask of the scatter scaling, on the GPU.

Drop-ins for ``Python/openSSS/TailsMask.py``: ``GenerateTailsMask`` (:28-82),
``MaskGenerator`` (:165-228, no span) and ``MaskGeneratorMashing``
(:311-376, span).  Every detector pair D1 <= D2 on the AccelerationFactor
grid of a ring pair is traced through the μ-map (or the activity map); a
sinogram cell is a tail when the float64 factor 1/exp(-sum mu L) is exactly 1
(or, in the activity variant, when sum lam L is not positive).  The traces
run in ``sss_tail_mask``.

As in the reference, the activity threshold of :204-206 / :350-352 is not
applied: its result is overwritten by the flatten of the unthresholded map
on the next line.
"""
from __future__ import annotations

import os

import numpy as np

from . import _native


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _run_mask(ActivityMap, AttenuationMap, ImageSize, Geometry, SinogramCoordinates, plane_rings,
              UseAttenuation, AccelerationFactor, context=None):
    image = np.asarray(AttenuationMap if UseAttenuation else ActivityMap)
    if image.ndim != 3:
        raise ValueError('maps must be 3-D (X, Y, Z)')
    geom = _f64(Geometry)
    n_rings, n_det = geom.shape[0], geom.shape[1]
    lor = np.ascontiguousarray(np.asarray(SinogramCoordinates, dtype=np.int16))
    if lor.shape != (n_det, n_det, 2):
        raise ValueError('SinogramCoordinates must be (N, N, 2)')
    pr = np.ascontiguousarray(np.asarray(plane_rings, dtype=np.int32).reshape(-1, 2))
    n_planes = pr.shape[0]
    grid = np.ascontiguousarray(np.asarray(image.shape, dtype=np.int32))
    bounds = _f64(ImageSize)
    flat = _f64(image.flatten(order='F'))
    mask = np.zeros((n_det + 1, n_det // 2, n_planes), dtype=np.uint8)
    ctx = context or _native.default_context()
    lib = _native.load_library()
    accel = 1 if AccelerationFactor is None else int(AccelerationFactor)
    _native.check(lib.sss_tail_mask(ctx.handle, _native.ptr(grid), _native.ptr(bounds), _native.ptr(flat),
                                    int(bool(UseAttenuation)), n_rings, n_det, _native.ptr(geom),
                                    _native.ptr(lor), n_planes, _native.ptr(pr), accel, _native.ptr(mask)),
                  'sss_tail_mask')
    return mask


def nospan_plane_rings(SinogramIndex, n_rings):
    """(NR^2, 2) ring pairs of the output planes without span: plane s holds
    the ring pair (Ring1, Ring2) with SinogramIndex[Ring1, Ring2] = s (the
    reshape and argsort reorder of TailsMask.py:215-219)."""
    sino = np.asarray(SinogramIndex)[:n_rings, :n_rings].astype(np.int64)
    pr = np.full((n_rings * n_rings, 2), -1, dtype=np.int32)
    r1, r2 = np.meshgrid(np.arange(n_rings), np.arange(n_rings), indexing='ij')
    pr[sino.ravel(), 0] = r1.ravel()
    pr[sino.ravel(), 1] = r2.ravel()
    if (pr < 0).any():
        raise ValueError('SinogramIndex must be a permutation of 0 .. NR^2 - 1')
    return pr


def span_plane_rings(MashedSinogramIndices):
    """(n_planes, 2) rings of the extended geometry for each span plane:
    int(2 * averaged ring) (TailsMask.py:357-359)."""
    msi = np.asarray(MashedSinogramIndices, dtype=np.float64)
    return np.stack([(msi[:, 0] * 2).astype(np.int64), (msi[:, 1] * 2).astype(np.int64)], axis=1).astype(np.int32)


def MaskGenerator(ActivityMap, AttenuationMap, ImageSize, Geometry, SinogramCoordinates, SinogramIndex,
                  UseAttenuation, AccelerationFactor=1, *, context=None):
    """uint8 (N+1, N//2, NR^2) tail mask without span (TailsMask.py:165-228)."""
    n_rings = np.asarray(Geometry).shape[0]
    return _run_mask(ActivityMap, AttenuationMap, ImageSize, Geometry, SinogramCoordinates,
                     nospan_plane_rings(SinogramIndex, n_rings), UseAttenuation, AccelerationFactor,
                     context)


def MaskGeneratorMashing(ActivityMap, AttenuationMap, ImageSize, Geometry, SinogramCoordinates,
                         MashedSinogramIndices, UseAttenuation, AccelerationFactor=1, *, context=None):
    """uint8 (N+1, N//2, n_planes) tail mask with span (TailsMask.py:311-376);
    Geometry is the extended (2 NR - 1 ring) geometry."""
    return _run_mask(ActivityMap, AttenuationMap, ImageSize, Geometry, SinogramCoordinates,
                     span_plane_rings(MashedSinogramIndices), UseAttenuation, AccelerationFactor,
                     context)


def GenerateTailsMask(SavePath, ActivityMapDownscaled, AttenuationMapDownscaled, ImageSize, Geometry,
                      extendedGeometry, LORCoordinates, SinogramIndex, AccelerationFactor,
                      UseAttenuation=True, Overwrite=False, *, context=None):
    """Tail mask with the reference's cache file ``AttenuationMask.npy``
    (TailsMask.py:28-82): computed when missing or Overwrite, else loaded."""
    path = os.path.join(SavePath, 'AttenuationMask.npy')
    if os.path.isfile(path) and not Overwrite:
        return np.load(path)
    print('Start generating AttenuationMask')
    if np.asarray(extendedGeometry).shape[0] != 0:
        print('With span')
        mask = MaskGeneratorMashing(ActivityMapDownscaled, AttenuationMapDownscaled, ImageSize,
                                    extendedGeometry, LORCoordinates, SinogramIndex, UseAttenuation,
                                    AccelerationFactor, context=context)
    else:
        print('No span')
        mask = MaskGenerator(ActivityMapDownscaled, AttenuationMapDownscaled, ImageSize, Geometry,
                             LORCoordinates, SinogramIndex, UseAttenuation, AccelerationFactor,
                             context=context)
    np.save(path, mask)
    print('completed!\n')
    return mask
