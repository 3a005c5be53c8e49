"""Backprojection tail fit of the scatter scale, on the GPU.

Drop-ins for ``Python/openSSS/BackProjection.py``:
``BackProjectionScatterRandFit`` (:115-214, no span) and
``BackProjectionScatterRandFitMashing`` (:302-392, span).  The masked
scatter, prompts and randoms of every surviving LOR (masked scatter != 0) on
the AccelerationFactor detector grid are spread along the LOR's voxels
(value / n without span, value * L / sum L with span) into prompt, scatter,
randoms and sensitivity images; the scale is the least-squares slope of
max(prompts_bp - randoms_bp, 0) on scatter_bp, where x_bp = image / sensitivity
and 0 where the sensitivity is 0.

The images and the two sums sum(S T), sum(S S) are float64 here (the
reference accumulates float32 images and solves the one-column least squares
in float32 LAPACK, scaled by 1e10, :205-210); the traces and the mirrored
plane rule (a LOR from ring Ring1 to Ring2 reads plane SinogramOrder[Ring1,
Ring2], i.e. SinogramIndex[Ring2, Ring1], :157-169) follow the reference.
"""
from __future__ import annotations

import numpy as np

from . import _native
from .tails_mask import span_plane_rings


def _f32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def tail_fit(PromptSinogram, ScatterSinogram, RandomSinogram, MaskSinogram, GridSize, ImageSize, Geometry,
             SinogramCoordinates, entries, AccelerationFactor, length_weighted, *, context=None,
             return_images=False):
    """Run ``sss_tail_fit``; entries is (n, 3): ring of D1, ring of D2, value plane.
    Returns (scale f64, sum(S T), sum(S S)[, images (4, X, Y, Z) f64])."""
    geom = np.ascontiguousarray(np.asarray(Geometry, dtype=np.float64))
    n_rings, n_det = geom.shape[0], geom.shape[1]
    lor = np.ascontiguousarray(np.asarray(SinogramCoordinates, dtype=np.int16))
    if lor.shape != (n_det, n_det, 2):
        raise ValueError('SinogramCoordinates must be (N, N, 2)')
    shape = (n_det + 1, n_det // 2, -1)
    s = _f32(np.asarray(ScatterSinogram).reshape(shape))
    n_planes = s.shape[2]
    p = _f32(np.asarray(PromptSinogram, dtype=np.float32).reshape(shape))
    r = np.broadcast_to(np.asarray(RandomSinogram, dtype=np.float32), s.shape)
    r = _f32(r)
    m = _f32(np.broadcast_to(np.asarray(MaskSinogram, dtype=np.float32), s.shape))
    if p.shape != s.shape:
        raise ValueError('prompt and scatter sinograms differ in shape')
    ent = np.ascontiguousarray(np.asarray(entries, dtype=np.int32).reshape(-1, 3))
    grid = np.ascontiguousarray(np.asarray(GridSize, dtype=np.int32).reshape(3))
    bounds = np.ascontiguousarray(np.asarray(ImageSize, dtype=np.float64).reshape(6))
    sums = np.zeros(2, dtype=np.float64)
    images = np.zeros(4 * int(np.prod(grid)), dtype=np.float64) if return_images else None
    ctx = context or _native.default_context()
    lib = _native.load_library()
    accel = 1 if AccelerationFactor is None else int(AccelerationFactor)
    _native.check(lib.sss_tail_fit(ctx.handle, _native.ptr(grid), _native.ptr(bounds), n_rings, n_det,
                                   _native.ptr(geom), _native.ptr(lor), ent.shape[0], _native.ptr(ent),
                                   n_planes, accel, int(bool(length_weighted)), _native.ptr(s),
                                   _native.ptr(p), _native.ptr(r), _native.ptr(m), _native.ptr(sums),
                                   _native.ptr(images) if images is not None else None),
                  'sss_tail_fit')
    scale = sums[0] / sums[1] if sums[1] != 0 else 0.0
    if return_images:
        imgs = images.reshape(4, -1)
        imgs = np.stack([im.reshape(tuple(grid), order='F') for im in imgs])
        return scale, sums[0], sums[1], imgs
    return scale, sums[0], sums[1]


def nospan_entries(SinogramOrder, n_rings):
    """Entries (Ring1, Ring2, SinogramOrder[Ring1, Ring2]) for every ring pair."""
    order = np.asarray(SinogramOrder, dtype=np.int64).reshape(n_rings, n_rings)
    r1, r2 = np.meshgrid(np.arange(n_rings), np.arange(n_rings), indexing='ij')
    return np.stack([r1.ravel(), r2.ravel(), order.ravel()], axis=1).astype(np.int32)


def span_entries(MashedSinogramIndices):
    """Entries (int(2 R1), int(2 R2), plane i) of the span planes."""
    pr = span_plane_rings(MashedSinogramIndices)
    return np.concatenate([pr, np.arange(pr.shape[0], dtype=np.int32)[:, None]], axis=1)


def BackProjectionScatterRandFit(PromptSinogram, ScatterSinogram, RandomSinogram, MaskSinogram, GridSize,
                                 ImageSize, Geometry, SinogramCoordinates, SinogramOrder,
                                 AccelerationFactor, *, context=None):
    """Scale of the scatters to the prompt tails without span
    (BackProjection.py:115-214).  Returns a float32 array of one value, as the
    reference's ``lstsq`` solution times 1e10."""
    n_rings = np.asarray(Geometry).shape[0]
    scale, _, _ = tail_fit(PromptSinogram, ScatterSinogram, RandomSinogram, MaskSinogram, GridSize,
                           ImageSize, Geometry, SinogramCoordinates, nospan_entries(SinogramOrder, n_rings),
                           AccelerationFactor, False, context=context)
    return np.array([scale], dtype=np.float32)


def BackProjectionScatterRandFitMashing(PromptSinogram, ScatterSinogram, RandomSinogram, MaskSinogram,
                                        GridSize, ImageSize, Geometry, SinogramCoordinates,
                                        AccelerationFactor, MashedSinogramIndices, *, context=None):
    """Scale of the scatters to the prompt tails with span
    (BackProjection.py:302-392); Geometry is the extended geometry and plane
    i is traced between its rings int(2 R1), int(2 R2) (:355-360).  Returns a
    float32 array of one value."""
    scale, _, _ = tail_fit(PromptSinogram, ScatterSinogram, RandomSinogram, MaskSinogram, GridSize,
                           ImageSize, Geometry, SinogramCoordinates, span_entries(MashedSinogramIndices),
                           AccelerationFactor, True, context=context)
    return np.array([scale], dtype=np.float32)


def tail_fit_device(GridSize, ImageSize, Geometry, SinogramCoordinates, entries, n_value_planes, AccelerationFactor,
                    length_weighted, scatter_ptr, prompts_ptr, randoms_ptr, mask_ptr, *, context=None,
                    sums_ptr=None, images_ptr=None, geometry_buffers=None):
    """The tail fit on device-resident sinograms (``sss_tail_fit_noalloc``):
    scatter/prompts/randoms/mask are device addresses of float32 (N+1, N//2,
    n_value_planes) sinograms on the context's device.  geometry_buffers:
    the DeviceBuffers of (geometry, coordinates, entries) from
    ``fit_geometry_buffers`` (uploaded once, reused per TOF bin).  Queued on
    the context's stream; returns (scale, sum(S T), sum(S S)) after a
    synchronisation, or nothing when sums_ptr is given (the caller reads the
    two doubles there)."""
    ctx = context or _native.default_context()
    lib = _native.load_library()
    geom = np.asarray(Geometry)
    n_rings, n_det = geom.shape[0], geom.shape[1]
    bufs = geometry_buffers or fit_geometry_buffers(Geometry, SinogramCoordinates, entries, context=ctx)
    grid = np.ascontiguousarray(np.asarray(GridSize, dtype=np.int32).reshape(3))
    bounds = np.ascontiguousarray(np.asarray(ImageSize, dtype=np.float64).reshape(6))
    own_sums = sums_ptr is None
    sums_buf = _native.DeviceBuffer(ctx, 16) if own_sums else None
    accel = 1 if AccelerationFactor is None else int(AccelerationFactor)
    n_entries = int(np.asarray(entries).reshape(-1, 3).shape[0])
    try:
        _native.check(lib.sss_tail_fit_noalloc(
            ctx.handle, _native.ptr(grid), _native.ptr(bounds), n_rings, n_det, bufs[0].ptr, bufs[1].ptr, n_entries,
            bufs[2].ptr, int(n_value_planes), accel, int(bool(length_weighted)), _native.ctypes.c_void_p(scatter_ptr),
            _native.ctypes.c_void_p(prompts_ptr), _native.ctypes.c_void_p(randoms_ptr),
            _native.ctypes.c_void_p(mask_ptr), sums_buf.ptr if own_sums else _native.ctypes.c_void_p(sums_ptr),
            None if images_ptr is None else _native.ctypes.c_void_p(images_ptr)), 'sss_tail_fit_noalloc')
        if not own_sums:
            return None
        sums = sums_buf.download(np.empty(2, np.float64))
    finally:
        if sums_buf is not None:
            sums_buf.free()
        if geometry_buffers is None:
            for b in bufs:
                b.free()
    scale = sums[0] / sums[1] if sums[1] != 0 else 0.0
    return scale, float(sums[0]), float(sums[1])


def fit_geometry_buffers(Geometry, SinogramCoordinates, entries, *, context=None):
    """Device copies of the geometry (f64), the coordinates (int16) and the
    fit entries (int32 (n, 3)) for ``tail_fit_device``."""
    ctx = context or _native.default_context()
    geom = np.ascontiguousarray(np.asarray(Geometry, dtype=np.float64))
    lor = np.ascontiguousarray(np.asarray(SinogramCoordinates, dtype=np.int16))
    ent = np.ascontiguousarray(np.asarray(entries, dtype=np.int32).reshape(-1, 3))
    return (_native.DeviceBuffer.from_array(ctx, geom), _native.DeviceBuffer.from_array(ctx, lor),
            _native.DeviceBuffer.from_array(ctx, ent if ent.size else np.zeros(3, np.int32)))
