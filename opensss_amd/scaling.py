"""Scale of the scatter estimate to the prompts, per TOF bin.

Drop-in for ``Python/openSSS/Scaling.py:27-121`` (``ScaleScattersToPrompts``):
the same files are read from and written to ``SavePath`` --
``PromptsSinogram.npz`` (or ``PromptsSinogramMashed.npz`` with span),
optional ``NormSinogram.npz`` and ``RandomsSinogram.npz``,
``AttenuationMask.npy`` and ``SSS_mashed_bin{b}.npz`` -- and the float32
scale factors go to ``ScaleFactors.npy``.

Every bin is fitted in one ``sss_scale_to_prompts`` call: the division of the
prompts by the LOR counts (``:90``), the randoms times the normalisation
(``:69-71``) and the mask threshold (``:75-76``) run inside the fit kernels,
and each LOR is traced once per group of eight bins.  Bin b's sums are the
same bits as ``sss_tail_fit`` of that bin with the prompts divided
beforehand (see :mod:`opensss_amd.backprojection` for the fit itself).
Prompts are read as float32 (uint16 counts, as the list-mode histogrammer
writes them, go to the device as they are).
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import _native
from .backprojection import nospan_entries, span_entries


def _load_npz(path):
    with np.load(path) as z:
        return z['arr_0']


def _counts_arg(LORCounts, cell_shape):
    """(array or None, per_plane flag) for the prompts divisor."""
    if LORCounts is None:
        return None, 0
    c = np.asarray(LORCounts, dtype=np.float32)
    if c.ndim == 0:
        return np.full(cell_shape[2], c, dtype=np.float32), 1
    if c.ndim == 1:
        if c.shape[0] != cell_shape[2]:
            raise ValueError(f'LORCounts has {c.shape[0]} planes, the prompts {cell_shape[2]}')
        return np.ascontiguousarray(c), 1
    try:
        return np.ascontiguousarray(np.broadcast_to(c, cell_shape)), 0
    except ValueError:
        raise ValueError(f'LORCounts of shape {c.shape} does not broadcast to the prompts {cell_shape} '
                         '(operands could not be broadcast together, as in Scaling.py:90)') from None


def fit_entries(SinogramIndex, Geometry, ExtendedGeometry, SpanFlag):
    """(geometry the LORs are traced in, fit entries, length weighted) of the
    no-span (Scaling.py:92-98) or span (:99-104) branch."""
    if not SpanFlag:
        NR = np.asarray(SinogramIndex).shape[0]
        order = np.asarray(SinogramIndex)[0:NR, 0:NR].T
        return np.asarray(Geometry), nospan_entries(order, np.asarray(Geometry).shape[0]), False
    return np.asarray(ExtendedGeometry), span_entries(SinogramIndex), True


def scale_to_prompts(Prompts, Scatters, Randoms, Normalization, Mask, LORCounts, SinogramIndex,
                     SinogramCoordinates, DesiredDimensions, FittingSize, Geometry, ExtendedGeometry,
                     AccelerationFactor, SpanFlag, *, context=None, return_sums=False):
    """In-memory form of the scaling loop: Prompts (N+1, N/2, P, B), Scatters
    (B, N+1, N/2, P) float32, Randoms and Normalization (N+1, N/2, P) or
    None, Mask (N+1, N/2, P) (cells >= 1 are fitted), LORCounts (N+1, N/2, P),
    (P,), a scalar or None.  Returns float32 (B,) scale factors (and the
    (B, 2) float64 sums when return_sums)."""
    prompts = np.asarray(Prompts)
    if prompts.ndim != 4:
        raise ValueError('Prompts must be (N+1, N/2, P, B)')
    cell_shape = prompts.shape[:3]
    n_bins = prompts.shape[3]
    if prompts.dtype == np.uint16:
        p_arr, p_type = np.ascontiguousarray(prompts), _native.SSS_PROMPTS_U16
    else:
        p_arr, p_type = np.ascontiguousarray(prompts, dtype=np.float32), _native.SSS_PROMPTS_F32
    scat = np.ascontiguousarray(np.asarray(Scatters, dtype=np.float32))
    if scat.shape != (n_bins,) + cell_shape:
        raise ValueError(f'Scatters must be {(n_bins,) + cell_shape}, got {scat.shape}')
    m = np.asarray(Mask)
    m = np.ascontiguousarray(m if m.dtype == np.uint8 else (m >= 1).astype(np.uint8))
    if m.shape != cell_shape:
        raise ValueError('Mask and prompts differ in shape')
    r = None if Randoms is None else np.ascontiguousarray(np.broadcast_to(np.asarray(Randoms, np.float32),
                                                                          cell_shape))
    nrm = None
    if r is not None and Normalization is not None:
        nrm = np.ascontiguousarray(np.broadcast_to(np.asarray(Normalization, np.float32), cell_shape))
    counts, per_plane = _counts_arg(LORCounts, cell_shape)

    geom, entries, weighted = fit_entries(SinogramIndex, Geometry, ExtendedGeometry, SpanFlag)
    geom = np.ascontiguousarray(geom, dtype=np.float64)
    n_rings, n_det = geom.shape[0], geom.shape[1]
    lor = np.ascontiguousarray(np.asarray(SinogramCoordinates, dtype=np.int16))
    if lor.shape != (n_det, n_det, 2) or cell_shape[:2] != (n_det + 1, n_det // 2):
        raise ValueError('SinogramCoordinates, geometry and sinograms disagree')
    ent = np.ascontiguousarray(np.asarray(entries, dtype=np.int32).reshape(-1, 3))
    grid = np.ascontiguousarray(np.asarray(DesiredDimensions, dtype=np.int32).reshape(3))
    bounds = np.ascontiguousarray(np.asarray(FittingSize, dtype=np.float64).reshape(6))
    sums = np.zeros((n_bins, 2), dtype=np.float64)
    ctx = context or _native.default_context()
    lib = _native.load_library()
    accel = 1 if AccelerationFactor is None else int(AccelerationFactor)
    _native.check(lib.sss_scale_to_prompts(
        ctx.handle, _native.ptr(grid), _native.ptr(bounds), n_rings, n_det, _native.ptr(geom), _native.ptr(lor),
        ent.shape[0], _native.ptr(ent), cell_shape[2], accel, int(weighted), n_bins, _native.ptr(scat),
        _native.ptr(p_arr), p_type, None if counts is None else _native.ptr(counts), per_plane,
        None if r is None else _native.ptr(r), None if nrm is None else _native.ptr(nrm), _native.ptr(m), 0,
        _native.ptr(sums)), 'sss_scale_to_prompts')
    with np.errstate(divide='ignore', invalid='ignore'):
        factors = np.where(sums[:, 1] != 0, sums[:, 0] / np.where(sums[:, 1] != 0, sums[:, 1], 1.0), 0.0)
    factors = factors.astype(np.float32)
    return (factors, sums) if return_sums else factors


def ScaleScattersToPrompts(SinogramIndex, SinogramCoordinates, SavePath, DesiredDimensions, FittingSize,
                           Geometry, ExtendedGeometry, AccelerationFactor, SpanFlag, LORCounts, *,
                           context=None):
    """float32 (TOF bins,) scale factors; also saved as ``ScaleFactors.npy``."""
    prompts = _load_npz(os.path.join(SavePath, 'PromptsSinogramMashed.npz' if SpanFlag
                                     else 'PromptsSinogram.npz'))
    n_bins = prompts.shape[-1]
    norm_path = os.path.join(SavePath, 'NormSinogram.npz')
    rand_path = os.path.join(SavePath, 'RandomsSinogram.npz')
    norm = _load_npz(norm_path) if os.path.exists(norm_path) else None
    randoms = _load_npz(rand_path) if os.path.exists(rand_path) else None
    mask = np.load(os.path.join(SavePath, 'AttenuationMask.npy'))
    scatters = np.stack([_load_npz(os.path.join(SavePath, f'SSS_mashed_bin{b}.npz')).astype(np.float32)
                         for b in range(n_bins)])

    print("Scaling for all bins")
    begin = time.time()
    factors = scale_to_prompts(prompts, scatters, randoms, norm, mask, LORCounts, SinogramIndex,
                               SinogramCoordinates, DesiredDimensions, FittingSize, Geometry, ExtendedGeometry,
                               AccelerationFactor, SpanFlag, context=context)
    print("Time for Scaling: {:.2f} seconds".format(time.time() - begin))
    np.save(os.path.join(SavePath, 'ScaleFactors.npy'), factors)
    return factors


def scale_to_prompts_device(DesiredDimensions, FittingSize, Geometry, SinogramCoordinates, entries,
                            n_value_planes, AccelerationFactor, length_weighted, n_bins, scatter_ptr, prompts_ptr,
                            prompts_u16, mask_ptr, sums_ptr, *, counts_ptr=None, counts_per_plane=False,
                            randoms_ptr=None, norm_ptr=None, geometry_buffers=None, context=None):
    """``sss_scale_to_prompts`` on device-resident sinograms (addresses on the
    context's device, layouts as there; geometry_buffers from
    ``backprojection.fit_geometry_buffers``).  Queued on the context's
    stream; the (n_bins, 2) float64 sums land at sums_ptr."""
    from .backprojection import fit_geometry_buffers
    ctx = context or _native.default_context()
    lib = _native.load_library()
    geom = np.asarray(Geometry)
    n_rings, n_det = geom.shape[0], geom.shape[1]
    bufs = geometry_buffers or fit_geometry_buffers(Geometry, SinogramCoordinates, entries, context=ctx)
    grid = np.ascontiguousarray(np.asarray(DesiredDimensions, dtype=np.int32).reshape(3))
    bounds = np.ascontiguousarray(np.asarray(FittingSize, dtype=np.float64).reshape(6))
    vp = _native.ctypes.c_void_p
    n_entries = int(np.asarray(entries).reshape(-1, 3).shape[0])
    accel = 1 if AccelerationFactor is None else int(AccelerationFactor)
    try:
        _native.check(lib.sss_scale_to_prompts(
            ctx.handle, _native.ptr(grid), _native.ptr(bounds), n_rings, n_det, bufs[0].ptr, bufs[1].ptr, n_entries,
            bufs[2].ptr, int(n_value_planes), accel, int(bool(length_weighted)), int(n_bins), vp(scatter_ptr),
            vp(prompts_ptr), _native.SSS_PROMPTS_U16 if prompts_u16 else _native.SSS_PROMPTS_F32,
            None if counts_ptr is None else vp(counts_ptr), int(bool(counts_per_plane)),
            None if randoms_ptr is None else vp(randoms_ptr), None if norm_ptr is None else vp(norm_ptr),
            vp(mask_ptr), _native.SSS_FIT_DEVICE, vp(sums_ptr)), 'sss_scale_to_prompts')
    finally:
        if geometry_buffers is None:
            ctx.synchronize()
            for b in bufs:
                b.free()
