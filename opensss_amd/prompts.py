"""Synthetic prompts for the tail fit: forward projection and Poisson counts.

The reference reads measured prompts (CASToR list mode, ``Timber/Datafile.py``);
a synthetic study needs them generated.  ``ForwardProject`` traces every LOR
of the sinogram, as the tail mask does (``TailsMask.py:139-145``), and gives
the trues sum(lam L) exp(-sum(mu L)); ``PoissonSample`` draws counts from a
mean sinogram; ``synthetic_prompts`` assembles trues + scatter + uniform
randoms at a total count and returns the scatter's true scale, which the tail
fit should recover.  The traces and draws run in ``sss_forward_project`` and
``sss_poisson``.

Plane convention (as the tail mask and the interpolated scatter): without
span, plane SinogramIndex[R1, R2] holds the LORs from detector d1 of ring R1
to d2 of ring R2, d1 <= d2; with span, plane p holds those between rings
int(2 m) of the extended geometry, m the plane's averaged rings.
"""
from __future__ import annotations

import numpy as np

from . import _native
from .tails_mask import nospan_plane_rings, span_plane_rings


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def ForwardProject(ActivityMap, AttenuationMap, ImageSize, Geometry, SinogramCoordinates, SinogramIndex=None,
                   MashedSinogramIndices=None, AccelerationFactor=1, *, context=None):
    """float32 (N+1, N//2, planes) trues; AttenuationMap None: no attenuation."""
    lam = np.asarray(ActivityMap)
    if lam.ndim != 3:
        raise ValueError('ActivityMap must be 3-D (X, Y, Z)')
    geom = _f64(Geometry)
    n_rings, n_det = geom.shape[0], geom.shape[1]
    lor = np.ascontiguousarray(np.asarray(SinogramCoordinates, dtype=np.int16))
    if lor.shape != (n_det, n_det, 2):
        raise ValueError('SinogramCoordinates must be (N, N, 2)')
    if MashedSinogramIndices is not None:
        pr = span_plane_rings(MashedSinogramIndices)
    else:
        pr = nospan_plane_rings(SinogramIndex, n_rings)
    pr = np.ascontiguousarray(pr, dtype=np.int32)
    grid = np.ascontiguousarray(np.asarray(lam.shape, dtype=np.int32))
    bounds = _f64(ImageSize)
    lam_f = _f64(lam.flatten(order='F'))
    mu_f = None if AttenuationMap is None else _f64(np.asarray(AttenuationMap).flatten(order='F'))
    out = np.zeros((n_det + 1, n_det // 2, pr.shape[0]), dtype=np.float32)
    ctx = context or _native.default_context()
    lib = _native.load_library()
    accel = 1 if AccelerationFactor is None else int(AccelerationFactor)
    _native.check(lib.sss_forward_project(ctx.handle, _native.ptr(grid), _native.ptr(bounds), _native.ptr(lam_f),
                                          None if mu_f is None else _native.ptr(mu_f), n_rings, n_det,
                                          _native.ptr(geom), _native.ptr(lor), pr.shape[0], _native.ptr(pr), accel,
                                          _native.ptr(out)), 'sss_forward_project')
    return out


def PoissonSample(mean, scale=1.0, seed=0, *, context=None):
    """float32 Poisson counts of mean * scale, same shape; the same seed gives
    the same counts."""
    m = np.ascontiguousarray(np.asarray(mean, dtype=np.float32))
    out = np.empty_like(m)
    ctx = context or _native.default_context()
    lib = _native.load_library()
    _native.check(lib.sss_poisson(ctx.handle, _native.ptr(m), m.size, float(scale), int(seed) & (2 ** 64 - 1),
                                  _native.ptr(out)), 'sss_poisson')
    return out


def synthetic_prompts(trues, scatter, scatter_fraction=0.3, randoms_fraction=0.1, total_counts=2e7, seed=0,
                      noise=True, *, context=None):
    """Prompts = Poisson(c (trues + s scatter + r)), with s such that the
    scatter counts are scatter_fraction of the trues counts, r uniform with
    randoms_fraction of the trues counts, and c scaling the total mean to
    total_counts.  Returns (prompts f32, randoms f32 (the mean c r per cell),
    k) where k = c s is the scale of `scatter` inside the prompts."""
    T = np.asarray(trues, dtype=np.float64)
    S = np.asarray(scatter, dtype=np.float64)
    if T.shape != S.shape:
        raise ValueError('trues and scatter differ in shape')
    st, ss = T.sum(), S.sum()
    if not (st > 0 and ss > 0):
        raise ValueError('trues and scatter must have positive totals')
    s = scatter_fraction * st / ss
    r = randoms_fraction * st / T.size
    mean = T + s * S + r
    c = float(total_counts) / mean.sum()
    randoms = np.full(T.shape, c * r, dtype=np.float32)
    if noise:
        prompts = PoissonSample(mean.astype(np.float32), scale=c, seed=seed, context=context)
    else:
        prompts = (c * mean).astype(np.float32)
    return prompts, randoms, c * s
