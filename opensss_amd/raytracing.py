"""GPU voxel tracer with the reference's outputs.

``RayTracing3DTOF`` is the drop-in for ``Python/openSSS/RayTracing3DTOF.py:29-180``:
per-voxel intersection lengths (f64), 1-based linear voxel indices (int32,
X + (Y-1)*Ny + (Z-1)*Ny*Nx) and TOF sample positions (f64,
s_k = s_{k-1} + L_k / 2).  ``RayTracing3DTOFBatch`` traces many segments in one
launch and is what callers should use.
"""
from __future__ import annotations

import numpy as np

from . import _native


def _grid(GridSize, GridBounds):
    gs = np.ascontiguousarray(np.asarray(GridSize).astype(np.int32).reshape(3))
    gb = np.ascontiguousarray(np.asarray(GridBounds, dtype=np.float64).reshape(6))
    return gs, gb


def RayTracing3DTOFBatch(GridSize, GridBounds, Lines, context=None):
    """Trace an (n, 6) array of segments.

    Returns ``(counts int32 (n,), lengths f64 (n, cap), indices int32 (n, cap),
    samples f64 (n, cap))`` with cap = X + Y + Z; row i is valid up to counts[i].
    """
    gs, gb = _grid(GridSize, GridBounds)
    segs = np.ascontiguousarray(np.asarray(Lines, dtype=np.float64).reshape(-1, 6))
    n = segs.shape[0]
    cap = int(gs.sum())
    counts = np.zeros(n, dtype=np.int32)
    L = np.zeros((n, cap))
    idx = np.zeros((n, cap), dtype=np.int32)
    s = np.zeros((n, cap))
    ctx = context or _native.default_context()
    lib = _native.load_library()
    _native.check(lib.sss_trace_segments(ctx.handle, _native.ptr(gs), _native.ptr(gb),
                                          _native.ptr(segs), n, _native.ptr(counts),
                                          _native.ptr(L), _native.ptr(idx), _native.ptr(s)),
                  'sss_trace_segments')
    return counts, L, idx, s


def RayTracing3DTOF(GridSize, GridBounds, LineCoordinates, context=None):
    """Same signature and return as the reference: (lengths, indices, samples)."""
    counts, L, idx, s = RayTracing3DTOFBatch(GridSize, GridBounds,
                                             np.asarray(LineCoordinates).reshape(1, 6), context)
    n = int(counts[0])
    return L[0, :n].copy(), idx[0, :n].copy(), s[0, :n].copy()


def LineIntegrals(GridSize, GridBounds, AttenuationMap_flat, ActivityMap_flat, Lines, context=None):
    """Per segment: voxel count, sum(mu L) and sum(lam L) (float64, sequential)."""
    gs, gb = _grid(GridSize, GridBounds)
    mu = np.ascontiguousarray(AttenuationMap_flat, dtype=np.float64)
    lam = None if ActivityMap_flat is None else np.ascontiguousarray(ActivityMap_flat, dtype=np.float64)
    segs = np.ascontiguousarray(np.asarray(Lines, dtype=np.float64).reshape(-1, 6))
    n = segs.shape[0]
    counts = np.zeros(n, dtype=np.int32)
    m = np.zeros(n)
    l = np.zeros(n)
    ctx = context or _native.default_context()
    lib = _native.load_library()
    _native.check(lib.sss_line_integrals(ctx.handle, _native.ptr(gs), _native.ptr(gb), _native.ptr(mu),
                                         None if lam is None else _native.ptr(lam), _native.ptr(segs),
                                         n, _native.ptr(counts), _native.ptr(m), _native.ptr(l)),
                  'sss_line_integrals')
    return counts, m, l
