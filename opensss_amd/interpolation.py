"""Sampled scatter sinogram -> full sinogram, on the GPU.

Drop-in for ``Python/openSSS/Interpolation.py:32-399``.  The reference runs,
per TOF bin and sampled ring pair, ``griddata(nearest)`` over the sampled
(radial, angular) cells (:95-131), clamps negatives (:163), then a
``RegularGridInterpolator(nearest)`` over the sampled rings (:256, :378).
Both nearest-neighbour searches depend only on the geometry, so they are
evaluated once, on the host with the same scipy calls (so ties resolve the
same way), into two index maps; the device then gathers every output cell of
every bin in one HBM-bound pass (``sss_interpolate``).
"""
from __future__ import annotations

import os
import time

import numpy as np
from scipy.interpolate import RegularGridInterpolator, griddata

from . import _native
from .tables import detectors_used, rings_used


def radial_angular_map(SinogramCoordinates, dets):
    """(N+1, N//2) int32: compact cell of the nearest sampled detector pair.

    Points are every (Detectors[i], Detectors[j]) pair in the reference's
    order (:122-127); ``griddata(method='nearest')`` with the point ids as
    values gives the same neighbour the reference's per-plane call picks.
    Compact cell of the pair (i, j) is that of (min, max): the sampled
    sinogram holds one value per unordered pair.
    """
    lor = np.asarray(SinogramCoordinates)
    N = lor.shape[0]
    ndu = len(dets)
    d1 = np.repeat(np.asarray(dets), ndu)
    d2 = np.tile(np.asarray(dets), ndu)
    pts = np.stack([lor[d1, d2, 1], lor[d1, d2, 0]], axis=1).astype(np.float64)
    gx, gy = np.mgrid[0:N + 1, 0:N // 2]
    nearest = griddata(pts, np.arange(pts.shape[0]), (gx, gy), method='nearest').astype(np.int64)
    i = nearest // ndu
    j = nearest % ndu
    lo = np.minimum(i, j)
    hi = np.maximum(i, j)
    cell = lo * ndu - lo * (lo - 1) // 2 + (hi - lo)
    return np.ascontiguousarray(cell.astype(np.int32))


def _nearest_ring_index(rings, x):
    """Index into ``rings`` that ``RegularGridInterpolator(nearest)`` picks (-1 outside)."""
    r = np.asarray(rings, dtype=np.float64)
    rgi = RegularGridInterpolator((r,), np.arange(len(r), dtype=np.float64), method='nearest',
                                  bounds_error=False, fill_value=-1.0)
    return rgi(np.asarray(x, dtype=np.float64).reshape(-1, 1)).astype(np.int64)


def plane_map_nospan(SinogramIndex, rings):
    """Output plane s = SinogramIndex[R1, R2] reads sampled plane
    nearest(R2) + nearest(R1) * NRU (the transpose and reorder of :251-310)."""
    sino = np.asarray(SinogramIndex, dtype=np.int64)
    NR = sino.shape[0]
    nru = len(rings)
    near = _nearest_ring_index(rings, np.arange(NR))
    out = np.full(NR * NR, -1, dtype=np.int32)
    for r1 in range(NR):
        for r2 in range(NR):
            a, b = near[r1], near[r2]
            out[sino[r1, r2]] = -1 if (a < 0 or b < 0) else b + a * nru
    return out


def plane_map_span(MashedSinogramIndices, rings):
    """Span plane s queries the ring grid at (mean ring2, mean ring1) (:384):
    sampled plane nearest(mean ring2) + nearest(mean ring1) * NRU."""
    msi = np.asarray(MashedSinogramIndices)
    nru = len(rings)
    ra = _nearest_ring_index(rings, msi[:, 0])
    rb = _nearest_ring_index(rings, msi[:, 1])
    return np.where((ra >= 0) & (rb >= 0), ra + rb * nru, -1).astype(np.int32)


def dense_to_compact(Scatters, SinogramCoordinates, dets):
    """(B, N+1, N//2, NRU^2) sampled sinogram -> compact (B, NRU^2, NDU(NDU+1)/2)."""
    lor = np.asarray(SinogramCoordinates)
    i1, i2 = np.triu_indices(len(dets))
    ang = lor[dets[i1], dets[i2], 0]
    rad = lor[dets[i1], dets[i2], 1]
    sc = np.asarray(Scatters, dtype=np.float32)
    return np.ascontiguousarray(np.transpose(sc[:, rad, ang, :], (0, 2, 1)))


class Interpolator:
    """Geometry-only maps on the device; interpolates any sampled estimate."""

    def __init__(self, NrRings, NrRingsUsed, NrDetectors, NrDetectorsUsed, SinogramCoordinates,
                 SinogramIndex, SpanFlag, context=None):
        self.ctx = context or _native.default_context()
        self._lib = _native.load_library()
        self.rings = rings_used(NrRings, NrRingsUsed)
        self.dets = detectors_used(NrDetectors, NrDetectorsUsed)
        self.n_det = NrDetectors
        self.nru = NrRingsUsed
        self.lor = np.asarray(SinogramCoordinates)
        cell_map = radial_angular_map(SinogramCoordinates, self.dets)
        plane_map = (plane_map_span(SinogramIndex, self.rings) if SpanFlag
                     else plane_map_nospan(SinogramIndex, self.rings))
        self.n_cells_out = cell_map.size
        self.n_out_planes = plane_map.size
        self.shape_out = (NrDetectors + 1, NrDetectors // 2, self.n_out_planes)
        self.d_cell = _native.DeviceBuffer.from_array(self.ctx, cell_map)
        self.d_plane = _native.DeviceBuffer.from_array(self.ctx, plane_map)

    @property
    def n_cells_sampled(self):
        return len(self.dets) * (len(self.dets) + 1) // 2

    def run_device(self, sampled_dev_ptr, n_bins, out_dev=None, sum_dev=None):
        _native.check(self._lib.sss_interpolate(
            self.ctx.handle, _native.ctypes.c_void_p(sampled_dev_ptr), int(n_bins), self.nru * self.nru,
            self.n_cells_sampled, self.d_cell.ptr, self.n_cells_out, self.d_plane.ptr,
            self.n_out_planes, None if out_dev is None else _native.ctypes.c_void_p(out_dev),
            None if sum_dev is None else _native.ctypes.c_void_p(sum_dev)), 'sss_interpolate')

    def run(self, compact: np.ndarray, per_bin: bool = True):
        """compact (B, NRU^2, cells) float32 -> (per-bin (B, N+1, N/2, P) or None, sum (N+1, N/2, P))."""
        c = np.ascontiguousarray(compact, dtype=np.float32)
        B = c.shape[0]
        src = _native.DeviceBuffer.from_array(self.ctx, c)
        n_out = self.n_cells_out * self.n_out_planes
        out_d = _native.DeviceBuffer(self.ctx, 4 * n_out * B) if per_bin else None
        sum_d = _native.DeviceBuffer(self.ctx, 4 * n_out)
        try:
            self.run_device(src.ptr.value, B, None if out_d is None else out_d.ptr.value, sum_d.ptr.value)
            total = sum_d.download(np.empty(self.shape_out, dtype=np.float32))
            bins = out_d.download(np.empty((B,) + self.shape_out, dtype=np.float32)) if per_bin else None
        finally:
            src.free()
            sum_d.free()
            if out_d is not None:
                out_d.free()
        return bins, total


def InterpolateScatters(Scatters, NrRings, NrRingsUsed, NrDetectors, NrDetectorsUsed,
                        SinogramCoordinates, SinogramIndex, SavePath, SpanFlag, *, SaveBins=True,
                        context=None):
    """Reference signature (:32-93): returns the float32 sum over TOF bins of the
    interpolated sinograms and writes ``SSS_mashed_bin{b}.npz`` per bin."""
    start = time.time()
    print("Start Interpolating with data reduction" if SpanFlag else "Start Interpolating")
    ip = Interpolator(NrRings, NrRingsUsed, NrDetectors, NrDetectorsUsed, SinogramCoordinates,
                      SinogramIndex, SpanFlag, context=context)
    compact = dense_to_compact(Scatters, SinogramCoordinates, ip.dets)
    save = SavePath is not None and SaveBins
    bins, total = ip.run(compact, per_bin=save)
    if save:
        for b in range(bins.shape[0]):
            np.savez_compressed(f'{SavePath}/SSS_mashed_bin{b}', bins[b])
    elapsed = time.time() - start
    print("Time for Interpolation: {:.2f} seconds".format(elapsed))
    if SavePath is not None:
        with open(os.path.join(SavePath, 'Simulation_time.txt'), 'w') as fh:
            fh.write("Time for Interpolation: {:.2f} seconds\n".format(elapsed))
    return total
