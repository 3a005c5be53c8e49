"""TOF single scatter simulation on the GPU.

``SingleScatterSimulationTOF`` is the drop-in for
``Python/openSSS/SingleScatterSimulationTOF.py:36-208`` (same arguments, units,
return array and ``Scatters_SSS.npz`` side effect).  ``ScatterSubject`` is the
resident form used by batching and serving: the mu-map, scatter points and
half-path table stay in HBM and only the activity image changes per call.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import _native
from .tables import (detectors_used, energy_tables, group_by_slice, rings_used,
                     scatter_point_coordinates, strided_scatter_points, tof_kernel)

PRECISIONS = {'fp32': _native.SSS_PRECISION_FP32, 'fp64': _native.SSS_PRECISION_FP64}


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


# Subjects of one scanner and energy window share their set-up tables; these
# are pure functions of their arguments, so they are computed once per key.
_ENERGY_CACHE: dict = {}
_TOF_CACHE: dict = {}
_CACHE_MAX = 32


def _energy_tables(table, resolution, threshold):
    t = _f64(table)
    key = (t.shape, t.tobytes(), float(resolution), float(threshold))
    hit = _ENERGY_CACHE.get(key)
    if hit is None:
        ratio, eff, ref = energy_tables(t, resolution, threshold)
        ratio, eff = _f64(ratio), _f64(eff)
        ratio.setflags(write=False)
        eff.setflags(write=False)
        hit = (ratio, eff, ref)
        if len(_ENERGY_CACHE) >= _CACHE_MAX:
            _ENERGY_CACHE.clear()
        _ENERGY_CACHE[key] = hit
    return hit


_SAMPLING_CACHE: dict = {}


def _sampling(n_rings, nru, n_det, ndu):
    key = (int(n_rings), int(nru), int(n_det), int(ndu))
    hit = _SAMPLING_CACHE.get(key)
    if hit is None:
        rings = _i32(rings_used(key[0], key[1]))
        dets = _i32(detectors_used(key[2], key[3]))
        rings.setflags(write=False)
        dets.setflags(write=False)
        hit = (rings, dets)
        if len(_SAMPLING_CACHE) >= _CACHE_MAX:
            _SAMPLING_CACHE.clear()
        _SAMPLING_CACHE[key] = hit
    return hit


def _tof_kernel(n_bins, tof_range, resolution, window_tol):
    key = (int(n_bins), float(tof_range), float(resolution), float(window_tol))
    hit = _TOF_CACHE.get(key)
    if hit is None:
        hit = tof_kernel(*key)
        if len(_TOF_CACHE) >= _CACHE_MAX:
            _TOF_CACHE.clear()
        _TOF_CACHE[key] = hit
    return hit


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


class ScatterSubject:
    """One resident scatter-estimation instance (``sss_subject``)."""

    def __init__(self, AttenuationMap, ImageSize, Geometry, NormalVectors, DetectorSize,
                 AttenuationTable, EnergyResolution, EnergyThreshold, NrRingsUsed,
                 NrDetectorsUsed, TOFResolution, TOFRange, NrBins, *, scatter_points=None,
                 SampleStep=(1, 1, 1), precision='fp32', context=None, n_chunks=0,
                 upload_order='C', tile_rings=0):
        mu = np.asarray(AttenuationMap)
        if mu.ndim != 3:
            raise ValueError('AttenuationMap must be 3-D (X, Y, Z)')
        self.grid = tuple(int(v) for v in mu.shape)
        if self.grid[0] != self.grid[1]:
            raise ValueError('AttenuationMap must have X == Y (the reference voxel index uses '
                             'the Y size as the y stride)')
        geom = _f64(Geometry)
        if geom.ndim != 3 or geom.shape[2] != 3:
            raise ValueError('Geometry must be (rings, detectors, 3)')
        if precision not in PRECISIONS:
            raise ValueError(f'precision must be one of {sorted(PRECISIONS)}')
        self.precision = precision
        self.n_rings, self.n_det = geom.shape[0], geom.shape[1]
        self.nru, self.ndu = int(NrRingsUsed), int(NrDetectorsUsed)
        self.n_bins = int(NrBins)
        self.image_size = _f64(ImageSize)
        self.ctx = context or _native.default_context()
        self._lib = _native.load_library()

        ratio, eff, _ = _energy_tables(AttenuationTable, EnergyResolution, EnergyThreshold)
        # the FP64 validation path keeps every bin above 1e-16 of the nearest one
        self.tof = _tof_kernel(self.n_bins, float(TOFRange), float(TOFResolution),
                               1e-16 if precision == 'fp64' else 1e-9)
        self.rings, self.dets = _sampling(self.n_rings, self.nru, self.n_det, self.ndu)
        if scatter_points is None:
            scatter_points = strided_scatter_points(mu, SampleStep)
        pts, ptr, zs = group_by_slice(scatter_points)
        self.scatter_points, self.slice_ptr, self.slice_z = pts, _i32(ptr), zs
        self.sp_xyz = _f64(scatter_point_coordinates(pts, self.grid, self.image_size))
        # upload_order 'C': volumes go over in C order (no host transpose) and
        # the subject lays them out x fastest on the device; 'F': flattened
        # x fastest on the host (the reference's flatten(order='F'))
        if upload_order not in ('C', 'F'):
            raise ValueError("upload_order must be 'C' or 'F'")
        self._upload_order = upload_order
        self._keep = dict(mu=self._volume(mu), geom=geom, normals=_f64(NormalVectors),
                          ratio=ratio, eff=eff)

        d = _native.SubjectDesc()
        d.grid[:] = list(self.grid)
        d.bounds[:] = [float(v) for v in self.image_size]
        d.mu = self._keep['mu'].ctypes.data
        d.n_rings, d.n_det = self.n_rings, self.n_det
        d.geom, d.normals = geom.ctypes.data, self._keep['normals'].ctypes.data
        d.det_size[:] = [float(DetectorSize[0]), float(DetectorSize[1])]
        d.ratio, d.eff = self._keep['ratio'].ctypes.data, self._keep['eff'].ctypes.data
        d.n_energy = ratio.shape[0]
        d.eff_ref_row = 511 * 2 - 1
        d.nru, d.ndu = self.nru, self.ndu
        d.rings, d.dets = self.rings.ctypes.data, self.dets.ctypes.data
        d.n_sp = self.sp_xyz.shape[0]
        d.sp_xyz = self.sp_xyz.ctypes.data if d.n_sp else None
        d.n_slices = len(self.slice_ptr) - 1
        d.slice_ptr = self.slice_ptr.ctypes.data
        d.n_bins = self.n_bins
        d.tof_bin_width = self.tof.bin_width
        d.tof_kappa, d.tof_beta = self.tof.kappa, self.tof.beta
        d.tof_half_window = self.tof.half_window
        d.tof_t_lo, d.tof_t_hi = self.tof.t_lo, self.tof.t_hi
        d.tof_trivial = int(self.tof.trivial)
        d.precision = PRECISIONS[precision]
        d.n_chunks = int(n_chunks)
        d.tile_rings = int(tile_rings)
        d.volume_order = (_native.SSS_VOLUME_Z_FASTEST if upload_order == 'C'
                          else _native.SSS_VOLUME_X_FASTEST)
        h = _native.ctypes.c_void_p()
        _native.check(self._lib.sss_subject_create(self.ctx.handle, _native.ctypes.byref(d),
                                                   _native.ctypes.byref(h)), 'sss_subject_create')
        self.handle = h
        self._traced = False

    def _volume(self, v):
        if self._upload_order == 'C':
            return _f64(v)
        return _f64(np.asarray(v).flatten(order='F'))

    # -- lifecycle -------------------------------------------------------
    def close(self):
        if getattr(self, 'handle', None):
            # fails (and keeps the subject) while a live batch holds it
            _native.check(self._lib.sss_subject_destroy(self.handle), 'sss_subject_destroy')
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- work --------------------------------------------------------------
    def trace(self):
        _native.check(self._lib.sss_subject_trace(self.handle), 'sss_subject_trace')
        self._traced = True

    def trace_begin(self):
        """Queue the walk counts (asynchronous); finish with trace_end."""
        _native.check(self._lib.sss_subject_trace_begin(self.handle), 'sss_subject_trace_begin')

    def trace_end(self):
        """Wait for this subject's counts, size its tables, queue the walks."""
        _native.check(self._lib.sss_subject_trace_end(self.handle), 'sss_subject_trace_end')
        self._traced = True

    def set_tiling(self, tile_rings: int = 0, n_chunks: int = 0):
        """Re-tile the scatter launch (rings per tile block, scatter-point
        chunks; 0 = defaults) without redoing the walks."""
        _native.check(self._lib.sss_subject_set_tiling(self.handle, int(tile_rings), int(n_chunks)),
                      'sss_subject_set_tiling')

    def set_activity(self, ActivityMap):
        lam = np.asarray(ActivityMap)
        if tuple(lam.shape) != self.grid:
            raise ValueError(f'ActivityMap shape {lam.shape} != {self.grid}')
        flat = self._volume(lam)
        _native.check(self._lib.sss_subject_set_activity(self.handle, _native.ptr(flat)),
                      'sss_subject_set_activity')

    def set_activity_device(self, device_ptr: int):
        """Activity from device memory: X*Y*Z float64 in C order of (X, Y, Z)
        (a contiguous torch tensor of the grid's shape), or x fastest with
        upload_order 'F'."""
        _native.check(self._lib.sss_subject_set_activity_device(self.handle,
                                                                _native.ctypes.c_void_p(device_ptr)),
                      'sss_subject_set_activity_device')

    def scatter(self):
        if not self._traced:
            self.trace()
        _native.check(self._lib.sss_subject_scatter(self.handle), 'sss_subject_scatter')

    def run(self, ActivityMap=None):
        """Trace, (optionally) set activity, scatter; the plain one-instance path."""
        if ActivityMap is not None:
            self.set_activity(ActivityMap)
        _native.check(self._lib.sss_subject_run(self.handle), 'sss_subject_run')
        self._traced = True

    def synchronize(self):
        self.ctx.synchronize()

    # -- results -----------------------------------------------------------
    @property
    def info(self) -> dict:
        i = _native.SubjectInfo()
        _native.check(self._lib.sss_subject_info_get(self.handle, _native.ctypes.byref(i)),
                      'sss_subject_info_get')
        return {k: getattr(i, k) for k, _ in i._fields_}

    def work_stats(self):
        """(pairs passing the filters, voxel samples folded) of the current activity."""
        p = _native.ctypes.c_int64()
        v = _native.ctypes.c_int64()
        _native.check(self._lib.sss_subject_work_stats(self.handle, _native.ctypes.byref(p),
                                                       _native.ctypes.byref(v)),
                      'sss_subject_work_stats')
        return int(p.value), int(v.value)

    def scatter_kernel_ms(self) -> float:
        """HIP-event duration of the last scatter kernel on the context stream."""
        ms = _native.ctypes.c_float()
        _native.check(self._lib.sss_subject_scatter_kernel_ms(self.handle, _native.ctypes.byref(ms)),
                      'sss_subject_scatter_kernel_ms')
        return float(ms.value)

    @property
    def n_cells_plane(self):
        return self.ndu * (self.ndu + 1) // 2

    def output_compact(self) -> np.ndarray:
        """(B, NRU**2, NDU(NDU+1)/2) in the subject's precision; plane = r2 + r1*NRU."""
        dt = np.float64 if self.precision == 'fp64' else np.float32
        out = np.empty((self.n_bins, self.nru * self.nru, self.n_cells_plane), dtype=dt)
        _native.check(self._lib.sss_subject_output_compact(self.handle, _native.ptr(out)),
                      'sss_subject_output_compact')
        return out

    def output_compact_async(self, host_addr: int):
        """Queue the compact download into pinned host memory at host_addr
        (Context.host_alloc, nbytes = output_nbytes); wait() waits for it."""
        _native.check(self._lib.sss_subject_output_compact_async(self.handle, _native.ctypes.c_void_p(host_addr)),
                      'sss_subject_output_compact_async')

    def wait(self):
        _native.check(self._lib.sss_subject_wait(self.handle), 'sss_subject_wait')

    def done(self) -> bool:
        """True once the last queued download has finished."""
        v = _native.ctypes.c_int32()
        _native.check(self._lib.sss_subject_query(self.handle, _native.ctypes.byref(v)), 'sss_subject_query')
        return bool(v.value)

    @property
    def output_shape(self):
        return (self.n_bins, self.nru * self.nru, self.n_cells_plane)

    @property
    def output_dtype(self):
        return np.float64 if self.precision == 'fp64' else np.float32

    @property
    def output_nbytes(self) -> int:
        return int(np.prod(self.output_shape)) * np.dtype(self.output_dtype).itemsize

    def output_dense(self, SinogramCoordinates) -> np.ndarray:
        """float32 (B, N+1, N//2, NRU**2) as the reference returns it."""
        lor = np.ascontiguousarray(np.asarray(SinogramCoordinates, dtype=np.int16))
        if lor.shape != (self.n_det, self.n_det, 2):
            raise ValueError('SinogramCoordinates must be (N, N, 2)')
        out = np.empty((self.n_bins, self.n_det + 1, self.n_det // 2, self.nru * self.nru),
                       dtype=np.float32)
        _native.check(self._lib.sss_subject_output_dense(self.handle, _native.ptr(lor),
                                                         _native.ptr(out)),
                      'sss_subject_output_dense')
        return out


def compact_cell_index(ndu: int):
    """(i1, i2) of every compact cell, i1 <= i2, in storage order."""
    i1, i2 = np.triu_indices(ndu)
    return i1, i2


def SingleScatterSimulationTOF(ActivityMap, AttenuationMap, ImageSize, Geometry,
                               SinogramCoordinates, NormalVectors, DetectorSize, AttenuationTable,
                               EnergyResolution, EnergyThreshold, NrRingsUsed, NrDetectorsUsed,
                               SampleStep, TOFResolution, TOFRange, NrBins, SavePath, *,
                               ScatterPoints=None, SaveOutput=True, precision='fp32',
                               context=None):
    """3-D TOF single scatter simulation (Watson), reference signature and units.

    Returns float32 (NrBins, N+1, N//2, NrRingsUsed**2).  ``ScatterPoints``
    (n, 3 voxel indices) replaces the strided ``SampleStep`` grid when given;
    points are grouped per z slice for the per-slice normalisation either way.
    """
    subj = ScatterSubject(AttenuationMap, ImageSize, Geometry, NormalVectors, DetectorSize,
                          AttenuationTable, EnergyResolution, EnergyThreshold, NrRingsUsed,
                          NrDetectorsUsed, TOFResolution, TOFRange, NrBins,
                          scatter_points=ScatterPoints, SampleStep=SampleStep,
                          precision=precision, context=context)
    try:
        print("Start Scattering Simulation")
        t0 = time.time()
        subj.run(ActivityMap)
        scatters = subj.output_dense(SinogramCoordinates)
        print("Time for SSS: {:.2f} seconds".format(time.time() - t0))
    finally:
        subj.close()
    if SavePath is not None and SaveOutput:
        np.savez_compressed(os.path.join(SavePath, 'Scatters_SSS.npz'), scatters)
    return scatters
