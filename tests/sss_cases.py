"""Shared helpers for the scatter-estimate parity tests: run a workload through
the GPU subject and through the oracle, and the error measures the parity bar
is stated in (max relative error on cells above 1% of the peak)."""
import os

import numpy as np

from opensss_amd.tables import detectors_used, rings_used

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WATER = np.load(os.path.join(ROOT, 'opensss_amd', 'data', 'attenuation_water.npy'))
ORACLE_THREADS = int(os.environ.get('SSS_ORACLE_THREADS', '16'))


def rel_err_above(got, want, frac=0.01):
    """(max rel err on cells >= frac * peak, max abs err / peak over all cells)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    peak = np.abs(want).max()
    if peak == 0:
        return float(np.abs(got).max()), float(np.abs(got).max())
    big = np.abs(want) >= frac * peak
    rel = np.abs(got[big] - want[big]) / np.abs(want[big])
    return float(rel.max()), float(np.abs(got - want).max() / peak)


def gpu_subject(w, precision='fp32', points=None, ptr=None, n_chunks=0):
    from opensss_amd.sss import ScatterSubject
    pts = w.scatter_points if points is None else points
    return ScatterSubject(w.attenuation, w.image_size, w.geometry, w.normals, w.det_size, WATER,
                          w.energy_resolution, w.energy_threshold, w.n_rings_used, w.n_det_used,
                          w.tof_resolution, w.tof_range, w.n_bins, scatter_points=pts,
                          precision=precision, n_chunks=n_chunks)


def gpu_dense(w, precision='fp32', points=None, lam=None):
    s = gpu_subject(w, precision, points)
    try:
        s.run(w.activity if lam is None else lam)
        return s.output_dense(w.lor_coordinates)
    finally:
        s.close()


def gpu_compact(w, precision='fp32', points=None, lam=None, n_chunks=0):
    s = gpu_subject(w, precision, points, n_chunks=n_chunks)
    try:
        s.run(w.activity if lam is None else lam)
        return s.output_compact()
    finally:
        s.close()


def oracle_dense(oracle, w, points=None, ptr=None, lam=None, n_bins=None, tof_res=None):
    from opensss_amd.tables import group_by_slice
    pts = w.scatter_points if points is None else points
    pts, ptr, _ = group_by_slice(pts)
    return oracle.sss(w.activity if lam is None else lam, w.attenuation, w.image_size, w.geometry,
                      w.lor_coordinates, w.normals, w.det_size, WATER, w.energy_resolution,
                      w.energy_threshold, rings_used(w.geometry.shape[0], w.n_rings_used),
                      detectors_used(w.geometry.shape[1], w.n_det_used), w.tof_range,
                      w.tof_resolution if tof_res is None else tof_res,
                      w.n_bins if n_bins is None else n_bins, pts, ptr, n_threads=ORACLE_THREADS)


def oracle_compact(oracle, w, points=None, lam=None):
    """Oracle result rearranged into the GPU compact layout (B, NRU^2, cells)."""
    from opensss_amd.tables import group_by_slice
    pts = w.scatter_points if points is None else points
    pts, ptr, _ = group_by_slice(pts)
    dets = detectors_used(w.geometry.shape[1], w.n_det_used)
    out, cell_of = oracle.sss(w.activity if lam is None else lam, w.attenuation, w.image_size,
                              w.geometry, w.lor_coordinates, w.normals, w.det_size, WATER,
                              w.energy_resolution, w.energy_threshold,
                              rings_used(w.geometry.shape[0], w.n_rings_used), dets, w.tof_range,
                              w.tof_resolution, w.n_bins, pts, ptr, n_threads=ORACLE_THREADS,
                              dense=False)
    N = w.geometry.shape[1]
    i1, i2 = np.triu_indices(len(dets))
    ang = w.lor_coordinates[dets[i1], dets[i2], 0].astype(np.int64)
    rad = w.lor_coordinates[dets[i1], dets[i2], 1].astype(np.int64)
    return out[:, :, cell_of[rad * (N // 2) + ang]]
