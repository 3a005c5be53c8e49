"""Host-side lookup tables and index sets for the scatter estimate.

Everything here is set-up work done once per call (or once per subject) on
the host, in float64, with the reference's own formulas, so that the device
receives bit-identical tables:

* ``energy_tables``   -- ``SingleScatterSimulationTOF.py:89-140``
* ``tof_kernel``      -- ``SingleScatterSimulationTOF.py:142-150,260-295``
* ``rings_used`` / ``detectors_used`` -- ``:152-167``
* ``strided_scatter_points`` / ``scatter_point_coordinates`` -- ``:176-182,381-394``
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np
from scipy.special import erf

ENERGY_REFERENCE_KEV = 511
ENERGY_REFERENCE_JOULE = 511.0e3 * 1.6e-19
ELECTRON_MASS = 9.10938356e-31
LIGHT_SPEED_M_S = 299792458.0
LIGHT_SPEED_CM_PS = 299792458.0 * 1e-10
TOF_TABLE_HALF_WIDTH_PS = 4000
TOF_TABLE_COLUMNS = 2 * TOF_TABLE_HALF_WIDTH_PS + 1
KN_GAMMA = ENERGY_REFERENCE_JOULE / (ELECTRON_MASS * LIGHT_SPEED_M_S ** 2)


def energy_tables(attenuation_table: np.ndarray, energy_resolution: float,
                  energy_threshold: float):
    """Water attenuation ratio and detection efficiency per 0.5 keV energy row.

    Returns ``(ratio f64 (E,), efficiency f64 (E,), reference_row int)``:
    ratio = mu_water(E)/mu_water(511 keV) (``:117-120``) and
    efficiency = 0.5 (1 - erf((threshold - E) / (511 res / 2 / sqrt(ln 2)))) (``:139-140``).
    Only a lower energy threshold exists in the reference.
    """
    table = np.asarray(attenuation_table, dtype=np.float64)
    rows = np.where(table[:, 0] == ENERGY_REFERENCE_KEV)[0]
    if rows.size != 1:
        raise ValueError("attenuation table must hold exactly one 511 keV row")
    ref = int(rows[0])
    ratio = table[:, 1] / table[ref, 1]
    eff = 0.5 * (1 - erf((energy_threshold - table[:, 0])
                         / (ENERGY_REFERENCE_KEV * energy_resolution / 2 / np.sqrt(np.log(2)))))
    return np.ascontiguousarray(ratio), np.ascontiguousarray(eff), ref


def tof_table(n_bins: int, bin_width: float, resolution: float) -> np.ndarray:
    """The reference's FP64 TOF table T[b, t + 4000], t in [-4000, 4000] ps.

    Used by the oracle and by the host checks below; the device evaluates the
    same normalised Gaussians analytically (see ``TofKernel``).
    """
    t = np.arange(-TOF_TABLE_HALF_WIDTH_PS, TOF_TABLE_HALF_WIDTH_PS + 1, dtype=np.float64)
    shifts = ((np.arange(1, n_bins + 1) - (n_bins + 1) / 2) * bin_width).reshape(-1, 1)
    g = np.exp(-((t.reshape(1, -1) - shifts) ** 2) / (resolution ** 2 / 4 / np.log(2)))
    s = np.sum(g, axis=0)
    zero = s == 0
    with np.errstate(invalid='ignore', divide='ignore'):
        p = g / s
    p[:, zero] = 0
    return p


@dataclass(frozen=True)
class TofKernel:
    """Parameters of the analytic, windowed TOF kernel evaluated on the device.

    For an integer time offset t the reference weight of bin b is the softmax
    of l_b(t) = -(t - s_b)^2 / w over the bins, w = R^2 / (4 ln 2),
    s_b = (b - (B-1)/2) W.  Relative to the nearest bin j0 (u = t - s_j0) the
    logit difference of bin j0+d is d*kappa*u - d^2*beta with kappa = 2W/w and
    beta = W^2/w, so only bins with |d| <= half_window carry weight above
    ``window_tol`` of the nearest one.  Columns where every Gaussian underflows
    in float64 are zero in the reference; they form the complement of
    [t_lo, t_hi] (checked on the host).
    """
    n_bins: int
    bin_width: float
    w: float
    kappa: float
    beta: float
    half_window: int
    t_lo: int
    t_hi: int
    trivial: bool          # one bin and no zero columns: every weight is exactly 1


def tof_kernel(n_bins: int, tof_range: float, resolution: float,
               window_tol: float = 1e-9) -> TofKernel:
    """Window tolerance: a bin whose weight is below window_tol of the nearest
    bin's is dropped.  Every dropped weight is at most window_tol of its
    sample's total, so a cell of one bin changes by at most
    window_tol * n_bins * (peak / cell) -- 1.5e-6 for 15 bins on cells above
    1% of the peak."""
    bin_width = 2.0 * tof_range / n_bins
    w = resolution ** 2 / 4.0 / math.log(2.0)
    kappa = 2.0 * bin_width / w
    beta = bin_width ** 2 / w
    table = tof_table(n_bins, bin_width, resolution)
    colsum = table.sum(axis=0)
    nz = np.nonzero(colsum != 0)[0]
    if nz.size == 0:
        t_lo, t_hi = 1, 0
    else:
        t_lo = int(nz[0]) - TOF_TABLE_HALF_WIDTH_PS
        t_hi = int(nz[-1]) - TOF_TABLE_HALF_WIDTH_PS
        if nz.size != nz[-1] - nz[0] + 1:
            raise NotImplementedError("TOF table with interior all-zero columns is not supported")
    if n_bins == 1:
        h = 0
    else:
        h = 0
        # weight of bin j0 +- (h+1) relative to bin j0 is at most exp(-h (h+1) beta)
        while h < n_bins - 1 and h * (h + 1) * beta < -math.log(window_tol):
            h += 1
    trivial = (n_bins == 1 and t_lo == -TOF_TABLE_HALF_WIDTH_PS and t_hi == TOF_TABLE_HALF_WIDTH_PS)
    return TofKernel(n_bins, bin_width, w, kappa, beta, h, t_lo, t_hi, trivial)


def rings_used(n_rings: int, n_rings_used: int) -> np.ndarray:
    """``floor(linspace(0, NR-1, NRU) + 0.5)`` (``:153``)."""
    return np.floor(np.linspace(0, n_rings - 1, n_rings_used) + 0.5).astype(np.int64)


def detectors_used(n_det: int, n_det_used: int) -> np.ndarray:
    """Cumulative floor of N/NDU starting at 0 (``:159-167``); identical for every ring."""
    step = n_det / n_det_used
    dets = np.zeros(n_det_used, dtype=np.int64)
    for d in range(1, n_det_used):
        dets[d] = int(np.floor(dets[d - 1] + step))
    return dets


def strided_scatter_points(attenuation_map: np.ndarray, sample_step) -> np.ndarray:
    """Voxel indices (ix, iy, iz) of the reference's strided scatter points.

    Order is the reference's: z slices, then y, then x (``:182,387-394``);
    a voxel is a scatter point when mu > 0 (the tissue index is > 0 exactly then).
    """
    mu = np.asarray(attenuation_map)
    X, Y, Z = mu.shape
    sx, sy, sz = (int(v) for v in sample_step)
    pts = []
    for iz in range(0, Z, sz):
        for iy in range(0, Y, sy):
            for ix in range(0, X, sx):
                if mu[ix, iy, iz] > 0:
                    pts.append((ix, iy, iz))
    return np.asarray(pts, dtype=np.int32).reshape(-1, 3)


def scatter_point_coordinates(points: np.ndarray, grid_shape, image_size) -> np.ndarray:
    """Scatter-point positions (cm) with the reference's conventions.

    x and y sit on the upper voxel edge, both using the x half-width
    (``:176-177,389,394``); z is the voxel centre (``:87,179,381``).
    """
    X, Y, Z = (int(v) for v in grid_shape)
    image_size = np.asarray(image_size, dtype=np.float64)
    radius = -image_size[0]
    width = -image_size[2]
    z_vox = (image_size[5] - image_size[2]) / Z
    z_start = z_vox / 2 - width
    p = np.asarray(points, dtype=np.int64)
    out = np.empty((p.shape[0], 3), dtype=np.float64)
    out[:, 0] = ((p[:, 0] + 1) * 2 * radius / X) - radius
    out[:, 1] = ((p[:, 1] + 1) * 2 * radius / Y) - radius
    out[:, 2] = z_start + p[:, 2] * z_vox
    return out


def group_by_slice(points: np.ndarray):
    """Stable-sort scatter points by z index; return (points, slice_ptr, slice_z).

    The reference normalises per z slice (``:555-557``), so every z index forms
    one normalisation group; order inside a group is preserved.
    """
    p = np.asarray(points, dtype=np.int32).reshape(-1, 3)
    order = np.argsort(p[:, 2], kind='stable')
    p = p[order]
    zs, starts = np.unique(p[:, 2], return_index=True)
    ptr = np.append(starts, p.shape[0]).astype(np.int32)
    return np.ascontiguousarray(p), ptr, zs.astype(np.int32)
