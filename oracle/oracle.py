"""ORACLE — test infrastructure only.

Python side of the CPU float64 restatement of the reference TOF-SSS path
(``oracle/sss_oracle.c``).  Imported by ``tests/``, ``__graft_entry__.smoke()``
and ``bench.py``'s cpu_baseline leg, and only there, as the checker; the
shipped package never imports it.

Functions mirror the reference entry points they restate:

* ``trace``          -> ``Python/openSSS/RayTracing3DTOF.py:29-180``
* ``sss``            -> ``Python/openSSS/SingleScatterSimulationTOF.py:36-208,300-559``
                        (explicit, slice-grouped scatter-point list; dense output)
* ``line_integrals`` -> the per-pair trace + sums of ``TailsMask.py:139-148``
* ``backproject``    -> ``BackProjection.py:31-113`` (mode 0) / ``:217-300`` (mode 1)
* ``sinogram_coordinates_loops`` -> ``Sinogram.py:27-141`` as plain loops
* ``tail_fit``       -> ``BackProjection.py:115-214`` / ``:302-392``
* ``scale_scatters_to_prompts`` -> the scaling loop of ``Scaling.py:60-113``
* ``parse_castor_ids``, ``export_prompts_loops``, ``inject_scatters_loops``
  -> ``Timber/Coincidence.py:45-159``, ``Timber/Datafile.py:66-113,161-230``
* ``crop_and_downscale`` -> ``Utils.py:212-267`` (scikit-image's resize as
  it states it, through scipy.ndimage: parity with scikit-image unpinned)
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, 'build', 'liboracle.so')
_lib = None

LIGHT_CM_PS = 299792458.0 * 1e-10


def build(force: bool = False) -> str:
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(['make', '-C', _HERE, '-B' if force else 'all'], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def use_native_build() -> str:
    """Build the oracle at -O3 -march=native on this machine (always afresh:
    a copy built elsewhere may target another CPU) and load that build from
    now on.  Used by the timed CPU baseline only; the parity tests use the
    portable build."""
    global _LIB_PATH, _lib
    subprocess.run(['make', '-C', _HERE, '-B', 'native'], check=True, stdout=subprocess.DEVNULL)
    _LIB_PATH = os.path.join(_HERE, 'build', 'liboracle_native.so')
    _lib = None
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.ora_trace.restype = ctypes.c_int
        _lib.ora_sss.restype = ctypes.c_int
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class _SssArgs(ctypes.Structure):
    _fields_ = [
        ('grid', ctypes.c_int32 * 3), ('bounds', ctypes.c_double * 6),
        ('mu', ctypes.c_void_p), ('lam', ctypes.c_void_p),
        ('n_rings', ctypes.c_int32), ('n_det', ctypes.c_int32),
        ('geom', ctypes.c_void_p), ('normals', ctypes.c_void_p),
        ('det_size', ctypes.c_double * 2),
        ('ratio', ctypes.c_void_p), ('eff', ctypes.c_void_p),
        ('n_energy', ctypes.c_int32), ('ref_row', ctypes.c_int32),
        ('tof', ctypes.c_void_p), ('n_bins', ctypes.c_int32),
        ('nru', ctypes.c_int32), ('ndu', ctypes.c_int32),
        ('rings', ctypes.c_void_p), ('dets', ctypes.c_void_p),
        ('sino', ctypes.c_void_p), ('cell_of', ctypes.c_void_p), ('n_cells', ctypes.c_int32),
        ('n_sp', ctypes.c_int32), ('sp_xyz', ctypes.c_void_p),
        ('n_slices', ctypes.c_int32), ('slice_ptr', ctypes.c_void_p),
        ('n_threads', ctypes.c_int32), ('stat_oor', ctypes.c_void_p),
    ]


def trace(grid_size, grid_bounds, line):
    """Returns (L f64, idx int32 1-based, s f64) exactly like RayTracing3DTOF."""
    gs = np.ascontiguousarray(np.asarray(grid_size, dtype=np.int32))
    gb = np.ascontiguousarray(np.asarray(grid_bounds, dtype=np.float64))
    lc = np.ascontiguousarray(np.asarray(line, dtype=np.float64))
    cap = int(gs.sum())
    L = np.zeros(cap)
    idx = np.zeros(cap, dtype=np.int32)
    s = np.zeros(cap)
    n = lib().ora_trace(_p(gs), _p(gb), _p(lc), _p(L), _p(idx), _p(s), ctypes.c_int(cap))
    if n < 0:
        raise RuntimeError('trace exceeded capacity')
    return L[:n].copy(), idx[:n].copy(), s[:n].copy()


def tof_table(n_bins, bin_width, resolution):
    """The reference's TOFEfficiencyTable (``SingleScatterSimulationTOF.py:260-295``)."""
    t = np.arange(-4000, 4001, dtype=np.float64)
    shifts = ((np.arange(1, n_bins + 1) - (n_bins + 1) / 2) * bin_width).reshape(-1, 1)
    p = t.reshape(1, -1) - shifts
    p = np.exp(-(p ** 2) / (resolution ** 2 / 4 / np.log(2)))
    zero = np.sum(p, axis=0) == 0
    with np.errstate(invalid='ignore', divide='ignore'):
        p = p / np.sum(p, axis=0)
    p[:, zero] = 0
    return p


def energy_tables(attenuation_table, energy_resolution, energy_threshold):
    from scipy.special import erf
    tab = np.asarray(attenuation_table, dtype=np.float64)
    ref = int(np.where(tab[:, 0] == 511)[0][0])
    ratio = tab[:, 1] / tab[ref, 1]
    eff = 0.5 * (1 - erf((energy_threshold - tab[:, 0]) / (511 * energy_resolution / 2 / np.sqrt(np.log(2)))))
    return ratio, eff


def sampled_cells(sino, dets):
    """Compact id per (radial, angular) cell hit by a sampled pair (D1 <= D2)."""
    N = sino.shape[0]
    cell_of = np.full((N + 1) * (N // 2), -1, dtype=np.int32)
    n = 0
    for i1 in range(len(dets)):
        for i2 in range(i1, len(dets)):
            ang = int(sino[dets[i1], dets[i2], 0])
            rad = int(sino[dets[i1], dets[i2], 1])
            c = rad * (N // 2) + ang
            if cell_of[c] < 0:
                cell_of[c] = n
                n += 1
    return cell_of, n


def sss(activity, attenuation, image_size, geometry, sino, normals, det_size, attenuation_table,
        energy_resolution, energy_threshold, rings, dets, tof_range, tof_resolution, n_bins,
        sp_points, slice_ptr, n_threads=1, dense=True, return_stats=False):
    """Oracle TOF-SSS.  ``sp_points`` (n, 3) int voxel indices grouped by z (``slice_ptr``).

    Returns float32 (B, N+1, N//2, NRU**2) when ``dense`` else the compact
    (B, NRU**2, n_cells) array plus the (radial, angular) -> cell map.
    """
    X, Y, Z = attenuation.shape
    mu = np.ascontiguousarray(np.asarray(attenuation, dtype=np.float64).flatten(order='F'))
    lam = np.ascontiguousarray(np.asarray(activity, dtype=np.float64).flatten(order='F'))
    geom = np.ascontiguousarray(geometry, dtype=np.float64)
    nv = np.ascontiguousarray(normals, dtype=np.float64)
    ratio, eff = energy_tables(attenuation_table, energy_resolution, energy_threshold)
    ratio = np.ascontiguousarray(ratio)
    eff = np.ascontiguousarray(eff)
    bin_width = 2 * tof_range / n_bins
    tof = np.ascontiguousarray(tof_table(n_bins, bin_width, tof_resolution))
    rings = np.ascontiguousarray(np.asarray(rings, dtype=np.int32))
    dets = np.ascontiguousarray(np.asarray(dets, dtype=np.int32))
    sino16 = np.ascontiguousarray(np.asarray(sino, dtype=np.int16))
    cell_of, n_cells = sampled_cells(sino16, dets)
    isz = np.asarray(image_size, dtype=np.float64)
    radius = -isz[0]
    width = -isz[2]
    zvox = (isz[5] - isz[2]) / Z
    zstart = zvox / 2 - width
    p = np.asarray(sp_points, dtype=np.int64).reshape(-1, 3)
    xyz = np.empty((p.shape[0], 3))
    xyz[:, 0] = ((p[:, 0] + 1) * 2 * radius / X) - radius
    xyz[:, 1] = ((p[:, 1] + 1) * 2 * radius / Y) - radius
    xyz[:, 2] = zstart + p[:, 2] * zvox
    xyz = np.ascontiguousarray(xyz)
    sptr = np.ascontiguousarray(np.asarray(slice_ptr, dtype=np.int32))
    oor = np.zeros(1, dtype=np.int64)
    a = _SssArgs()
    a.grid[:] = [X, Y, Z]
    a.bounds[:] = [float(v) for v in isz]
    a.mu, a.lam = mu.ctypes.data, lam.ctypes.data
    a.n_rings, a.n_det = geom.shape[0], geom.shape[1]
    a.geom, a.normals = geom.ctypes.data, nv.ctypes.data
    a.det_size[:] = [float(det_size[0]), float(det_size[1])]
    a.ratio, a.eff = ratio.ctypes.data, eff.ctypes.data
    a.n_energy = ratio.shape[0]
    a.ref_row = 511 * 2 - 1
    a.tof, a.n_bins = tof.ctypes.data, n_bins
    a.nru, a.ndu = len(rings), len(dets)
    a.rings, a.dets = rings.ctypes.data, dets.ctypes.data
    a.sino, a.cell_of, a.n_cells = sino16.ctypes.data, cell_of.ctypes.data, n_cells
    a.n_sp, a.sp_xyz = xyz.shape[0], xyz.ctypes.data
    a.n_slices, a.slice_ptr = len(sptr) - 1, sptr.ctypes.data
    a.n_threads = n_threads
    a.stat_oor = oor.ctypes.data
    nru = len(rings)
    out = np.zeros((n_bins, nru * nru, n_cells), dtype=np.float32)
    rc = lib().ora_sss(ctypes.byref(a), _p(out))
    if rc != 0:
        raise RuntimeError(f'ora_sss failed: {rc}')
    stats = {'offsets_out_of_range': int(oor[0])}
    if dense:
        N = geom.shape[1]
        res = np.zeros((n_bins, N + 1, N // 2, nru * nru), dtype=np.float32)
        cells = np.nonzero(cell_of >= 0)[0]
        rad, ang = cells // (N // 2), cells % (N // 2)
        res[:, rad, ang, :] = np.transpose(out[:, :, cell_of[cells]], (0, 2, 1))
        return (res, stats) if return_stats else res
    return (out, cell_of, stats) if return_stats else (out, cell_of)


def line_integrals(grid_size, grid_bounds, mu_flat_f, lam_flat_f, segs):
    gs = np.ascontiguousarray(np.asarray(grid_size, dtype=np.int32))
    gb = np.ascontiguousarray(np.asarray(grid_bounds, dtype=np.float64))
    mu = np.ascontiguousarray(mu_flat_f, dtype=np.float64)
    lam = np.ascontiguousarray(lam_flat_f, dtype=np.float64)
    segs = np.ascontiguousarray(segs, dtype=np.float64).reshape(-1, 6)
    n = np.zeros(segs.shape[0], dtype=np.int32)
    m = np.zeros(segs.shape[0])
    l = np.zeros(segs.shape[0])
    lib().ora_line_integrals(_p(gs), _p(gb), _p(mu), _p(lam), _p(segs),
                             ctypes.c_int64(segs.shape[0]), _p(n), _p(m), _p(l))
    return n, m, l


def backproject(grid_size, grid_bounds, segs, v_prompt, v_scatter, v_random, mode):
    gs = np.ascontiguousarray(np.asarray(grid_size, dtype=np.int32))
    gb = np.ascontiguousarray(np.asarray(grid_bounds, dtype=np.float64))
    segs = np.ascontiguousarray(segs, dtype=np.float64).reshape(-1, 6)
    vp = np.ascontiguousarray(v_prompt, dtype=np.float32)
    vs = np.ascontiguousarray(v_scatter, dtype=np.float32)
    vr = np.ascontiguousarray(v_random, dtype=np.float32)
    nvox = int(np.prod(gs))
    imgs = [np.zeros(nvox, dtype=np.float32) for _ in range(4)]
    lib().ora_backproject(_p(gs), _p(gb), _p(segs), ctypes.c_int64(segs.shape[0]),
                          _p(vp), _p(vs), _p(vr), ctypes.c_int(mode),
                          *[_p(i) for i in imgs])
    return imgs


def sinogram_coordinates_loops(n_sectors_trans, n_sectors_axial, n_modules_axial, n_modules_trans,
                               n_crystals_trans, n_crystals_axial, min_sector_difference=0,
                               id_shift=True):
    """Plain-loop restatement of ``Sinogram.py:55-141`` (small scanners only)."""
    NR = n_sectors_axial * n_modules_axial * n_crystals_axial
    N = n_sectors_trans * n_modules_trans * n_crystals_trans
    mcd = min_sector_difference * n_modules_trans * n_crystals_trans
    radial_size = N - 2 * (mcd - 1) - 1
    shift = (n_modules_trans * n_crystals_trans) // 2 if id_shift else 0
    lor = np.zeros((N, N, 2)) - 1
    for d1 in range(N):
        c1 = d1 % N - shift
        for d2 in range(N):
            c2 = d2 % N - shift
            if c1 < 0:
                c1 += N
            if c2 < 0:
                c2 += N
            ida, idb = (c1, c2) if c1 < c2 else (c2, c1)
            if idb - ida < mcd:
                continue
            if ida + idb >= (3 * N) // 2 or ida + idb < N // 2:
                radial = -N // 2 if ida == idb else ((idb - ida - 1) // 2) - ((N - (idb - ida + 1)) // 2)
            else:
                radial = N // 2 if ida == idb else ((N - (idb - ida + 1)) // 2) - ((idb - ida - 1) // 2)
            if ida + idb < N // 2:
                angular = (2 * ida + N + radial) // 2
            elif ida + idb >= (3 * N) // 2:
                angular = (2 * ida - N + radial) // 2
            else:
                angular = (2 * ida - radial) // 2
            lor[d1, d2, 0] = angular
            lor[d1, d2, 1] = radial + radial_size // 2
    sino = np.zeros((NR, NR))
    for r1 in range(NR):
        for r2 in range(NR):
            if r1 == r2:
                idx = r1
            elif r2 < r1:
                d = r1 - r2
                idx = (2 * d - 1) * NR - d * (d - 1) + r2
            else:
                d = r2 - r1
                idx = (2 * d - 1) * NR - d * (d - 1) + (NR - d) + r1
            sino[r1, r2] = idx
    return lor.astype(np.int16), sino.astype(np.int16), shift


# ---------------------------------------------------------------------------
# Sinogram stages after the scatter estimate (small cases only: plain loops)
# ---------------------------------------------------------------------------

def _nearest_fill(F):
    """fill_with_interp1d (Interpolation.py:133-164) for a nearest-neighbour
    griddata result: no NaN can occur, so only the clamp at 0 applies."""
    F = F.copy()
    F[F < 0] = 0
    return F


def interpolate_bins(Scatters, rings, dets, SinogramCoordinates, SinogramIndex=None,
                     MashedSinogramIndices=None):
    """Restatement of InterpolateAllBins / InterpolateAllBinsSpan
    (Interpolation.py:197-319, :321-399), one bin at a time:
    griddata(nearest) of each sampled plane over the (radial, angular) grid,
    clamp at 0, then RegularGridInterpolator(nearest) over the sampled rings
    evaluated at every output ring pair (without span) or at the averaged
    rings of every span plane.  Returns (B, N+1, N//2, P) float32."""
    from scipy.interpolate import RegularGridInterpolator, griddata
    S = np.asarray(Scatters, dtype=np.float32)
    B, NR1, NA, NP = S.shape
    nru = len(rings)
    lor = np.asarray(SinogramCoordinates)
    gx, gy = np.mgrid[0:NR1, 0:NA]
    out = []
    for b in range(B):
        cur = S[b].copy()
        for i in range(NP):
            rad = lor[np.ix_(dets, dets, [1])][:, :, 0]
            ang = lor[np.ix_(dets, dets, [0])][:, :, 0]
            pts = np.array([(y, x) for y, x in zip(rad.flatten(), ang.flatten())])
            vals = cur[rad.astype(int), ang.astype(int), i].flatten()
            cur[:, :, i] = _nearest_fill(griddata(pts, vals, (gx, gy), method='nearest'))
        cur = np.transpose(cur.reshape(NR1, NA, nru, nru), (0, 1, 3, 2))
        rgi = RegularGridInterpolator((np.arange(NR1), np.arange(NA), np.asarray(rings, float),
                                       np.asarray(rings, float)), cur, method='nearest',
                                      bounds_error=False, fill_value=0)
        if MashedSinogramIndices is None:
            NR = np.asarray(SinogramIndex).shape[0]
            xi = np.meshgrid(np.arange(NR1), np.arange(NA), np.arange(NR), np.arange(NR), indexing='ij')
            full = rgi(np.stack([x.flatten() for x in xi], axis=-1)).reshape(NR1, NA, NR * NR)
            order = np.argsort(np.asarray(SinogramIndex)[:NR, :NR].T.flatten())
            out.append(full[:, :, order].astype(np.float32))
        else:
            msi = np.asarray(MashedSinogramIndices, dtype=np.float64)
            planes = []
            for i in range(msi.shape[0]):
                xi = np.meshgrid(np.arange(NR1), np.arange(NA), [msi[i, 0]], [msi[i, 1]], indexing='ij')
                planes.append(rgi(np.stack([x.flatten() for x in xi], axis=-1)).reshape(NR1, NA))
            out.append(np.stack(planes, axis=-1).astype(np.float32))
    return np.stack(out)


def _lor_segment(geom, ra, d1, rb, d2):
    """Segment of a trace loop: detector x, y of its ring, ring z from detector 0."""
    return [geom[ra, d1, 0], geom[ra, d1, 1], geom[ra, 0, 2],
            geom[rb, d2, 0], geom[rb, d2, 1], geom[rb, 0, 2]]


def tail_mask(ActivityMap, AttenuationMap, ImageSize, Geometry, SinogramCoordinates, UseAttenuation,
              AccelerationFactor, SinogramIndex=None, MashedSinogramIndices=None):
    """Restatement of MaskGenerator (TailsMask.py:84-228) and
    MaskGeneratorMashing (:230-376); traces through ora_line_integrals."""
    geom = np.asarray(Geometry, dtype=np.float64)
    NR, N = geom.shape[0], geom.shape[1]
    lor = np.asarray(SinogramCoordinates)
    grid = np.asarray(np.asarray(ActivityMap).shape)
    mu = np.asarray(AttenuationMap, dtype=np.float64).flatten(order='F')
    lam = np.asarray(ActivityMap, dtype=np.float64).flatten(order='F')
    af = AccelerationFactor or 1

    def masks(pairs):
        """pairs: list of (Ring1, Ring2); returns uint8 (N+1, N//2, len(pairs))."""
        segs, where = [], []
        for p, (r1, r2) in enumerate(pairs):
            for d1 in range(0, N, af):
                for d2 in range(0, N, af):
                    if (d1 == d2 and r2 == r1) or d1 > d2:
                        continue
                    segs.append(_lor_segment(geom, r1, d1, r2, d2))
                    where.append((lor[d1, d2, 1], lor[d1, d2, 0], p))
        out = np.zeros((N + 1, N // 2, len(pairs)), dtype=np.uint8)
        if not segs:
            return out
        n, m, l = line_integrals(grid, ImageSize, mu, lam, np.array(segs))
        for k, (rad, ang, p) in enumerate(where):
            if n[k] != 0:
                act = l[k]
                att = 1 / np.exp(-m[k])
            else:
                act, att = 0, 1
            att = 0 if att != 1 else 1
            act = 0 if act > 0 else 1
            out[rad, ang, p] += np.uint8(att if UseAttenuation else act)
        return out

    if MashedSinogramIndices is None:
        # Sinograms[:, :, Ring2, Ring1], reshaped and reordered (:209-219)
        sino = np.zeros((N + 1, N // 2, NR, NR), dtype=np.uint8)
        for r1 in range(NR):
            sino[:, :, :, r1] = masks([(r1, r2) for r2 in range(NR)])
        order = np.argsort(np.asarray(SinogramIndex)[:NR, :NR].T.flatten())
        return sino.reshape(N + 1, N // 2, NR * NR)[:, :, order]
    msi = np.asarray(MashedSinogramIndices)
    return masks([(int(msi[i][0] * 2), int(msi[i][1] * 2)) for i in range(msi.shape[0])])


def tail_fit(PromptSinogram, ScatterSinogram, RandomSinogram, MaskSinogram, GridSize, ImageSize, Geometry,
             SinogramCoordinates, AccelerationFactor, SinogramOrder=None, MashedSinogramIndices=None):
    """Restatement of BackProjectionScatterRandFit (BackProjection.py:31-214)
    and ...Mashing (:217-392): float32 images per worker (per Ring1, or per
    span plane) summed in float32, float32 ratios, float32 lstsq scaled by
    1e10.  Returns (scale float32, images dict)."""
    geom = np.asarray(Geometry, dtype=np.float64)
    NR, N = geom.shape[0], geom.shape[1]
    lor = np.asarray(SinogramCoordinates)
    gs = np.asarray(GridSize, dtype=np.int32)
    af = AccelerationFactor or 1
    S = np.asarray(MaskSinogram, np.float32) * np.asarray(ScatterSinogram, np.float32)
    R = np.asarray(MaskSinogram, np.float32) * np.asarray(RandomSinogram, np.float32)
    P = np.asarray(MaskSinogram, np.float32) * np.float32(PromptSinogram)
    nvox = int(np.prod(gs))
    acc = [np.zeros(nvox, np.float32) for _ in range(4)]

    def worker(r1, r2_list, plane_of, mode):
        segs, vp, vs, vr = [], [], [], []
        for d1 in range(0, N, af):
            for r2 in r2_list:
                for d2 in range(0, N, af):
                    if (d1 == d2 and r2 == r1) or d1 > d2:
                        continue
                    rad, ang = lor[d1, d2, 1], lor[d1, d2, 0]
                    pl = plane_of(r2)
                    segs.append(_lor_segment(geom, r1, d1, r2, d2))
                    vp.append(P[rad, ang, pl]); vs.append(S[rad, ang, pl]); vr.append(R[rad, ang, pl])
        if segs:
            imgs = backproject(gs, ImageSize, np.array(segs), np.array(vp), np.array(vs), np.array(vr), mode)
            for a_, b_ in zip(acc, imgs):
                a_ += b_

    if MashedSinogramIndices is None:
        order = np.asarray(SinogramOrder).reshape(NR, NR)
        for r1 in range(NR):
            worker(r1, range(NR), lambda r2, r1=r1: order[r1, r2], 0)
    else:
        msi = np.asarray(MashedSinogramIndices)
        for i in range(msi.shape[0]):
            worker(int(msi[i, 0] * 2), [int(msi[i, 1] * 2)], lambda r2, i=i: i, 1)
    prompt_img, scatter_img, random_img, sens = acc
    eps = 0 if MashedSinogramIndices is None else 1e-15
    with np.errstate(divide='ignore', invalid='ignore'):
        sbp = scatter_img / (sens + np.float32(eps))
        abp = prompt_img / (sens + np.float32(eps))
        rbp = random_img / (sens + np.float32(eps))
    for x in (sbp, abp, rbp):
        x[np.isnan(x)] = 0
        x[np.isinf(x)] = 0
    tails = abp - rbp
    tails[tails < 0] = 0
    a, _, _, _ = np.linalg.lstsq(sbp[:, np.newaxis] * np.float32(1e10), tails, rcond=None)
    scale = a * np.float32(1e10)
    return scale, {'prompt': prompt_img, 'scatter': scatter_img, 'random': random_img, 'sens': sens}


def scale_scatters_to_prompts(Prompts, Scatters, Randoms, Normalization, Mask, LORCounts, SinogramIndex,
                              SinogramCoordinates, DesiredDimensions, FittingSize, Geometry, ExtendedGeometry,
                              AccelerationFactor, SpanFlag):
    """Restatement of the scaling loop of ``Scaling.py:60-113`` on in-memory
    arrays: Randoms * Normalization (:69-71, zeros without randoms), float32
    (mask >= 1) (:75-76), and per bin Prompts[..., b] / LORCounts (:90) into
    the no-span (:92-98, mirrored plane order SinogramIndex.T) or span
    (:99-104, extended geometry) fit.  Returns float32 (B,)."""
    prompts = np.asarray(Prompts)
    n_bins = prompts.shape[-1]
    cell_shape = prompts.shape[:-1]
    norm = np.ones(cell_shape, np.single) if Normalization is None else np.asarray(Normalization, np.float32)
    randoms = (np.zeros(cell_shape, np.single) if Randoms is None
               else np.asarray(Randoms, np.float32) * norm)
    mask = np.float32(np.asarray(Mask) >= 1)
    out = np.zeros(n_bins, dtype=np.float32)
    for b in range(n_bins):
        bin_prompts = prompts[:, :, :, b] / LORCounts
        if not SpanFlag:
            NR = np.asarray(SinogramIndex).shape[0]
            order = np.asarray(SinogramIndex)[0:NR, 0:NR].T
            f, _ = tail_fit(bin_prompts, Scatters[b], randoms, mask, DesiredDimensions, FittingSize, Geometry,
                            SinogramCoordinates, AccelerationFactor, SinogramOrder=order)
        else:
            f, _ = tail_fit(bin_prompts, Scatters[b], randoms, mask, DesiredDimensions, FittingSize,
                            ExtendedGeometry, SinogramCoordinates, AccelerationFactor,
                            MashedSinogramIndices=SinogramIndex)
        out[b] = f[0]
    return out


# ---- CASToR list mode (Timber/Coincidence.py, Timber/Datafile.py) ----

def listmode_event_format(attenuation, scatter, random, normalization, tof, max_lines):
    """struct format and positions (scatter, random, norm, TOF, first ID) of
    a list-mode event, Coincidence.py:45-95."""
    lines_flag = max_lines > 1
    fmt = '<1L'
    nc = attenuation + scatter + random + normalization + tof
    if nc > 0:
        fmt += f'{nc}f'
    if lines_flag:
        fmt += '1I'
    fmt += f'{2 * max_lines}L'
    base = 1 + attenuation
    pos = (base if scatter else None,
           base + scatter if random else None,
           base + scatter + random if normalization else None,
           base + scatter + random + normalization if tof else None,
           base + scatter + random + normalization + tof + lines_flag)
    return fmt, pos


def parse_castor_ids(event, pos, ncpr, shift=0):
    """ParseCASToRIDs (Coincidence.py:97-159) on one unpacked event."""
    a, b = event[pos[4]], event[pos[4] + 1]
    tof = event[pos[3]]
    lo, hi = (a, b) if a < b else (b, a)
    tof_lo = tof if a < b else -tof
    det_lo, det_hi = lo % ncpr, hi % ncpr
    d1, d2 = (det_lo, det_hi) if det_lo < det_hi else (det_hi, det_lo)
    s_lo, s_hi = lo % ncpr - shift, hi % ncpr - shift
    s_lo += ncpr if s_lo < 0 else 0
    s_hi += ncpr if s_hi < 0 else 0
    if s_lo < s_hi:
        return d1, lo // ncpr, d2, hi // ncpr, tof_lo
    return d1, hi // ncpr, d2, lo // ncpr, -tof_lo


def export_prompts_loops(data, flags, duration, LORCoordinates, SinogramIndex, TOF_bins, TOF_range, mash=1,
                         Shift=0):
    """The event loop of ExportPrompts (Datafile.py:66-113) on the datafile
    bytes; flags = (attenuation, scatter, random, normalization, tof,
    max_lines).  Returns (prompts uint16, randoms f32 or None, norm f32 or None)."""
    import struct
    fmt, pos = listmode_event_format(*flags)
    lor = np.asarray(LORCoordinates)
    ncpr = lor.shape[0] * mash
    size = [int(np.max(lor[:, :, 1])) + 1, int(np.max(lor[:, :, 0])) + 1, int(np.max(SinogramIndex)) + 1]
    prompts = np.zeros(size + [TOF_bins], dtype=np.uint16)
    randoms = np.zeros(size, dtype=np.single) if flags[2] else None
    norm = np.ones(size, dtype=np.single) if flags[3] else None
    for ev in struct.iter_unpack(fmt, data):
        d1, r1, d2, r2, tof = parse_castor_ids(ev, pos, ncpr, Shift)
        c0 = lor[d1 // mash, d2 // mash, 0]
        c1 = lor[d1 // mash, d2 // mash, 1]
        sino = SinogramIndex[r1, r2]
        if sino == -1 or c0 == -1 or c1 == -1:
            continue
        if randoms is not None:
            randoms[c1, c0, sino] = ev[0] * duration / TOF_bins
        if norm is not None:
            norm[c1, c0, sino] = ev[1]
        if abs(tof) > TOF_range:
            continue
        tb = (tof + TOF_range) / (2 * TOF_range / TOF_bins)
        if tb < 0 or tb >= TOF_bins:
            continue
        with np.errstate(over='ignore'):   # uint16 counts wrap, as the reference's
            prompts[c1, c0, sino, int(tb)] += np.uint16(1)
    return prompts, randoms, norm


def inject_scatters_loops(data, flags, duration, LORCoordinates, RingTable, scatter_sinograms, Scale, TOF_range,
                          mash=1, Shift=0):
    """The event loop of InjectScatters (Datafile.py:161-230) on the datafile
    bytes: returns the bytes of the new datafile."""
    import struct
    fmt, pos = listmode_event_format(*flags)
    out_flags = (flags[0], 1, flags[2], flags[3], flags[4], flags[5])
    out_fmt, out_pos = listmode_event_format(*out_flags)
    lor = np.asarray(LORCoordinates)
    ncpr = lor.shape[0] * mash
    n_bins = len(Scale)
    scale = np.array(Scale, dtype=np.float32, copy=True)
    scale /= (2 * TOF_range / n_bins)
    sf = 1. / (0.299792458 / 2)
    out = bytearray()
    for ev in struct.iter_unpack(fmt, data):
        d1, r1, d2, r2, tof = parse_castor_ids(ev, pos, ncpr, Shift)
        if abs(tof) > TOF_range:
            continue
        c0 = lor[d1 // mash, d2 // mash, 0]
        c1 = lor[d1 // mash, d2 // mash, 1]
        sino = RingTable[r1, r2]
        if sino == -1 or c0 == -1 or c1 == -1:
            continue
        tb = (tof + TOF_range) / (2 * TOF_range / n_bins)
        if tb < 0 or tb >= n_bins:
            continue
        tb = int(tb)
        rate = scale[tb] * sf / duration * scatter_sinograms[tb, c1, c0, sino]
        out += struct.pack(out_fmt, *ev[:out_pos[0]], rate, *ev[out_pos[0]:])
    return bytes(out)


# ---- image preparation (Utils.py:212-267) ----

def crop_and_downscale(image, voxel_size, crop_size, downscaled_dimensions, anti_aliasing=True, clip=True):
    """Restatement of CropAndDownscale (Utils.py:212-267): the crop loop as
    written there, then scikit-image's resize(order=3, mode='edge',
    preserve_range=True) as that library states it, through the scipy.ndimage
    routines it calls: gaussian_filter (sigma = max(0, (f - 1) / 2), mode
    'nearest'), zoom (order 3, mode 'nearest', grid_mode=True) and the clip
    to the input's range.  float32 in, float32 out."""
    from scipy import ndimage as ndi
    cropped = crop(image, voxel_size, crop_size)
    factors = np.divide(cropped.shape, downscaled_dimensions)
    filtered = cropped
    if anti_aliasing:
        filtered = ndi.gaussian_filter(cropped, np.maximum(0, (factors - 1) / 2), mode='nearest')
    out = ndi.zoom(filtered, [1 / f for f in factors], order=3, mode='nearest', grid_mode=True)
    if clip:
        np.clip(out, np.min(cropped), np.max(cropped), out=out)
    return out


def crop(image, voxel_size, crop_size):
    """The crop / zero-pad loop of CropAndDownscale (Utils.py:243-262)."""
    image = np.asarray(image, dtype=np.float32)
    half = np.ceil(np.array(crop_size) / np.array(voxel_size) / 2).astype(int)
    center = np.array(image.shape) // 2
    cropped = image
    for dim in range(3):
        start, end = center[dim] - half[dim], center[dim] + half[dim]
        if start < 0 and end > image.shape[dim]:
            pw = [(0, 0)] * 3
            pw[dim] = (half[dim] - center[dim], half[dim] - center[dim])
            cropped = np.pad(cropped, pw, mode='constant', constant_values=0)
        else:
            sl = [slice(None)] * 3
            sl[dim] = slice(start, end)
            cropped = cropped[tuple(sl)]
    return cropped
