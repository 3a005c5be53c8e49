"""Scanner geometry helpers, host side.

Drop-ins for ``Python/openSSS/Geometry.py``: ``ImportCASToRGeometry`` (:27-99),
``PrepareGeometryAndSinograms`` (:101-159), ``expandRings`` (:161-197) and
``mashDetectors`` (:200-235).  Units: cm for positions (the CASToR ``.glut``
holds mm and is divided by 10), mm for the crystal face size.
"""
from __future__ import annotations

import os
import re

import numpy as np

from .data_reduction import SinogramLUT
from .sinogram import SinogramCoordinates


def _header_int(header: str, key: str) -> int:
    return int(re.findall(r'.*' + key + r'.*: (\d*).*', header)[0])


def _header_float(header: str, key: str) -> float:
    return float(re.findall(r'.*' + key + r'.*: (\d*[.]?\d*).*', header)[0])


def ImportCASToRGeometry(scanner: str, ScannerPath: str):
    """Read ``<scanner>.geom`` (text header) and ``<scanner>.glut`` (float32 LUT).

    Returns the reference's 9-tuple: geometry (NR, N, 3) cm, normals (NR, N, 3),
    axial sectors, transaxial sectors, axial modules, transaxial modules,
    axial crystals, transaxial crystals and ``[size_axial, size_trans]`` (mm).
    """
    with open(os.path.join(ScannerPath, f'{scanner}.geom'), 'r') as fh:
        header = fh.read()
    n_elem = _header_int(header, 'number of elements')
    n_sec_ax = _header_int(header, 'number of rsectors axial')
    n_mod_ax = _header_int(header, 'number of modules axial')
    n_cry_ax = _header_int(header, 'number of crystals axial')
    size_ax = _header_float(header, 'crystals size axial')
    n_sec_tx = _header_int(header, 'number of rsectors')
    n_mod_tx = _header_int(header, 'number of modules transaxial')
    n_cry_tx = _header_int(header, 'number of crystals transaxial')
    size_tx = _header_float(header, 'crystals size transaxial')
    n_rings = n_mod_ax * n_cry_ax * n_sec_ax
    lut = np.fromfile(os.path.join(ScannerPath, f'{scanner}.glut'),
                      dtype=np.float32).reshape(n_elem, 6)
    geom = lut[:, 0:3].astype(np.float64).reshape(n_rings, n_elem // n_rings, 3) / 10
    normals = lut[:, 3:6].astype(np.float64).reshape(n_rings, n_elem // n_rings, 3) / 10
    return (geom, normals, n_sec_ax, n_sec_tx, n_mod_ax, n_mod_tx, n_cry_ax, n_cry_tx,
            np.array([size_ax, size_tx]))


def expandRings(Geometry: np.ndarray, NormalVectors: np.ndarray):
    """Insert the half-way ring between every pair of rings (2NR-1 rings).

    The normals of every expanded ring are those of ring 0, as in
    ``Geometry.py:187-189``.
    """
    n_cry = Geometry.shape[1]
    total = Geometry.shape[0] * 2 - 1
    geom = np.zeros((total, n_cry, 3))
    geom[0:total:2] = Geometry
    geom[1:total:2] = (Geometry[:-1] + Geometry[1:]) / 2
    normals = np.zeros_like(geom)
    normals[:, :, :] = NormalVectors[0, :, :]
    return geom, normals


def mashDetectors(Geometry: np.ndarray, NormalVectors: np.ndarray, combineSize: int):
    """Average groups of ``combineSize`` neighbouring crystals (nan-mean)."""
    n_cry = Geometry.shape[1]
    if n_cry % combineSize != 0:
        raise Exception("Number of detectors not divisible by combineSize: ", n_cry, combineSize)
    shape = (Geometry.shape[0], n_cry // combineSize, combineSize, 3)
    geom = np.nanmean(Geometry.reshape(shape), axis=2)
    normals = np.nanmean(NormalVectors.reshape(shape), axis=2)
    return geom, normals


def PrepareGeometryAndSinograms(ScannerName: str, Mash: int, Span: int, MRD: int,
                                ScannerPath: str, Shift: bool = True):
    """The reference's 10-tuple (``Geometry.py:101-159``) from a CASToR scanner."""
    (geom, normals, n_sec_ax, n_sec_tx, n_mod_ax, n_mod_tx, n_cry_ax, n_cry_tx,
     det_size) = ImportCASToRGeometry(ScannerName, ScannerPath)
    n_rings = geom.shape[0]
    if Mash > 1:
        geom, normals = mashDetectors(geom, normals, Mash)
        lor, sino, shift = SinogramCoordinates(1, n_sec_ax, n_mod_ax, 1, geom.shape[1],
                                               n_cry_ax, ID_Shidt=False)
    else:
        lor, sino, shift = SinogramCoordinates(n_sec_tx, n_sec_ax, n_mod_ax, n_mod_tx,
                                               n_cry_tx, n_cry_ax, ID_Shidt=Shift)
    if Span > 1:
        sino, _, lut, counts = SinogramLUT(n_rings, Span, MRD)
        ext_geom, ext_normals = expandRings(geom, normals)
    else:
        lut = np.array([])
        ext_geom = np.array([])
        ext_normals = np.array([])
        counts = np.array([])
    return geom, normals, det_size, lor, sino, lut, ext_geom, ext_normals, counts, shift


def cylinder_geometry(n_rings: int, n_det: int, radius_cm: float, ring_pitch_cm: float):
    """Synthetic cylindrical scanner: detector d at angle 2*pi*d/N, rings centred on z=0.

    Returns ``(Geometry (NR, N, 3) cm, NormalVectors (NR, N, 3))`` with inward normals.
    """
    phi = 2.0 * np.pi * np.arange(n_det) / n_det
    z = (np.arange(n_rings) + 0.5) * ring_pitch_cm - n_rings * ring_pitch_cm / 2.0
    geom = np.empty((n_rings, n_det, 3))
    geom[:, :, 0] = radius_cm * np.cos(phi)[None, :]
    geom[:, :, 1] = radius_cm * np.sin(phi)[None, :]
    geom[:, :, 2] = z[:, None]
    normals = np.empty_like(geom)
    normals[:, :, 0] = -np.cos(phi)[None, :]
    normals[:, :, 1] = -np.sin(phi)[None, :]
    normals[:, :, 2] = 0.0
    return geom, normals
