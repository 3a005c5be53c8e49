"""Seeded synthetic scanners, phantoms and scatter-point lists.

These build the workloads the benchmarks and tests run on: a cylindrical
scanner, an elliptical water body with a lung cylinder and hot spheres, and a
seeded uniform scatter-point sample inside mu > 0.03 cm^-1.  Units follow the
reference: cm for geometry and images, cm^-1 for mu, mm for the crystal face,
ps for TOF.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .geometry import cylinder_geometry, mashDetectors
from .sinogram import SinogramCoordinates
from .tables import group_by_slice

WATER_MU = 0.096       # cm^-1 at 511 keV
LUNG_MU = 0.029
BONE_MU = 0.150


@dataclass
class Workload:
    """Everything one TOF-SSS call needs, in the reference's argument conventions."""
    name: str
    activity: np.ndarray          # (X, Y, Z) float64
    attenuation: np.ndarray       # (X, Y, Z) float64, cm^-1
    image_size: np.ndarray        # (6,) cm
    geometry: np.ndarray          # (NR, N, 3) cm
    normals: np.ndarray           # (NR, N, 3)
    lor_coordinates: np.ndarray   # (N, N, 2) int16
    sinogram_index: np.ndarray    # (NR, NR) int16
    det_size: np.ndarray          # (2,) mm
    energy_resolution: float
    energy_threshold: float
    n_rings_used: int
    n_det_used: int
    tof_resolution: float
    tof_range: float
    n_bins: int
    scatter_points: np.ndarray    # (n, 3) int32 voxel indices, grouped by z
    slice_ptr: np.ndarray
    extra: dict = field(default_factory=dict)

    @property
    def grid(self):
        return tuple(int(v) for v in self.attenuation.shape)


def voxel_centres(shape, image_size):
    X, Y, Z = shape
    isz = np.asarray(image_size, dtype=np.float64)
    dx = (isz[3] - isz[0]) / X
    dy = (isz[4] - isz[1]) / Y
    dz = (isz[5] - isz[2]) / Z
    x = isz[0] + (np.arange(X) + 0.5) * dx
    y = isz[1] + (np.arange(Y) + 0.5) * dy
    z = isz[2] + (np.arange(Z) + 0.5) * dz
    return np.meshgrid(x, y, z, indexing='ij')


def nema_like_phantom(shape, image_size, body_axes=(15.0, 11.5), lung_radius=2.5,
                      sphere_diameters=(1.0, 1.3, 1.7, 2.2, 2.8, 3.7), sphere_ring=5.7,
                      sphere_contrast=4.0, inserts=None):
    """Elliptical water body over the whole axial extent, centred lung cylinder
    (mu 0.029, no activity), hot spheres on a ring at z = 0."""
    xs, ys, zs = voxel_centres(shape, image_size)
    a, b = body_axes
    body = (xs / a) ** 2 + (ys / b) ** 2 <= 1.0
    mu = np.where(body, WATER_MU, 0.0)
    lam = np.where(body, 1.0, 0.0)
    if lung_radius > 0:
        lung = xs ** 2 + ys ** 2 <= lung_radius ** 2
        mu = np.where(lung & body, LUNG_MU, mu)
        lam = np.where(lung, 0.0, lam)
    for k, dmt in enumerate(sphere_diameters):
        ang = 2 * np.pi * k / max(len(sphere_diameters), 1)
        cx, cy = sphere_ring * np.cos(ang), sphere_ring * np.sin(ang)
        sph = (xs - cx) ** 2 + (ys - cy) ** 2 + zs ** 2 <= (dmt / 2) ** 2
        lam = np.where(sph & body, sphere_contrast, lam)
    if inserts:
        for (cx, cy, r, m) in inserts:
            ins = ((xs - cx) ** 2 + (ys - cy) ** 2 <= r ** 2) & body
            mu = np.where(ins, m, mu)
    return np.ascontiguousarray(lam, dtype=np.float64), np.ascontiguousarray(mu, dtype=np.float64)


def sample_scatter_points(mu, n, seed, mu_min=0.03):
    """Seeded uniform sample (without replacement) of voxels with mu > mu_min,
    returned grouped by z (the reference's normalisation slices)."""
    cand = np.argwhere(mu > mu_min)
    rng = np.random.default_rng(seed)
    n = min(int(n), cand.shape[0])
    pick = rng.choice(cand.shape[0], size=n, replace=False)
    pts = cand[np.sort(pick)]
    # reference order inside a slice: y then x
    order = np.lexsort((pts[:, 0], pts[:, 1], pts[:, 2]))
    pts = pts[order].astype(np.int32)
    pts, ptr, _ = group_by_slice(pts)
    return pts, ptr


def _workload(name, n_rings, n_det, ring_pitch, shape, voxel, n_sp, seed, nru, ndu, n_bins,
              tof_res, tof_range, radius=41.0, **phantom_kw):
    geom, normals = cylinder_geometry(n_rings, n_det, radius, ring_pitch)
    half = np.array(shape) * np.array(voxel) / 2.0
    image_size = np.array([-half[0], -half[1], -half[2], half[0], half[1], half[2]])
    lam, mu = nema_like_phantom(shape, image_size, **phantom_kw)
    pts, ptr = sample_scatter_points(mu, n_sp, seed)
    lor, sino, _ = SinogramCoordinates(n_det, 1, 1, 1, 1, n_rings)
    return Workload(name=name, activity=lam, attenuation=mu, image_size=image_size,
                    geometry=geom, normals=normals, lor_coordinates=lor, sinogram_index=sino,
                    det_size=np.array([4.0, 4.0]), energy_resolution=0.10,
                    energy_threshold=435.0, n_rings_used=nru, n_det_used=ndu,
                    tof_resolution=tof_res, tof_range=tof_range, n_bins=n_bins,
                    scatter_points=pts, slice_ptr=ptr)


def config1(seed=1):
    """Plain non-TOF SSS: 32 sampled detectors x 8 rings, 48x48x24 of 8x8x10.7 mm,
    500 sp, B = 1.  The scanner has 192 detectors per ring (a 193 x 96 x 64
    sinogram, the layout of the tail fit's prompts); the SSS samples every
    sixth detector, at the angles of a 32-detector ring."""
    return _workload('sss-nontof-32x8-48x48x24-500sp', 8, 192, 3.2, (48, 48, 24),
                     (0.8, 0.8, 1.07), 500, seed, 8, 32, 1, 1e9, 1000.0)


def config2(seed=2, n_sp=2000):
    """TOF SSS, Vision-like: 760 det x 80 rings (72 x 16 used), 96x96x48 of
    4x4x5.3 mm, 2000 sp, 15 bins of a 214 ps FWHM kernel over +-500 mm."""
    return _workload('tofsss-760x80-72x16used-96x96x48-2000sp-15bins', 80, 760, 0.32,
                     (96, 96, 48), (0.4, 0.4, 0.53), n_sp, seed, 16, 72, 15, 214.0,
                     3335.64)


def small_tof(seed=3, n_sp=40, n_bins=7, shape=(24, 24, 12), n_det=48, n_rings=8, nru=4, ndu=16):
    """Small TOF case for fast parity tests (seconds in the oracle)."""
    vox = (38.4 / shape[0], 38.4 / shape[1], 25.6 / shape[2])
    return _workload(f'small-tof-{n_det}x{n_rings}-{shape[0]}x{shape[1]}x{shape[2]}-{n_sp}sp-{n_bins}bins',
                     n_rings, n_det, 25.6 / n_rings, shape, vox, n_sp, seed, nru, ndu, n_bins,
                     214.0, 3335.64)


def mashed(w, mash):
    """The workload on its scanner with `mash` neighbouring crystals merged
    (Geometry.py:137-141 with Mash > 1: nan-mean geometry and normals, and
    the sinogram coordinates of one transaxial sector of N/mash crystals
    without the ID shift).  The sampled detectors are kept at most N/mash."""
    from dataclasses import replace
    geom, normals = mashDetectors(w.geometry, w.normals, mash)
    n_rings, n_det = geom.shape[0], geom.shape[1]
    lor, sino, _ = SinogramCoordinates(1, 1, 1, 1, n_det, n_rings, ID_Shidt=False)
    return replace(w, name=f'{w.name}-mash{mash}', geometry=geom, normals=normals, lor_coordinates=lor,
                   sinogram_index=sino, n_det_used=min(w.n_det_used, n_det))


def random_instance(seed):
    """One instance of the uneven batch: grid 32x32x16 .. 128x128x64, 300-4000 sp,
    32-96 det x 8-24 rings, 1-21 TOF bins, random ellipse 200-350 mm, an insert of
    mu in {0.029, 0.096, 0.150} and 0-8 spheres of 8-40 mm at contrast 2-8."""
    rng = np.random.default_rng(seed)
    g = int(rng.choice([32, 48, 64, 96, 128]))
    shape = (g, g, max(16, g // 2))
    vox = (38.4 / shape[0], 38.4 / shape[1], 25.6 / shape[2])
    n_det = int(rng.choice([32, 48, 64, 72, 96]))
    n_rings = int(rng.choice([8, 12, 16, 20, 24]))
    n_bins = int(rng.integers(1, 22))
    n_sp = int(rng.integers(300, 4001))
    ax = rng.uniform(10.0, 17.5, size=2)
    ax = (max(ax), min(ax))
    n_sph = int(rng.integers(0, 9))
    diam = tuple(rng.uniform(0.8, 4.0, size=n_sph))
    contrast = float(rng.uniform(2.0, 8.0))
    ins_mu = float(rng.choice([LUNG_MU, WATER_MU, BONE_MU]))
    ins = [(rng.uniform(-4, 4), rng.uniform(-4, 4), rng.uniform(1.0, 3.0), ins_mu)]
    nru = n_rings if n_rings <= 16 else 16
    ndu = n_det if n_det <= 64 else 64
    return _workload(f'random-{seed}', n_rings, n_det, 25.6 / n_rings, shape, vox, n_sp,
                     seed, nru, ndu, n_bins, 214.0, 3335.64, body_axes=ax, lung_radius=0.0,
                     sphere_diameters=diam, sphere_contrast=contrast, sphere_ring=6.0,
                     inserts=ins)
