"""Host-side helpers against their plain-loop restatements and the invariants
the reference's integer tables must satisfy (Sinogram.py, DataReduction.py,
Geometry.py, the table set-up of SingleScatterSimulationTOF.py:82-171)."""
import math
import os

import numpy as np
import pytest

from opensss_amd import data_reduction as dr
from opensss_amd import geometry, sinogram, tables


@pytest.mark.parametrize('args', [(32, 1, 1, 1, 1, 8), (8, 1, 1, 2, 3, 4), (6, 2, 1, 1, 5, 3),
                                  (4, 1, 2, 2, 4, 2, 1), (10, 1, 1, 1, 3, 2, 0, False)])
def test_sinogram_coordinates_match_loops(oracle, args):
    got = sinogram.SinogramCoordinates(*args)
    want = oracle.sinogram_coordinates_loops(*args)
    assert got[0].dtype == np.int16 and got[1].dtype == np.int16
    np.testing.assert_array_equal(got[0], want[0])
    np.testing.assert_array_equal(got[1], want[1])
    assert got[2] == want[2]


@pytest.mark.parametrize('n', [8, 32, 72, 380, 760])
def test_lor_coordinates_bijection(n):
    lor, _, _ = sinogram.SinogramCoordinates(n, 1, 1, 1, 1, 2)
    np.testing.assert_array_equal(lor, lor.transpose(1, 0, 2))
    d1, d2 = np.triu_indices(n)
    cells = lor[d1, d2, 1].astype(np.int64) * (n // 2) + lor[d1, d2, 0]
    assert lor[..., 0].min() >= 0 and lor[..., 0].max() < n // 2
    assert lor[..., 1].min() >= 0 and lor[..., 1].max() <= n
    assert len(np.unique(cells)) == len(cells) == (n + 1) * (n // 2)


@pytest.mark.parametrize('nr', [1, 2, 8, 16, 80])
def test_sinogram_index_is_permutation(nr):
    _, sino, _ = sinogram.SinogramCoordinates(4, 1, 1, 1, 1, nr)
    assert sorted(sino.ravel().tolist()) == list(range(nr * nr))
    assert all(sino[r, r] == r for r in range(nr))


def test_span_lut_vision_segment0():
    """Span 19, MRD 9 on 80 rings: only segment 0 (2*80-1 = 159 planes)."""
    planes, ring_lut, lut, counts = dr.SinogramLUT(80, 19, 9)
    assert planes.shape == (159, 2)
    assert planes.dtype == np.float32
    # every allowed ring pair lands in a plane whose mean ring sum is its own ring sum
    for r1 in range(80):
        for r2 in range(80):
            p = lut[r1, r2]
            if abs(r1 - r2) > 9:
                assert p == -1
                continue
            assert abs((planes[p, 0] + planes[p, 1]) - (r1 + r2)) < 1e-4
    assert counts.sum() == sum(1 for a in range(80) for b in range(80) if abs(a - b) <= 9)


def test_span_lut_full_mrd_plane_count_and_symmetry():
    planes, _, lut, counts = dr.SinogramLUT(16, 5, None)
    assert int(counts.sum()) == 16 * 16
    # mirrored ring pairs (r2 - r1 = -d) live in a different plane than +d except d = 0
    for r1 in range(16):
        for r2 in range(16):
            if abs(r1 - r2) >= 3:
                assert lut[r1, r2] != lut[r2, r1]
    assert planes.shape[0] == int(np.count_nonzero(counts))


def _michelogram_planes(nr, span, mrd):
    """Span plane of every ring pair (0-based), built as the usual
    Michelogram order and nothing else: segment 0 holds |r2 - r1| < (span+1)/2,
    segment s >= 1 holds |r2 - r1| in [s span - (span-1)/2, s span + (span-1)/2];
    segment 0 first, then per segment its r2 < r1 block before its r2 > r1
    block; inside a block one plane per ring sum r1 + r2, in increasing sum.
    Ring pairs beyond mrd have no plane (-1)."""
    half = (span + 1) // 2
    lut = np.full((nr, nr), -1, dtype=np.int64)
    blocks = [(0, 0)]                  # (segment, sign) in plane order
    s = 1
    while s * span - (span - 1) // 2 <= nr - 1:
        blocks += [(s, -1), (s, +1)]
        s += 1
    base = 0
    for seg, sign in blocks:
        dmin = 0 if seg == 0 else seg * span - (span - 1) // 2
        dmax = half - 1 if seg == 0 else seg * span + (span - 1) // 2
        lo = dmin if seg else 0
        for r1 in range(nr):
            for r2 in range(nr):
                d = r2 - r1
                if mrd is not None and abs(d) > mrd:
                    continue
                if not dmin <= abs(d) <= dmax or (seg and (d > 0) != (sign > 0)):
                    continue
                lut[r1, r2] = base + (r1 + r2 - lo)
        base += 2 * (nr - lo) - 1     # distinct ring sums of the block
    return lut


@pytest.mark.parametrize('nr,span,mrd', [(16, 3, None), (16, 5, None), (12, 3, 7), (24, 5, 13),
                                         (9, 3, None), (80, 19, 9)])
def test_span_lut_matches_michelogram_order(nr, span, mrd):
    """Every ring pair's plane, in every segment, against an independent
    enumeration of the segment/block order; the averaged rings are the means
    of each plane's ring pairs."""
    planes, _, lut, counts = dr.SinogramLUT(nr, span, mrd)
    want = _michelogram_planes(nr, span, mrd)
    assert np.array_equal(lut.astype(np.int64), want)
    # the averaged rings list only the planes some ring pair reaches (an mrd
    # inside a segment leaves every other ring sum of its last block empty),
    # while the LUT keeps the full numbering, as the reference's does
    full = np.nonzero(counts > 0)[0]
    assert planes.shape[0] == len(full)
    for k, p in enumerate(full):
        r1, r2 = np.nonzero(want == p)
        assert counts[p] == len(r1) > 0
        assert abs(planes[k, 0] - r2.mean()) < 1e-5 and abs(planes[k, 1] - r1.mean()) < 1e-5
    assert int(counts.sum()) == int(np.count_nonzero(want >= 0))


def test_sum_lor_counts():
    geom, _ = geometry.cylinder_geometry(4, 8, 10.0, 1.0)
    _, sino, _ = sinogram.SinogramCoordinates(8, 1, 1, 1, 1, 4)
    c = dr.SumLORsCounts(geom, np.array([]), sino, 2, 1, 0)
    assert c.shape == (9, 4, 16)
    assert np.all(c == 4.0)


def test_expand_and_mash():
    geom, normals = geometry.cylinder_geometry(3, 8, 10.0, 1.0)
    g2, n2 = geometry.expandRings(geom, normals)
    assert g2.shape == (5, 8, 3)
    np.testing.assert_allclose(g2[1], (geom[0] + geom[1]) / 2)
    np.testing.assert_array_equal(n2[4], normals[0])
    gm, nm = geometry.mashDetectors(geom, normals, 2)
    np.testing.assert_allclose(gm[:, 1], geom[:, 2:4].mean(axis=1))
    with pytest.raises(Exception):
        geometry.mashDetectors(geom, normals, 3)


def test_casto_r_geometry_reader(tmp_path):
    geom, normals = geometry.cylinder_geometry(4, 12, 41.0, 0.4)
    # CASToR header order: the transaxial "number of rsectors" line comes first,
    # so the reference's first-match regex picks it and not the axial one
    hdr = ('number of elements: 48\nnumber of rsectors: 12\nnumber of rsectors axial: 1\n'
           'number of modules axial: 2\nnumber of crystals axial: 2\ncrystals size axial: 4.0\n'
           'number of modules transaxial: 1\nnumber of crystals transaxial: 1\n'
           'crystals size transaxial: 4.0\n')
    (tmp_path / 'toy.geom').write_text(hdr)
    lut = np.concatenate([geom.reshape(-1, 3) * 10, normals.reshape(-1, 3) * 10], axis=1)
    lut.astype(np.float32).tofile(tmp_path / 'toy.glut')
    out = geometry.ImportCASToRGeometry('toy', str(tmp_path))
    np.testing.assert_allclose(out[0], geom, rtol=1e-6, atol=1e-5)
    assert out[2:8] == (1, 12, 2, 1, 2, 1)
    res = geometry.PrepareGeometryAndSinograms('toy', 1, 1, 3, str(tmp_path))
    assert len(res) == 10 and res[3].shape == (12, 12, 2)


def test_rings_and_detectors_used():
    assert tables.rings_used(80, 16).tolist() == [int(math.floor(x + 0.5)) for x in np.linspace(0, 79, 16)]
    d = tables.detectors_used(760, 72)
    assert d[0] == 0 and np.all(np.diff(d) == 10) and d[-1] == 710
    d = tables.detectors_used(32, 32)
    assert d.tolist() == list(range(32))


def test_energy_tables_closed_form():
    tab = np.load(os.path.join(os.path.dirname(os.path.dirname(__file__)), 'opensss_amd', 'data',
                               'attenuation_water.npy'))
    ratio, eff, ref = tables.energy_tables(tab, 0.10, 435.0)
    assert ref == 1021 and tab[ref, 0] == 511.0
    assert ratio[ref] == 1.0
    assert abs(eff[np.where(tab[:, 0] == 435.0)[0][0]] - 0.5) < 1e-15
    assert eff[np.where(tab[:, 0] == 200.0)[0][0]] == 0.0   # erf saturates: the skip rule
    assert np.all(np.diff(eff) >= 0)


@pytest.mark.parametrize('B,R,rng', [(15, 214.0, 3335.64), (7, 300.0, 1000.0), (3, 214.0, 3335.64),
                                     (21, 214.0, 3335.64), (1, 1e9, 1000.0), (4, 40.0, 500.0)])
def test_tof_table_partition_and_symmetry(B, R, rng):
    W = 2 * rng / B
    T = tables.tof_table(B, W, R)
    s = T.sum(axis=0)
    nz = s != 0
    np.testing.assert_allclose(s[nz], 1.0, rtol=1e-13)
    np.testing.assert_allclose(T, T[::-1, ::-1], rtol=1e-12, atol=1e-300)


@pytest.mark.parametrize('B,R,rng', [(15, 214.0, 3335.64), (7, 300.0, 1000.0), (3, 214.0, 3335.64),
                                     (21, 214.0, 3335.64), (4, 150.0, 500.0)])
def test_windowed_kernel_reproduces_table(B, R, rng):
    """The device's window form d*kappa*u - d^2*beta around the nearest bin,
    evaluated here in float64, equals the table to the window tolerance."""
    k = tables.tof_kernel(B, rng, R)
    T = tables.tof_table(B, k.bin_width, R)
    t = np.arange(-4000, 4001, dtype=np.float64)
    mid = (B - 1) / 2
    j0 = np.clip(np.rint(t / k.bin_width + mid), 0, B - 1)
    u = t - (j0 - mid) * k.bin_width
    W = np.zeros_like(T)
    for d in range(-k.half_window, k.half_window + 1):
        j = j0 + d
        ok = (j >= 0) & (j < B)
        e = np.exp(d * k.kappa * u - d * d * k.beta)
        W[j[ok].astype(int), np.nonzero(ok)[0]] = e[ok]
    W /= W.sum(axis=0)
    inside = (t >= k.t_lo) & (t <= k.t_hi)
    assert np.abs(W[:, inside] - T[:, inside]).max() < 1e-8


def test_scatter_point_coordinates():
    pts = np.array([[0, 0, 0], [47, 47, 23], [10, 3, 5]])
    isz = np.array([-19.2, -19.2, -12.84, 19.2, 19.2, 12.84])
    xyz = tables.scatter_point_coordinates(pts, (48, 48, 24), isz)
    np.testing.assert_allclose(xyz[0], [-19.2 + 0.8, -19.2 + 0.8, -12.84 + 0.535], atol=1e-12)
    np.testing.assert_allclose(xyz[1, :2], [19.2, 19.2], atol=1e-12)


def test_strided_points_order():
    mu = np.zeros((4, 4, 3))
    mu[1, 2, 0] = mu[3, 0, 0] = mu[0, 1, 2] = 1
    pts = tables.strided_scatter_points(mu, (1, 1, 1))
    assert pts.tolist() == [[3, 0, 0], [1, 2, 0], [0, 1, 2]]
    pts = tables.strided_scatter_points(mu, (2, 2, 2))
    assert pts.tolist() == [[0, 1, 2]] or pts.tolist() == []
