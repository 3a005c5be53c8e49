"""GPU parity: libsss_hip.so against the oracle on the same seeded inputs.

Bar: bit-exact voxel walks (indices, lengths, TOF samples); scatter sinogram
within 1e-4 max relative error on cells above 1% of the peak for the FP32
path (1e-6 for the FP64 path) and 1e-6 of the peak elsewhere.
"""
import numpy as np
import pytest

from opensss_amd import phantoms
from tests.sss_cases import (WATER, gpu_compact, gpu_dense, gpu_subject, oracle_compact,
                             oracle_dense, rel_err_above)

pytestmark = pytest.mark.gpu

FP32_TOL = 1e-4
FP64_TOL = 1e-6


def _segments(rng, n, bounds, scale=1.6):
    lo, hi = np.array(bounds[:3]), np.array(bounds[3:])
    c, h = (lo + hi) / 2, (hi - lo) / 2
    a = c + rng.uniform(-1, 1, size=(n, 3)) * h * scale
    b = c + rng.uniform(-1, 1, size=(n, 3)) * h * scale
    return np.concatenate([a, b], axis=1)


@pytest.mark.parametrize('grid,bounds', [
    ((8, 8, 6), (-4.0, -4.0, -3.0, 4.0, 4.0, 3.0)),
    ((48, 48, 24), (-19.2, -19.2, -12.84, 19.2, 19.2, 12.84)),
    ((96, 96, 48), (-19.2, -19.2, -12.72, 19.2, 19.2, 12.72)),
])
def test_tracer_bitwise_vs_oracle(oracle, grid, bounds):
    from opensss_amd.raytracing import RayTracing3DTOFBatch
    rng = np.random.default_rng(5)
    segs = _segments(rng, 2000, bounds)
    # special cases: axis parallel, diagonal ties, start on a face, miss
    segs = np.concatenate([segs, np.array([
        [bounds[0] - 1, 0.5, 0.5, bounds[3] + 1, 0.5, 0.5],
        [0.0, 0.0, 0.0, bounds[3], bounds[4], 0.0],
        [0.0, 0.0, 0.1, 30.0, 5.0, 1.0],
        [bounds[0] - 5, bounds[4] + 5, 0, bounds[0] - 5, bounds[4] + 9, 0]])], axis=0)
    counts, L, idx, s = RayTracing3DTOFBatch(grid, bounds, segs)
    for i, seg in enumerate(segs):
        oL, oi, os_ = oracle.trace(grid, bounds, seg)
        n = counts[i]
        assert n == len(oi), i
        assert np.array_equal(idx[i, :n], oi), i
        assert np.array_equal(L[i, :n], oL), i
        assert np.array_equal(s[i, :n], os_), i


def test_single_segment_drop_in(oracle):
    from opensss_amd.raytracing import RayTracing3DTOF
    grid = np.array([10, 10, 6])
    bounds = np.array([-5.0, -5.0, -3.0, 5.0, 5.0, 3.0])
    line = np.array([-4.3, 2.1, -1.0, 12.0, -7.0, 2.5])
    L, idx, s = RayTracing3DTOF(grid, bounds, line)
    oL, oi, os_ = oracle.trace(grid, bounds, line)
    assert idx.dtype == np.int32 and np.array_equal(idx, oi)
    assert np.array_equal(L, oL) and np.array_equal(s, os_)


def test_line_integrals_vs_oracle(oracle):
    from opensss_amd.raytracing import LineIntegrals
    rng = np.random.default_rng(9)
    grid, bounds = (24, 24, 12), (-6.0, -6.0, -3.0, 6.0, 6.0, 3.0)
    mu = rng.uniform(0, 0.1, size=grid).flatten(order='F')
    lam = rng.uniform(0, 2, size=grid).flatten(order='F')
    segs = _segments(rng, 500, bounds)
    n, m, l = LineIntegrals(grid, bounds, mu, lam, segs)
    on, om, ol = oracle.line_integrals(grid, bounds, mu, lam, segs)
    assert np.array_equal(n, on)
    assert np.array_equal(m, om) and np.array_equal(l, ol)


def _small_case_workload():
    from tests.test_oracle_sss import small_case
    c = small_case()
    return phantoms.Workload(
        name='small-case', activity=c['lam'], attenuation=c['mu'], image_size=c['bounds'],
        geometry=c['geom'], normals=c['normals'], lor_coordinates=c['lor'],
        sinogram_index=c['sino'], det_size=c['det_size'], energy_resolution=c['e_res'],
        energy_threshold=c['e_thr'], n_rings_used=c['nru'], n_det_used=c['ndu'],
        tof_resolution=c['tof_res'], tof_range=c['tof_range'], n_bins=c['n_bins'],
        scatter_points=c['pts'], slice_ptr=c['ptr'])


@pytest.mark.parametrize('precision,tol', [('fp32', FP32_TOL), ('fp64', FP64_TOL)])
def test_small_case_dense_vs_oracle(oracle, precision, tol):
    w = _small_case_workload()
    got = gpu_dense(w, precision)
    want = oracle_dense(oracle, w)
    rel, absn = rel_err_above(got, want)
    print(f'small-case {precision}: max rel err (cells > 1% peak) {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= tol and absn <= 1e-6


@pytest.mark.parametrize('precision,tol', [('fp32', FP32_TOL), ('fp64', FP64_TOL)])
def test_small_tof_workload_vs_oracle(oracle, precision, tol):
    w = phantoms.small_tof()
    got = gpu_compact(w, precision)
    want = oracle_compact(oracle, w)
    rel, absn = rel_err_above(got, want)
    print(f'{w.name} {precision}: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= tol and absn <= 1e-6


# (bins, TOF resolution ps, TOF range ps) -> TOF fold of the FP32 and FP64 launches
# (0 one exp2 per window bin, 1 one exp2 with end-bin masks, 2 one exp2 with the
# window centre inside the bins, 4 and 5 the relative sliding windows, 6 the
# sliding window of 2 with the records shared per quad).  The FP64
# path drops bins below 1e-16 of the nearest one (FP32: 1e-9), so its window
# is wider at 214 ps.
TOF_VARIANTS = {
    'h1-15bins': ((15, 214.0, 3335.64), 6, 0),
    'h1-14bins-even': ((14, 214.0, 3335.64), 6, 0),
    'h2-21bins': ((21, 214.0, 3335.64), 4, 0),
    'h2-17bins': ((17, 214.0, 3335.64), 4, 0),
    'h1-8bins': ((8, 214.0, 3335.64), 4, 0),
    'h1-7bins': ((7, 214.0, 3335.64), 4, 0),
    'h1-6bins-wide': ((6, 214.0, 3335.64), 5, 0),
    'h1-2bins': ((2, 214.0, 3335.64), 5, 0),
    'h0-1bin-zero-columns': ((1, 214.0, 3335.64), 0, 0),
    'h1-3bins-narrow': ((3, 214.0, 3335.64), 5, 0),
    'h1-3bins-masked': ((3, 600.0, 2417.0), 4, 1),
}


@pytest.mark.parametrize('case', list(TOF_VARIANTS))
def test_tof_kernel_variants_vs_oracle(oracle, case):
    """15 and 14 bins at 214 ps: one exp2 per sample with a sliding window (odd
    and even bin counts, bin centres at integer / half-integer multiples of W);
    21 and 17 bins: h = 2 (FP32: relative sliding window, one exp2); 8 and 7
    bins: wide bins (beta2 61, 79), the relative window's one-exp2 form; 6, 3
    and 2 bins at 214 ps: bins so wide (beta2 108 to 3e3) that only its
    two-exp2 form keeps to the float range; one bin: zero TOF columns beyond
    +-3508 ps; 3 bins of 1611 ps at 600 ps: offsets far beyond the outer
    bins, so the window keeps masked end bins."""
    import dataclasses
    (n_bins, res, rng), fold32, fold64 = TOF_VARIANTS[case]
    w = phantoms.small_tof(seed=12, n_bins=n_bins, n_sp=30)
    w = dataclasses.replace(w, tof_resolution=res, tof_range=rng)
    want = oracle_compact(oracle, w)
    for precision, fold, tol in (('fp32', fold32, FP32_TOL), ('fp64', fold64, FP64_TOL)):
        s = gpu_subject(w, precision)
        try:
            s.run(w.activity)
            got = s.output_compact()
            assert s.info['tof_fold'] == fold, (precision, s.info['tof_fold'])
        finally:
            s.close()
        rel, absn = rel_err_above(got, want)
        print(f'{w.name} {case} {precision}: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
        # the oracle stores its per-slice sums in float32 as the reference does
        assert rel <= tol and absn <= 1e-6


def test_config1_nontof_full_vs_oracle(oracle):
    w = phantoms.config1()
    got = gpu_compact(w, 'fp32')
    want = oracle_compact(oracle, w)
    rel, absn = rel_err_above(got, want)
    print(f'{w.name}: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= FP32_TOL and absn <= 1e-6


def test_config2_scatter_point_subset_vs_oracle(oracle):
    w = phantoms.config2()
    pts = w.scatter_points[::250]            # 8 points over several slices
    got = gpu_compact(w, 'fp32', points=pts)
    want = oracle_compact(oracle, w, points=pts)
    rel, absn = rel_err_above(got, want)
    print(f'{w.name} subset {len(pts)} sp: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= FP32_TOL and absn <= 1e-6


def test_config2_full_fp32_vs_fp64():
    w = phantoms.config2()
    a = gpu_compact(w, 'fp32')
    b = gpu_compact(w, 'fp64')
    rel, absn = rel_err_above(a, b)
    print(f'{w.name} FP32 vs FP64: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= FP32_TOL and absn <= 1e-6


def test_chunking_does_not_change_results():
    w = phantoms.small_tof(n_sp=60)
    a = gpu_compact(w, 'fp64', n_chunks=1)
    b = gpu_compact(w, 'fp64', n_chunks=5)
    rel, absn = rel_err_above(a, b)
    assert rel <= 1e-12 and absn <= 1e-12


def test_resident_subject_activity_refresh(oracle):
    """Same subject, two activity images: each matches a fresh run, and
    scaling the activity scales the estimate (the counts only see zeros)."""
    w = phantoms.small_tof(seed=8)
    s = gpu_subject(w, 'fp32')
    try:
        s.run(w.activity)
        first = s.output_compact()
        s.set_activity(2.5 * w.activity)
        s.scatter()
        second = s.output_compact()
        lam3 = w.activity * np.random.default_rng(1).uniform(0.9, 1.1, size=w.activity.shape)
        s.set_activity(lam3)
        s.scatter()
        third = s.output_compact()
    finally:
        s.close()
    rel, _ = rel_err_above(second, 2.5 * first)
    assert rel <= 1e-5
    fresh = gpu_compact(w, 'fp32', lam=lam3)
    assert np.array_equal(third, fresh)
    rel, absn = rel_err_above(third, oracle_compact(oracle, w, lam=lam3))
    assert rel <= FP32_TOL and absn <= 1e-6


@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_weights_only_refresh_and_its_fallback(precision):
    """A resident subject given a sequence of activity images: rescaled ones
    keep the zero terms (the fold table's weights are rewritten in place),
    others zero some voxels or bring them back (the table is rebuilt).  Every
    result equals a fresh subject's, bit for bit."""
    w = phantoms.small_tof(seed=9, n_sp=30)
    rng = np.random.default_rng(3)
    lam = w.activity
    holes = lam.copy()
    holes[::3, :, :] = 0.0                      # zero terms appear
    images = [lam, lam * rng.uniform(0.9, 1.1, size=lam.shape), holes,
              holes * 1.7, lam * 0.5, lam]
    s = gpu_subject(w, precision)
    try:
        s.run(images[0])
        for k, img in enumerate(images):
            if k:
                s.set_activity(img)
                s.scatter()
            got = s.output_compact()
            want = gpu_compact(w, precision, lam=img)
            assert np.array_equal(got, want), f'image {k}'
    finally:
        s.close()


def _swap_voxels_on_the_same_paths(oracle, w, n_pairs=4):
    """Voxel pairs (u, v), both active, that the same half paths cross with
    positive length, found by tracing every half path of the subject: zeroing
    u in one image and v in the next keeps every path's count of non-zero
    terms but moves its zero term."""
    from opensss_amd.tables import detectors_used, rings_used, scatter_point_coordinates
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    rings = rings_used(NR, w.n_rings_used)
    dets = detectors_used(N, w.n_det_used)
    xyz = scatter_point_coordinates(w.scatter_points, w.attenuation.shape, w.image_size)
    grid = np.array(w.attenuation.shape)
    lam = w.activity.ravel(order='F')
    visits = {}
    path = 0
    for p in xyz:
        for r in rings:
            for d in dets:
                det = [w.geometry[r, d, 0], w.geometry[r, d, 1], w.geometry[r, 0, 2]]
                L, idx, _ = oracle.trace(grid, w.image_size, np.concatenate([p, det]))
                for li, vi in zip(L, idx - 1):
                    if li > 0 and lam[vi] > 0:
                        visits.setdefault(int(vi), set()).add(path)
                path += 1
    by_set = {}
    for v, paths in visits.items():
        by_set.setdefault(frozenset(paths), []).append(v)
    pairs = [vs[:2] for vs in by_set.values() if len(vs) >= 2]
    return pairs[:n_pairs]


@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_weights_only_refresh_sees_moved_zero_terms(oracle, precision):
    """Images A and B with the same count of non-zero terms on every half
    path but the zero terms in other voxels: the weights-only refresh must
    notice the moved zeros (a fold record whose term became zero) and rebuild
    the fold table; each result equals a fresh subject's, bit for bit."""
    w = phantoms.small_tof(seed=9, n_sp=6)
    pairs = _swap_voxels_on_the_same_paths(oracle, w)
    assert pairs, 'no voxel pair shares its half paths'
    X, Y, Z = w.attenuation.shape
    a = w.activity.copy()
    b = w.activity.copy()
    for u, v in pairs:
        a[np.unravel_index(u, (X, Y, Z), order='F')] = 0.0
        b[np.unravel_index(v, (X, Y, Z), order='F')] = 0.0
    images = [a, a * 1.1, b, b * 0.9]
    s = gpu_subject(w, precision)
    try:
        s.run(images[0])
        for k, img in enumerate(images):
            if k:
                s.set_activity(img)
                s.scatter()
            got = s.output_compact()
            want = gpu_compact(w, precision, lam=img)
            assert np.array_equal(got, want), f'image {k}'
    finally:
        s.close()


@pytest.mark.parametrize('n_bins', [1, 7])
def test_volume_upload_orders_agree(n_bins):
    """Volumes sent in C order (transposed on the device, through a 32 x 32
    tile per workgroup) give the results of volumes flattened x fastest on
    the host, bit for bit; grid sizes not multiples of 32, two activity
    images in a row (the second reuses the device staging slot)."""
    from opensss_amd.sss import ScatterSubject
    w = phantoms.small_tof(seed=4, n_bins=n_bins, shape=(40, 40, 21))
    lam2 = w.activity * np.random.default_rng(2).uniform(0.5, 1.5, size=w.activity.shape)
    outs = {}
    for order in ('C', 'F'):
        s = ScatterSubject(w.attenuation, w.image_size, w.geometry, w.normals, w.det_size, WATER,
                           w.energy_resolution, w.energy_threshold, w.n_rings_used, w.n_det_used,
                           w.tof_resolution, w.tof_range, w.n_bins, scatter_points=w.scatter_points,
                           upload_order=order)
        try:
            s.run(w.activity)
            first = s.output_compact()
            s.set_activity(lam2)
            s.scatter()
            outs[order] = (first, s.output_compact())
        finally:
            s.close()
    assert np.array_equal(outs['C'][0], outs['F'][0])
    assert np.array_equal(outs['C'][1], outs['F'][1])
    assert not np.array_equal(outs['C'][0], outs['C'][1])


def test_drop_in_strided_matches_oracle(oracle, tmp_path):
    from opensss_amd.sss import SingleScatterSimulationTOF
    from opensss_amd.tables import strided_scatter_points
    w = phantoms.small_tof(seed=4)
    step = np.array([5, 4, 3])
    out = SingleScatterSimulationTOF(w.activity, w.attenuation, w.image_size, w.geometry,
                                     w.lor_coordinates, w.normals, w.det_size, WATER,
                                     w.energy_resolution, w.energy_threshold, w.n_rings_used,
                                     w.n_det_used, step, w.tof_resolution, w.tof_range, w.n_bins,
                                     str(tmp_path))
    N = w.geometry.shape[1]
    assert out.dtype == np.float32
    assert out.shape == (w.n_bins, N + 1, N // 2, w.n_rings_used ** 2)
    assert (tmp_path / 'Scatters_SSS.npz').exists()
    want = oracle_dense(oracle, w, points=strided_scatter_points(w.attenuation, step))
    rel, absn = rel_err_above(out, want)
    assert rel <= FP32_TOL and absn <= 1e-6


def test_invalid_inputs_raise():
    from opensss_amd import _native
    from opensss_amd.sss import ScatterSubject
    w = phantoms.small_tof()
    with pytest.raises(ValueError):
        ScatterSubject(w.attenuation[:, :-1], w.image_size, w.geometry, w.normals, w.det_size,
                       WATER, 0.1, 435.0, 4, 16, 214.0, 3335.64, 7)
    with pytest.raises(_native.SssError):
        ScatterSubject(w.attenuation, w.image_size, w.geometry, w.normals, w.det_size, WATER,
                       0.1, 435.0, 99, 16, 214.0, 3335.64, 7)
    with pytest.raises(_native.SssError):
        ScatterSubject(w.attenuation, w.image_size, w.geometry, w.normals, w.det_size, WATER,
                       0.1, 435.0, 4, 16, 214.0, 3335.64, 40)
    s = gpu_subject(w)
    try:
        with pytest.raises(_native.SssError):
            s.scatter()          # no activity yet
    finally:
        s.close()


def test_work_stats_count_the_reference_work(oracle):
    """Pairs kept by the reference's filters and the voxel samples of their
    active half paths (the algorithmic work count), against a brute-force
    count over the oracle's traces: the reference's energy index
    (acos/degrees/cos round trip) and efficiency test, Counts' activity tests
    and every voxel of each active half path."""
    import math
    from opensss_amd.tables import (energy_tables, scatter_point_coordinates, detectors_used,
                                    rings_used)
    w = phantoms.small_tof(seed=5, n_sp=12)
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    rings = rings_used(NR, w.n_rings_used)
    dets = detectors_used(N, w.n_det_used)
    _, eff, ref = energy_tables(WATER, w.energy_resolution, w.energy_threshold)
    xyz = scatter_point_coordinates(w.scatter_points, w.attenuation.shape, w.image_size)
    lam = w.activity.ravel(order='F')
    mu = w.attenuation.ravel(order='F')
    grid = np.array(w.attenuation.shape)
    NRU, NDU = len(rings), len(dets)
    det = np.array([[[w.geometry[r, d, 0], w.geometry[r, d, 1], w.geometry[r, 0, 2]] for d in dets]
                    for r in rings])
    pairs = samples = 0
    for p in xyz:
        n = np.zeros((NRU, NDU), int)
        act = np.zeros((NRU, NDU))
        att = np.zeros((NRU, NDU))
        dist = np.zeros((NRU, NDU))
        for a in range(NRU):
            for b in range(NDU):
                L, idx, _ = oracle.trace(grid, w.image_size, np.concatenate([p, det[a, b]]))
                n[a, b] = len(L)
                act[a, b] = np.sum(lam[idx - 1] * L)
                att[a, b] = math.exp(-np.sum(mu[idx - 1] * L))
                dist[a, b] = math.sqrt(np.sum((det[a, b] - p) ** 2))
        for r1 in range(NRU):
            for i1 in range(NDU):
                for r2 in range(NRU):
                    for i2 in range(NDU):
                        if i1 > i2 or (i1 == i2 and r1 == r2):
                            continue
                        if att[r1, i1] == 0 or att[r2, i2] == 0 or (act[r1, i1] == 0 and act[r2, i2] == 0):
                            continue
                        s1q, s2q = dist[r1, i1] ** 2, dist[r2, i2] ** 2
                        lor = np.sum((det[r1, i1] - det[r2, i2]) ** 2)
                        theta = 180 - math.degrees(math.acos(min(1.0, max(-1.0, (s1q + s2q - lor)
                                                                           / (2 * math.sqrt(s1q * s2q))))))
                        e = math.floor(511.0 / (1 + (1 - math.cos(math.radians(theta)))) + 0.5)
                        k = min(max(int(e * 2) - 1, 0), len(eff) - 1)
                        if eff[k] * eff[ref] == 0:
                            continue
                        pairs += 1
                        samples += (n[r1, i1] if act[r1, i1] != 0 else 0) + (n[r2, i2] if act[r2, i2] != 0 else 0)
    s = gpu_subject(w, 'fp32')
    try:
        s.run(w.activity)
        got_pairs, got_samples = s.work_stats()
    finally:
        s.close()
    assert (got_pairs, got_samples) == (pairs, samples)
    assert 0 < pairs < w.n_rings_used ** 2 * w.n_det_used ** 2 * len(xyz)


@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_voxel_test_of_resident_images(precision):
    """The voxel test that lets a rescaled image skip the per-sample checks:
    images that keep the zero voxels take it; an image with a voxel far
    below the safe level (a term that is zero in float32, not in float64), a
    negative voxel, or a negative zero must fall back to the exact checks.
    Every result equals a fresh subject's, bit for bit."""
    w = phantoms.small_tof(seed=9, n_sp=24)
    lam = w.activity
    X, Y, Z = lam.shape
    vox = np.argwhere(lam > 0)[len(np.argwhere(lam > 0)) // 2]
    tiny = lam.copy()
    tiny[tuple(vox)] = 1e-300
    neg = lam.copy()
    neg[tuple(vox)] = -0.5
    negzero = np.where(lam == 0.0, -0.0, lam)
    images = [lam, lam * 1.05, lam * 0.7, tiny, tiny * 1.3, lam * 0.9, neg, lam, negzero, lam * 1.2]
    s = gpu_subject(w, precision)
    try:
        s.run(images[0])
        for k, img in enumerate(images):
            if k:
                s.set_activity(img)
                s.scatter()
            got = s.output_compact()
            want = gpu_compact(w, precision, lam=img)
            assert np.array_equal(got, want), f'image {k}'
    finally:
        s.close()
