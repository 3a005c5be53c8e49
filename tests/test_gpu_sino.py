"""GPU parity of the sinogram stages after the scatter estimate against the
oracle's restatements of the reference loops (small scanner: 48 detectors x
8 rings, 24x24x12 maps).

Bar: interpolation and tail mask bit-exact (they copy values / test exact
zeros); tail fit scale within 1e-5 of the reference arithmetic (float32
images and lstsq, whose own rounding is ~1e-6) and images within 1e-5.
"""
import numpy as np
import pytest

from opensss_amd import phantoms
from opensss_amd.data_reduction import SinogramLUT
from opensss_amd.geometry import expandRings
from opensss_amd.tables import detectors_used, rings_used

pytestmark = pytest.mark.gpu


def _case():
    return phantoms.small_tof(seed=21)


def _compact_to_dense(compact, lor, dets, n_det):
    """Compact (B, NRU^2, cells) -> reference dense (B, N+1, N//2, NRU^2)."""
    B, P, _ = compact.shape
    i1, i2 = np.triu_indices(len(dets))
    dense = np.zeros((B, n_det + 1, n_det // 2, P), dtype=np.float32)
    rad = lor[dets[i1], dets[i2], 1]
    ang = lor[dets[i1], dets[i2], 0]
    dense[:, rad, ang, :] = np.transpose(compact, (0, 2, 1))
    return dense


def _random_compact(rng, B, nru, ndu):
    c = rng.normal(1.0, 1.0, size=(B, nru * nru, ndu * (ndu + 1) // 2)).astype(np.float32)
    # same-ring diagonal cells are never computed by the scatter estimate
    diag = np.array([i * ndu - i * (i - 1) // 2 for i in range(ndu)])
    for r in range(nru):
        c[:, r + r * nru, diag] = 0
    return c


@pytest.mark.parametrize('span', [False, True])
def test_interpolation_matches_reference_loops(oracle, span):
    from opensss_amd.interpolation import Interpolator
    w = _case()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    rings = rings_used(NR, w.n_rings_used)
    dets = detectors_used(N, w.n_det_used)
    rng = np.random.default_rng(3)
    compact = _random_compact(rng, 3, w.n_rings_used, w.n_det_used)
    dense = _compact_to_dense(compact, w.lor_coordinates, dets, N)
    if span:
        msi = SinogramLUT(NR, 3, 4)[0]
        ip = Interpolator(NR, w.n_rings_used, N, w.n_det_used, w.lor_coordinates, msi, True)
        want = oracle.interpolate_bins(dense, rings, dets, w.lor_coordinates, MashedSinogramIndices=msi)
    else:
        ip = Interpolator(NR, w.n_rings_used, N, w.n_det_used, w.lor_coordinates, w.sinogram_index, False)
        want = oracle.interpolate_bins(dense, rings, dets, w.lor_coordinates, SinogramIndex=w.sinogram_index)
    bins, total = ip.run(compact)
    assert bins.shape == want.shape
    assert np.array_equal(bins, want)
    assert np.allclose(total, want.sum(axis=0), rtol=1e-6, atol=0)


def test_interpolate_scatters_drop_in_writes_bins(oracle, tmp_path):
    from opensss_amd.interpolation import InterpolateScatters
    w = _case()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    dets = detectors_used(N, w.n_det_used)
    compact = _random_compact(np.random.default_rng(4), 2, w.n_rings_used, w.n_det_used)
    dense = _compact_to_dense(compact, w.lor_coordinates, dets, N)
    total = InterpolateScatters(dense, NR, w.n_rings_used, N, w.n_det_used, w.lor_coordinates,
                                w.sinogram_index, str(tmp_path), False)
    want = oracle.interpolate_bins(dense, rings_used(NR, w.n_rings_used), dets, w.lor_coordinates,
                                   SinogramIndex=w.sinogram_index)
    assert total.dtype == np.float32 and total.shape == want.shape[1:]
    for b in range(2):
        assert np.array_equal(np.load(tmp_path / f'SSS_mashed_bin{b}.npz')['arr_0'], want[b])
    assert (tmp_path / 'Simulation_time.txt').exists()


@pytest.mark.parametrize('use_att,accel,span', [(True, 1, False), (False, 1, False), (True, 2, False),
                                                (True, 1, True), (False, 3, True)])
def test_tail_mask_matches_reference_loops(oracle, use_att, accel, span):
    from opensss_amd.tails_mask import MaskGenerator, MaskGeneratorMashing
    w = _case()
    NR = w.geometry.shape[0]
    if span:
        geom, _ = expandRings(w.geometry, w.normals)
        msi = SinogramLUT(NR, 3, 4)[0]
        got = MaskGeneratorMashing(w.activity, w.attenuation, w.image_size, geom, w.lor_coordinates, msi,
                                   use_att, accel)
        want = oracle.tail_mask(w.activity, w.attenuation, w.image_size, geom, w.lor_coordinates, use_att,
                                accel, MashedSinogramIndices=msi)
    else:
        got = MaskGenerator(w.activity, w.attenuation, w.image_size, w.geometry, w.lor_coordinates,
                            w.sinogram_index, use_att, accel)
        want = oracle.tail_mask(w.activity, w.attenuation, w.image_size, w.geometry, w.lor_coordinates,
                                use_att, accel, SinogramIndex=w.sinogram_index)
    assert got.dtype == np.uint8 and got.shape == want.shape
    assert np.array_equal(got, want)
    assert 0 < got.sum() < got.size


def _fit_inputs(w, n_planes, seed):
    rng = np.random.default_rng(seed)
    N = w.geometry.shape[1]
    shape = (N + 1, N // 2, n_planes)
    scatter = rng.uniform(0.5, 2.0, size=shape).astype(np.float32)
    randoms = rng.uniform(0.0, 0.3, size=shape).astype(np.float32)
    prompts = (1.7 * scatter + randoms + rng.normal(0, 0.05, size=shape)).astype(np.float32)
    mask = (rng.uniform(size=shape) < 0.6).astype(np.float32)
    return prompts, scatter, randoms, mask


@pytest.mark.parametrize('accel', [1, 2])
def test_tail_fit_nospan_matches_reference_loops(oracle, accel):
    from opensss_amd.backprojection import BackProjectionScatterRandFit, nospan_entries, tail_fit
    w = _case()
    NR = w.geometry.shape[0]
    P, S, R, M = _fit_inputs(w, NR * NR, 5)
    grid, bounds = np.array([12, 12, 6]), w.image_size
    order = w.sinogram_index[:NR, :NR].T
    want, imgs = oracle.tail_fit(P, S, R, M, grid, bounds, w.geometry, w.lor_coordinates, accel,
                                 SinogramOrder=order)
    got = BackProjectionScatterRandFit(P, S, R, M, grid, bounds, w.geometry, w.lor_coordinates, order, accel)
    assert got.dtype == np.float32 and got.shape == (1,)
    assert abs(got[0] - want[0]) <= 1e-5 * abs(want[0])
    _, _, _, gi = tail_fit(P, S, R, M, grid, bounds, w.geometry, w.lor_coordinates,
                           nospan_entries(order, NR), accel, False, return_images=True)
    for k, name in enumerate(['prompt', 'scatter', 'random', 'sens']):
        ref = imgs[name].reshape(tuple(grid), order='F')
        assert np.allclose(gi[k], ref, rtol=1e-5, atol=1e-6 * np.abs(ref).max()), name


def test_tail_fit_span_matches_reference_loops(oracle):
    from opensss_amd.backprojection import BackProjectionScatterRandFitMashing
    w = _case()
    NR = w.geometry.shape[0]
    geom, _ = expandRings(w.geometry, w.normals)
    msi = SinogramLUT(NR, 3, 4)[0]
    P, S, R, M = _fit_inputs(w, msi.shape[0], 6)
    grid, bounds = np.array([12, 12, 6]), w.image_size
    want, _ = oracle.tail_fit(P, S, R, M, grid, bounds, geom, w.lor_coordinates, 2,
                              MashedSinogramIndices=msi)
    got = BackProjectionScatterRandFitMashing(P, S, R, M, grid, bounds, geom, w.lor_coordinates, 2, msi)
    assert abs(got[0] - want[0]) <= 1e-5 * abs(want[0])


def test_tail_fit_recovers_exact_scale():
    """Prompts = k S + R on every LOR: the backprojected tails are k S_bp, so
    the fitted scale is k up to float64 rounding."""
    from opensss_amd.backprojection import nospan_entries, tail_fit
    w = _case()
    NR = w.geometry.shape[0]
    _, S, R, M = _fit_inputs(w, NR * NR, 7)
    k = 2.375
    # rounding k S + R to float32 perturbs each prompt by <= 6e-8 relative
    P = (np.float64(k) * S + R).astype(np.float32)
    scale, sxy, sxx = tail_fit(P, S, R, M, np.array([12, 12, 6]), w.image_size, w.geometry,
                               w.lor_coordinates, nospan_entries(w.sinogram_index[:NR, :NR].T, NR), 1, False)
    assert sxx > 0
    assert abs(scale - k) <= 1e-6 * k


def test_scale_scatters_to_prompts_drop_in(oracle, tmp_path):
    """File-based driver (Scaling.py:27-121): two TOF bins, randoms and
    normalisation files present, per-bin scale against the oracle fit."""
    from opensss_amd.scaling import ScaleScattersToPrompts
    w = _case()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    rng = np.random.default_rng(8)
    shape = (N + 1, N // 2, NR * NR)
    scat = [rng.uniform(0.5, 2.0, size=shape).astype(np.float32) for _ in range(2)]
    randoms = rng.uniform(0.0, 0.3, size=shape).astype(np.float32)
    norm = rng.uniform(0.8, 1.2, size=shape).astype(np.float32)
    lor_counts = np.float32(2.0)
    prompts = np.stack([(1.3 + 0.4 * b) * scat[b] + randoms * norm for b in range(2)], axis=-1) * lor_counts
    mask = (rng.uniform(size=shape) < 0.5).astype(np.uint8)
    np.savez_compressed(tmp_path / 'PromptsSinogram.npz', prompts.astype(np.float32))
    np.savez_compressed(tmp_path / 'RandomsSinogram.npz', randoms)
    np.savez_compressed(tmp_path / 'NormSinogram.npz', norm)
    np.save(tmp_path / 'AttenuationMask.npy', mask)
    for b in range(2):
        np.savez_compressed(tmp_path / f'SSS_mashed_bin{b}.npz', scat[b])
    grid = np.array([12, 12, 6])
    got = ScaleScattersToPrompts(w.sinogram_index, w.lor_coordinates, str(tmp_path), grid, w.image_size,
                                 w.geometry, np.zeros((0,)), 2, False, lor_counts)
    assert got.dtype == np.float32 and got.shape == (2,)
    assert np.array_equal(np.load(tmp_path / 'ScaleFactors.npy'), got)
    for b in range(2):
        want, _ = oracle.tail_fit(prompts[..., b] / lor_counts, scat[b], randoms * norm,
                                  np.float32(mask >= 1), grid, w.image_size, w.geometry, w.lor_coordinates, 2,
                                  SinogramOrder=w.sinogram_index[:NR, :NR].T)
        assert abs(got[b] - want[0]) <= 1e-5 * abs(want[0])


@pytest.mark.parametrize('span', [False, True])
def test_tail_fit_is_deterministic(span):
    """The fit's images are 64-bit fixed-point sums of integer atomics and its
    normal equations are reduced in a fixed order: two runs give the same
    bits (scale, sums and images), whatever order the LORs land in."""
    from opensss_amd.backprojection import nospan_entries, span_entries, tail_fit
    w = _case()
    NR = w.geometry.shape[0]
    if span:
        geom, _ = expandRings(w.geometry, w.normals)
        msi = SinogramLUT(NR, 3, 4)[0]
        entries, n_planes, lw = span_entries(msi), msi.shape[0], True
    else:
        geom = w.geometry
        entries, n_planes, lw = nospan_entries(w.sinogram_index[:NR, :NR].T, NR), NR * NR, False
    P, S, R, M = _fit_inputs(w, n_planes, 11)
    runs = [tail_fit(P, S, R, M, np.array([24, 24, 12]), w.image_size, geom, w.lor_coordinates, entries, 1, lw,
                     return_images=True) for _ in range(3)]
    for r in runs[1:]:
        assert r[0] == runs[0][0] and r[1] == runs[0][1] and r[2] == runs[0][2]
        assert np.array_equal(r[3], runs[0][3])
    assert runs[0][2] > 0


def _scaling_inputs(w, span, n_bins, seed, u16):
    """Per-bin scatter, prompts (N+1, N/2, P, B), randoms, norm, uint8 mask
    (values 0..2) and the span case's per-plane LOR counts (4 and 8, mash 2)."""
    from opensss_amd.data_reduction import SumLORsCounts
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    if span:
        msi, _, lut, counts = SinogramLUT(NR, 3, 4)
        P = msi.shape[0]
        lor_counts = SumLORsCounts(w.geometry, lut, w.sinogram_index, 2, 3, counts.shape[0])[:, :, :P]
        sino = msi
    else:
        P = NR * NR
        lor_counts = SumLORsCounts(w.geometry, None, w.sinogram_index, 2, 1, P)
        sino = w.sinogram_index
    rng = np.random.default_rng(seed)
    shape = (N + 1, N // 2, P)
    S = rng.uniform(0.5, 2.0, size=(n_bins,) + shape).astype(np.float32)
    if n_bins > 3:
        S[3] = 0.0                                  # a bin with no scatter: its sums are 0
    S[:, rng.uniform(size=shape) < 0.2] = 0.0       # LORs skipped in every bin
    R = rng.uniform(0.0, 0.3, size=shape).astype(np.float32)
    Nm = rng.uniform(0.8, 1.2, size=shape).astype(np.float32)
    M = rng.integers(0, 3, size=shape).astype(np.uint8)
    if u16:
        prompts = rng.integers(0, 60, size=shape + (n_bins,)).astype(np.uint16)
    else:
        prompts = rng.uniform(0.0, 40.0, size=shape + (n_bins,)).astype(np.float32)
    return sino, S, prompts, R, Nm, M, lor_counts


@pytest.mark.parametrize('span,u16,counts_kind', [(False, True, 'plane'), (False, False, 'cell'),
                                                   (True, True, 'cell'), (True, False, 'plane'),
                                                   (True, True, None)])
def test_scale_to_prompts_is_the_per_bin_fit_bit_for_bit(span, u16, counts_kind):
    """sss_scale_to_prompts over 11 bins (two groups of 8, the second partial)
    against sss_tail_fit per bin on the same sinograms with the prompts
    divided by the LOR counts in numpy float32, the randoms multiplied by
    the normalisation and the mask thresholded beforehand: the same sums,
    bit for bit."""
    from opensss_amd.backprojection import nospan_entries, span_entries, tail_fit
    from opensss_amd.scaling import scale_to_prompts
    w = _case()
    NR = w.geometry.shape[0]
    n_bins = 11
    sino, S, prompts, R, Nm, M, lor_counts = _scaling_inputs(w, span, n_bins, 5, u16)
    if span:
        geom, _ = expandRings(w.geometry, w.normals)
        entries = span_entries(sino)
    else:
        geom = w.geometry
        entries = nospan_entries(w.sinogram_index[:NR, :NR].T, NR)
    counts_arg = None if counts_kind is None else (lor_counts[0, 0, :] if counts_kind == 'plane' else lor_counts)
    grid = np.array([24, 24, 12])
    factors, sums = scale_to_prompts(prompts, S, R, Nm, M, counts_arg, sino, w.lor_coordinates, grid,
                                     w.image_size, w.geometry, geom, 2, span, return_sums=True)
    Mf = np.float32(M >= 1)
    for b in range(n_bins):
        pb = prompts[..., b] / (np.float32(1) if counts_kind is None else lor_counts)
        scale, sxy, sxx = tail_fit(pb.astype(np.float32), S[b], R * Nm, Mf, grid, w.image_size, geom,
                                   w.lor_coordinates, entries, 2, span)
        assert sums[b, 0] == sxy and sums[b, 1] == sxx, b
        assert factors[b] == np.float32(scale)
    assert np.all(sums[3] == 0) and factors[3] == 0
    assert np.all(sums[np.arange(n_bins) != 3, 1] > 0)


def test_scale_scatters_to_prompts_recovers_k_and_matches_oracle(oracle, tmp_path):
    """File-based drop-in on a span case with float32 prompts =
    (k_b S_b + R Norm) LORCounts: every bin recovers k_b and agrees with the
    oracle's scaling loop to 1e-5."""
    from opensss_amd.scaling import ScaleScattersToPrompts
    w = _case()
    sino, S, _, R, Nm, M, lor_counts = _scaling_inputs(w, True, 3, 9, False)
    ks = [0.75, 1.25, 2.5]
    prompts = np.stack([((np.float64(k) * S[b] + R * Nm) * lor_counts).astype(np.float32)
                        for b, k in enumerate(ks)], axis=-1)
    geom, _ = expandRings(w.geometry, w.normals)
    np.savez_compressed(tmp_path / 'PromptsSinogramMashed.npz', prompts)
    np.savez_compressed(tmp_path / 'RandomsSinogram.npz', R)
    np.savez_compressed(tmp_path / 'NormSinogram.npz', Nm)
    np.save(tmp_path / 'AttenuationMask.npy', M)
    for b in range(3):
        np.savez_compressed(tmp_path / f'SSS_mashed_bin{b}.npz', S[b])
    grid = np.array([12, 12, 6])
    got = ScaleScattersToPrompts(sino, w.lor_coordinates, str(tmp_path), grid, w.image_size, w.geometry, geom, 2,
                                 True, lor_counts)
    assert np.array_equal(np.load(tmp_path / 'ScaleFactors.npy'), got)
    want = oracle.scale_scatters_to_prompts(prompts, S, R, Nm, M, lor_counts, sino, w.lor_coordinates, grid,
                                            w.image_size, w.geometry, geom, 2, True)
    for b, k in enumerate(ks):
        assert abs(got[b] - k) <= 2e-5 * k
        assert abs(got[b] - want[b]) <= 1e-5 * abs(want[b])


def test_scale_to_prompts_device_pointers():
    """SSS_FIT_DEVICE: the same sums from device-resident sinograms
    (asynchronous call, sums read back after a synchronisation)."""
    from opensss_amd import _native
    from opensss_amd.backprojection import fit_geometry_buffers, span_entries
    from opensss_amd.scaling import scale_to_prompts, scale_to_prompts_device
    w = _case()
    n_bins = 9
    sino, S, prompts, R, Nm, M, lor_counts = _scaling_inputs(w, True, n_bins, 13, True)
    geom, _ = expandRings(w.geometry, w.normals)
    grid = np.array([24, 24, 12])
    _, want = scale_to_prompts(prompts, S, R, Nm, M, lor_counts[0, 0, :], sino, w.lor_coordinates, grid,
                               w.image_size, w.geometry, geom, 2, True, return_sums=True)
    ctx = _native.default_context()
    entries = span_entries(sino)
    bufs = fit_geometry_buffers(geom, w.lor_coordinates, entries, context=ctx)
    dev = [_native.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a))
           for a in (S, prompts, np.ascontiguousarray(lor_counts[0, 0, :]), R, Nm, M)]
    sums = _native.DeviceBuffer(ctx, 16 * n_bins)
    try:
        scale_to_prompts_device(grid, w.image_size, geom, w.lor_coordinates, entries, sino.shape[0], 2, True,
                                n_bins, dev[0].ptr.value, dev[1].ptr.value, True, dev[5].ptr.value,
                                sums.ptr.value, counts_ptr=dev[2].ptr.value, counts_per_plane=True,
                                randoms_ptr=dev[3].ptr.value, norm_ptr=dev[4].ptr.value, geometry_buffers=bufs,
                                context=ctx)
        ctx.synchronize()
        got = sums.download(np.empty((n_bins, 2), np.float64))
    finally:
        for b in dev + list(bufs) + [sums]:
            b.free()
    assert np.array_equal(got, want)
