"""Config 2's sinogram stages at full size on the GPU: the 15-bin scatter
estimate of the 760 x 80 scanner (72 x 16 sampled) interpolated to the
761 x 380 x 159 x 15 sinogram (span 19, MRD 9), the span tail mask on the
159-ring extended geometry, and the span tail fit on device-resident
sinograms.  The oracle's loops are far too slow at this size, so the checks
are the properties that fix each stage:

* interpolation: at every sampled detector pair and every span plane the
  output is the sampled value of the plane of the nearest sampled rings
  (brute-force nearest, ties to the lower ring), clamped at 0; a sample of
  other cells holds the value of a nearest sampled (radial, angular) point;
  the sum over bins is the bins added in order;
* mask: for a box of mu, a LOR is a tail iff it misses the box (vectorised
  slab test on every LOR of the accelerated grid);
* fit: prompts = k S + R give back k within 1e-5, and two runs agree bit
  for bit.
"""
import numpy as np
import pytest

from opensss_amd import _native, phantoms
from opensss_amd.data_reduction import SinogramLUT
from opensss_amd.geometry import expandRings
from opensss_amd.tables import detectors_used, rings_used

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def config2_stages():
    from opensss_amd.interpolation import Interpolator
    from tests.sss_cases import gpu_compact
    w = phantoms.config2()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    msi = SinogramLUT(NR, 19, 9)[0]
    assert msi.shape == (159, 2)
    compact = gpu_compact(w, 'fp32')
    ctx = _native.default_context()
    ip = Interpolator(NR, w.n_rings_used, N, w.n_det_used, w.lor_coordinates, msi, True, context=ctx)
    B = compact.shape[0]
    n_out = ip.n_cells_out * ip.n_out_planes
    src = _native.DeviceBuffer.from_array(ctx, compact)
    out = _native.DeviceBuffer(ctx, 4 * n_out * B)
    tot = _native.DeviceBuffer(ctx, 4 * n_out)
    ip.run_device(src.ptr.value, B, out.ptr.value, tot.ptr.value)
    ctx.synchronize()
    yield dict(w=w, msi=msi, compact=compact, ip=ip, out=out, tot=tot, B=B, ctx=ctx)
    for b in (src, out, tot):
        b.free()


def _bin(st, b):
    ip = st['ip']
    n_out = ip.n_cells_out * ip.n_out_planes
    host = np.empty(n_out, np.float32)
    _native.check(_native.load_library().sss_buffer_download(st['ctx'].handle, host.ctypes.data_as(_native.ctypes.c_void_p),
                                                              _native.ctypes.c_void_p(st['out'].ptr.value + 4 * n_out * b),
                                                              host.nbytes), 'download')
    return host.reshape(ip.shape_out)


def _nearest_lower(grid, q):
    d = np.abs(np.asarray(grid, dtype=np.float64) - float(q))
    return int(np.flatnonzero(d == d.min())[0])


def test_full_interpolation_copies_the_nearest_sampled_values(config2_stages):
    st = config2_stages
    w, msi, compact = st['w'], st['msi'], st['compact']
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    nru = w.n_rings_used
    rings = rings_used(NR, nru)
    dets = detectors_used(N, w.n_det_used)
    lor = w.lor_coordinates
    i1, i2 = np.triu_indices(len(dets))
    rad_s, ang_s = lor[dets[i1], dets[i2], 1].astype(np.int64), lor[dets[i1], dets[i2], 0].astype(np.int64)
    src_plane = np.array([_nearest_lower(rings, msi[p, 0]) + _nearest_lower(rings, msi[p, 1]) * nru
                          for p in range(msi.shape[0])])
    # span 19 with MRD 9 keeps segment 0 only: planes of ring difference
    # below 10, whose averaged rings are closest to sampled rings r, r
    assert len(set(src_plane.tolist())) == nru
    rng = np.random.default_rng(0)
    probe = rng.integers(0, (N + 1) * (N // 2), size=1500)
    pr, pa = probe // (N // 2), probe % (N // 2)
    d2 = (pr[:, None] - rad_s[None, :]) ** 2 + (pa[:, None] - ang_s[None, :]) ** 2
    acc = np.zeros(st['ip'].shape_out, np.float32)
    for b in range(st['B']):
        out = _bin(st, b)
        assert out.shape == (N + 1, N // 2, 159)
        want = np.maximum(compact[b][src_plane][:, :], 0)            # (159, cells)
        assert np.array_equal(out[rad_s, ang_s, :], want.T), b
        if b == 7:
            for k in range(len(probe)):
                near = np.flatnonzero(d2[k] == d2[k].min())
                vals = out[pr[k], pa[k], :]
                assert any(np.array_equal(vals, want[:, c]) for c in near), k
        acc += out
    tot = st['tot'].download(np.empty(st['ip'].shape_out, np.float32))
    assert np.array_equal(tot, acc)
    assert tot.max() > 0


def _hits_box(p0, p1, lo, hi):
    """Vectorised: positive-length intersection of segments p0 -> p1 with the box."""
    d = p1 - p0
    t0 = np.zeros(len(p0))
    t1 = np.ones(len(p0))
    for k in range(3):
        dk = d[:, k]
        par = np.abs(dk) < 1e-15
        with np.errstate(divide='ignore', invalid='ignore'):
            a = (lo[k] - p0[:, k]) / dk
            b = (hi[k] - p0[:, k]) / dk
        t0 = np.where(par, t0, np.maximum(t0, np.minimum(a, b)))
        t1 = np.where(par, t1, np.minimum(t1, np.maximum(a, b)))
        inside = (p0[:, k] > lo[k]) & (p0[:, k] < hi[k])
        t1 = np.where(par & ~inside, -1.0, t1)
    return t1 - t0 > 1e-12


def test_full_span_mask_is_the_box_hull(config2_stages):
    from opensss_amd.tails_mask import MaskGeneratorMashing
    st = config2_stages
    w, msi = st['w'], st['msi']
    geom, _ = expandRings(w.geometry, w.normals)
    N = geom.shape[1]
    X, Y, Z = w.attenuation.shape
    mu = np.zeros((X, Y, Z))
    mu[30:70, 20:61, 10:35] = 0.096
    b = np.asarray(w.image_size, dtype=np.float64)
    vox = (b[3:] - b[:3]) / np.array([X, Y, Z])
    lo, hi = b[:3] + vox * np.array([30, 20, 10]), b[:3] + vox * np.array([70, 61, 35])
    accel = 2
    mask = MaskGeneratorMashing(np.zeros_like(mu), mu, w.image_size, geom, w.lor_coordinates, msi, True, accel)
    lor = w.lor_coordinates
    d1, d2 = np.triu_indices(N)
    keep = (d1 % accel == 0) & (d2 % accel == 0)
    d1, d2 = d1[keep], d2[keep]
    seen = np.zeros(mask.shape, bool)
    for p in range(msi.shape[0]):
        ra, rb = int(np.float32(msi[p, 0]) * 2), int(np.float32(msi[p, 1]) * 2)
        sel = ~((d1 == d2) & (ra == rb))
        a, c = d1[sel], d2[sel]
        p0 = np.stack([geom[ra, a, 0], geom[ra, a, 1], np.full(len(a), geom[ra, 0, 2])], axis=1)
        p1 = np.stack([geom[rb, c, 0], geom[rb, c, 1], np.full(len(c), geom[rb, 0, 2])], axis=1)
        want = np.where(_hits_box(p0, p1, lo, hi), 0, 1).astype(np.uint8)
        rr, aa = lor[a, c, 1], lor[a, c, 0]
        assert np.array_equal(mask[rr, aa, p], want), p
        seen[rr, aa, p] = True
    assert not mask[~seen].any()
    assert 0 < mask.sum() < seen.sum()


def test_full_span_fit_recovers_k(config2_stages):
    from opensss_amd.backprojection import fit_geometry_buffers, span_entries, tail_fit_device
    from opensss_amd.tails_mask import MaskGeneratorMashing
    st = config2_stages
    w, msi, ctx = st['w'], st['msi'], st['ctx']
    geom, _ = expandRings(w.geometry, w.normals)
    S = _bin(st, 7)
    M = np.float32(MaskGeneratorMashing(w.activity, w.attenuation, w.image_size, geom, w.lor_coordinates, msi,
                                        True, 2) >= 1)
    assert 0.05 < M.mean() < 0.95
    R = np.full(S.shape, np.float32(0.05 * S.mean()), np.float32)
    k = 0.4375
    P = (np.float64(k) * S + R).astype(np.float32)
    bufs = [_native.DeviceBuffer.from_array(ctx, a) for a in (S, P, R, M)]
    gb = fit_geometry_buffers(geom, w.lor_coordinates, span_entries(msi), context=ctx)
    try:
        runs = [tail_fit_device(w.attenuation.shape, w.image_size, geom, w.lor_coordinates, span_entries(msi),
                                msi.shape[0], 2, True, *(b.ptr.value for b in bufs), context=ctx,
                                geometry_buffers=gb) for _ in range(2)]
    finally:
        for b in bufs + list(gb):
            b.free()
    scale = runs[0][0]
    print(f'config 2 span fit: k {k}, fitted {scale:.9f}')
    assert runs[0] == runs[1]
    assert abs(scale - k) <= 1e-5 * k
