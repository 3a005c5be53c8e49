"""sss_tail_fit_noalloc (device buffers, static scratch, no allocation)
against sss_tail_fit, through the C ABI: the same sums and images bit for
bit, on two calls in a row, the second on a smaller fit grid and fewer
planes (the shared scratch keeps the first call's contents and must be
cleared); a grid above the scratch's limit is refused."""
import ctypes

import numpy as np
import pytest

from opensss_amd import _native, phantoms
from opensss_amd.backprojection import nospan_entries, span_entries, tail_fit
from opensss_amd.data_reduction import SinogramLUT
from opensss_amd.geometry import expandRings

pytestmark = pytest.mark.gpu


def _inputs(w, n_planes, seed):
    rng = np.random.default_rng(seed)
    N = w.geometry.shape[1]
    shape = (N + 1, N // 2, n_planes)
    S = rng.uniform(0.5, 2.0, size=shape).astype(np.float32)
    R = rng.uniform(0.0, 0.3, size=shape).astype(np.float32)
    P = (1.7 * S + R + rng.normal(0, 0.05, size=shape)).astype(np.float32)
    M = (rng.uniform(size=shape) < 0.6).astype(np.float32)
    return P, S, R, M


def _noalloc(ctx, grid, bounds, geom, lor, entries, n_planes, accel, lw, P, S, R, M):
    lib = _native.load_library()
    keep = []

    def dev(a):
        b = _native.DeviceBuffer.from_array(ctx, np.ascontiguousarray(a))
        keep.append(b)
        return b.ptr

    g = np.ascontiguousarray(grid, dtype=np.int32)
    bd = np.ascontiguousarray(bounds, dtype=np.float64)
    nvox = int(np.prod(g))
    sums_d = _native.DeviceBuffer(ctx, 16)
    img_d = _native.DeviceBuffer(ctx, 4 * nvox * 8)
    keep += [sums_d, img_d]
    rc = lib.sss_tail_fit_noalloc(ctx.handle, _native.ptr(g), _native.ptr(bd), geom.shape[0], geom.shape[1],
                                  dev(geom.astype(np.float64)), dev(lor.astype(np.int16)), entries.shape[0],
                                  dev(entries.astype(np.int32)), n_planes, accel, int(lw), dev(S), dev(P), dev(R),
                                  dev(M), sums_d.ptr, img_d.ptr)
    _native.check(rc, 'sss_tail_fit_noalloc')
    ctx.synchronize()
    sums = sums_d.download(np.empty(2, np.float64))
    imgs = img_d.download(np.empty(4 * nvox, np.float64))
    for b in keep:
        b.free()
    return sums, imgs


def test_noalloc_fit_matches_the_allocating_entry_point():
    ctx = _native.default_context()
    w = phantoms.small_tof(seed=21)
    NR = w.geometry.shape[0]
    # call 1: span entries on a 24 x 24 x 12 fit grid
    geom_ext, _ = expandRings(w.geometry, w.normals)
    msi = SinogramLUT(NR, 3, 4)[0]
    ent = span_entries(msi)
    P, S, R, M = _inputs(w, msi.shape[0], 1)
    grid1 = np.array([24, 24, 12])
    want = tail_fit(P, S, R, M, grid1, w.image_size, geom_ext, w.lor_coordinates, ent, 1, True, return_images=True)
    sums, imgs = _noalloc(ctx, grid1, w.image_size, geom_ext, w.lor_coordinates, ent, msi.shape[0], 1, True,
                          P, S, R, M)
    assert sums[0] == want[1] and sums[1] == want[2] and sums[1] > 0
    want_imgs = np.stack([im.ravel(order='F') for im in want[3]]).ravel()
    assert np.array_equal(imgs, want_imgs)
    # call 2: fewer planes (no span) on a smaller grid, through the same scratch
    order = w.sinogram_index[:NR, :NR].T
    ent2 = nospan_entries(order, NR)[: NR * NR // 2]
    P2, S2, R2, M2 = _inputs(w, NR * NR, 2)
    grid2 = np.array([12, 12, 6])
    want2 = tail_fit(P2, S2, R2, M2, grid2, w.image_size, w.geometry, w.lor_coordinates, ent2, 2, False,
                     return_images=True)
    sums2, imgs2 = _noalloc(ctx, grid2, w.image_size, w.geometry, w.lor_coordinates, ent2, NR * NR, 2, False,
                            P2, S2, R2, M2)
    assert sums2[0] == want2[1] and sums2[1] == want2[2] and sums2[1] > 0
    want_imgs2 = np.stack([im.ravel(order='F') for im in want2[3]]).ravel()
    assert np.array_equal(imgs2, want_imgs2)


def test_noalloc_fit_refuses_grids_above_its_scratch():
    ctx = _native.default_context()
    lib = _native.load_library()
    g = np.array([128, 128, 65], dtype=np.int32)       # 1,064,960 voxels > 2^20
    bd = np.array([-1.0, -1.0, -1.0, 1.0, 1.0, 1.0])
    dummy = _native.DeviceBuffer(ctx, 1 << 12)
    rc = lib.sss_tail_fit_noalloc(ctx.handle, _native.ptr(g), _native.ptr(bd), 1, 4, dummy.ptr, dummy.ptr, 0,
                                  None, 1, 1, 0, dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr, dummy.ptr, None)
    dummy.free()
    assert rc == -4
    assert b'exceeds the scratch' in lib.sss_last_error()
