"""GPU image preparation (CropAndDownscale, PrepareImages) against the
oracle's restatement through scipy.ndimage.  Floating point: the GPU sums the
same float64 terms in another order (Gaussian taps paired as scipy pairs
them; spline coefficients by the same recursion; 64 interpolation terms), so
the bar is agreement within 2e-6 of the image's largest value (float32 has
6e-8 relative resolution: a few ulps); on these cases all but a handful of
voxels come out bit for bit."""
import numpy as np
import pytest

from opensss_amd import images

pytestmark = pytest.mark.gpu

TOL = 2e-6


@pytest.mark.parametrize('shape,vox,size,dims,aa,clip', [
    ((64, 64, 40), (2.0, 2.0, 3.0), (100.0, 100.0, 90.0), (12, 12, 7), True, True),   # slice, smooth
    ((30, 30, 17), (4.0, 4.0, 4.0), (180.0, 160.0, 60.0), (20, 18, 9), True, True),  # zero pad x, y
    ((48, 40, 24), (1.0, 1.0, 2.0), (48.0, 40.0, 48.0), (20, 60, 12), True, False),  # upscale along y
    ((32, 32, 16), (2.0, 2.0, 2.0), (64.0, 64.0, 32.0), (11, 11, 6), False, True),   # no anti-aliasing
])
def test_crop_and_downscale_matches_oracle(oracle, shape, vox, size, dims, aa, clip):
    rng = np.random.default_rng(sum(shape))
    img = rng.gamma(2.0, 0.5, size=shape).astype(np.float32)
    img[shape[0] // 3:shape[0] // 2, :, :] += 3.0       # an edge, so the clip has something to do
    want = oracle.crop_and_downscale(img, vox, size, dims, anti_aliasing=aa, clip=clip)
    got = images.CropAndDownscale(img, vox, size, dims, clip=clip, anti_aliasing=aa)
    assert got.shape == want.shape and got.dtype == np.float32
    err = np.max(np.abs(got.astype(np.float64) - want)) / np.max(np.abs(want))
    print(f'{shape} -> {dims}: max err {err:.2e} of the max, identical {np.mean(got == want):.3f}')
    assert err <= TOL
    assert np.mean(got == want) >= 0.99


def test_prepare_images_from_interfile(oracle, tmp_path):
    """Interfile round trip and the attenuation floor (mu < 0.001 -> 0)."""
    rng = np.random.default_rng(6)
    shape = (40, 40, 20)
    act = rng.uniform(0, 5, shape).astype(np.float32)
    mu = np.where(rng.uniform(size=shape) < 0.5, rng.uniform(0, 0.002, shape), rng.uniform(0.05, 0.1, shape))
    mu = mu.astype(np.float32)
    for name, im in (('act', act), ('mu', mu)):
        im.transpose(2, 1, 0).astype(np.float32).tofile(tmp_path / f'{name}.i33')
        (tmp_path / f'{name}.h33').write_text(
            ''.join(f'matrix size [{i + 1}] := {shape[i]}\nscaling factor (mm/pixel) [{i + 1}] := 2.5\n'
                    for i in range(3)) + f'name of data file := {name}.i33\n')
    a, m, size, dims = images.PrepareImages(str(tmp_path / 'act.h33'), str(tmp_path / 'mu.h33'),
                                            [80.0, 80.0, None], np.array([8.0, 8.0, 10.0]))
    assert list(dims) == [10, 10, 5] and np.allclose(size, [80, 80, 50])
    wa = oracle.crop_and_downscale(act, [2.5] * 3, size, dims)
    wm = oracle.crop_and_downscale(mu, [2.5] * 3, size, dims)
    wm[wm < np.float32(0.001)] = 0
    assert np.max(np.abs(a - wa)) <= TOL * np.max(wa)
    assert np.max(np.abs(m - wm)) <= TOL * np.max(wm)
    assert np.array_equal(m == 0, wm == 0)
