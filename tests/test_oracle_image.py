"""Pins of the oracle's image preparation (CPU) and the host side of the
crop: the resize of a constant image is that constant; without smoothing at
factor 1 it returns the image; away from the borders a quadratic along an
axis comes back as x^2 + (discrete variance of the Gaussian) at the grid-mode
source coordinate x = (o + 1/2) f - 1/2, with sigma = (f - 1) / 2 (cubic
B-splines reproduce polynomials up to degree 3, a symmetric kernel adds its
variance to x^2); the host crop window equals the reference loop's crop
(symmetric zero pad, slices with Python's negative-start rule); the host's
Gaussian taps are scipy's response to a delta."""
import numpy as np
import pytest

from opensss_amd.images import crop_window, gaussian_taps


def test_constant_image_stays_constant(oracle):
    img = np.full((30, 30, 14), 0.75, np.float32)
    out = oracle.crop_and_downscale(img, [2, 2, 3], [50, 50, 36], [8, 8, 5])
    assert out.shape == (8, 8, 5) and np.allclose(out, 0.75, rtol=0, atol=1e-7)


def test_identity_at_factor_one(oracle):
    img = np.random.default_rng(1).uniform(0, 1, (16, 16, 8)).astype(np.float32)
    out = oracle.crop_and_downscale(img, [1, 1, 1], [16, 16, 8], [16, 16, 8], anti_aliasing=False)
    assert np.max(np.abs(out - img)) <= 2e-6


@pytest.mark.parametrize('f', [2, 3])
def test_quadratic_gets_the_gaussian_variance(oracle, f):
    n, m = 60 * f, 60
    x = np.arange(n, dtype=np.float64)
    img = np.broadcast_to((x ** 2 / n)[:, None, None], (n, 4, 4)).astype(np.float32).copy()
    out = oracle.crop_and_downscale(img, [1, 1, 1], [n, 4, 4], [m, 4, 4], clip=False)
    sigma = (f - 1) / 2
    taps = gaussian_taps(sigma)
    r = (taps.size - 1) // 2
    var = float(np.sum(taps * np.arange(-r, r + 1) ** 2))
    assert abs(var - sigma ** 2) < 0.15 * sigma ** 2   # the sampled Gaussian's variance
    o = np.arange(m)
    xs = (o + 0.5) * f - 0.5
    want = (xs ** 2 + var) / n
    inner = slice(10, m - 10)
    assert np.max(np.abs(out[inner, 1, 1] - want[inner])) <= 2e-6 * want.max()


@pytest.mark.parametrize('shape,vox,size', [
    ((40, 36, 20), (2, 2, 3), (60, 60, 45)),     # slices on every axis
    ((10, 12, 9), (2, 2, 2), (30, 30, 30)),      # zero pad on every axis (odd and even)
    ((20, 20, 7), (1, 1, 1), (10, 26, 3)),       # slice, pad, slice
    ((20, 20, 5), (1, 1, 1), (10, 10, 8)),       # window starts below 0, ends inside: Python slice rule
])
def test_crop_window_is_the_reference_loop(oracle, shape, vox, size):
    img = np.random.default_rng(4).uniform(0, 1, shape).astype(np.float32)
    want = oracle.crop(img, vox, size)
    starts, lengths = crop_window(shape, vox, size)
    assert tuple(lengths) == want.shape
    got = np.zeros(lengths, np.float32)
    idx = [np.arange(lengths[d]) + starts[d] for d in range(3)]
    ok = [(i >= 0) & (i < shape[d]) for d, i in enumerate(idx)]
    sub = img[np.ix_(np.clip(idx[0], 0, shape[0] - 1), np.clip(idx[1], 0, shape[1] - 1),
                     np.clip(idx[2], 0, shape[2] - 1))]
    mask = ok[0][:, None, None] & ok[1][None, :, None] & ok[2][None, None, :]
    got[mask] = sub[mask]
    assert np.array_equal(got, want)


@pytest.mark.parametrize('sigma', [0.5, 1.0, 3.25])
def test_gaussian_taps_are_scipys(sigma):
    from scipy.ndimage import gaussian_filter1d
    taps = gaussian_taps(sigma)
    r = (taps.size - 1) // 2
    delta = np.zeros(4 * r + 3)
    delta[2 * r + 1] = 1.0
    resp = gaussian_filter1d(delta, sigma, mode='nearest')
    assert np.array_equal(resp[r + 1:3 * r + 2], taps)
    assert gaussian_taps(0.0) is None
