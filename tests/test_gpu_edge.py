"""GPU parity on degenerate inputs, against the oracle.

- an all-zero activity image: every pair is skipped (ActPath 0 on both
  half paths), so the estimate is zero; the fold table is empty, and the
  same subject then takes a real image and gives a fresh subject's output;
- one hot voxel: most half paths have no non-zero term, the others a
  single record (one group of padding);
- a single scatter point, and scatter points whose half paths all miss the
  emission (a box of activity away from them).

Each case runs the TOF folds the 7-bin (relative window) and 15-bin (quad
sliding window) subjects take in FP32, and FP64, and the one-bin kernel of
a non-TOF subject (1e9 ps resolution: every TOF weight is 1).
"""
import dataclasses

import numpy as np
import pytest

from opensss_amd import phantoms
from tests.sss_cases import gpu_compact, gpu_subject, oracle_compact, rel_err_above

pytestmark = pytest.mark.gpu

TOL = {'fp32': 1e-4, 'fp64': 1e-6}
BINS = [1, 7, 15]


def _case(n_bins, n_sp=12):
    w = phantoms.small_tof(seed=21, n_bins=n_bins, n_sp=n_sp)
    return dataclasses.replace(w, tof_resolution=1e9) if n_bins == 1 else w


@pytest.mark.parametrize('n_bins', BINS)
@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_zero_image_then_real_image(n_bins, precision):
    w = _case(n_bins)
    s = gpu_subject(w, precision)
    try:
        s.run(np.zeros_like(w.activity))
        zero = s.output_compact()
        assert not np.any(zero), 'an all-zero image scatters nothing'
        s.run(w.activity)
        after = s.output_compact()
    finally:
        s.close()
    fresh = gpu_compact(w, precision)
    assert np.any(fresh)
    assert np.array_equal(after, fresh), 'the subject recovers from an empty fold table'


@pytest.mark.parametrize('n_bins', BINS)
@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_one_hot_voxel_vs_oracle(oracle, n_bins, precision):
    w = _case(n_bins)
    lam = np.zeros_like(w.activity)
    hot = tuple(s // 2 for s in lam.shape)
    lam[hot] = 3.0
    want = oracle_compact(oracle, w, lam=lam)
    got = gpu_compact(w, precision, lam=lam)
    assert np.any(want)
    rel, absn = rel_err_above(got, want)
    print(f'{w.name} one hot voxel {precision}: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= TOL[precision] and absn <= 1e-6


@pytest.mark.parametrize('n_bins', BINS)
@pytest.mark.parametrize('precision', ['fp32', 'fp64'])
def test_single_scatter_point_vs_oracle(oracle, n_bins, precision):
    w = _case(n_bins)
    pts = w.scatter_points[len(w.scatter_points) // 2:][:1]
    want = oracle_compact(oracle, w, points=pts)
    got = gpu_compact(w, precision, points=pts)
    assert np.any(want)
    rel, absn = rel_err_above(got, want)
    print(f'{w.name} one scatter point {precision}: max rel err {rel:.3e}, max abs/peak {absn:.3e}')
    assert rel <= TOL[precision] and absn <= 1e-6


@pytest.mark.parametrize('n_bins', [7, 15])
def test_activity_away_from_every_path(oracle, n_bins):
    """Activity in one corner box that no half path of the chosen points
    crosses: the oracle and the GPU both give zero, or agree where a path
    does cross it."""
    w = _case(n_bins)
    lam = np.zeros_like(w.activity)
    lam[:3, :3, :2] = 1.0
    w = dataclasses.replace(w, activity=lam)
    want = oracle_compact(oracle, w)
    for precision in ('fp32', 'fp64'):
        got = gpu_compact(w, precision)
        if not np.any(want):
            assert not np.any(got)
        else:
            rel, absn = rel_err_above(got, want)
            assert rel <= TOL[precision] and absn <= 1e-6
