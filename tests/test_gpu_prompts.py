"""Synthetic prompts on the GPU (sss_forward_project, sss_poisson) and the
config-1 tail fit they make possible.

* forward projection against the oracle's line integrals (trues = sum(lam L)
  exp(-sum(mu L)) per LOR) and against closed-form chords through a uniform
  box (no oracle involved);
* Poisson counts: the same seed gives the same counts, other seeds others;
  means, variances and a chi-square test against the Poisson pmf;
* config 1 end to end: SSS -> interpolation to the 193 x 96 x 64 sinogram ->
  tail mask -> prompts of trues + 0.3 x scatter + uniform randoms at 2e7
  counts -> tail fit.  Noiseless prompts give back the scatter's true scale
  within 1e-5; Poisson prompts give the oracle's fit of the same counts
  within 1e-5 and the true scale within their noise.
"""
import numpy as np
import pytest

from opensss_amd import phantoms
from opensss_amd.tables import detectors_used, rings_used

pytestmark = pytest.mark.gpu


def _lor_segments(geom, sino_index, n_det, accel):
    NR = geom.shape[0]
    segs, where = [], []
    for r1 in range(NR):
        for r2 in range(NR):
            for d1 in range(0, n_det, accel):
                for d2 in range(d1, n_det, accel):
                    if d1 == d2 and r1 == r2:
                        continue
                    segs.append([geom[r1, d1, 0], geom[r1, d1, 1], geom[r1, 0, 2],
                                 geom[r2, d2, 0], geom[r2, d2, 1], geom[r2, 0, 2]])
                    where.append((d1, d2, sino_index[r1, r2]))
    return np.array(segs), where


def test_forward_project_matches_oracle_line_integrals(oracle):
    from opensss_amd.prompts import ForwardProject
    w = phantoms.small_tof(seed=4)
    N = w.geometry.shape[1]
    got = ForwardProject(w.activity, w.attenuation, w.image_size, w.geometry, w.lor_coordinates,
                         SinogramIndex=w.sinogram_index, AccelerationFactor=2)
    segs, where = _lor_segments(w.geometry, w.sinogram_index, N, 2)
    n, m, l = oracle.line_integrals(np.array(w.attenuation.shape), w.image_size,
                                    w.attenuation.ravel(order='F'), w.activity.ravel(order='F'), segs)
    want = np.zeros_like(got)
    lor = w.lor_coordinates
    for k, (d1, d2, p) in enumerate(where):
        want[lor[d1, d2, 1], lor[d1, d2, 0], p] = np.float32(l[k] * np.exp(-m[k]))
    assert np.allclose(got, want, rtol=1e-6, atol=0)
    assert got.max() > 0 and np.count_nonzero(got) > 1000


def test_forward_project_uniform_box_is_the_chord():
    """lam = 1 on a box of whole voxels, no attenuation: each LOR's trues are
    the length of its chord through the box, computed analytically."""
    from opensss_amd.prompts import ForwardProject
    w = phantoms.small_tof(seed=4)
    X, Y, Z = w.attenuation.shape
    lam = np.zeros((X, Y, Z))
    lam[5:19, 7:15, 2:9] = 1.0
    b = np.asarray(w.image_size, dtype=np.float64)
    vox = (b[3:] - b[:3]) / np.array([X, Y, Z])
    lo, hi = b[:3] + vox * np.array([5, 7, 2]), b[:3] + vox * np.array([19, 15, 9])
    got = ForwardProject(lam, None, w.image_size, w.geometry, w.lor_coordinates, SinogramIndex=w.sinogram_index,
                         AccelerationFactor=3)
    segs, where = _lor_segments(w.geometry, w.sinogram_index, w.geometry.shape[1], 3)
    lor = w.lor_coordinates
    hit = 0
    for seg, (d1, d2, p) in zip(segs, where):
        pa, pb = seg[:3], seg[3:]
        d = pb - pa
        t0, t1 = 0.0, 1.0
        for k in range(3):
            if abs(d[k]) < 1e-15:
                if not lo[k] <= pa[k] <= hi[k]:
                    t0, t1 = 1.0, 0.0
                continue
            a_, b_ = (lo[k] - pa[k]) / d[k], (hi[k] - pa[k]) / d[k]
            t0, t1 = max(t0, min(a_, b_)), min(t1, max(a_, b_))
        chord = max(0.0, t1 - t0) * np.linalg.norm(d)
        assert abs(got[lor[d1, d2, 1], lor[d1, d2, 0], p] - chord) <= 2e-6 * max(1.0, chord), (d1, d2, p)
        hit += chord > 0
    assert hit > 100


def test_poisson_counts():
    from scipy import stats
    from opensss_amd.prompts import PoissonSample
    n = 400_000
    means = [0.3, 4.0, 25.0, 400.0]
    for mu in means:
        mean = np.full(n, mu, dtype=np.float32)
        a = PoissonSample(mean, seed=7)
        b = PoissonSample(mean, seed=7)
        c = PoissonSample(mean, seed=8)
        assert np.array_equal(a, b) and not np.array_equal(a, c)
        assert np.all(a == np.floor(a)) and a.min() >= 0
        # mean and variance within 5 standard errors
        se_mean = np.sqrt(mu / n)
        assert abs(a.mean() - mu) < 5 * se_mean, (mu, a.mean())
        se_var = mu * np.sqrt(2.0 / n) * np.sqrt(1 + 1 / (2 * mu))
        assert abs(a.var() - mu) < 5 * se_var, (mu, a.var())
        # chi-square over the central counts (expected >= 50 each), tail bins pooled
        ks = np.arange(int(a.max()) + 2)
        pmf = stats.poisson.pmf(ks, mu) * n
        keep = pmf >= 50
        obs = np.bincount(a.astype(np.int64), minlength=len(ks))[:len(ks)]
        o, e = obs[keep].astype(np.float64), pmf[keep]
        o_rest, e_rest = n - o.sum(), n - e.sum()
        if e_rest >= 20:                       # both tails pooled into one bin
            o, e = np.append(o, o_rest), np.append(e, e_rest)
        else:                                  # too few expected: into the last bin
            o[-1] += o_rest
            e[-1] += e_rest
        chi2 = ((o - e) ** 2 / e).sum()
        assert stats.chi2.sf(chi2, len(o) - 1) > 1e-4, (mu, chi2, len(o))
    # a scaled mean, per cell
    mean = np.linspace(0, 50, 1000, dtype=np.float32)
    a = PoissonSample(np.tile(mean, 200), scale=2.0, seed=3).reshape(200, 1000)
    assert np.all(a[:, 0] == 0)
    assert abs(a[:, -1].mean() - 100.0) < 5 * np.sqrt(100.0 / 200)


def _config1_sinograms():
    """Config 1: SSS, interpolation to the full sinogram, mask, trues."""
    from opensss_amd.interpolation import InterpolateScatters
    from opensss_amd.prompts import ForwardProject
    from opensss_amd.tails_mask import MaskGenerator
    from tests.sss_cases import gpu_dense
    w = phantoms.config1()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    sampled = gpu_dense(w, 'fp32')
    S = InterpolateScatters(sampled, NR, w.n_rings_used, N, w.n_det_used, w.lor_coordinates, w.sinogram_index,
                            None, False)
    mask = MaskGenerator(w.activity, w.attenuation, w.image_size, w.geometry, w.lor_coordinates, w.sinogram_index,
                         True, 2)
    T = ForwardProject(w.activity, w.attenuation, w.image_size, w.geometry, w.lor_coordinates,
                       SinogramIndex=w.sinogram_index)
    return w, S, np.float32(mask >= 1), T


def test_config1_tail_fit_recovers_the_scatter_scale(oracle):
    from opensss_amd.backprojection import BackProjectionScatterRandFit
    from opensss_amd.prompts import synthetic_prompts
    w, S, M, T = _config1_sinograms()
    NR, N = w.geometry.shape[0], w.geometry.shape[1]
    assert S.shape == T.shape == M.shape == (N + 1, N // 2, NR * NR) == (193, 96, 64)
    assert 0.05 < M.mean() < 0.95
    # trues vanish on the tails (the emission lies inside the attenuating body)
    assert T[M > 0].max() <= 1e-9 * T.max()
    order = w.sinogram_index[:NR, :NR].T
    grid = np.array(w.attenuation.shape)
    P0, R, k = synthetic_prompts(T, S, noise=False)
    fit0 = BackProjectionScatterRandFit(P0, S, R, M, grid, w.image_size, w.geometry, w.lor_coordinates, order, 2)
    assert abs(fit0[0] - k) <= 1e-5 * k, (fit0[0], k)
    P, R, k = synthetic_prompts(T, S, seed=11)
    assert abs(P.sum() - 2e7) < 5 * np.sqrt(2e7)
    fit = BackProjectionScatterRandFit(P, S, R, M, grid, w.image_size, w.geometry, w.lor_coordinates, order, 2)
    want, _ = oracle.tail_fit(P, S, R, M, grid, w.image_size, w.geometry, w.lor_coordinates, 2,
                              SinogramOrder=order)
    print(f'config 1 tail fit: k {k:.6e}, noiseless {fit0[0]:.6e}, Poisson {fit[0]:.6e}, oracle {want[0]:.6e}')
    assert abs(fit[0] - want[0]) <= 1e-5 * abs(want[0])
    assert abs(fit[0] - k) <= 0.05 * k
