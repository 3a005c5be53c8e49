"""The float64 estimate behind energy_row_fast (opensss_amd/csrc/sss_scatter.hip),
modelled on the CPU.

The FP32 kernels take a pair's energy row from a quick estimate of
511/(2 + c) + 1/2 (c: the law-of-cosines cosine at the scatter point):
2 d1 d2 for the root 2 sqrt(d1^2 d2^2), and products with reciprocals
(v_rcp_f64 refined by two Newton steps: within 2 ulp of 1/x) for the two
quotients.  Where the estimate's fraction lies more than 1e-9 from an
integer it takes floor(estimate); otherwise it runs the exact sequence.
The row is right if the estimate is within 1e-9 of the exact sequence's
value, which this model checks with a wide margin over scanner geometries
(including points next to a detector), with every reciprocal perturbed by
up to 2 ulp.  The numerator (d1^2 + d2^2) - lor2 is rounded as in the exact
sequence (contraction is off there), so the two differ only through the
root and the reciprocals; a fused multiply-add in it would cost most of the
margin for points next to a detector, where it cancels.  The
exact sequence is the one SingleScatterSimulationTOF.py:216-236
(CalcAngleScatter) reduces to: cos(180 - acos(c)) = -c, so
511/(1 + (1 - cos(theta))) = 511/(2 + c); the acos/cos form is checked to
round the same way away from ties too.
"""
import numpy as np

MARGIN = 1e-9   # kErowMargin in sss_scatter.hip


def _geometry(rng, n):
    phi1, phi2 = rng.uniform(0, 2 * np.pi, n), rng.uniform(0, 2 * np.pi, n)
    det1 = np.stack([41 * np.cos(phi1), 41 * np.sin(phi1), rng.uniform(-12.8, 12.8, n)], 1)
    det2 = np.stack([41 * np.cos(phi2), 41 * np.sin(phi2), rng.uniform(-12.8, 12.8, n)], 1)
    sp = np.stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(-13, 13, n)], 1)
    # a tenth of the points close to the first detector (d1 from 0.01 cm)
    near = rng.random(n) < 0.1
    toward = det1 - sp
    sp[near] = det1[near] - toward[near] / np.linalg.norm(toward[near], axis=1)[:, None] * \
        rng.uniform(0.01, 2.0, near.sum())[:, None]
    d1 = np.sqrt(((det1 - sp) ** 2).sum(1))
    d2 = np.sqrt(((det2 - sp) ** 2).sum(1))
    dx = det1 - det2
    lor2 = (dx[:, 0] * dx[:, 0] + dx[:, 1] * dx[:, 1]) + dx[:, 2] * dx[:, 2]
    return d1, d2, lor2


def _exact(d1, d2, lor2):
    s1q, s2q = d1 * d1, d2 * d2
    c = np.clip((s1q + s2q - lor2) / (2.0 * np.sqrt(s1q * s2q)), -1.0, 1.0)
    return 511.0 / (2.0 + c) + 0.5, c


def _estimate(d1, d2, lor2, rng, ulps):
    eps = np.finfo(np.float64).eps

    def rcp(x):
        return (1.0 / x) * (1.0 + ulps * eps * rng.uniform(-1, 1, x.shape))

    den = (2.0 * d1) * d2
    num = (d1 * d1 + d2 * d2) - lor2
    c = np.clip(num * rcp(den), -1.0, 1.0)
    return 511.0 * rcp(2.0 + c) + 0.5


def test_estimate_is_far_inside_the_margin():
    rng = np.random.default_rng(7)
    d1, d2, lor2 = _geometry(rng, 2_000_000)
    v_exact, _ = _exact(d1, d2, lor2)
    v = _estimate(d1, d2, lor2, rng, ulps=2.0)
    worst = float(np.abs(v - v_exact).max())
    assert worst < MARGIN / 100, worst
    # away from the margin both floor to the same energy
    f = v - np.floor(v)
    far = (f > MARGIN) & (f < 1 - MARGIN)
    assert far.mean() > 0.999
    assert np.array_equal(np.floor(v[far]), np.floor(v_exact[far]))


def test_reference_angle_form_rounds_the_same_away_from_ties():
    rng = np.random.default_rng(8)
    d1, d2, lor2 = _geometry(rng, 1_000_000)
    v_exact, c = _exact(d1, d2, lor2)
    # CalcAngleScatter: theta = 180 - acos(c) 180/pi; E = 511/(1 + (1 - cos(theta pi/180)))
    theta = 180 - np.arccos(c) * (180.0 / np.pi)
    v_ref = 511.0 / (1 + (511.0 / 511.0) * (1 - np.cos(theta * (np.pi / 180.0)))) + 0.5
    assert np.abs(v_ref - v_exact).max() < 1e-9
    f = v_exact - np.floor(v_exact)
    far = (f > MARGIN) & (f < 1 - MARGIN)
    assert np.array_equal(np.floor(v_ref[far]), np.floor(v_exact[far]))


def test_a_planted_error_is_caught():
    # an estimate that drops the 1/2 before the floor would move every row;
    # one that uses d1 + d2 for the root would move many: both fail the bound
    rng = np.random.default_rng(9)
    d1, d2, lor2 = _geometry(rng, 100_000)
    v_exact, _ = _exact(d1, d2, lor2)
    c_bad = np.clip(((d1 * d1 + d2 * d2) - lor2) / (d1 + d2) ** 2 * 2, -1, 1)
    v_bad = 511.0 / (2.0 + c_bad) + 0.5
    assert np.abs(v_bad - v_exact).max() > MARGIN
    assert np.abs((v_exact - 0.5) - v_exact).max() > MARGIN
