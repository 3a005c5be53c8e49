// Fused single-scatter sum (the bodySSS LOR loop, SingleScatterSimulationTOF.py:433-557).
//
// Work decomposition: one wavefront owns an 8 x 8 tile of cells, rows
// (r1, i1) x columns (r2, i2), each a block of up to 4 rings x 2 detectors,
// for one chunk of scatter points (a few consecutive points of one
// normalisation slice); the (tile, chunk) items are listed per wave in an
// XCD-aware order built on the host.  Each lane accumulates its own cell
// over the chunk's scatter points and writes the chunk's sum and pair count;
// k_chunk_reduce then applies the per-slice normalisation (:555-557) and
// adds the slices.  No global atomics and a fixed reduction order: results
// do not depend on what else runs.
//
// The lanes of a wave take the chunk's scatter points in step; a lane whose
// pair the reference filters out (attenuation 0, no activity on either path,
// detection efficiency 0 -- about half of the pairs, mostly backscatter
// beyond the energy threshold) idles through that point's folds.
//
// Per (scatter point, cell) the lane evaluates, in float64: the distances,
// the Compton energy index floor(511/(2 + cos_arg) + 0.5) (the value of the
// reference's acos/degrees/cos round trip at :495-501) and the pair's TOF
// time shift; in R (float or double): Klein-Nishina, efficiency, solid angle,
// attenuation scaling and the TOF fold.
//
// The fold (:516-530) visits the voxel samples of both half paths.  The
// reference's integer picosecond offset floor((d1 - d2)/c - +2 s_k/c + 0.5)
// is formed in 32.32 fixed point (one 64-bit add, the high word is the floor)
// from the pair's shift and the sample's time, both rounded to 2^-32 ps.
// Each sample adds into the nearest TOF bin and its +-h neighbours with the
// normalised Gaussian weights (the columns of the reference's TOF table,
// evaluated analytically).  Along a half path the nearest bin changes only
// every few tens of samples, so the weights collect in a register window
// that is added into the lane's LDS column (one row per bin, padded by h on
// both sides) when the nearest bin changes.  The column holds the chunk's
// sums.
#include "sss_internal.h"

namespace sss {
namespace {

template <typename R>
__device__ __forceinline__ R fast_exp2(R x)
{
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_exp2f(x);
    else return exp2(x);
}

template <typename R>
__device__ __forceinline__ R fast_rcp(R x)
{
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_rcpf(x);
    else return 1.0 / x;
}

template <typename R>
__device__ __forceinline__ R clampr(R x, R lo, R hi)
{
    if constexpr (sizeof(R) == 4) return __builtin_amdgcn_fmed3f(x, lo, hi);
    else return x < lo ? lo : (x > hi ? hi : x);
}

// LDS add without a returned value (the lane owns its column: no conflicts).
template <typename R>
__device__ __forceinline__ void lds_add(R* p, R v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Cell of one lane of a tile: row lane/8 is (ring r1_0 + row >> k, detector
// i1_0 + row & (2^k - 1)), column lane%8 likewise for (r2, i2).  valid: the
// cell exists and is one the reference computes (D1 < D2, or D1 == D2 on
// different rings, SingleScatterSimulationTOF.py:466-469).  d1, d2: sampled
// detector index ring * NDU + detector, clamped into range for invalid lanes.
struct LaneCell {
    int r1, i1, r2, i2, d1, d2;
    bool valid;
};

__device__ __forceinline__ LaneCell lane_cell(const ScatterArgs& a, int4 tl, int lane)
{
    const int row = lane / kTileCols, cl = lane % kTileCols;
    const int k = a.tile_db_log2, m = (1 << k) - 1;
    LaneCell c;
    c.r1 = tl.x + (row >> k);
    c.i1 = tl.z + (row & m);
    c.r2 = tl.y + (cl >> k);
    c.i2 = tl.w + (cl & m);
    c.valid = c.r1 < a.nru && c.r2 < a.nru && c.i1 < a.ndu && c.i2 < a.ndu &&
              (c.i1 < c.i2 || (c.i1 == c.i2 && c.r1 != c.r2));
    c.d1 = min(c.r1, a.nru - 1) * a.ndu + min(c.i1, a.ndu - 1);
    c.d2 = min(c.r2, a.nru - 1) * a.ndu + min(c.i2, a.ndu - 1);
    return c;
}

// Energy row of a pair (law of cosines at the scatter point, FP64 in the
// reference's operation order, no fused multiply-add); carg, s1q, s2q are
// kept for the float terms.
__device__ __forceinline__ int energy_row(const ScatterArgs& a, double dist1, double dist2, double lor2,
                                          double& carg, double& s1q, double& s2q)
{
#pragma clang fp contract(off)
    s1q = dist1 * dist1;
    s2q = dist2 * dist2;
    carg = (s1q + s2q - lor2) / (2.0 * sqrt(s1q * s2q));
    carg = carg < -1.0 ? -1.0 : (carg > 1.0 ? 1.0 : carg);
    const double e_kev = floor(511.0 / (2.0 + carg) + 0.5);
    const int erow = (int)(e_kev * 2.0) - 1;
    return erow < 0 ? 0 : (erow >= a.n_energy ? a.n_energy - 1 : erow);
}

// Geometry and coefficients of one (scatter point, cell) pair.
template <typename R>
struct Pair {
    R coef1, coef2;          // G eff KN A1 A2^r, G eff KN A2 A1^r
    int64_t shift;           // ((d1 - d2)/c + 1/2) ps in 32.32 fixed point
    int n1, n2;              // samples folded on each path (0 when its activity is 0)
    int64_t b1, b2;          // first sample of each path
    R act1, act2;            // activity integrals (one-bin path)
    float counts;
};

// The reference's per-pair filters and scalar terms (:451-540).  Returns
// false for pairs it skips.
template <typename R>
__device__ __forceinline__ bool pair_eval(const ScatterArgs& a, int sp, int d1, int d2, double lor2,
                                          double eff_ref, Pair<R>& p)
{
    const int64_t per_sp = (int64_t)a.nru * a.ndu;
    const int64_t h1 = sp * per_sp + d1;
    const int64_t h2 = sp * per_sp + d2;
    const double att1 = a.hp_att[h1], att2 = a.hp_att[h2];
    const double act1 = a.hp_act[h1], act2 = a.hp_act[h2];
    if (att1 == 0.0 || att2 == 0.0 || (act1 == 0.0 && act2 == 0.0)) return false;
    const double dist1 = a.hp_dist[h1], dist2 = a.hp_dist[h2];
    // law of cosines at the scatter point, energy of the scattered photon
    // (the reference's operation order, no fused multiply-add)
    double carg, s1q, s2q;
    const int erow = energy_row(a, dist1, dist2, lor2, carg, s1q, s2q);
    const double eff = a.eff[erow] * eff_ref;
    if (eff == 0.0) return false;
    const R ratio = (R)a.ratio[erow];
    const R log2e = (R)1.4426950408889634;
    const R a1s = fast_exp2(-ratio * (R)a.hp_lnatt[h1] * log2e);
    const R a2s = fast_exp2(-ratio * (R)a.hp_lnatt[h2] * log2e);
    const R c = (R)carg;   // cos(angle at S) = -cos(scattering angle)
    const R rho = (R)1 / ((R)1 + (R)a.kn_gamma * ((R)1 + c));
    const R kn = rho * rho * (rho + (R)1 / rho - ((R)1 - c * c));
    const R geo = (R)a.g_const * (R)a.hp_cos[h1] * (R)a.hp_cos[h2] / ((R)s1q * (R)s2q);
    const R pre = geo * (R)eff * kn;
    p.coef1 = pre * (R)att1 * a2s;
    p.coef2 = pre * (R)att2 * a1s;
    p.shift = (int64_t)rint(((dist1 - dist2) * (1.0 / kLightCmPerPs) + 0.5) * kFixOne);
    p.n1 = act1 != 0.0 ? a.hp_n[h1] : 0;
    p.n2 = act2 != 0.0 ? a.hp_n[h2] : 0;
    p.b1 = a.hp_off[h1];
    p.b2 = a.hp_off[h2];
    p.act1 = (R)act1;
    p.act2 = (R)act2;
    p.counts = (act1 != 0.0 && act2 != 0.0) ? 2.0f : 1.0f;
    return true;
}

// energy_row from a quick float64 estimate: each quotient is a product with
// v_rcp_f64 refined by two Newton steps, and the root sqrt(s1q s2q) is
// d1 d2 (the numerator is energy_row's, bit for bit), so 511/(2 + c) + 1/2
// comes out within about 1e-12 of energy_row's value.  Where its fraction lies more than kErowMargin from an integer both
// floor to the same energy; the other lanes (near ties, a zero or non-finite
// distance) run energy_row's sequence.  The cosine kept for the float terms
// is the estimate's (a few float64 ulps from energy_row's).
constexpr double kErowMargin = 1e-9;

__device__ __forceinline__ double rcp_refined(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
}

__device__ __forceinline__ int energy_row_fast(const ScatterArgs& a, double dist1, double dist2, double lor2,
                                               double& carg, double& s1q, double& s2q)
{
    // the numerator rounds as energy_row's (no fused multiply-add)
#pragma clang fp contract(off)
    const double den = (2.0 * dist1) * dist2;
    const double num = (dist1 * dist1 + dist2 * dist2) - lor2;
    const double c = __builtin_fmin(__builtin_fmax(num * rcp_refined(den), -1.0), 1.0);
    const double v = 511.0 * rcp_refined(2.0 + c) + 0.5;
    const double k = __builtin_floor(v);
    const double f = v - k;
    if (!(den > 0.0 && den < 1e300 && f > kErowMargin && f < 1.0 - kErowMargin))
        return energy_row(a, dist1, dist2, lor2, carg, s1q, s2q);
    carg = c;
    s1q = dist1 * dist1;
    s2q = dist2 * dist2;
    const int erow = (int)k * 2 - 1;
    return erow < 0 ? 0 : (erow >= a.n_energy ? a.n_energy - 1 : erow);
}

// The float terms of pair_eval from the two records, the energy row's
// efficiency (times the 511 keV row's) and attenuation ratio.
__device__ __forceinline__ void pair_coef(const ScatterArgs& a, const PairHalf& r1, const PairHalf& r2,
                                          double carg, double s1q, double s2q, float eff, float ratio,
                                          Pair<float>& p)
{
    const float log2e = 1.4426950408889634f;
    const float a1s = fast_exp2(-ratio * r1.lnatt * log2e);
    const float a2s = fast_exp2(-ratio * r2.lnatt * log2e);
    const float c = (float)carg;
    // rho = 1 / q, so 1 / rho is q itself; the two quotients use the
    // hardware reciprocal (1 ulp), well inside the FP32 path's tolerance
    const float q = 1.0f + (float)a.kn_gamma * (1.0f + c);
    const float rho = fast_rcp(q);
    const float kn = rho * rho * (rho + q - (1.0f - c * c));
    const float geo = (float)a.g_const * r1.cosabs * r2.cosabs * fast_rcp((float)s1q * (float)s2q);
    const float pre = geo * eff * kn;
    p.coef1 = pre * r1.att * a2s;
    p.coef2 = pre * r2.att * a1s;
    p.counts = ((r1.n_flags & r2.n_flags) & kHalfActive) ? 2.0f : 1.0f;
}

// pair_eval for the FP32 fold, from the two half paths' PairHalf records
// (two 16-byte loads each, then the energy tables): the same filters and
// the same float terms.
__device__ __forceinline__ bool pair_terms_rec(const ScatterArgs& a, const PairHalf& r1, const PairHalf& r2,
                                               double lor2, double eff_ref, Pair<float>& p)
{
    const int32_t f = r1.n_flags | r2.n_flags;
    if ((f & kHalfAttZero) || !(f & kHalfActive)) return false;
    double carg, s1q, s2q;
    const int erow = energy_row_fast(a, r1.dist, r2.dist, lor2, carg, s1q, s2q);
    const double eff = a.eff[erow] * eff_ref;
    if (eff == 0.0) return false;
    pair_coef(a, r1, r2, carg, s1q, s2q, (float)eff, a.ratio[erow], p);
    p.shift = (int64_t)rint(((r1.dist - r2.dist) * (1.0 / kLightCmPerPs) + 0.5) * kFixOne);
    p.n1 = (r1.n_flags & kHalfActive) ? (r1.n_flags & kHalfCountMask) : 0;
    p.n2 = (r2.n_flags & kHalfActive) ? (r2.n_flags & kHalfCountMask) : 0;
    p.b1 = r1.off;
    p.b2 = r2.off;
    p.act1 = p.act2 = 0.0f;   // the one-bin path reads pair_eval's
    return true;
}

__device__ __forceinline__ bool pair_eval_rec(const ScatterArgs& a, int sp, int d1, int d2, double lor2,
                                              double eff_ref, Pair<float>& p)
{
    const int64_t per_sp = (int64_t)a.nru * a.ndu;
    return pair_terms_rec(a, a.hp_rec[sp * per_sp + d1], a.hp_rec[sp * per_sp + d2], lor2, eff_ref, p);
}

// Entry of the energy table in LDS for an energy row: rows kEnergyRowLo..
// kEnergyRowHi (clamped into the table) in order, then row 0, which only a
// NaN cosine reaches.
__device__ __forceinline__ int energy_slot(const ScatterArgs& a, int erow)
{
    return (erow >= a.etab_lo && erow <= a.etab_hi) ? erow - a.etab_lo : kEnergyRows - 1;
}

// pair_eval for one-bin subjects in FP32: records loaded by the caller, the
// energy terms from the LDS table (x: efficiency times the 511 keV row's
// efficiency in float, negative where that product is 0; y: the ratio).
__device__ __forceinline__ bool pair_eval_one_bin(const ScatterArgs& a, const float2* et, const PairHalf& r1,
                                                  const PairHalf& r2, double lor2, Pair<float>& p)
{
    const int32_t f = r1.n_flags | r2.n_flags;
    if ((f & kHalfAttZero) || !(f & kHalfActive)) return false;
    double carg, s1q, s2q;
    const int erow = energy_row_fast(a, r1.dist, r2.dist, lor2, carg, s1q, s2q);
    const float2 e = et[energy_slot(a, erow)];
    if (e.x < 0.0f) return false;
    pair_coef(a, r1, r2, carg, s1q, s2q, e.x, e.y, p);
    p.act1 = (float)r1.act;
    p.act2 = (float)r2.act;
    return true;
}

// TOF kernel constants in R (for float: rounded once on the host, so they
// stay in scalar registers).
template <typename R>
struct TofConst {
    R inv_w;      // 1/W: nearest bin j0 = rint(t / W + mid)
    R kap;        // kappa2
    R mid, jmax;  // (B-1)/2, B-1
    R wk, mid_wk; // W kappa2, mid W kappa2
    R c;          // 2^-beta2
    R inv_c;      // 2^beta2
};

template <typename R>
__device__ __forceinline__ TofConst<R> tof_const(const ScatterArgs& a)
{
    TofConst<R> d;
    if constexpr (sizeof(R) == 4) {
        d.inv_w = a.f_inv_w; d.kap = a.f_kap; d.mid = a.f_mid; d.jmax = a.f_jmax;
        d.wk = a.f_wk; d.mid_wk = a.f_mid_wk; d.c = a.f_c; d.inv_c = a.f_inv_c;
    } else {
        d.inv_w = a.inv_w_bins; d.kap = a.kappa2; d.mid = a.mid; d.jmax = (R)(a.n_bins - 1);
        d.wk = a.bin_w * a.kappa2; d.mid_wk = a.mid * a.bin_w * a.kappa2; d.c = a.c_beta;
        d.inv_c = 1.0 / a.c_beta;
    }
    return d;
}

// TOF weights of one voxel sample at offset tf ps:
// w[H + j] = wsample * T[j0 + j, t] for |j| <= h, zero outside the bins or
// the window; returns the window's centre bin j0 (as R).
//
// TM (TOF mode): kTofGeneral -- exp2 per window bin; kTofOneExp -- h = 1,
// one exp2, centre = nearest bin, edge bins masked; kTofOneExpInner -- h = 1,
// one exp2, centre clamped to [1, B-2] so all three window bins exist (no
// masks): when the nearest bin is an end bin the window still holds it and
// its neighbour, and a dropped bin is >= 2 bins from the nearest one, so the
// truncation stays within the window tolerance; the host enables it only
// when 2^(2 ku - beta2) stays finite for every offset of the subject.
template <typename R, int H, int TM>
__device__ __forceinline__ R tof_weights(const ScatterArgs& a, const TofConst<R>& d, R hf, R wsample,
                                         R (&w)[2 * H + 1])
{
    if constexpr (TM == kTofOneExpInner) {
        const R jf = clampr(rint(hf * d.inv_w + d.mid), (R)1, d.jmax - (R)1);
        const R ku = hf * d.kap + (d.mid_wk - jf * d.wk);
        const R E = fast_exp2(ku);
        const R cE = d.c * E;
        const R D = cE * E + (E + d.c);
        const R g = wsample * fast_rcp(D);
        w[1] = g * E;
        w[2] = w[1] * cE;
        w[0] = g * d.c;
        return jf;
    }
    const R jf = clampr(rint(hf * d.inv_w + d.mid), (R)0, d.jmax);
    // ku = kappa2 (t - (j0 - mid) W)
    R ku = hf * d.kap + (d.mid_wk - jf * d.wk);
    if constexpr (TM == kTofOneExp) {
        // h = 1: e+ = c E, e- = c / E with E = 2^ku, c = 2^-beta2; with
        // D = E (1 + e+ + e-) the weights are E/D, c E^2/D and c/D.
        ku = clampr(ku, (R)-100, (R)100);
        const R E = fast_exp2(ku);
        const R cp = (jf < d.jmax) ? d.c * E : (R)0;
        const R cm = (jf > (R)0) ? d.c : (R)0;
        const R D = cp * E + (E + cm);
        const R g = wsample * fast_rcp(D);
        w[1] = g * E;
        w[2] = w[1] * cp;
        w[0] = g * cm;
    } else {
        const int h = a.half_window;
        R z = (R)1;
        w[H] = (R)1;
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            R ep = (R)0, em = (R)0;
            if (j <= h) {
                const R dd = (R)(j * j) * (R)a.beta2;
                ep = (jf + (R)j <= d.jmax) ? fast_exp2((R)j * ku - dd) : (R)0;
                em = (jf - (R)j >= (R)0) ? fast_exp2(-(R)j * ku - dd) : (R)0;
            }
            w[H + j] = ep;
            w[H - j] = em;
            z += ep + em;
        }
        const R g = wsample * fast_rcp(z);
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) w[i] *= g;
    }
    return jf;
}

// Register window over the 2H+1 TOF bins around the current nearest bin; it
// is added into the lane's LDS column (rows b + H) when the nearest bin
// changes and at the end of a slice.
template <typename R, int H>
struct BinWindow {
    R v[2 * H + 1];
    R cur;          // nearest bin of the samples in the window

    __device__ __forceinline__ void reset(R j)
    {
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) v[i] = (R)0;
        cur = j;
    }
    __device__ __forceinline__ void spill(R* col) const
    {
        // window bins cur-H .. cur+H live in rows cur .. cur+2H
        R* base = col + (int)cur * kWave;
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) lds_add(base + i * kWave, v[i]);
    }
    __device__ __forceinline__ void add(R* col, R j0, const R (&w)[2 * H + 1])
    {
        if (j0 != cur) {
            spill(col);
            reset(j0);
        }
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) v[i] += w[i];
    }
};

// One voxel sample of a half path (see fold_path).
template <typename R, int H, int TM, bool RANGE, int SIGN>
__device__ __forceinline__ void fold_sample(const ScatterArgs& a, const TofConst<R>& d, R* __restrict__ col,
                                            BinWindow<R, H>& win, int64_t x, int64_t nq, R av, R coef)
{
    const int hi = (int)((x + nq) >> 32);   // high word: floor
    R ws = av * coef;
    R tf;
    if constexpr (RANGE) {
        // a zero-weight sample is moved to t_lo so its exponentials stay finite
        int t = SIGN < 0 ? hi : -hi;
        const bool ok = (uint32_t)(t - a.t_lo) <= (uint32_t)(a.t_hi - a.t_lo);
        ws = ok ? ws : (R)0;
        t = ok ? t : a.t_lo;
        tf = (R)t;
    } else {
        tf = SIGN < 0 ? (R)hi : -(R)hi;
    }
    R w[2 * H + 1];
    const R j0 = tof_weights<R, H, TM>(a, d, tf, ws, w);
    win.add(col, j0, w);
}

// Fold-table access: a half path's records start at a group boundary; group
// j of the path is three 16-byte words {nq0, nq1}, {nq2, nq3}, {a0..a3}.
__device__ __forceinline__ const uint4* fold_groups(const ScatterArgs& a, int64_t base)
{
    return reinterpret_cast<const uint4*>(a.s_fold) + (base / kSampleGroup) * 3;
}

__device__ __forceinline__ int64_t nq_lo(const uint4& q) { return (int64_t)(((uint64_t)q.y << 32) | q.x); }
__device__ __forceinline__ int64_t nq_hi(const uint4& q) { return (int64_t)(((uint64_t)q.w << 32) | q.z); }

// Fold the voxel samples of one half path.  SIGN = -1 (path to D1): offset
// t = floor(X - q) = hi32(shift + nq); SIGN = +1 (path to D2): t = floor(X + q)
// = -ceil(-X - q) = -hi32((2^32 - 1 - shift) + nq).  Whole groups of 4 (the
// padding records weigh 0); the next group is loaded before the current one
// is folded, and the loads run one group past the path (the table has spare
// records after its end).  With RANGE, offsets outside the non-zero TOF
// columns [t_lo, t_hi] weigh 0.
template <typename R, int H, int TM, bool RANGE, int SIGN>
__device__ __forceinline__ void fold_path(const ScatterArgs& a, const TofConst<R>& d, R* __restrict__ col,
                                          BinWindow<R, H>& win, int64_t base, int n, int64_t shift,
                                          R coef)
{
    const uint4* __restrict__ g = fold_groups(a, base);
    const double* __restrict__ ap = a.s_a64 + base;
    int64_t x = SIGN < 0 ? shift : ((int64_t)0xffffffffLL - shift);
    asm volatile("" : "+v"(x));   // keep x + nq as one 64-bit add per sample
    uint4 c0 = g[0], c1 = g[1], c2 = g[2];
    for (int k = n; k > 0; k -= kSampleGroup, g += 3, ap += kSampleGroup) {
        const uint4 q01 = c0, q23 = c1, av = c2;
        c0 = g[3];
        c1 = g[4];
        c2 = g[5];
        R a0, a1, a2, a3;
        if constexpr (sizeof(R) == 8) {
            a0 = ap[0]; a1 = ap[1]; a2 = ap[2]; a3 = ap[3];
        } else {
            a0 = __uint_as_float(av.x); a1 = __uint_as_float(av.y);
            a2 = __uint_as_float(av.z); a3 = __uint_as_float(av.w);
        }
        fold_sample<R, H, TM, RANGE, SIGN>(a, d, col, win, x, nq_lo(q01), a0, coef);
        fold_sample<R, H, TM, RANGE, SIGN>(a, d, col, win, x, nq_hi(q01), a1, coef);
        fold_sample<R, H, TM, RANGE, SIGN>(a, d, col, win, x, nq_lo(q23), a2, coef);
        fold_sample<R, H, TM, RANGE, SIGN>(a, d, col, win, x, nq_hi(q23), a3, coef);
    }
}

// Sliding window of the float h = 1 fold (kTofSlide) over the current half
// path, without the path's coefficient (applied when a bin leaves the
// window): bin cur - 1 holds c A, bin cur holds M and bin cur + 1 holds c P,
// c = 2^-beta2 (the side bins' common factor is applied when they are added
// into LDS or become the centre).
//
// The centre is the bin nearest to t, kept in [1, B-2]: bin b owns the
// integer offsets [L(b), L(b+1)), L(b) = ceil((b - mid - 1/2) W) (a tie may
// go either way: both nearest bins are in the window).  Along a half path t
// only moves one way (s_k never decreases), so the centre only changes once
// the high word of the fixed-point offset drops below one integer
// threshold: one integer compare per sample.  A centre's thresholds and ku
// offset are computed in registers when the centre moves (slide_thr_f,
// slide_kc_f): no LDS read, so a move waits for nothing (tables in LDS
// measured 2% slower).
struct SlideWindow {
    float A, M, P;
    int cur;       // centre bin
    int thr;       // the centre changes once hi < thr (INT_MIN: it cannot)
    float kc;      // kappa2 (mid - cur) W: ku = kappa2 t + kc
    float* wp;     // LDS row of bin cur - 1 in the lane's column
};


// Adds the window, times the path coefficient (ccoef = c coef), into its
// three LDS rows and clears it.
__device__ __forceinline__ void slide_flush(SlideWindow& win, float coef, float ccoef)
{
    lds_add(win.wp, win.A * ccoef);
    lds_add(win.wp + kWave, win.M * coef);
    lds_add(win.wp + 2 * kWave, win.P * ccoef);
    win.A = win.M = win.P = 0.0f;
}


// The inner window's thresholds and ku offsets from registers: L(b) =
// ceil((b - mid - 1/2) W) in float (within a float rounding of a tie the
// centre may go either way: both nearest bins are in the window), kc(b) =
// kappa2 (mid - b) W.
template <int SIGN, bool REL = false>
__device__ __forceinline__ int slide_thr_f(const ScatterArgs& a, int b)
{
    // centres: [1, B-2] for the inner window, [0, B-1] for the relative one
    constexpr int kLo = REL ? 0 : 1, kHiOff = REL ? 1 : 2;
    if constexpr (SIGN < 0) {
        return b <= kLo ? INT_MIN : (int)__builtin_ceilf(__builtin_fmaf((float)b, a.f_w, a.f_l0));
    } else {
        return b >= a.n_bins - kHiOff ? INT_MIN
                                      : 1 - (int)__builtin_ceilf(__builtin_fmaf((float)b, a.f_w, a.f_l1));
    }
}

__device__ __forceinline__ float slide_kc_f(const ScatterArgs& a, int b)
{
    // the constants' rounding parts added back: as accurate as rounding the
    // exact value once
    const float fb = (float)b;
    return __builtin_fmaf(-fb, a.f_wk, a.f_mid_wk) + __builtin_fmaf(-fb, a.f_wk_lo, a.f_mid_wk_lo);
}

// One sample.  ku = kappa2 (t - s_cur), E = 2^ku, D = c E^2 + E + c; the
// weights of bins cur - 1, cur, cur + 1 are a c/D, a E/D and a c E^2/D
// (tof_weights' kTofOneExpInner).  A centre move by one bin (nearly every
// move: a voxel step moves t by far less than a bin) adds the bin leaving the
// window into LDS and shifts the other two; a longer move (after a run of
// zero terms) flushes the window.
// av(): the sample's emission term, read after any centre move (the quad
// fold's DPP broadcast then sits next to the multiply that takes it).
template <int SIGN, typename AV>
__device__ __forceinline__ void slide_step(const ScatterArgs& a, const TofConst<float>& d, float* col,
                                           SlideWindow& win, int hi, AV av, float coef, float ccoef,
                                           float skap)
{
    if (hi < win.thr) {
        int b = win.cur + SIGN;
        int thr = slide_thr_f<SIGN>(a, b);
        if (hi >= thr) {
            if constexpr (SIGN < 0) {
                lds_add(win.wp + 2 * kWave, win.P * ccoef);
                win.P = win.M * d.inv_c;
                win.M = win.A * d.c;
                win.A = 0.0f;
                win.wp -= kWave;
            } else {
                lds_add(win.wp, win.A * ccoef);
                win.A = win.M * d.inv_c;
                win.M = win.P * d.c;
                win.P = 0.0f;
                win.wp += kWave;
            }
        } else {
            do {
                b += SIGN;
                thr = slide_thr_f<SIGN>(a, b);
            } while (hi < thr);
            slide_flush(win, coef, ccoef);
            win.wp = col + (b - 1 + a.slide_row0) * kWave;
        }
        win.cur = b;
        win.thr = thr;
        win.kc = slide_kc_f(a, b);
    }
    const float ku = __builtin_fmaf((float)hi, skap, win.kc);
    const float E = fast_exp2(ku);
    const float D = __builtin_fmaf(E, __builtin_fmaf(d.c, E, 1.0f), d.c);
    float g = av() * fast_rcp(D);
    // the term's one use: the quad fold's DPP broadcast folds into this multiply
    asm volatile("" : "+v"(g));
    const float gE = g * E;
    win.A += g;
    win.M += gE;
    win.P = __builtin_fmaf(gE, E, win.P);
}

template <int SIGN>
__device__ __forceinline__ void slide_sample(const ScatterArgs& a, const TofConst<float>& d,
                                             float* col, SlideWindow& win, int64_t x,
                                             int64_t nq, float av, float coef, float ccoef, float skap)
{
    slide_step<SIGN>(a, d, col, win, (int)((x + nq) >> 32), [av] { return av; }, coef, ccoef, skap);
}

// Centre bin of the path's first sample (offset t, high word hi): the float
// estimate, then the thresholds.
template <int SIGN>
__device__ __forceinline__ int slide_first(const ScatterArgs& a, int hi)
{
    const int t = SIGN < 0 ? hi : -hi;
    int b = (int)rintf((float)t * a.f_inv_w + a.f_mid);
    b = min(max(b, 1), a.n_bins - 2);
    // move the centre towards the path's direction of travel while t is past
    // its threshold, and back while the previous centre's threshold holds
    while (hi < slide_thr_f<SIGN>(a, b)) b += SIGN;
    while (b - SIGN >= 1 && b - SIGN <= a.n_bins - 2 && hi >= slide_thr_f<SIGN>(a, b - SIGN)) b -= SIGN;
    return b;
}

// fold_path for the sliding window.  The path's records come in whole
// groups of 4 (the padding records weigh 0), so the loop has no tests; two
// groups are in registers, the one being folded and the next, and the group
// after is loaded once one is done (loads run two groups past the path).
// The window is placed at the path's first sample and added into LDS at its
// end.
template <int SIGN>
__device__ __forceinline__ void slide_group(const ScatterArgs& a, const TofConst<float>& d,
                                            float* col, SlideWindow& win, int64_t x, const uint4& q01,
                                            const uint4& q23, const uint4& av, float coef, float ccoef, float skap)
{
    slide_sample<SIGN>(a, d, col, win, x, nq_lo(q01), __uint_as_float(av.x), coef, ccoef, skap);
    slide_sample<SIGN>(a, d, col, win, x, nq_hi(q01), __uint_as_float(av.y), coef, ccoef, skap);
    slide_sample<SIGN>(a, d, col, win, x, nq_lo(q23), __uint_as_float(av.z), coef, ccoef, skap);
    slide_sample<SIGN>(a, d, col, win, x, nq_hi(q23), __uint_as_float(av.w), coef, ccoef, skap);
}

template <int SIGN>
__device__ __forceinline__ void slide_path(const ScatterArgs& a, const TofConst<float>& d,
                                           float* __restrict__ col, SlideWindow& win, int64_t base, int n,
                                           int64_t shift, float coef)
{
    const uint4* __restrict__ g = fold_groups(a, base);
    int64_t x = SIGN < 0 ? shift : ((int64_t)0xffffffffLL - shift);
    asm volatile("" : "+v"(x));   // keep x + nq as one 64-bit add per sample
    const float skap = SIGN < 0 ? d.kap : -d.kap;
    const float ccoef = coef * d.c;
    uint4 c0 = g[0], c1 = g[1], c2 = g[2];
    uint4 n0 = g[3], n1 = g[4], n2 = g[5];
    {
        const int hi = (int)((x + nq_lo(c0)) >> 32);
        win.cur = slide_first<SIGN>(a, hi);
        win.thr = slide_thr_f<SIGN>(a, win.cur);
        win.kc = slide_kc_f(a, win.cur);
        win.wp = col + (win.cur - 1 + a.slide_row0) * kWave;
    }
    for (int k = n;; ) {
        slide_group<SIGN>(a, d, col, win, x, c0, c1, c2, coef, ccoef, skap);
        c0 = g[6]; c1 = g[7]; c2 = g[8];
        k -= kSampleGroup;
        g += 3;
        if (k <= 0) break;
        slide_group<SIGN>(a, d, col, win, x, n0, n1, n2, coef, ccoef, skap);
        n0 = g[6]; n1 = g[7]; n2 = g[8];
        k -= kSampleGroup;
        g += 3;
        if (k <= 0) break;
    }
    slide_flush(win, coef, ccoef);
}

// Quad-shared sliding window (kTofSlideQuad): slide_path's arithmetic with
// the fold records loaded once per quad.  The 8 lanes of a tile row read the
// same 48-byte groups of their row path; here the four lanes of a quad each
// load a third of every group (three of its 12 words, one 12-byte load
// instead of three 16-byte loads) and take the other words from each other
// by DPP quad broadcasts.  The four lanes of a quad must run the fold
// together: the caller keeps whole quads active through it (same path, same
// length).
struct QuadWords {
    uint32_t x, y, z;
};

// Word v of quad lane K, for every lane of the quad.
template <int K>
__device__ __forceinline__ uint32_t quad_bcast(uint32_t v)
{
    return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, K * 0x55, 0xf, 0xf, true);
}

// Or of a flag over the quad.
__device__ __forceinline__ int quad_any(int f)
{
    f |= __builtin_amdgcn_mov_dpp(f, 0xb1, 0xf, 0xf, true);   // quad_perm [1,0,3,2]
    f |= __builtin_amdgcn_mov_dpp(f, 0x4e, 0xf, 0xf, true);   // quad_perm [2,3,0,1]
    return f;
}

// High word of x + nq for the 32.32 offset nq = (qhi of quad lane KH, qlo
// of quad lane KL): the add with carry of the low words, then of the high
// words, each reading its quad lane's word through DPP (the compiler does
// not fold a DPP move into an add with carry).  The s_nop gives the two wait
// states a DPP read needs after a VALU write of its source.
template <int KL, int KH>
__device__ __forceinline__ int quad_hi(uint32_t xlo, uint32_t xhi, uint32_t qlo, uint32_t qhi)
{
    uint32_t lo, hi;
    asm volatile("s_nop 1\n\t"
                 "v_add_co_u32_dpp %0, vcc, %2, %3 quad_perm:[%6,%6,%6,%6] row_mask:0xf bank_mask:0xf bound_ctrl:1\n\t"
                 "v_addc_co_u32_dpp %1, vcc, %4, %5, vcc quad_perm:[%7,%7,%7,%7] row_mask:0xf bank_mask:0xf bound_ctrl:1"
                 : "=&v"(lo), "=v"(hi)
                 : "v"(qlo), "v"(xlo), "v"(qhi), "v"(xhi), "i"(KL), "i"(KH)
                 : "vcc");
    (void)lo;
    return (int)hi;
}

// Sample J of a group from the quad's words: lane j holds words 3j..3j+2 of
// {nq0 lo, nq0 hi, nq1 lo, nq1 hi, nq2 lo, nq2 hi, nq3 lo, nq3 hi, a0, a1, a2, a3}.
template <int SIGN, int J>
__device__ __forceinline__ void quad_sample(const ScatterArgs& a, const TofConst<float>& d, float* col,
                                            SlideWindow& win, uint32_t xlo, uint32_t xhi, const QuadWords& w,
                                            float coef, float ccoef, float skap)
{
    int hi;
    uint32_t aw;
    if constexpr (J == 0) {
        hi = quad_hi<0, 0>(xlo, xhi, w.x, w.y); aw = w.z;
    } else if constexpr (J == 1) {
        hi = quad_hi<0, 1>(xlo, xhi, w.z, w.x); aw = w.x;
    } else if constexpr (J == 2) {
        hi = quad_hi<1, 1>(xlo, xhi, w.y, w.z); aw = w.y;
    } else {
        hi = quad_hi<2, 2>(xlo, xhi, w.x, w.y); aw = w.z;
    }
    constexpr int KA = J == 0 ? 2 : 3;
    slide_step<SIGN>(a, d, col, win, hi, [aw] { return __uint_as_float(quad_bcast<KA>(aw)); }, coef, ccoef,
                     skap);
}

template <int SIGN>
__device__ __forceinline__ void quad_group(const ScatterArgs& a, const TofConst<float>& d, float* col,
                                           SlideWindow& win, uint32_t xlo, uint32_t xhi, const QuadWords& w,
                                           float coef, float ccoef, float skap)
{
    quad_sample<SIGN, 0>(a, d, col, win, xlo, xhi, w, coef, ccoef, skap);
    quad_sample<SIGN, 1>(a, d, col, win, xlo, xhi, w, coef, ccoef, skap);
    quad_sample<SIGN, 2>(a, d, col, win, xlo, xhi, w, coef, ccoef, skap);
    quad_sample<SIGN, 3>(a, d, col, win, xlo, xhi, w, coef, ccoef, skap);
}

// slide_path for a quad: qj = lane % 4.  Groups are kFoldGroupWords = 4
// QuadWords apart; two groups are in registers and the loads run two groups
// past the path, as in slide_path.
template <int SIGN>
__device__ __forceinline__ void quad_path(const ScatterArgs& a, const TofConst<float>& d,
                                          float* __restrict__ col, SlideWindow& win, int64_t base, int n,
                                          int64_t shift, float coef, int qj)
{
    constexpr int kStride = kFoldGroupWords / 3;
    const QuadWords* __restrict__ g =
        reinterpret_cast<const QuadWords*>(a.s_fold + (base / kSampleGroup) * kFoldGroupWords) + qj;
    const int64_t x = SIGN < 0 ? shift : ((int64_t)0xffffffffLL - shift);
    const uint32_t xlo = (uint32_t)x, xhi = (uint32_t)((uint64_t)x >> 32);
    const float skap = SIGN < 0 ? d.kap : -d.kap;
    const float ccoef = coef * d.c;
    QuadWords c = g[0];
    QuadWords nx = g[kStride];
    {
        const int hi = quad_hi<0, 0>(xlo, xhi, c.x, c.y);
        win.cur = slide_first<SIGN>(a, hi);
        win.thr = slide_thr_f<SIGN>(a, win.cur);
        win.kc = slide_kc_f(a, win.cur);
        win.wp = col + (win.cur - 1 + a.slide_row0) * kWave;
    }
    for (int k = n;; ) {
        quad_group<SIGN>(a, d, col, win, xlo, xhi, c, coef, ccoef, skap);
        c = g[2 * kStride];
        k -= kSampleGroup;
        g += kStride;
        if (k <= 0) break;
        quad_group<SIGN>(a, d, col, win, xlo, xhi, nx, coef, ccoef, skap);
        nx = g[2 * kStride];
        k -= kSampleGroup;
        g += kStride;
        if (k <= 0) break;
    }
    slide_flush(win, coef, ccoef);
}

// One scatter point of a kTofSlideQuad wave item.  Every lane loads its
// cell's two records (clamped detector indices keep invalid cells' loads in
// bounds) and evaluates its pair.  The row path is folded by quads: the four
// lanes of a quad share the cell row, so they fold one half path, and lanes
// whose pair the reference skips fold with coefficient 0 and shift 0 so that
// the quad stays whole.  The column path (8 lanes of a tile column, 8 apart,
// which DPP quad broadcasts cannot join) is folded per lane as slide_path
// does it.  Sharing the loads in one of the two folds balances the kernel:
// with per-lane loads in both it is bound by the texture address and data
// path; with quads in both (the column fold on the transposed cells) it is
// bound by the VALU, which the DPP adds with carry cost one instruction a
// sample more (38.2 and 38.8 ms at config 2, 36.3 ms as here).
__device__ __forceinline__ void quad_point(const ScatterArgs& a, const TofConst<float>& tc, float* col,
                                           SlideWindow& sl, int sp, int d1, int d2, double lor2, double eff_ref,
                                           bool valid, int qj, float& cnt)
{
    const int64_t per_sp = (int64_t)a.nru * a.ndu;
    const PairHalf r1 = a.hp_rec[sp * per_sp + d1];
    const PairHalf r2 = a.hp_rec[sp * per_sp + d2];
    Pair<float> p;
    bool found = false;
    if (valid) found = pair_terms_rec(a, r1, r2, lor2, eff_ref, p);
    if (found) cnt += p.counts;
    const int n1 = (r1.n_flags & kHalfActive) ? (r1.n_flags & kHalfCountMask) : 0;
    if (n1 > 0 && quad_any(found ? 1 : 0))
        quad_path<-1>(a, tc, col, sl, r1.off, n1, found ? p.shift : 0, found ? p.coef1 : 0.0f, qj);
    if (found && p.n2 > 0) slide_path<+1>(a, tc, col, sl, p.b2, p.n2, p.shift, p.coef2);
}

// Sliding window in relative form (kTofSlideRel, float, h = H in {1, 2}).
// The centre is the nearest bin, b in [0, B-1] (thresholds as above, INT_MIN
// only at the two ends).  Relative to the centre the weights of bins
// cur + d are 2^(d ku - d^2 beta2), ku = kappa2 (t - s_cur); bins past the
// ends are masked by per-centre factors.  One exp2 per sample in each form
// (see rel_sample): h = 1 with beta2 <= 95 and h = 2 with beta2 < 30 scale
// the weights to keep every term a normal float; h = 1 with wider bins
// keeps only the side that ku points to.  The window holds
// the weights without the path coefficient, which is applied when a bin
// leaves the window.
template <int H>
struct RelWindow {
    float A[2 * H + 1];   // bins cur - H .. cur + H
    int cur;
    int thr;              // the centre changes once hi < thr
    float kc;             // ku = skap hi + kc
    float mp1, mm1;       // 1 where bin cur +- 1 exists, else 0 (h = 1, wide bins)
    float cp1, cm1;       // 2^-beta2 where bin cur +- 1 exists, else 0
    float c4p, c4m;       // 2^(-4 beta2) where bin cur +- 2 exists, else 0 (h = 2)
    float* wp;            // LDS row of bin cur - H
};

template <int H, bool WIDE>
__device__ __forceinline__ void rel_centre(const ScatterArgs& a, RelWindow<H>& w, int b)
{
    w.cur = b;
    const float kc = slide_kc_f(a, b);
    if constexpr (WIDE) {
        w.kc = kc;
        w.mp1 = b + 1 < a.n_bins ? 1.0f : 0.0f;
        w.mm1 = b > 0 ? 1.0f : 0.0f;
    } else {
        w.kc = kc;
        w.cp1 = b + 1 < a.n_bins ? a.f_c : 0.0f;
        w.cm1 = b > 0 ? a.f_c : 0.0f;
        if constexpr (H == 2) {
            w.c4p = b + 2 < a.n_bins ? a.f_c4 : 0.0f;
            w.c4m = b > 1 ? a.f_c4 : 0.0f;
        }
    }
}

template <int H>
__device__ __forceinline__ void rel_flush(RelWindow<H>& w, float coef)
{
#pragma unroll
    for (int i = 0; i < 2 * H + 1; ++i) {
        lds_add(w.wp + i * kWave, w.A[i] * coef);
        w.A[i] = 0.0f;
    }
}

// av(): the sample's emission term, read after any centre move (as slide_step).
template <int H, int SIGN, bool WIDE, typename AV>
__device__ __forceinline__ void rel_step(const ScatterArgs& a, float* col, RelWindow<H>& w, int hi, AV av,
                                         float coef, float skap)
{
    if (hi < w.thr) {
        int b = w.cur + SIGN;
        int thr = slide_thr_f<SIGN, true>(a, b);
        if (hi >= thr) {
            // a move by one bin: the bin leaving the window goes into LDS
            if constexpr (SIGN < 0) {
                lds_add(w.wp + 2 * H * kWave, w.A[2 * H] * coef);
#pragma unroll
                for (int i = 2 * H; i > 0; --i) w.A[i] = w.A[i - 1];
                w.A[0] = 0.0f;
                w.wp -= kWave;
            } else {
                lds_add(w.wp, w.A[0] * coef);
#pragma unroll
                for (int i = 0; i < 2 * H; ++i) w.A[i] = w.A[i + 1];
                w.A[2 * H] = 0.0f;
                w.wp += kWave;
            }
        } else {
            do {
                b += SIGN;
                thr = slide_thr_f<SIGN, true>(a, b);
            } while (hi < thr);
            rel_flush(w, coef);
            w.wp = col + (b - H + a.slide_row0) * kWave;
        }
        w.thr = thr;
        rel_centre<H, WIDE>(a, w, b);
    }
    const float fh = (float)hi;
    if constexpr (WIDE) {
        // beta2 > 95: the two side weights multiply to 2^(-2 beta2), so only
        // the side ku points to can matter; the other weighs under
        // 2^-beta2 < 2^-95 of the centre and is dropped.  One exp2:
        // 2^(|ku| - beta2), capped before the end-bin mask.
        const float ku = __builtin_fmaf(fh, skap, w.kc);
        const bool up = ku > 0.0f;
        const float t = fast_exp2(fminf(__builtin_fabsf(ku) - a.f_beta2, 30.0f)) * (up ? w.mp1 : w.mm1);
        float g = av() * fast_rcp(1.0f + t);
        asm volatile("" : "+v"(g));
        const float gt = g * t;
        w.A[1] += g;
        w.A[2] += up ? gt : 0.0f;
        w.A[0] += up ? 0.0f : gt;
    } else if constexpr (H == 1) {
        // one exp2: the weights c, E, c E^2 over D = c + E + c E^2 (E = 2^ku,
        // c = 2^-beta2); with the nearest-bin centre |ku| <= beta2, and
        // beta2 <= 95 keeps every term a normal float; past the end bins ku
        // is held to +-(beta2 + 30), where the far side weighs under
        // 2^-(2 beta2 + 30)
        const float ku = __builtin_amdgcn_fmed3f(__builtin_fmaf(fh, skap, w.kc), -a.f_kclamp1, a.f_kclamp1);
        const float E = fast_exp2(ku);
        const float tp = (w.cp1 * E) * E;
        const float D = (w.cm1 + E) + tp;
        float g = av() * fast_rcp(D);
        asm volatile("" : "+v"(g));
        w.A[0] = __builtin_fmaf(g, w.cm1, w.A[0]);
        w.A[1] = __builtin_fmaf(g, E, w.A[1]);
        w.A[2] = __builtin_fmaf(g, tp, w.A[2]);
    } else {
        // h = 2 with one exp2: the weights 2^(d ku - d^2 beta2) times E^2,
        // E = 2^ku: c^4, c E, E^2, c E^3, c^4 E^4 (c = 2^-beta2).  With the
        // nearest-bin centre |ku| <= beta2, and beta2 < 30 keeps every term
        // a normal float; past the end bins ku is held to +-(beta2 + 10),
        // where the far side's terms weigh under 2^-(2 beta2 + 10).
        // D by Horner in E, the weights g E^k by a chain of products (the
        // masked constants multiply last, so a masked side never forms
        // g E^4 unscaled: every factor stays finite)
        const float ku = __builtin_amdgcn_fmed3f(__builtin_fmaf(fh, skap, w.kc), -a.f_kclamp, a.f_kclamp);
        const float E = fast_exp2(ku);
        const float D = __builtin_fmaf(
            E, __builtin_fmaf(E, __builtin_fmaf(E, __builtin_fmaf(E, w.c4p, w.cp1), 1.0f), w.cm1), w.c4m);
        const float g = av() * fast_rcp(D);
        const float g1 = g * E;
        const float g2 = g1 * E;
        const float g3 = g2 * E;
        w.A[0] = __builtin_fmaf(g, w.c4m, w.A[0]);
        w.A[1] = __builtin_fmaf(g1, w.cm1, w.A[1]);
        w.A[2] += g2;
        w.A[3] = __builtin_fmaf(g3, w.cp1, w.A[3]);
        w.A[4] = __builtin_fmaf(g3, E * w.c4p, w.A[4]);
    }
}

template <int H, int SIGN, bool WIDE>
__device__ __forceinline__ void rel_sample(const ScatterArgs& a, float* col,
                                           RelWindow<H>& w, int64_t x, int64_t nq, float av, float coef,
                                           float skap)
{
    rel_step<H, SIGN, WIDE>(a, col, w, (int)((x + nq) >> 32), [av] { return av; }, coef, skap);
}

// Nearest bin of the path's first sample: the float estimate, then the
// thresholds (INT_MIN at the ends stops both loops).
template <int SIGN>
__device__ __forceinline__ int rel_first(const ScatterArgs& a, int hi)
{
    const int t = SIGN < 0 ? hi : -hi;
    int b = (int)rintf((float)t * a.f_inv_w + a.f_mid);
    b = min(max(b, 0), a.n_bins - 1);
    while (hi < slide_thr_f<SIGN, true>(a, b)) b += SIGN;
    while (b - SIGN >= 0 && b - SIGN <= a.n_bins - 1 && hi >= slide_thr_f<SIGN, true>(a, b - SIGN)) b -= SIGN;
    return b;
}

template <int H, int SIGN, bool WIDE>
__device__ __forceinline__ void rel_group(const ScatterArgs& a, float* col,
                                          RelWindow<H>& w, int64_t x, const uint4& q01, const uint4& q23,
                                          const uint4& av, float coef, float skap)
{
    rel_sample<H, SIGN, WIDE>(a, col, w, x, nq_lo(q01), __uint_as_float(av.x), coef, skap);
    rel_sample<H, SIGN, WIDE>(a, col, w, x, nq_hi(q01), __uint_as_float(av.y), coef, skap);
    rel_sample<H, SIGN, WIDE>(a, col, w, x, nq_lo(q23), __uint_as_float(av.z), coef, skap);
    rel_sample<H, SIGN, WIDE>(a, col, w, x, nq_hi(q23), __uint_as_float(av.w), coef, skap);
}

// The window start of a relative-window path at high word hi.
template <int H, int SIGN, bool WIDE>
__device__ __forceinline__ void rel_start(const ScatterArgs& a, float* col, RelWindow<H>& w, int hi)
{
    const int b = rel_first<SIGN>(a, hi);
    w.thr = slide_thr_f<SIGN, true>(a, b);
    w.wp = col + (b - H + a.slide_row0) * kWave;
    rel_centre<H, WIDE>(a, w, b);
#pragma unroll
    for (int i = 0; i < 2 * H + 1; ++i) w.A[i] = 0.0f;
}

// fold_path for kTofSlideRel: groups of 4 records, two groups in registers
// (as slide_path).
template <int H, int SIGN, bool WIDE>
__device__ __forceinline__ void rel_path(const ScatterArgs& a, float* __restrict__ col,
                                         RelWindow<H>& w, int64_t base, int n, int64_t shift, float coef)
{
    const uint4* __restrict__ g = fold_groups(a, base);
    int64_t x = SIGN < 0 ? shift : ((int64_t)0xffffffffLL - shift);
    asm volatile("" : "+v"(x));   // keep x + nq as one 64-bit add per sample
    const float skap = SIGN < 0 ? a.f_kap : -a.f_kap;
    uint4 c0 = g[0], c1 = g[1], c2 = g[2];
    uint4 n0 = g[3], n1 = g[4], n2 = g[5];
    rel_start<H, SIGN, WIDE>(a, col, w, (int)((x + nq_lo(c0)) >> 32));
    for (int k = n;; ) {
        rel_group<H, SIGN, WIDE>(a, col, w, x, c0, c1, c2, coef, skap);
        c0 = g[6]; c1 = g[7]; c2 = g[8];
        k -= kSampleGroup;
        g += 3;
        if (k <= 0) break;
        rel_group<H, SIGN, WIDE>(a, col, w, x, n0, n1, n2, coef, skap);
        n0 = g[6]; n1 = g[7]; n2 = g[8];
        k -= kSampleGroup;
        g += 3;
        if (k <= 0) break;
    }
    rel_flush(w, coef);
}

// rel_path for a quad (quad_path's loads and broadcasts, rel_step's window).
template <int H, int SIGN, bool WIDE, int J>
__device__ __forceinline__ void quad_rel_sample(const ScatterArgs& a, float* col, RelWindow<H>& w, uint32_t xlo,
                                                uint32_t xhi, const QuadWords& q, float coef, float skap)
{
    int hi;
    uint32_t aw;
    if constexpr (J == 0) {
        hi = quad_hi<0, 0>(xlo, xhi, q.x, q.y); aw = q.z;
    } else if constexpr (J == 1) {
        hi = quad_hi<0, 1>(xlo, xhi, q.z, q.x); aw = q.x;
    } else if constexpr (J == 2) {
        hi = quad_hi<1, 1>(xlo, xhi, q.y, q.z); aw = q.y;
    } else {
        hi = quad_hi<2, 2>(xlo, xhi, q.x, q.y); aw = q.z;
    }
    constexpr int KA = J == 0 ? 2 : 3;
    rel_step<H, SIGN, WIDE>(a, col, w, hi, [aw] { return __uint_as_float(quad_bcast<KA>(aw)); }, coef, skap);
}

template <int H, int SIGN, bool WIDE>
__device__ __forceinline__ void quad_rel_group(const ScatterArgs& a, float* col, RelWindow<H>& w, uint32_t xlo,
                                               uint32_t xhi, const QuadWords& q, float coef, float skap)
{
    quad_rel_sample<H, SIGN, WIDE, 0>(a, col, w, xlo, xhi, q, coef, skap);
    quad_rel_sample<H, SIGN, WIDE, 1>(a, col, w, xlo, xhi, q, coef, skap);
    quad_rel_sample<H, SIGN, WIDE, 2>(a, col, w, xlo, xhi, q, coef, skap);
    quad_rel_sample<H, SIGN, WIDE, 3>(a, col, w, xlo, xhi, q, coef, skap);
}

template <int H, int SIGN, bool WIDE>
__device__ __forceinline__ void quad_rel_path(const ScatterArgs& a, float* __restrict__ col, RelWindow<H>& w,
                                              int64_t base, int n, int64_t shift, float coef, int qj)
{
    constexpr int kStride = kFoldGroupWords / 3;
    const QuadWords* __restrict__ g =
        reinterpret_cast<const QuadWords*>(a.s_fold + (base / kSampleGroup) * kFoldGroupWords) + qj;
    const int64_t x = SIGN < 0 ? shift : ((int64_t)0xffffffffLL - shift);
    const uint32_t xlo = (uint32_t)x, xhi = (uint32_t)((uint64_t)x >> 32);
    const float skap = SIGN < 0 ? a.f_kap : -a.f_kap;
    QuadWords c = g[0];
    QuadWords nx = g[kStride];
    rel_start<H, SIGN, WIDE>(a, col, w, quad_hi<0, 0>(xlo, xhi, c.x, c.y));
    for (int k = n;; ) {
        quad_rel_group<H, SIGN, WIDE>(a, col, w, xlo, xhi, c, coef, skap);
        c = g[2 * kStride];
        k -= kSampleGroup;
        g += kStride;
        if (k <= 0) break;
        quad_rel_group<H, SIGN, WIDE>(a, col, w, xlo, xhi, nx, coef, skap);
        nx = g[2 * kStride];
        k -= kSampleGroup;
        g += kStride;
        if (k <= 0) break;
    }
    rel_flush(w, coef);
}

// quad_point for the relative windows: the row path on quads, the column
// path per lane.
template <int H, bool WIDE>
__device__ __forceinline__ void quad_rel_point(const ScatterArgs& a, float* col, RelWindow<H>& w, int sp,
                                               int d1, int d2, double lor2, double eff_ref, bool valid, int qj,
                                               float& cnt)
{
    const int64_t per_sp = (int64_t)a.nru * a.ndu;
    const PairHalf r1 = a.hp_rec[sp * per_sp + d1];
    const PairHalf r2 = a.hp_rec[sp * per_sp + d2];
    Pair<float> p;
    bool found = false;
    if (valid) found = pair_terms_rec(a, r1, r2, lor2, eff_ref, p);
    if (found) cnt += p.counts;
    const int n1 = (r1.n_flags & kHalfActive) ? (r1.n_flags & kHalfCountMask) : 0;
    if (n1 > 0 && quad_any(found ? 1 : 0))
        quad_rel_path<H, -1, WIDE>(a, col, w, r1.off, n1, found ? p.shift : 0, found ? p.coef1 : 0.0f, qj);
    if (found && p.n2 > 0) rel_path<H, +1, WIDE>(a, col, w, p.b2, p.n2, p.shift, p.coef2);
}

// One wave's work item (tile, chunk) of the scatter sum.  LDS: one column
// per lane; the one-bin FP32 path reads the energy table, which the caller
// has placed at the start of LDS.  Padding items (tile -1) return at once.
template <typename R, int HM, int TM, bool TRIVIAL, bool RANGE>
__device__ __forceinline__ void scatter_item(const ScatterArgs& a, unsigned char* smem_raw, int wave,
                                             int lane, int2 item)
{
    constexpr bool QUAD = TM == kTofSlideQuad;
    constexpr bool SLIDE = QUAD;
    constexpr bool REL = TM == kTofSlideRel || TM == kTofSlideRelWide;
    constexpr bool WIDE = TM == kTofSlideRelWide;
    static_assert(!SLIDE || (sizeof(R) == 4 && HM == 1 && !RANGE), "sliding window: float, h = 1, no range test");
    static_assert(!REL || (sizeof(R) == 4 && (HM == 1 || HM == 2) && !RANGE && (!WIDE || HM == 1)),
                  "relative sliding window: float, h = 1 or 2 (wide: h = 1), no range test");
    constexpr bool ONE_BIN_REC = TRIVIAL && sizeof(R) == 4;
    const float2* et = reinterpret_cast<const float2*>(smem_raw);
    if (item.x < 0) return;
    const int tile = item.x;
    const int chunk = item.y;
    const int B = a.n_bins;
    const int rows = B + 2 * HM;   // bins -HM .. B-1+HM (the window's reach)
    R* col = reinterpret_cast<R*>(smem_raw) +
             (size_t)wave * rows * kWave + lane;
    if constexpr (!TRIVIAL) {
        for (int r = 0; r < rows; ++r) col[r * kWave] = (R)0;
    }

    const LaneCell lc = lane_cell(a, a.tiles[tile], lane);
    const int d1 = lc.d1, d2 = lc.d2;
    const bool valid = lc.valid;
    double lor2;
    {
        const double dx = a.detpos[3 * d1 + 0] - a.detpos[3 * d2 + 0];
        const double dy = a.detpos[3 * d1 + 1] - a.detpos[3 * d2 + 1];
        const double dz = a.detpos[3 * d1 + 2] - a.detpos[3 * d2 + 2];
        lor2 = ((dx * dx) + (dy * dy)) + (dz * dz);
    }
    const double eff_ref = a.eff[a.eff_ref_row];
    const TofConst<R> tc = tof_const<R>(a);

    R num = (R)0;   // one-bin path
    float cnt = 0.0f;
    BinWindow<R, HM> win;
    win.reset((R)0);
    SlideWindow sl;
    sl.A = sl.M = sl.P = 0.0f;
    RelWindow<(HM == 2 ? 2 : 1)> rw;
    // The lanes take the chunk's scatter points in step: the lanes of a tile
    // row (column) then read one half path at the same time, and neighbouring
    // cells mostly keep or drop the same points, so fewer lane slots idle
    // than with a cursor per lane (a CPU model of config 2: 89% busy against
    // 86%).
    const int p_begin = a.chunk_sp[chunk];
    const int p_end = a.chunk_sp[chunk + 1];
    if constexpr (ONE_BIN_REC) {
        // The pair terms are short, so the records of the next two points
        // are in flight while one pair is evaluated: three record pairs
        // take turns (unrolled by three, so no register copy waits for a
        // load), and the energy terms come from LDS (no global load queues
        // behind the records).  Clamped detector indices keep the loads of
        // invalid lanes in bounds, and loads past the chunk read its last
        // point again.  The points are loaded in order, so the record row is
        // a scalar pointer stepped once per load (the lanes add their
        // detectors' 32-bit byte offsets).
        const int64_t row_bytes = (int64_t)a.nru * a.ndu * (int64_t)sizeof(PairHalf);
        const int last = p_end - 1;
        const unsigned char* row = reinterpret_cast<const unsigned char*>(a.hp_rec) + p_begin * row_bytes;
        int row_sp = p_begin;
        const uint32_t o1 = (uint32_t)d1 * (uint32_t)sizeof(PairHalf);
        const uint32_t o2 = (uint32_t)d2 * (uint32_t)sizeof(PairHalf);
        auto load = [&](int, PairHalf& r1, PairHalf& r2) {
            r1 = *reinterpret_cast<const PairHalf*>(row + o1);
            r2 = *reinterpret_cast<const PairHalf*>(row + o2);
            if (row_sp < last) row += row_bytes;
            ++row_sp;
        };
        auto eval = [&](int sp, const PairHalf& r1, const PairHalf& r2) {
            Pair<float> p;
            if (valid && sp < p_end && pair_eval_one_bin(a, et, r1, r2, lor2, p)) {
                cnt += p.counts;
                num += p.coef1 * p.act1 + p.coef2 * p.act2;
            }
        };
        if (p_begin < p_end) {
            PairHalf u1, u2, v1, v2, w1, w2;
            load(p_begin, u1, u2);
            load(p_begin + 1, v1, v2);
            for (int sp = p_begin; sp < p_end; sp += 3) {
                load(sp + 2, w1, w2);
                eval(sp, u1, u2);
                load(sp + 3, u1, u2);
                eval(sp + 1, v1, v2);
                load(sp + 4, v1, v2);
                eval(sp + 2, w1, w2);
            }
        }
    }
    if constexpr (QUAD) {
        // every lane takes every point (whole quads through the row fold)
        for (int sp = p_begin; sp < p_end; ++sp)
            quad_point(a, tc, col, sl, sp, d1, d2, lor2, eff_ref, valid, lane % 4, cnt);
    }
    // h = 1 relative windows fold the row path on quads; the h = 2 window's
    // sample costs more VALU than the quad's offset broadcast saves in loads
    // (random instances with 17-21 bins 3-6% slower on quads, 5-11 bins 2-7%
    // faster)
    constexpr bool QREL = REL && HM == 1;
    if constexpr (QREL) {
        // every lane takes every point (whole quads through the row fold)
        for (int sp = p_begin; sp < p_end; ++sp)
            quad_rel_point<1, WIDE>(a, col, rw, sp, d1, d2, lor2, eff_ref, valid, lane % 4, cnt);
    }
    if constexpr (!ONE_BIN_REC && !QUAD && !QREL) {
        for (int sp = p_begin; sp < p_end; ++sp) {
            Pair<R> p;
            bool found = false;
            if (valid) {
                if constexpr (sizeof(R) == 4 && !TRIVIAL) found = pair_eval_rec(a, sp, d1, d2, lor2, eff_ref, p);
                else found = pair_eval<R>(a, sp, d1, d2, lor2, eff_ref, p);
            }
            if (found) {
                cnt += p.counts;
                if constexpr (TRIVIAL) {
                    num += p.coef1 * p.act1 + p.coef2 * p.act2;
                } else if constexpr (REL) {
                    if (p.n1 > 0) rel_path<HM, -1, WIDE>(a, col, rw, p.b1, p.n1, p.shift, p.coef1);
                    if (p.n2 > 0) rel_path<HM, +1, WIDE>(a, col, rw, p.b2, p.n2, p.shift, p.coef2);
                } else if (a.t_hi >= a.t_lo) {   // else every TOF weight is zero
                    fold_path<R, HM, TM, RANGE, -1>(a, tc, col, win, p.b1, p.n1, p.shift, p.coef1);
                    fold_path<R, HM, TM, RANGE, +1>(a, tc, col, win, p.b2, p.n2, p.shift, p.coef2);
                }
            }
        }
    }
    // the chunk's sums and pair count into the lane's slot (the reduce maps
    // slots to cells); window reach rows (bins beyond the ends) are dropped
    if constexpr (!TRIVIAL && !SLIDE && !REL) win.spill(col);
    if (!valid) return;
    const int64_t n_slots = (int64_t)a.n_tiles * kWave;
    const int64_t slot = (int64_t)tile * kWave + lane;
    R* dst = reinterpret_cast<R*>(a.part_sum) + (int64_t)chunk * B * n_slots + slot;
    if constexpr (TRIVIAL) {
        dst[0] = num;
    } else {
        for (int b = 0; b < B; ++b) dst[b * n_slots] = col[(b + HM) * kWave];
    }
    a.part_cnt[(int64_t)chunk * n_slots + slot] = cnt;
}

// At least 6 waves per SIMD (<= 80 VGPRs): the values that do not fit are
// spilled outside the sample loops (per slice and per pair only).  Measured
// on config 2: 4 waves 240 ms, 5 waves 203 ms, 6 waves 189 ms.  One work
// item per wave.
template <typename R, int HM, int TM, bool TRIVIAL, bool RANGE>
__global__ __launch_bounds__(kScatterWaves * kWave, 6) void k_scatter(ScatterArgs a)
{
    // 16-byte aligned: the columns hold doubles (8-byte LDS atomics) and
    // the energy table float2
    extern __shared__ __align__(16) unsigned char smem_raw[];
    // the wave index is uniform: held in a scalar register, so the work item,
    // chunk and scatter-point loop are scalar too
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int lane = threadIdx.x % kWave;
    if constexpr (TRIVIAL && sizeof(R) == 4) {
        float2* et = reinterpret_cast<float2*>(smem_raw);
        for (int i = threadIdx.x; i < kEnergyRows; i += blockDim.x) et[i] = a.etab[i];
        __syncthreads();   // every wave of the workgroup, before any exits
    }
    scatter_item<R, HM, TM, TRIVIAL, RANGE>(a, smem_raw, wave, lane, a.items[blockIdx.x * kScatterWaves + wave]);
}

// Read-only device data that every lane of a wave reads at one address:
// through the constant address space, so the loads are scalar (s_load) and
// the values sit in scalar registers, as kernel arguments do.
template <typename T>
using ConstPtr = const __attribute__((address_space(4))) T*;

template <typename T>
__device__ __forceinline__ ConstPtr<T> as_const(const T* p)
{
    return (ConstPtr<T>)p;
}

// p[i] read through the constant address space, one dword at a time (the
// copy constructors take generic references).
template <typename T>
__device__ __forceinline__ T load_const(const T* p, int64_t i)
{
    static_assert(sizeof(T) % 4 == 0, "whole dwords");
    T v;
    ConstPtr<uint32_t> src = as_const(reinterpret_cast<const uint32_t*>(p + i));
    uint32_t* dst = reinterpret_cast<uint32_t*>(&v);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) dst[k] = src[k];
    return v;
}

// Several subjects' scatter sums in one persistent launch (kernel variant
// given by the template; every member of the launch uses it).  The work is
// whole workgroup blocks of the members' own item lists (kScatterWaves waves
// of one subject), in 8 queues, one per XCD: block b of a member's list was
// laid out for XCD b % 8 (its tiles share rings there), and a workgroup takes
// blocks from its own XCD's queue (workgroups reach the XCDs round-robin by
// id, and the grid is a multiple of 8), then from the others once its own is
// empty.  Each queue lists its members' blocks member by member, the members
// with the most work per block first, so the short blocks fill the end of
// the launch.  Results are those of k_scatter: a block's waves do exactly
// the work they do there.
template <typename R, int HM, int TM, bool TRIVIAL, bool RANGE>
__global__ __launch_bounds__(kScatterWaves * kWave, 6) void k_scatter_queue(BatchQueue q)
{
    // after the static s_next: aligned as in k_scatter
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_next;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int lane = threadIdx.x % kWave;
    const int home = blockIdx.x % kXcdCount;
    ConstPtr<int32_t> q_begin = as_const(q.q_begin);
    for (int v = 0; v < kXcdCount; ++v) {
        const int x = (home + v) % kXcdCount;
        const int begin = q_begin[x];
        const int n = q_begin[x + 1] - begin;
        for (;;) {
            if (threadIdx.x == 0) s_next = atomicAdd(q.counters + x, 1);
            __syncthreads();
            const int k = __builtin_amdgcn_readfirstlane(s_next);
            __syncthreads();   // every wave has read s_next before it is written again
            if (k >= n) break;
            const int2 bl = load_const(q.blocks, begin + k);
            const ScatterArgs a = load_const(q.args, bl.x);
            if constexpr (TRIVIAL && sizeof(R) == 4) {
                float2* et = reinterpret_cast<float2*>(smem_raw);
                for (int i = threadIdx.x; i < kEnergyRows; i += blockDim.x) et[i] = a.etab[i];
                __syncthreads();
            }
            scatter_item<R, HM, TM, TRIVIAL, RANGE>(a, smem_raw, wave, lane,
                                                    a.items[bl.y * kScatterWaves + wave]);
            __syncthreads();   // LDS (energy table, columns) is reused by the next block
        }
    }
}

// Counts the (scatter point, LOR) pairs that pass the reference's filters and
// the voxel samples of their half paths with activity: the work the
// algorithmic FLOP count F = sum_pairs [2 B (n1 [act1 != 0] + n2 [act2 != 0])
// + 64] is made of (n: all samples of the half path, zero terms included).
__global__ __launch_bounds__(256) void k_pair_stats(ScatterArgs a, unsigned long long* out)
{
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int tile = blockIdx.x * (blockDim.x / kWave) + wave;
    if (tile >= a.n_tiles) return;
    const LaneCell lc = lane_cell(a, a.tiles[tile], lane);
    if (!lc.valid) return;
    const int d1 = lc.d1, d2 = lc.d2;
    double lor2;
    {
        const double dx = a.detpos[3 * d1 + 0] - a.detpos[3 * d2 + 0];
        const double dy = a.detpos[3 * d1 + 1] - a.detpos[3 * d2 + 1];
        const double dz = a.detpos[3 * d1 + 2] - a.detpos[3 * d2 + 2];
        lor2 = ((dx * dx) + (dy * dy)) + (dz * dz);
    }
    const double eff_ref = a.eff[a.eff_ref_row];
    unsigned long long pairs = 0, samples = 0;
    const int64_t per_sp = (int64_t)a.nru * a.ndu;
    const int p_end = a.chunk_sp[a.n_chunks];
    for (int sp = a.chunk_sp[0]; sp < p_end; ++sp) {
        Pair<float> p;
        if (!pair_eval_rec(a, sp, d1, d2, lor2, eff_ref, p)) continue;
        ++pairs;
        // every voxel sample of a path with activity, as the reference folds
        // them (the fold here skips the zero terms)
        const int64_t h1 = sp * per_sp + d1, h2 = sp * per_sp + d2;
        if (a.hp_rec[h1].n_flags & kHalfActive) samples += (unsigned long long)a.hp_n_all[h1];
        if (a.hp_rec[h2].n_flags & kHalfActive) samples += (unsigned long long)a.hp_n_all[h2];
    }
    atomicAdd(out + 0, pairs);
    atomicAdd(out + 1, samples);
}

// One thread per lane slot (all bins, so the count is read once per chunk):
// sum the chunks of each slice, divide by the slice's pair count (0/0 -> 0,
// SingleScatterSimulationTOF.py:201,555-557), add the slices in order
// (compensated in FP32) and write the slot's cell of the compact output.
// Chunks of one slice are consecutive.
template <typename R>
__device__ __forceinline__ void kahan_add(R& s, R& c, R v)
{
    if constexpr (sizeof(R) == 4) {
        const R y = v - c;
        const R t = s + y;
        c = (t - s) - y;
        s = t;
    } else {
        s += v;
    }
}

// Bins per thread of the chunk reduce; blockIdx.y picks the group.  Several
// groups per slot keep enough waves in flight for the chunk loop's loads.
constexpr int kReduceBins = 4;

// One lane slot's cell, bins [bg kReduceBins, bg kReduceBins + kReduceBins).
template <typename R>
__device__ __forceinline__ void reduce_slot(const ScatterArgs& a, int64_t slot, int bg)
{
    constexpr int BMAX = kReduceBins;
    const int64_t n_slots = (int64_t)a.n_tiles * kWave;
    if (slot >= n_slots) return;
    const int b0 = bg * kReduceBins;
    const int B = a.n_bins;
    if (b0 >= B) return;
    const LaneCell lc = lane_cell(a, a.tiles[slot / kWave], (int)(slot % kWave));
    if (!lc.valid) return;
    const int nb = min(kReduceBins, B - b0);
    const R* __restrict__ part_sum = reinterpret_cast<const R*>(a.part_sum) + (int64_t)b0 * n_slots;
    R total[BMAX], comp[BMAX], num[BMAX];
#pragma unroll
    for (int b = 0; b < BMAX; ++b) total[b] = comp[b] = num[b] = (R)0;
    float cnt = 0.0f;
    int slice = a.n_chunks > 0 ? a.chunk_slice[0] : 0;
    // kReduceAhead chunks' partials are loaded before any is added (indices
    // clamped into range; the extra values are not used), so a slot with
    // many chunks and few bins (small subjects) keeps loads in flight
    constexpr int kReduceAhead = 4;
    for (int c0 = 0; c0 < a.n_chunks; c0 += kReduceAhead) {
        R v[kReduceAhead][BMAX];
        float vc[kReduceAhead];
#pragma unroll
        for (int u = 0; u < kReduceAhead; ++u) {
            const int c = min(c0 + u, a.n_chunks - 1);
            const R* src = part_sum + (int64_t)c * B * n_slots + slot;
#pragma unroll
            for (int b = 0; b < BMAX; ++b) v[u][b] = src[min(b, nb - 1) * n_slots];
            vc[u] = a.part_cnt[(int64_t)c * n_slots + slot];
        }
#pragma unroll
        for (int u = 0; u < kReduceAhead; ++u) {
            const int c = c0 + u;
            if (c >= a.n_chunks) break;
            if (a.chunk_slice[c] != slice) {
                if (cnt > 0.0f) {
#pragma unroll
                    for (int b = 0; b < BMAX; ++b) kahan_add(total[b], comp[b], num[b] / (R)cnt);
                }
#pragma unroll
                for (int b = 0; b < BMAX; ++b) num[b] = (R)0;
                cnt = 0.0f;
                slice = a.chunk_slice[c];
            }
#pragma unroll
            for (int b = 0; b < BMAX; ++b)
                if (b < nb) num[b] += v[u][b];
            cnt += vc[u];
        }
    }
    const int64_t n_cells = (int64_t)a.nru * a.nru * a.n_cells_plane;
    const int64_t cell = (int64_t)lc.i1 * a.ndu - (int64_t)lc.i1 * (lc.i1 - 1) / 2 + (lc.i2 - lc.i1);
    const int64_t pc = (int64_t)(lc.r2 + lc.r1 * a.nru) * a.n_cells_plane + cell;
    R* out = reinterpret_cast<R*>(a.out) + (int64_t)b0 * n_cells;
#pragma unroll
    for (int b = 0; b < BMAX; ++b) {
        if (cnt > 0.0f) kahan_add(total[b], comp[b], num[b] / (R)cnt);
        if (b < nb) out[b * n_cells + pc] = total[b];
    }
}

template <typename R>
__global__ void k_chunk_reduce(ScatterArgs a)
{
    reduce_slot<R>(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y);
}

// Several subjects' chunk reduces in one launch: block x reduces slots
// [256 m.y, 256 m.y + 256) of subject m.x, m = blocks[x]; blockIdx.y is the
// bin group (groups past a subject's bins return).
template <typename R>
__global__ void k_chunk_reduce_queue(const ScatterArgs* args, const int2* blocks)
{
    const int2 m = load_const(blocks, blockIdx.x);
    const ScatterArgs a = load_const(args, m.x);
    reduce_slot<R>(a, (int64_t)m.y * blockDim.x + threadIdx.x, blockIdx.y);
}

template <typename R>
hipError_t launch_reduce_typed(const ScatterArgs& a, hipStream_t st)
{
    const int64_t n_slots = (int64_t)a.n_tiles * kWave;
    const dim3 grid((unsigned)((n_slots + 255) / 256), (unsigned)((a.n_bins + kReduceBins - 1) / kReduceBins));
    hipLaunchKernelGGL(k_chunk_reduce<R>, grid, dim3(256), 0, st, a);
    return hipGetLastError();
}

template <typename R>
__global__ void k_dense_scatter(const R* __restrict__ compact, int n_bins, int nru, int ndu,
                                int n_det, const int32_t* __restrict__ dets,
                                const int16_t* __restrict__ lor, float* dense)
{
    const int64_t ncell = (int64_t)ndu * (ndu + 1) / 2;
    const int64_t planes = (int64_t)nru * nru;
    const int64_t total = (int64_t)n_bins * planes * ncell;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int64_t cell = i % ncell;
    const int64_t plane = (i / ncell) % planes;
    const int64_t b = i / (ncell * planes);
    // invert the triangular cell index: i1 rows of lengths ndu, ndu-1, ...
    int i1 = 0;
    int64_t rem = cell;
    while (rem >= ndu - i1) { rem -= ndu - i1; ++i1; }
    const int i2 = i1 + (int)rem;
    const int D1 = dets[i1], D2 = dets[i2];
    const int ang = lor[((int64_t)D1 * n_det + D2) * 2 + 0];
    const int rad = lor[((int64_t)D1 * n_det + D2) * 2 + 1];
    if (ang < 0 || rad < 0) return;
    const int64_t na = n_det / 2;
    dense[((b * (n_det + 1) + rad) * na + ang) * planes + plane] = (float)compact[i];
}

// The kernel variant of a subject's scatter sum: calls
// f.template run<R, HM, TM, TRIVIAL, RANGE>() for the template that serves it.
template <typename F>
hipError_t dispatch_variant(const ScatterArgs& a, bool fp64, bool trivial, F& f);

template <typename R, bool RANGE, typename F>
hipError_t dispatch_hm(const ScatterArgs& a, F& f)
{
    const int h = a.half_window;
    if (h == 0) return f.template run<R, 0, kTofGeneral, false, RANGE>();
    if constexpr (sizeof(R) == 4 && !RANGE) {
        if (a.tof_mode == kTofSlideRel) {
            if (h == 1) return f.template run<R, 1, kTofSlideRel, false, false>();
            if (h == 2) return f.template run<R, 2, kTofSlideRel, false, false>();
        }
        if (a.tof_mode == kTofSlideRelWide && h == 1) return f.template run<R, 1, kTofSlideRelWide, false, false>();
    }
    if (h == 1) {
        if constexpr (!RANGE) {
            if (a.tof_mode == kTofOneExpInner) {
                if constexpr (sizeof(R) == 4)
                    return f.template run<R, 1, kTofSlideQuad, false, false>();
                else
                    return f.template run<R, 1, kTofOneExpInner, false, false>();
            }
        }
        if (a.tof_mode == kTofOneExp) return f.template run<R, 1, kTofOneExp, false, RANGE>();
        return f.template run<R, 1, kTofGeneral, false, RANGE>();
    }
    if (h <= 2) return f.template run<R, 2, kTofGeneral, false, RANGE>();
    if (h <= 4) return f.template run<R, 4, kTofGeneral, false, RANGE>();
    return f.template run<R, 8, kTofGeneral, false, RANGE>();
}

template <typename R, typename F>
hipError_t dispatch_typed(const ScatterArgs& a, bool trivial, F& f)
{
    if (trivial) return f.template run<R, 1, kTofGeneral, true, false>();
    return a.range_check ? dispatch_hm<R, true>(a, f) : dispatch_hm<R, false>(a, f);
}

template <typename F>
hipError_t dispatch_variant(const ScatterArgs& a, bool fp64, bool trivial, F& f)
{
    return fp64 ? dispatch_typed<double>(a, trivial, f) : dispatch_typed<float>(a, trivial, f);
}

// Dynamic LDS of one scatter workgroup: the energy table (one-bin FP32), or
// one column of B + 2 HM rows per lane.
template <typename R, int HM, bool TRIVIAL>
size_t scatter_lds(int n_bins)
{
    return TRIVIAL ? (sizeof(R) == 4 ? kEnergyRows * sizeof(float2) : 0)
                   : (size_t)kScatterWaves * (n_bins + 2 * HM) * kWave * sizeof(R);
}

struct LaunchOne {
    const ScatterArgs& a;
    hipStream_t st;
    template <typename R, int HM, int TM, bool TRIVIAL, bool RANGE>
    hipError_t run()
    {
        // one workgroup per kScatterWaves work items (the item list is padded to that)
        const dim3 grid((unsigned)(a.n_items / kScatterWaves));
        const size_t lds = scatter_lds<R, HM, TRIVIAL>(a.n_bins);
        hipLaunchKernelGGL((k_scatter<R, HM, TM, TRIVIAL, RANGE>), grid, dim3(kScatterWaves * kWave), lds, st, a);
        return hipGetLastError();
    }
};

struct VariantId {
    int id = -1;
    template <typename R, int HM, int TM, bool TRIVIAL, bool RANGE>
    hipError_t run()
    {
        id = (sizeof(R) == 8 ? 1 : 0) + (TRIVIAL ? 2 : 0) + (RANGE ? 4 : 0) + 8 * HM + 128 * TM;
        return hipSuccess;
    }
};

struct QueueResident {
    int max_bins;
    int per_cu = 0;
    template <typename R, int HM, int TM, bool TRIVIAL, bool RANGE>
    hipError_t run()
    {
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_scatter_queue<R, HM, TM, TRIVIAL, RANGE>,
                                                            kScatterWaves * kWave,
                                                            scatter_lds<R, HM, TRIVIAL>(max_bins));
    }
};

struct LaunchQueue {
    const BatchQueue& q;
    int max_bins;
    int grid;
    hipStream_t st;
    template <typename R, int HM, int TM, bool TRIVIAL, bool RANGE>
    hipError_t run()
    {
        const size_t lds = scatter_lds<R, HM, TRIVIAL>(max_bins);
        hipLaunchKernelGGL((k_scatter_queue<R, HM, TM, TRIVIAL, RANGE>), dim3((unsigned)grid),
                           dim3(kScatterWaves * kWave), lds, st, q);
        return hipGetLastError();
    }
};

}  // namespace

hipError_t launch_scatter(const ScatterArgs& a, bool fp64, bool trivial, hipStream_t st)
{
    if (a.n_tiles == 0 || a.n_items < kScatterWaves) return hipSuccess;
    LaunchOne f{a, st};
    return dispatch_variant(a, fp64, trivial, f);
}

int scatter_variant(const ScatterArgs& a, bool fp64, bool trivial)
{
    VariantId f;
    (void)dispatch_variant(a, fp64, trivial, f);
    return f.id;
}

hipError_t scatter_queue_grid(const ScatterArgs& a, bool fp64, bool trivial, int max_bins, int n_cu, int n_blocks,
                              int* grid)
{
    QueueResident f{max_bins};
    hipError_t e = dispatch_variant(a, fp64, trivial, f);
    if (e != hipSuccess) return e;
    // every resident workgroup, a multiple of 8 so that workgroup id % 8 is
    // the XCD; no more than the blocks to run (rounded up to 8)
    int g = std::max(1, f.per_cu) * std::max(1, n_cu);
    g = std::min(g, (n_blocks + kXcdCount - 1) / kXcdCount * kXcdCount);
    *grid = std::max(kXcdCount, g / kXcdCount * kXcdCount);
    return hipSuccess;
}

hipError_t launch_scatter_queue(const ScatterArgs& a, bool fp64, bool trivial, const BatchQueue& q, int max_bins,
                                int grid, hipStream_t st)
{
    LaunchQueue f{q, max_bins, grid, st};
    return dispatch_variant(a, fp64, trivial, f);
}

hipError_t launch_pair_stats(const ScatterArgs& a, unsigned long long* out, hipStream_t st)
{
    if (a.n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_stats, dim3((a.n_tiles + 3) / 4), dim3(256), 0, st, a, out);
    return hipGetLastError();
}

hipError_t launch_chunk_reduce(const ScatterArgs& a, bool fp64, hipStream_t st)
{
    if (a.n_tiles == 0) return hipSuccess;
    if (a.n_bins > kMaxBins) return hipErrorInvalidValue;
    return fp64 ? launch_reduce_typed<double>(a, st) : launch_reduce_typed<float>(a, st);
}

int reduce_blocks(const ScatterArgs& a)
{
    return (int)(((int64_t)a.n_tiles * kWave + 255) / 256);
}

hipError_t launch_chunk_reduce_queue(const ScatterArgs* args, const int2* blocks, int n_blocks, int max_bins,
                                     bool fp64, hipStream_t st)
{
    if (n_blocks == 0) return hipSuccess;
    if (max_bins > kMaxBins) return hipErrorInvalidValue;
    const dim3 grid((unsigned)n_blocks, (unsigned)((max_bins + kReduceBins - 1) / kReduceBins));
    if (fp64)
        hipLaunchKernelGGL(k_chunk_reduce_queue<double>, grid, dim3(256), 0, st, args, blocks);
    else
        hipLaunchKernelGGL(k_chunk_reduce_queue<float>, grid, dim3(256), 0, st, args, blocks);
    return hipGetLastError();
}

hipError_t launch_dense_scatter_typed(const void* compact, bool fp64, int n_bins, int nru, int ndu,
                                      int n_det, const int32_t* dets, const int16_t* lor,
                                      float* dense, hipStream_t st)
{
    const int64_t total = (int64_t)n_bins * nru * nru * ((int64_t)ndu * (ndu + 1) / 2);
    if (total == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((total + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(k_dense_scatter<double>, dim3(blocks), dim3(256), 0, st,
                           (const double*)compact, n_bins, nru, ndu, n_det, dets, lor, dense);
    else
        hipLaunchKernelGGL(k_dense_scatter<float>, dim3(blocks), dim3(256), 0, st,
                           (const float*)compact, n_bins, nru, ndu, n_det, dets, lor, dense);
    return hipGetLastError();
}

}  // namespace sss
