// Fused single-scatter sum (the bodySSS LOR loop, SingleScatterSimulationTOF.py:433-557).
//
// Work decomposition: one wavefront owns an 8 x 8 tile of cells (i1, i2) of
// one ring-pair plane (r1, r2) for a contiguous range of normalisation
// slices (a chunk).  Each lane accumulates its own cell over the scatter
// points of a slice, then adds sum/count of the slice (the per-slice
// normalisation of :555-557) into its running output.  No global atomics and
// a fixed reduction order: results do not depend on what else runs.
//
// Every lane walks the slice's scatter points with its own cursor and skips
// the pairs the reference filters out (attenuation 0, no activity on either
// path, detection efficiency 0 -- about half of them are backscatter beyond
// the energy threshold), so the TOF fold only ever runs on lanes with work.
//
// Per (scatter point, cell) the lane evaluates, in float64: the distances,
// the Compton energy index floor(511/(2 + cos_arg) + 0.5) (the value of the
// reference's acos/degrees/cos round trip at :495-501) and the pair's TOF
// time shift; in R (float or double): Klein-Nishina, efficiency, solid angle,
// attenuation scaling and the TOF fold.  The fold visits every voxel sample
// of both half paths in one loop, rounds the time offset to the reference's
// integer picosecond in float64 (:520,526) and adds the sample into the nearest TOF bin and
// its +-h neighbours with the normalised Gaussian weights (the columns of the
// reference's TOF table, evaluated analytically).  Fold accumulators live in
// LDS, one column per lane, so per-lane bin indices need no register
// indexing.  Sums over scatter points and slices are compensated (Kahan) in
// the FP32 path.
#include "sss_internal.h"

namespace sss {
namespace {

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

// Geometry and coefficients of one (scatter point, cell) pair.
template <typename R>
struct Pair {
    R coef1, coef2;          // G eff KN A1 A2^r, G eff KN A2 A1^r
    double shift;            // (d1 - d2)/c + 4000.5 ps
    int n1, n2;             // samples folded on each path (0 when its activity is 0)
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
    const double s1q = dist1 * dist1, s2q = dist2 * dist2;
    double carg = (s1q + s2q - lor2) / (2.0 * sqrt(s1q * s2q));
    carg = carg < -1.0 ? -1.0 : (carg > 1.0 ? 1.0 : carg);
    const double e_kev = floor(511.0 / (2.0 + carg) + 0.5);
    int erow = (int)(e_kev * 2.0) - 1;
    erow = erow < 0 ? 0 : (erow >= a.n_energy ? a.n_energy - 1 : erow);
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
    p.shift = (dist1 - dist2) * (1.0 / kLightCmPerPs) + (kTofHalf + 0.5);
    p.n1 = act1 != 0.0 ? a.hp_n[h1] : 0;
    p.n2 = act2 != 0.0 ? a.hp_n[h2] : 0;
    p.b1 = a.hp_off[h1];
    p.b2 = a.hp_off[h2];
    p.act1 = (R)act1;
    p.act2 = (R)act2;
    p.counts = (act1 != 0.0 && act2 != 0.0) ? 2.0f : 1.0f;
    return true;
}

// TOF weights of one voxel sample at integer offset off (t = off - 4000 ps):
// w[H + d] = wsample * T[j0 + d, t] for |d| <= h, zero outside the bins or
// the window; returns the nearest bin j0.
template <typename R, int H, bool SINGLE_EXP>
__device__ __forceinline__ int tof_weights(const ScatterArgs& a, uint32_t off, R wsample, int n_bins,
                                           R (&w)[2 * H + 1])
{
    const R tf = (R)((int)off - kTofHalf);
    const R jf = clampr(rint(tf * (R)a.inv_w_bins + (R)a.mid), (R)0, (R)(n_bins - 1));
    const int j0 = (int)jf;
    const R u = tf - (jf - (R)a.mid) * (R)a.bin_w;
    R ku = u * (R)a.kappa2;
    if constexpr (SINGLE_EXP) {
        // h = 1: e+ = c E, e- = c / E with E = 2^ku, c = 2^-beta2; with
        // D = E (1 + e+ + e-) the weights are E/D, c E^2/D and c/D.
        ku = clampr(ku, (R)-100, (R)100);
        const R E = fast_exp2(ku);
        const R c = (R)a.c_beta;
        const R cp = (j0 < n_bins - 1) ? c * E : (R)0;
        const R cm = (j0 > 0) ? c : (R)0;
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
        for (int d = 1; d <= H; ++d) {
            R ep = (R)0, em = (R)0;
            if (d <= h) {
                const R dd = (R)(d * d) * (R)a.beta2;
                ep = (j0 + d < n_bins) ? fast_exp2((R)d * ku - dd) : (R)0;
                em = (j0 - d >= 0) ? fast_exp2(-(R)d * ku - dd) : (R)0;
            }
            w[H + d] = ep;
            w[H - d] = em;
            z += ep + em;
        }
        const R g = wsample * fast_rcp(z);
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) w[i] *= g;
    }
    return j0;
}

// Register window over the 2H+1 TOF bins around the current nearest bin.
// Along one half path the nearest bin changes only every few tens of
// samples (a bin is hundreds of ps, a voxel sample tens), so consecutive
// samples add into the same registers; when the nearest bin changes the
// window is added once into the lane's LDS column (rows b + H, padded so no
// bounds test is needed; a plain read-add-write, the lane owns its column).
template <typename R, int H>
struct BinWindow {
    R v[2 * H + 1];
    int cur;        // nearest bin of the samples in the window

    __device__ __forceinline__ void reset(int j)
    {
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) v[i] = (R)0;
        cur = j;
    }
    __device__ __forceinline__ void spill(R* col) const
    {
        // window bins cur-H .. cur+H live in rows cur .. cur+2H
        R* base = col + cur * kWave;
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) base[i * kWave] += v[i];
    }
    __device__ __forceinline__ void add(R* col, int j0, const R (&w)[2 * H + 1])
    {
        if (j0 != cur) {
            spill(col);
            reset(j0);
        }
#pragma unroll
        for (int i = 0; i < 2 * H + 1; ++i) v[i] += w[i];
    }
};

// Fold the voxel samples of one half path into the lane's column:
// offset floor(shift + sign * q) (SIGN = -1: path to D1, P - q; +1: path to
// D2, P + q), weight a_k * coef * T[b, t].  Offsets outside the non-zero TOF
// columns weigh 0.  The samples two ahead are in flight.
template <typename R, int H, bool SINGLE_EXP, int SIGN>
__device__ __forceinline__ void fold_path(const ScatterArgs& a, R* __restrict__ col, int64_t base, int n,
                                          double shift, R coef, int n_bins)
{
    const uint32_t lo = (uint32_t)(a.t_lo + kTofHalf);
    const uint32_t span = (uint32_t)(a.t_hi - a.t_lo);
    const double* __restrict__ qp = a.s_q + base;
    BinWindow<R, H> win;
    win.reset(0);
    double q0 = 0.0, q1 = 0.0;
    R a0 = (R)0, a1 = (R)0;
    auto load = [&](int k, double& q, R& av) {
        q = qp[k];
        if constexpr (sizeof(R) == 4) av = a.s_a[base + k];
        else av = a.s_a64[base + k];
    };
    if (n > 0) load(0, q0, a0);
    if (n > 1) load(1, q1, a1);
    for (int k = 0; k < n; ++k) {
        const double q = q0;
        const R av = a0;
        q0 = q1;
        a0 = a1;
        if (k + 2 < n) load(k + 2, q1, a1);
        // integer picosecond offset, float64 as the reference (:520,526)
        const double fl = floor(SIGN < 0 ? shift - q : shift + q);
        const uint32_t off = (uint32_t)fl;   // saturates below 0
        const bool ok = (fl >= 0.0) && (off - lo) <= span;
        R w[2 * H + 1];
        const int j0 = tof_weights<R, H, SINGLE_EXP>(a, ok ? off : lo, ok ? av * coef : (R)0, n_bins, w);
        win.add(col, j0, w);
    }
    win.spill(col);
}

// Fold both half paths of a pair (path 1: offsets P - q; path 2: P + q).
template <typename R, int H, bool SINGLE_EXP>
__device__ __forceinline__ void fold_pair(const ScatterArgs& a, R* __restrict__ col, const Pair<R>& p,
                                          int n_bins)
{
    if (a.t_hi < a.t_lo) return;   // no non-zero TOF column at all: every weight is zero
    fold_path<R, H, SINGLE_EXP, -1>(a, col, p.b1, p.n1, p.shift, p.coef1, n_bins);
    fold_path<R, H, SINGLE_EXP, +1>(a, col, p.b2, p.n2, p.shift, p.coef2, n_bins);
}

template <typename R, int BMAX, int HM, bool SINGLE_EXP, bool TRIVIAL>
__global__ __launch_bounds__(256, 3) void k_scatter(ScatterArgs a)
{
    extern __shared__ unsigned char smem_raw[];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int tile = blockIdx.x * (blockDim.x / kWave) + wave;
    if (tile >= a.n_tiles) return;
    const int chunk = blockIdx.y;
    const int B = a.n_bins;
    const int rows = B + 2 * HM;   // bins -HM .. B-1+HM (the window's reach)
    R* col = reinterpret_cast<R*>(smem_raw) + (size_t)wave * rows * kWave + lane;
    if constexpr (!TRIVIAL) {
        for (int r = 0; r < rows; ++r) col[r * kWave] = (R)0;
    }

    const int4 tl = a.tiles[tile];
    const int r1 = tl.x, r2 = tl.y;
    const int ndu = a.ndu, nru = a.nru;
    const int i1 = tl.z + lane / kTileCols;
    const int i2 = tl.w + lane % kTileCols;
    const bool valid = (i1 < ndu) && (i2 < ndu) && (i1 < i2 || (i1 == i2 && r1 != r2));
    const int d1 = r1 * ndu + (i1 < ndu ? i1 : ndu - 1);
    const int d2 = r2 * ndu + (i2 < ndu ? i2 : ndu - 1);
    double lor2;
    {
        const double dx = a.detpos[3 * d1 + 0] - a.detpos[3 * d2 + 0];
        const double dy = a.detpos[3 * d1 + 1] - a.detpos[3 * d2 + 1];
        const double dz = a.detpos[3 * d1 + 2] - a.detpos[3 * d2 + 2];
        lor2 = ((dx * dx) + (dy * dy)) + (dz * dz);
    }
    const double eff_ref = a.eff[a.eff_ref_row];

    R out[BMAX], outc[BMAX], num[BMAX], numc[BMAX];
#pragma unroll
    for (int b = 0; b < BMAX; ++b) { out[b] = outc[b] = (R)0; }

    const int s_begin = a.chunk_slices[chunk];
    const int s_end = a.chunk_slices[chunk + 1];
    for (int sl = s_begin; sl < s_end; ++sl) {
#pragma unroll
        for (int b = 0; b < BMAX; ++b) { num[b] = numc[b] = (R)0; }
        float cnt = 0.0f;
        const int p_end = a.slice_ptr[sl + 1];
        int cursor = valid ? a.slice_ptr[sl] : p_end;
        while (true) {
            // next scatter point of this lane's cell that the reference keeps
            Pair<R> p;
            bool found = false;
            while (cursor < p_end) {
                found = pair_eval<R>(a, cursor, d1, d2, lor2, eff_ref, p);
                ++cursor;
                if (found) break;
            }
            if (!__any(found)) break;
            if (found) {
                cnt += p.counts;
                if constexpr (TRIVIAL) {
                    kahan_add(num[0], numc[0], p.coef1 * p.act1 + p.coef2 * p.act2);
                } else {
                    fold_pair<R, HM, SINGLE_EXP>(a, col, p, B);
#pragma unroll
                    for (int b = 0; b < BMAX; ++b) {
                        if (b < B) {
                            R* q = col + (b + HM) * kWave;
                            const R v = *q;
                            *q = (R)0;
                            kahan_add(num[b], numc[b], v);
                        }
                    }
                }
            }
        }
        if (cnt > 0.0f) {
#pragma unroll
            for (int b = 0; b < BMAX; ++b)
                if (b < B) kahan_add(out[b], outc[b], num[b] / (R)cnt);
        }
    }
    if (!valid) return;
    const int plane = r2 + r1 * nru;
    const int64_t cell = (int64_t)i1 * ndu - (int64_t)i1 * (i1 - 1) / 2 + (i2 - i1);
    const int64_t planes = (int64_t)nru * nru;
    R* dst = reinterpret_cast<R*>(a.out);
#pragma unroll
    for (int b = 0; b < BMAX; ++b) {
        if (b < B)
            dst[(((int64_t)chunk * B + b) * planes + plane) * a.n_cells_plane + cell] = out[b];
    }
}

// Counts the (scatter point, LOR) pairs that pass the reference's filters and
// the voxel samples their TOF folds visit: the work the algorithmic FLOP count
// F = sum_pairs [2 B (n1 [act1 != 0] + n2 [act2 != 0]) + 64] is made of.
__global__ __launch_bounds__(256) void k_pair_stats(ScatterArgs a, unsigned long long* out)
{
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int tile = blockIdx.x * (blockDim.x / kWave) + wave;
    if (tile >= a.n_tiles) return;
    const int4 tl = a.tiles[tile];
    const int r1 = tl.x, r2 = tl.y;
    const int ndu = a.ndu;
    const int i1 = tl.z + lane / kTileCols;
    const int i2 = tl.w + lane % kTileCols;
    const bool valid = (i1 < ndu) && (i2 < ndu) && (i1 < i2 || (i1 == i2 && r1 != r2));
    if (!valid) return;
    const int d1 = r1 * ndu + i1;
    const int d2 = r2 * ndu + i2;
    double lor2;
    {
        const double dx = a.detpos[3 * d1 + 0] - a.detpos[3 * d2 + 0];
        const double dy = a.detpos[3 * d1 + 1] - a.detpos[3 * d2 + 1];
        const double dz = a.detpos[3 * d1 + 2] - a.detpos[3 * d2 + 2];
        lor2 = ((dx * dx) + (dy * dy)) + (dz * dz);
    }
    const double eff_ref = a.eff[a.eff_ref_row];
    unsigned long long pairs = 0, samples = 0;
    const int p_end = a.slice_ptr[a.chunk_slices[a.n_chunks]];
    for (int sp = a.slice_ptr[a.chunk_slices[0]]; sp < p_end; ++sp) {
        Pair<float> p;
        if (!pair_eval<float>(a, sp, d1, d2, lor2, eff_ref, p)) continue;
        ++pairs;
        samples += (unsigned long long)(p.n1 + p.n2);
    }
    atomicAdd(out + 0, pairs);
    atomicAdd(out + 1, samples);
}

template <typename R>
__global__ void k_chunk_reduce(const R* __restrict__ partial, int n_chunks, int64_t n, R* out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    R s = partial[i];
    for (int c = 1; c < n_chunks; ++c) s += partial[(int64_t)c * n + i];
    out[i] = s;
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

template <typename R, int BMAX, int HM, bool SINGLE_EXP, bool TRIVIAL>
hipError_t launch_variant(const ScatterArgs& a, hipStream_t st)
{
    const int waves_per_block = 4;
    dim3 grid((a.n_tiles + waves_per_block - 1) / waves_per_block, a.n_chunks);
    size_t lds = TRIVIAL ? 0 : (size_t)waves_per_block * (a.n_bins + 2 * HM) * kWave * sizeof(R);
    hipLaunchKernelGGL((k_scatter<R, BMAX, HM, SINGLE_EXP, TRIVIAL>), grid, dim3(waves_per_block * kWave),
                       lds, st, a);
    return hipGetLastError();
}

template <typename R, int BMAX>
hipError_t launch_hm(const ScatterArgs& a, hipStream_t st)
{
    const int h = a.half_window;
    if (h == 0) return launch_variant<R, BMAX, 0, false, false>(a, st);
    if (h == 1) {
        if (a.single_exp) return launch_variant<R, BMAX, 1, true, false>(a, st);
        return launch_variant<R, BMAX, 1, false, false>(a, st);
    }
    if (h <= 2) return launch_variant<R, BMAX, 2, false, false>(a, st);
    if (h <= 4) return launch_variant<R, BMAX, 4, false, false>(a, st);
    return launch_variant<R, BMAX, 8, false, false>(a, st);
}

template <typename R>
hipError_t launch_typed(const ScatterArgs& a, bool trivial, hipStream_t st)
{
    if (trivial) return launch_variant<R, 1, 1, false, true>(a, st);
    if (a.n_bins <= 4) return launch_hm<R, 4>(a, st);
    if (a.n_bins <= 8) return launch_hm<R, 8>(a, st);
    if (a.n_bins <= 16) return launch_hm<R, 16>(a, st);
    return launch_hm<R, 32>(a, st);
}

}  // namespace

hipError_t launch_scatter(const ScatterArgs& a, bool fp64, bool trivial, hipStream_t st)
{
    if (a.n_tiles == 0) return hipSuccess;
    return fp64 ? launch_typed<double>(a, trivial, st) : launch_typed<float>(a, trivial, st);
}

hipError_t launch_pair_stats(const ScatterArgs& a, unsigned long long* out, hipStream_t st)
{
    if (a.n_tiles == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_stats, dim3((a.n_tiles + 3) / 4), dim3(256), 0, st, a, out);
    return hipGetLastError();
}

hipError_t launch_chunk_reduce_typed(const void* partial, int n_chunks, int64_t n, void* out,
                                     bool fp64, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (fp64)
        hipLaunchKernelGGL(k_chunk_reduce<double>, dim3(blocks), dim3(256), 0, st,
                           (const double*)partial, n_chunks, n, (double*)out);
    else
        hipLaunchKernelGGL(k_chunk_reduce<float>, dim3(blocks), dim3(256), 0, st,
                           (const float*)partial, n_chunks, n, (float*)out);
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
