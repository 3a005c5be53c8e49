// Fused single-scatter sum (the bodySSS LOR loop, SingleScatterSimulationTOF.py:433-557).
//
// Work decomposition: one wavefront owns an 8 x 8 tile of cells (i1, i2) of
// one ring-pair plane (r1, r2) and keeps that tile for a contiguous range of
// normalisation slices (a chunk).  Each lane accumulates its cell over every
// scatter point of a slice, then adds sum/count of the slice (the per-slice
// normalisation of :555-557) into its running output.  No atomics on global
// memory and a fixed reduction order: results do not depend on what else
// runs on the device.
//
// Per (scatter point, cell) the lane evaluates, in float64: the distances,
// the Compton energy index floor(511/(2 + cos_arg) + 0.5) (the same value as
// the reference's acos/degrees/cos round trip at :495-501) and the pair's
// TOF time shift; in R (float or double): Klein-Nishina, efficiency, solid
// angle, attenuation scaling and the TOF fold.  The fold visits every voxel
// sample of both half paths, rounds the time offset to the reference's
// integer picosecond (:520,526), and adds the sample into the nearest TOF
// bin and its +-h neighbours with the normalised Gaussian weights (the rows
// of the reference's TOF table, evaluated analytically).  Fold accumulators
// live in LDS, one column per lane, so per-lane bin indices need no
// register indexing.  Sums over scatter points and slices are compensated
// (Kahan) in the float32 path.
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

// Adds coef * a_k * T[b, t_k] for every sample of one half path into the
// lane's LDS column (rows b + h).  sign = -1 for the D1 path, +1 for D2.
template <typename R, int HM>
__device__ __forceinline__ void tof_fold(const ScatterArgs& a, R* __restrict__ col, int64_t base,
                                         int n, double shift, double sign, R coef, int n_bins)
{
    const int h = a.half_window;
    const R inv_w = (R)a.inv_w_bins, mid = (R)a.mid, bw = (R)a.bin_w;
    const R kappa2 = (R)a.kappa2, beta2 = (R)a.beta2;
    const R bmax = (R)(n_bins - 1);
    for (int k = 0; k < n; ++k) {
        const double q = a.s_q[base + k];
        const R av = (R)a.s_a[base + k];
        const double fl = floor(shift + sign * q);
        bool ok = (fl >= 0.0) && (fl <= 2.0 * kTofHalf);
        const int t = (ok ? (int)fl : kTofHalf) - kTofHalf;
        ok = ok && (t >= a.t_lo) && (t <= a.t_hi);
        const R tf = (R)t;
        R jf = rint(tf * inv_w + mid);
        jf = jf < (R)0 ? (R)0 : (jf > bmax ? bmax : jf);
        const int j0 = (int)jf;
        const R ku = kappa2 * (tf - (jf - mid) * bw);
        R e[2 * HM + 1];
        R z = (R)1;
        e[HM] = (R)1;
#pragma unroll
        for (int d = 1; d <= HM; ++d) {
            R ep = (R)0, em = (R)0;
            if (d <= h) {
                const R dd = (R)(d * d) * beta2;
                ep = (j0 + d < n_bins) ? fast_exp2((R)d * ku - dd) : (R)0;
                em = (j0 - d >= 0) ? fast_exp2(-(R)d * ku - dd) : (R)0;
            }
            e[HM + d] = ep;
            e[HM - d] = em;
            z += ep + em;
        }
        const R wgt = ok ? coef * av * fast_rcp(z) : (R)0;
        R* row = col + (j0 + h) * kWave;
        atomicAdd(row, wgt);
#pragma unroll
        for (int d = 1; d <= HM; ++d) {
            if (d <= h) {
                atomicAdd(row + d * kWave, wgt * e[HM + d]);
                atomicAdd(row - d * kWave, wgt * e[HM - d]);
            }
        }
    }
}

template <typename R, int BMAX, int HM, bool TRIVIAL>
__global__ __launch_bounds__(256) void k_scatter(ScatterArgs a)
{
    extern __shared__ unsigned char smem_raw[];
    const int wave = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    const int tile = blockIdx.x * (blockDim.x / kWave) + wave;
    if (tile >= a.n_tiles) return;
    const int chunk = blockIdx.y;
    const int B = a.n_bins;
    const int rows = B + 2 * a.half_window;
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
    const int i1c = i1 < ndu ? i1 : ndu - 1;
    const int i2c = i2 < ndu ? i2 : ndu - 1;
    const int d1 = r1 * ndu + i1c;
    const int d2 = r2 * ndu + i2c;
    double lor2;
    {
        const double dx = a.detpos[3 * d1 + 0] - a.detpos[3 * d2 + 0];
        const double dy = a.detpos[3 * d1 + 1] - a.detpos[3 * d2 + 1];
        const double dz = a.detpos[3 * d1 + 2] - a.detpos[3 * d2 + 2];
        lor2 = ((dx * dx) + (dy * dy)) + (dz * dz);
    }
    const double eff_ref = a.eff[a.eff_ref_row];
    const int64_t per_sp = (int64_t)nru * ndu;
    const double inv_c = 1.0 / kLightCmPerPs;

    R out[BMAX], outc[BMAX], num[BMAX], numc[BMAX];
#pragma unroll
    for (int b = 0; b < BMAX; ++b) { out[b] = outc[b] = (R)0; }

    const int s_begin = a.chunk_slices[chunk];
    const int s_end = a.chunk_slices[chunk + 1];
    for (int sl = s_begin; sl < s_end; ++sl) {
#pragma unroll
        for (int b = 0; b < BMAX; ++b) { num[b] = numc[b] = (R)0; }
        float cnt = 0.0f;
        const int p_begin = a.slice_ptr[sl];
        const int p_end = a.slice_ptr[sl + 1];
        for (int sp = p_begin; sp < p_end; ++sp) {
            const int64_t h1 = sp * per_sp + d1;
            const int64_t h2 = sp * per_sp + d2;
            const double dist1 = a.hp_dist[h1], dist2 = a.hp_dist[h2];
            const double att1 = a.hp_att[h1], att2 = a.hp_att[h2];
            const double act1 = a.hp_act[h1], act2 = a.hp_act[h2];
            bool ok = valid && att1 != 0.0 && att2 != 0.0 && !(act1 == 0.0 && act2 == 0.0);
            // law of cosines at the scatter point, energy of the scattered photon
            const double s1q = dist1 * dist1, s2q = dist2 * dist2;
            double carg = (s1q + s2q - lor2) / (2.0 * sqrt(s1q * s2q));
            carg = carg < -1.0 ? -1.0 : (carg > 1.0 ? 1.0 : carg);
            const double e_kev = floor(511.0 / (2.0 + carg) + 0.5);
            int erow = (int)(e_kev * 2.0) - 1;
            erow = erow < 0 ? 0 : (erow >= a.n_energy ? a.n_energy - 1 : erow);
            const double eff = a.eff[erow] * eff_ref;
            ok = ok && (eff != 0.0);
            if (!ok) continue;
            const R ratio = (R)a.ratio[erow];
            const R ln1 = (R)a.hp_lnatt[h1], ln2 = (R)a.hp_lnatt[h2];
            const R log2e = (R)1.4426950408889634;
            const R a1s = fast_exp2(-ratio * ln1 * log2e);
            const R a2s = fast_exp2(-ratio * ln2 * log2e);
            const R c = (R)carg;
            const R rho = (R)1 / ((R)1 + (R)a.kn_gamma * ((R)1 + c));
            const R kn = rho * rho * (rho + (R)1 / rho - ((R)1 - c * c));
            const R geo = (R)a.g_const * (R)a.hp_cos[h1] * (R)a.hp_cos[h2] / ((R)s1q * (R)s2q);
            const R pre = geo * (R)eff * kn;
            const R coef1 = pre * (R)att1 * a2s;
            const R coef2 = pre * (R)att2 * a1s;
            if constexpr (TRIVIAL) {
                const R v = coef1 * (R)act1 + coef2 * (R)act2;
                kahan_add(num[0], numc[0], v);
            } else {
                const double shift = (dist1 - dist2) * inv_c + (kTofHalf + 0.5);
                if (act1 != 0.0) tof_fold<R, HM>(a, col, a.hp_off[h1], a.hp_n[h1], shift, -1.0, coef1, B);
                if (act2 != 0.0) tof_fold<R, HM>(a, col, a.hp_off[h2], a.hp_n[h2], shift, 1.0, coef2, B);
                const int h = a.half_window;
#pragma unroll
                for (int b = 0; b < BMAX; ++b) {
                    if (b < B) {
                        R* p = col + (b + h) * kWave;
                        const R v = *p;
                        *p = (R)0;
                        kahan_add(num[b], numc[b], v);
                    }
                }
            }
            cnt += (act1 != 0.0 && act2 != 0.0) ? 2.0f : 1.0f;
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
    const int ndu = a.ndu, nru = a.nru;
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
    const int64_t per_sp = (int64_t)nru * ndu;
    unsigned long long pairs = 0, samples = 0;
    const int p_end = a.slice_ptr[a.chunk_slices[a.n_chunks]];
    for (int sp = a.slice_ptr[a.chunk_slices[0]]; sp < p_end; ++sp) {
        const int64_t h1 = sp * per_sp + d1;
        const int64_t h2 = sp * per_sp + d2;
        const double dist1 = a.hp_dist[h1], dist2 = a.hp_dist[h2];
        const double act1 = a.hp_act[h1], act2 = a.hp_act[h2];
        if (a.hp_att[h1] == 0.0 || a.hp_att[h2] == 0.0 || (act1 == 0.0 && act2 == 0.0)) continue;
        const double s1q = dist1 * dist1, s2q = dist2 * dist2;
        double carg = (s1q + s2q - lor2) / (2.0 * sqrt(s1q * s2q));
        carg = carg < -1.0 ? -1.0 : (carg > 1.0 ? 1.0 : carg);
        int erow = (int)(floor(511.0 / (2.0 + carg) + 0.5) * 2.0) - 1;
        erow = erow < 0 ? 0 : (erow >= a.n_energy ? a.n_energy - 1 : erow);
        if (a.eff[erow] * eff_ref == 0.0) continue;
        ++pairs;
        samples += (act1 != 0.0 ? (unsigned long long)a.hp_n[h1] : 0ull) +
                   (act2 != 0.0 ? (unsigned long long)a.hp_n[h2] : 0ull);
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

template <typename R, int BMAX, int HM, bool TRIVIAL>
hipError_t launch_variant(const ScatterArgs& a, hipStream_t st)
{
    const int waves_per_block = 4;
    dim3 grid((a.n_tiles + waves_per_block - 1) / waves_per_block, a.n_chunks);
    size_t lds = TRIVIAL ? 0 : (size_t)waves_per_block * (a.n_bins + 2 * a.half_window) * kWave * sizeof(R);
    hipLaunchKernelGGL((k_scatter<R, BMAX, HM, TRIVIAL>), grid, dim3(waves_per_block * kWave), lds, st, a);
    return hipGetLastError();
}

template <typename R, int BMAX>
hipError_t launch_hm(const ScatterArgs& a, hipStream_t st)
{
    const int h = a.half_window;
    if (h <= 1) return launch_variant<R, BMAX, 1, false>(a, st);
    if (h <= 2) return launch_variant<R, BMAX, 2, false>(a, st);
    if (h <= 4) return launch_variant<R, BMAX, 4, false>(a, st);
    return launch_variant<R, BMAX, 8, false>(a, st);
}

template <typename R>
hipError_t launch_typed(const ScatterArgs& a, bool trivial, hipStream_t st)
{
    if (trivial) return launch_variant<R, 1, 1, true>(a, st);
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
