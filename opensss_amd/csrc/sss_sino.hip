// Full-sinogram stages after the scatter estimate:
//
//   k_interp      Interpolation.py:95-399: sampled (compact) sinogram -> every
//                 (radial, angular, plane) cell of every TOF bin through two
//                 geometry-only index maps (nearest sampled detector pair,
//                 nearest sampled ring pair), clamped at 0, plus the sum over bins.
//   k_tail_mask   TailsMask.py:84-163,230-309: detector pairs traced through
//                 mu (or lambda); a cell is a tail when the float64 attenuation
//                 factor 1/exp(-sum mu L) is exactly 1 (or the activity is 0).
//   k_fit_bp      BackProjection.py:31-113,217-300: masked scatter, prompts and
//                 randoms backprojected along every surviving LOR into 64-bit
//                 fixed-point images plus the sensitivity image (integer
//                 atomics: the same bits in every run).
//   k_fit_reduce  BackProjection.py:191-210: per-voxel ratios and the float64
//   k_fit_final   normal equations sum(x y), sum(x x) of the one-parameter
//                 fit, reduced in a fixed order.
#include "sss_walk.h"

#include <algorithm>

namespace sss {
namespace {

__global__ void k_interp(const float* __restrict__ sampled, int n_bins, int n_sampled_planes,
                         int64_t cells_per_sampled_plane, const int32_t* __restrict__ cell_map,
                         int64_t n_cells_out, const int32_t* __restrict__ plane_map, int n_out_planes,
                         float* __restrict__ out, float* __restrict__ sum)
{
    const int64_t n = n_cells_out * n_out_planes;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t c = i / n_out_planes;
    const int s = (int)(i - c * n_out_planes);
    const int32_t src_cell = cell_map[c];
    const int32_t src_plane = plane_map[s];
    const bool ok = src_cell >= 0 && src_plane >= 0;
    float acc = 0.0f;
    for (int b = 0; b < n_bins; ++b) {
        float v = 0.0f;
        if (ok) {
            v = sampled[((int64_t)b * n_sampled_planes + src_plane) * cells_per_sampled_plane + src_cell];
            v = v < 0.0f ? 0.0f : v;
        }
        if (out) out[(int64_t)b * n + i] = v;
        acc += v;
    }
    if (sum) sum[i] = acc;
}

// Segment between detector d1 of geometry ring ra and d2 of ring rb; the
// ring's z is that of its detector 0, as every reference trace loop does.
__device__ __forceinline__ void lor_line(const double* geom, int n_det, int ra, int d1, int rb, int d2,
                                         double* lc)
{
    lc[0] = geom[((int64_t)ra * n_det + d1) * 3 + 0];
    lc[1] = geom[((int64_t)ra * n_det + d1) * 3 + 1];
    lc[2] = geom[((int64_t)ra * n_det + 0) * 3 + 2];
    lc[3] = geom[((int64_t)rb * n_det + d2) * 3 + 0];
    lc[4] = geom[((int64_t)rb * n_det + d2) * 3 + 1];
    lc[5] = geom[((int64_t)rb * n_det + 0) * 3 + 2];
}

__global__ void k_tail_mask(Grid g, const double* __restrict__ image, int use_attenuation,
                            const double* __restrict__ geom, int n_det, const int16_t* __restrict__ lor,
                            int n_planes, const int32_t* __restrict__ plane_rings, int accel, int nd,
                            uint8_t* __restrict__ mask)
{
    const int64_t per_plane = (int64_t)nd * nd;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per_plane * n_planes) return;
    const int p = (int)(i / per_plane);
    const int64_t rem = i - (int64_t)p * per_plane;
    const int d1 = (int)(rem / nd) * accel;
    const int d2 = (int)(rem % nd) * accel;
    const int ra = plane_rings[2 * p], rb = plane_rings[2 * p + 1];
    if ((d1 == d2 && ra == rb) || d1 > d2) return;
    double lc[6];
    lor_line(geom, n_det, ra, d1, rb, d2, lc);
    Walk w;
    double sum = 0.0;
    int got = 0;
    if (walk_setup(g, lc, w))
        got = walk_run(g, w, [&](int, int vox, double L, double) { sum += image[vox] * L; });
    uint8_t v;
    if (use_attenuation) {
        // 1/exp(-sum) in float64; a LOR that misses the grid has factor 1
        const double att = got > 0 ? 1.0 / exp(-sum) : 1.0;
        v = att == 1.0 ? 1 : 0;
    } else {
        v = (got > 0 && sum > 0.0) ? 0 : 1;
    }
    const int ang = lor[((int64_t)d1 * n_det + d2) * 2 + 0];
    const int rad = lor[((int64_t)d1 * n_det + d2) * 2 + 1];
    if (ang < 0 || rad < 0) return;
    mask[((int64_t)rad * (n_det / 2) + ang) * n_planes + p] = v;
}

// Fixed-point images of the tail fit.  Each contribution v to the prompt,
// scatter and randoms images is added as the integer llrint(v 2^f) with
// 64-bit integer atomics; integer addition is associative, so the images and
// the fitted scale are the same bits in every run whatever order the LORs
// land in.  f is chosen per image from the largest masked value |v|max of its
// sinogram and a bound on the LORs: a walk visits a voxel at most once and a
// LOR adds at most |v|max to it (value / n, or value L / sum L), so no voxel
// sum exceeds |v|max n_lors < 2^61 2^-f.  The sensitivity image counts LORs.
__device__ __forceinline__ double fit_scale(const unsigned int* vmax, int i, double n_lors)
{
    const double m = (double)__uint_as_float(vmax[i]);
    if (!(m > 0.0) || !isfinite(m)) return 1.0;   // all zero, or a non-finite sinogram
    int e;
    (void)frexp(m * n_lors, &e);   // m n_lors < 2^e
    return ldexp(1.0, 61 - e);
}

__device__ __forceinline__ void fit_add(unsigned long long* p, double v, double scale)
{
    const long long q = llrint(v * scale);
    if (q != 0) atomicAdd(p, (unsigned long long)q);
}

// Every fit image and the n_vmax maxima to zero (a kernel, so the whole fit
// can be captured in a graph: no memset nodes).  n_vmax <= the grid's threads.
__global__ void k_fit_clear(unsigned long long* __restrict__ img, int64_t n, unsigned int* __restrict__ vmax,
                            int n_vmax)
{
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i0 < n_vmax) vmax[i0] = 0u;
    for (int64_t i = i0; i < n; i += (int64_t)gridDim.x * blockDim.x) img[i] = 0ull;
}

// Largest |masked value| of the prompt, scatter and randoms sinograms (float
// bits of non-negative floats order as unsigned integers, so the atomic max
// is exact and order-independent).
__global__ void k_fit_max(const float* __restrict__ scatter, const float* __restrict__ prompts,
                          const float* __restrict__ randoms, const float* __restrict__ maskv, int64_t n,
                          unsigned int* __restrict__ vmax)
{
    unsigned int mp = 0, ms = 0, mr = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float m = maskv[i];
        mp = max(mp, __float_as_uint(fabsf(m * prompts[i])));
        ms = max(ms, __float_as_uint(fabsf(m * scatter[i])));
        mr = max(mr, __float_as_uint(fabsf(m * randoms[i])));
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        mp = max(mp, (unsigned int)__shfl_xor((int)mp, o));
        ms = max(ms, (unsigned int)__shfl_xor((int)ms, o));
        mr = max(mr, (unsigned int)__shfl_xor((int)mr, o));
    }
    if ((threadIdx.x % kWave) == 0) {
        atomicMax(vmax + 0, mp);
        atomicMax(vmax + 1, ms);
        atomicMax(vmax + 2, mr);
    }
}

// plane_rings: (n_entries, 3) = ring of d1, ring of d2, plane the values are read from.
__global__ void k_fit_bp(Grid g, const double* __restrict__ geom, int n_rings_geom, int n_det,
                         const int16_t* __restrict__ lor, int n_entries,
                         const int32_t* __restrict__ plane_rings, int n_value_planes, int accel, int nd,
                         int length_weighted, double n_lors, const float* __restrict__ scatter,
                         const float* __restrict__ prompts, const float* __restrict__ randoms,
                         const float* __restrict__ maskv, const unsigned int* __restrict__ vmax,
                         unsigned long long* __restrict__ img)
{
    const int64_t per_plane = (int64_t)nd * nd;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per_plane * n_entries) return;
    const int e = (int)(i / per_plane);
    const int64_t rem = i - (int64_t)e * per_plane;
    const int d1 = (int)(rem / nd) * accel;
    const int d2 = (int)(rem % nd) * accel;
    const int ra = plane_rings[3 * e], rb = plane_rings[3 * e + 1], vp = plane_rings[3 * e + 2];
    if ((d1 == d2 && ra == rb) || d1 > d2) return;
    // entries and coordinates may come from device memory unchecked: out of
    // range ones are skipped
    if ((unsigned)ra >= (unsigned)n_rings_geom || (unsigned)rb >= (unsigned)n_rings_geom ||
        (unsigned)vp >= (unsigned)n_value_planes)
        return;
    const int ang = lor[((int64_t)d1 * n_det + d2) * 2 + 0];
    const int rad = lor[((int64_t)d1 * n_det + d2) * 2 + 1];
    if (ang < 0 || rad < 0 || ang >= n_det / 2 || rad > n_det) return;
    const int64_t cidx = ((int64_t)rad * (n_det / 2) + ang) * n_value_planes + vp;
    const float m = maskv[cidx];
    const float sv = m * scatter[cidx];           // float32 products as the reference's masking
    if (sv == 0.0f) return;
    const float pv = m * prompts[cidx];
    const float rv = m * randoms[cidx];
    double lc[6];
    lor_line(geom, n_det, ra, d1, rb, d2, lc);
    Walk w;
    if (!walk_setup(g, lc, w)) return;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const double sc_p = fit_scale(vmax, 0, n_lors), sc_s = fit_scale(vmax, 1, n_lors),
                 sc_r = fit_scale(vmax, 2, n_lors);
    unsigned long long* const ip = img;
    unsigned long long* const is = img + nvox;
    unsigned long long* const ir = img + 2 * nvox;
    unsigned long long* const isens = img + 3 * nvox;
    if (!length_weighted) {
        // value / n per voxel, the division in float32 (BackProjection.py:108-111)
        const int n = walk_run(g, w, [&](int, int, double, double) {});
        if (n <= 0) return;
        const double ps = (double)(pv / (float)n), ss = (double)(sv / (float)n), rs = (double)(rv / (float)n);
        walk_run(g, w, [&](int, int vox, double, double) {
            fit_add(ip + vox, ps, sc_p);
            fit_add(is + vox, ss, sc_s);
            fit_add(ir + vox, rs, sc_r);
            atomicAdd(isens + vox, 1ull);
        });
    } else {
        // value * L / sum(L) per voxel in float64 (BackProjection.py:294-297)
        double lsum = 0.0;
        const int n = walk_run(g, w, [&](int, int, double L, double) { lsum += L; });
        if (n <= 0) return;
        walk_run(g, w, [&](int, int vox, double L, double) {
            fit_add(ip + vox, (double)pv * L / lsum, sc_p);
            fit_add(is + vox, (double)sv * L / lsum, sc_s);
            fit_add(ir + vox, (double)rv * L / lsum, sc_r);
            atomicAdd(isens + vox, 1ull);
        });
    }
}

// Per-voxel ratios and the two normal-equation sums, per block of a fixed
// grid (kFitBlocks x 256 threads, voxels strided over the grid): each block
// writes its partial sums; k_fit_final adds them in a fixed order.
constexpr int kFitThreads = 256;

// shared (may be null): randoms and sensitivity images added to img's
// (integers, so the sums are those of one image holding both).
__global__ __launch_bounds__(kFitThreads) void k_fit_reduce(const unsigned long long* __restrict__ img, int64_t nvox,
                                                           double n_lors, const unsigned int* __restrict__ vmax,
                                                           double* __restrict__ part,
                                                           const unsigned long long* __restrict__ shared)
{
    auto at = [&](int k, int64_t v) {
        unsigned long long a = img[k * nvox + v];
        if (shared && k >= 2) a += shared[(k - 2) * nvox + v];
        return (double)(long long)a;
    };
    __shared__ double sxy[kFitThreads], sxx[kFitThreads];
    const double inv_p = 1.0 / fit_scale(vmax, 0, n_lors), inv_s = 1.0 / fit_scale(vmax, 1, n_lors),
                 inv_r = 1.0 / fit_scale(vmax, 2, n_lors);
    double xy = 0.0, xx = 0.0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (int64_t)gridDim.x * blockDim.x) {
        const double sens = at(3, v);
        if (sens == 0.0) continue;   // 0/0 and x/0 are set to 0 in the reference
        const double sbp = at(1, v) * inv_s / sens;
        const double abp = at(0, v) * inv_p / sens;
        const double rbp = at(2, v) * inv_r / sens;
        double tails = abp - rbp;
        tails = tails < 0.0 ? 0.0 : tails;
        xy += sbp * tails;
        xx += sbp * sbp;
    }
    sxy[threadIdx.x] = xy;
    sxx[threadIdx.x] = xx;
    __syncthreads();
    for (int s = kFitThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sxy[threadIdx.x] += sxy[threadIdx.x + s];
            sxx[threadIdx.x] += sxx[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x + 0] = sxy[0];
        part[2 * blockIdx.x + 1] = sxx[0];
    }
}

// One block: the partial sums in a fixed tree order.
__global__ __launch_bounds__(kFitBlocks) void k_fit_final(const double* __restrict__ part, double* __restrict__ sums)
{
    __shared__ double sxy[kFitBlocks], sxx[kFitBlocks];
    sxy[threadIdx.x] = part[2 * threadIdx.x + 0];
    sxx[threadIdx.x] = part[2 * threadIdx.x + 1];
    __syncthreads();
    for (int s = kFitBlocks / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sxy[threadIdx.x] += sxy[threadIdx.x + s];
            sxx[threadIdx.x] += sxx[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sums[0] = sxy[0];
        sums[1] = sxx[0];
    }
}

// The fixed-point images as float64 (prompts, scatter, randoms, sensitivity).
__global__ void k_fit_images(const unsigned long long* __restrict__ img, int64_t nvox, double n_lors,
                             const unsigned int* __restrict__ vmax, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4 * nvox) return;
    const int k = (int)(i / nvox);
    const double inv = k == 3 ? 1.0 : 1.0 / fit_scale(vmax, k, n_lors);
    out[i] = (double)(long long)img[i] * inv;
}

// ---- the fit of every TOF bin at once (Scaling.py:85-113) ----
// Each LOR is traced once per group of kBinGroup bins instead of once per
// bin; bin b's images are the planar block img + 4 b nvox, laid out as the
// one-bin fit's, followed by the shared randoms and sensitivity images
// (img + 4 n_bins nvox, 2 nvox).
constexpr int kBinGroup = 8;

__device__ __forceinline__ float bin_prompt(const BinFitIn& in, int64_t c, int b)
{
    const int64_t i = c * in.n_bins + b;
    float p = in.prompts_u16 ? (float)static_cast<const uint16_t*>(in.prompts)[i]
                             : static_cast<const float*>(in.prompts)[i];
    // Prompts / LORCounts in float32 (Scaling.py:90)
    if (in.counts) p = p / in.counts[in.counts_per_plane ? c % in.n_value_planes : c];
    return p;
}

__device__ __forceinline__ float bin_randoms(const BinFitIn& in, int64_t c)
{
    if (!in.randoms) return 0.0f;
    const float r = in.randoms[c];
    return in.norm ? r * in.norm[c] : r;   // Randoms * Normalization (Scaling.py:69-71)
}

__device__ __forceinline__ float bin_mask(const BinFitIn& in, int64_t c)
{
    return in.mask[c] >= 1 ? 1.0f : 0.0f;   // float32(mask >= 1) (Scaling.py:75-76)
}

// Largest |masked value| per bin (vmax + 3 b: prompts, scatter, randoms);
// blockIdx.y is the group of bins.
__global__ void k_fitb_max(BinFitIn in, unsigned int* __restrict__ vmax)
{
    const int b0 = blockIdx.y * kBinGroup;
    unsigned int mp[kBinGroup], ms[kBinGroup], mr = 0;
#pragma unroll
    for (int j = 0; j < kBinGroup; ++j) mp[j] = ms[j] = 0u;
    for (int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; c < in.n_cells;
         c += (int64_t)gridDim.x * blockDim.x) {
        const float m = bin_mask(in, c);
        mr = max(mr, __float_as_uint(fabsf(m * bin_randoms(in, c))));
#pragma unroll
        for (int j = 0; j < kBinGroup; ++j) {
            const int b = b0 + j;
            if (b < in.n_bins) {
                mp[j] = max(mp[j], __float_as_uint(fabsf(m * bin_prompt(in, c, b))));
                ms[j] = max(ms[j], __float_as_uint(fabsf(m * in.scatter[(int64_t)b * in.n_cells + c])));
            }
        }
    }
    for (int o = kWave / 2; o > 0; o >>= 1) {
        mr = max(mr, (unsigned int)__shfl_xor((int)mr, o));
#pragma unroll
        for (int j = 0; j < kBinGroup; ++j) {
            mp[j] = max(mp[j], (unsigned int)__shfl_xor((int)mp[j], o));
            ms[j] = max(ms[j], (unsigned int)__shfl_xor((int)ms[j], o));
        }
    }
    if ((threadIdx.x % kWave) == 0) {
#pragma unroll
        for (int j = 0; j < kBinGroup; ++j) {
            const int b = b0 + j;
            if (b < in.n_bins) {
                atomicMax(vmax + 3 * b + 0, mp[j]);
                atomicMax(vmax + 3 * b + 1, ms[j]);
                atomicMax(vmax + 3 * b + 2, mr);
            }
        }
    }
}

// k_fit_bp for every bin: one thread per (entry, d1, d2); the LOR's voxels
// are walked once for their count and length sum, then once per group of
// bins in which some bin has a nonzero masked scatter value.  Per bin the
// integers added are those of k_fit_bp on that bin's sinograms.
__global__ __launch_bounds__(128) void k_fitb_bp(Grid g, const double* __restrict__ geom, int n_rings_geom,
                                                 int n_det, const int16_t* __restrict__ lor, int n_entries,
                                                 const int32_t* __restrict__ plane_rings, int accel, int nd,
                                                 int length_weighted, double n_lors, BinFitIn in,
                                                 const unsigned int* __restrict__ vmax,
                                                 unsigned long long* __restrict__ img)
{
    const int64_t per_plane = (int64_t)nd * nd;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= per_plane * n_entries) return;
    const int e = (int)(i / per_plane);
    const int64_t rem = i - (int64_t)e * per_plane;
    const int d1 = (int)(rem / nd) * accel;
    const int d2 = (int)(rem % nd) * accel;
    const int ra = plane_rings[3 * e], rb = plane_rings[3 * e + 1], vp = plane_rings[3 * e + 2];
    if ((d1 == d2 && ra == rb) || d1 > d2) return;
    if ((unsigned)ra >= (unsigned)n_rings_geom || (unsigned)rb >= (unsigned)n_rings_geom ||
        (unsigned)vp >= (unsigned)in.n_value_planes)
        return;
    const int ang = lor[((int64_t)d1 * n_det + d2) * 2 + 0];
    const int rad = lor[((int64_t)d1 * n_det + d2) * 2 + 1];
    if (ang < 0 || rad < 0 || ang >= n_det / 2 || rad > n_det) return;
    const int64_t cidx = ((int64_t)rad * (n_det / 2) + ang) * in.n_value_planes + vp;
    const float m = bin_mask(in, cidx);
    const float rv = m * bin_randoms(in, cidx);
    double lc[6];
    lor_line(geom, n_det, ra, d1, rb, d2, lc);
    Walk w;
    bool set = false, walked = false;
    int n = 0;
    double lsum = 0.0;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    // A LOR with a nonzero masked scatter value in every bin adds the same
    // randoms integers (one randoms maximum, so one scale, for all bins) and
    // the same sensitivity count to every bin: those go once into the shared
    // images, which the reduction adds to each bin's own.
    bool full = true;
    for (int b = 0; b < in.n_bins && full; ++b) full = m * in.scatter[(int64_t)b * in.n_cells + cidx] != 0.0f;
    unsigned long long* const shared = img + 4 * nvox * in.n_bins;
    for (int b0 = 0; b0 < in.n_bins; b0 += kBinGroup) {
        float sv[kBinGroup], pv[kBinGroup];
        unsigned act = 0;
#pragma unroll
        for (int j = 0; j < kBinGroup; ++j) {
            const int b = b0 + j;
            sv[j] = b < in.n_bins ? m * in.scatter[(int64_t)b * in.n_cells + cidx] : 0.0f;
            if (sv[j] != 0.0f) act |= 1u << j;
        }
        if (!act) continue;
        if (!walked) {
            walked = true;
            set = walk_setup(g, lc, w);
            if (set) n = walk_run(g, w, [&](int, int, double L, double) { lsum += L; });
        }
        if (!set || n <= 0) return;
        const bool own_rs = !full;           // randoms and sensitivity per bin
        const bool add_shared = full && b0 == 0;
#pragma unroll
        for (int j = 0; j < kBinGroup; ++j) pv[j] = (act >> j) & 1u ? m * bin_prompt(in, cidx, b0 + j) : 0.0f;
        if (!length_weighted) {
            // value / n per voxel, the division in float32: one integer per bin and image
            long long q[kBinGroup][3];
#pragma unroll
            for (int j = 0; j < kBinGroup; ++j) {
                const unsigned int* vm = vmax + 3 * (b0 + j);
                q[j][0] = q[j][1] = q[j][2] = 0;
                if ((act >> j) & 1u) {
                    q[j][0] = llrint((double)(pv[j] / (float)n) * fit_scale(vm, 0, n_lors));
                    q[j][1] = llrint((double)(sv[j] / (float)n) * fit_scale(vm, 1, n_lors));
                    q[j][2] = llrint((double)(rv / (float)n) * fit_scale(vm, 2, n_lors));
                }
            }
            const long long qr = add_shared ? q[0][2] : 0;
            walk_run(g, w, [&](int, int vox, double, double) {
                if (add_shared) {
                    if (qr) atomicAdd(shared + vox, (unsigned long long)qr);
                    atomicAdd(shared + nvox + vox, 1ull);
                }
#pragma unroll
                for (int j = 0; j < kBinGroup; ++j) {
                    if (!((act >> j) & 1u)) continue;
                    unsigned long long* const base = img + (int64_t)(b0 + j) * 4 * nvox + vox;
                    if (q[j][0]) atomicAdd(base, (unsigned long long)q[j][0]);
                    if (q[j][1]) atomicAdd(base + nvox, (unsigned long long)q[j][1]);
                    if (own_rs) {
                        if (q[j][2]) atomicAdd(base + 2 * nvox, (unsigned long long)q[j][2]);
                        atomicAdd(base + 3 * nvox, 1ull);
                    }
                }
            });
        } else {
            // value * L / sum(L) per voxel in float64
            double sc[kBinGroup][3];
#pragma unroll
            for (int j = 0; j < kBinGroup; ++j) {
                const unsigned int* vm = vmax + 3 * (b0 + j);
                sc[j][0] = sc[j][1] = sc[j][2] = 0.0;
                if ((act >> j) & 1u) {
                    sc[j][0] = fit_scale(vm, 0, n_lors);
                    sc[j][1] = fit_scale(vm, 1, n_lors);
                    sc[j][2] = fit_scale(vm, 2, n_lors);
                }
            }
            const double scr = sc[0][2];
            walk_run(g, w, [&](int, int vox, double L, double) {
                if (add_shared) {
                    fit_add(shared + vox, (double)rv * L / lsum, scr);
                    atomicAdd(shared + nvox + vox, 1ull);
                }
#pragma unroll
                for (int j = 0; j < kBinGroup; ++j) {
                    if (!((act >> j) & 1u)) continue;
                    unsigned long long* const base = img + (int64_t)(b0 + j) * 4 * nvox + vox;
                    fit_add(base, (double)pv[j] * L / lsum, sc[j][0]);
                    fit_add(base + nvox, (double)sv[j] * L / lsum, sc[j][1]);
                    if (own_rs) {
                        fit_add(base + 2 * nvox, (double)rv * L / lsum, sc[j][2]);
                        atomicAdd(base + 3 * nvox, 1ull);
                    }
                }
            });
        }
    }
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_interp(const float* sampled, int n_bins, int n_sampled_planes, int64_t cells_s,
                         const int32_t* cell_map, int64_t n_cells_out, const int32_t* plane_map,
                         int n_out_planes, float* out, float* sum, hipStream_t st)
{
    const int64_t n = n_cells_out * n_out_planes;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_interp, dim3(blocks_for(n, 256)), dim3(256), 0, st, sampled, n_bins,
                       n_sampled_planes, cells_s, cell_map, n_cells_out, plane_map, n_out_planes, out, sum);
    return hipGetLastError();
}

hipError_t launch_tail_mask(const Grid& g, const double* image, int use_attenuation, const double* geom,
                            int n_det, const int16_t* lor, int n_planes, const int32_t* plane_rings,
                            int accel, uint8_t* mask, hipStream_t st)
{
    const int nd = (n_det + accel - 1) / accel;
    const int64_t n = (int64_t)nd * nd * n_planes;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_tail_mask, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, image, use_attenuation,
                       geom, n_det, lor, n_planes, plane_rings, accel, nd, mask);
    return hipGetLastError();
}

FitScratch fit_scratch_carve(void* base, int64_t nvox)
{
    unsigned char* p = static_cast<unsigned char*>(base);
    FitScratch f;
    f.img = reinterpret_cast<unsigned long long*>(p);
    p += (size_t)(4 * nvox) * sizeof(unsigned long long);
    f.part = reinterpret_cast<double*>(p);
    p += (size_t)(2 * kFitBlocks) * sizeof(double);
    f.vmax = reinterpret_cast<unsigned int*>(p);
    return f;
}

hipError_t launch_fit(const Grid& g, const double* geom, int n_rings_geom, int n_det, const int16_t* lor, int n_entries,
                      const int32_t* plane_rings, int n_value_planes, int accel, int length_weighted,
                      const float* scatter, const float* prompts, const float* randoms, const float* mask,
                      FitScratch scratch, double* sums, double* images_out, hipStream_t st)
{
    const int nd = (n_det + accel - 1) / accel;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const int64_t n_sino = (int64_t)(n_det + 1) * (n_det / 2) * n_value_planes;
    // every LOR the fit can trace: entries x detector pairs d1 <= d2
    const double n_lors = (double)n_entries * ((double)nd * (nd + 1) / 2) + 1.0;
    hipLaunchKernelGGL(k_fit_clear, dim3(1024), dim3(256), 0, st, scratch.img, 4 * nvox, scratch.vmax, 3);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (n_sino > 0) {
        hipLaunchKernelGGL(k_fit_max, dim3((unsigned)std::min<int64_t>(1024, blocks_for(n_sino, 256))), dim3(256), 0,
                           st, scatter, prompts, randoms, mask, n_sino, scratch.vmax);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int64_t n = (int64_t)nd * nd * n_entries;
    if (n > 0) {
        hipLaunchKernelGGL(k_fit_bp, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, geom, n_rings_geom, n_det, lor,
                           n_entries, plane_rings, n_value_planes, accel, nd, length_weighted, n_lors, scatter,
                           prompts, randoms, mask, scratch.vmax, scratch.img);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_fit_reduce, dim3(kFitBlocks), dim3(kFitThreads), 0, st, scratch.img, nvox, n_lors,
                       scratch.vmax, scratch.part, nullptr);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_fit_final, dim3(1), dim3(kFitBlocks), 0, st, scratch.part, sums);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (images_out) {
        hipLaunchKernelGGL(k_fit_images, dim3(blocks_for(4 * nvox, 256)), dim3(256), 0, st, scratch.img, nvox,
                           n_lors, scratch.vmax, images_out);
        e = hipGetLastError();
    }
    return e;
}

FitScratch fit_bins_scratch_carve(void* base, int64_t nvox, int n_bins)
{
    unsigned char* p = static_cast<unsigned char*>(base);
    FitScratch f;
    f.img = reinterpret_cast<unsigned long long*>(p);
    p += (size_t)((4 * n_bins + 2) * nvox) * sizeof(unsigned long long);
    f.part = reinterpret_cast<double*>(p);
    p += (size_t)(2 * kFitBlocks) * n_bins * sizeof(double);
    f.vmax = reinterpret_cast<unsigned int*>(p);
    return f;
}

hipError_t launch_fit_bins(const Grid& g, const double* geom, int n_rings_geom, int n_det, const int16_t* lor,
                           int n_entries, const int32_t* plane_rings, int accel, int length_weighted,
                           const BinFitIn& in, FitScratch scratch, double* sums, hipStream_t st)
{
    const int nd = (n_det + accel - 1) / accel;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const double n_lors = (double)n_entries * ((double)nd * (nd + 1) / 2) + 1.0;   // as launch_fit
    // the clear grid (1024 x 256 threads) covers the 3 n_bins maxima for n_bins <= 87381
    hipLaunchKernelGGL(k_fit_clear, dim3(1024), dim3(256), 0, st, scratch.img, (4 * in.n_bins + 2) * nvox,
                       scratch.vmax, 3 * in.n_bins);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const unsigned groups = (unsigned)((in.n_bins + kBinGroup - 1) / kBinGroup);
    if (in.n_cells > 0) {
        hipLaunchKernelGGL(k_fitb_max, dim3((unsigned)std::min<int64_t>(1024, blocks_for(in.n_cells, 256)), groups),
                           dim3(256), 0, st, in, scratch.vmax);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    const int64_t n = (int64_t)nd * nd * n_entries;
    if (n > 0) {
        hipLaunchKernelGGL(k_fitb_bp, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, geom, n_rings_geom, n_det, lor,
                           n_entries, plane_rings, accel, nd, length_weighted, n_lors, in, scratch.vmax, scratch.img);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    for (int b = 0; b < in.n_bins; ++b) {
        double* part = scratch.part + (int64_t)2 * kFitBlocks * b;
        hipLaunchKernelGGL(k_fit_reduce, dim3(kFitBlocks), dim3(kFitThreads), 0, st, scratch.img + 4 * nvox * b,
                           nvox, n_lors, scratch.vmax + 3 * b, part, scratch.img + 4 * nvox * in.n_bins);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_fit_final, dim3(1), dim3(kFitBlocks), 0, st, part, sums + 2 * b);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    return e;
}

}  // namespace sss
