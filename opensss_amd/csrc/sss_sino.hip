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
//                 randoms backprojected along every surviving LOR into float64
//                 images plus the sensitivity image.
//   k_fit_reduce  BackProjection.py:191-210: per-voxel ratios and the float64
//                 normal equations sum(x y), sum(x x) of the one-parameter fit.
#include "sss_walk.h"

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

// plane_rings: (n_entries, 3) = ring of d1, ring of d2, plane the values are read from.
__global__ void k_fit_bp(Grid g, const double* __restrict__ geom, int n_det,
                         const int16_t* __restrict__ lor, int n_entries,
                         const int32_t* __restrict__ plane_rings, int n_value_planes, int accel, int nd,
                         int length_weighted, const float* __restrict__ scatter,
                         const float* __restrict__ prompts, const float* __restrict__ randoms,
                         const float* __restrict__ maskv, double* __restrict__ img)
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
    const int ang = lor[((int64_t)d1 * n_det + d2) * 2 + 0];
    const int rad = lor[((int64_t)d1 * n_det + d2) * 2 + 1];
    if (ang < 0 || rad < 0) return;
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
    if (!length_weighted) {
        // value / n per voxel, the division in float32 (BackProjection.py:108-111)
        const int n = walk_run(g, w, [&](int, int, double, double) {});
        if (n <= 0) return;
        const double ps = (double)(pv / (float)n), ss = (double)(sv / (float)n), rs = (double)(rv / (float)n);
        walk_run(g, w, [&](int, int vox, double, double) {
            atomicAdd(img + 0 * nvox + vox, ps);
            atomicAdd(img + 1 * nvox + vox, ss);
            atomicAdd(img + 2 * nvox + vox, rs);
            atomicAdd(img + 3 * nvox + vox, 1.0);
        });
    } else {
        // value * L / sum(L) per voxel in float64 (BackProjection.py:294-297)
        double lsum = 0.0;
        const int n = walk_run(g, w, [&](int, int, double L, double) { lsum += L; });
        if (n <= 0) return;
        walk_run(g, w, [&](int, int vox, double L, double) {
            atomicAdd(img + 0 * nvox + vox, (double)pv * L / lsum);
            atomicAdd(img + 1 * nvox + vox, (double)sv * L / lsum);
            atomicAdd(img + 2 * nvox + vox, (double)rv * L / lsum);
            atomicAdd(img + 3 * nvox + vox, 1.0);
        });
    }
}

__global__ void k_fit_reduce(const double* __restrict__ img, int64_t nvox, double* __restrict__ sums)
{
    __shared__ double sxy[256], sxx[256];
    double xy = 0.0, xx = 0.0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < nvox;
         v += (int64_t)gridDim.x * blockDim.x) {
        const double sens = img[3 * nvox + v];
        if (sens == 0.0) continue;   // 0/0 and x/0 are set to 0 in the reference
        const double sbp = img[1 * nvox + v] / sens;
        const double abp = img[0 * nvox + v] / sens;
        const double rbp = img[2 * nvox + v] / sens;
        double tails = abp - rbp;
        tails = tails < 0.0 ? 0.0 : tails;
        xy += sbp * tails;
        xx += sbp * sbp;
    }
    sxy[threadIdx.x] = xy;
    sxx[threadIdx.x] = xx;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sxy[threadIdx.x] += sxy[threadIdx.x + s];
            sxx[threadIdx.x] += sxx[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        atomicAdd(sums + 0, sxy[0]);
        atomicAdd(sums + 1, sxx[0]);
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

hipError_t launch_fit(const Grid& g, const double* geom, int n_det, const int16_t* lor, int n_entries,
                      const int32_t* plane_rings, int n_value_planes, int accel, int length_weighted,
                      const float* scatter, const float* prompts, const float* randoms, const float* mask,
                      double* img, double* sums, hipStream_t st)
{
    const int nd = (n_det + accel - 1) / accel;
    const int64_t n = (int64_t)nd * nd * n_entries;
    if (n > 0)
        hipLaunchKernelGGL(k_fit_bp, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, geom, n_det, lor,
                           n_entries, plane_rings, n_value_planes, accel, nd, length_weighted, scatter,
                           prompts, randoms, mask, img);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    hipLaunchKernelGGL(k_fit_reduce, dim3(256), dim3(256), 0, st, img, nvox, sums);
    return hipGetLastError();
}

}  // namespace sss
