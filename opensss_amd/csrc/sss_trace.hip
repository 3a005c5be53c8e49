// Voxel ray tracer (float64) and the half-path table kernels.
//
// The walk restates Python/openSSS/RayTracing3DTOF.py:29-180 (and its box
// test :204-269) with the same float64 operation order: slab clip with the
// 1e-20 guards, t clamped to [0, 1], start/end voxels by ceil and clamp,
// Amanatides-Woo stepping with ties resolved x < y (strict), x < z (strict),
// else y < z (strict), else z; length = |p_prev - p_cross|; TOF sample
// s_k = s_{k-1} + L_k / 2.  Fused multiply-add contraction is disabled in
// this file so every crossing parameter and length is bit-identical to the
// reference's float64 arithmetic.
#include "sss_internal.h"

#pragma clang fp contract(off)

namespace sss {
namespace {

struct Walk {
    double X, Y, Z, xEnd, yEnd, zEnd;
    double xStep, yStep, zStep;
    double xTD, yTD, zTD, xTM, yTM, zTM;
    double xVec, yVec, zVec, tMin, tMax;
    int n;  // voxels the walk visits: |dX| + |dY| + |dZ| + 1
};

__device__ __forceinline__ bool box_intersect(const double* gb, const double* lc, double& tmin,
                                              double& tmax)
{
    double x_div = 1 / (lc[3] - lc[0] + 1e-20);
    double t_min, t_max;
    if (x_div >= 0) { t_min = (gb[0] - lc[0]) * x_div; t_max = (gb[3] - lc[0]) * x_div; }
    else            { t_min = (gb[3] - lc[0]) * x_div; t_max = (gb[0] - lc[0]) * x_div; }
    double y_div = 1 / (lc[4] - lc[1] + 1e-20);
    double y_t_min, y_t_max;
    if (y_div >= 0) { y_t_min = (gb[1] - lc[1]) * y_div; y_t_max = (gb[4] - lc[1]) * y_div; }
    else            { y_t_min = (gb[4] - lc[1]) * y_div; y_t_max = (gb[1] - lc[1]) * y_div; }
    if (t_min > y_t_max || y_t_min > t_max) return false;
    if (y_t_min > t_min) t_min = y_t_min;
    if (y_t_max < t_max) t_max = y_t_max;
    double z_div = 1 / (lc[5] - lc[2] + 1e-20);
    double z_t_min, z_t_max;
    if (z_div >= 0) { z_t_min = (gb[2] - lc[2]) * z_div; z_t_max = (gb[5] - lc[2]) * z_div; }
    else            { z_t_min = (gb[5] - lc[2]) * z_div; z_t_max = (gb[2] - lc[2]) * z_div; }
    if (t_min > z_t_max || z_t_min > t_max) return false;
    if (z_t_min > t_min) t_min = z_t_min;
    if (z_t_max < t_max) t_max = z_t_max;
    if ((t_min >= 1 && t_max >= 1) || (t_min <= 0 && t_max <= 0)) return false;
    tmin = t_min;
    tmax = t_max;
    return true;
}

__device__ __forceinline__ void axis_setup(double start, double end, double vec, double lo,
                                           double vox, double nvox, double tMin, double tMax,
                                           double& V, double& VEnd, double& step, double& td,
                                           double& tm)
{
    V = fmax(1.0, fmin(nvox, ceil((start - lo) / vox)));
    VEnd = fmax(1.0, fmin(nvox, ceil((end - lo) / vox)));
    if (vec > 0) { step = 1; td = vox / vec; tm = tMin + (lo + V * vox - start) / vec; }
    else if (vec < 0) { step = -1; td = vox / -vec; tm = tMin + (lo + (V - 1) * vox - start) / vec; }
    else { step = 0; tm = tMax; td = tMax; }
}

__device__ __forceinline__ bool walk_setup(const Grid& g, const double* lc, Walk& w)
{
    double tMin, tMax;
    if (!box_intersect(g.b, lc, tMin, tMax)) { w.n = 0; return false; }
    tMin = tMin > 0 ? tMin : 0.0;
    tMax = tMax < 1 ? tMax : 1.0;
    const double xv = (g.b[3] - g.b[0]) / g.nx;
    const double yv = (g.b[4] - g.b[1]) / g.ny;
    const double zv = (g.b[5] - g.b[2]) / g.nz;
    w.xVec = lc[3] - lc[0];
    w.yVec = lc[4] - lc[1];
    w.zVec = lc[5] - lc[2];
    const double xs = lc[0] + w.xVec * tMin, ys = lc[1] + w.yVec * tMin, zs = lc[2] + w.zVec * tMin;
    const double xe = lc[0] + w.xVec * tMax, ye = lc[1] + w.yVec * tMax, ze = lc[2] + w.zVec * tMax;
    axis_setup(xs, xe, w.xVec, g.b[0], xv, (double)g.nx, tMin, tMax, w.X, w.xEnd, w.xStep, w.xTD, w.xTM);
    axis_setup(ys, ye, w.yVec, g.b[1], yv, (double)g.ny, tMin, tMax, w.Y, w.yEnd, w.yStep, w.yTD, w.yTM);
    axis_setup(zs, ze, w.zVec, g.b[2], zv, (double)g.nz, tMin, tMax, w.Z, w.zEnd, w.zStep, w.zTD, w.zTM);
    w.tMin = tMin;
    w.tMax = tMax;
    w.n = (int)(fabs(w.xEnd - w.X) + fabs(w.yEnd - w.Y) + fabs(w.zEnd - w.Z)) + 1;
    return true;
}

// Voxel linear index of the reference, 0-based: X + (Y-1) Ny + (Z-1) Ny Nx - 1.
__device__ __forceinline__ int voxel_index(const Grid& g, double X, double Y, double Z)
{
    return (int)(X + (Y - 1) * g.ny + (Z - 1) * g.ny * g.nx) - 1;
}

// Runs the walk, calling f(k, voxel, length, sample) for each visited voxel;
// returns the visited count (always w.n).  Every walk that reaches its end
// voxel visits exactly w.n voxels.  When a segment runs (almost) inside a
// voxel face the float64 end voxel can disagree with where the steps lead
// and the reference loop never terminates; the walk is then stopped after
// w.n voxels, the last one closed at tMax, and *truncated is set.
template <typename F>
__device__ __forceinline__ int walk_run(const Grid& g, Walk w, F&& f, bool* truncated = nullptr)
{
    int cur = voxel_index(g, w.X, w.Y, w.Z);
    double px = w.xVec * w.tMin, py = w.yVec * w.tMin, pz = w.zVec * w.tMin;
    double prev = 0;
    int k = 0;
    while (w.X != w.xEnd || w.Y != w.yEnd || w.Z != w.zEnd) {
        if (k + 1 >= w.n) {
            if (truncated) *truncated = true;
            break;
        }
        double len;
        if (w.xTM < w.yTM) {
            if (w.xTM < w.zTM) { w.X += w.xStep; len = w.xTM; w.xTM += w.xTD; }
            else               { w.Z += w.zStep; len = w.zTM; w.zTM += w.zTD; }
        } else {
            if (w.yTM < w.zTM) { w.Y += w.yStep; len = w.yTM; w.yTM += w.yTD; }
            else               { w.Z += w.zStep; len = w.zTM; w.zTM += w.zTD; }
        }
        const double dx = px - w.xVec * len, dy = py - w.yVec * len, dz = pz - w.zVec * len;
        const double L = sqrt(dx * dx + dy * dy + dz * dz);
        const double s = prev + L / 2;
        prev = s;
        f(k, cur, L, s);
        px = w.xVec * len;
        py = w.yVec * len;
        pz = w.zVec * len;
        ++k;
        cur = voxel_index(g, w.X, w.Y, w.Z);
    }
    const double dx = px - w.xVec * w.tMax, dy = py - w.yVec * w.tMax, dz = pz - w.zVec * w.tMax;
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    f(k, cur, L, prev + L / 2);
    return k + 1;
}

__device__ __forceinline__ void halfpath_line(const double* sp_xyz, const double* detpos, int nru,
                                              int ndu, int64_t h, double* lc, int& det)
{
    const int64_t per_sp = (int64_t)nru * ndu;
    const int64_t sp = h / per_sp;
    det = (int)(h - sp * per_sp);     // ring * NDU + detector
    lc[0] = sp_xyz[3 * sp + 0];
    lc[1] = sp_xyz[3 * sp + 1];
    lc[2] = sp_xyz[3 * sp + 2];
    lc[3] = detpos[3 * det + 0];
    lc[4] = detpos[3 * det + 1];
    lc[5] = detpos[3 * det + 2];
}

__global__ void k_halfpath_count(Grid g, const double* __restrict__ sp_xyz, int n_sp,
                                 const double* __restrict__ detpos,
                                 const double* __restrict__ detnorm, int nru, int ndu,
                                 HalfPathTable t)
{
    const int64_t n_hp = (int64_t)n_sp * nru * ndu;
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n_hp) return;
    double lc[6];
    int det;
    halfpath_line(sp_xyz, detpos, nru, ndu, h, lc, det);
    Walk w;
    walk_setup(g, lc, w);
    t.n[h] = w.n;
    // ScatterVector length and detector-normal angle (SingleScatterSimulationTOF.py:424-430,442-444)
    const double sv0 = lc[3] - lc[0], sv1 = lc[4] - lc[1], sv2 = lc[5] - lc[2];
    t.dist[h] = sqrt(sv0 * sv0 + sv1 * sv1 + sv2 * sv2);
    const double* nv = detnorm + 3 * det;
    const double dot = sv0 * nv[0] + sv1 * nv[1] + sv2 * nv[2];
    const double nn = sqrt(nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2]);
    t.cosabs[h] = (float)fabs(dot / (nn * t.dist[h]));
}

__global__ void k_halfpath_fill(Grid g, const double* __restrict__ sp_xyz, int n_sp,
                                const double* __restrict__ detpos, int nru, int ndu,
                                const double* __restrict__ mu, HalfPathTable t, SampleTable s)
{
    const int64_t n_hp = (int64_t)n_sp * nru * ndu;
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n_hp) return;
    double lc[6];
    int det;
    halfpath_line(sp_xyz, detpos, nru, ndu, h, lc, det);
    Walk w;
    double mu_int = 0;
    int status = 0;
    if (walk_setup(g, lc, w)) {
        const int64_t base = t.off[h];
        const double two_over_c = 2.0 / kLightCmPerPs;
        bool truncated = false;
        walk_run(g, w, [&](int k, int vox, double L, double smp) {
            s.vox[base + k] = vox;
            s.len[base + k] = L;
            s.q[base + k] = smp * two_over_c;
            mu_int += mu[vox] * L;
        }, &truncated);
        if (truncated) status = 1;
    }
    t.lnatt[h] = mu_int;
    t.att[h] = exp(-mu_int);
    t.status[h] = status;
}

__global__ void k_activity_refresh(HalfPathTable t, SampleTable s, int64_t n_hp,
                                   const double* __restrict__ lam)
{
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n_hp) return;
    const int64_t base = t.off[h];
    const int n = t.n[h];
    double act = 0;
    for (int k = 0; k < n; ++k) {
        const double ar = lam[s.vox[base + k]] * s.len[base + k];
        s.a[base + k] = (float)ar;
        act += ar;
    }
    t.act[h] = act;
}

__global__ void k_trace_segments(Grid g, const double* __restrict__ segs, int64_t n, int cap,
                                 int32_t* counts, double* L, int32_t* idx, double* smp)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Walk w;
    if (!walk_setup(g, segs + 6 * i, w) || w.n > cap) {
        counts[i] = w.n > cap ? -1 : 0;   // -1: more voxels than X+Y+Z (not reachable)
        return;
    }
    const int64_t base = i * cap;
    const int got = walk_run(g, w, [&](int k, int vox, double len, double s) {
        L[base + k] = len;
        idx[base + k] = vox + 1;
        smp[base + k] = s;
    });
    counts[i] = got;
}

__global__ void k_line_integrals(Grid g, const double* __restrict__ segs, int64_t n,
                                 const double* __restrict__ mu, const double* __restrict__ lam,
                                 int32_t* counts, double* mu_int, double* lam_int)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Walk w;
    double m = 0, l = 0;
    int got = 0;
    if (walk_setup(g, segs + 6 * i, w)) {
        got = walk_run(g, w, [&](int, int vox, double len, double) {
            m += mu[vox] * len;
            if (lam) l += lam[vox] * len;
        });
    }
    counts[i] = got;
    mu_int[i] = m;
    lam_int[i] = l;
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_halfpath_count(const Grid& g, const double* sp_xyz, int n_sp, const double* detpos,
                                 const double* detnorm, int nru, int ndu, HalfPathTable t,
                                 hipStream_t st)
{
    const int64_t n = (int64_t)n_sp * nru * ndu;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_halfpath_count, dim3(blocks_for(n, 256)), dim3(256), 0, st, g, sp_xyz,
                       n_sp, detpos, detnorm, nru, ndu, t);
    return hipGetLastError();
}

hipError_t launch_halfpath_fill(const Grid& g, const double* sp_xyz, int n_sp, const double* detpos,
                                int nru, int ndu, const double* mu, HalfPathTable t, SampleTable s,
                                hipStream_t st)
{
    const int64_t n = (int64_t)n_sp * nru * ndu;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_halfpath_fill, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, sp_xyz, n_sp,
                       detpos, nru, ndu, mu, t, s);
    return hipGetLastError();
}

hipError_t launch_activity_refresh(const HalfPathTable t, SampleTable s, int64_t n_hp,
                                   const double* lam, hipStream_t st)
{
    if (n_hp == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activity_refresh, dim3(blocks_for(n_hp, 128)), dim3(128), 0, st, t, s, n_hp,
                       lam);
    return hipGetLastError();
}

hipError_t launch_trace_segments(const Grid& g, const double* segs, int64_t n, int cap,
                                 int32_t* counts, double* L, int32_t* idx, double* s,
                                 hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_trace_segments, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, segs, n, cap,
                       counts, L, idx, s);
    return hipGetLastError();
}

hipError_t launch_line_integrals(const Grid& g, const double* segs, int64_t n, const double* mu,
                                 const double* lam, int32_t* counts, double* mu_int,
                                 double* lam_int, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_line_integrals, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, segs, n, mu,
                       lam, counts, mu_int, lam_int);
    return hipGetLastError();
}

}  // namespace sss
