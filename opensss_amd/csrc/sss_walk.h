// Float64 voxel walk shared by every kernel that traces segments.
//
// Restates Python/openSSS/RayTracing3DTOF.py:29-180 (box test :204-269,
// index :182-202) with the same float64 operation order: slab clip with the
// 1e-20 guards, t clamped to [0, 1], start/end voxels by ceil and clamp,
// Amanatides-Woo stepping with ties resolved x < y (strict), x < z (strict),
// else y < z (strict), else z; length = |p_prev - p_cross|; TOF sample
// s_k = s_{k-1} + L_k / 2.  Fused multiply-add contraction is off for these
// functions (the setting sticks to the operations when they are inlined), so
// crossing parameters and lengths are bit-identical to float64 numpy/numba.
#pragma once

#include "sss_internal.h"

namespace sss {

#pragma clang fp contract(off)


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
// returns the visited count.  Every walk that reaches its end voxel visits
// exactly w.n voxels.  When a segment runs (almost) inside a voxel face the
// float64 end voxel can disagree with where the steps lead and the reference
// loop never terminates (it writes past its arrays); the walk is then stopped
// after w.n voxels, or before a step that would leave the grid, the last
// voxel is closed at tMax, and *truncated is set.
template <typename F>
__device__ __forceinline__ int walk_run(const Grid& g, Walk w, F&& f, bool* truncated = nullptr)
{
    int cur = voxel_index(g, w.X, w.Y, w.Z);
    // index steps of a move along x, y, z (voxel_index is linear in X, Y, Z
    // and exact in float64, so the running index equals it at every voxel)
    const int dx_idx = (int)w.xStep, dy_idx = (int)w.yStep * g.ny, dz_idx = (int)w.zStep * g.ny * g.nx;
    double px = w.xVec * w.tMin, py = w.yVec * w.tMin, pz = w.zVec * w.tMin;
    double prev = 0;
    int k = 0;
    while (w.X != w.xEnd || w.Y != w.yEnd || w.Z != w.zEnd) {
        if (k + 1 >= w.n) {
            if (truncated) *truncated = true;
            break;
        }
        // axis of the next crossing, ties as the reference: x only if strictly
        // first, then y if strictly before z, else z
        const int axis = (w.xTM < w.yTM) ? (w.xTM < w.zTM ? 0 : 2) : (w.yTM < w.zTM ? 1 : 2);
        const double nX = axis == 0 ? w.X + w.xStep : w.X;
        const double nY = axis == 1 ? w.Y + w.yStep : w.Y;
        const double nZ = axis == 2 ? w.Z + w.zStep : w.Z;
        if (nX < 1 || nX > g.nx || nY < 1 || nY > g.ny || nZ < 1 || nZ > g.nz) {
            if (truncated) *truncated = true;
            break;
        }
        // selects rather than branches: lanes of a wave step along
        // different axes, and every value is the one the branch would give
        const double len = axis == 0 ? w.xTM : (axis == 1 ? w.yTM : w.zTM);
        w.xTM = axis == 0 ? w.xTM + w.xTD : w.xTM;
        w.yTM = axis == 1 ? w.yTM + w.yTD : w.yTM;
        w.zTM = axis == 2 ? w.zTM + w.zTD : w.zTM;
        w.X = nX;
        w.Y = nY;
        w.Z = nZ;
        const int next = cur + (axis == 0 ? dx_idx : (axis == 1 ? dy_idx : dz_idx));
        const double dx = px - w.xVec * len, dy = py - w.yVec * len, dz = pz - w.zVec * len;
        const double L = sqrt(dx * dx + dy * dy + dz * dz);
        const double s = prev + L / 2;
        prev = s;
        f(k, cur, L, s);
        px = w.xVec * len;
        py = w.yVec * len;
        pz = w.zVec * len;
        ++k;
        cur = next;
    }
    const double dx = px - w.xVec * w.tMax, dy = py - w.yVec * w.tMax, dz = pz - w.zVec * w.tMax;
    const double L = sqrt(dx * dx + dy * dy + dz * dz);
    f(k, cur, L, prev + L / 2);
    return k + 1;
}


#pragma clang fp contract(fast)

}  // namespace sss
