// Synthetic prompts: forward projection of the emission through the
// attenuation, and Poisson sampling of a mean sinogram.
//
//   k_forward_project  per LOR (detector pair d1 <= d2 on the accelerated grid
//                      of a ring pair), the trace of TailsMask.py:139-145:
//                      trues = sum(lam L) exp(-sum(mu L)) in float64, stored
//                      as float32 in the (N+1, N/2, planes) layout.
//   k_poisson          one Poisson draw per sinogram cell from a
//                      counter-based generator (Philox4x32-10, Salmon et al.,
//                      SC'11), so the result depends on (seed, cell) only:
//                      inversion by sequential search below mean 10, the
//                      transformed rejection method PTRS (Hoermann, 1993)
//                      above.
#include "sss_walk.h"

namespace sss {
namespace {

__global__ void k_forward_project(Grid g, const double* __restrict__ lam, const double* __restrict__ mu,
                                  const double* __restrict__ geom, int n_det, const int16_t* __restrict__ lor,
                                  int n_planes, const int32_t* __restrict__ plane_rings, int accel, int nd,
                                  float* __restrict__ out)
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
    const int ang = lor[((int64_t)d1 * n_det + d2) * 2 + 0];
    const int rad = lor[((int64_t)d1 * n_det + d2) * 2 + 1];
    if (ang < 0 || rad < 0) return;
    double lc[6];
    lc[0] = geom[((int64_t)ra * n_det + d1) * 3 + 0];
    lc[1] = geom[((int64_t)ra * n_det + d1) * 3 + 1];
    lc[2] = geom[((int64_t)ra * n_det + 0) * 3 + 2];
    lc[3] = geom[((int64_t)rb * n_det + d2) * 3 + 0];
    lc[4] = geom[((int64_t)rb * n_det + d2) * 3 + 1];
    lc[5] = geom[((int64_t)rb * n_det + 0) * 3 + 2];
    Walk w;
    double sl = 0.0, sm = 0.0;
    if (walk_setup(g, lc, w))
        walk_run(g, w, [&](int, int vox, double L, double) {
#pragma clang fp contract(off)   // sums in the reference's order and rounding (no fused multiply-add)
            sl += lam[vox] * L;
            if (mu) sm += mu[vox] * L;
        });
    out[((int64_t)rad * (n_det / 2) + ang) * n_planes + p] = (float)(sl * exp(-sm));
}

// Philox4x32-10: counter (c0..c3), key (k0, k1).
struct U4 {
    uint32_t x, y, z, w;
};

__device__ __forceinline__ U4 philox(U4 c, uint32_t k0, uint32_t k1)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
        k0 += W0;
        k1 += W1;
    }
    return c;
}

// Uniform in (0, 1) from 53 random bits.
__device__ __forceinline__ double unit53(uint32_t a, uint32_t b)
{
    const uint64_t m = ((uint64_t)(a >> 5) << 26) | (uint64_t)(b >> 6);   // 27 + 26 bits
    return ((double)m + 0.5) * 0x1p-53;
}

// Stream of uniforms for one cell: draw j uses counter (cell, j).
struct CellRng {
    uint64_t cell;
    uint32_t k0, k1, j;
    U4 buf;
    int left;

    __device__ __forceinline__ double next()
    {
        if (left == 0) {
            buf = philox(U4{(uint32_t)cell, (uint32_t)(cell >> 32), j++, 0u}, k0, k1);
            left = 2;
        }
        const double u = left == 2 ? unit53(buf.x, buf.y) : unit53(buf.z, buf.w);
        --left;
        return u;
    }
};

__device__ double poisson_draw(double mu, CellRng& rng)
{
    if (!(mu > 0.0)) return 0.0;
    if (mu < 10.0) {
        // inversion: the smallest k with F(k) >= u
        const double u = rng.next();
        double p = exp(-mu), s = p;
        int k = 0;
        while (u > s && k < 1000) {
            ++k;
            p *= mu / k;
            s += p;
        }
        return (double)k;
    }
    // PTRS
    const double smu = sqrt(mu);
    const double b = 0.931 + 2.53 * smu;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    const double lmu = log(mu);
    for (int it = 0; it < 1000; ++it) {
        const double U = rng.next() - 0.5;
        const double V = rng.next();
        const double us = 0.5 - fabs(U);
        const double k = floor((2.0 * a / us + b) * U + mu + 0.43);
        if (us >= 0.07 && V <= vr) return k;
        if (k < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -mu + k * lmu - lgamma(k + 1.0)) return k;
    }
    return floor(mu);   // not reached in practice (acceptance rate above 0.9)
}

__global__ void k_poisson(const float* __restrict__ mean, int64_t n, double scale, uint32_t k0, uint32_t k1,
                          float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    CellRng rng{(uint64_t)i, k0, k1, 0u, U4{0, 0, 0, 0}, 0};
    out[i] = (float)poisson_draw((double)mean[i] * scale, rng);
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_forward_project(const Grid& g, const double* lam, const double* mu, const double* geom, int n_det,
                                  const int16_t* lor, int n_planes, const int32_t* plane_rings, int accel,
                                  float* out, hipStream_t st)
{
    const int nd = (n_det + accel - 1) / accel;
    const int64_t n = (int64_t)nd * nd * n_planes;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_forward_project, dim3(blocks_for(n, 128)), dim3(128), 0, st, g, lam, mu, geom, n_det, lor,
                       n_planes, plane_rings, accel, nd, out);
    return hipGetLastError();
}

hipError_t launch_poisson(const float* mean, int64_t n, double scale, uint64_t seed, float* out, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_poisson, dim3(blocks_for(n, 256)), dim3(256), 0, st, mean, n, scale, (uint32_t)seed,
                       (uint32_t)(seed >> 32), out);
    return hipGetLastError();
}

}  // namespace sss
