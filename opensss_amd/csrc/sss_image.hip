// Image preparation before the scatter path (Python/openSSS/Utils.py:212-267,
// CropAndDownscale): the crop (or zero pad) around the image centre, then the
// resize of scikit-image (resize(order=3, mode='edge', anti_aliasing=True,
// preserve_range=True, clip=True)), as that library states it:
//
//   k_crop          the cropped window of the input, zero outside it
//   k_gauss_axis    anti-aliasing Gaussian along one axis (sigma = (f - 1) / 2
//                   for a downscale factor f, taps from the host), edge
//                   values repeated, float64 sums stored as float32
//   k_spline_pad / k_spline_axis
//                   cubic B-spline coefficients (Unser's recursive filter,
//                   pole sqrt(3) - 2, half-sample symmetric start) of the
//                   image padded by 12 repeated edge voxels, float64
//   k_zoom          cubic B-spline value at each output voxel's centre mapped
//                   back ((o + 1/2) in/out - 1/2), float32
//   k_minmax / k_clip  the output clipped to the cropped image's range, and
//                   values below a threshold set to 0 (Utils.py:72)
//
// Volumes are x-fastest (the Fortran order of (X, Y, Z)).
#include "sss_internal.h"

#include <algorithm>
#include <cmath>

namespace sss {
namespace {

#pragma clang fp contract(off)

constexpr int kSplinePad = 12;

struct Vol {
    int64_t n[3];
    __host__ __device__ int64_t size() const { return n[0] * n[1] * n[2]; }
    __host__ __device__ int64_t at(int64_t x, int64_t y, int64_t z) const { return x + n[0] * (y + n[1] * z); }
};

__device__ __forceinline__ void coords(const Vol& v, int64_t i, int64_t* c)
{
    c[0] = i % v.n[0];
    const int64_t r = i / v.n[0];
    c[1] = r % v.n[1];
    c[2] = r / v.n[1];
}

__global__ void k_crop(const float* __restrict__ in, Vol vin, Vol vc, int64_t s0, int64_t s1, int64_t s2,
                       float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= vc.size()) return;
    int64_t c[3];
    coords(vc, i, c);
    const int64_t x = c[0] + s0, y = c[1] + s1, z = c[2] + s2;
    const bool inside = x >= 0 && y >= 0 && z >= 0 && x < vin.n[0] && y < vin.n[1] && z < vin.n[2];
    out[i] = inside ? in[vin.at(x, y, z)] : 0.0f;
}

// out = sum_j w[j] in[i + j - r] along `axis`, indices clamped to the line;
// the symmetric taps paired as w[j] (in[i - j] + in[i + j]), far taps first.
__global__ void k_gauss_axis(const float* __restrict__ in, Vol v, int axis, const double* __restrict__ w, int r,
                             float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= v.size()) return;
    int64_t c[3];
    coords(v, i, c);
    const int64_t len = v.n[axis];
    const int64_t stride = axis == 0 ? 1 : (axis == 1 ? v.n[0] : v.n[0] * v.n[1]);
    const int64_t p = c[axis];
    const float* line = in + (i - p * stride);
    double acc = (double)line[p * stride] * w[r];
    for (int j = r; j >= 1; --j) {
        const int64_t a = p - j < 0 ? 0 : p - j;
        const int64_t b = p + j > len - 1 ? len - 1 : p + j;
        acc += ((double)line[a * stride] + (double)line[b * stride]) * w[r + j];
    }
    out[i] = (float)acc;
}

// The float32 image padded by kSplinePad repeated edge voxels on every side, float64.
__global__ void k_spline_pad(const float* __restrict__ in, Vol v, Vol vp, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= vp.size()) return;
    int64_t c[3];
    coords(vp, i, c);
    int64_t s[3];
    for (int k = 0; k < 3; ++k) {
        const int64_t q = c[k] - kSplinePad;
        s[k] = q < 0 ? 0 : (q > v.n[k] - 1 ? v.n[k] - 1 : q);
    }
    out[i] = (double)in[v.at(s[0], s[1], s[2])];
}

// One line per thread: c_k = gain x_k; causal pass c_k += z c_{k-1} from a
// start that sums the half-sample symmetric extension, anticausal pass
// c_k = z (c_{k+1} - c_k) from c_{n-1} z / (z - 1).
__global__ void k_spline_axis(double* __restrict__ data, Vol v, int axis)
{
    const int64_t len = v.n[axis];
    const int64_t lines = v.size() / len;
    const int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= lines || len < 2) return;
    int64_t stride, base;
    if (axis == 0) {
        stride = 1;
        base = l * len;
    } else if (axis == 1) {
        stride = v.n[0];
        base = (l % v.n[0]) + (l / v.n[0]) * v.n[0] * v.n[1];
    } else {
        stride = v.n[0] * v.n[1];
        base = l;
    }
    double* c = data + base;
    const double z = sqrt(3.0) - 2.0;
    const double gain = (1.0 - z) * (1.0 - 1.0 / z);
    for (int64_t k = 0; k < len; ++k) c[k * stride] *= gain;
    // causal start: c_0 + sum over the mirrored line (symmetric about -1/2 and n - 1/2)
    {
        const double zn = pow(z, (double)len);
        const double c0 = c[0];
        double s = c[0] + zn * c[(len - 1) * stride];
        double zi = z;
        for (int64_t k = 1; k < len; ++k) {
            s += zi * (c[k * stride] + zn * c[(len - 1 - k) * stride]);
            zi *= z;
        }
        s *= z / (1.0 - zn * zn);
        c[0] = s + c0;
    }
    for (int64_t k = 1; k < len; ++k) c[k * stride] += z * c[(k - 1) * stride];
    c[(len - 1) * stride] *= z / (z - 1.0);
    for (int64_t k = len - 2; k >= 0; --k) c[k * stride] = z * (c[(k + 1) * stride] - c[k * stride]);
}

__device__ __forceinline__ void cubic_weights(double t, double* w)
{
    const double t2 = t * t, t3 = t2 * t, u = 1.0 - t;
    w[0] = u * u * u / 6.0;
    w[1] = (3.0 * t3 - 6.0 * t2 + 4.0) / 6.0;
    w[2] = (-3.0 * t3 + 3.0 * t2 + 3.0 * t + 1.0) / 6.0;
    w[3] = t3 / 6.0;
}

// Output voxel o along an axis of in/out lengths: source (o + 1/2) in/out - 1/2
// in the unpadded image, + kSplinePad in the coefficients; taps clamped.
__global__ void k_zoom(const double* __restrict__ coef, Vol vp, Vol vc, Vol vo, float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= vo.size()) return;
    int64_t c[3];
    coords(vo, i, c);
    double w[3][4];
    int64_t idx[3][4];
    for (int k = 0; k < 3; ++k) {
        const double zoom = (double)vc.n[k] / (double)vo.n[k];
        double cc = ((double)c[k] + 0.5) * zoom - 0.5;
        cc += (double)kSplinePad;
        const double lim = (double)(vp.n[k] - 1);
        cc = cc < 0.0 ? 0.0 : (cc > lim ? lim : cc);
        const double f = floor(cc);
        cubic_weights(cc - f, w[k]);
        const int64_t start = (int64_t)f - 1;
        for (int t = 0; t < 4; ++t) {
            const int64_t q = start + t;
            idx[k][t] = q < 0 ? 0 : (q > vp.n[k] - 1 ? vp.n[k] - 1 : q);
        }
    }
    double acc = 0.0;
    for (int a = 0; a < 4; ++a) {
        for (int b = 0; b < 4; ++b) {
            const double wzy = w[2][a] * w[1][b];
            const double* row = coef + (idx[2][a] * vp.n[1] + idx[1][b]) * vp.n[0];
            for (int t = 0; t < 4; ++t) acc += row[idx[0][t]] * (wzy * w[0][t]);
        }
    }
    out[i] = (float)acc;
}

// float bits mapped to unsigned integers in the order of the floats
__device__ __forceinline__ unsigned int ordered(float f)
{
    const unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unordered(unsigned int u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void k_minmax(const float* __restrict__ in, int64_t n, unsigned int* __restrict__ mm)
{
    unsigned int lo = 0xffffffffu, hi = 0u;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float f = in[i];
        if (f != f) continue;   // NaN: ignored (the range of the other values)
        const unsigned int o = ordered(f);
        lo = min(lo, o);
        hi = max(hi, o);
    }
    for (int s = kWave / 2; s > 0; s >>= 1) {
        lo = min(lo, (unsigned int)__shfl_xor((int)lo, s));
        hi = max(hi, (unsigned int)__shfl_xor((int)hi, s));
    }
    if ((threadIdx.x % kWave) == 0) {
        atomicMin(mm + 0, lo);
        atomicMax(mm + 1, hi);
    }
}

__global__ void k_minmax_init(unsigned int* mm)
{
    if (threadIdx.x == 0) {
        mm[0] = 0xffffffffu;
        mm[1] = 0u;
    }
}

__global__ void k_clip(float* __restrict__ out, int64_t n, const unsigned int* __restrict__ mm, int clip,
                       float zero_below)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v = out[i];
    if (clip && mm[0] <= mm[1]) {
        const float lo = unordered(mm[0]), hi = unordered(mm[1]);
        v = v < lo ? lo : (v > hi ? hi : v);
    }
    if (v < zero_below) v = 0.0f;
    out[i] = v;
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

int64_t resize_scratch_bytes(const int64_t crop[3])
{
    const int64_t nc = crop[0] * crop[1] * crop[2];
    const int64_t np_ = (crop[0] + 2 * kSplinePad) * (crop[1] + 2 * kSplinePad) * (crop[2] + 2 * kSplinePad);
    return 2 * nc * (int64_t)sizeof(float) + np_ * (int64_t)sizeof(double) + 256;
}

hipError_t launch_crop_and_downscale(const float* in, const int64_t in_shape[3], const int64_t crop_start[3],
                                     const int64_t crop[3], const int64_t out_shape[3], const double* const weights[3],
                                     const int radius[3], int clip, float zero_below, void* scratch, float* out,
                                     hipStream_t st)
{
    Vol vin{{in_shape[0], in_shape[1], in_shape[2]}};
    Vol vc{{crop[0], crop[1], crop[2]}};
    Vol vp{{crop[0] + 2 * kSplinePad, crop[1] + 2 * kSplinePad, crop[2] + 2 * kSplinePad}};
    Vol vo{{out_shape[0], out_shape[1], out_shape[2]}};
    unsigned char* p = static_cast<unsigned char*>(scratch);
    float* a = reinterpret_cast<float*>(p);
    float* b = a + vc.size();
    double* coef = reinterpret_cast<double*>(b + vc.size());
    unsigned int* mm = reinterpret_cast<unsigned int*>(coef + vp.size());
    const int64_t nc = vc.size();
    hipLaunchKernelGGL(k_crop, dim3(blocks_for(nc, 256)), dim3(256), 0, st, in, vin, vc, crop_start[0], crop_start[1],
                       crop_start[2], a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // the clip range is that of the cropped image before smoothing
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(64), 0, st, mm);
    hipLaunchKernelGGL(k_minmax, dim3((unsigned)std::min<int64_t>(1024, blocks_for(nc, 256))), dim3(256), 0, st, a,
                       nc, mm);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    float* cur = a;
    float* nxt = b;
    for (int k = 0; k < 3; ++k) {
        if (radius[k] < 0) continue;
        hipLaunchKernelGGL(k_gauss_axis, dim3(blocks_for(nc, 256)), dim3(256), 0, st, cur, vc, k, weights[k],
                           radius[k], nxt);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        std::swap(cur, nxt);
    }
    hipLaunchKernelGGL(k_spline_pad, dim3(blocks_for(vp.size(), 256)), dim3(256), 0, st, cur, vc, vp, coef);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int k = 0; k < 3; ++k) {
        const int64_t lines = vp.size() / vp.n[k];
        hipLaunchKernelGGL(k_spline_axis, dim3(blocks_for(lines, 64)), dim3(64), 0, st, coef, vp, k);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_zoom, dim3(blocks_for(vo.size(), 256)), dim3(256), 0, st, coef, vp, vc, vo, out);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_clip, dim3(blocks_for(vo.size(), 256)), dim3(256), 0, st, out, vo.size(), mm, clip,
                       zero_below);
    return hipGetLastError();
}

}  // namespace sss

extern "C" int sss_crop_and_downscale(sss_ctx* ctx, const float* image, const sss_resize_desc* d, float* out)
{
    using namespace sss;
    if (!ctx || !image || !d || !out) return fail(SSS_ERR_INVALID, "sss_crop_and_downscale: NULL argument");
    for (int k = 0; k < 3; ++k) {
        if (d->in_shape[k] < 1 || d->crop_shape[k] < 1 || d->out_shape[k] < 1)
            return fail(SSS_ERR_INVALID, "sss_crop_and_downscale: shapes must be positive");
        if (d->radius[k] >= 0 && !d->weights[k])
            return fail(SSS_ERR_INVALID, "sss_crop_and_downscale: smoothing axis without taps");
        if (d->radius[k] > 4096) return fail(SSS_ERR_UNSUPPORTED, "sss_crop_and_downscale: Gaussian radius above 4096");
    }
    SSS_HIP(hipSetDevice(ctx_device(ctx)));
    hipStream_t st = ctx_stream(ctx);
    const int64_t n_in = d->in_shape[0] * d->in_shape[1] * d->in_shape[2];
    const int64_t n_out = d->out_shape[0] * d->out_shape[1] * d->out_shape[2];
    int64_t n_w = 0;
    for (int k = 0; k < 3; ++k) n_w += d->radius[k] >= 0 ? 2 * d->radius[k] + 1 : 0;
    void *d_in = nullptr, *d_out = nullptr, *d_w = nullptr, *d_scratch = nullptr;
    int rc;
    if ((rc = ctx_pool_get(ctx, (size_t)n_in * 4, &d_in)) || (rc = ctx_pool_get(ctx, (size_t)n_out * 4, &d_out)) ||
        (rc = ctx_pool_get(ctx, (size_t)std::max<int64_t>(n_w, 1) * 8, &d_w)) ||
        (rc = ctx_pool_get(ctx, (size_t)resize_scratch_bytes(d->crop_shape), &d_scratch))) {
        ctx_pool_put(ctx, d_in); ctx_pool_put(ctx, d_out); ctx_pool_put(ctx, d_w); ctx_pool_put(ctx, d_scratch);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_in, image, (size_t)n_in * 4, hipMemcpyHostToDevice, st);
    const double* w_dev[3] = {nullptr, nullptr, nullptr};
    int radius[3];
    int64_t off = 0;
    for (int k = 0; k < 3; ++k) {
        radius[k] = d->radius[k];
        if (radius[k] < 0) continue;
        w_dev[k] = static_cast<const double*>(d_w) + off;
        if (e == hipSuccess)
            e = hipMemcpyAsync(static_cast<double*>(d_w) + off, d->weights[k], (size_t)(2 * radius[k] + 1) * 8,
                               hipMemcpyHostToDevice, st);
        off += 2 * radius[k] + 1;
    }
    if (e == hipSuccess)
        e = launch_crop_and_downscale(static_cast<const float*>(d_in), d->in_shape, d->crop_start, d->crop_shape,
                                      d->out_shape, w_dev, radius, d->clip, d->zero_below, d_scratch,
                                      static_cast<float*>(d_out), st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n_out * 4, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    ctx_pool_put(ctx, d_in); ctx_pool_put(ctx, d_out); ctx_pool_put(ctx, d_w); ctx_pool_put(ctx, d_scratch);
    if (e != hipSuccess) return hip_fail(e, "sss_crop_and_downscale");
    return SSS_OK;
}
