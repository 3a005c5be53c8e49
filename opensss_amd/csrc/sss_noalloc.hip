// The tail fit on caller-owned device buffers with no allocation.
//
// sss_tail_fit allocates device copies of its inputs and its scratch (the
// fixed-point images and the reduction partials) on every call.  Callers that
// cannot allocate (code under HIP graph capture, tight serving loops) call
// sss_tail_fit_noalloc instead: inputs and outputs are device buffers they
// own, and the scratch is the file-scope array below, one copy per GPU, held
// until the process ends.  Its device address is looked up once per device
// (hipGetSymbolAddress, at context creation with that device current) and
// only that address is ever handed to kernels; the array's name is a host
// stand-in in host code.
//
// Every call on one GPU shares that GPU's array: calls that overlap on one
// GPU (several contexts, or several host threads) race.  Each call clears the
// images and maxima it accumulates into (a kernel, no memset).
#include <mutex>

#include "sss_internal.h"

namespace sss {
namespace {

constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[fit_scratch_bytes(SSS_TAIL_FIT_NOALLOC_MAX_VOXELS)];

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace

int noalloc_scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(SSS_ERR_UNSUPPORTED, "device id above 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (g_scratch_ptr[device]) return SSS_OK;
    // the caller has made `device` current
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return fail(SSS_ERR_HIP, "hipGetSymbolAddress(scratch) returned NULL");
    g_scratch_ptr[device] = p;
    return SSS_OK;
}

void* noalloc_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    return g_scratch_ptr[device];
}

}  // namespace sss

using namespace sss;

extern "C" int sss_tail_fit_noalloc(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], int n_rings_geom,
                                    int n_det, const double* geom, const int16_t* lor_coords, int n_entries,
                                    const int32_t* plane_rings, int n_value_planes, int accel, int length_weighted,
                                    const float* scatter, const float* prompts, const float* randoms,
                                    const float* mask, double* sums_out, double* images_out)
{
    if (!ctx || !grid || !bounds || !geom || !lor_coords || (n_entries > 0 && !plane_rings) || !scatter ||
        !prompts || !randoms || !mask || !sums_out || n_det < 2 || n_rings_geom < 1 || n_entries < 0 ||
        n_value_planes < 1 || accel < 1)
        return fail(SSS_ERR_INVALID, "sss_tail_fit_noalloc: bad argument");
    // the grid is read from host memory (plain values), the rest stays on the device
    const int32_t g3[3] = {grid[0], grid[1], grid[2]};
    double b6[6];
    for (int i = 0; i < 6; ++i) b6[i] = bounds[i];
    if (g3[0] <= 0 || g3[1] <= 0 || g3[2] <= 0 || g3[0] != g3[1])
        return fail(SSS_ERR_INVALID, "sss_tail_fit_noalloc: grid must be positive with Nx == Ny");
    if (!(b6[3] > b6[0] && b6[4] > b6[1] && b6[5] > b6[2]))
        return fail(SSS_ERR_INVALID, "sss_tail_fit_noalloc: bounds must satisfy max > min on every axis");
    const int64_t nvox = (int64_t)g3[0] * g3[1] * g3[2];
    if (nvox > SSS_TAIL_FIT_NOALLOC_MAX_VOXELS)
        return fail(SSS_ERR_UNSUPPORTED, "sss_tail_fit_noalloc: fit grid of " + std::to_string(nvox) +
                                             " voxels exceeds the scratch of " +
                                             std::to_string((long long)SSS_TAIL_FIT_NOALLOC_MAX_VOXELS) +
                                             " voxels; use sss_tail_fit");
    const int device = ctx_device(ctx);
    SSS_HIP(hipSetDevice(device));
    void* base = noalloc_scratch(device);
    if (!base) return fail(SSS_ERR_HIP, "sss_tail_fit_noalloc: scratch not set up for this device");
    Grid g;
    g.nx = g3[0];
    g.ny = g3[1];
    g.nz = g3[2];
    for (int i = 0; i < 6; ++i) g.b[i] = b6[i];
    SSS_HIP(launch_fit(g, geom, n_rings_geom, n_det, lor_coords, n_entries, plane_rings, n_value_planes, accel,
                       length_weighted, scatter, prompts, randoms, mask, fit_scratch_carve(base, nvox), sums_out,
                       images_out, static_cast<hipStream_t>(sss_ctx_stream(ctx))));
    return SSS_OK;
}
