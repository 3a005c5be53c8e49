// CASToR list-mode events on either side of the scatter path:
//
//   k_lm_histogram  Timber/Datafile.py:29-115 (ExportPrompts): every event's
//                   crystal pair -> (radial, angular, plane, TOF bin) of the
//                   prompts sinogram, plus the randoms and normalisation
//                   sinograms, in which the last event of a cell wins.
//   k_lm_keep / k_lm_inject  Timber/Datafile.py:117-240 (InjectScatters):
//                   the events kept, in file order, each with its scatter rate
//                   inserted at the scatter-correction word.
//
// Events are records of 4-byte words (Coincidence.py:45-95: time, the float
// corrections, the optional line count, the crystal ID pairs); the crystal
// pair, TOF and ring order follow ParseCASToRIDs (Coincidence.py:97-159).
#include "sss_internal.h"

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

namespace sss {
namespace {

#pragma clang fp contract(off)   // the reference's float64 / float32 operation order

struct EventLor {
    int d1, r1, d2, r2;
    double tof;
};

// ParseCASToRIDs: the detectors come from the IDs without the shift, the ring
// order (and the TOF sign) from the shifted transaxial indices.
__device__ __forceinline__ EventLor parse_ids(const uint32_t* ev, const ListmodeLayout& lay, int64_t ncpr,
                                              int64_t shift)
{
    const uint32_t c0 = ev[lay.id_word], c1 = ev[lay.id_word + 1];
    const double tof_info = (double)__uint_as_float(ev[lay.tof_word]);
    const bool lt = c0 < c1;
    const int64_t id1 = lt ? c0 : c1, id2 = lt ? c1 : c0;
    const double tof_tmp = lt ? tof_info : -tof_info;
    const int64_t t1 = id1 % ncpr, t2 = id2 % ncpr;
    EventLor e;
    e.d1 = (int)(t1 < t2 ? t1 : t2);
    e.d2 = (int)(t1 < t2 ? t2 : t1);
    int64_t s1 = id1 % ncpr - shift, s2 = id2 % ncpr - shift;
    if (s1 < 0) s1 += ncpr;
    if (s2 < 0) s2 += ncpr;
    if (s1 < s2) {
        e.r1 = (int)(id1 / ncpr);
        e.r2 = (int)(id2 / ncpr);
        e.tof = tof_tmp;
    } else {
        e.r1 = (int)(id2 / ncpr);
        e.r2 = (int)(id1 / ncpr);
        e.tof = -tof_tmp;
    }
    return e;
}

// Sinogram cell of an event: (radial, angular, plane) from the coordinates of
// the mashed detector pair and the ring-pair table; -1 entries skip the event.
// Returns 1 (cell), 0 (skipped) or -1 (an index the reference's arrays do
// not have: it raises IndexError there).
__device__ __forceinline__ int event_cell(const EventLor& e, const ListmodeTables& t, int64_t* cell)
{
    const int m1 = e.d1 / t.mash, m2 = e.d2 / t.mash;
    if (m1 >= t.n_det_lor || m2 >= t.n_det_lor || e.r1 >= t.n_rings_index || e.r2 >= t.n_rings_index) return -1;
    const int ang = t.lor[((int64_t)m1 * t.n_det_lor + m2) * 2 + 0];
    const int rad = t.lor[((int64_t)m1 * t.n_det_lor + m2) * 2 + 1];
    const int plane = t.ring_index[(int64_t)e.r1 * t.n_rings_index + e.r2];
    if (plane == -1 || rad == -1 || ang == -1) return 0;
    if (plane < 0 || rad < 0 || ang < 0 || rad >= t.n_rad || ang >= t.n_ang || plane >= t.n_planes) return -1;
    *cell = ((int64_t)rad * t.n_ang + ang) * t.n_planes + plane;
    return 1;
}

// int((TOF + R) / (2 R / B)) when |TOF| <= R and the quotient lies in [0, B), else -1.
__device__ __forceinline__ int tof_bin(double tof, double range, int n_bins)
{
    if (fabs(tof) > range) return -1;
    const double tb = (tof + range) / (2.0 * range / (double)n_bins);
    if (tb < 0.0 || tb >= (double)n_bins) return -1;
    return (int)tb;
}

__global__ void k_lm_histogram(const uint32_t* __restrict__ events, int64_t n_events, int64_t first_index,
                               ListmodeLayout lay, ListmodeTables t, int64_t ncpr, int64_t shift, int n_bins,
                               double tof_range, double duration, uint32_t* __restrict__ counts,
                               unsigned long long* __restrict__ randoms_last,
                               unsigned long long* __restrict__ norm_last, int* __restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_events) return;
    const uint32_t* ev = events + i * lay.words;
    const EventLor e = parse_ids(ev, lay, ncpr, shift);
    int64_t cell;
    const int got = event_cell(e, t, &cell);
    if (got < 0) {
        atomicOr(bad, 1);
        return;
    }
    if (got == 0) return;
    // last event of the cell wins: (file index + 1) in the high word, the
    // float32 value's bits in the low word, kept by an atomic maximum
    const unsigned long long tag = (unsigned long long)(first_index + i + 1) << 32;
    if (randoms_last) {
        const float r = (float)((double)ev[0] * duration / (double)n_bins);
        atomicMax(randoms_last + cell, tag | __float_as_uint(r));
    }
    if (norm_last) atomicMax(norm_last + cell, tag | ev[1]);
    const int b = tof_bin(e.tof, tof_range, n_bins);
    if (b < 0) return;
    atomicAdd(counts + cell * n_bins + b, 1u);
}

// uint16 prompts (the reference's uint16 += 1 wraps), randoms (0 where no
// event) and normalisation (1 where no event).
__global__ void k_lm_finish(const uint32_t* __restrict__ counts, int64_t n_counts, uint16_t* __restrict__ prompts,
                            const unsigned long long* __restrict__ randoms_last,
                            const unsigned long long* __restrict__ norm_last, int64_t n_cells,
                            float* __restrict__ randoms, float* __restrict__ norm)
{
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (int64_t i = i0; i < n_counts; i += (int64_t)gridDim.x * blockDim.x) prompts[i] = (uint16_t)(counts[i] & 0xffffu);
    for (int64_t c = i0; c < n_cells; c += (int64_t)gridDim.x * blockDim.x) {
        if (randoms) {
            const unsigned long long v = randoms_last[c];
            randoms[c] = v ? __uint_as_float((uint32_t)v) : 0.0f;
        }
        if (norm) {
            const unsigned long long v = norm_last[c];
            norm[c] = v ? __uint_as_float((uint32_t)v) : 1.0f;
        }
    }
}

// InjectScatters, pass 1: keep[i] = 1 for an event that is written out, and
// its scatter rate factor[b] * scatter[b][cell] in float32.
__global__ void k_lm_keep(const uint32_t* __restrict__ events, int64_t n_events, ListmodeLayout lay,
                          ListmodeTables t, int64_t ncpr, int64_t shift, int n_bins, double tof_range,
                          const float* __restrict__ factor, const float* __restrict__ scatter,
                          int64_t n_cells, int32_t* __restrict__ keep, float* __restrict__ rate,
                          int* __restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_events) return;
    const uint32_t* ev = events + i * lay.words;
    const EventLor e = parse_ids(ev, lay, ncpr, shift);
    keep[i] = 0;
    if (fabs(e.tof) > tof_range) return;   // before the cell lookup, as the reference
    int64_t cell;
    const int got = event_cell(e, t, &cell);
    if (got < 0) {
        atomicOr(bad, 1);
        return;
    }
    if (got == 0) return;
    const int b = tof_bin(e.tof, tof_range, n_bins);
    if (b < 0) return;
    keep[i] = 1;
    rate[i] = factor[b] * scatter[(int64_t)b * n_cells + cell];
}

// Pass 2: kept events to their place (pos = exclusive scan of keep), the rate
// inserted at word lay.scatter_word.
__global__ void k_lm_inject(const uint32_t* __restrict__ events, int64_t n_events, ListmodeLayout lay,
                            const int32_t* __restrict__ keep, const int64_t* __restrict__ pos,
                            const float* __restrict__ rate, uint32_t* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_events || !keep[i]) return;
    const uint32_t* ev = events + i * lay.words;
    uint32_t* o = out + pos[i] * (lay.words + 1);
    for (int w = 0; w < lay.scatter_word; ++w) o[w] = ev[w];
    o[lay.scatter_word] = __float_as_uint(rate[i]);
    for (int w = lay.scatter_word; w < lay.words; ++w) o[w + 1] = ev[w];
}

// Exclusive scan of keep into pos, one block per kScanChunk events, then the
// block totals scanned by one block and added back (three launches).
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr int64_t kScanChunk = (int64_t)kScanThreads * kScanPer;

__global__ __launch_bounds__(kScanThreads) void k_scan_blocks(const int32_t* __restrict__ keep, int64_t n,
                                                              int64_t* __restrict__ pos,
                                                              int64_t* __restrict__ block_sum)
{
    __shared__ int64_t s[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanChunk + (int64_t)threadIdx.x * kScanPer;
    int64_t local[kScanPer];
    int64_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        local[k] = acc;
        acc += (base + k < n) ? keep[base + k] : 0;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
        const int64_t v = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    const int64_t before = s[threadIdx.x] - acc;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k)
        if (base + k < n) pos[base + k] = before + local[k];
    if (threadIdx.x == kScanThreads - 1) block_sum[blockIdx.x] = s[threadIdx.x];
}

// One block: exclusive scan of the block totals in place, the grand total at [n_blocks].
__global__ __launch_bounds__(kScanThreads) void k_scan_totals(int64_t* __restrict__ block_sum, int64_t n_blocks)
{
    __shared__ int64_t s[kScanThreads];
    __shared__ int64_t carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (int64_t b0 = 0; b0 < n_blocks; b0 += kScanThreads) {
        const int64_t i = b0 + threadIdx.x;
        const int64_t v = i < n_blocks ? block_sum[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kScanThreads; o <<= 1) {
            const int64_t u = threadIdx.x >= (unsigned)o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += u;
            __syncthreads();
        }
        if (i < n_blocks) block_sum[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == kScanThreads - 1) carry += s[threadIdx.x];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_sum[n_blocks] = carry;
}

__global__ void k_scan_add(int64_t* __restrict__ pos, int64_t n, const int64_t* __restrict__ block_sum)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) pos[i] += block_sum[i / kScanChunk];
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

hipError_t launch_lm_histogram(const uint32_t* events, int64_t n_events, int64_t first_index,
                               const ListmodeLayout& lay, const ListmodeTables& t, int64_t ncpr, int64_t shift,
                               int n_bins, double tof_range, double duration, uint32_t* counts,
                               unsigned long long* randoms_last, unsigned long long* norm_last, int* bad,
                               hipStream_t st)
{
    if (n_events <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lm_histogram, dim3(blocks_for(n_events, 256)), dim3(256), 0, st, events, n_events,
                       first_index, lay, t, ncpr, shift, n_bins, tof_range, duration, counts, randoms_last, norm_last,
                       bad);
    return hipGetLastError();
}

hipError_t launch_lm_finish(const uint32_t* counts, int64_t n_counts, uint16_t* prompts,
                            const unsigned long long* randoms_last, const unsigned long long* norm_last,
                            int64_t n_cells, float* randoms, float* norm, hipStream_t st)
{
    const int64_t n = n_counts > n_cells ? n_counts : n_cells;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lm_finish, dim3((unsigned)std::min<int64_t>(4096, blocks_for(n, 256))), dim3(256), 0, st,
                       counts, n_counts, prompts, randoms_last, norm_last, n_cells, randoms, norm);
    return hipGetLastError();
}

int64_t lm_scan_blocks(int64_t n_events) { return (n_events + kScanChunk - 1) / kScanChunk; }

// keep, rate, pos (exclusive scan of keep) and block_sum[lm_scan_blocks(n)]
// = the number of events kept.
hipError_t launch_lm_select(const uint32_t* events, int64_t n_events, const ListmodeLayout& lay,
                            const ListmodeTables& t, int64_t ncpr, int64_t shift, int n_bins, double tof_range,
                            const float* factor, const float* scatter, int64_t n_cells, int32_t* keep, float* rate,
                            int64_t* pos, int64_t* block_sum, int* bad, hipStream_t st)
{
    const int64_t nb = lm_scan_blocks(n_events);
    hipError_t e = hipSuccess;
    if (n_events > 0) {
        hipLaunchKernelGGL(k_lm_keep, dim3(blocks_for(n_events, 256)), dim3(256), 0, st, events, n_events, lay, t,
                           ncpr, shift, n_bins, tof_range, factor, scatter, n_cells, keep, rate, bad);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(kScanThreads), 0, st, keep, n_events, pos,
                           block_sum);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanThreads), 0, st, block_sum, nb);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (n_events > 0) {
        hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n_events, 256)), dim3(256), 0, st, pos, n_events, block_sum);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_lm_write(const uint32_t* events, int64_t n_events, const ListmodeLayout& lay, const int32_t* keep,
                           const int64_t* pos, const float* rate, uint32_t* out, hipStream_t st)
{
    if (n_events <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_lm_inject, dim3(blocks_for(n_events, 256)), dim3(256), 0, st, events, n_events, lay, keep, pos,
                       rate, out);
    return hipGetLastError();
}

}  // namespace sss

// ---- C ABI ----
namespace {

using namespace sss;

constexpr int64_t kLmChunk = int64_t(1) << 24;   // events per host upload

int lm_check(const char* fn, sss_ctx* ctx, const void* events, int64_t n_events, const sss_listmode_layout* lay,
             int64_t ncpr, int mash, int n_det_lor, const int16_t* lor, int nri, const int16_t* ri, int n_rad,
             int n_ang, int n_planes, int n_bins, double tof_range, bool inject)
{
    const std::string f(fn);
    if (!ctx || !lay || !lor || !ri || n_events < 0 || (n_events > 0 && !events))
        return fail(SSS_ERR_INVALID, f + ": NULL argument");
    if (ncpr < 1 || mash < 1 || n_det_lor < 1 || nri < 1 || n_rad < 1 || n_ang < 1 || n_planes < 1 || n_bins < 1)
        return fail(SSS_ERR_INVALID, f + ": sizes must be positive");
    if (!(tof_range > 0.0) || !std::isfinite(tof_range))
        return fail(SSS_ERR_INVALID, f + ": the TOF range must be positive and finite");
    if (lay->words < 3 || lay->tof_word < 0 || lay->tof_word >= lay->words || lay->id_word < 0 ||
        lay->id_word + 1 >= lay->words || (inject && (lay->scatter_word < 0 || lay->scatter_word > lay->words)))
        return fail(SSS_ERR_INVALID, f + ": event layout outside its record");
    if (n_events >= (int64_t)0xffffffffLL) return fail(SSS_ERR_UNSUPPORTED, f + ": 2^32 - 1 events or more");
    return SSS_OK;
}

struct Blocks {
    sss_ctx* ctx;
    std::vector<void*> held;
    explicit Blocks(sss_ctx* c) : ctx(c) {}
    ~Blocks()
    {
        for (void* p : held) ctx_pool_put(ctx, p);
    }
    template <typename T>
    int get(T** p, size_t count)
    {
        void* v = nullptr;
        const int rc = ctx_pool_get(ctx, count * sizeof(T), &v);
        if (rc == SSS_OK) held.push_back(v);
        *p = static_cast<T*>(v);
        return rc;
    }
};

}  // namespace

extern "C" int sss_listmode_histogram(sss_ctx* ctx, const void* events, int64_t n_events,
                                      const sss_listmode_layout* layout, int64_t crystals_per_ring, int64_t shift,
                                      int mash, int n_det_lor, const int16_t* lor_coords, int n_rings_index,
                                      const int16_t* ring_index, int n_rad, int n_ang, int n_planes, int n_bins,
                                      double tof_range, double duration, int flags, uint16_t* prompts_out,
                                      float* randoms_out, float* norm_out)
{
    int rc = lm_check("sss_listmode_histogram", ctx, events, n_events, layout, crystals_per_ring, mash, n_det_lor,
                      lor_coords, n_rings_index, ring_index, n_rad, n_ang, n_planes, n_bins, tof_range, false);
    if (rc) return rc;
    const bool dev = (flags & SSS_LM_DEVICE) != 0, want_r = (flags & SSS_LM_RANDOMS) != 0,
               want_n = (flags & SSS_LM_NORM) != 0;
    if ((flags & ~(SSS_LM_DEVICE | SSS_LM_RANDOMS | SSS_LM_NORM)) || !prompts_out || (want_r && !randoms_out) ||
        (want_n && !norm_out))
        return fail(SSS_ERR_INVALID, "sss_listmode_histogram: bad flags or output");
    SSS_HIP(hipSetDevice(ctx_device(ctx)));
    hipStream_t st = ctx_stream(ctx);
    const int64_t n_cells = (int64_t)n_rad * n_ang * n_planes, n_counts = n_cells * n_bins;
    const size_t nlor = (size_t)n_det_lor * n_det_lor * 2, nri = (size_t)n_rings_index * n_rings_index;
    const int words = layout->words;
    Blocks bl(ctx);
    uint32_t *d_counts = nullptr, *d_ev = nullptr;
    unsigned long long *d_rl = nullptr, *d_nl = nullptr;
    int* d_bad = nullptr;
    int16_t *d_lor = nullptr, *d_ri = nullptr;
    uint16_t* d_prompts = nullptr;
    float *d_r = nullptr, *d_n = nullptr;
    const int64_t chunk = dev ? n_events : std::min<int64_t>(n_events, kLmChunk);
    if ((rc = bl.get(&d_counts, (size_t)n_counts)) || (want_r && (rc = bl.get(&d_rl, (size_t)n_cells))) ||
        (want_n && (rc = bl.get(&d_nl, (size_t)n_cells))) || (rc = bl.get(&d_bad, 1)))
        return rc;
    if (!dev && ((rc = bl.get(&d_lor, nlor)) || (rc = bl.get(&d_ri, nri)) ||
                 (rc = bl.get(&d_ev, (size_t)std::max<int64_t>(chunk, 1) * words)) ||
                 (rc = bl.get(&d_prompts, (size_t)n_counts)) || (want_r && (rc = bl.get(&d_r, (size_t)n_cells))) ||
                 (want_n && (rc = bl.get(&d_n, (size_t)n_cells)))))
        return rc;
    hipError_t e = hipMemsetAsync(d_counts, 0, (size_t)n_counts * sizeof(uint32_t), st);
    if (e == hipSuccess && d_rl) e = hipMemsetAsync(d_rl, 0, (size_t)n_cells * 8, st);
    if (e == hipSuccess && d_nl) e = hipMemsetAsync(d_nl, 0, (size_t)n_cells * 8, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, sizeof(int), st);
    if (!dev) {
        if (e == hipSuccess) e = hipMemcpyAsync(d_lor, lor_coords, nlor * 2, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_ri, ring_index, nri * 2, hipMemcpyHostToDevice, st);
    }
    ListmodeLayout lay{layout->words, layout->tof_word, layout->id_word, layout->scatter_word};
    ListmodeTables t{dev ? lor_coords : d_lor, n_det_lor, mash, dev ? ring_index : d_ri, n_rings_index,
                     n_rad, n_ang, n_planes};
    for (int64_t first = 0; e == hipSuccess && first < n_events; first += chunk) {
        const int64_t n = std::min<int64_t>(chunk, n_events - first);
        const uint32_t* src = static_cast<const uint32_t*>(events) + first * words;
        if (!dev) {
            e = hipMemcpyAsync(d_ev, src, (size_t)n * words * 4, hipMemcpyHostToDevice, st);
            src = d_ev;
        }
        if (e == hipSuccess)
            e = launch_lm_histogram(src, n, first, lay, t, crystals_per_ring, shift, n_bins, tof_range, duration,
                                    d_counts, d_rl, d_nl, d_bad, st);
    }
    if (e == hipSuccess)
        e = launch_lm_finish(d_counts, n_counts, dev ? prompts_out : d_prompts, d_rl, d_nl, n_cells,
                             want_r ? (dev ? randoms_out : d_r) : nullptr, want_n ? (dev ? norm_out : d_n) : nullptr, st);
    int bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (!dev) {
        if (e == hipSuccess) e = hipMemcpyAsync(prompts_out, d_prompts, (size_t)n_counts * 2, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && want_r) e = hipMemcpyAsync(randoms_out, d_r, (size_t)n_cells * 4, hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && want_n) e = hipMemcpyAsync(norm_out, d_n, (size_t)n_cells * 4, hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "sss_listmode_histogram");
    if (bad)
        return fail(SSS_ERR_INVALID, "sss_listmode_histogram: an event's detector, ring or sinogram coordinate is "
                                     "outside the tables");
    return SSS_OK;
}

extern "C" int sss_listmode_inject(sss_ctx* ctx, const void* events, int64_t n_events,
                                   const sss_listmode_layout* layout, int64_t crystals_per_ring, int64_t shift,
                                   int mash, int n_det_lor, const int16_t* lor_coords, int n_rings_index,
                                   const int16_t* ring_index, int n_rad, int n_ang, int n_planes, int n_bins,
                                   double tof_range, const float* factor, const float* scatter, int flags,
                                   void* events_out, int64_t* n_out)
{
    int rc = lm_check("sss_listmode_inject", ctx, events, n_events, layout, crystals_per_ring, mash, n_det_lor,
                      lor_coords, n_rings_index, ring_index, n_rad, n_ang, n_planes, n_bins, tof_range, true);
    if (rc) return rc;
    if ((flags & ~SSS_LM_DEVICE) || !factor || !scatter || !n_out || (n_events > 0 && !events_out))
        return fail(SSS_ERR_INVALID, "sss_listmode_inject: bad flags or output");
    const bool dev = (flags & SSS_LM_DEVICE) != 0;
    *n_out = 0;
    SSS_HIP(hipSetDevice(ctx_device(ctx)));
    hipStream_t st = ctx_stream(ctx);
    const int64_t n_cells = (int64_t)n_rad * n_ang * n_planes;
    const size_t nlor = (size_t)n_det_lor * n_det_lor * 2, nri = (size_t)n_rings_index * n_rings_index;
    const int words = layout->words;
    const int64_t chunk = std::max<int64_t>(1, dev ? n_events : std::min<int64_t>(n_events, kLmChunk));
    Blocks bl(ctx);
    uint32_t *d_ev = nullptr, *d_out = nullptr;
    int16_t *d_lor = nullptr, *d_ri = nullptr;
    float *d_f = nullptr, *d_s = nullptr, *d_rate = nullptr;
    int32_t* d_keep = nullptr;
    int64_t *d_pos = nullptr, *d_bs = nullptr;
    int* d_bad = nullptr;
    const int64_t nb = lm_scan_blocks(chunk);
    if ((rc = bl.get(&d_rate, (size_t)chunk)) || (rc = bl.get(&d_keep, (size_t)chunk)) ||
        (rc = bl.get(&d_pos, (size_t)chunk)) || (rc = bl.get(&d_bs, (size_t)nb + 1)) || (rc = bl.get(&d_bad, 1)))
        return rc;
    if (!dev && ((rc = bl.get(&d_lor, nlor)) || (rc = bl.get(&d_ri, nri)) || (rc = bl.get(&d_f, (size_t)n_bins)) ||
                 (rc = bl.get(&d_s, (size_t)n_cells * n_bins)) || (rc = bl.get(&d_ev, (size_t)chunk * words)) ||
                 (rc = bl.get(&d_out, (size_t)chunk * (words + 1)))))
        return rc;
    hipError_t e = hipMemsetAsync(d_bad, 0, sizeof(int), st);
    if (!dev) {
        if (e == hipSuccess) e = hipMemcpyAsync(d_lor, lor_coords, nlor * 2, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_ri, ring_index, nri * 2, hipMemcpyHostToDevice, st);
        if (e == hipSuccess) e = hipMemcpyAsync(d_f, factor, (size_t)n_bins * 4, hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_s, scatter, (size_t)n_cells * n_bins * 4, hipMemcpyHostToDevice, st);
    }
    ListmodeLayout lay{layout->words, layout->tof_word, layout->id_word, layout->scatter_word};
    ListmodeTables t{dev ? lor_coords : d_lor, n_det_lor, mash, dev ? ring_index : d_ri, n_rings_index,
                     n_rad, n_ang, n_planes};
    int64_t total = 0;
    for (int64_t first = 0; e == hipSuccess && first < n_events; first += chunk) {
        const int64_t n = std::min<int64_t>(chunk, n_events - first);
        const uint32_t* src = static_cast<const uint32_t*>(events) + first * words;
        uint32_t* dst = dev ? static_cast<uint32_t*>(events_out) + total * (words + 1) : d_out;
        if (!dev) {
            e = hipMemcpyAsync(d_ev, src, (size_t)n * words * 4, hipMemcpyHostToDevice, st);
            src = d_ev;
        }
        if (e == hipSuccess)
            e = launch_lm_select(src, n, lay, t, crystals_per_ring, shift, n_bins, tof_range, dev ? factor : d_f,
                                 dev ? scatter : d_s, n_cells, d_keep, d_rate, d_pos, d_bs, d_bad, st);
        int64_t kept = 0;
        if (e == hipSuccess)
            e = hipMemcpyAsync(&kept, d_bs + lm_scan_blocks(n), sizeof(int64_t), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = launch_lm_write(src, n, lay, d_keep, d_pos, d_rate, dst, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e == hipSuccess && !dev && kept > 0)
            e = hipMemcpy(static_cast<uint32_t*>(events_out) + total * (words + 1), d_out,
                          (size_t)kept * (words + 1) * 4, hipMemcpyDeviceToHost);
        total += kept;
    }
    int bad = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "sss_listmode_inject");
    if (bad)
        return fail(SSS_ERR_INVALID, "sss_listmode_inject: an event's detector, ring or sinogram coordinate is "
                                     "outside the tables");
    *n_out = total;
    return SSS_OK;
}
