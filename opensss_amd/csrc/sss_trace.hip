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
#include "sss_walk.h"

#include <algorithm>

// distances, angles and the line-integral sums below keep float64 operation
// order too (no fused multiply-add)
#pragma clang fp contract(off)

namespace sss {
namespace {

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
    int n = 0;
    double lc[6];
    int det = 0;
    if (h < n_hp) {
        halfpath_line(sp_xyz, detpos, nru, ndu, h, lc, det);
        Walk w;
        walk_setup(g, lc, w);
        n = w.n;
    }
    // records of this wave's block of half paths: kWave x its longest walk
    int m = n;
    for (int o = kWave / 2; o > 0; o /= 2) m = max(m, __shfl_xor(m, o, kWave));
    if (h >= n_hp) return;
    if (h % kWave == 0) t.blk[h / kWave] = m * kWave;
    t.n[h] = n;
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
    const int64_t base = t.blk_off[h / kWave] + h % kWave;
    t.off[h] = base;
    if (walk_setup(g, lc, w)) {
        const double two_over_c = 2.0 / kLightCmPerPs;
        bool truncated = false;
        // the mu gather of a step is added one step later, so its latency
        // overlaps the next step (same terms, same order: the first added
        // term is 0 * 0)
        double mu_prev = 0, len_prev = 0;
        float lmin = __builtin_inff();
        const int got = walk_run(g, w, [&](int k, int vox, double L, double smp) {
            const int64_t i = base + (int64_t)k * kWave;
            // q * 2^32 is exact in float64; rint drops bits below 2^-32 ps
            s.nq[i] = -(int64_t)rint(smp * two_over_c * kFixOne);
            s.vox[i] = vox;
            const float lf = (float)L;
            if (s.len32) s.len32[i] = lf;
            else s.len[i] = L;
            if (lf > 0.0f) lmin = fminf(lmin, lf);
            mu_int += mu_prev * len_prev;
            mu_prev = mu[vox];
            len_prev = L;
        }, &truncated);
        mu_int += mu_prev * len_prev;
        if (truncated) status = 1;
        // positive floats order as their bits (the float of a float64 length
        // may round up: the voxel test keeps a 2^10 margin on it)
        if (t.len_min && lmin < __builtin_inff()) atomicMin(t.len_min, __float_as_uint(lmin));
        t.n[h] = got;   // <= the count the offsets were built from
    }
    t.lnatt[h] = mu_int;
    t.att[h] = exp(-mu_int);
    t.status[h] = status;
}

// Emission terms lam[v] L per sample, their float64 sum per half path
// (ActivityRay and ActPath, SingleScatterSimulationTOF.py:417-418) and the
// number of non-zero terms (the samples the fold visits).  A workgroup of
// kRefreshWaves waves takes one block of 64 half paths: wave w the samples
// k = w, w + kRefreshWaves, ... of every path of the block (each step reads
// 64 consecutive interleaved records), and the waves' partial sums are added
// in wave order.  The terms are non-negative, so the exact-zero tests on the
// sum do not depend on the order.
constexpr int kRefreshWaves = 4;

// Intersection length of sample entry i in the precision the subject keeps
// (the same for every thread: a uniform branch).
__device__ __forceinline__ double sample_len(const SampleTable& s, int64_t i)
{
    return s.len32 ? (double)s.len32[i] : s.len[i];
}

// kActSum: ActPath only (one-bin subjects); kActRefresh: also the FP32 (and
// FP64) terms per sample and the non-zero counts; kActCheck: ActPath and the
// non-zero counts, compared with the counts the fold table was built from
// (any difference sets *changed).
enum ActMode { kActSum, kActRefresh, kActCheck };

template <int MODE>
__device__ __forceinline__ void activity_block(HalfPathTable t, SampleTable s, int64_t n_hp,
                                               const double* __restrict__ lam, int32_t* changed = nullptr,
                                               int64_t blk = -1)
{
    if (blk < 0) blk = blockIdx.x;
    constexpr bool FOLD = MODE == kActRefresh;
    constexpr bool COUNT = MODE != kActSum;
    __shared__ double part_act[kRefreshWaves][kWave];
    __shared__ int part_nz[kRefreshWaves][kWave];
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    const int64_t h = blk * kWave + l;
    double act = 0;
    int nz = 0;
    if (h < n_hp) {
        const int64_t base = t.off[h];
        const int n = t.n[h];
        // the voxel index and length of the next step are loaded before
        // this step's gather, so the two loads of a step do not wait in turn
        int vox_next = 0;
        double len_next = 0;
        if (w < n) {
            vox_next = s.vox[base + (int64_t)w * kWave];
            len_next = sample_len(s, base + (int64_t)w * kWave);
        }
        for (int k = w; k < n; k += kRefreshWaves) {
            const int64_t i = base + (int64_t)k * kWave;
            const int vox = vox_next;
            const double len = len_next;
            if (k + kRefreshWaves < n) {
                vox_next = s.vox[i + kRefreshWaves * kWave];
                len_next = sample_len(s, i + kRefreshWaves * kWave);
            }
            const double ar = lam[vox] * len;
            act += ar;
            if constexpr (FOLD) {
                s.a[i] = (float)ar;
                if (s.a64) s.a64[i] = ar;
            }
            // the fold's own precision decides which terms are zero
            if constexpr (COUNT) nz += s.fold_a64 ? (ar != 0.0) : ((float)ar != 0.0f);
        }
    }
    part_act[w][l] = act;
    part_nz[w][l] = nz;
    __syncthreads();
    if (w == 0 && h < n_hp) {
        double sum = part_act[0][l];
        int cnt = part_nz[0][l];
#pragma unroll
        for (int j = 1; j < kRefreshWaves; ++j) {
            sum += part_act[j][l];
            cnt += part_nz[j][l];
        }
        t.act[h] = sum;
        if constexpr (FOLD) t.n_fold[h] = cnt;
        if constexpr (MODE == kActCheck) {
            if (cnt != t.n_fold[h]) atomicOr(changed, 1);
        }
    }
}

__global__ __launch_bounds__(kRefreshWaves * kWave) void k_activity_refresh(HalfPathTable t, SampleTable s,
                                                                            int64_t n_hp,
                                                                            const double* __restrict__ lam,
                                                                            const int32_t* run_if)
{
    if (run_if && *run_if == 0) return;   // the same value for every thread
    activity_block<kActRefresh>(t, s, n_hp, lam);
}

__global__ __launch_bounds__(kRefreshWaves * kWave) void k_activity_check(HalfPathTable t, SampleTable s,
                                                                          int64_t n_hp,
                                                                          const double* __restrict__ lam,
                                                                          int32_t* changed, const int32_t* run_if)
{
    if (run_if && *run_if == 0) return;   // the same value for every thread
    activity_block<kActCheck>(t, s, n_hp, lam, changed);
}

// A new activity image has the fold table's terms when every voxel is zero
// exactly where the table's image was, and every other voxel is a finite
// value >= 2^-148 / (smallest positive float32 sample length): then every
// product lam L of such a length is at least 2^-148, so its float32 term is
// not zero (these kernels keep float32 denormals: (float)(lam L) is zero
// only for lam L <= 2^-150) and its float64 term neither, and the same
// samples keep non-zero terms and the same half paths a non-zero integral.
// Walks can produce lengths of 1e-30 cm (a ray that starts on a voxel
// plane), so the level must not be much above the float32 underflow.  tab
// and fast[1] come from the table's image (k_vox_tab).  Any voxel that fails
// sets fast[0].
__device__ __forceinline__ double vox_lam_safe(const uint32_t* len_min)
{
    return ldexp(1.0, -148) / (double)__uint_as_float(*len_min);
}

__global__ void k_vox_check(const double* __restrict__ lam, int64_t n_vox, const uint8_t* __restrict__ tab,
                            const uint32_t* __restrict__ len_min, int32_t* fast)
{
    const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (fast[1]) {
        if (i0 == 0) fast[0] = 1;
        return;
    }
    const double safe = vox_lam_safe(len_min);
    bool bad = false;
    for (int64_t v = i0; v < n_vox; v += (int64_t)gridDim.x * blockDim.x) {
        const double l = lam[v];
        bad |= tab[v] ? !(isfinite(l) && l >= safe) : !(l == 0.0);
    }
    if (bad) fast[0] = 1;
}

// The table's image: its non-zero voxels, and fast[1] = 1 when one of them is
// not a finite value >= the safe level.  One workgroup, so the flag is
// written once, after every voxel was seen.
constexpr int kVoxTabThreads = 1024;

__global__ __launch_bounds__(kVoxTabThreads) void k_vox_tab(const double* __restrict__ lam, int64_t n_vox,
                                                            uint8_t* __restrict__ tab,
                                                            const uint32_t* __restrict__ len_min, int32_t* fast,
                                                            const int32_t* run_if)
{
    if (run_if && *run_if == 0) return;   // the same value for every thread
    const double safe = vox_lam_safe(len_min);
    int bad = 0;
    for (int64_t v = threadIdx.x; v < n_vox; v += kVoxTabThreads) {
        const double l = lam[v];
        const bool nz = !(l == 0.0);
        tab[v] = nz ? 1 : 0;
        if (nz && !(isfinite(l) && l >= safe)) bad = 1;
    }
    bad = __syncthreads_or(bad);
    if (threadIdx.x == 0) fast[1] = bad ? 1 : 0;
}

// Weights-only refresh of the fold table (same layout: the same terms are
// non-zero): record r gets the FP32 (and FP64) term of the sample it was
// copied from, lam[voxel] * length as k_activity_refresh forms it.  A
// record whose term became zero sets *changed (with the count check in
// k_activity_check, equal counts and all records non-zero mean the same set
// of terms).
__global__ __launch_bounds__(256) void k_fold_weights(SampleTable s, const int64_t* __restrict__ n_records,
                                                      int64_t cap, const double* __restrict__ lam,
                                                      int32_t* changed)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= cap || r >= *n_records) return;
    const int2 m = s.fold_map[r];
    if (m.x < 0) return;   // padding record: weight 0 already
    const double len = s.fold_a64 ? s.len[m.y] : (double)__int_as_float(m.y);
    const double ar = lam[m.x] * len;
    const bool keep = s.fold_a64 ? ar != 0.0 : (float)ar != 0.0f;
    if (!keep) atomicOr(changed, 1);
    s.fold[(r / kSampleGroup) * kFoldGroupWords + 2 * kSampleGroup + r % kSampleGroup] = __float_as_uint((float)ar);
    if (s.fold_a64) s.fold_a64[r] = ar;
}

// k_activity_refresh for one-bin subjects, which have no fold table: only
// the integral.
__global__ __launch_bounds__(kRefreshWaves * kWave) void k_activity_sum(HalfPathTable t, SampleTable s,
                                                                        int64_t n_hp,
                                                                        const double* __restrict__ lam)
{
    activity_block<kActSum>(t, s, n_hp, lam);
}

constexpr int kCompactLanes = 16;   // lanes per half path in k_fold_compact

// Record r of the fold table (packed groups of kSampleGroup records).
__device__ __forceinline__ void fold_store(SampleTable s, int64_t r, int64_t nq, float a)
{
    uint32_t* g = s.fold + (r / kSampleGroup) * kFoldGroupWords;
    const int i = (int)(r % kSampleGroup);
    g[2 * i] = (uint32_t)nq;
    g[2 * i + 1] = (uint32_t)((uint64_t)nq >> 32);
    g[2 * kSampleGroup + i] = __float_as_uint(a);
}

// Copies the samples with lam L != 0 of each half path, in order, to the
// fold table at off_fold (scanned from the padded n_fold), then pads the
// half path to a whole group with copies of its last kept record of weight
// 0.  One group of kCompactLanes lanes per half path; a lane's place among
// the kept samples of one pass comes from the wave's ballot.
// Workgroups of 256 threads take 16 half paths, so the 64 paths of one
// interleaved block (which share cache lines) take 4 workgroups.  Those 4 go
// to the same XCD (workgroup g runs on XCD g % 8), so the block's lines are
// fetched into one L2 once rather than into 4.
constexpr int kXcds = 8;
constexpr int kCompactParts = kWave * kCompactLanes / 256;

__global__ __launch_bounds__(256) void k_fold_compact(HalfPathTable t, SampleTable s, int64_t n_hp,
                                                      const int32_t* run_if)
{
    if (run_if && *run_if == 0) return;   // the same value for every thread
    const int64_t g = blockIdx.x;
    const int64_t slot = g / kXcds;
    const int64_t path_block = (slot / kCompactParts) * kXcds + g % kXcds;
    const int64_t h = path_block * kWave + (slot % kCompactParts) * (256 / kCompactLanes) +
                      threadIdx.x / kCompactLanes;
    const int j = threadIdx.x % kCompactLanes;
    const int g0 = (threadIdx.x % kWave) - j;   // the group's first lane in the wave
    if (h >= n_hp) return;
    const int64_t base = t.off[h];
    const int n = t.n[h];
    const int64_t dst = t.off_fold[h];
    int kept = 0;
    int64_t last_nq = 0;
    for (int k0 = 0; k0 < n; k0 += kCompactLanes) {
        const int k = k0 + j;
        int64_t nq = 0;
        float a = 0.0f;
        double a64 = 0.0;
        if (k < n) {
            const int64_t i = base + (int64_t)k * kWave;
            nq = s.nq[i];
            a = s.a[i];
            if (s.fold_a64) a64 = s.a64[i];
        }
        // the same test as the count in k_activity_refresh
        const bool keep = k < n && (s.fold_a64 ? a64 != 0.0 : a != 0.0f);
        const uint64_t wave_mask = __ballot(keep);
        const uint32_t gm = (uint32_t)(wave_mask >> g0) & ((1u << kCompactLanes) - 1);
        if (keep) {
            const int at = kept + __popc(gm & ((1u << j) - 1));
            fold_store(s, dst + at, nq, a);
            if (s.fold_a64) s.fold_a64[dst + at] = a64;
            if (s.fold_map) {
                const int64_t i = base + (int64_t)k * kWave;
                s.fold_map[dst + at] = make_int2(s.vox[i], s.fold_a64 ? (int32_t)i : __float_as_int(s.len32[i]));
            }
        }
        if (gm) last_nq = __shfl(nq, g0 + 31 - __clz(gm), kWave);
        kept += __popc(gm);
    }
    const int padded = (kept + kSampleGroup - 1) / kSampleGroup * kSampleGroup;
    if (kept + j < padded) {
        fold_store(s, dst + kept + j, last_nq, 0.0f);
        if (s.fold_a64) s.fold_a64[dst + kept + j] = 0.0;
        if (s.fold_map) s.fold_map[dst + kept + j] = make_int2(-1, 0);
    }
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
                                   const double* lam, hipStream_t st, const int32_t* run_if)
{
    if (n_hp == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activity_refresh, dim3(blocks_for(n_hp, kWave)), dim3(kRefreshWaves * kWave), 0, st, t, s,
                       n_hp, lam, run_if);
    return hipGetLastError();
}

hipError_t launch_activity_check(const HalfPathTable t, SampleTable s, int64_t n_hp, const double* lam,
                                 int32_t* changed, hipStream_t st, const int32_t* run_if)
{
    if (n_hp == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activity_check, dim3(blocks_for(n_hp, kWave)), dim3(kRefreshWaves * kWave), 0, st, t, s,
                       n_hp, lam, changed, run_if);
    return hipGetLastError();
}

hipError_t launch_fold_weights(SampleTable s, const int64_t* n_records, int64_t cap, const double* lam,
                               int32_t* changed, hipStream_t st)
{
    if (cap == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fold_weights, dim3(blocks_for(cap, 256)), dim3(256), 0, st, s, n_records, cap, lam,
                       changed);
    return hipGetLastError();
}

// The FP32 pair evaluation's record of half path h (after the fold table's
// offsets are scanned).
__device__ __forceinline__ void pair_half_record(const HalfPathTable& t, int64_t h, bool one_bin)
{
    PairHalf r;
    r.dist = t.dist[h];
    const double act = t.act[h];
    if (one_bin) r.act = act;
    else r.off = t.off_fold[h];
    const double att = t.att[h];
    r.n_flags = (one_bin ? 0 : t.n_fold[h]) | (att == 0.0 ? kHalfAttZero : 0) | (act != 0.0 ? kHalfActive : 0);
    r.cosabs = t.cosabs[h];
    r.lnatt = (float)t.lnatt[h];
    r.att = (float)att;
    t.rec[h] = r;
}

__global__ void k_pair_half(HalfPathTable t, int64_t n_hp, bool one_bin, const int32_t* run_if_a,
                            const int32_t* run_if_b)
{
    if (run_if_a && run_if_b && *run_if_a == 0 && *run_if_b == 0) return;
    const int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= n_hp) return;
    pair_half_record(t, h, one_bin);
}

// k_activity_sum then k_pair_half for many one-bin subjects in one launch:
// workgroup i takes block blocks[i].y of 64 half paths of member
// blocks[i].x, with the same arithmetic as the two kernels; the lane that
// stored a path's ActPath then writes its pair record.
__global__ __launch_bounds__(kRefreshWaves * kWave) void k_one_bin_refresh(const OneBinRefresh* __restrict__ m,
                                                                           const int2* __restrict__ blocks)
{
    const int2 b = blocks[blockIdx.x];
    const OneBinRefresh& r = m[b.x];
    activity_block<kActSum>(r.t, r.s, r.n_hp, r.lam, nullptr, b.y);
    const int64_t h = (int64_t)b.y * kWave + threadIdx.x % kWave;
    if (threadIdx.x < kWave && h < r.n_hp) pair_half_record(r.t, h, true);
}

hipError_t launch_pair_half(const HalfPathTable t, int64_t n_hp, bool one_bin, hipStream_t st,
                            const int32_t* run_if_a, const int32_t* run_if_b)
{
    if (n_hp == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pair_half, dim3(blocks_for(n_hp, 256)), dim3(256), 0, st, t, n_hp, one_bin, run_if_a,
                       run_if_b);
    return hipGetLastError();
}

hipError_t launch_vox_check(const double* lam, int64_t n_vox, const uint8_t* tab, const uint32_t* len_min,
                            int32_t* fast, hipStream_t st)
{
    if (n_vox == 0) return hipSuccess;
    hipLaunchKernelGGL(k_vox_check, dim3((unsigned)std::min<int64_t>(1024, blocks_for(n_vox, 256))), dim3(256), 0, st,
                       lam, n_vox, tab, len_min, fast);
    return hipGetLastError();
}

hipError_t launch_vox_tab(const double* lam, int64_t n_vox, uint8_t* tab, const uint32_t* len_min, int32_t* fast,
                          hipStream_t st, const int32_t* run_if)
{
    hipLaunchKernelGGL(k_vox_tab, dim3(1), dim3(kVoxTabThreads), 0, st, lam, n_vox, tab, len_min, fast, run_if);
    return hipGetLastError();
}

// Volume in C order of an (X, Y, Z) array (z fastest) to the x-fastest
// layout the kernels read: a 32 x 32 (x, z) tile per workgroup at one y,
// through LDS, so both the reads and the writes are coalesced.
__global__ __launch_bounds__(256) void k_volume_x_fastest(const double* __restrict__ src, double* __restrict__ dst,
                                                          int X, int Y, int Z)
{
    __shared__ double tile[32][33];
    const int z0 = blockIdx.x * 32, x0 = blockIdx.y * 32, y = blockIdx.z;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int j = ty; j < 32; j += 8) {
        const int x = x0 + j, z = z0 + tx;
        if (x < X && z < Z) tile[j][tx] = src[((int64_t)x * Y + y) * Z + z];
    }
    __syncthreads();
    for (int j = ty; j < 32; j += 8) {
        const int z = z0 + j, x = x0 + tx;
        if (x < X && z < Z) dst[x + (int64_t)X * (y + (int64_t)Y * z)] = tile[tx][j];
    }
}

hipError_t launch_volume_x_fastest(const double* src, double* dst, int X, int Y, int Z, hipStream_t st)
{
    if ((int64_t)X * Y * Z == 0) return hipSuccess;
    const dim3 grid((unsigned)((Z + 31) / 32), (unsigned)((X + 31) / 32), (unsigned)Y);
    hipLaunchKernelGGL(k_volume_x_fastest, grid, dim3(256), 0, st, src, dst, X, Y, Z);
    return hipGetLastError();
}

hipError_t launch_one_bin_refresh(const OneBinRefresh* m, const int2* blocks, int n_blocks, hipStream_t st)
{
    if (n_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_one_bin_refresh, dim3((unsigned)n_blocks), dim3(kRefreshWaves * kWave), 0, st, m, blocks);
    return hipGetLastError();
}

hipError_t launch_activity_sum(const HalfPathTable t, SampleTable s, int64_t n_hp, const double* lam,
                               hipStream_t st)
{
    if (n_hp == 0) return hipSuccess;
    hipLaunchKernelGGL(k_activity_sum, dim3(blocks_for(n_hp, kWave)), dim3(kRefreshWaves * kWave), 0, st, t, s,
                       n_hp, lam);
    return hipGetLastError();
}

hipError_t launch_fold_compact(const HalfPathTable t, SampleTable s, int64_t n_hp, hipStream_t st,
                               const int32_t* run_if)
{
    if (n_hp == 0) return hipSuccess;
    const int64_t path_blocks = (n_hp + kWave - 1) / kWave;
    const int64_t groups = (path_blocks + kXcds - 1) / kXcds;   // of kXcds path blocks
    hipLaunchKernelGGL(k_fold_compact, dim3((unsigned)(groups * kXcds * kCompactParts)), dim3(256), 0, st,
                       t, s, n_hp, run_if);
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
