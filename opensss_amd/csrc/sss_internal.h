// Internal declarations shared by the HIP translation units of libsss_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <string>

#include "../../include/sss_hip.h"

namespace sss {

constexpr int kTofHalf = 4000;            // TOF table half width, ps
constexpr double kLightCmPerPs = 299792458.0 * 1e-10;
constexpr int kWave = 64;
// A wave owns 8 x 8 cells: 8 rows (ring r1, detector i1) x 8 columns (ring
// r2, detector i2), each a block of 2^k rings x 8/2^k detectors.  Path
// lengths vary far less across neighbouring rings than across neighbouring
// detectors, so ring blocks keep the lanes' trip counts close (a CPU model
// of config 2 with the lanes taking the scatter points in step: 88% of lane
// slots busy with 4 rings x 2 detectors, 78% with 2 x 4, 61% with 1 x 8).
constexpr int kTileRows = 8;
constexpr int kTileCols = 8;
constexpr int kTileMaxRings = 4;
constexpr int kScatterWaves = 4;          // waves per workgroup of the scatter launch
// Energy rows a pair can reach: e = floor(511 / (2 + cos) + 1/2) keV for
// cos in [-1, 1] is 170..511 keV, row 2e - 1.  The LDS energy table of the
// one-bin scatter holds these rows and row 0 (reached by a NaN cosine only).
constexpr int kEnergyRowLo = 2 * 170 - 1;
constexpr int kEnergyRowHi = 2 * 511 - 1;
constexpr int kEnergyRows = kEnergyRowHi - kEnergyRowLo + 2;
constexpr int kAutoChunkPoints = 64;      // at most this many scatter points per chunk by default
constexpr int kMaxBins = 32;

// TOF weight evaluation of the fold (see tof_weights in sss_scatter.hip)
constexpr int kTofGeneral = 0;            // one exp2 per window bin
constexpr int kTofOneExp = 1;             // h = 1: one exp2, end bins masked
constexpr int kTofOneExpInner = 2;        // h = 1: one exp2, window centre kept inside the bins
// float only: kTofOneExpInner with a window that slides along a half path
// (the centre bin moves one bin at a time, in one direction); launched in its
// kTofSlideQuad form
constexpr int kTofSlide = 3;
// float only, h = 1 or 2: a sliding window centred on the nearest bin with
// the side weights in relative form, 2^(+-ku - beta2) (two exp2 per sample):
// no float range limit on beta2 (h = 1) or on the offsets' reach
constexpr int kTofSlideRel = 4;
// float only, h = 1 with beta2 above 95 (bins far wider than the kernel):
// kTofSlideRel's window keeping only the side weight that ku points to,
// 2^(|ku| - beta2) (the other is below 2^-95 of the centre)
constexpr int kTofSlideRelWide = 5;
// float only: kTofSlide with the row path's fold records shared by the four
// lanes of a quad (each lane loads a third of a group, DPP broadcasts the rest)
constexpr int kTofSlideQuad = 6;
constexpr int kMaxHalfWindow = 8;

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define SSS_HIP(call)                                         \
    do {                                                      \
        hipError_t e_ = (call);                               \
        if (e_ != hipSuccess) return ::sss::hip_fail(e_, #call); \
    } while (0)

// Geometry of the voxel grid, float64 as the reference.
struct Grid {
    int nx, ny, nz;
    double b[6];
};

// Per-half-path scalars (structure of arrays, one entry per (sp, ring, det)).
struct HalfPathTable {
    int64_t* off;      // first sample; sample k at off + k * kWave (see SampleTable)
    int32_t* n;        // samples
    int32_t* blk;      // per block of kWave half paths: kWave * its largest n
    int64_t* blk_off;  // per block: its first record (scan of blk)
    int64_t* off_fold; // first record in the fold table
    int32_t* n_fold;   // samples with lam L != 0 (the ones the fold visits)
    double* dist;      // |D - S|, cm
    double* att;       // exp(-sum mu L)
    double* lnatt;     // sum mu L
    double* act;       // sum lam L (f64, exact-zero tests)
    float* cosabs;     // |cos| of the angle to the detector normal
    int32_t* status;   // 0 ok, 1 walk overflow
    struct PairHalf* rec;   // per activity image: what the FP32 pair evaluation reads
    uint32_t* len_min; // smallest positive stored sample length (float bits), set by the fill
};

// One half path as the FP32 pair evaluation reads it: two 16-byte loads.
// The zero tests of the float64 attenuation and activity are kept as flags,
// so they stay exact.  A one-bin subject has no fold table: its records hold
// the activity integral in place of the first fold record.
struct alignas(16) PairHalf {
    double dist;       // |D - S|, cm
    union {
        int64_t off;   // first fold record (TOF subjects)
        double act;    // sum lam L (one-bin subjects)
    };
    int32_t n_flags;   // fold records | kHalfAttZero | kHalfActive
    float cosabs;
    float lnatt;       // sum mu L
    float att;         // exp(-sum mu L)
};
constexpr int32_t kHalfAttZero = 1 << 30;   // exp(-sum mu L) == 0
constexpr int32_t kHalfActive = 1 << 29;    // sum lam L != 0
constexpr int32_t kHalfCountMask = kHalfActive - 1;

// TOF time offsets are carried in 32.32 fixed point: one 64-bit add and the
// high word give floor(X - q) exactly for X, q quantised to 2^-32 ps.  The
// reference's offsets fall within 1e-6 ps of an integer for about one
// sample in a thousand (the fractions of (d1 - d2)/c and 2 s_k/c are often
// nearly complementary), so the fraction needs far more than 20 bits.
constexpr double kFixOne = 4294967296.0;   // 2^32
// In the fold table each half path's records are padded to a multiple of
// kSampleGroup with zero-weight copies of its last record (the FP32 fold runs
// whole groups).
constexpr int kSampleGroup = 4;
constexpr int kSamplePad = 8;              // fold records after the last one (prefetch reads up to 7 ahead)

// Per-sample arrays.  nq/vox/len: every voxel sample of every half path,
// built by the tracer; a/a64: its emission term, rewritten per activity
// image.  The half paths h of one block of kWave (h / kWave) are
// interleaved: sample k of h is entry blk_off[h / kWave] + k * kWave +
// h % kWave, so a wave that walks the block's paths in step writes and reads
// kWave consecutive entries.  fold/fold_a64: per activity image, the samples
// with lam L != 0 only (their terms are exactly zero otherwise), offsets
// HalfPathTable::off_fold, padded per half path to whole groups, kSamplePad
// spare records at the end.  The fold table packs each group of 4 records
// into 12 words (three 16-byte loads): nq of records 0..3 (low, high word),
// then a of records 0..3.
struct SampleTable {
    int64_t* nq;       // -round(2^32 q), q = 2 s_k / c in ps
    int32_t* vox;      // 0-based voxel (reference linear index - 1)
    double* len;       // intersection length, cm (FP64 subjects)
    float* len32;      // the same in FP32 (FP32 subjects, which read 8 bytes a sample instead of 12)
    float* a;          // lam * len (FP32 fold input)
    double* a64;       // lam * len in float64 (FP64 validation path only, else null)
    uint32_t* fold;    // kFoldGroupWords per group of kSampleGroup records
    double* fold_a64;  // FP64 path only
    int2* fold_map;    // per fold record {voxel, FP32 subjects: the sample's
                       // float length bits; FP64 subjects: its sample entry}
                       // ({-1, 0}: padding), so a new activity image with the
                       // same zero terms only rewrites the weights, reading
                       // one coalesced word per record
};
constexpr int kFoldGroupWords = 3 * kSampleGroup;

struct ScatterArgs {
    // sizes
    int n_sp, nru, ndu, n_bins, n_energy, eff_ref_row;
    int n_tiles, n_chunks;
    int64_t n_cells_plane;       // NDU (NDU+1)/2
    // tables
    const PairHalf* hp_rec;      // per half path, FP32 pair evaluation
    const float2* etab;          // kEnergyRows: (eff * eff_511 or -1 where 0, ratio)
    int etab_lo, etab_hi;        // rows kEnergyRowLo/Hi clamped into the tables
    const int64_t* hp_off;       // first fold record of each half path
    const int32_t* hp_n;         // fold records (samples with lam L != 0)
    const int32_t* hp_n_all;     // all samples of each half path
    const double* hp_dist;
    const double* hp_att;
    const double* hp_lnatt;
    const double* hp_act;
    const float* hp_cos;
    const uint32_t* s_fold;      // fold table (packed groups): fixed-point TOF offsets, FP32 emission terms
    const double* s_a64;         // fold table FP64 emission terms (FP64 path)
    const double* detpos;        // (NRU, NDU, 3)
    const double* eff;           // n_energy
    const float* ratio;          // n_energy
    const int4* tiles;           // (r1_0, r2_0, i1_0, i2_0) block origins
    int tile_db_log2;            // log2 of the detectors per row/column block
    const int2* items;           // per wave of the scatter launch: (tile or -1, chunk)
    int n_items;
    const int32_t* slice_ptr;    // n_slices + 1
    const int32_t* chunk_sp;     // n_chunks + 1: scatter-point range of each chunk (inside one slice)
    // TOF kernel: bins b at s_b = (b - mid) W; weights of bin j0 + d relative to
    // the nearest bin j0 are 2^(d ku - d^2 beta2), ku = kappa2 (t - s_j0)
    double inv_w_bins, mid, bin_w, kappa2, beta2;  // log2 scaled kappa, beta
    double c_beta;                                 // 2^-beta2
    int half_window, t_lo, t_hi;
    int tof_mode;                                  // kTofGeneral / kTofOneExp / kTofOneExpInner
    int range_check;                               // offsets may leave [t_lo, t_hi]: test each sample
    // the same constants rounded to float once on the host (uniform registers)
    float f_inv_w, f_kap, f_mid, f_jmax, f_wk, f_mid_wk, f_c, f_inv_c;
    // sliding windows (kTofSlide, kTofSlideRel): LDS row of bin b is
    // b + slide_row0; centre thresholds L(b) = ceil((b - mid - 1/2) W) and
    // ku offsets kappa2 (mid - b) W are computed from these float constants
    int slide_row0;
    float f_beta2;               // beta2
    float f_w, f_l0, f_l1;       // W, (-mid - 1/2) W, (1/2 - mid) W: L(b) = ceil(b W + f_l0)
    float f_wk_lo, f_mid_wk_lo;  // W kappa2 - f_wk, mid W kappa2 - f_mid_wk (the rounding parts)
    float f_c4, f_kclamp;        // 2^(-4 beta2); the h = 2 relative window's ku bound, beta2 + 10
    float f_kclamp1;             // the h = 1 relative window's ku bound, beta2 + 30
    double g_const;              // (Dx Dy 1e-2)^2 / (4 pi)
    double kn_gamma;
    // output: per chunk, the un-normalised sums and the pair counts of every
    // lane slot (tile * 64 + lane, so a wave writes 64 contiguous values per
    // bin); the chunk reduce divides per slice, sums the slices and scatters
    // the slots into the compact output
    void* part_sum;              // [chunk][B][n_tiles * 64], float or double
    float* part_cnt;             // [chunk][n_tiles * 64]
    const int32_t* chunk_slice;  // n_chunks: slice of each chunk
    void* out;                   // compact output [B][NRU^2][cells]
};

// Largest |TOF offset| in ps the fold can produce for a subject: the
// detector-distance difference of any pair plus the largest sample time
// (s_k <= half the clipped path, so q <= box diagonal / c), plus rounding.
double tof_offset_bound(const Grid& g, const double* sp_xyz, int n_sp, const double* detpos,
                        int n_det_pos);

// Launchers (defined in the .hip files).
hipError_t launch_halfpath_count(const Grid& g, const double* sp_xyz, int n_sp, const double* detpos,
                                 const double* detnorm, int nru, int ndu, HalfPathTable t,
                                 hipStream_t st);
hipError_t launch_halfpath_fill(const Grid& g, const double* sp_xyz, int n_sp, const double* detpos,
                                int nru, int ndu, const double* mu, HalfPathTable t, SampleTable s,
                                hipStream_t st);
// Emission terms lam L of every sample, their sums and the fold-table counts
// (then the caller scans the padded counts into off_fold), and the copy of
// the non-zero terms into the fold table.
hipError_t launch_activity_check(const HalfPathTable t, SampleTable s, int64_t n_hp, const double* lam,
                                 int32_t* changed, hipStream_t st, const int32_t* run_if = nullptr);
// Voxel test of a new activity image against the one the fold table was
// built from (fast[0]: set on any difference; fast[1]: the table image was
// unfit for the test; tab: its non-zero voxels), and the record of that image
// when the table is rebuilt (run_if as below).
hipError_t launch_vox_check(const double* lam, int64_t n_vox, const uint8_t* tab, const uint32_t* len_min,
                            int32_t* fast, hipStream_t st);
hipError_t launch_vox_tab(const double* lam, int64_t n_vox, uint8_t* tab, const uint32_t* len_min, int32_t* fast,
                          hipStream_t st, const int32_t* run_if);
hipError_t launch_fold_weights(SampleTable s, const int64_t* n_records, int64_t cap, const double* lam,
                               int32_t* changed, hipStream_t st);
// run_if: when not null, the kernel does its work only if *run_if != 0 (the
// fallback of a weights-only refresh)
hipError_t launch_activity_refresh(const HalfPathTable t, SampleTable s, int64_t n_hp,
                                   const double* lam, hipStream_t st, const int32_t* run_if = nullptr);
hipError_t launch_volume_x_fastest(const double* src, double* dst, int X, int Y, int Z, hipStream_t st);
hipError_t launch_activity_sum(const HalfPathTable t, SampleTable s, int64_t n_hp,
                               const double* lam, hipStream_t st);
hipError_t launch_fold_compact(const HalfPathTable t, SampleTable s, int64_t n_hp, hipStream_t st,
                               const int32_t* run_if = nullptr);
// run_if_a, run_if_b: when both are given, the records are rebuilt only if either is non-zero
hipError_t launch_pair_half(const HalfPathTable t, int64_t n_hp, bool one_bin, hipStream_t st,
                            const int32_t* run_if_a = nullptr, const int32_t* run_if_b = nullptr);
// The activity refresh of a one-bin subject (ActPath per half path, then its
// pair record) for many subjects in one launch: blocks[i] = {member, block of
// 64 half paths}, the member's tables in m.
struct OneBinRefresh {
    HalfPathTable t;
    SampleTable s;
    int64_t n_hp;
    const double* lam;
};
hipError_t launch_one_bin_refresh(const OneBinRefresh* m, const int2* blocks, int n_blocks, hipStream_t st);
hipError_t launch_trace_segments(const Grid& g, const double* segs, int64_t n, int cap,
                                 int32_t* counts, double* L, int32_t* idx, double* s,
                                 hipStream_t st);
hipError_t launch_line_integrals(const Grid& g, const double* segs, int64_t n, const double* mu,
                                 const double* lam, int32_t* counts, double* mu_int,
                                 double* lam_int, hipStream_t st);
hipError_t launch_scatter(const ScatterArgs& a, bool fp64, bool trivial, hipStream_t st);
hipError_t launch_pair_stats(const ScatterArgs& a, unsigned long long* out, hipStream_t st);
hipError_t launch_chunk_reduce(const ScatterArgs& a, bool fp64, hipStream_t st);

// Several subjects in one persistent scatter launch (k_scatter_queue).  The
// work is workgroup blocks {subject, block of its item list}; block b of a
// subject's list was laid out for XCD b % 8, so the blocks sit in 8 queues,
// one per XCD, back to back in `blocks` (queue x: [q_begin[x], q_begin[x+1])).
constexpr int kXcdCount = 8;
struct BatchQueue {
    const ScatterArgs* args;   // per subject
    const int2* blocks;
    const int32_t* q_begin;    // kXcdCount + 1
    int32_t* counters;         // kXcdCount, zero before the launch
};
// Kernel variant of a subject's scatter (subjects of one variant can share a
// launch), the persistent grid of a queue launch, and the launch itself
// (a: any subject of the launch, for the variant).
int scatter_variant(const ScatterArgs& a, bool fp64, bool trivial);
hipError_t scatter_queue_grid(const ScatterArgs& a, bool fp64, bool trivial, int max_bins, int n_cu, int n_blocks,
                              int* grid);
hipError_t launch_scatter_queue(const ScatterArgs& a, bool fp64, bool trivial, const BatchQueue& q, int max_bins,
                                int grid, hipStream_t st);
// Chunk reduces of several subjects: blocks[x] = {subject, block of 256 of
// its lane slots}; reduce_blocks gives a subject's block count.
int reduce_blocks(const ScatterArgs& a);
hipError_t launch_chunk_reduce_queue(const ScatterArgs* args, const int2* blocks, int n_blocks, int max_bins,
                                     bool fp64, hipStream_t st);
// n int32 values set to v by a kernel (graph captures record no memset)
hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st);
hipError_t launch_interp(const float* sampled, int n_bins, int n_sampled_planes, int64_t cells_s,
                         const int32_t* cell_map, int64_t n_cells_out, const int32_t* plane_map,
                         int n_out_planes, float* out, float* sum, hipStream_t st);
hipError_t launch_tail_mask(const Grid& g, const double* image, int use_attenuation, const double* geom,
                            int n_det, const int16_t* lor, int n_planes, const int32_t* plane_rings,
                            int accel, uint8_t* mask, hipStream_t st);
// Device scratch of the tail fit: 64-bit fixed-point images (prompts,
// scatter, randoms; sensitivity as a count), 4 * nvox; the per-block partial
// sums of the reduction; the three largest masked values (float bits).
struct FitScratch {
    unsigned long long* img;
    double* part;
    unsigned int* vmax;
};
constexpr int kFitBlocks = 256;              // blocks of the fit's fixed-order reduction
constexpr size_t fit_scratch_bytes(int64_t nvox)
{
    return (size_t)(4 * nvox) * sizeof(unsigned long long) + (size_t)(2 * kFitBlocks) * sizeof(double) +
           4 * sizeof(unsigned int) + 256;
}
FitScratch fit_scratch_carve(void* base, int64_t nvox);   // base 256-byte aligned, fit_scratch_bytes long
// sums (device, 2 doubles) = (sum S_bp T, sum S_bp^2); images_out (device,
// 4 * nvox doubles, may be null) = the four images in float64.
hipError_t launch_fit(const Grid& g, const double* geom, int n_rings_geom, int n_det, const int16_t* lor, int n_entries,
                      const int32_t* plane_rings, int n_value_planes, int accel, int length_weighted,
                      const float* scatter, const float* prompts, const float* randoms, const float* mask,
                      FitScratch scratch, double* sums, double* images_out, hipStream_t st);
// The sinograms of a tail fit over every TOF bin (device pointers, cell
// c = (radial * N/2 + angular) * P + plane): bin b fits scatter[b * n_cells +
// c] against prompts[c * n_bins + b] / counts (float32 division; counts null
// = 1, one per plane, or one per cell), randoms[c] * norm[c] (null randoms =
// 0, null norm = 1) and the cells with mask[c] >= 1.
struct BinFitIn {
    const float* scatter;
    const void* prompts;
    int prompts_u16;          // prompts are uint16 (else float32)
    const float* counts;
    int counts_per_plane;
    const float* randoms;
    const float* norm;
    const uint8_t* mask;
    int n_bins;
    int64_t n_cells;
    int n_value_planes;
};
constexpr size_t fit_bins_scratch_bytes(int64_t nvox, int n_bins)
{
    return (size_t)((4 * n_bins + 2) * nvox) * sizeof(unsigned long long) +
           (size_t)(2 * kFitBlocks) * n_bins * sizeof(double) + (size_t)(3 * n_bins + 1) * sizeof(unsigned int) + 512;
}
FitScratch fit_bins_scratch_carve(void* base, int64_t nvox, int n_bins);
// sums (device, 2 n_bins doubles): bin b's two sums at [2b], [2b + 1], the
// same bits as launch_fit of that bin's sinograms (prompts already divided).
hipError_t launch_fit_bins(const Grid& g, const double* geom, int n_rings_geom, int n_det, const int16_t* lor,
                           int n_entries, const int32_t* plane_rings, int accel, int length_weighted,
                           const BinFitIn& in, FitScratch scratch, double* sums, hipStream_t st);
hipError_t launch_forward_project(const Grid& g, const double* lam, const double* mu, const double* geom, int n_det,
                                  const int16_t* lor, int n_planes, const int32_t* plane_rings, int accel,
                                  float* out, hipStream_t st);
hipError_t launch_poisson(const float* mean, int64_t n, double scale, uint64_t seed, float* out, hipStream_t st);

// A CASToR list-mode event record: `words` 4-byte words, the TOF float at
// tof_word, the first crystal ID pair at id_word; InjectScatters writes the
// scatter rate at scatter_word (the later words move up by one).
struct ListmodeLayout {
    int words, tof_word, id_word, scatter_word;
};
// Tables an event is binned through (device pointers): the sinogram
// coordinates of the n_det_lor detectors (mashed, (N, N, 2) int16) and the
// (n_rings_index, n_rings_index) int16 ring-pair -> plane table; the
// sinogram is (n_rad, n_ang, n_planes).
struct ListmodeTables {
    const int16_t* lor;
    int n_det_lor, mash;
    const int16_t* ring_index;
    int n_rings_index;
    int n_rad, n_ang, n_planes;
};
hipError_t launch_lm_histogram(const uint32_t* events, int64_t n_events, int64_t first_index,
                               const ListmodeLayout& lay, const ListmodeTables& t, int64_t ncpr, int64_t shift,
                               int n_bins, double tof_range, double duration, uint32_t* counts,
                               unsigned long long* randoms_last, unsigned long long* norm_last, int* bad,
                               hipStream_t st);
hipError_t launch_lm_finish(const uint32_t* counts, int64_t n_counts, uint16_t* prompts,
                            const unsigned long long* randoms_last, const unsigned long long* norm_last,
                            int64_t n_cells, float* randoms, float* norm, hipStream_t st);
int64_t lm_scan_blocks(int64_t n_events);
hipError_t launch_lm_select(const uint32_t* events, int64_t n_events, const ListmodeLayout& lay,
                            const ListmodeTables& t, int64_t ncpr, int64_t shift, int n_bins, double tof_range,
                            const float* factor, const float* scatter, int64_t n_cells, int32_t* keep, float* rate,
                            int64_t* pos, int64_t* block_sum, int* bad, hipStream_t st);
hipError_t launch_lm_write(const uint32_t* events, int64_t n_events, const ListmodeLayout& lay, const int32_t* keep,
                           const int64_t* pos, const float* rate, uint32_t* out, hipStream_t st);

// Crop and resize of an x-fastest volume (Utils.py:212-267); scratch of
// resize_scratch_bytes(crop) bytes (256-byte aligned).
int64_t resize_scratch_bytes(const int64_t crop[3]);
hipError_t launch_crop_and_downscale(const float* in, const int64_t in_shape[3], const int64_t crop_start[3],
                                     const int64_t crop[3], const int64_t out_shape[3], const double* const weights[3],
                                     const int radius[3], int clip, float zero_below, void* scratch, float* out,
                                     hipStream_t st);

// Device address of this library's no-allocation fit scratch on the current
// device (looked up once per device at context creation).
int noalloc_scratch_init(int device);
void* noalloc_scratch(int device);
int ctx_device(const sss_ctx* ctx);
hipStream_t ctx_stream(const sss_ctx* ctx);
// A device block from the context's cache (SSS_OK or SSS_ERR_NOMEM with the
// error set); ctx_pool_put hands it back (its next user works on the
// context's stream, after what was queued with it).
int ctx_pool_get(sss_ctx* ctx, size_t bytes, void** p);
void ctx_pool_put(sss_ctx* ctx, void* p);

// What one subject brings to a batch launch (sss_batch.hip).  The subject's
// tables must stay as they are while a batch holds it: `generation` changes
// whenever its walks are redone, and a batch rebuilds its lists then.
struct BatchMember {
    ScatterArgs args;
    bool fp64 = false, trivial = false;
    int variant = -1;
    int n_bins = 0;
    int n_blocks = 0;          // workgroup blocks of its scatter item list
    size_t out_bytes = 0;      // compact output
    const void* d_out = nullptr;
    double work = 0;           // estimated scatter work (for the queue order)
    int generation = 0;
};
// Checks that the subject is traced and has a host activity image on its
// context, brings its activity refresh to the form every later image takes
// (queued refreshes of the current image), and describes it.
int subject_batch_prepare(sss_subject* s, sss_ctx* ctx, BatchMember* m);
int subject_generation(const sss_subject* s);
// Batch membership count: a subject in a live batch is not destroyed.
void subject_batch_ref(sss_subject* s, int delta);
// Copies a host activity image into the subject's pinned staging copy
// (after any queued upload from it has finished).
int subject_stage_activity(sss_subject* s, const double* lam);
// Queued on the subject's context stream: the staged image to the device
// (transposed when the subject takes C-order volumes), then the activity
// refresh.
int subject_queue_activity(sss_subject* s);
// The upload half of subject_queue_activity (staged image to the device, in
// the kernels' layout), for a one-bin subject whose refresh a batch merges.
int subject_queue_upload(sss_subject* s);
// A one-bin subject's tables as launch_one_bin_refresh reads them.
int subject_one_bin_refresh(const sss_subject* s, OneBinRefresh* r);

hipError_t launch_dense_scatter_typed(const void* compact, bool fp64, int n_bins, int nru, int ndu,
                                      int n_det, const int32_t* dets, const int16_t* lor,
                                      float* dense, hipStream_t st);

}  // namespace sss
