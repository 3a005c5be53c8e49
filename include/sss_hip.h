/*
 * sss_hip.h — C ABI of the MI355X TOF single-scatter-simulation library
 * (libsss_hip.so).  Plain C types only; every function returns an int status
 * (SSS_OK = 0, negative on error) and never lets an exception cross the ABI.
 * sss_last_error() returns the message of the last failure on the calling
 * thread.  A context owns one HIP stream on one device; calls on one context
 * must not overlap from several threads.
 *
 * The reference (r-santo/openSSS) is pure Python; these entry points replace
 * the numba kernels behind its Python functions:
 *
 *   sss_trace_segments     Python/openSSS/RayTracing3DTOF.py:29-180
 *                          (RayTracing3DTOF, one segment per row)
 *   sss_subject_create     Python/openSSS/SingleScatterSimulationTOF.py:82-198
 *   + sss_subject_trace    (set-up tables, half paths of :400-430: traces,
 *                           attenuation integrals, detector angles)
 *   sss_subject_set_activity  the activity gathers of :417-418 per request
 *   sss_subject_scatter    bodySSS :433-557 and the slice sum :193-201
 *   sss_subject_output_dense  the dense (B, N+1, N/2, NRU^2) layout of :171,551
 *   sss_interpolate        Python/openSSS/Interpolation.py:95-399
 *   sss_line_integrals     Python/openSSS/TailsMask.py:84-163,230-309 traces
 *   sss_tail_mask          Python/openSSS/TailsMask.py:28-376 (the per-LOR mask loops)
 *   sss_tail_fit           Python/openSSS/BackProjection.py:31-392 (back-projections
 *                          and the scale of Scaling.py:27-121)
 *   sss_scale_to_prompts   Python/openSSS/Scaling.py:60-113 (every TOF bin,
 *                          with the LORCounts division and the mask threshold)
 *   sss_listmode_histogram Python/Timber/Datafile.py:29-115 (ExportPrompts)
 *   sss_listmode_inject    Python/Timber/Datafile.py:117-240 (InjectScatters)
 *   sss_crop_and_downscale Python/openSSS/Utils.py:212-267 (CropAndDownscale)
 */
#ifndef SSS_HIP_H
#define SSS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SSS_OK 0
#define SSS_ERR_INVALID -1      /* bad argument or shape */
#define SSS_ERR_HIP -2          /* HIP runtime error */
#define SSS_ERR_NOMEM -3        /* device or host allocation failed */
#define SSS_ERR_UNSUPPORTED -4  /* valid input outside what the kernels implement */
#define SSS_ERR_TRACE -5        /* a voxel walk did not end where its end voxel says */

#define SSS_PRECISION_FP32 0    /* FP32 fold and sums; FP64 geometry, energy index, TOF offsets */
#define SSS_PRECISION_FP64 1    /* all-FP64 validation path */

typedef struct sss_ctx sss_ctx;
typedef struct sss_subject sss_subject;

/* One scatter-estimation instance: scanner, mu-map and scatter points (host
 * pointers, copied at create).  Arrays use the reference's layouts: images
 * are the Fortran-order flattening of (X, Y, Z) (ActivityMap.flatten("F")),
 * geometry/normals are C-order (NR, N, 3) in cm, scatter points are
 * positions in cm grouped by normalisation slice (slice_ptr, n_slices+1). */
typedef struct {
    int32_t grid[3];
    double bounds[6];            /* ImageSize [xmin ymin zmin xmax ymax zmax] */
    const double* mu;            /* X*Y*Z, cm^-1 */
    int32_t n_rings, n_det;
    const double* geom;
    const double* normals;
    double det_size[2];          /* mm */
    const double* ratio;         /* n_energy water attenuation ratios */
    const double* eff;           /* n_energy detection efficiencies */
    int32_t n_energy;
    int32_t eff_ref_row;         /* row of 511 keV in the efficiency product (1021) */
    int32_t nru, ndu;
    const int32_t* rings;        /* NRU ring ids */
    const int32_t* dets;         /* NDU detector ids, strictly increasing */
    int32_t n_sp;
    const double* sp_xyz;        /* n_sp x 3 */
    int32_t n_slices;
    const int32_t* slice_ptr;
    int32_t n_bins;
    double tof_bin_width;        /* ps */
    double tof_kappa;            /* 2 W / w, w = R^2 / (4 ln 2) */
    double tof_beta;             /* W^2 / w */
    int32_t tof_half_window;     /* bins either side of the nearest bin */
    int32_t tof_t_lo, tof_t_hi;  /* non-zero TOF columns, ps */
    int32_t tof_trivial;         /* one bin, all weights 1 */
    int32_t precision;           /* SSS_PRECISION_* */
    int32_t n_chunks;            /* target number of scatter-point chunks (each inside one
                                    slice) for the scatter launch; 0 = auto */
    int32_t tile_rings;          /* rings per block of a scatter tile's rows and columns
                                    (1, 2, 4 or 8; the block holds 8 / tile_rings detectors);
                                    0 = the default (4, or NRU when smaller) */
    int32_t volume_order;        /* layout of mu here and of every activity image (host or
                                    device) given to the subject: SSS_VOLUME_X_FASTEST (index
                                    x + X y + X Y z, the reference's flatten(order='F')) or
                                    SSS_VOLUME_Z_FASTEST (C order of an (X, Y, Z) array; the
                                    subject transposes on the device) */
} sss_subject_desc;

enum { SSS_VOLUME_X_FASTEST = 0, SSS_VOLUME_Z_FASTEST = 1 };

typedef struct {
    int64_t n_halfpaths;
    int64_t n_samples;           /* sample records: each block of 64 half paths holds 64 x its
                                    longest walk */
    int64_t n_cells;             /* compact cells per bin = NRU^2 * NDU(NDU+1)/2 */
    int64_t n_pairs_enumerated;  /* n_sp * LORs per scatter point */
    int64_t evals;               /* n_pairs_enumerated * n_bins */
    int64_t device_bytes;        /* resident device memory of the subject */
    int32_t n_tiles, n_chunks;
    int32_t n_truncated_walks;   /* half paths whose walk stopped early (voxel-face runs) */
    int32_t tof_fold;            /* TOF weights of the scatter launch: 0 one exp2 per window bin,
                                    1 one exp2 (h = 1, end bins masked), 2 one exp2 with the
                                    window centre inside the bins, 3 as 2 with a window that
                                    slides along each half path (FP32), 4 a sliding window
                                    on the nearest bin, one exp2 per sample (FP32, h = 1 or
                                    2), 5 as 4 keeping only the side weight that the
                                    offset points to (FP32, h = 1, bins far wider than the
                                    kernel), 6 as 3 with the fold records of each cell row's
                                    half path shared by the four lanes of a quad (FP32; what
                                    the FP32 launch of 3's subjects runs) */
} sss_subject_info;

const char* sss_last_error(void);
int sss_version(void);

int sss_ctx_create(int device, sss_ctx** out);
int sss_ctx_destroy(sss_ctx* ctx);
int sss_ctx_synchronize(sss_ctx* ctx);
/* The context's hipStream_t, for callers that time or order work on it. */
void* sss_ctx_stream(sss_ctx* ctx);
/* Device memory cache: a context keeps the blocks its destroyed subjects
 * freed and hands them to later subjects (no counterpart in the reference;
 * for callers that estimate many instances in a row).  release_cache frees
 * the idle blocks; cached_bytes reports their size. */
int sss_ctx_release_cache(sss_ctx* ctx);
int sss_ctx_cached_bytes(sss_ctx* ctx, int64_t* bytes);
/* Free and total device memory of the context's GPU (hipMemGetInfo), for
 * callers that admit work by its device-memory footprint. */
int sss_ctx_memory(sss_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes);
/* Pinned host buffers of a context, kept for reuse after sss_host_free
 * (hipHostMalloc/hipHostFree synchronise the device). */
int sss_host_alloc(sss_ctx* ctx, int64_t bytes, void** out);
int sss_host_free(sss_ctx* ctx, void* p);

/* Trace n_segs segments (6 doubles each: x1 y1 z1 x2 y2 z2) through the grid.
 * Outputs have cap = X+Y+Z slots per segment: lengths (f64), 1-based voxel
 * indices (int32, X + (Y-1)*Ny + (Z-1)*Ny*Nx) and TOF sample positions (f64,
 * s_k = s_{k-1} + L_k/2).  counts[i] = voxels of segment i (0 if it misses).
 * Host pointers. */
int sss_trace_segments(sss_ctx* ctx, const int32_t grid[3], const double bounds[6],
                       const double* segs, int64_t n_segs, int32_t* counts,
                       double* lengths, int32_t* indices, double* samples);

/* Per segment: voxel count, sum(mu L) and sum(lam L) (f64).  Host pointers.
 * lam may be NULL. */
int sss_line_integrals(sss_ctx* ctx, const int32_t grid[3], const double bounds[6],
                       const double* mu, const double* lam, const double* segs,
                       int64_t n_segs, int32_t* counts, double* mu_int, double* lam_int);

int sss_subject_create(sss_ctx* ctx, const sss_subject_desc* desc, sss_subject** out);
int sss_subject_destroy(sss_subject* s);
/* Trace every scatter-point -> detector half path and build the resident
 * table (voxel ids, lengths, TOF sample times, attenuation, angles). */
int sss_subject_trace(sss_subject* s);
/* sss_subject_trace in two halves, so a caller with several subjects on
 * several contexts waits once for all of them: begin queues the walk counts
 * (asynchronous); end waits for this subject's counts, sizes the sample
 * tables and queues the walks (asynchronous after that wait).  When end
 * returns SSS_ERR_NOMEM the trace stays begun: free device memory (other
 * subjects, sss_ctx_release_cache) and call end again. */
int sss_subject_trace_begin(sss_subject* s);
int sss_subject_trace_end(sss_subject* s);
/* Upload lam (X*Y*Z f64, host, in the subject's volume_order) and refresh
 * the per-sample emission terms. */
int sss_subject_set_activity(sss_subject* s, const double* lam);
/* Same from a device pointer (X*Y*Z f64, volume_order), no host copy. */
int sss_subject_set_activity_device(sss_subject* s, const void* lam_dev);
/* Scatter sum into the subject's device output (compact layout, float32
 * [B][NRU^2][NDU(NDU+1)/2], cell (i1 <= i2) at i1*NDU - i1(i1-1)/2 + i2 - i1). */
int sss_subject_scatter(sss_subject* s);
/* Trace (if needed) + scatter in one call: the plain one-instance path. */
int sss_subject_run(sss_subject* s);
int sss_subject_info_get(sss_subject* s, sss_subject_info* out);
/* Re-tiles the subject's scatter launch without redoing its walks:
 * tile_rings and n_chunks as in sss_subject_desc (0: the defaults).  The
 * tile shape does not change any output value; n_chunks changes the
 * grouping in which a slice's chunk sums are added (FP32 rounding only).
 * For tile-shape tuners (no counterpart in the reference). */
int sss_subject_set_tiling(sss_subject* s, int tile_rings, int n_chunks);
/* Pairs passing the reference's filters (A != 0, some activity, efficiency
 * != 0; SingleScatterSimulationTOF.py:451-509) and the voxel samples of their
 * half paths with activity (n1 [Act1 != 0] + n2 [Act2 != 0], every voxel as
 * the reference's fold visits them; the fold here skips the zero terms), for
 * the algorithmic work count.  Synchronous. */
int sss_subject_work_stats(sss_subject* s, int64_t* pairs, int64_t* samples);
/* Duration (HIP events on the context stream) of the last scatter kernel. */
int sss_subject_scatter_kernel_ms(sss_subject* s, float* ms);
/* Device pointer of the compact output (valid until destroy). */
const void* sss_subject_output_device(const sss_subject* s);
int sss_subject_output_compact(sss_subject* s, float* host_out);
/* The compact download queued on the context's stream (host_out should come
 * from sss_host_alloc for the copy to run asynchronously); sss_subject_wait
 * waits for it and for the work queued before it on this subject.  For
 * callers that keep several subjects in flight (no counterpart in the
 * reference). */
int sss_subject_output_compact_async(sss_subject* s, float* host_out);
int sss_subject_wait(sss_subject* s);
/* done = 1 when the last queued download (and the work before it) finished. */
int sss_subject_query(sss_subject* s, int32_t* done);
/* Dense reference layout (B, N+1, N/2, NRU^2) float32 on the host, using
 * lor_coords (N, N, 2) int16 = (angular, radial). */
int sss_subject_output_dense(sss_subject* s, const int16_t* lor_coords, float* host_out);

/* Batches: several traced subjects of one context driven as one unit, for
 * the repeated estimates of reconstruction (a new activity image per
 * iteration for every subject).  No counterpart in the reference, where each
 * SingleScatterSimulationTOF call (SingleScatterSimulationTOF.py:36-208) is
 * one instance that is traced and summed anew.
 *
 * One batch launch queues, in order: every member's staged activity upload
 * and refresh; one persistent scatter launch per kernel variant present over
 * all members of that variant (their workgroup blocks in 8 per-XCD work
 * queues, the members with the most work per block first); one chunk-reduce
 * launch per precision; every member's compact output download into pinned
 * host memory.  sss_batch_graph_launch captures that sequence in a HIP graph
 * on its first call (and again after a member's walks were redone) and
 * replays the graph; sss_batch_submit queues the same sequence directly.
 * Each member's output equals its own sss_subject_scatter output bit for bit.
 *
 * Members: traced, with a host activity image (sss_subject_set_activity),
 * all on `ctx`; they must outlive the batch (sss_subject_destroy fails for a
 * member of a live batch), and while a launch is in flight only the batch
 * functions may use them. */
typedef struct sss_batch sss_batch;
int sss_batch_create(sss_ctx* ctx, sss_subject* const* members, int n_members, sss_batch** out);
/* Member i's next activity image (host, X*Y*Z f64 in its volume_order),
 * copied into its pinned staging before the call returns (after the batch's
 * previous launch has read the last one). */
int sss_batch_set_activity(sss_batch* b, int i, const double* lam);
int sss_batch_submit(sss_batch* b);
int sss_batch_graph_launch(sss_batch* b);
/* Waits for the last launch. */
int sss_batch_wait(sss_batch* b);
/* Member i's compact output of the last launch (its layout and type as
 * sss_subject_output_compact; waits for the launch). */
int sss_batch_output(sss_batch* b, int i, void* host_out);
/* Duration of the last launch (HIP events around it on the context stream). */
int sss_batch_elapsed_ms(sss_batch* b, float* ms);
/* Scatter launches per batch launch (kernel variants present), and whether
 * a captured graph is current (1) or not (0). */
int sss_batch_info(sss_batch* b, int32_t* n_scatter_launches, int32_t* graph_ready);
int sss_batch_destroy(sss_batch* b);

/* Device buffers on the context's device (synchronous copies). */
int sss_buffer_alloc(sss_ctx* ctx, int64_t bytes, void** dev);
int sss_buffer_free(sss_ctx* ctx, void* dev);
int sss_buffer_upload(sss_ctx* ctx, void* dev, const void* host, int64_t bytes);
int sss_buffer_download(sss_ctx* ctx, void* host, const void* dev, int64_t bytes);

/* Sampled -> full sinogram (Python/openSSS/Interpolation.py:95-399), device
 * pointers, asynchronous on the context stream.  sampled is the compact
 * [B][n_sampled_planes][cells] estimate; cell_map[c] (c = radial * n_ang +
 * angular of the output) is the compact cell of the nearest sampled detector
 * pair, plane_map[s] the sampled plane of the nearest sampled ring pair of
 * output plane s (-1: zero).  Writes out[b][c][s] = max(value, 0) (may be
 * NULL) and sum[c][s] = sum over bins (may be NULL). */
int sss_interpolate(sss_ctx* ctx, const float* sampled, int n_bins, int n_sampled_planes,
                    int64_t cells_per_sampled_plane, const int32_t* cell_map, int64_t n_cells_out,
                    const int32_t* plane_map, int n_out_planes, float* out, float* sum);

/* Tail mask (Python/openSSS/TailsMask.py:84-163 no span, :230-309 span), host
 * pointers.  For output plane p the LORs run from detector d1 of geometry ring
 * plane_rings[2p] to d2 of ring plane_rings[2p+1], d1 <= d2 on the
 * accel-strided detector grid; mask[(radial * N/2 + angular) * n_planes + p]
 * = 1 where 1/exp(-sum mu L) == 1 in float64 (use_attenuation) or where
 * sum lam L <= 0, else 0; cells no LOR reaches stay 0. */
int sss_tail_mask(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], const double* image,
                  int use_attenuation, int n_rings_geom, int n_det, const double* geom,
                  const int16_t* lor_coords, int n_planes, const int32_t* plane_rings, int accel,
                  uint8_t* mask_out);

/* Backprojection tail fit (Python/openSSS/BackProjection.py:31-392), host
 * pointers.  Entry e traces d1 of ring plane_rings[3e] to d2 of ring
 * plane_rings[3e+1] and reads its values from plane plane_rings[3e+2] of the
 * (N+1, N/2, n_value_planes) sinograms; LORs whose masked scatter is 0 are
 * skipped.  Values are spread value/n per voxel, or value*L/sum(L) when
 * length_weighted, into float64 images (prompts, scatter, randoms,
 * sensitivity).  sums_out = (sum S_bp T, sum S_bp^2) over voxels with
 * S_bp = scatter/sens, T = max(prompts/sens - randoms/sens, 0); the fitted
 * scale is sums_out[0] / sums_out[1].  images_out (4 * X*Y*Z) may be NULL. */
int sss_tail_fit(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], int n_rings_geom, int n_det,
                 const double* geom, const int16_t* lor_coords, int n_entries, const int32_t* plane_rings,
                 int n_value_planes, int accel, int length_weighted, const float* scatter,
                 const float* prompts, const float* randoms, const float* mask, double* sums_out,
                 double* images_out);

/* The scaling loop over every TOF bin (Python/openSSS/Scaling.py:60-113:
 * Prompts[..., b] / LORCounts, Randoms * Normalization, float32(mask >= 1),
 * then BackProjectionScatterRandFit or ...Mashing per bin) in one call.
 * Geometry, entries and weighting are those of sss_tail_fit.  With cells
 * c = (radial * N/2 + angular) * n_value_planes + plane:
 *   scatter     float32 [n_bins][cells]: the per-bin interpolated sinograms
 *   prompts     [cells][n_bins], uint16 (SSS_PROMPTS_U16, what the list-mode
 *               histogrammer writes) or float32 (SSS_PROMPTS_F32)
 *   lor_counts  float32 divisor of the prompts: NULL (none), one per plane
 *               (lor_counts_per_plane != 0; SumLORsCounts is constant over a
 *               plane's cells) or one per cell
 *   randoms     float32 [cells] or NULL (0); norm float32 [cells] or NULL (1)
 *   mask        uint8 [cells]; cells with mask >= 1 are fitted
 * sums_out[2b], sums_out[2b+1] are bin b's two sums, the same bits as
 * sss_tail_fit on that bin's sinograms with the prompts divided in float32
 * beforehand.  Each LOR is traced once per group of 8 bins.  flags
 * SSS_FIT_DEVICE: every array pointer (geom, lor_coords, plane_rings, the
 * sinograms, sums_out) is device memory on the context's device and the call
 * is asynchronous on the context's stream (entries out of range are skipped);
 * otherwise they are host pointers and the call returns when sums_out is
 * written.  Scratch: 32 * n_bins * X*Y*Z bytes from the context's cache. */
#define SSS_PROMPTS_F32 0
#define SSS_PROMPTS_U16 1
#define SSS_FIT_DEVICE 1
#define SSS_MAX_FIT_BINS 1024
int sss_scale_to_prompts(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], int n_rings_geom, int n_det,
                         const double* geom, const int16_t* lor_coords, int n_entries, const int32_t* plane_rings,
                         int n_value_planes, int accel, int length_weighted, int n_bins, const float* scatter,
                         const void* prompts, int prompts_type, const float* lor_counts, int lor_counts_per_plane,
                         const float* randoms, const float* norm, const uint8_t* mask, int flags,
                         double* sums_out);

/* CASToR list-mode events (Python/Timber).  An event record is `words`
 * 4-byte words (Timber/Coincidence.py:45-95): the time (uint32, word 0), the
 * float corrections, the optional line count, then the crystal ID pairs; the
 * first pair is at id_word and the TOF (float32, ps) at tof_word.  An event's
 * crystal pair is parsed as ParseCASToRIDs (Coincidence.py:97-159) with
 * crystals_per_ring (the coordinates' N times mash) and shift, and binned to
 * (radial, angular, plane) = (lor[d1/mash][d2/mash][1], [..][0],
 * ring_index[ring1][ring2]) of an (n_rad, n_ang, n_planes) sinogram; -1
 * entries drop the event.  TOF bin int((TOF + R) / (2R / B)) for |TOF| <= R.
 * Event counts are below 2^32 - 1.  Host pointers, or device pointers on the
 * context's device with SSS_LM_DEVICE; both calls return when done.  A
 * detector, ring or coordinate outside the tables (the reference's
 * IndexError) is SSS_ERR_INVALID.
 *
 * sss_listmode_histogram: Timber/Datafile.py:29-115 (ExportPrompts): prompts
 * (n_rad, n_ang, n_planes, n_bins) uint16 counts (wrapping at 2^16, as the
 * reference's uint16 sinogram); with SSS_LM_RANDOMS, randoms (n_rad, n_ang,
 * n_planes) float32 = word 0 * duration / n_bins of the cell's last event
 * (0 where none); with SSS_LM_NORM, norm = word 1 of the cell's last event as
 * float32 (1 where none).  Every binned event sets randoms and norm, also one
 * outside the TOF range.
 * sss_listmode_inject: Timber/Datafile.py:117-240 (InjectScatters): the
 * events with |TOF| <= R, a cell and a TOF bin b, in file order, each copied
 * with the float32 rate factor[b] * scatter[b][cell] inserted at word
 * scatter_word (later words move up one); events_out holds up to n_events
 * records of words + 1 words, *n_out receives the count. */
typedef struct {
    int32_t words;         /* 4-byte words per event record */
    int32_t tof_word;      /* word of the TOF (float32, ps) */
    int32_t id_word;       /* word of the first crystal ID (uint32) */
    int32_t scatter_word;  /* sss_listmode_inject: word the scatter rate is written at */
} sss_listmode_layout;
#define SSS_LM_RANDOMS 1
#define SSS_LM_NORM 2
#define SSS_LM_DEVICE 4
int sss_listmode_histogram(sss_ctx* ctx, const void* events, int64_t n_events, const sss_listmode_layout* layout,
                           int64_t crystals_per_ring, int64_t shift, int mash, int n_det_lor,
                           const int16_t* lor_coords, int n_rings_index, const int16_t* ring_index, int n_rad,
                           int n_ang, int n_planes, int n_bins, double tof_range, double duration, int flags,
                           uint16_t* prompts_out, float* randoms_out, float* norm_out);
int sss_listmode_inject(sss_ctx* ctx, const void* events, int64_t n_events, const sss_listmode_layout* layout,
                        int64_t crystals_per_ring, int64_t shift, int mash, int n_det_lor,
                        const int16_t* lor_coords, int n_rings_index, const int16_t* ring_index, int n_rad,
                        int n_ang, int n_planes, int n_bins, double tof_range, const float* factor,
                        const float* scatter, int flags, void* events_out, int64_t* n_out);

/* Image preparation (Python/openSSS/Utils.py:212-267, CropAndDownscale, and
 * the attenuation floor of PrepareImages :72), host pointers, volumes
 * x-fastest (the Fortran order of (X, Y, Z)).  The cropped volume's voxel c
 * is input voxel c + crop_start (0 outside the input: the reference's crop
 * window or zero pad, resolved by the caller).  It is then resized as
 * scikit-image's resize(order=3, mode='edge', anti_aliasing=True,
 * preserve_range=True) states it: along each axis with radius[k] >= 0 a
 * Gaussian of the 2 radius[k] + 1 taps weights[k] (edge values repeated,
 * float64 sums stored float32); cubic B-spline coefficients of the result
 * padded by 12 repeated edge voxels (float64); at each output voxel the
 * spline's value at (o + 1/2) crop/out - 1/2; with clip, the output clipped
 * to the cropped volume's range; values below zero_below set to 0. */
typedef struct {
    int64_t in_shape[3];
    int64_t crop_start[3];
    int64_t crop_shape[3];
    int64_t out_shape[3];
    int32_t radius[3];          /* Gaussian half width per axis, -1: no smoothing along it */
    const double* weights[3];   /* host, 2 radius + 1 taps */
    int32_t clip;
    float zero_below;
} sss_resize_desc;
int sss_crop_and_downscale(sss_ctx* ctx, const float* image, const sss_resize_desc* desc, float* out);

/* Synthetic prompts (no counterpart in the reference, which reads measured
 * prompts; the per-LOR trace follows Python/openSSS/TailsMask.py:139-145).
 * sss_forward_project: for output plane p, the LORs from detector d1 of ring
 * plane_rings[2p] to d2 of ring plane_rings[2p+1], d1 <= d2 on the
 * accel-strided grid (as sss_tail_mask), get trues = sum(lam L) exp(-sum(mu L))
 * (float64, stored float32) at [(radial * N/2 + angular) * n_planes + p];
 * mu may be NULL (no attenuation); other cells are 0.  Host pointers.
 * sss_poisson: out[i] = a Poisson draw of mean mean[i] * scale, from a
 * counter-based generator keyed by (seed, i): the same seed gives the same
 * counts.  Host pointers. */
int sss_forward_project(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], const double* lam,
                        const double* mu, int n_rings_geom, int n_det, const double* geom, const int16_t* lor_coords,
                        int n_planes, const int32_t* plane_rings, int accel, float* trues_out);
int sss_poisson(sss_ctx* ctx, const float* mean, int64_t n, double scale, uint64_t seed, float* out);

/* sss_tail_fit on device buffers, allocating nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc).  Every
 * pointer except grid and bounds (host) is device memory on the context's
 * device: geom, lor_coords, plane_rings, the four sinograms as in
 * sss_tail_fit; sums_out (2 doubles) and images_out (4 * X*Y*Z doubles, may
 * be NULL) receive what sss_tail_fit returns, bit for bit.  Asynchronous on
 * the context's stream.  Entries or coordinates out of range are skipped.
 * The fit's scratch (fixed-point images, reduction partials) is a static
 * device array of the library, one per GPU, sized for fit grids of at most
 * SSS_TAIL_FIT_NOALLOC_MAX_VOXELS voxels (larger grids: SSS_ERR_UNSUPPORTED).
 * Calls that overlap on one GPU (from several contexts or threads) share
 * that array and race; each call clears it before use. */
#define SSS_TAIL_FIT_NOALLOC_MAX_VOXELS (1 << 20)
int sss_tail_fit_noalloc(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], int n_rings_geom,
                         int n_det, const double* geom, const int16_t* lor_coords, int n_entries,
                         const int32_t* plane_rings, int n_value_planes, int accel, int length_weighted,
                         const float* scatter, const float* prompts, const float* randoms, const float* mask,
                         double* sums_out, double* images_out);

#ifdef __cplusplus
}
#endif

#endif /* SSS_HIP_H */
