// C ABI, contexts and resident scatter subjects of libsss_hip.so.
//
// A subject holds, in HBM, everything one scatter estimate reuses across
// calls: the mu-map, scatter points, sampled-detector positions, the
// half-path table built by the tracer (voxel ids, lengths, TOF sample
// times, attenuation integrals, detector angles) and the output buffers.
// Only the activity image changes per request.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include <algorithm>
#include <map>
#include <type_traits>
#include <unordered_map>
#include <climits>
#include <cmath>
#include <new>
#include <string>
#include <vector>

#include "sss_internal.h"

namespace sss {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int fail(int code, const std::string& msg)
{
    g_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char* what)
{
    g_error = std::string(what) + ": " + hipGetErrorString(e);
    // a caller can free memory and retry an out-of-memory failure
    return e == hipErrorOutOfMemory ? SSS_ERR_NOMEM : SSS_ERR_HIP;
}

namespace {

// Number of half paths whose voxel walk was truncated (status 1).
__global__ void k_count_status(const int32_t* __restrict__ st, int64_t n, int32_t* out)
{
    int32_t m = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m += st[i];
    if (m) atomicAdd(out, m);
}

// Records a half path occupies in the sample table.
struct PaddedRecords {
    __host__ __device__ int64_t operator()(int32_t n) const
    {
        return (int64_t)((n + kSampleGroup - 1) / kSampleGroup) * kSampleGroup;
    }
};

template <typename T>
int dev_alloc(T** p, size_t count, int64_t* tally)
{
    *p = nullptr;
    if (count == 0) return SSS_OK;
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(SSS_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    }
    if (tally) *tally += (int64_t)(count * sizeof(T));
    return SSS_OK;
}

template <typename T>
void dev_free(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

}  // namespace

double tof_offset_bound(const Grid& g, const double* sp_xyz, int n_sp, const double* detpos,
                        int n_det_pos)
{
    const double dx = g.b[3] - g.b[0], dy = g.b[4] - g.b[1], dz = g.b[5] - g.b[2];
    const double qmax = std::sqrt(dx * dx + dy * dy + dz * dz) / kLightCmPerPs;
    double dd = 0;
    for (int p = 0; p < n_sp; ++p) {
        // squared distances: sqrt is monotonic and correctly rounded, so the
        // extremes of r are the roots of the extremes of r^2
        double lo = HUGE_VAL, hi = 0;
        for (int i = 0; i < n_det_pos; ++i) {
            const double ex = detpos[3 * i + 0] - sp_xyz[3 * p + 0];
            const double ey = detpos[3 * i + 1] - sp_xyz[3 * p + 1];
            const double ez = detpos[3 * i + 2] - sp_xyz[3 * p + 2];
            const double r2 = ex * ex + ey * ey + ez * ez;
            lo = std::min(lo, r2);
            hi = std::max(hi, r2);
        }
        dd = std::max(dd, std::sqrt(hi) - std::sqrt(lo));
    }
    return dd / kLightCmPerPs + qmax + 1.0;
}

// Wave work items of the scatter launch.  Workgroups go to the 8 XCDs of an
// MI355X round-robin by workgroup id, and each XCD has its own L2; a plane
// (r1, r2) reads the half paths of rings r1 and r2, so each XCD is given the
// planes of one pair of ring-group blocks {G_i x G_j, G_j x G_i} (rings in 4
// groups; the diagonal blocks pair up as {G0 x G0, G1 x G1}, {G2 x G2, G3 x G3})
// and reads half the rings.  That split is even only when the ring blocks
// fall into 4 equal groups (NRU = 16 with 4-ring blocks); otherwise (NRU = 8
// puts all tiles on 3 XCDs, one of them with half the work) the tiles, in
// their order (ring-block pair, then detector blocks), are cut into 8
// contiguous runs of equal size, so each XCD still reads few ring blocks.
// Within an XCD the items run chunk-major, so the waves resident at one time
// share one chunk's scatter points.
std::vector<int2> xcd_work_items(const std::vector<int4>& tiles, int nru, int n_chunks,
                                 int waves_per_block)
{
    constexpr int kXcd = 8;
    auto group = [nru](int r) { return std::min(3, r * 4 / std::max(1, nru)); };
    auto xcd_of = [](int g1, int g2) {
        if (g1 == g2) return 6 + g1 / 2;
        const int i = std::min(g1, g2), j = std::max(g1, g2);
        static const int pair_id[4][4] = {{-1, 0, 1, 2}, {0, -1, 3, 4}, {1, 3, -1, 5}, {2, 4, 5, -1}};
        return pair_id[i][j];
    };
    std::vector<std::vector<int>> lists(kXcd);
    for (size_t t = 0; t < tiles.size(); ++t)
        lists[xcd_of(group(tiles[t].x), group(tiles[t].y))].push_back((int)t);
    const size_t even = (tiles.size() + kXcd - 1) / kXcd;
    size_t largest = 0;
    for (const auto& l : lists) largest = std::max(largest, l.size());
    if (largest * 8 > even * 9) {   // more than 12.5% above an even share
        for (auto& l : lists) l.clear();
        for (size_t t = 0; t < tiles.size(); ++t) lists[t * kXcd / tiles.size()].push_back((int)t);
    }
    size_t per_xcd = 0;
    for (const auto& l : lists) per_xcd = std::max(per_xcd, l.size() * (size_t)n_chunks);
    const size_t blocks_per_xcd = (per_xcd + waves_per_block - 1) / waves_per_block;
    std::vector<int2> items(blocks_per_xcd * kXcd * waves_per_block, make_int2(-1, 0));
    for (int x = 0; x < kXcd; ++x) {
        size_t k = 0;
        for (int c = 0; c < n_chunks; ++c)
            for (int t : lists[x]) {
                const size_t blk = (k / waves_per_block) * kXcd + x;   // workgroup id on XCD x
                items[blk * waves_per_block + k % waves_per_block] = make_int2(t, c);
                ++k;
            }
    }
    return items;
}

}  // namespace sss

using namespace sss;

namespace sss {
namespace {

// Device memory cache of one context.  Blocks freed by its subjects are kept
// for later subjects: hipMalloc of the GB-sized tables costs milliseconds and
// hipFree synchronises the device.  Every user of a context's blocks works on
// the context's stream, so a block's next use is ordered after its last one.
struct DevicePool {
    std::multimap<size_t, void*> idle;          // size -> block
    std::unordered_map<void*, size_t> owned;    // every block of the pool
    size_t idle_bytes = 0;

    static size_t rounded(size_t b)
    {
        const size_t q = b < (size_t(1) << 20) ? 256 : (size_t(2) << 20);
        return (b + q - 1) / q * q;
    }
    hipError_t get(size_t bytes, void** p)
    {
        const size_t b = rounded(bytes);
        auto it = idle.lower_bound(b);
        if (it != idle.end() && it->first <= b + b / 4 + (size_t(4) << 20)) {
            *p = it->second;
            idle_bytes -= it->first;
            idle.erase(it);
            return hipSuccess;
        }
        hipError_t e = hipMalloc(p, b);
        if (e != hipSuccess && !idle.empty()) {   // out of memory: drop the cache, try once more
            (void)hipGetLastError();
            release_idle();
            e = hipMalloc(p, b);
        }
        if (e == hipSuccess) owned[*p] = b;
        return e;
    }
    void put(void* p)
    {
        auto it = owned.find(p);
        if (it == owned.end()) return;
        idle.emplace(it->second, p);
        idle_bytes += it->second;
    }
    void release_idle()
    {
        for (auto& kv : idle) {
            (void)hipFree(kv.second);
            owned.erase(kv.second);
        }
        idle.clear();
        idle_bytes = 0;
    }
};

}  // namespace
}  // namespace sss

struct sss_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    sss::DevicePool pool;
    // pinned host words for the subjects' sample counts (hipHostMalloc and
    // hipHostFree synchronise the device, so they are not done per subject)
    int64_t* pinned = nullptr;
    std::vector<int> pinned_free;
    // pinned host buffers handed out by sss_host_alloc, kept for reuse
    std::multimap<size_t, void*> host_idle;
    std::unordered_map<void*, size_t> host_owned;
    // events of destroyed subjects, kept for later subjects ([0] timing,
    // [1] hipEventDisableTiming)
    std::vector<hipEvent_t> events[2];
};

int sss::ctx_device(const sss_ctx* ctx) { return ctx->device; }
hipStream_t sss::ctx_stream(const sss_ctx* ctx) { return ctx->stream; }

// An event from the context's cache, or a new one.
static hipError_t ctx_event(sss_ctx* c, bool timing, hipEvent_t* e)
{
    auto& v = c->events[timing ? 0 : 1];
    if (!v.empty()) {
        *e = v.back();
        v.pop_back();
        return hipSuccess;
    }
    return timing ? hipEventCreate(e) : hipEventCreateWithFlags(e, hipEventDisableTiming);
}

static void ctx_event_put(sss_ctx* c, bool timing, hipEvent_t& e)
{
    if (e) c->events[timing ? 0 : 1].push_back(e);
    e = nullptr;
}

static constexpr int kPinnedSlots = 4096;

namespace sss {
namespace {

template <typename T>
int pool_alloc(sss_ctx* c, T** p, size_t count, int64_t* tally)
{
    *p = nullptr;
    if (count == 0) return SSS_OK;
    void* v = nullptr;
    hipError_t e = c->pool.get(count * sizeof(T), &v);
    if (e != hipSuccess) return fail(SSS_ERR_NOMEM, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    *p = static_cast<T*>(v);
    if (tally) *tally += (int64_t)(count * sizeof(T));
    return SSS_OK;
}

template <typename T>
void pool_free(sss_ctx* c, T*& p)
{
    if (p) c->pool.put(static_cast<void*>(const_cast<std::remove_const_t<T>*>(p)));
    p = nullptr;
}

// Pinned host blocks of a context, kept for reuse (hipHostMalloc/hipHostFree
// synchronise the device).
int host_block(sss_ctx* c, size_t bytes, void** out)
{
    const size_t b = std::max<size_t>(4096, (bytes + 4095) / 4096 * 4096);
    auto it = c->host_idle.lower_bound(b);
    if (it != c->host_idle.end() && it->first <= 2 * b) {
        *out = it->second;
        c->host_idle.erase(it);
        return SSS_OK;
    }
    SSS_HIP(hipSetDevice(c->device));
    void* p = nullptr;
    SSS_HIP(hipHostMalloc(&p, b, hipHostMallocDefault));
    c->host_owned[p] = b;
    *out = p;
    return SSS_OK;
}

void host_block_free(sss_ctx* c, void* p)
{
    auto it = c->host_owned.find(p);
    if (it != c->host_owned.end()) c->host_idle.emplace(it->second, p);
}

}  // namespace

int ctx_pool_get(sss_ctx* ctx, size_t bytes, void** p)
{
    unsigned char* b = nullptr;
    const int rc = pool_alloc(ctx, &b, bytes > 0 ? bytes : 1, nullptr);
    *p = b;
    return rc;
}

void ctx_pool_put(sss_ctx* ctx, void* p)
{
    if (p) ctx->pool.put(p);
}

}  // namespace sss

struct sss_subject {
    sss_ctx* ctx = nullptr;
    Grid grid{};
    int n_rings = 0, n_det = 0, nru = 0, ndu = 0, n_sp = 0, n_slices = 0;
    int n_bins = 0, n_energy = 0, eff_ref_row = 0, precision = 0;
    bool trivial = false;
    int half_window = 0, t_lo = 0, t_hi = 0;
    bool range_check = true;          // TOF offsets may leave the non-zero table columns
    double tof_bound = 0;             // largest |TOF offset| any sample can reach, ps
    double bin_w = 0, kappa = 0, beta = 0, g_const = 0;
    int64_t n_vox = 0, n_hp = 0, n_samples = 0, samples_cap = 0, n_cells_plane = 0, out_elems = 0;
    int64_t n_cells = 0;                  // cells of all planes
    int n_tiles = 0, n_chunks = 1;
    int tile_db_log2 = 3;                 // log2 detectors per tile row/column block
    int64_t device_bytes = 0;
    bool traced = false, has_activity = false;

    double* d_mu = nullptr;
    double* d_lam = nullptr;
    double* d_sp = nullptr;
    double* d_detpos = nullptr;
    double* d_detnorm = nullptr;
    unsigned char* d_upload = nullptr;    // one block holding every array uploaded at create
    bool z_fastest = false;               // volumes arrive in C order of (X, Y, Z)
    double* d_vol_in = nullptr;           // z fastest: mu's upload slot, later each activity upload
    double* d_mu_own = nullptr;           // z fastest: mu, x fastest
    double* d_eff = nullptr;
    float* d_ratio = nullptr;
    float2* d_etab = nullptr;             // kEnergyRows energy terms of the one-bin scatter
    int etab_lo = 0, etab_hi = 0;
    int4* d_tiles = nullptr;
    int2* d_items = nullptr;          // scatter launch: (tile, chunk) per wave, XCD-aware order
    int n_items = 0;
    int32_t* d_slice_ptr = nullptr;
    int32_t* d_chunk_sp = nullptr;        // n_chunks + 1 scatter-point bounds
    int32_t* d_chunk_slice = nullptr;     // slice of each chunk
    int32_t* d_dets = nullptr;
    int32_t* d_flag = nullptr;
    int32_t* d_fold_changed = nullptr;    // set when a new activity changes which terms are zero
    // voxel test of a new activity image: [0] failed, [1] the table's image
    // is unfit for it, [2] smallest positive sample length (float bits)
    int32_t* d_fast = nullptr;
    uint8_t* d_tab = nullptr;             // non-zero voxels of the fold table's image
    int64_t fold_cap = 0;                 // fold records allocated
    bool fold_valid = false;              // the fold table's layout matches the current tables
    int refreshes = 0;                    // activity images since the last walks
    void* d_scan_tmp = nullptr;
    size_t scan_tmp_bytes = 0;
    HalfPathTable hp{};
    SampleTable smp{};
    void* d_out = nullptr;
    void* d_part_sum = nullptr;           // [chunk][B][slot], slot = tile * 64 + lane
    float* d_part_cnt = nullptr;          // [chunk][slot]
    int64_t n_slots = 0;
    std::vector<int32_t> dets;
    std::vector<int32_t> slice_ptr_h;     // n_slices + 1 scatter-point bounds (host)
    unsigned char* d_tiling = nullptr;    // tiles, items and chunks after sss_subject_set_tiling
    size_t part_sum_cap = 0, part_cnt_cap = 0;   // elements allocated for the chunk partials
    // pinned host sources of the queued uploads (kept until destroy)
    void* h_upload = nullptr;
    double* h_lam = nullptr;
    hipEvent_t ev_lam = nullptr;          // after the last activity upload
    bool lam_queued = false;
    hipEvent_t ev_upload = nullptr;       // after the uploads queued at create
    hipEvent_t ev_scatter[2] = {nullptr, nullptr};   // around the last scatter kernel
    bool ev_recorded = false;
    // two-phase trace: the sample count reaches the host through pinned memory
    int64_t* h_total = nullptr;
    hipEvent_t ev_count = nullptr;
    hipEvent_t ev_out = nullptr;          // after the last output download
    bool trace_begun = false;
    int generation = 0;                   // walks done (a batch holding the subject rebuilds on change)
    int batch_refs = 0;                   // live batches holding the subject

    ~sss_subject()
    {
        for (auto& e : ev_scatter) ctx_event_put(ctx, true, e);
        if (h_total) ctx->pinned_free.push_back((int)(h_total - ctx->pinned));
        if (h_upload) host_block_free(ctx, h_upload);
        if (h_lam) host_block_free(ctx, h_lam);
        ctx_event_put(ctx, false, ev_count);
        ctx_event_put(ctx, false, ev_out);
        ctx_event_put(ctx, false, ev_lam);
        ctx_event_put(ctx, false, ev_upload);
        sss_ctx* c = ctx;
        pool_free(c, d_upload);   // d_mu (x fastest) .. d_dets live in it
        pool_free(c, d_mu_own);
        pool_free(c, d_lam); pool_free(c, d_flag); pool_free(c, d_fold_changed); pool_free(c, d_scan_tmp);
        pool_free(c, d_fast); pool_free(c, d_tab);
        pool_free(c, hp.off); pool_free(c, hp.n); pool_free(c, hp.off_fold); pool_free(c, hp.n_fold);
        pool_free(c, hp.rec); pool_free(c, hp.blk); pool_free(c, hp.blk_off); pool_free(c, hp.dist);
        pool_free(c, hp.att); pool_free(c, hp.lnatt); pool_free(c, hp.act); pool_free(c, hp.cosabs);
        pool_free(c, hp.status);
        pool_free(c, smp.nq); pool_free(c, smp.vox); pool_free(c, smp.len); pool_free(c, smp.len32); pool_free(c, smp.a);
        pool_free(c, smp.a64); pool_free(c, smp.fold); pool_free(c, smp.fold_a64); pool_free(c, smp.fold_map);
        pool_free(c, d_out); pool_free(c, d_part_sum); pool_free(c, d_part_cnt); pool_free(c, d_tiling);
    }
};

namespace sss {
namespace {

// Scatter-launch layout of a subject: wave tiles, scatter-point chunks and
// the XCD-aware work items.
struct Tiling {
    std::vector<int4> tiles;
    std::vector<int32_t> chunk_sp, chunk_slice;
    std::vector<int2> items;
    int tile_db_log2 = 0;
};

}  // namespace
}  // namespace sss

static void make_tiling(const sss_subject* s, int tile_rings, int n_chunks, Tiling* t)
{
    // wave tiles: rows and columns are blocks of rb rings x db detectors
    // (rb * db = 8, rb the largest power of two <= min(4, NRU) by default);
    // detector blocks cover the upper triangle i1 <= i2
    int rb = 1;
    if (tile_rings > 0) {
        rb = tile_rings;
    } else {
        while (rb * 2 <= std::min(kTileMaxRings, s->nru)) rb *= 2;
    }
    const int db = kTileRows / rb;
    t->tile_db_log2 = 0;
    while ((1 << t->tile_db_log2) < db) ++t->tile_db_log2;
    t->tiles.clear();
    const int ndb = (s->ndu + db - 1) / db;
    for (int r1 = 0; r1 < s->nru; r1 += rb)
        for (int r2 = 0; r2 < s->nru; r2 += rb)
            for (int ti = 0; ti < ndb; ++ti)
                for (int tj = ti; tj < ndb; ++tj) t->tiles.push_back(make_int4(r1, r2, ti * db, tj * db));
    // scatter-point chunks: groups of at most G consecutive points of one
    // slice (n_chunks > 0: about n_chunks groups in all).  Every chunk costs
    // one partial-sum and one count array (at most 16 GiB in all).
    const size_t elem = s->precision == SSS_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    const int64_t n_slots = (int64_t)t->tiles.size() * kWave;
    const int64_t chunk_bytes = std::max<int64_t>(1, n_slots * (s->n_bins * (int64_t)elem + (int64_t)sizeof(float)));
    const int64_t max_chunks = std::max<int64_t>(1, (int64_t(16) << 30) / chunk_bytes);
    int64_t group = n_chunks > 0 ? (s->n_sp + n_chunks - 1) / n_chunks : kAutoChunkPoints;
    group = std::max<int64_t>({group, 1, (s->n_sp + max_chunks - 1) / max_chunks});
    t->chunk_sp.clear();
    t->chunk_slice.clear();
    const int n_slices = (int)s->slice_ptr_h.size() - 1;
    for (int sl = 0; sl < n_slices; ++sl) {
        const int64_t b = s->slice_ptr_h[sl], m = s->slice_ptr_h[sl + 1] - b;
        const int64_t k = (m + group - 1) / group;
        for (int64_t j = 0; j < k; ++j) {
            t->chunk_sp.push_back((int32_t)(b + j * m / k));
            t->chunk_slice.push_back(sl);
        }
    }
    t->chunk_sp.push_back(s->n_sp);
    t->items = xcd_work_items(t->tiles, s->nru, (int)t->chunk_slice.size(), kScatterWaves);
}

static void apply_tiling_sizes(sss_subject* s, const Tiling& t)
{
    s->tile_db_log2 = t.tile_db_log2;
    s->n_tiles = (int)t.tiles.size();
    s->n_slots = (int64_t)s->n_tiles * kWave;
    s->n_chunks = (int)t.chunk_slice.size();
    s->n_items = (int)t.items.size();
}

extern "C" {

const char* sss_last_error(void) { return g_error.c_str(); }

int sss_version(void) { return 1; }

int sss_ctx_create(int device, sss_ctx** out)
{
    if (!out) return fail(SSS_ERR_INVALID, "sss_ctx_create: out is NULL");
    *out = nullptr;
    int n = 0;
    SSS_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SSS_ERR_INVALID, "sss_ctx_create: no such device");
    SSS_HIP(hipSetDevice(device));
    sss_ctx* c = new (std::nothrow) sss_ctx;
    if (!c) return fail(SSS_ERR_NOMEM, "sss_ctx_create: out of host memory");
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess)
        e = hipHostMalloc(reinterpret_cast<void**>(&c->pinned), kPinnedSlots * sizeof(int64_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return hip_fail(e, "sss_ctx_create");
    }
    for (int i = kPinnedSlots - 1; i >= 0; --i) c->pinned_free.push_back(i);

    if (int rc = noalloc_scratch_init(device)) {
        (void)hipHostFree(c->pinned);
        (void)hipStreamDestroy(c->stream);
        delete c;
        return rc;
    }
    *out = c;
    return SSS_OK;
}

int sss_ctx_destroy(sss_ctx* ctx)
{
    if (!ctx) return SSS_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    ctx->pool.release_idle();
    for (auto& v : ctx->events)
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
    for (auto& kv : ctx->host_owned) (void)hipHostFree(kv.first);
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return SSS_OK;
}

int sss_ctx_release_cache(sss_ctx* ctx)
{
    if (!ctx) return fail(SSS_ERR_INVALID, "sss_ctx_release_cache: NULL context");
    SSS_HIP(hipSetDevice(ctx->device));
    SSS_HIP(hipStreamSynchronize(ctx->stream));
    ctx->pool.release_idle();
    return SSS_OK;
}

int sss_host_alloc(sss_ctx* ctx, int64_t bytes, void** out)
{
    if (!ctx || !out || bytes < 0) return fail(SSS_ERR_INVALID, "sss_host_alloc: bad argument");
    return host_block(ctx, (size_t)bytes, out);
}

int sss_host_free(sss_ctx* ctx, void* p)
{
    if (!ctx) return fail(SSS_ERR_INVALID, "sss_host_free: NULL context");
    if (!ctx->host_owned.count(p)) return fail(SSS_ERR_INVALID, "sss_host_free: not from sss_host_alloc");
    host_block_free(ctx, p);
    return SSS_OK;
}

int sss_ctx_cached_bytes(sss_ctx* ctx, int64_t* bytes)
{
    if (!ctx || !bytes) return fail(SSS_ERR_INVALID, "sss_ctx_cached_bytes: NULL argument");
    *bytes = (int64_t)ctx->pool.idle_bytes;
    return SSS_OK;
}

int sss_ctx_memory(sss_ctx* ctx, int64_t* free_bytes, int64_t* total_bytes)
{
    if (!ctx || !free_bytes || !total_bytes) return fail(SSS_ERR_INVALID, "sss_ctx_memory: NULL argument");
    SSS_HIP(hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    SSS_HIP(hipMemGetInfo(&f, &t));
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return SSS_OK;
}

int sss_ctx_synchronize(sss_ctx* ctx)
{
    if (!ctx) return fail(SSS_ERR_INVALID, "sss_ctx_synchronize: NULL context");
    SSS_HIP(hipSetDevice(ctx->device));
    SSS_HIP(hipStreamSynchronize(ctx->stream));
    return SSS_OK;
}

void* sss_ctx_stream(sss_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

static int check_grid(const int32_t grid[3], const double bounds[6], Grid* g)
{
    if (!grid || !bounds) return fail(SSS_ERR_INVALID, "grid/bounds is NULL");
    if (grid[0] <= 0 || grid[1] <= 0 || grid[2] <= 0)
        return fail(SSS_ERR_INVALID, "grid dimensions must be positive");
    if (grid[0] != grid[1])
        return fail(SSS_ERR_INVALID,
                    "grid must have Nx == Ny: the reference voxel index uses Ny as the y stride");
    if (!(bounds[3] > bounds[0] && bounds[4] > bounds[1] && bounds[5] > bounds[2]))
        return fail(SSS_ERR_INVALID, "bounds must satisfy max > min on every axis");
    g->nx = grid[0];
    g->ny = grid[1];
    g->nz = grid[2];
    for (int i = 0; i < 6; ++i) g->b[i] = bounds[i];
    return SSS_OK;
}

int sss_trace_segments(sss_ctx* ctx, const int32_t grid[3], const double bounds[6],
                       const double* segs, int64_t n_segs, int32_t* counts, double* lengths,
                       int32_t* indices, double* samples)
{
    if (!ctx) return fail(SSS_ERR_INVALID, "sss_trace_segments: NULL context");
    Grid g;
    int rc = check_grid(grid, bounds, &g);
    if (rc) return rc;
    if (n_segs < 0 || (n_segs > 0 && (!segs || !counts || !lengths || !indices || !samples)))
        return fail(SSS_ERR_INVALID, "sss_trace_segments: bad buffers");
    if (n_segs == 0) return SSS_OK;
    SSS_HIP(hipSetDevice(ctx->device));
    const int cap = g.nx + g.ny + g.nz;
    double *d_segs = nullptr, *d_L = nullptr, *d_s = nullptr;
    int32_t *d_c = nullptr, *d_i = nullptr;
    int64_t tally = 0;
    const size_t slots = (size_t)n_segs * cap;
    if ((rc = dev_alloc(&d_segs, (size_t)n_segs * 6, &tally)) ||
        (rc = dev_alloc(&d_L, slots, &tally)) || (rc = dev_alloc(&d_s, slots, &tally)) ||
        (rc = dev_alloc(&d_c, (size_t)n_segs, &tally)) || (rc = dev_alloc(&d_i, slots, &tally))) {
        dev_free(d_segs); dev_free(d_L); dev_free(d_s); dev_free(d_c); dev_free(d_i);
        return rc;
    }
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(d_segs, segs, (size_t)n_segs * 6 * sizeof(double),
                                  hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_trace_segments(g, d_segs, n_segs, cap, d_c, d_L, d_i, d_s, st);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, d_c, (size_t)n_segs * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(lengths, d_L, slots * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(indices, d_i, slots * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(samples, d_s, slots * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_segs); dev_free(d_L); dev_free(d_s); dev_free(d_c); dev_free(d_i);
    if (e != hipSuccess) return hip_fail(e, "sss_trace_segments");
    for (int64_t i = 0; i < n_segs; ++i)
        if (counts[i] < 0) return fail(SSS_ERR_TRACE, "sss_trace_segments: a voxel walk did not end at its end voxel");
    return SSS_OK;
}

int sss_line_integrals(sss_ctx* ctx, const int32_t grid[3], const double bounds[6],
                       const double* mu, const double* lam, const double* segs, int64_t n_segs,
                       int32_t* counts, double* mu_int, double* lam_int)
{
    if (!ctx) return fail(SSS_ERR_INVALID, "sss_line_integrals: NULL context");
    Grid g;
    int rc = check_grid(grid, bounds, &g);
    if (rc) return rc;
    if (!mu || n_segs < 0 || (n_segs > 0 && (!segs || !counts || !mu_int || !lam_int)))
        return fail(SSS_ERR_INVALID, "sss_line_integrals: bad buffers");
    if (n_segs == 0) return SSS_OK;
    SSS_HIP(hipSetDevice(ctx->device));
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    double *d_mu = nullptr, *d_lam = nullptr, *d_segs = nullptr, *d_m = nullptr, *d_l = nullptr;
    int32_t* d_c = nullptr;
    int64_t tally = 0;
    if ((rc = dev_alloc(&d_mu, nvox, &tally)) || (lam && (rc = dev_alloc(&d_lam, nvox, &tally))) ||
        (rc = dev_alloc(&d_segs, (size_t)n_segs * 6, &tally)) ||
        (rc = dev_alloc(&d_m, (size_t)n_segs, &tally)) || (rc = dev_alloc(&d_l, (size_t)n_segs, &tally)) ||
        (rc = dev_alloc(&d_c, (size_t)n_segs, &tally))) {
        dev_free(d_mu); dev_free(d_lam); dev_free(d_segs); dev_free(d_m); dev_free(d_l); dev_free(d_c);
        return rc;
    }
    hipStream_t st = ctx->stream;
    hipError_t e = hipMemcpyAsync(d_mu, mu, nvox * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && lam) e = hipMemcpyAsync(d_lam, lam, nvox * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_segs, segs, (size_t)n_segs * 6 * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_line_integrals(g, d_segs, n_segs, d_mu, d_lam, d_c, d_m, d_l, st);
    if (e == hipSuccess) e = hipMemcpyAsync(counts, d_c, (size_t)n_segs * sizeof(int32_t), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(mu_int, d_m, (size_t)n_segs * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(lam_int, d_l, (size_t)n_segs * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_mu); dev_free(d_lam); dev_free(d_segs); dev_free(d_m); dev_free(d_l); dev_free(d_c);
    if (e != hipSuccess) return hip_fail(e, "sss_line_integrals");
    for (int64_t i = 0; i < n_segs; ++i)
        if (counts[i] < 0) return fail(SSS_ERR_TRACE, "sss_line_integrals: a voxel walk did not end at its end voxel");
    return SSS_OK;
}

int sss_subject_create(sss_ctx* ctx, const sss_subject_desc* d, sss_subject** out)
{
    if (!ctx || !d || !out) return fail(SSS_ERR_INVALID, "sss_subject_create: NULL argument");
    *out = nullptr;
    Grid g;
    int rc = check_grid(d->grid, d->bounds, &g);
    if (rc) return rc;
    if (!d->mu || !d->geom || !d->normals || !d->ratio || !d->eff || !d->rings || !d->dets ||
        !d->slice_ptr || (d->n_sp > 0 && !d->sp_xyz))
        return fail(SSS_ERR_INVALID, "sss_subject_create: NULL array");
    if (d->n_rings <= 0 || d->n_det <= 1 || d->nru <= 0 || d->ndu <= 0 || d->nru > d->n_rings ||
        d->ndu > d->n_det)
        return fail(SSS_ERR_INVALID, "sss_subject_create: bad scanner sizes");
    if (d->n_bins < 1 || d->n_bins > kMaxBins)
        return fail(SSS_ERR_UNSUPPORTED, "sss_subject_create: 1 <= n_bins <= 32 is supported");
    if (d->tof_half_window < 0 || d->tof_half_window > kMaxHalfWindow)
        return fail(SSS_ERR_UNSUPPORTED, "sss_subject_create: TOF half window above 8 bins");
    if (d->n_energy <= 0 || d->eff_ref_row < 0 || d->eff_ref_row >= d->n_energy)
        return fail(SSS_ERR_INVALID, "sss_subject_create: bad energy tables");
    if (d->precision != SSS_PRECISION_FP32 && d->precision != SSS_PRECISION_FP64)
        return fail(SSS_ERR_INVALID, "sss_subject_create: bad precision");
    if (d->volume_order != SSS_VOLUME_X_FASTEST && d->volume_order != SSS_VOLUME_Z_FASTEST)
        return fail(SSS_ERR_INVALID, "sss_subject_create: bad volume_order");
    if (d->tile_rings != 0 && d->tile_rings != 1 && d->tile_rings != 2 && d->tile_rings != 4 && d->tile_rings != 8)
        return fail(SSS_ERR_INVALID, "sss_subject_create: tile_rings must be 0, 1, 2, 4 or 8");
    for (int r = 0; r < d->nru; ++r)
        if (d->rings[r] < 0 || d->rings[r] >= d->n_rings)
            return fail(SSS_ERR_INVALID, "sss_subject_create: ring id out of range");
    for (int i = 0; i < d->ndu; ++i) {
        if (d->dets[i] < 0 || d->dets[i] >= d->n_det)
            return fail(SSS_ERR_INVALID, "sss_subject_create: detector id out of range");
        if (i > 0 && d->dets[i] <= d->dets[i - 1])
            return fail(SSS_ERR_INVALID, "sss_subject_create: detector ids must increase");
    }
    if (d->n_slices < 0 || d->slice_ptr[0] != 0 || d->slice_ptr[d->n_slices] != d->n_sp)
        return fail(SSS_ERR_INVALID, "sss_subject_create: slice_ptr must run from 0 to n_sp");
    for (int s = 0; s < d->n_slices; ++s)
        if (d->slice_ptr[s + 1] < d->slice_ptr[s])
            return fail(SSS_ERR_INVALID, "sss_subject_create: slice_ptr must not decrease");

    SSS_HIP(hipSetDevice(ctx->device));
    sss_subject* s = new (std::nothrow) sss_subject;
    if (!s) return fail(SSS_ERR_NOMEM, "sss_subject_create: out of host memory");
    s->ctx = ctx;
    s->grid = g;
    s->n_rings = d->n_rings; s->n_det = d->n_det; s->nru = d->nru; s->ndu = d->ndu;
    s->n_sp = d->n_sp; s->n_slices = d->n_slices; s->n_bins = d->n_bins;
    s->n_energy = d->n_energy; s->eff_ref_row = d->eff_ref_row; s->precision = d->precision;
    s->trivial = d->tof_trivial != 0 && d->n_bins == 1;
    s->half_window = d->tof_half_window; s->t_lo = d->tof_t_lo; s->t_hi = d->tof_t_hi;
    s->bin_w = d->tof_bin_width; s->kappa = d->tof_kappa; s->beta = d->tof_beta;
    const double ds = d->det_size[0] * d->det_size[1] * 1e-2;
    s->g_const = ds * ds / (4.0 * M_PI);
    s->n_vox = (int64_t)g.nx * g.ny * g.nz;
    s->z_fastest = d->volume_order == SSS_VOLUME_Z_FASTEST;
    s->n_hp = (int64_t)d->n_sp * d->nru * d->ndu;
    s->n_cells_plane = (int64_t)d->ndu * (d->ndu + 1) / 2;
    s->out_elems = (int64_t)d->n_bins * d->nru * d->nru * s->n_cells_plane;
    s->dets.assign(d->dets, d->dets + d->ndu);

    // sampled detector positions; z of a ring is taken from its detector 0
    // (SingleScatterSimulationTOF.py:402,406-407)
    std::vector<double> detpos((size_t)d->nru * d->ndu * 3), detnorm((size_t)d->nru * d->ndu * 3);
    for (int r = 0; r < d->nru; ++r)
        for (int i = 0; i < d->ndu; ++i) {
            const size_t ring = (size_t)d->rings[r], det = (size_t)d->dets[i];
            const size_t o = ((size_t)r * d->ndu + i) * 3;
            detpos[o + 0] = d->geom[(ring * d->n_det + det) * 3 + 0];
            detpos[o + 1] = d->geom[(ring * d->n_det + det) * 3 + 1];
            detpos[o + 2] = d->geom[(ring * d->n_det + 0) * 3 + 2];
            for (int c = 0; c < 3; ++c) detnorm[o + c] = d->normals[(ring * d->n_det + det) * 3 + c];
        }
    // the per-sample window test can go when every offset provably stays
    // inside a table whose columns are all non-zero
    const double tof_bound = tof_offset_bound(g, d->sp_xyz, d->n_sp, detpos.data(), d->nru * d->ndu);
    s->tof_bound = tof_bound;
    s->range_check = !(s->t_lo == -kTofHalf && s->t_hi == kTofHalf && tof_bound < kTofHalf - 2);
    const size_t elem = s->precision == SSS_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    s->n_cells = (int64_t)d->nru * d->nru * s->n_cells_plane;
    s->slice_ptr_h.assign(d->slice_ptr, d->slice_ptr + d->n_slices + 1);
    Tiling tl;
    make_tiling(s, d->tile_rings, d->n_chunks, &tl);
    const std::vector<int4>& tiles = tl.tiles;
    const std::vector<int32_t>& chunk_sp = tl.chunk_sp;
    const std::vector<int32_t>& chunk_slice = tl.chunk_slice;
    const std::vector<int2>& items = tl.items;
    apply_tiling_sizes(s, tl);
    std::vector<float> ratio32(d->ratio, d->ratio + d->n_energy);

    // energy terms of the one-bin scatter: the rows a pair can reach, then row 0
    std::vector<float2> etab(kEnergyRows);
    s->etab_lo = std::min(kEnergyRowLo, d->n_energy - 1);
    s->etab_hi = std::min(kEnergyRowHi, d->n_energy - 1);
    {
        const double eff_ref = d->eff[d->eff_ref_row];
        for (int k = 0; k < kEnergyRows; ++k) {
            const int row = k == kEnergyRows - 1 ? 0 : std::min(s->etab_lo + k, s->etab_hi);
            const double e = d->eff[row] * eff_ref;
            etab[k] = make_float2(e == 0.0 ? -1.0f : (float)e, ratio32[row]);
        }
    }
    // Every array uploaded here shares one device block and one pinned
    // staging block: one copy per subject.  Each array's slot has room for
    // one element more than it holds (256-byte aligned).
    struct Upload {
        void** dst;
        const void* src;
        size_t bytes;
    };
    const Upload ups[] = {
        {reinterpret_cast<void**>(s->z_fastest ? &s->d_vol_in : &s->d_mu), d->mu, (size_t)s->n_vox * sizeof(double)},
        {reinterpret_cast<void**>(&s->d_sp), d->sp_xyz, (size_t)d->n_sp * 3 * sizeof(double)},
        {reinterpret_cast<void**>(&s->d_detpos), detpos.data(), detpos.size() * sizeof(double)},
        {reinterpret_cast<void**>(&s->d_detnorm), detnorm.data(), detnorm.size() * sizeof(double)},
        {reinterpret_cast<void**>(&s->d_eff), d->eff, (size_t)d->n_energy * sizeof(double)},
        {reinterpret_cast<void**>(&s->d_ratio), ratio32.data(), ratio32.size() * sizeof(float)},
        {reinterpret_cast<void**>(&s->d_etab), etab.data(), etab.size() * sizeof(float2)},
        {reinterpret_cast<void**>(&s->d_tiles), tiles.data(), tiles.size() * sizeof(int4)},
        {reinterpret_cast<void**>(&s->d_items), items.data(), items.size() * sizeof(int2)},
        {reinterpret_cast<void**>(&s->d_slice_ptr), d->slice_ptr, ((size_t)d->n_slices + 1) * sizeof(int32_t)},
        {reinterpret_cast<void**>(&s->d_chunk_sp), chunk_sp.data(), chunk_sp.size() * sizeof(int32_t)},
        {reinterpret_cast<void**>(&s->d_chunk_slice), chunk_slice.data(), chunk_slice.size() * sizeof(int32_t)},
        {reinterpret_cast<void**>(&s->d_dets), d->dets, (size_t)d->ndu * sizeof(int32_t)},
    };
    auto slot_bytes = [](size_t b) { return (b + 16 + 255) / 256 * 256; };
    size_t up_bytes = 0;
    for (const Upload& u : ups) up_bytes += slot_bytes(u.bytes);

    int64_t& t = s->device_bytes;
    if ((s->z_fastest && (rc = pool_alloc(ctx, &s->d_mu_own, (size_t)s->n_vox, &t))) ||
        (rc = pool_alloc(ctx, &s->d_upload, up_bytes, &t)) || (rc = pool_alloc(ctx, &s->d_lam, (size_t)s->n_vox, &t)) ||
        (rc = pool_alloc(ctx, &s->d_flag, 1, &t)) || (rc = pool_alloc(ctx, &s->d_fold_changed, 1, &t)) ||
        (rc = pool_alloc(ctx, &s->d_fast, 4, &t)) || (rc = pool_alloc(ctx, &s->d_tab, (size_t)s->n_vox + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.off, (size_t)s->n_hp + 1, &t)) || (rc = pool_alloc(ctx, &s->hp.n, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.off_fold, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.n_fold, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.rec, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.blk, (size_t)(s->n_hp + kWave - 1) / kWave + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.blk_off, (size_t)(s->n_hp + kWave - 1) / kWave + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.dist, (size_t)s->n_hp + 1, &t)) || (rc = pool_alloc(ctx, &s->hp.att, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.lnatt, (size_t)s->n_hp + 1, &t)) || (rc = pool_alloc(ctx, &s->hp.act, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.cosabs, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->hp.status, (size_t)s->n_hp + 1, &t)) ||
        (rc = pool_alloc(ctx, reinterpret_cast<unsigned char**>(&s->d_out), (size_t)s->out_elems * elem + 1, &t)) ||
        (rc = pool_alloc(ctx, reinterpret_cast<unsigned char**>(&s->d_part_sum),
                        (size_t)s->n_slots * s->n_bins * elem * s->n_chunks + 1, &t)) ||
        (rc = pool_alloc(ctx, &s->d_part_cnt, (size_t)s->n_slots * s->n_chunks + 1, &t))) {
        delete s;
        return rc;
    }
    s->hp.len_min = reinterpret_cast<uint32_t*>(s->d_fast + 2);
    s->part_sum_cap = (size_t)s->n_slots * s->n_bins * elem * s->n_chunks + 1;
    s->part_cnt_cap = (size_t)s->n_slots * s->n_chunks + 1;
    hipStream_t st = ctx->stream;
    // The upload is queued behind whatever the context's stream holds, from
    // the pinned block that the subject keeps until destroy: nothing here
    // waits for the stream.
    if ((rc = host_block(ctx, up_bytes, &s->h_upload))) {
        delete s;
        return rc;
    }
    {
        size_t off = 0;
        unsigned char* stage = static_cast<unsigned char*>(s->h_upload);
        for (const Upload& u : ups) {
            *u.dst = s->d_upload + off;
            if (u.bytes) std::memcpy(stage + off, u.src, u.bytes);
            off += slot_bytes(u.bytes);
        }
    }
    // cells no lane owns (i1 == i2 on same-ring planes) are never written: zero once
    hipError_t e = hipMemsetAsync(s->d_out, 0, (size_t)s->out_elems * elem, st);
    if (e == hipSuccess) e = hipMemcpyAsync(s->d_upload, s->h_upload, up_bytes, hipMemcpyHostToDevice, st);
    if (s->z_fastest) {
        s->d_mu = s->d_mu_own;
        if (e == hipSuccess) e = launch_volume_x_fastest(s->d_vol_in, s->d_mu, g.nx, g.ny, g.nz, st);
    }
    if (e == hipSuccess) e = ctx_event(ctx, true, &s->ev_scatter[0]);
    if (e == hipSuccess) e = ctx_event(ctx, true, &s->ev_scatter[1]);
    if (e == hipSuccess) e = ctx_event(ctx, false, &s->ev_count);
    if (e == hipSuccess) e = ctx_event(ctx, false, &s->ev_out);
    if (e == hipSuccess) e = ctx_event(ctx, false, &s->ev_lam);
    if (e == hipSuccess) e = ctx_event(ctx, false, &s->ev_upload);
    if (e == hipSuccess) e = hipEventRecord(s->ev_upload, st);
    if (e == hipSuccess && ctx->pinned_free.empty()) e = hipErrorOutOfMemory;   // kPinnedSlots live subjects
    if (e == hipSuccess) {
        s->h_total = ctx->pinned + ctx->pinned_free.back();
        ctx->pinned_free.pop_back();
    }
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(st);   // queued uploads read the subject's host copies
        delete s;
        return hip_fail(e, "sss_subject_create: upload");
    }
    *out = s;
    return SSS_OK;
}

int sss_subject_set_tiling(sss_subject* s, int tile_rings, int n_chunks)
{
    if (!s) return fail(SSS_ERR_INVALID, "sss_subject_set_tiling: NULL subject");
    if (tile_rings != 0 && tile_rings != 1 && tile_rings != 2 && tile_rings != 4 && tile_rings != 8)
        return fail(SSS_ERR_INVALID, "sss_subject_set_tiling: tile_rings must be 0, 1, 2, 4 or 8");
    if (n_chunks < 0) return fail(SSS_ERR_INVALID, "sss_subject_set_tiling: n_chunks must be >= 0");
    SSS_HIP(hipSetDevice(s->ctx->device));
    sss_ctx* c = s->ctx;
    hipStream_t st = c->stream;
    Tiling t;
    make_tiling(s, tile_rings, n_chunks, &t);
    // one device block: tiles | items | chunk_sp | chunk_slice (256-byte slots)
    auto slot = [](size_t b) { return (b + 16 + 255) / 256 * 256; };
    const size_t b_tiles = t.tiles.size() * sizeof(int4), b_items = t.items.size() * sizeof(int2);
    const size_t b_sp = t.chunk_sp.size() * sizeof(int32_t), b_sl = t.chunk_slice.size() * sizeof(int32_t);
    const size_t o_items = slot(b_tiles), o_sp = o_items + slot(b_items), o_sl = o_sp + slot(b_sp);
    std::vector<unsigned char> img(o_sl + slot(b_sl), 0);
    std::memcpy(img.data(), t.tiles.data(), b_tiles);
    std::memcpy(img.data() + o_items, t.items.data(), b_items);
    std::memcpy(img.data() + o_sp, t.chunk_sp.data(), b_sp);
    std::memcpy(img.data() + o_sl, t.chunk_slice.data(), b_sl);
    unsigned char* blk = nullptr;
    if (int rc = pool_alloc(c, &blk, img.size(), &s->device_bytes)) return rc;
    const size_t elem = s->precision == SSS_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    const size_t n_slots = t.tiles.size() * kWave, n_ch = t.chunk_slice.size();
    const size_t need_sum = n_slots * s->n_bins * elem * n_ch + 1, need_cnt = n_slots * n_ch + 1;
    unsigned char* sum = nullptr;
    float* cnt = nullptr;
    if (need_sum > s->part_sum_cap) {
        if (int rc = pool_alloc(c, &sum, need_sum, &s->device_bytes)) {
            pool_free(c, blk);
            return rc;
        }
    }
    if (need_cnt > s->part_cnt_cap) {
        if (int rc = pool_alloc(c, &cnt, need_cnt, &s->device_bytes)) {
            pool_free(c, blk);
            pool_free(c, sum);
            return rc;
        }
    }
    hipError_t e = hipMemcpyAsync(blk, img.data(), img.size(), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);   // img is pageable and local
    if (e != hipSuccess) {
        pool_free(c, blk);
        pool_free(c, sum);
        pool_free(c, cnt);
        return hip_fail(e, "sss_subject_set_tiling: upload");
    }
    // the old blocks go back to the cache; their queued readers ran above
    pool_free(c, s->d_tiling);
    s->d_tiling = blk;
    if (sum) {
        pool_free(c, s->d_part_sum);
        s->d_part_sum = sum;
        s->part_sum_cap = need_sum;
    }
    if (cnt) {
        pool_free(c, s->d_part_cnt);
        s->d_part_cnt = cnt;
        s->part_cnt_cap = need_cnt;
    }
    s->d_tiles = reinterpret_cast<int4*>(blk);
    s->d_items = reinterpret_cast<int2*>(blk + o_items);
    s->d_chunk_sp = reinterpret_cast<int32_t*>(blk + o_sp);
    s->d_chunk_slice = reinterpret_cast<int32_t*>(blk + o_sl);
    apply_tiling_sizes(s, t);
    ++s->generation;   // batches holding the subject rebuild their lists
    return SSS_OK;
}

int sss_subject_destroy(sss_subject* s)
{
    if (!s) return SSS_OK;
    if (s->batch_refs > 0) return fail(SSS_ERR_INVALID, "sss_subject_destroy: the subject is in a live batch");
    (void)hipSetDevice(s->ctx->device);
    // the device blocks go back to the context's cache: their next users are
    // queued behind this subject's work on the same stream; the pinned count
    // may still be the target of a queued copy
    if (s->trace_begun) (void)hipEventSynchronize(s->ev_count);
    // the queued uploads read host memory the subject owns
    if (s->ev_upload) (void)hipEventSynchronize(s->ev_upload);
    if (s->lam_queued) (void)hipEventSynchronize(s->ev_lam);
    delete s;
    return SSS_OK;
}

// Exclusive scan of n_items + 1 record counts (slot n_items is 0, so
// out[n_items] is the total) into int64 offsets.
extern "C++" template <typename It>
static int scan_counts(sss_subject* s, It counts, int64_t* out, int64_t n_items, hipStream_t st)
{
    size_t need = 0;
    SSS_HIP(rocprim::exclusive_scan(nullptr, need, counts, out, (int64_t)0, (size_t)n_items + 1,
                                    rocprim::plus<int64_t>(), st));
    if (need > s->scan_tmp_bytes) {
        pool_free(s->ctx, s->d_scan_tmp);
        unsigned char* tmp = nullptr;
        int rc = pool_alloc(s->ctx, &tmp, need, &s->device_bytes);
        if (rc) return rc;
        s->d_scan_tmp = tmp;
        s->scan_tmp_bytes = need;
    }
    SSS_HIP(rocprim::exclusive_scan(s->d_scan_tmp, need, counts, out, (int64_t)0, (size_t)n_items + 1,
                                    rocprim::plus<int64_t>(), st));
    return SSS_OK;
}

// Per activity image: emission terms, their sums and the fold table of the
// non-zero terms.
static int refresh_fold(sss_subject* s)
{
    hipStream_t st = s->ctx->stream;
    if (s->trivial) {   // one bin: the pair terms need the integrals only
        SSS_HIP(launch_activity_sum(s->hp, s->smp, s->n_hp, s->d_lam, st));
        SSS_HIP(launch_pair_half(s->hp, s->n_hp, true, st));
        return SSS_OK;
    }
    // When the fold table was built for these walks, the new image most often
    // has the same zero terms (served requests rescale the activity): then
    // only its weights are rewritten.  The checks set d_fold_changed on any
    // difference, and the full refresh and compaction below run only then
    // (they test the flag on the device: no wait here).
    const int32_t* run_if = nullptr;
    // the record -> sample map costs the compaction 4 bytes a record: it is
    // kept from the second image after the walks on (a subject that takes
    // one image per trace, as one reference call does, never needs it)
    SampleTable smp = s->smp;
    if (s->refreshes++ == 0) smp.fold_map = nullptr;
    if (s->fold_valid && smp.fold_map) {
        // A voxel test against the table's image decides first: when it
        // passes, the same samples keep non-zero terms and the same half
        // paths non-zero integrals, so the per-sample count check and the
        // pair records are skipped (on the device, no wait here)
        // flags cleared by a kernel: batches capture this sequence in a graph
        SSS_HIP(launch_fill_i32(s->d_fold_changed, 1, 0, st));
        SSS_HIP(launch_fill_i32(s->d_fast, 1, 0, st));
        SSS_HIP(launch_vox_check(s->d_lam, s->n_vox, s->d_tab, s->hp.len_min, s->d_fast, st));
        SSS_HIP(launch_activity_check(s->hp, smp, s->n_hp, s->d_lam, s->d_fold_changed, st, s->d_fast));
        SSS_HIP(launch_fold_weights(smp, s->hp.off_fold + s->n_hp, s->fold_cap, s->d_lam, s->d_fold_changed, st));
        run_if = s->d_fold_changed;
    }
    SSS_HIP(launch_activity_refresh(s->hp, smp, s->n_hp, s->d_lam, st, run_if));
    const auto padded = rocprim::make_transform_iterator(s->hp.n_fold, PaddedRecords{});
    if (int rc = scan_counts(s, padded, s->hp.off_fold, s->n_hp, st)) return rc;
    SSS_HIP(launch_fold_compact(s->hp, smp, s->n_hp, st, run_if));
    // the image a rebuilt table comes from, for the next images' voxel test
    // (which needs the record map, so the first image after the walks, whose
    // table keeps none, does not need it either)
    if (smp.fold_map) SSS_HIP(launch_vox_tab(s->d_lam, s->n_vox, s->d_tab, s->hp.len_min, s->d_fast, st, run_if));
    SSS_HIP(launch_pair_half(s->hp, s->n_hp, false, st, run_if ? s->d_fast : nullptr, run_if));
    s->fold_valid = smp.fold_map != nullptr;
    return SSS_OK;
}

int sss_subject_trace_begin(sss_subject* s)
{
    if (!s) return fail(SSS_ERR_INVALID, "sss_subject_trace_begin: NULL subject");
    SSS_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    s->traced = false;
    s->trace_begun = true;
    *s->h_total = 0;
    if (s->n_hp == 0) return SSS_OK;
    SSS_HIP(launch_halfpath_count(s->grid, s->d_sp, s->n_sp, s->d_detpos, s->d_detnorm, s->nru, s->ndu,
                                  s->hp, st));
    // the last slot holds a zero count so the scan's last entry is the total
    const int64_t n_blk = (s->n_hp + kWave - 1) / kWave;
    SSS_HIP(hipMemsetAsync(s->hp.blk + n_blk, 0, sizeof(int32_t), st));
    SSS_HIP(hipMemsetAsync(s->hp.n_fold + s->n_hp, 0, sizeof(int32_t), st));
    if (int rc = scan_counts(s, s->hp.blk, s->hp.blk_off, n_blk, st)) return rc;
    SSS_HIP(hipMemcpyAsync(s->h_total, s->hp.blk_off + n_blk, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SSS_HIP(hipEventRecord(s->ev_count, st));
    return SSS_OK;
}

int sss_subject_trace_end(sss_subject* s)
{
    if (!s) return fail(SSS_ERR_INVALID, "sss_subject_trace_end: NULL subject");
    if (!s->trace_begun) return fail(SSS_ERR_INVALID, "sss_subject_trace_end: call sss_subject_trace_begin first");
    SSS_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    if (s->n_hp == 0) {
        s->trace_begun = false;
        s->n_samples = 0;
        s->traced = true;
        return SSS_OK;
    }
    SSS_HIP(hipEventSynchronize(s->ev_count));
    const int64_t total = *s->h_total;
    if (total > s->samples_cap || !s->smp.nq) {
        sss_ctx* c = s->ctx;
        pool_free(c, s->smp.nq); pool_free(c, s->smp.vox); pool_free(c, s->smp.len); pool_free(c, s->smp.len32);
        pool_free(c, s->smp.a);
        pool_free(c, s->smp.a64); pool_free(c, s->smp.fold); pool_free(c, s->smp.fold_a64);
        pool_free(c, s->smp.fold_map);
        s->samples_cap = 0;
        s->fold_cap = 0;
        // the fold table holds at most every sample plus the group padding,
        // and spare records after the last one (the fold prefetches ahead)
        size_t fold_cap = (size_t)total + (size_t)(kSampleGroup - 1) * s->n_hp + kSamplePad;
        fold_cap = (fold_cap + kSampleGroup - 1) / kSampleGroup * kSampleGroup;
        const bool fp64 = s->precision == SSS_PRECISION_FP64;
        const bool fold = !s->trivial;   // one-bin subjects sum the samples, no fold table
        int64_t* tally = &s->device_bytes;
        int rc;
        if ((rc = pool_alloc(c, &s->smp.nq, (size_t)total, tally)) ||
            (rc = pool_alloc(c, &s->smp.vox, (size_t)total, tally)) ||
            (fp64 && (rc = pool_alloc(c, &s->smp.len, (size_t)total, tally))) ||
            (!fp64 && (rc = pool_alloc(c, &s->smp.len32, (size_t)total, tally))) ||
            (fold && (rc = pool_alloc(c, &s->smp.a, (size_t)total, tally))) ||
            (fold && (rc = pool_alloc(c, &s->smp.fold, fold_cap / kSampleGroup * kFoldGroupWords, tally))) ||
            (fold && fp64 && ((rc = pool_alloc(c, &s->smp.a64, (size_t)total, tally)) ||
                              (rc = pool_alloc(c, &s->smp.fold_a64, fold_cap, tally)))) ||
            // sample entries of the fold records (int32 entries), for
            // weights-only refreshes
            (fold && total < INT32_MAX && (rc = pool_alloc(c, &s->smp.fold_map, fold_cap, tally))))
            return rc;
        s->fold_cap = fold ? (int64_t)fold_cap : 0;
        s->samples_cap = total;
    }
    s->n_samples = total;
    ++s->generation;
    s->fold_valid = false;   // new walks: the next activity builds the fold table in full
    s->refreshes = 0;
    // the fill keeps the smallest positive length (0x7f7f7f7f: 3.4e38 as float bits)
    SSS_HIP(hipMemsetAsync(s->hp.len_min, 0x7f, sizeof(uint32_t), st));
    SSS_HIP(launch_halfpath_fill(s->grid, s->d_sp, s->n_sp, s->d_detpos, s->nru, s->ndu, s->d_mu, s->hp,
                                 s->smp, st));
    // truncated walks are counted (read by sss_subject_info_get), not an error
    SSS_HIP(hipMemsetAsync(s->d_flag, 0, sizeof(int32_t), st));
    hipLaunchKernelGGL(k_count_status, dim3(256), dim3(256), 0, st, s->hp.status, s->n_hp, s->d_flag);
    SSS_HIP(hipGetLastError());
    // the walks are queued: a failure before this point (device memory)
    // leaves the trace begun, so the caller may free memory and call
    // sss_subject_trace_end again
    s->trace_begun = false;
    s->traced = true;
    if (s->has_activity) return refresh_fold(s);
    return SSS_OK;
}

int sss_subject_trace(sss_subject* s)
{
    if (int rc = sss_subject_trace_begin(s)) return rc;
    return sss_subject_trace_end(s);
}

static int refresh_activity(sss_subject* s)
{
    s->has_activity = true;
    if (s->traced) return refresh_fold(s);
    return SSS_OK;
}

int sss_subject_set_activity(sss_subject* s, const double* lam)
{
    if (!s || !lam) return fail(SSS_ERR_INVALID, "sss_subject_set_activity: NULL argument");
    SSS_HIP(hipSetDevice(s->ctx->device));
    // the caller's buffer is copied before the call returns; the queued upload
    // reads the subject's copy (the previous upload from it has finished)
    if (s->lam_queued) SSS_HIP(hipEventSynchronize(s->ev_lam));
    if (!s->h_lam) {
        void* p = nullptr;
        if (int rc = host_block(s->ctx, (size_t)s->n_vox * sizeof(double), &p)) return rc;
        s->h_lam = static_cast<double*>(p);
    }
    std::memcpy(s->h_lam, lam, (size_t)s->n_vox * sizeof(double));
    SSS_HIP(hipMemcpyAsync(s->z_fastest ? s->d_vol_in : s->d_lam, s->h_lam, (size_t)s->n_vox * sizeof(double),
                           hipMemcpyHostToDevice, s->ctx->stream));
    if (s->z_fastest)
        SSS_HIP(launch_volume_x_fastest(s->d_vol_in, s->d_lam, s->grid.nx, s->grid.ny, s->grid.nz, s->ctx->stream));
    SSS_HIP(hipEventRecord(s->ev_lam, s->ctx->stream));
    s->lam_queued = true;
    return refresh_activity(s);
}

int sss_subject_set_activity_device(sss_subject* s, const void* lam_dev)
{
    if (!s || !lam_dev) return fail(SSS_ERR_INVALID, "sss_subject_set_activity_device: NULL argument");
    SSS_HIP(hipSetDevice(s->ctx->device));
    if (s->z_fastest)
        SSS_HIP(launch_volume_x_fastest(static_cast<const double*>(lam_dev), s->d_lam, s->grid.nx, s->grid.ny,
                                        s->grid.nz, s->ctx->stream));
    else
        SSS_HIP(hipMemcpyAsync(s->d_lam, lam_dev, (size_t)s->n_vox * sizeof(double), hipMemcpyDeviceToDevice,
                               s->ctx->stream));
    return refresh_activity(s);
}

static ScatterArgs scatter_args(const sss_subject* s)
{
    ScatterArgs a{};
    a.n_sp = s->n_sp; a.nru = s->nru; a.ndu = s->ndu; a.n_bins = s->n_bins;
    a.n_energy = s->n_energy; a.eff_ref_row = s->eff_ref_row;
    a.n_tiles = s->n_tiles; a.n_chunks = s->n_chunks; a.n_cells_plane = s->n_cells_plane;
    a.hp_rec = s->hp.rec;
    a.etab = s->d_etab;
    a.etab_lo = s->etab_lo;
    a.etab_hi = s->etab_hi;
    a.hp_off = s->hp.off_fold; a.hp_n = s->hp.n_fold; a.hp_n_all = s->hp.n; a.hp_dist = s->hp.dist; a.hp_att = s->hp.att;
    a.hp_lnatt = s->hp.lnatt; a.hp_act = s->hp.act; a.hp_cos = s->hp.cosabs;
    a.s_fold = s->smp.fold; a.s_a64 = s->smp.fold_a64;
    a.detpos = s->d_detpos; a.eff = s->d_eff; a.ratio = s->d_ratio; a.tiles = s->d_tiles;
    a.items = s->d_items; a.n_items = s->n_items; a.tile_db_log2 = s->tile_db_log2;
    a.slice_ptr = s->d_slice_ptr; a.chunk_sp = s->d_chunk_sp;
    const double log2e = 1.4426950408889634;
    a.inv_w_bins = 1.0 / s->bin_w;
    a.mid = (s->n_bins - 1) / 2.0;
    a.bin_w = s->bin_w;
    a.kappa2 = s->kappa * log2e;
    a.beta2 = s->beta * log2e;
    a.c_beta = std::exp2(-a.beta2);
    a.half_window = s->half_window; a.t_lo = s->t_lo; a.t_hi = s->t_hi;
    // one exp2 per sample needs c E^2 = 2^(2 ku - beta2) finite; inside the
    // bins ku <= beta2.  With the window centre kept in [1, B-2], ku reaches
    // kappa2 (|t| - s_(B-2)) at the largest offset |t| of the subject.
    a.tof_mode = kTofGeneral;
    if (s->half_window == 1 && a.beta2 <= 60.0) {
        a.tof_mode = kTofOneExp;
        if (s->n_bins >= 3 && !s->range_check) {
            const double s_inner = (s->n_bins - 2 - a.mid) * s->bin_w;
            const double umax = std::max(0.5 * s->bin_w, s->tof_bound - s_inner);
            if (2.0 * a.kappa2 * umax - a.beta2 < 120.0 && a.kappa2 * umax < 120.0)
                a.tof_mode = kTofOneExpInner;
        }
    }
    // float subjects whose fold is not the inner sliding window: the relative
    // sliding window (h = 1 at any beta2; h = 2 while 2^(-4 beta2) is a
    // normal float)
    const bool rel = s->precision != SSS_PRECISION_FP64 && !s->range_check && a.tof_mode != kTofOneExpInner &&
                     (s->half_window == 1 || (s->half_window == 2 && a.beta2 < 30.0));
    // one exp2 per sample while every term stays a normal float (h = 1:
    // beta2 + 30 <= 125), else the two-exp2 form
    if (rel) a.tof_mode = (s->half_window == 1 && a.beta2 > 95.0) ? kTofSlideRelWide : kTofSlideRel;
    a.range_check = s->range_check ? 1 : 0;
    a.f_inv_w = (float)a.inv_w_bins;
    a.f_kap = (float)a.kappa2;
    a.f_mid = (float)a.mid;
    a.f_jmax = (float)(s->n_bins - 1);
    a.f_wk = (float)(a.bin_w * a.kappa2);
    a.f_mid_wk = (float)(a.mid * a.bin_w * a.kappa2);
    a.f_c = (float)a.c_beta;
    a.f_inv_c = (float)std::exp2(a.beta2);
    a.slide_row0 = s->half_window;
    a.f_beta2 = (float)a.beta2;
    a.f_w = (float)s->bin_w;
    a.f_l0 = (float)((-a.mid - 0.5) * s->bin_w);
    a.f_l1 = (float)((0.5 - a.mid) * s->bin_w);
    a.f_c4 = (float)std::exp2(-4.0 * a.beta2);
    a.f_kclamp = (float)(a.beta2 + 10.0);
    a.f_kclamp1 = (float)(a.beta2 + 30.0);
    a.f_wk_lo = (float)(a.bin_w * a.kappa2 - (double)a.f_wk);
    a.f_mid_wk_lo = (float)(a.mid * a.bin_w * a.kappa2 - (double)a.f_mid_wk);
    a.g_const = s->g_const;
    a.kn_gamma = (511.0e3 * 1.6e-19) / (9.10938356e-31 * (299792458.0 * 299792458.0));
    a.part_sum = s->d_part_sum;
    a.part_cnt = s->d_part_cnt;
    a.chunk_slice = s->d_chunk_slice;
    a.out = s->d_out;
    return a;
}

int sss_subject_scatter(sss_subject* s)
{
    if (!s) return fail(SSS_ERR_INVALID, "sss_subject_scatter: NULL subject");
    if (!s->traced) return fail(SSS_ERR_INVALID, "sss_subject_scatter: call sss_subject_trace first");
    if (!s->has_activity) return fail(SSS_ERR_INVALID, "sss_subject_scatter: no activity image set");
    SSS_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const bool fp64 = s->precision == SSS_PRECISION_FP64;
    const ScatterArgs a = scatter_args(s);
    SSS_HIP(hipEventRecord(s->ev_scatter[0], st));
    SSS_HIP(launch_scatter(a, fp64, s->trivial, st));
    SSS_HIP(hipEventRecord(s->ev_scatter[1], st));
    s->ev_recorded = true;
    SSS_HIP(launch_chunk_reduce(a, fp64, st));
    return SSS_OK;
}

int sss_subject_scatter_kernel_ms(sss_subject* s, float* ms)
{
    if (!s || !ms) return fail(SSS_ERR_INVALID, "sss_subject_scatter_kernel_ms: NULL argument");
    if (!s->ev_recorded) return fail(SSS_ERR_INVALID, "sss_subject_scatter_kernel_ms: no scatter run yet");
    SSS_HIP(hipSetDevice(s->ctx->device));
    SSS_HIP(hipEventSynchronize(s->ev_scatter[1]));
    SSS_HIP(hipEventElapsedTime(ms, s->ev_scatter[0], s->ev_scatter[1]));
    return SSS_OK;
}

int sss_subject_work_stats(sss_subject* s, int64_t* pairs, int64_t* samples)
{
    if (!s || !pairs || !samples) return fail(SSS_ERR_INVALID, "sss_subject_work_stats: NULL argument");
    if (!s->traced || !s->has_activity)
        return fail(SSS_ERR_INVALID, "sss_subject_work_stats: trace and set the activity first");
    SSS_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    unsigned long long* d = nullptr;
    int rc = dev_alloc(&d, 2, nullptr);
    if (rc) return rc;
    unsigned long long h[2] = {0, 0};
    hipError_t e = hipMemsetAsync(d, 0, sizeof(h), st);
    if (e == hipSuccess) e = launch_pair_stats(scatter_args(s), d, st);
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d);
    if (e != hipSuccess) return hip_fail(e, "sss_subject_work_stats");
    *pairs = (int64_t)h[0];
    *samples = (int64_t)h[1];
    return SSS_OK;
}

int sss_subject_run(sss_subject* s)
{
    int rc = sss_subject_trace(s);
    if (rc) return rc;
    return sss_subject_scatter(s);
}

int sss_subject_info_get(sss_subject* s, sss_subject_info* out)
{
    if (!s || !out) return fail(SSS_ERR_INVALID, "sss_subject_info_get: NULL argument");
    out->n_halfpaths = s->n_hp;
    out->n_samples = s->n_samples;
    out->n_cells = (int64_t)s->nru * s->nru * s->n_cells_plane;
    const int64_t lors = (int64_t)s->nru * (s->ndu * (int64_t)(s->ndu - 1) / 2) +
                         (int64_t)s->nru * (s->nru - 1) * ((int64_t)s->ndu * (s->ndu + 1) / 2);
    out->n_pairs_enumerated = lors * s->n_sp;
    out->evals = out->n_pairs_enumerated * s->n_bins;
    out->device_bytes = s->device_bytes;
    out->n_tiles = s->n_tiles;
    out->n_chunks = s->n_chunks;
    int32_t trunc = 0;
    if (s->traced) {
        SSS_HIP(hipSetDevice(s->ctx->device));
        SSS_HIP(hipMemcpyAsync(&trunc, s->d_flag, sizeof(int32_t), hipMemcpyDeviceToHost, s->ctx->stream));
        SSS_HIP(hipStreamSynchronize(s->ctx->stream));
    }
    out->n_truncated_walks = trunc;
    const ScatterArgs a = scatter_args(s);
    out->tof_fold = (s->trivial || a.half_window == 0) ? kTofGeneral : a.tof_mode;
    if (out->tof_fold == kTofOneExpInner && s->precision != SSS_PRECISION_FP64)
        out->tof_fold = kTofSlideQuad;
    return SSS_OK;
}

const void* sss_subject_output_device(const sss_subject* s) { return s ? s->d_out : nullptr; }

int sss_subject_output_compact_async(sss_subject* s, float* host_out)
{
    if (!s || !host_out) return fail(SSS_ERR_INVALID, "sss_subject_output_compact_async: NULL argument");
    SSS_HIP(hipSetDevice(s->ctx->device));
    const size_t elem = s->precision == SSS_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    SSS_HIP(hipMemcpyAsync(host_out, s->d_out, (size_t)s->out_elems * elem, hipMemcpyDeviceToHost,
                           s->ctx->stream));
    SSS_HIP(hipEventRecord(s->ev_out, s->ctx->stream));
    return SSS_OK;
}

int sss_subject_query(sss_subject* s, int32_t* done)
{
    if (!s || !done) return fail(SSS_ERR_INVALID, "sss_subject_query: NULL argument");
    SSS_HIP(hipSetDevice(s->ctx->device));
    const hipError_t e = hipEventQuery(s->ev_out);
    if (e == hipErrorNotReady) {
        *done = 0;
        return SSS_OK;
    }
    SSS_HIP(e);
    *done = 1;
    return SSS_OK;
}

int sss_subject_wait(sss_subject* s)
{
    if (!s) return fail(SSS_ERR_INVALID, "sss_subject_wait: NULL subject");
    SSS_HIP(hipSetDevice(s->ctx->device));
    SSS_HIP(hipEventSynchronize(s->ev_out));
    return SSS_OK;
}

int sss_subject_output_compact(sss_subject* s, float* host_out)
{
    if (!s || !host_out) return fail(SSS_ERR_INVALID, "sss_subject_output_compact: NULL argument");
    SSS_HIP(hipSetDevice(s->ctx->device));
    const size_t elem = s->precision == SSS_PRECISION_FP64 ? sizeof(double) : sizeof(float);
    SSS_HIP(hipMemcpyAsync(host_out, s->d_out, (size_t)s->out_elems * elem, hipMemcpyDeviceToHost,
                           s->ctx->stream));
    // wait for this subject's work only, not for what later calls queued
    SSS_HIP(hipEventRecord(s->ev_out, s->ctx->stream));
    SSS_HIP(hipEventSynchronize(s->ev_out));
    return SSS_OK;
}

int sss_subject_output_dense(sss_subject* s, const int16_t* lor_coords, float* host_out)
{
    if (!s || !lor_coords || !host_out) return fail(SSS_ERR_INVALID, "sss_subject_output_dense: NULL argument");
    SSS_HIP(hipSetDevice(s->ctx->device));
    hipStream_t st = s->ctx->stream;
    const size_t n_dense = (size_t)s->n_bins * (s->n_det + 1) * (s->n_det / 2) * s->nru * s->nru;
    const size_t n_lor = (size_t)s->n_det * s->n_det * 2;
    float* d_dense = nullptr;
    int16_t* d_lor = nullptr;
    int64_t tally = 0;
    int rc;
    if ((rc = dev_alloc(&d_dense, n_dense, &tally)) || (rc = dev_alloc(&d_lor, n_lor, &tally))) {
        dev_free(d_dense); dev_free(d_lor);
        return rc;
    }
    hipError_t e = hipMemsetAsync(d_dense, 0, n_dense * sizeof(float), st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lor, lor_coords, n_lor * sizeof(int16_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_dense_scatter_typed(s->d_out, s->precision == SSS_PRECISION_FP64, s->n_bins, s->nru,
                                       s->ndu, s->n_det, s->d_dets, d_lor, d_dense, st);
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, d_dense, n_dense * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_dense); dev_free(d_lor);
    if (e != hipSuccess) return hip_fail(e, "sss_subject_output_dense");
    return SSS_OK;
}

/* ---- device buffers ---------------------------------------------------- */

int sss_buffer_alloc(sss_ctx* ctx, int64_t bytes, void** dev)
{
    if (!ctx || !dev || bytes < 0) return fail(SSS_ERR_INVALID, "sss_buffer_alloc: bad argument");
    SSS_HIP(hipSetDevice(ctx->device));
    unsigned char* p = nullptr;
    int rc = dev_alloc(&p, (size_t)(bytes > 0 ? bytes : 1), nullptr);
    if (rc) return rc;
    *dev = p;
    return SSS_OK;
}

int sss_buffer_free(sss_ctx* ctx, void* dev)
{
    if (!ctx) return fail(SSS_ERR_INVALID, "sss_buffer_free: NULL context");
    if (!dev) return SSS_OK;
    SSS_HIP(hipSetDevice(ctx->device));
    SSS_HIP(hipStreamSynchronize(ctx->stream));
    SSS_HIP(hipFree(dev));
    return SSS_OK;
}

int sss_buffer_upload(sss_ctx* ctx, void* dev, const void* host, int64_t bytes)
{
    if (!ctx || (bytes > 0 && (!dev || !host)) || bytes < 0)
        return fail(SSS_ERR_INVALID, "sss_buffer_upload: bad argument");
    if (bytes == 0) return SSS_OK;
    SSS_HIP(hipSetDevice(ctx->device));
    SSS_HIP(hipMemcpyAsync(dev, host, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    SSS_HIP(hipStreamSynchronize(ctx->stream));
    return SSS_OK;
}

int sss_buffer_download(sss_ctx* ctx, void* host, const void* dev, int64_t bytes)
{
    if (!ctx || (bytes > 0 && (!dev || !host)) || bytes < 0)
        return fail(SSS_ERR_INVALID, "sss_buffer_download: bad argument");
    if (bytes == 0) return SSS_OK;
    SSS_HIP(hipSetDevice(ctx->device));
    SSS_HIP(hipMemcpyAsync(host, dev, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    SSS_HIP(hipStreamSynchronize(ctx->stream));
    return SSS_OK;
}

/* ---- full-sinogram stages ---------------------------------------------- */

int sss_interpolate(sss_ctx* ctx, const float* sampled, int n_bins, int n_sampled_planes,
                    int64_t cells_per_sampled_plane, const int32_t* cell_map, int64_t n_cells_out,
                    const int32_t* plane_map, int n_out_planes, float* out, float* sum)
{
    if (!ctx || !sampled || !cell_map || !plane_map || (!out && !sum) || n_bins < 1 ||
        n_sampled_planes < 1 || cells_per_sampled_plane < 1 || n_cells_out < 0 || n_out_planes < 0)
        return fail(SSS_ERR_INVALID, "sss_interpolate: bad argument");
    SSS_HIP(hipSetDevice(ctx->device));
    SSS_HIP(launch_interp(sampled, n_bins, n_sampled_planes, cells_per_sampled_plane, cell_map,
                          n_cells_out, plane_map, n_out_planes, out, sum, ctx->stream));
    return SSS_OK;
}

static int check_planes(const int32_t* plane_rings, int n, int stride, int n_rings_geom, int n_value_planes)
{
    for (int p = 0; p < n; ++p) {
        const int ra = plane_rings[stride * p], rb = plane_rings[stride * p + 1];
        if (ra < 0 || ra >= n_rings_geom || rb < 0 || rb >= n_rings_geom)
            return fail(SSS_ERR_INVALID, "plane ring out of range");
        if (stride == 3 && (plane_rings[3 * p + 2] < 0 || plane_rings[3 * p + 2] >= n_value_planes))
            return fail(SSS_ERR_INVALID, "value plane out of range");
    }
    return SSS_OK;
}

int sss_tail_mask(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], const double* image,
                  int use_attenuation, int n_rings_geom, int n_det, const double* geom,
                  const int16_t* lor_coords, int n_planes, const int32_t* plane_rings, int accel,
                  uint8_t* mask_out)
{
    if (!ctx || !image || !geom || !lor_coords || !plane_rings || !mask_out || n_det < 2 ||
        n_rings_geom < 1 || n_planes < 0 || accel < 1)
        return fail(SSS_ERR_INVALID, "sss_tail_mask: bad argument");
    Grid g;
    int rc = check_grid(grid, bounds, &g);
    if (rc) return rc;
    if ((rc = check_planes(plane_rings, n_planes, 2, n_rings_geom, 0))) return rc;
    SSS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    const size_t ngeom = (size_t)n_rings_geom * n_det * 3, nlor = (size_t)n_det * n_det * 2;
    const size_t nmask = (size_t)(n_det + 1) * (n_det / 2) * n_planes;
    double *d_img = nullptr, *d_geom = nullptr;
    int16_t* d_lor = nullptr;
    int32_t* d_pr = nullptr;
    uint8_t* d_mask = nullptr;
    int64_t tally = 0;
    if ((rc = dev_alloc(&d_img, nvox, &tally)) || (rc = dev_alloc(&d_geom, ngeom, &tally)) ||
        (rc = dev_alloc(&d_lor, nlor, &tally)) || (rc = dev_alloc(&d_pr, (size_t)2 * n_planes + 1, &tally)) ||
        (rc = dev_alloc(&d_mask, nmask + 1, &tally))) {
        dev_free(d_img); dev_free(d_geom); dev_free(d_lor); dev_free(d_pr); dev_free(d_mask);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_img, image, nvox * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_geom, geom, ngeom * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lor, lor_coords, nlor * sizeof(int16_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n_planes > 0)
        e = hipMemcpyAsync(d_pr, plane_rings, (size_t)2 * n_planes * sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_mask, 0, nmask + 1, st);
    if (e == hipSuccess)
        e = launch_tail_mask(g, d_img, use_attenuation, d_geom, n_det, d_lor, n_planes, d_pr, accel, d_mask, st);
    if (e == hipSuccess) e = hipMemcpyAsync(mask_out, d_mask, nmask, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_img); dev_free(d_geom); dev_free(d_lor); dev_free(d_pr); dev_free(d_mask);
    if (e != hipSuccess) return hip_fail(e, "sss_tail_mask");
    return SSS_OK;
}

int sss_forward_project(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], const double* lam,
                        const double* mu, int n_rings_geom, int n_det, const double* geom, const int16_t* lor_coords,
                        int n_planes, const int32_t* plane_rings, int accel, float* trues_out)
{
    if (!ctx || !lam || !geom || !lor_coords || (n_planes > 0 && !plane_rings) || !trues_out || n_det < 2 ||
        n_rings_geom < 1 || n_planes < 0 || accel < 1)
        return fail(SSS_ERR_INVALID, "sss_forward_project: bad argument");
    Grid g;
    int rc = check_grid(grid, bounds, &g);
    if (rc) return rc;
    if ((rc = check_planes(plane_rings, n_planes, 2, n_rings_geom, 0))) return rc;
    SSS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    const size_t ngeom = (size_t)n_rings_geom * n_det * 3, nlor = (size_t)n_det * n_det * 2;
    const size_t nout = (size_t)(n_det + 1) * (n_det / 2) * n_planes;
    double *d_lam = nullptr, *d_mu = nullptr, *d_geom = nullptr;
    int16_t* d_lor = nullptr;
    int32_t* d_pr = nullptr;
    float* d_out = nullptr;
    int64_t tally = 0;
    if ((rc = dev_alloc(&d_lam, nvox, &tally)) || (mu && (rc = dev_alloc(&d_mu, nvox, &tally))) ||
        (rc = dev_alloc(&d_geom, ngeom, &tally)) || (rc = dev_alloc(&d_lor, nlor, &tally)) ||
        (rc = dev_alloc(&d_pr, (size_t)2 * n_planes + 1, &tally)) || (rc = dev_alloc(&d_out, nout + 1, &tally))) {
        dev_free(d_lam); dev_free(d_mu); dev_free(d_geom); dev_free(d_lor); dev_free(d_pr); dev_free(d_out);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_lam, lam, nvox * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && mu) e = hipMemcpyAsync(d_mu, mu, nvox * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_geom, geom, ngeom * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lor, lor_coords, nlor * sizeof(int16_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n_planes > 0)
        e = hipMemcpyAsync(d_pr, plane_rings, (size_t)2 * n_planes * sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(d_out, 0, nout * sizeof(float), st);
    if (e == hipSuccess)
        e = launch_forward_project(g, d_lam, d_mu, d_geom, n_det, d_lor, n_planes, d_pr, accel, d_out, st);
    if (e == hipSuccess) e = hipMemcpyAsync(trues_out, d_out, nout * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_lam); dev_free(d_mu); dev_free(d_geom); dev_free(d_lor); dev_free(d_pr); dev_free(d_out);
    if (e != hipSuccess) return hip_fail(e, "sss_forward_project");
    return SSS_OK;
}

int sss_poisson(sss_ctx* ctx, const float* mean, int64_t n, double scale, uint64_t seed, float* out)
{
    if (!ctx || n < 0 || (n > 0 && (!mean || !out)) || !(scale >= 0.0))
        return fail(SSS_ERR_INVALID, "sss_poisson: bad argument");
    if (n == 0) return SSS_OK;
    SSS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    float *d_in = nullptr, *d_out = nullptr;
    int64_t tally = 0;
    int rc;
    if ((rc = dev_alloc(&d_in, (size_t)n, &tally)) || (rc = dev_alloc(&d_out, (size_t)n, &tally))) {
        dev_free(d_in); dev_free(d_out);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_in, mean, (size_t)n * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = launch_poisson(d_in, n, scale, seed, d_out, st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_in); dev_free(d_out);
    if (e != hipSuccess) return hip_fail(e, "sss_poisson");
    return SSS_OK;
}

int sss_tail_fit(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], int n_rings_geom, int n_det,
                 const double* geom, const int16_t* lor_coords, int n_entries, const int32_t* plane_rings,
                 int n_value_planes, int accel, int length_weighted, const float* scatter,
                 const float* prompts, const float* randoms, const float* mask, double* sums_out,
                 double* images_out)
{
    if (!ctx || !geom || !lor_coords || !plane_rings || !scatter || !prompts || !randoms || !mask ||
        !sums_out || n_det < 2 || n_rings_geom < 1 || n_entries < 0 || n_value_planes < 1 || accel < 1)
        return fail(SSS_ERR_INVALID, "sss_tail_fit: bad argument");
    Grid g;
    int rc = check_grid(grid, bounds, &g);
    if (rc) return rc;
    if ((rc = check_planes(plane_rings, n_entries, 3, n_rings_geom, n_value_planes))) return rc;
    SSS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const size_t nvox = (size_t)g.nx * g.ny * g.nz;
    const size_t ngeom = (size_t)n_rings_geom * n_det * 3, nlor = (size_t)n_det * n_det * 2;
    const size_t nsino = (size_t)(n_det + 1) * (n_det / 2) * n_value_planes;
    double *d_geom = nullptr, *d_img = nullptr, *d_sums = nullptr;
    int16_t* d_lor = nullptr;
    int32_t* d_pr = nullptr;
    float *d_s = nullptr, *d_p = nullptr, *d_r = nullptr, *d_m = nullptr;
    unsigned char* d_scratch = nullptr;
    int64_t tally = 0;
    if ((rc = dev_alloc(&d_geom, ngeom, &tally)) || (images_out && (rc = dev_alloc(&d_img, 4 * nvox, &tally))) ||
        (rc = dev_alloc(&d_sums, 2, &tally)) || (rc = dev_alloc(&d_lor, nlor, &tally)) ||
        (rc = dev_alloc(&d_pr, (size_t)3 * n_entries + 1, &tally)) || (rc = dev_alloc(&d_s, nsino, &tally)) ||
        (rc = dev_alloc(&d_p, nsino, &tally)) || (rc = dev_alloc(&d_r, nsino, &tally)) ||
        (rc = dev_alloc(&d_m, nsino, &tally)) || (rc = dev_alloc(&d_scratch, fit_scratch_bytes((int64_t)nvox), &tally))) {
        dev_free(d_geom); dev_free(d_img); dev_free(d_sums); dev_free(d_lor); dev_free(d_pr);
        dev_free(d_s); dev_free(d_p); dev_free(d_r); dev_free(d_m); dev_free(d_scratch);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(d_geom, geom, ngeom * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_lor, lor_coords, nlor * sizeof(int16_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess && n_entries > 0)
        e = hipMemcpyAsync(d_pr, plane_rings, (size_t)3 * n_entries * sizeof(int32_t), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_s, scatter, nsino * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_p, prompts, nsino * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_r, randoms, nsino * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(d_m, mask, nsino * sizeof(float), hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_fit(g, d_geom, n_rings_geom, n_det, d_lor, n_entries, d_pr, n_value_planes, accel, length_weighted, d_s,
                       d_p, d_r, d_m, fit_scratch_carve(d_scratch, (int64_t)nvox), d_sums, d_img, st);
    if (e == hipSuccess) e = hipMemcpyAsync(sums_out, d_sums, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && images_out)
        e = hipMemcpyAsync(images_out, d_img, 4 * nvox * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    dev_free(d_geom); dev_free(d_img); dev_free(d_sums); dev_free(d_lor); dev_free(d_pr);
    dev_free(d_s); dev_free(d_p); dev_free(d_r); dev_free(d_m); dev_free(d_scratch);
    if (e != hipSuccess) return hip_fail(e, "sss_tail_fit");
    return SSS_OK;
}

int sss_scale_to_prompts(sss_ctx* ctx, const int32_t grid[3], const double bounds[6], int n_rings_geom, int n_det,
                         const double* geom, const int16_t* lor_coords, int n_entries, const int32_t* plane_rings,
                         int n_value_planes, int accel, int length_weighted, int n_bins, const float* scatter,
                         const void* prompts, int prompts_type, const float* lor_counts, int lor_counts_per_plane,
                         const float* randoms, const float* norm, const uint8_t* mask, int flags,
                         double* sums_out)
{
    if (!ctx || !geom || !lor_coords || (n_entries > 0 && !plane_rings) || !scatter || !prompts || !mask ||
        !sums_out || n_det < 2 || n_rings_geom < 1 || n_entries < 0 || n_value_planes < 1 || accel < 1 ||
        n_bins < 1 || n_bins > SSS_MAX_FIT_BINS ||
        (prompts_type != SSS_PROMPTS_F32 && prompts_type != SSS_PROMPTS_U16) || (flags & ~SSS_FIT_DEVICE))
        return fail(SSS_ERR_INVALID, "sss_scale_to_prompts: bad argument");
    Grid g;
    int rc = check_grid(grid, bounds, &g);
    if (rc) return rc;
    const bool on_device = (flags & SSS_FIT_DEVICE) != 0;
    // device entries cannot be read here: the kernels skip out-of-range ones
    if (!on_device && (rc = check_planes(plane_rings, n_entries, 3, n_rings_geom, n_value_planes))) return rc;
    SSS_HIP(hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int64_t nvox = (int64_t)g.nx * g.ny * g.nz;
    const size_t n_cells = (size_t)(n_det + 1) * (n_det / 2) * n_value_planes;
    const size_t n_counts = lor_counts_per_plane ? (size_t)n_value_planes : n_cells;
    const size_t psize = prompts_type == SSS_PROMPTS_U16 ? sizeof(uint16_t) : sizeof(float);
    const size_t ngeom = (size_t)n_rings_geom * n_det * 3, nlor = (size_t)n_det * n_det * 2;

    BinFitIn in{};
    in.prompts_u16 = prompts_type == SSS_PROMPTS_U16;
    in.counts_per_plane = lor_counts_per_plane ? 1 : 0;
    in.n_bins = n_bins;
    in.n_cells = (int64_t)n_cells;
    in.n_value_planes = n_value_planes;
    const double* d_geom = geom;
    const int16_t* d_lor = lor_coords;
    const int32_t* d_pr = plane_rings;
    double* d_sums = sums_out;
    unsigned char* d_scratch = nullptr;
    // host mode: device copies of every input, from the context's cache
    double *h_geom = nullptr, *h_sums = nullptr;
    int16_t* h_lor = nullptr;
    int32_t* h_pr = nullptr;
    float *h_s = nullptr, *h_c = nullptr, *h_r = nullptr, *h_n = nullptr;
    unsigned char* h_p = nullptr;
    uint8_t* h_m = nullptr;
    auto release = [&]() {
        pool_free(ctx, h_geom); pool_free(ctx, h_sums); pool_free(ctx, h_lor); pool_free(ctx, h_pr);
        pool_free(ctx, h_s); pool_free(ctx, h_c); pool_free(ctx, h_r); pool_free(ctx, h_n);
        pool_free(ctx, h_p); pool_free(ctx, h_m); pool_free(ctx, d_scratch);
    };
    int64_t tally = 0;
    if ((rc = pool_alloc(ctx, &d_scratch, fit_bins_scratch_bytes(nvox, n_bins), &tally))) {
        release();
        return rc;
    }
    hipError_t e = hipSuccess;
    if (on_device) {
        in.scatter = scatter;
        in.prompts = prompts;
        in.counts = lor_counts;
        in.randoms = randoms;
        in.norm = norm;
        in.mask = mask;
    } else {
        if ((rc = pool_alloc(ctx, &h_geom, ngeom, &tally)) || (rc = pool_alloc(ctx, &h_lor, nlor, &tally)) ||
            (rc = pool_alloc(ctx, &h_pr, (size_t)3 * n_entries + 1, &tally)) ||
            (rc = pool_alloc(ctx, &h_sums, (size_t)2 * n_bins, &tally)) ||
            (rc = pool_alloc(ctx, &h_s, n_cells * n_bins, &tally)) ||
            (rc = pool_alloc(ctx, &h_p, n_cells * n_bins * psize, &tally)) ||
            (lor_counts && (rc = pool_alloc(ctx, &h_c, n_counts, &tally))) ||
            (randoms && (rc = pool_alloc(ctx, &h_r, n_cells, &tally))) ||
            (randoms && norm && (rc = pool_alloc(ctx, &h_n, n_cells, &tally))) ||
            (rc = pool_alloc(ctx, &h_m, n_cells, &tally))) {
            release();
            return rc;
        }
        auto up = [&](void* dst, const void* src, size_t bytes) {
            if (e == hipSuccess && bytes) e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st);
        };
        up(h_geom, geom, ngeom * sizeof(double));
        up(h_lor, lor_coords, nlor * sizeof(int16_t));
        if (n_entries > 0) up(h_pr, plane_rings, (size_t)3 * n_entries * sizeof(int32_t));
        up(h_s, scatter, n_cells * n_bins * sizeof(float));
        up(h_p, prompts, n_cells * n_bins * psize);
        if (h_c) up(h_c, lor_counts, n_counts * sizeof(float));
        if (h_r) up(h_r, randoms, n_cells * sizeof(float));
        if (h_n) up(h_n, norm, n_cells * sizeof(float));
        up(h_m, mask, n_cells);
        d_geom = h_geom;
        d_lor = h_lor;
        d_pr = h_pr;
        d_sums = h_sums;
        in.scatter = h_s;
        in.prompts = h_p;
        in.counts = h_c;
        in.randoms = h_r;
        in.norm = h_n;
        in.mask = h_m;
    }
    if (e == hipSuccess)
        e = launch_fit_bins(g, d_geom, n_rings_geom, n_det, d_lor, n_entries, d_pr, accel, length_weighted, in,
                            fit_bins_scratch_carve(d_scratch, nvox, n_bins), d_sums, st);
    if (!on_device) {
        if (e == hipSuccess)
            e = hipMemcpyAsync(sums_out, d_sums, (size_t)2 * n_bins * sizeof(double), hipMemcpyDeviceToHost, st);
        if (e == hipSuccess) e = hipStreamSynchronize(st);
    }
    // the blocks go back to the context's cache: their next user works on
    // this stream, after the kernels queued above
    release();
    if (e != hipSuccess) return hip_fail(e, "sss_scale_to_prompts");
    return SSS_OK;
}

}  // extern "C"

namespace sss {

namespace {
__global__ void k_fill_i32(int32_t* __restrict__ p, int64_t n, int32_t v)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
}  // namespace

hipError_t launch_fill_i32(int32_t* p, int64_t n, int32_t v, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 1024);
    hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    return hipGetLastError();
}

int subject_batch_prepare(sss_subject* s, sss_ctx* ctx, BatchMember* m)
{
    if (s->ctx != ctx) return fail(SSS_ERR_INVALID, "batch: every subject must be on the batch's context");
    if (!s->traced || s->trace_begun) return fail(SSS_ERR_INVALID, "batch: every subject must be traced");
    if (!s->has_activity || !s->h_lam)
        return fail(SSS_ERR_INVALID, "batch: every subject needs a host activity image (sss_subject_set_activity)");
    SSS_HIP(hipSetDevice(ctx->device));
    // the first two images after the walks build the fold table (the second
    // one with its record -> sample map); every later image takes the same
    // queued sequence, which is what a batch replays
    while (!s->trivial && s->refreshes < 2)
        if (int rc = refresh_fold(s)) return rc;
    const bool fp64 = s->precision == SSS_PRECISION_FP64;
    m->args = scatter_args(s);
    m->fp64 = fp64;
    m->trivial = s->trivial;
    m->variant = scatter_variant(m->args, fp64, s->trivial);
    m->n_bins = s->n_bins;
    m->n_blocks = s->n_items / kScatterWaves;
    m->out_bytes = (size_t)s->out_elems * (fp64 ? sizeof(double) : sizeof(float));
    m->d_out = s->d_out;
    // pairs x samples: one-bin subjects evaluate pairs; the others fold every
    // sample of a half path once per detector cell it pairs with
    const double cells = (double)s->nru * s->ndu;
    m->work = (s->trivial ? (double)s->n_hp : (double)s->n_samples) * cells;
    m->generation = s->generation;
    return SSS_OK;
}

int subject_generation(const sss_subject* s) { return s->generation; }

void subject_batch_ref(sss_subject* s, int delta) { s->batch_refs += delta; }

int subject_stage_activity(sss_subject* s, const double* lam)
{
    SSS_HIP(hipSetDevice(s->ctx->device));
    if (s->lam_queued) {
        SSS_HIP(hipEventSynchronize(s->ev_lam));
        s->lam_queued = false;
    }
    if (!s->h_lam) {
        void* p = nullptr;
        if (int rc = host_block(s->ctx, (size_t)s->n_vox * sizeof(double), &p)) return rc;
        s->h_lam = static_cast<double*>(p);
    }
    std::memcpy(s->h_lam, lam, (size_t)s->n_vox * sizeof(double));
    return SSS_OK;
}

int subject_queue_upload(sss_subject* s)
{
    hipStream_t st = s->ctx->stream;
    SSS_HIP(hipMemcpyAsync(s->z_fastest ? s->d_vol_in : s->d_lam, s->h_lam, (size_t)s->n_vox * sizeof(double),
                           hipMemcpyHostToDevice, st));
    if (s->z_fastest)
        SSS_HIP(launch_volume_x_fastest(s->d_vol_in, s->d_lam, s->grid.nx, s->grid.ny, s->grid.nz, st));
    s->has_activity = true;
    return SSS_OK;
}

int subject_queue_activity(sss_subject* s)
{
    if (int rc = subject_queue_upload(s)) return rc;
    return refresh_fold(s);
}

int subject_one_bin_refresh(const sss_subject* s, OneBinRefresh* r)
{
    if (!s->trivial) return fail(SSS_ERR_INVALID, "batch: a merged one-bin refresh needs a one-bin subject");
    r->t = s->hp;
    r->s = s->smp;
    r->n_hp = s->n_hp;
    r->lam = s->d_lam;
    return SSS_OK;
}

}  // namespace sss
