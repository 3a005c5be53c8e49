// Batches of traced subjects on one context (sss_batch_* in sss_hip.h).
//
// A batch launch runs every member's activity upload and refresh (the
// one-bin members' refreshes merged into one launch), then one
// persistent scatter launch per kernel variant present (k_scatter_queue: the
// members' workgroup blocks in 8 per-XCD queues), one chunk-reduce launch per
// precision and the output downloads.  sss_batch_graph_launch captures that
// sequence in a HIP graph once and replays it: the work of a launch is fixed
// by the members' tables, and only the staged activity images (pinned host
// memory that the graph's upload nodes read) change between launches.  The
// refresh inside the sequence decides on the device whether a new image
// needs the fold table rebuilt (the member's flags), so the captured kernels
// serve every image.  Flags and queue counters are cleared by kernels: the
// graph holds no memset node.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "sss_internal.h"

using namespace sss;

struct sss_batch {
    struct Group {
        int rep = -1;        // a member of this kernel variant
        int max_bins = 0;
        int n_blocks = 0;
        int grid = 0;
        BatchQueue q{};
    };
    struct Reduce {
        bool fp64 = false;
        int n_blocks = 0;
        int max_bins = 0;
        const int2* blocks = nullptr;
    };
    sss_ctx* ctx = nullptr;
    std::vector<sss_subject*> members;
    std::vector<BatchMember> m;
    std::vector<Group> groups;
    std::vector<Reduce> reduces;
    const ScatterArgs* d_args = nullptr;
    const OneBinRefresh* d_one_bin = nullptr;   // one-bin members' tables
    const int2* d_one_bin_blocks = nullptr;     // {one-bin member, block of 64 half paths}
    int n_one_bin_blocks = 0;
    int32_t* d_counters = nullptr;
    int n_counters = 0;
    void* d_mem = nullptr;           // args, queues, reduce maps, counters
    std::vector<void*> h_out;        // pinned compact outputs
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool launched = false;
    bool built = false;

    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
};

namespace {

bool stale(const sss_batch* b)
{
    if (!b->built) return true;
    for (size_t i = 0; i < b->members.size(); ++i)
        if (subject_generation(b->members[i]) != b->m[i].generation) return true;
    return false;
}

// Member descriptions, kernel-variant groups, queues and reduce maps; the
// device block holding them is written once here (synchronously: a rebuild
// first waits for the launches that read the old one).
int build(sss_batch* b)
{
    hipStream_t st = ctx_stream(b->ctx);
    SSS_HIP(hipSetDevice(ctx_device(b->ctx)));
    SSS_HIP(hipStreamSynchronize(st));
    b->drop_graph();
    b->built = false;
    const int n = (int)b->members.size();
    b->m.assign(n, BatchMember{});
    for (int i = 0; i < n; ++i)
        if (int rc = subject_batch_prepare(b->members[i], b->ctx, &b->m[i])) return rc;

    // kernel variants in order of first appearance
    std::vector<std::vector<int>> of_variant;
    std::map<int, int> variant_index;
    for (int i = 0; i < n; ++i) {
        auto it = variant_index.find(b->m[i].variant);
        if (it == variant_index.end()) {
            it = variant_index.emplace(b->m[i].variant, (int)of_variant.size()).first;
            of_variant.emplace_back();
        }
        of_variant[it->second].push_back(i);
    }
    // host image of the device block: args | per group q_begin | blocks |
    // reduce maps | counters
    std::vector<unsigned char> img;
    auto place = [&img](size_t bytes) {
        const size_t off = (img.size() + 255) / 256 * 256;
        img.resize(off + bytes);
        return off;
    };
    const size_t args_off = place((size_t)n * sizeof(ScatterArgs));
    for (int i = 0; i < n; ++i) std::memcpy(img.data() + args_off + i * sizeof(ScatterArgs), &b->m[i].args, sizeof(ScatterArgs));
    struct GroupPlan {
        size_t qb_off, blk_off;
        std::vector<int32_t> q_begin;
        std::vector<int2> blocks;
    };
    std::vector<GroupPlan> plans(of_variant.size());
    b->groups.assign(of_variant.size(), sss_batch::Group{});
    for (size_t g = 0; g < of_variant.size(); ++g) {
        std::vector<int> order = of_variant[g];
        // most work per workgroup block first (a dynamic queue then ends on
        // the short blocks)
        auto per_block = [&](int i) { return b->m[i].work / std::max(1, b->m[i].n_blocks); };
        std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return per_block(x) > per_block(y); });
        GroupPlan& p = plans[g];
        p.q_begin.assign(kXcdCount + 1, 0);
        for (int x = 0; x < kXcdCount; ++x) {
            p.q_begin[x] = (int32_t)p.blocks.size();
            for (int i : order)
                for (int j = x; j < b->m[i].n_blocks; j += kXcdCount) p.blocks.push_back(make_int2(i, j));
        }
        p.q_begin[kXcdCount] = (int32_t)p.blocks.size();
        p.qb_off = place((kXcdCount + 1) * sizeof(int32_t));
        p.blk_off = place(std::max<size_t>(1, p.blocks.size()) * sizeof(int2));
        std::memcpy(img.data() + p.qb_off, p.q_begin.data(), p.q_begin.size() * sizeof(int32_t));
        if (!p.blocks.empty()) std::memcpy(img.data() + p.blk_off, p.blocks.data(), p.blocks.size() * sizeof(int2));
        sss_batch::Group& G = b->groups[g];
        G.rep = order.front();
        G.n_blocks = (int)p.blocks.size();
        for (int i : order) G.max_bins = std::max(G.max_bins, b->m[i].n_bins);
    }
    // chunk reduces, one launch per precision
    std::vector<std::pair<size_t, std::vector<int2>>> rplans;
    b->reduces.clear();
    for (int fp64 = 0; fp64 < 2; ++fp64) {
        std::vector<int2> blocks;
        int max_bins = 0;
        for (int i = 0; i < n; ++i) {
            if (b->m[i].fp64 != (fp64 != 0)) continue;
            const int nb = reduce_blocks(b->m[i].args);
            for (int k = 0; k < nb; ++k) blocks.push_back(make_int2(i, k));
            max_bins = std::max(max_bins, b->m[i].n_bins);
        }
        if (blocks.empty()) continue;
        const size_t off = place(blocks.size() * sizeof(int2));
        std::memcpy(img.data() + off, blocks.data(), blocks.size() * sizeof(int2));
        sss_batch::Reduce r;
        r.fp64 = fp64 != 0;
        r.n_blocks = (int)blocks.size();
        r.max_bins = max_bins;
        b->reduces.push_back(r);
        rplans.emplace_back(off, std::move(blocks));
    }
    // one-bin members' activity refreshes, merged into one launch
    std::vector<OneBinRefresh> one_bin;
    std::vector<int2> one_bin_blocks;
    for (int i = 0; i < n; ++i) {
        if (!b->m[i].trivial) continue;
        OneBinRefresh r;
        if (int rc = subject_one_bin_refresh(b->members[i], &r)) return rc;
        const int64_t nb = (r.n_hp + kWave - 1) / kWave;
        for (int64_t k = 0; k < nb; ++k) one_bin_blocks.push_back(make_int2((int)one_bin.size(), (int)k));
        one_bin.push_back(r);
    }
    const size_t ob_off = place(std::max<size_t>(1, one_bin.size()) * sizeof(OneBinRefresh));
    const size_t obb_off = place(std::max<size_t>(1, one_bin_blocks.size()) * sizeof(int2));
    if (!one_bin.empty()) std::memcpy(img.data() + ob_off, one_bin.data(), one_bin.size() * sizeof(OneBinRefresh));
    if (!one_bin_blocks.empty())
        std::memcpy(img.data() + obb_off, one_bin_blocks.data(), one_bin_blocks.size() * sizeof(int2));
    b->n_counters = (int)b->groups.size() * kXcdCount;
    const size_t cnt_off = place(std::max(1, b->n_counters) * sizeof(int32_t));

    if (b->d_mem) ctx_pool_put(b->ctx, b->d_mem);
    b->d_mem = nullptr;
    if (int rc = ctx_pool_get(b->ctx, img.size(), &b->d_mem)) return rc;
    SSS_HIP(hipMemcpyAsync(b->d_mem, img.data(), img.size(), hipMemcpyHostToDevice, st));
    SSS_HIP(hipStreamSynchronize(st));
    unsigned char* base = static_cast<unsigned char*>(b->d_mem);
    b->d_args = reinterpret_cast<const ScatterArgs*>(base + args_off);
    b->d_one_bin = reinterpret_cast<const OneBinRefresh*>(base + ob_off);
    b->d_one_bin_blocks = reinterpret_cast<const int2*>(base + obb_off);
    b->n_one_bin_blocks = (int)one_bin_blocks.size();
    b->d_counters = reinterpret_cast<int32_t*>(base + cnt_off);
    for (size_t r = 0; r < b->reduces.size(); ++r)
        b->reduces[r].blocks = reinterpret_cast<const int2*>(base + rplans[r].first);

    int n_cu = 0;
    SSS_HIP(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, ctx_device(b->ctx)));
    for (size_t g = 0; g < b->groups.size(); ++g) {
        sss_batch::Group& G = b->groups[g];
        const BatchMember& rep = b->m[G.rep];
        G.q.args = b->d_args;
        G.q.blocks = reinterpret_cast<const int2*>(base + plans[g].blk_off);
        G.q.q_begin = reinterpret_cast<const int32_t*>(base + plans[g].qb_off);
        G.q.counters = b->d_counters + g * kXcdCount;
        SSS_HIP(scatter_queue_grid(rep.args, rep.fp64, rep.trivial, G.max_bins, n_cu, G.n_blocks, &G.grid));
    }
    b->built = true;
    return SSS_OK;
}

// The launch sequence on the context stream (queued directly, or recorded
// into the graph under capture).
int queue_all(sss_batch* b)
{
    hipStream_t st = ctx_stream(b->ctx);
    for (size_t i = 0; i < b->members.size(); ++i) {
        // one-bin members: the upload here, their refreshes in one launch below
        const int rc = b->m[i].trivial ? subject_queue_upload(b->members[i]) : subject_queue_activity(b->members[i]);
        if (rc) return rc;
    }
    SSS_HIP(launch_one_bin_refresh(b->d_one_bin, b->d_one_bin_blocks, b->n_one_bin_blocks, st));
    SSS_HIP(launch_fill_i32(b->d_counters, b->n_counters, 0, st));
    for (const sss_batch::Group& G : b->groups) {
        if (G.n_blocks == 0) continue;
        const BatchMember& rep = b->m[G.rep];
        SSS_HIP(launch_scatter_queue(rep.args, rep.fp64, rep.trivial, G.q, G.max_bins, G.grid, st));
    }
    for (const sss_batch::Reduce& r : b->reduces)
        SSS_HIP(launch_chunk_reduce_queue(b->d_args, r.blocks, r.n_blocks, r.max_bins, r.fp64, st));
    for (size_t i = 0; i < b->members.size(); ++i)
        SSS_HIP(hipMemcpyAsync(b->h_out[i], b->m[i].d_out, b->m[i].out_bytes, hipMemcpyDeviceToHost, st));
    return SSS_OK;
}

int check_batch(sss_batch* b, const char* what)
{
    if (!b) return fail(SSS_ERR_INVALID, std::string(what) + ": NULL batch");
    SSS_HIP(hipSetDevice(ctx_device(b->ctx)));
    return SSS_OK;
}

int wait_last(sss_batch* b)
{
    if (b->launched) SSS_HIP(hipEventSynchronize(b->ev[1]));
    return SSS_OK;
}

}  // namespace

extern "C" {

int sss_batch_create(sss_ctx* ctx, sss_subject* const* members, int n_members, sss_batch** out)
{
    if (!ctx || !members || !out || n_members <= 0) return fail(SSS_ERR_INVALID, "sss_batch_create: bad argument");
    *out = nullptr;
    for (int i = 0; i < n_members; ++i) {
        if (!members[i]) return fail(SSS_ERR_INVALID, "sss_batch_create: NULL member");
        for (int j = 0; j < i; ++j)
            if (members[j] == members[i]) return fail(SSS_ERR_INVALID, "sss_batch_create: a subject is listed twice");
    }
    SSS_HIP(hipSetDevice(ctx_device(ctx)));
    sss_batch* b = new (std::nothrow) sss_batch;
    if (!b) return fail(SSS_ERR_NOMEM, "sss_batch_create: out of host memory");
    b->ctx = ctx;
    b->members.assign(members, members + n_members);
    for (sss_subject* s : b->members) subject_batch_ref(s, +1);
    int rc = build(b);
    hipError_t e = hipSuccess;
    if (!rc) {
        e = hipEventCreate(&b->ev[0]);
        if (e == hipSuccess) e = hipEventCreate(&b->ev[1]);
        if (e != hipSuccess) rc = hip_fail(e, "sss_batch_create: events");
    }
    for (int i = 0; !rc && i < n_members; ++i) {
        void* p = nullptr;
        rc = sss_host_alloc(ctx, (int64_t)b->m[i].out_bytes, &p);
        if (!rc) b->h_out.push_back(p);
    }
    if (rc) {
        const std::string msg = sss_last_error();
        (void)sss_batch_destroy(b);
        return fail(rc, msg);
    }
    *out = b;
    return SSS_OK;
}

int sss_batch_set_activity(sss_batch* b, int i, const double* lam)
{
    if (int rc = check_batch(b, "sss_batch_set_activity")) return rc;
    if (i < 0 || i >= (int)b->members.size() || !lam) return fail(SSS_ERR_INVALID, "sss_batch_set_activity: bad member or image");
    if (int rc = wait_last(b)) return rc;
    return subject_stage_activity(b->members[i], lam);
}

int sss_batch_submit(sss_batch* b)
{
    if (int rc = check_batch(b, "sss_batch_submit")) return rc;
    if (stale(b))
        if (int rc = build(b)) return rc;
    hipStream_t st = ctx_stream(b->ctx);
    SSS_HIP(hipEventRecord(b->ev[0], st));
    const int rc = queue_all(b);
    SSS_HIP(hipEventRecord(b->ev[1], st));
    b->launched = true;
    return rc;
}

int sss_batch_graph_launch(sss_batch* b)
{
    if (int rc = check_batch(b, "sss_batch_graph_launch")) return rc;
    if (stale(b))
        if (int rc = build(b)) return rc;
    hipStream_t st = ctx_stream(b->ctx);
    if (!b->exec) {
        SSS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
        const int rc = queue_all(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(st, &g);   // always ends the capture
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return hip_fail(e, "sss_batch_graph_launch: capture");
        b->graph = g;
        SSS_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    SSS_HIP(hipEventRecord(b->ev[0], st));
    SSS_HIP(hipGraphLaunch(b->exec, st));
    SSS_HIP(hipEventRecord(b->ev[1], st));
    b->launched = true;
    return SSS_OK;
}

int sss_batch_wait(sss_batch* b)
{
    if (int rc = check_batch(b, "sss_batch_wait")) return rc;
    return wait_last(b);
}

int sss_batch_output(sss_batch* b, int i, void* host_out)
{
    if (int rc = check_batch(b, "sss_batch_output")) return rc;
    if (i < 0 || i >= (int)b->members.size() || !host_out) return fail(SSS_ERR_INVALID, "sss_batch_output: bad member or buffer");
    if (!b->launched) return fail(SSS_ERR_INVALID, "sss_batch_output: no launch yet");
    if (int rc = wait_last(b)) return rc;
    std::memcpy(host_out, b->h_out[i], b->m[i].out_bytes);
    return SSS_OK;
}

int sss_batch_elapsed_ms(sss_batch* b, float* ms)
{
    if (int rc = check_batch(b, "sss_batch_elapsed_ms")) return rc;
    if (!ms || !b->launched) return fail(SSS_ERR_INVALID, "sss_batch_elapsed_ms: no launch yet");
    SSS_HIP(hipEventSynchronize(b->ev[1]));
    SSS_HIP(hipEventElapsedTime(ms, b->ev[0], b->ev[1]));
    return SSS_OK;
}

int sss_batch_info(sss_batch* b, int32_t* n_scatter_launches, int32_t* graph_ready)
{
    if (int rc = check_batch(b, "sss_batch_info")) return rc;
    if (n_scatter_launches) {
        int32_t k = 0;
        for (const auto& G : b->groups) k += G.n_blocks > 0;
        *n_scatter_launches = k;
    }
    if (graph_ready) *graph_ready = (b->exec && !stale(b)) ? 1 : 0;
    return SSS_OK;
}

int sss_batch_destroy(sss_batch* b)
{
    if (!b) return SSS_OK;
    (void)hipSetDevice(ctx_device(b->ctx));
    (void)hipStreamSynchronize(ctx_stream(b->ctx));
    b->drop_graph();
    for (hipEvent_t& e : b->ev)
        if (e) (void)hipEventDestroy(e);
    for (void* p : b->h_out) (void)sss_host_free(b->ctx, p);
    if (b->d_mem) ctx_pool_put(b->ctx, b->d_mem);
    for (sss_subject* s : b->members) subject_batch_ref(s, -1);
    delete b;
    return SSS_OK;
}

}  // extern "C"
