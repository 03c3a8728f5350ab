// Batched reconstruction: many scans of uneven sizes in one hipGraph.
//
// Each scan is the chain of the single-scan entry points -- preprocess,
// filter, stage, accumulate -- run over chunks of views, then finalised.
// All scans are captured once into one graph whose parallel branches each
// own a set of chunk buffers; replaying the graph recomputes every volume
// from whatever the bound input buffers hold at launch time.  The graph
// contains kernel nodes and event-ordering edges only (the accumulators are
// cleared by a fill kernel, never by a memset node).
//
// The reference runs one scan per process through its agent graph
// (JEngine/JEngine.cpp:185-313, parts loop :400-425); batching is an
// extension with no counterpart there.
#include "jhip_bp_plan.h"
#include "jhip_internal.h"

#include <algorithm>
#include <numeric>

namespace jh {

struct GraphScan {
    jh_pre_plan* pre;
    jh_filter_plan* filt;
    jh_bp_plan* bp;
    const uint16_t* raw;
    const uint16_t* dark;
    const float* gain;
    int dtype;
    int chunk;  // views per chunk (0 = not chosen yet)
    float* vol;
    int branch;
    bool operand;  // every chunk takes the operand path (decided once, at instantiation)
};

struct GraphBranch {
    DevBuf pre, filt, q;  // chunk buffers, sized for the largest chunk of the branch
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
};

static void check_status(jh_status s, const char* what)
{
    if (s != JH_OK) {
        const std::string m = jh_last_error();
        fail(s, std::string(what) + ": " + m);
    }
}

// Whether a chunk of n views of the scan can take the operand path: filter
// and staging fuse, and the preprocessing neither resizes nor changes the
// frame.  The row limit is monotonic in n, so when the scan's full chunk
// qualifies every shorter chunk (the last, N % chunk) does too; the graph
// decides once per scan from the full chunk and sizes the branch's operand
// buffer from that decision, so no chunk writes an operand the buffer was
// not sized for.
static bool operand_path(const GraphScan& s, int n)
{
    int fused = 0;
    check_status(jh_filter_stage_fusable(s.filt, s.bp, s.dtype, &fused), "fusable");
    const jh_pre_plan* p = s.pre;
    const int roiW = p->inW - p->bLeft - p->bRight, roiH = p->inH - p->bUp - p->bDown;
    return fused && roiW == p->outW && roiH == p->outH && p->outW == s.bp->W && p->outH == s.bp->V &&
           (size_t)s.bp->V * n <= 1048560u;
}

static double scan_cost(const GraphScan& s)
{
    // voxel-updates dominate; the per-pixel chain adds a term per detector pixel
    const double vox = (double)s.bp->X * s.bp->Y * s.bp->Z;
    return (vox + 64.0 * s.bp->W * s.bp->V) * s.bp->N;
}

}  // namespace jh

struct jh_graph {
    int device;
    std::vector<jh::GraphScan> scans;
    std::vector<jh::GraphBranch> branches;
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;

    void release_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        for (auto& b : branches) {
            if (b.done) (void)hipEventDestroy(b.done);
            if (b.stream) (void)hipStreamDestroy(b.stream);
        }
        branches.clear();
        if (fork) (void)hipEventDestroy(fork);
        if (origin) (void)hipStreamDestroy(origin);
        fork = nullptr;
        origin = nullptr;
    }
    ~jh_graph() { release_graph(); }
};

using namespace jh;

extern "C" {

jh_status jh_graph_create(jh_graph** g, int device)
{
    return guard([&] {
        require(g != nullptr, "jh_graph_create: handle pointer is NULL");
        *g = nullptr;
        int n = 0;
        JH_HIP(hipGetDeviceCount(&n));
        require(device >= 0 && device < n, "jh_graph_create: no such device");
        auto* p = new jh_graph();
        p->device = device;
        *g = p;
    });
}

jh_status jh_graph_add_scan(jh_graph* g, jh_pre_plan* pre, jh_filter_plan* filt, jh_bp_plan* bp,
                            const uint16_t* d_raw, const uint16_t* d_dark, const float* d_gain, int dtype,
                            int viewsPerChunk, float* d_vol, int* index)
{
    return guard([&] {
        require(g != nullptr && pre != nullptr && filt != nullptr && bp != nullptr,
                "jh_graph_add_scan: NULL handle");
        require(d_raw != nullptr && d_vol != nullptr, "jh_graph_add_scan: NULL buffer");
        require(g->exec == nullptr, "jh_graph_add_scan: the graph is already instantiated");
        require(dtype == JH_F32 || dtype == JH_BF16 || dtype == JH_F16, "jh_graph_add_scan: unknown dtype");
        require(viewsPerChunk >= 0, "jh_graph_add_scan: negative chunk");
        int fdev, fW, fV, fN;
        filter_plan_shape(filt, &fdev, &fW, &fV, &fN);
        require(pre->device == g->device && fdev == g->device && bp->device == g->device,
                "jh_graph_add_scan: plans live on another device");
        require(pre->outW == bp->W && pre->outH == bp->V, "jh_graph_add_scan: preprocess output != detector");
        require(fW == bp->W && fV == bp->V && fN == bp->N, "jh_graph_add_scan: filter plan does not match");
        for (const auto& s : g->scans)
            require(s.bp != bp && s.vol != d_vol, "jh_graph_add_scan: plan or volume used twice");
        GraphScan s{pre, filt, bp, d_raw, d_dark, d_gain, dtype, std::min(viewsPerChunk, bp->N), d_vol, -1, false};
        g->scans.push_back(s);
        if (index) *index = (int)g->scans.size() - 1;
    });
}

jh_status jh_graph_scan_tuning(jh_graph* g, int index, int* viewsPerChunk, int* zc, int* ty)
{
    return guard([&] {
        require(g != nullptr && index >= 0 && index < (int)g->scans.size(), "jh_graph_scan_tuning: bad scan");
        const GraphScan& s = g->scans[index];
        if (viewsPerChunk) *viewsPerChunk = s.chunk;
        if (zc) *zc = s.bp->zc;
        if (ty) *ty = s.bp->ty;
    });
}

jh_status jh_graph_autotune(jh_graph* g, int reps, jh_stream_t stream)
{
    return guard([&] {
        require(g != nullptr, "jh_graph_autotune: graph is NULL");
        require(g->exec == nullptr, "jh_graph_autotune: the graph is already instantiated");
        require(reps >= 1 && reps <= 100, "jh_graph_autotune: reps must be 1..100");
        DeviceScope ds(g->device);
        hipStream_t st = as_stream(stream);
        hipEvent_t e0, e1;
        JH_HIP(hipEventCreate(&e0));
        JH_HIP(hipEventCreate(&e1));
        try {
            for (auto& s : g->scans) {
                // The backprojector's timing depends on the geometry and the
                // tile, not on the sample values: time it on a zero-filled
                // staging buffer that spans the whole orbit.
                size_t qb = 0;
                check_status(jh_bp_stage_bytes(s.bp, s.dtype, s.bp->N, &qb), "stage bytes");
                DevBuf q(qb);
                JH_HIP(hipMemsetAsync(q.get(), 0, qb, st));
                check_status(jh_bp_autotune(s.bp, q.get(), s.dtype, s.bp->N, s.vol, reps, stream, nullptr),
                             "tile autotune");
                if (s.chunk > 0) continue;
                // View chunk: time accumulation per view over candidate chunk
                // sizes; each chunk costs one read-modify-write of the
                // accumulators, so larger chunks are never slower per view.
                // Keep the smallest chunk within 3 % of the best (less scratch
                // per branch, finer interleaving of the branches).
                check_status(jh_bp_reset(s.bp, stream), "reset");
                std::vector<int> cands;
                for (int c = 16; c < s.bp->N && c <= 256; c *= 2) cands.push_back(c);
                cands.push_back(std::min(s.bp->N, 512));
                std::vector<float> perView;
                for (int c : cands) {
                    check_status(jh_bp_accumulate(s.bp, q.get(), s.dtype, 0, c, stream), "warm-up");
                    JH_HIP(hipEventRecord(e0, st));
                    for (int r = 0; r < reps; ++r)
                        check_status(jh_bp_accumulate(s.bp, q.get(), s.dtype, 0, c, stream), "accumulate");
                    JH_HIP(hipEventRecord(e1, st));
                    JH_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    JH_HIP(hipEventElapsedTime(&ms, e0, e1));
                    perView.push_back(ms / (float)(reps * c));
                }
                const float best = *std::min_element(perView.begin(), perView.end());
                for (size_t i = 0; i < cands.size(); ++i)
                    if (perView[i] <= 1.03f * best) { s.chunk = cands[i]; break; }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

jh_status jh_graph_instantiate(jh_graph* g, int nbranches)
{
    return guard([&] {
        require(g != nullptr, "jh_graph_instantiate: graph is NULL");
        require(g->exec == nullptr, "jh_graph_instantiate: the graph is already instantiated");
        require(!g->scans.empty(), "jh_graph_instantiate: no scans");
        require(nbranches >= 1 && nbranches <= 16, "jh_graph_instantiate: branches must be 1..16");
        DeviceScope ds(g->device);
        const int nb = std::min<int>(nbranches, (int)g->scans.size());
        for (auto& s : g->scans) {
            if (s.chunk <= 0) s.chunk = std::min(s.bp->N, 128);
            s.operand = operand_path(s, s.chunk);
        }

        // Longest-first greedy assignment of scans to branches.
        std::vector<int> order(g->scans.size());
        std::iota(order.begin(), order.end(), 0);
        std::sort(order.begin(), order.end(),
                  [&](int a, int b) { return scan_cost(g->scans[a]) > scan_cost(g->scans[b]); });
        std::vector<double> load(nb, 0.0);
        for (int i : order) {
            const int b = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            g->scans[i].branch = b;
            load[b] += scan_cost(g->scans[i]);
        }

        try {
            g->branches.resize(nb);
            JH_HIP(hipStreamCreateWithFlags(&g->origin, hipStreamNonBlocking));
            JH_HIP(hipEventCreateWithFlags(&g->fork, hipEventDisableTiming));
            for (int b = 0; b < nb; ++b) {
                GraphBranch& br = g->branches[b];
                size_t pb = 0, fb = 0, qb = 0;
                for (const auto& s : g->scans) {
                    if (s.branch != b) continue;
                    const size_t frame = (size_t)s.bp->W * s.bp->V * sizeof(float);
                    pb = std::max(pb, frame * s.chunk);
                    if (s.operand) {  // the pre buffer holds the filter's operand of a full chunk
                        size_t ob = 0;
                        check_status(jh_filter_operand_bytes(s.filt, s.chunk, &ob), "operand bytes");
                        pb = std::max(pb, ob);
                    }
                    fb = std::max(fb, frame * s.chunk);
                    size_t q = 0;
                    check_status(jh_bp_stage_bytes(s.bp, s.dtype, s.chunk, &q), "stage bytes");
                    qb = std::max(qb, q);
                }
                br.pre.alloc(pb);
                br.filt.alloc(fb);
                br.q.alloc(qb);
                JH_HIP(hipStreamCreateWithFlags(&br.stream, hipStreamNonBlocking));
                JH_HIP(hipEventCreateWithFlags(&br.done, hipEventDisableTiming));
            }
            // Accumulators are allocated on first reset; do that before the
            // capture so no allocation happens inside it.
            for (auto& s : g->scans) check_status(jh_bp_reset(s.bp, g->origin), "reset");
            JH_HIP(hipStreamSynchronize(g->origin));

            JH_HIP(hipStreamBeginCapture(g->origin, hipStreamCaptureModeThreadLocal));
            try {
                JH_HIP(hipEventRecord(g->fork, g->origin));
                for (auto& br : g->branches) JH_HIP(hipStreamWaitEvent(br.stream, g->fork, 0));
                for (int i : order) {
                    GraphScan& s = g->scans[i];
                    GraphBranch& br = g->branches[s.branch];
                    jh_stream_t st = br.stream;
                    const size_t inFrame = (size_t)s.pre->inW * s.pre->inH;
                    check_status(jh_bp_reset(s.bp, st), "reset");
                    for (int v0 = 0; v0 < s.bp->N; v0 += s.chunk) {
                        const int n = std::min(s.chunk, s.bp->N - v0);
                        int fused = 0;
                        check_status(jh_filter_stage_fusable(s.filt, s.bp, s.dtype, &fused), "fusable");
                        if (s.operand) {
                            // preprocessing into the filter's pre-weighted operand, then filter and staging
                            check_status(jh_pre_run_operand(s.pre, s.filt, s.raw + inFrame * v0, s.dark, s.gain, v0, n,
                                                            br.pre.as<float>(), st),
                                         "preprocess into the operand");
                            check_status(jh_filter_run_operand_staged(s.filt, s.bp, br.pre.as<float>(), v0, n, s.dtype,
                                                                      br.q.get(), st),
                                         "filter and stage");
                            check_status(jh_bp_accumulate(s.bp, br.q.get(), s.dtype, v0, n, st), "accumulate");
                            continue;
                        }
                        check_status(jh_pre_run(s.pre, s.raw + inFrame * v0, s.dark, s.gain, br.pre.as<float>(), n, st),
                                     "preprocess");
                        if (fused) {  // filter and staging in one kernel
                            check_status(jh_filter_run_staged(s.filt, s.bp, br.pre.as<float>(), v0, n, s.dtype,
                                                              br.q.get(), st),
                                         "filter and stage");
                        } else {
                            check_status(jh_filter_run(s.filt, br.pre.as<float>(), br.filt.as<float>(), v0, n, st),
                                         "filter");
                            check_status(jh_bp_stage(s.bp, br.filt.as<float>(), v0, n, s.dtype, br.q.get(), st),
                                         "stage");
                        }
                        check_status(jh_bp_accumulate(s.bp, br.q.get(), s.dtype, v0, n, st), "accumulate");
                    }
                    check_status(jh_bp_finalize(s.bp, s.vol, 0, s.bp->Z, st), "finalize");
                }
                for (auto& br : g->branches) {
                    JH_HIP(hipEventRecord(br.done, br.stream));
                    JH_HIP(hipStreamWaitEvent(g->origin, br.done, 0));
                }
            } catch (...) {
                hipGraph_t broken = nullptr;
                (void)hipStreamEndCapture(g->origin, &broken);
                if (broken) (void)hipGraphDestroy(broken);
                throw;
            }
            JH_HIP(hipStreamEndCapture(g->origin, &g->graph));
            JH_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
        } catch (...) {
            g->release_graph();
            throw;
        }
    });
}

jh_status jh_graph_launch(jh_graph* g, jh_stream_t stream)
{
    return guard([&] {
        require(g != nullptr, "jh_graph_launch: graph is NULL");
        require(g->exec != nullptr, "jh_graph_launch: call jh_graph_instantiate first");
        DeviceScope ds(g->device);
        JH_HIP(hipGraphLaunch(g->exec, as_stream(stream)));
    });
}

jh_status jh_graph_destroy(jh_graph* g)
{
    return guard([&] {
        if (g == nullptr) return;
        DeviceScope ds(g->device);
        delete g;
    });
}

}  // extern "C"
