// Batched transforms: many NTTs of mixed fields, sizes and directions.
//
// The reference has no batch path (one synchronous best_fft per vector,
// cuda/fft_kernel.cu:117-147).  Here every transform contributes its passes
// to fused launches: pass level p of every transform of one field (and, for
// the 32-bit field, one tile shape) runs as ONE multi-job kernel launch
// whose blocks map to (transform, tile) through a tile table.  Levels run in
// order, so pass p of a transform always sees its pass p-1 complete.  The
// whole schedule is captured once into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "batch_impl.h"

namespace gsn {

// Per-field job collection: level -> shape -> jobs.
template <int FID>
struct Collector {
    using F = typename FieldInfo<FID>::Dev;
    std::map<std::tuple<int, int, int, int>, std::vector<PassJob<F>>> by_key;  // (level, r, w, e)

    void add(const PlanImpl<FID> &plan, const void *in, void *out, void *ws0, void *ws1)
    {
        std::vector<PassJob<F>> jobs;
        plan.make_jobs(in, out, ws0, ws1, jobs);
        for (size_t lvl = 0; lvl < jobs.size(); ++lvl) {
            const PassCfg &c = plan.cfg[lvl];
            // 753-bit tiles take their shape at run time: one group per level
            // and workgroup size (tile class)
            const auto key = FieldInfo<FID>::Dev::LIMBS > 1
                                 ? std::make_tuple((int)lvl, t753_log_tile(c), 0, 1)
                                 : std::make_tuple((int)lvl, c.log_r, c.log_w, c.log_e);
            by_key[key].push_back(jobs[lvl]);
        }
    }

    int emit(std::vector<Group> &out, const BatchMemory &mem)
    {
        for (auto &kv : by_key) {
            std::vector<PassJob<F>> &jobs = kv.second;
            Group g;
            g.field = FID;
            g.level = std::get<0>(kv.first);
            g.cfg = {std::get<1>(kv.first), std::get<2>(kv.first), std::get<3>(kv.first)};
            for (size_t i = 0; i < jobs.size(); ++i) {
                jobs[i].first_tile = (uint32_t)g.host_tiles.size();
                g.host_tiles.insert(g.host_tiles.end(), jobs[i].num_tiles, (uint32_t)i);
            }
            g.total_tiles = (uint32_t)g.host_tiles.size();
            if (g.total_tiles == 0)
                continue;
            const size_t jb = jobs.size() * sizeof(PassJob<F>), tb = g.host_tiles.size() * sizeof(uint32_t);
            g.host_jobs.assign((const char *)jobs.data(), (const char *)jobs.data() + jb);
            void *dj, *dt;
            if (mem.alloc) {
                dj = mem.alloc(mem.ctx, jb);
                dt = mem.alloc(mem.ctx, tb);
                if (!dj || !dt)
                    return set_error(GSN_ERR_OUT_OF_MEMORY, "batch arena exhausted");
                GSN_TRY(hip_status(hipMemcpyAsync(dj, g.host_jobs.data(), jb, hipMemcpyHostToDevice, mem.stream),
                                   "hipMemcpyAsync(jobs)"));
                GSN_TRY(hip_status(hipMemcpyAsync(dt, g.host_tiles.data(), tb, hipMemcpyHostToDevice, mem.stream),
                                   "hipMemcpyAsync(tile map)"));
            } else {
                GSN_TRY(g.jobs_buf.alloc(jb));
                GSN_TRY(g.tile_buf.alloc(tb));
                dj = g.jobs_buf.ptr;
                dt = g.tile_buf.ptr;
                GSN_TRY(hip_status(hipMemcpy(dj, g.host_jobs.data(), jb, hipMemcpyHostToDevice), "hipMemcpy(jobs)"));
                GSN_TRY(hip_status(hipMemcpy(dt, g.host_tiles.data(), tb, hipMemcpyHostToDevice),
                                   "hipMemcpy(tile map)"));
            }
            g.jobs = dj;
            g.tile_job = (const uint32_t *)dt;
            out.push_back(std::move(g));
        }
        return GSN_OK;
    }
};

namespace {
template <class F>
void launch_group(const Group &g, hipStream_t s)
{
    launch_pass_multi<F>((const PassJob<F> *)g.jobs, g.tile_job, g.total_tiles, g.cfg, s);
}
}  // namespace

BatchImpl::~BatchImpl()
{
    if (exec)
        (void)hipGraphExecDestroy(exec);
    if (graph)
        (void)hipGraphDestroy(graph);
    if (capture_stream)
        (void)hipStreamDestroy(capture_stream);
}

int BatchImpl::enqueue(hipStream_t s)
{
    for (const Copy &c : copies) {
        switch (c.field) {
        case GSN_FIELD_FP32: launch_scale<F32>((const uint32_t *)c.in, (uint32_t *)c.out, {}, 1, 0, s); break;
        case GSN_FIELD_MNT4753_FR: launch_scale<FMnt4>((const E753 *)c.in, (E753 *)c.out, {}, 1, 0, s); break;
        default: launch_scale<FMnt6>((const E753 *)c.in, (E753 *)c.out, {}, 1, 0, s); break;
        }
    }
    for (const Group &g : groups) {
        switch (g.field) {
        case GSN_FIELD_FP32: launch_group<F32>(g, s); break;
        case GSN_FIELD_MNT4753_FR: launch_group<FMnt4>(g, s); break;
        default: launch_group<FMnt6>(g, s); break;
        }
        GSN_TRY(hip_status(hipGetLastError(), "batch group launch"));
    }
    return GSN_OK;
}

int BatchImpl::launch(hipStream_t s)
{
    GSN_TRY(take_injected_fault());
    if (exec)
        return hip_status(hipGraphLaunch(exec, s), "hipGraphLaunch");
    return enqueue(s);
}

size_t batch_workspace_bytes(const gsn_batch_item *items, size_t count)
{
    size_t bytes = 0;
    for (size_t i = 0; i < count; ++i) {
        const std::vector<PassCfg> v = choose_passes(items[i].field, items[i].log_n, true);
        const int nbuf = v.size() >= 3 ? 2 : (v.size() == 2 ? 1 : 0);
        bytes += (size_t)nbuf * ((size_t)1 << items[i].log_n) * element_bytes(items[i].field);
        bytes = (bytes + 255) & ~(size_t)255;
    }
    return bytes;
}

int build_batch(BatchImpl &b, const gsn_batch_item *items, size_t count, int use_graph, const BatchMemory &mem)
{
    GSN_TRY(ensure_device());
    // plans and workspace sizes
    std::vector<PlanBase *> plans(count);
    size_t ws_bytes = 0;
    std::vector<size_t> ws_off(count);
    for (size_t i = 0; i < count; ++i) {
        const gsn_batch_item &it = items[i];
        if (!it.d_in || !it.d_out)
            return set_error(GSN_ERR_INVALID_ARGUMENT, "batch item " + std::to_string(i) + " has a NULL buffer");
        GSN_TRY(cached_plan(&plans[i], it.field, it.log_n, it.omega, it.inverse, true));
        const size_t eb = element_bytes(it.field);
        const size_t n = (size_t)1 << it.log_n;
        const int P = plans[i]->num_passes;
        const int nbuf = P >= 3 ? 2 : (P == 2 ? 1 : 0);
        ws_off[i] = ws_bytes;
        ws_bytes += (size_t)nbuf * n * eb;
        ws_bytes = (ws_bytes + 255) & ~(size_t)255;
    }
    char *ws_base = nullptr;
    if (mem.alloc) {
        ws_base = (char *)(ws_bytes ? mem.alloc(mem.ctx, ws_bytes) : nullptr);
        if (ws_bytes && !ws_base)
            return set_error(GSN_ERR_OUT_OF_MEMORY, "batch arena exhausted");
    } else {
        GSN_TRY(b.workspace.alloc(ws_bytes));
        ws_base = (char *)b.workspace.ptr;
    }
    Collector<GSN_FIELD_FP32> c32;
    Collector<GSN_FIELD_MNT4753_FR> c4;
    Collector<GSN_FIELD_MNT6753_FR> c6;
    for (size_t i = 0; i < count; ++i) {
        const gsn_batch_item &it = items[i];
        const size_t n = (size_t)1 << it.log_n;
        char *ws0 = ws_base + ws_off[i];
        char *ws1 = ws0 + n * element_bytes(it.field);
        if (plans[i]->num_passes == 0) {
            b.copies.push_back({it.field, it.d_in, it.d_out});
            continue;
        }
        switch (it.field) {
        case GSN_FIELD_FP32: c32.add(*static_cast<PlanImpl<GSN_FIELD_FP32> *>(plans[i]), it.d_in, it.d_out, ws0, ws1); break;
        case GSN_FIELD_MNT4753_FR:
            c4.add(*static_cast<PlanImpl<GSN_FIELD_MNT4753_FR> *>(plans[i]), it.d_in, it.d_out, ws0, ws1);
            break;
        default: c6.add(*static_cast<PlanImpl<GSN_FIELD_MNT6753_FR> *>(plans[i]), it.d_in, it.d_out, ws0, ws1); break;
        }
    }
    // 753-bit groups first within a level: they are the long ones
    GSN_TRY(c6.emit(b.groups, mem));
    GSN_TRY(c4.emit(b.groups, mem));
    GSN_TRY(c32.emit(b.groups, mem));
    std::stable_sort(b.groups.begin(), b.groups.end(),
                     [](const Group &x, const Group &y) { return x.level < y.level; });
    b.transforms = count;
    if (use_graph) {
        GSN_TRY(hip_status(hipStreamCreateWithFlags(&b.capture_stream, hipStreamNonBlocking), "hipStreamCreate"));
        GSN_TRY(hip_status(hipStreamBeginCapture(b.capture_stream, hipStreamCaptureModeThreadLocal),
                           "hipStreamBeginCapture"));
        const int st = b.enqueue(b.capture_stream);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b.capture_stream, &graph);
        GSN_TRY(st);
        GSN_TRY(hip_status(e, "hipStreamEndCapture"));
        b.graph = graph;
        GSN_TRY(hip_status(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate"));
    }
    return GSN_OK;
}

}  // namespace gsn

using namespace gsn;

struct gsn_batch {
    std::unique_ptr<BatchImpl> impl;
};

extern "C" {

int gsn_batch_create(gsn_batch **batch, const gsn_batch_item *items, size_t count, int use_graph)
{
    if (!batch || (!items && count))
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL batch or items");
    *batch = nullptr;
    std::unique_ptr<gsn_batch> b(new gsn_batch{std::unique_ptr<BatchImpl>(new BatchImpl())});
    GSN_TRY(build_batch(*b->impl, items, count, use_graph));
    *batch = b.release();
    return GSN_OK;
}

int gsn_batch_launch(gsn_batch *batch, void *stream)
{
    if (!batch)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL batch");
    GSN_TRY(ensure_device());
    return batch->impl->launch((hipStream_t)stream);
}

int gsn_batch_destroy(gsn_batch *batch)
{
    delete batch;
    return GSN_OK;
}

int gsn_batch_execute(const gsn_batch_item *items, size_t count, void *stream)
{
    gsn_batch *b = nullptr;
    GSN_TRY(gsn_batch_create(&b, items, count, 0));
    int st = gsn_batch_launch(b, stream);
    if (st == GSN_OK)
        st = hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
    gsn_batch_destroy(b);
    return st;
}

int gsn_batch_num_launches(const gsn_batch *batch)
{
    return batch ? (int)(batch->impl->groups.size() + batch->impl->copies.size()) : 0;
}

}  // extern "C"
