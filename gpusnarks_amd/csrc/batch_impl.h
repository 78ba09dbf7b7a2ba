// Batch schedule: fused multi-transform pass launches, level by level.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "gpusnarks.h"
#include "plan_impl.h"

namespace gsn {

// One fused launch: one pass level of many transforms of one field (and,
// for the 32-bit field, one tile shape).
struct Group {
    int field = 0;
    int level = 0;
    PassCfg cfg{};
    uint32_t total_tiles = 0;
    DeviceBuffer jobs_buf, tile_buf;  // owned storage (default memory)
    const void *jobs = nullptr;       // PassJob<F>[]
    const uint32_t *tile_job = nullptr;  // [total_tiles]: block -> job
    std::vector<char> host_jobs;      // host copies for asynchronous uploads
    std::vector<uint32_t> host_tiles;
};

struct Copy {  // n == 1 transforms
    int field;
    const void *in;
    void *out;
};

struct BatchImpl {
    std::vector<Group> groups;  // sorted by level
    std::vector<Copy> copies;
    DeviceBuffer workspace;     // owned workspace (when none was supplied)
    hipStream_t capture_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    size_t transforms = 0;

    ~BatchImpl();
    int enqueue(hipStream_t s);  // launches every group on s
    int launch(hipStream_t s);   // graph replay, or enqueue
};

// Where a batch's device memory comes from.  With alloc == nullptr the
// batch owns hipMalloc'd buffers and uploads its tables synchronously;
// otherwise the workspace and tables come from alloc (an arena the caller
// owns, no hipMalloc/hipFree) and the tables are uploaded on `stream` from
// host copies kept in the BatchImpl until it is destroyed.
struct BatchMemory {
    void *(*alloc)(void *ctx, size_t bytes) = nullptr;
    void *ctx = nullptr;
    hipStream_t stream = nullptr;
};

// Workspace bytes a batch of these items needs (256-byte aligned slices).
size_t batch_workspace_bytes(const gsn_batch_item *items, size_t count);
// Builds the schedule.
int build_batch(BatchImpl &b, const gsn_batch_item *items, size_t count, int use_graph,
                const BatchMemory &mem = BatchMemory());

}  // namespace gsn
