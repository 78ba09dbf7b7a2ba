// Internal declarations shared by the runtime's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <mutex>
#include <string>

#include "gpusnarks.h"

#define GSN_TRY(expr)                  \
    do {                               \
        const int gsn_st_ = (expr);    \
        if (gsn_st_ != 0)              \
            return gsn_st_;            \
    } while (0)

namespace gsn {

int set_error(int status, const std::string &msg);
int hip_status(hipError_t e, const char *what);
int take_injected_fault();
int ensure_device();
size_t element_bytes(int field);
int two_adicity(int field);
int log2_exact(size_t n, unsigned *log_n);

// Owning device allocation (hipMalloc / hipFree).
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    DeviceBuffer(DeviceBuffer &&o) noexcept : ptr(o.ptr), bytes(o.bytes)
    {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DeviceBuffer &operator=(DeviceBuffer &&o) noexcept;
    ~DeviceBuffer();
    int alloc(size_t n);  // grows; contents are not preserved
};

// A transform plan: field, size, direction, twiddles and workspace.
struct PlanBase {
    int field = 0;
    unsigned log_n = 0;
    int inverse = 0;
    int num_passes = 0;
    bool batch_mode = false;  // schedule of the batch table (fused launches)
    size_t device_bytes = 0;
    std::mutex host_mutex;  // serialises the host entry point's buffer
    DeviceBuffer host_buf;
    virtual ~PlanBase() = default;
    virtual int execute(const void *in, void *out, hipStream_t s) = 0;
    // Plain power-of-two plans only: execute() with the workspace in a
    // caller's device buffer, allocating nothing; workspace_bytes() is what
    // that buffer must hold.
    virtual size_t workspace_bytes() const { return 0; }
    virtual int execute_workspace(const void *in, void *out, void *ws, size_t ws_bytes, hipStream_t s)
    {
        (void)in, (void)out, (void)ws, (void)ws_bytes, (void)s;
        return set_error(GSN_ERR_INVALID_ARGUMENT,
                         "the allocation-free entry point takes plans of gsn_plan_create only");
    }
};

// Copies of caller-owned host buffers through pinned, double-buffered
// staging (synchronous with respect to the host buffer: staged_h2d may
// return before the device holds the data, but work later on `st` sees it;
// staged_d2h returns once the host buffer is filled).
int staged_h2d(void *dev, const void *host, size_t bytes, hipStream_t st);
int staged_d2h(void *host, const void *dev, size_t bytes, hipStream_t st);

// batch = true: a plan whose passes run fused with other transforms' (the
// batch table's schedule)
int create_plan(PlanBase **out, int field, unsigned log_n, const void *omega, int inverse, bool batch = false);
void clear_schedules();
void clear_plan_cache();
int cached_plan(PlanBase **out, int field, unsigned log_n, const void *omega, int inverse, bool batch = false);

}  // namespace gsn

// the C ABI's opaque plan handle
struct gsn_plan {
    gsn::PlanBase *impl;
};
