// Host <-> device copies of caller-owned (pageable) buffers for the host
// entry point gsn_fft_host, replacing the reference's per-call cudaMemcpy
// of the whole vector (cuda/fft_kernel.cu:129-146).
//
// The data moves through two pinned chunks (hipHostMalloc, allocated once
// and kept for the process): while the DMA engine copies one chunk, host
// threads fill (or drain) the other, and each chunk's CPU copy is split over
// a small pool of worker threads, since one core's memcpy from pageable
// memory is slower than the PCIe link.  Transfers below one chunk use a
// plain hipMemcpy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "gpusnarks.h"
#include "gsn_internal.h"

namespace gsn {

namespace {

constexpr size_t CHUNK = (size_t)8 << 20;
constexpr size_t MIN_SPLIT = (size_t)1 << 20;  // smallest piece a worker copies

// memcpy split over the caller and up to `workers` pool threads
class CopyPool {
  public:
    explicit CopyPool(unsigned workers)
    {
        for (unsigned i = 0; i < workers; ++i)
            threads_.emplace_back([this] { run(); });
    }
    void copy(void *dst, const void *src, size_t n)
    {
        const size_t parts = std::min<size_t>(threads_.size() + 1, std::max<size_t>(1, n / MIN_SPLIT));
        if (parts <= 1) {
            std::memcpy(dst, src, n);
            return;
        }
        const size_t piece = (n / parts + 4095) & ~(size_t)4095;
        std::unique_lock<std::mutex> lk(m_);
        size_t off = piece;  // the caller copies [0, piece)
        for (; off < n; off += piece)
            tasks_.push_back({(char *)dst + off, (const char *)src + off, std::min(piece, n - off)});
        pending_ = tasks_.size();
        lk.unlock();
        cv_.notify_all();
        std::memcpy(dst, src, std::min(piece, n));
        lk.lock();
        done_.wait(lk, [this] { return pending_ == 0; });
    }

  private:
    struct Task {
        char *dst;
        const char *src;
        size_t n;
    };
    void run()
    {
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [this] { return !tasks_.empty(); });
            const Task t = tasks_.back();
            tasks_.pop_back();
            lk.unlock();
            std::memcpy(t.dst, t.src, t.n);
            lk.lock();
            if (--pending_ == 0)
                done_.notify_all();
        }
    }
    std::vector<std::thread> threads_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    std::vector<Task> tasks_;
    size_t pending_ = 0;
};

struct Staging {
    std::mutex m;
    void *pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    CopyPool *pool = nullptr;

    int init()
    {
        if (pin[0])
            return GSN_OK;
        for (int b = 0; b < 2; ++b) {
            GSN_TRY(hip_status(hipHostMalloc(&pin[b], CHUNK, hipHostMallocDefault), "hipHostMalloc"));
            GSN_TRY(hip_status(hipEventCreateWithFlags(&ev[b], hipEventDisableTiming), "hipEventCreate"));
        }
        const unsigned hw = std::thread::hardware_concurrency();
        pool = new CopyPool(std::min(3u, hw > 1 ? hw - 1 : 0u));
        return GSN_OK;
    }
};

// never destroyed: the pinned chunks and the pool live until the process ends
Staging &staging()
{
    static Staging *s = new Staging();
    return *s;
}

}  // namespace

int staged_h2d(void *dev, const void *host, size_t bytes, hipStream_t st)
{
    if (bytes <= CHUNK)
        return hip_status(hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice), "hipMemcpy H2D");
    Staging &S = staging();
    std::lock_guard<std::mutex> lk(S.m);
    GSN_TRY(S.init());
    size_t off = 0;
    for (int i = 0; off < bytes; ++i) {
        const int b = i & 1;
        const size_t len = std::min(CHUNK, bytes - off);
        // the chunk's previous DMA (two chunks back) must have finished
        GSN_TRY(hip_status(hipEventSynchronize(S.ev[b]), "hipEventSynchronize"));
        S.pool->copy(S.pin[b], (const char *)host + off, len);
        GSN_TRY(hip_status(hipMemcpyAsync((char *)dev + off, S.pin[b], len, hipMemcpyHostToDevice, st),
                           "hipMemcpyAsync H2D"));
        GSN_TRY(hip_status(hipEventRecord(S.ev[b], st), "hipEventRecord"));
        off += len;
    }
    // the caller's later work on `st` is ordered after the copies; the
    // chunks are reused only after their events
    return GSN_OK;
}

int staged_d2h(void *host, const void *dev, size_t bytes, hipStream_t st)
{
    if (bytes <= CHUNK)
        return hip_status(hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost), "hipMemcpy D2H");
    Staging &S = staging();
    std::lock_guard<std::mutex> lk(S.m);
    GSN_TRY(S.init());
    const size_t chunks = (bytes + CHUNK - 1) / CHUNK;
    auto issue = [&](size_t i) {
        const int b = (int)(i & 1);
        const size_t off = i * CHUNK, len = std::min(CHUNK, bytes - off);
        GSN_TRY(hip_status(hipMemcpyAsync(S.pin[b], (const char *)dev + off, len, hipMemcpyDeviceToHost, st),
                           "hipMemcpyAsync D2H"));
        return hip_status(hipEventRecord(S.ev[b], st), "hipEventRecord");
    };
    GSN_TRY(issue(0));
    if (chunks > 1)
        GSN_TRY(issue(1));
    for (size_t i = 0; i < chunks; ++i) {
        const int b = (int)(i & 1);
        const size_t off = i * CHUNK, len = std::min(CHUNK, bytes - off);
        GSN_TRY(hip_status(hipEventSynchronize(S.ev[b]), "hipEventSynchronize"));
        S.pool->copy((char *)host + off, S.pin[b], len);
        if (i + 2 < chunks)
            GSN_TRY(issue(i + 2));
    }
    return GSN_OK;
}

}  // namespace gsn
