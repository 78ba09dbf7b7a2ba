// Serving: continuous batching of NTT requests.
//
// The reference serves one synchronous best_fft at a time
// (cuda/fft_kernel.cu:117-147).  Here a scheduler thread admits queued
// requests in arrival order into "waves" whenever a stream slot and device
// memory are free, fuses each wave into one batch schedule (one launch per
// pass level and field), and runs waves concurrently on several streams, so
// requests that arrive while others are on the GPU start as soon as
// capacity frees up.
//
// Device memory: one HBM arena of the admission budget, allocated at start;
// each wave carves a contiguous block (operands, workspace, job tables) and
// returns it when its completion event fires.  A request larger than the
// whole arena is rejected (GSN_ERR_ADMISSION); one that fits waits.  No
// hipMalloc/hipFree runs in the serving loop (hipFree would synchronise the
// device and serialise the waves).  Host-buffer requests stage through a
// pinned arena (hipHostMalloc) so copies are asynchronous.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "batch_impl.h"

namespace gsn {

namespace {

constexpr size_t ALIGN = 256;
size_t align_up(size_t x) { return (x + ALIGN - 1) & ~(ALIGN - 1); }

// First-fit allocator over [0, size) with coalescing.
class RangeAllocator {
  public:
    void reset(size_t size)
    {
        free_.clear();
        size_ = size;
        if (size)
            free_[0] = size;
    }
    size_t size() const { return size_; }
    // offset, or SIZE_MAX when no free block is large enough
    size_t alloc(size_t bytes)
    {
        bytes = align_up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rem = it->second - bytes;
                free_.erase(it);
                if (rem)
                    free_[off + bytes] = rem;
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes)
    {
        bytes = align_up(bytes ? bytes : 1);
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }

  private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0;
};

struct Request {
    uint64_t ticket = 0;
    int field = 0;
    unsigned log_n = 0;
    int inverse = 0;
    bool has_omega = false;
    uint32_t omega[24] = {};
    const void *in = nullptr;
    void *out = nullptr;
    bool device = false;  // operands already in device memory
    int status = GSN_ERR_NOT_READY;
    bool done = false;
    double t_submit = 0, t_start = 0, t_done = 0;
    size_t operand_bytes = 0;  // n * element bytes
    size_t device_bytes = 0;   // arena bytes this request needs in a wave
};

struct Wave {
    std::vector<Request *> reqs;
    size_t dev_off = 0, dev_size = 0, dev_used = 0;
    size_t pin_off = SIZE_MAX, pin_size = 0;
    int stream = 0;
    hipEvent_t ev = nullptr;
    std::unique_ptr<BatchImpl> batch;
    char *dev_base = nullptr;
};

void *wave_alloc(void *ctx, size_t bytes)
{
    Wave *w = (Wave *)ctx;
    bytes = align_up(bytes);
    if (w->dev_used + bytes > w->dev_size)
        return nullptr;
    void *p = w->dev_base + w->dev_used;
    w->dev_used += bytes;
    return p;
}

// One request alone: its plan's passes launched directly, with PassArgs by
// value and the wave's own workspace -- no job tables to upload, which a
// fused batch needs.
template <int FID>
int enqueue_single_t(PlanBase *base, const gsn_batch_item &it, Wave *w, hipStream_t s)
{
    using F = typename FieldInfo<FID>::Dev;
    const PlanImpl<FID> &plan = *static_cast<PlanImpl<FID> *>(base);
    const size_t n = (size_t)1 << it.log_n;
    if (plan.cfg.empty()) {  // n == 1: the canonical copy
        launch_scale<F>((const typename F::elem_t *)it.d_in, (typename F::elem_t *)it.d_out, {}, 1, 0, s);
        return hip_status(hipGetLastError(), "copy launch");
    }
    const int P = plan.num_passes;
    void *ws0 = nullptr, *ws1 = nullptr;
    if (P >= 2 && !(ws0 = wave_alloc(w, n * sizeof(typename F::elem_t))))
        return set_error(GSN_ERR_OUT_OF_MEMORY, "wave memory exhausted");
    if (P >= 3 && !(ws1 = wave_alloc(w, n * sizeof(typename F::elem_t))))
        return set_error(GSN_ERR_OUT_OF_MEMORY, "wave memory exhausted");
    std::vector<PassJob<F>> jobs;
    plan.make_jobs(it.d_in, it.d_out, ws0, ws1, jobs);
    for (size_t i = 0; i < jobs.size(); ++i) {
        launch_pass<F>(jobs[i], plan.cfg[i], s);
        GSN_TRY(hip_status(hipGetLastError(), "ntt pass launch"));
    }
    return GSN_OK;
}

int enqueue_single(const gsn_batch_item &it, Wave *w, hipStream_t s)
{
    PlanBase *p = nullptr;
    GSN_TRY(cached_plan(&p, it.field, it.log_n, it.omega, it.inverse, true));
    switch (it.field) {
    case GSN_FIELD_FP32: return enqueue_single_t<GSN_FIELD_FP32>(p, it, w, s);
    case GSN_FIELD_MNT4753_FR: return enqueue_single_t<GSN_FIELD_MNT4753_FR>(p, it, w, s);
    default: return enqueue_single_t<GSN_FIELD_MNT6753_FR>(p, it, w, s);
    }
}

// Device bytes a request needs in a wave: operand (host requests) +
// workspace + job tables, each rounded to the allocator alignment.
size_t request_bytes(int field, unsigned log_n, bool device)
{
    const size_t n = (size_t)1 << log_n, eb = element_bytes(field);
    const std::vector<PassCfg> v = choose_passes(field, log_n, true);
    const int nbuf = v.size() >= 3 ? 2 : (v.size() == 2 ? 1 : 0);
    size_t bytes = device ? 0 : align_up(n * eb);
    bytes += align_up((size_t)nbuf * n * eb);
    for (const PassCfg &c : v) {
        const size_t tiles = (n >> c.log_r) >> c.log_w;
        // a job table and a tile map per pass (worst case: a group of its own)
        bytes += align_up(sizeof(PassJob<FMnt6>)) + align_up(tiles * sizeof(uint32_t));
    }
    return bytes + ALIGN;
}

}  // namespace

class Server {
  public:
    explicit Server(const gsn_server_config &cfg) : cfg_(cfg) {}
    ~Server() { shutdown(); }

    int start()
    {
        GSN_TRY(ensure_device());
        if (cfg_.max_inflight <= 0)
            cfg_.max_inflight = 4096;
        if (cfg_.num_streams <= 0)
            cfg_.num_streams = 4;
        if (cfg_.max_batch <= 0)
            cfg_.max_batch = 256;
        size_t budget = cfg_.hbm_budget_bytes;
        if (budget == 0) {
            size_t fr = 0, total = 0;
            GSN_TRY(hip_status(hipMemGetInfo(&fr, &total), "hipMemGetInfo"));
            budget = fr / 2;
        }
        budget = align_up(budget);
        GSN_TRY(hip_status(hipMalloc(&arena_, budget), "hipMalloc(server arena)"));
        dev_.reset(budget);
        pin_bytes_ = budget < ((size_t)2 << 30) ? budget : ((size_t)2 << 30);
        GSN_TRY(hip_status(hipHostMalloc(&pinned_, pin_bytes_, hipHostMallocDefault), "hipHostMalloc"));
        pin_.reset(pin_bytes_);
        streams_.resize(cfg_.num_streams, nullptr);
        busy_.assign(cfg_.num_streams, false);
        for (auto &s : streams_)
            GSN_TRY(hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"));
        t0_ = std::chrono::steady_clock::now();
        thread_ = std::thread([this] { loop(); });
        return GSN_OK;
    }

    int submit(int field, unsigned log_n, int inverse, const void *omega, const void *in, void *out, bool device,
               uint64_t *ticket)
    {
        if (!ticket || !in || !out)
            return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL ticket or buffer");
        const size_t eb = element_bytes(field);
        if (!eb)
            return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id");
        if ((int)log_n > two_adicity(field))
            return set_error(GSN_ERR_SIZE_EXCEEDS_TWO_ADICITY, "log2(n) exceeds the field's 2-adicity");
        std::unique_ptr<Request> r(new Request());
        r->field = field;
        r->log_n = log_n;
        r->inverse = inverse ? 1 : 0;
        if (omega) {
            r->has_omega = true;
            std::memcpy(r->omega, omega, eb);
        }
        r->in = in;
        r->out = out;
        r->device = device;
        r->operand_bytes = ((size_t)1 << log_n) * eb;
        r->device_bytes = request_bytes(field, log_n, device);
        if (r->device_bytes > dev_.size() || (!device && r->operand_bytes > pin_bytes_))
            return set_error(GSN_ERR_ADMISSION, "request needs " + std::to_string(r->device_bytes) +
                                                    " device bytes, more than the admission budget " +
                                                    std::to_string(dev_.size()));
        std::lock_guard<std::mutex> lk(m_);
        if (stop_)
            return set_error(GSN_ERR_SHUTDOWN, "server is shutting down");
        r->ticket = ++next_ticket_;
        r->t_submit = now();
        *ticket = r->ticket;
        pending_.push_back(r.get());
        all_[r->ticket] = std::move(r);
        cv_.notify_all();
        return GSN_OK;
    }

    int poll(uint64_t ticket, int *done)
    {
        std::lock_guard<std::mutex> lk(m_);
        auto it = all_.find(ticket);
        if (it == all_.end())
            return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown ticket");
        if (done)
            *done = it->second->done ? 1 : 0;
        return it->second->done ? it->second->status : GSN_OK;
    }

    int wait(uint64_t ticket)
    {
        std::unique_lock<std::mutex> lk(m_);
        auto it = all_.find(ticket);
        if (it == all_.end())
            return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown ticket");
        Request *r = it->second.get();
        done_cv_.wait(lk, [r] { return r->done; });
        return r->status;
    }

    int times(uint64_t ticket, double *ts, double *tb, double *td)
    {
        std::lock_guard<std::mutex> lk(m_);
        auto it = all_.find(ticket);
        if (it == all_.end())
            return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown ticket");
        if (!it->second->done)
            return set_error(GSN_ERR_NOT_READY, "ticket has not completed");
        if (ts)
            *ts = it->second->t_submit;
        if (tb)
            *tb = it->second->t_start;
        if (td)
            *td = it->second->t_done;
        return GSN_OK;
    }

    void shutdown()
    {
        {
            std::lock_guard<std::mutex> lk(m_);
            if (stop_ && !thread_.joinable())
                return;
            stop_ = true;
            cv_.notify_all();
        }
        if (thread_.joinable())
            thread_.join();
        for (auto &s : streams_)
            if (s)
                (void)hipStreamDestroy(s);
        streams_.clear();
        if (arena_)
            (void)hipFree(arena_);
        arena_ = nullptr;
        if (pinned_)
            (void)hipHostFree(pinned_);
        pinned_ = nullptr;
    }

  private:
    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    // Scheduler thread: every HIP call of the serving loop happens here.
    void loop()
    {
        for (;;) {
            retire_finished();
            while (launch_wave()) {
            }
            std::unique_lock<std::mutex> lk(m_);
            if (stop_ && pending_.empty() && waves_.empty())
                break;
            if (!waves_.empty()) {
                // waves in flight: poll their events without sleeping (a
                // timed wait would round up to the kernel's timer slack,
                // tens of microseconds added to every request's latency)
                lk.unlock();
                std::this_thread::yield();
            } else if (pending_.empty())
                cv_.wait(lk, [this] { return stop_ || !pending_.empty(); });
            else
                cv_.wait_for(lk, std::chrono::microseconds(200));  // waiting for memory
        }
    }

    void finish(Request *r, int status)
    {
        r->status = status;
        r->t_done = now();
        r->done = true;
    }

    void retire_finished()
    {
        for (size_t i = 0; i < waves_.size();) {
            Wave *w = waves_[i].get();
            const hipError_t q = hipEventQuery(w->ev);
            if (q == hipErrorNotReady) {
                ++i;
                continue;
            }
            const int st = q == hipSuccess ? GSN_OK : hip_status(q, "wave completion");
            size_t pin_cursor = w->pin_off;
            {
                std::lock_guard<std::mutex> lk(m_);
                for (Request *r : w->reqs) {
                    if (!r->device && st == GSN_OK) {
                        std::memcpy(r->out, (char *)pinned_ + pin_cursor, r->operand_bytes);
                        pin_cursor += align_up(r->operand_bytes);
                    }
                    finish(r, st);
                }
                dev_.release(w->dev_off, w->dev_size);
                if (w->pin_off != SIZE_MAX)
                    pin_.release(w->pin_off, w->pin_size);
                busy_[w->stream] = false;
                inflight_ -= (int)w->reqs.size();
            }
            (void)hipEventDestroy(w->ev);
            waves_.erase(waves_.begin() + i);
            done_cv_.notify_all();
        }
    }

    // Admits the longest FIFO prefix of pending requests that fits, and
    // launches it as one wave.  Returns true if a wave was launched.
    bool launch_wave()
    {
        std::unique_ptr<Wave> w(new Wave());
        {
            std::lock_guard<std::mutex> lk(m_);
            if (pending_.empty())
                return false;
            int slot = -1;
            for (int s = 0; s < (int)busy_.size(); ++s)
                if (!busy_[s]) {
                    slot = s;
                    break;
                }
            if (slot < 0)
                return false;
            size_t dev_need = 0, pin_need = 0;
            std::vector<Request *> take;
            for (Request *r : pending_) {
                if ((int)take.size() >= cfg_.max_batch || inflight_ + (int)take.size() >= cfg_.max_inflight)
                    break;
                const size_t d2 = dev_need + r->device_bytes;
                const size_t p2 = pin_need + (r->device ? 0 : align_up(r->operand_bytes));
                const size_t doff = dev_.alloc(d2);
                if (doff == SIZE_MAX)
                    break;
                dev_.release(doff, d2);
                if (p2) {
                    const size_t poff = pin_.alloc(p2);
                    if (poff == SIZE_MAX)
                        break;
                    pin_.release(poff, p2);
                }
                dev_need = d2;
                pin_need = p2;
                take.push_back(r);
            }
            if (take.empty())
                return false;
            w->dev_off = dev_.alloc(dev_need);
            w->dev_size = dev_need;
            if (pin_need) {
                w->pin_off = pin_.alloc(pin_need);
                w->pin_size = pin_need;
            }
            w->reqs = take;
            w->stream = slot;
            busy_[slot] = true;
            inflight_ += (int)take.size();
            pending_.erase(pending_.begin(), pending_.begin() + take.size());
        }
        w->dev_base = (char *)arena_ + w->dev_off;
        const hipStream_t s = streams_[w->stream];
        int st = hip_status(hipEventCreateWithFlags(&w->ev, hipEventDisableTiming), "hipEventCreate");
        // operands: host requests get a slice of the wave block and are
        // staged through pinned memory
        std::vector<gsn_batch_item> items(w->reqs.size());
        size_t pin_cursor = w->pin_off;
        for (size_t i = 0; i < w->reqs.size() && st == GSN_OK; ++i) {
            Request *r = w->reqs[i];
            gsn_batch_item &it = items[i];
            it.field = r->field;
            it.log_n = r->log_n;
            it.inverse = r->inverse;
            it.omega = r->has_omega ? r->omega : nullptr;
            if (r->device) {
                it.d_in = r->in;
                it.d_out = r->out;
            } else {
                void *d = wave_alloc(w.get(), r->operand_bytes);
                char *p = (char *)pinned_ + pin_cursor;
                pin_cursor += align_up(r->operand_bytes);
                std::memcpy(p, r->in, r->operand_bytes);
                st = hip_status(hipMemcpyAsync(d, p, r->operand_bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
                it.d_in = d;
                it.d_out = d;
            }
        }
        if (st == GSN_OK && items.size() == 1) {
            st = enqueue_single(items[0], w.get(), s);
        } else if (st == GSN_OK) {
            w->batch.reset(new BatchImpl());
            BatchMemory mem;
            mem.alloc = wave_alloc;
            mem.ctx = w.get();
            mem.stream = s;
            st = build_batch(*w->batch, items.data(), items.size(), 0, mem);
            if (st == GSN_OK)
                st = w->batch->enqueue(s);
        }
        pin_cursor = w->pin_off;
        for (size_t i = 0; i < w->reqs.size() && st == GSN_OK; ++i) {
            Request *r = w->reqs[i];
            if (r->device)
                continue;
            st = hip_status(hipMemcpyAsync((char *)pinned_ + pin_cursor, items[i].d_out, r->operand_bytes,
                                           hipMemcpyDeviceToHost, s),
                            "hipMemcpyAsync D2H");
            pin_cursor += align_up(r->operand_bytes);
        }
        if (st == GSN_OK)
            st = hip_status(hipEventRecord(w->ev, s), "hipEventRecord");
        const double t = now();
        std::lock_guard<std::mutex> lk(m_);
        for (Request *r : w->reqs)
            r->t_start = t;
        if (st != GSN_OK) {
            // nothing usable was queued past the failure: wait for the stream,
            // then fail the wave's requests and give its memory back
            (void)hipStreamSynchronize(s);
            for (Request *r : w->reqs)
                finish(r, st);
            dev_.release(w->dev_off, w->dev_size);
            if (w->pin_off != SIZE_MAX)
                pin_.release(w->pin_off, w->pin_size);
            busy_[w->stream] = false;
            inflight_ -= (int)w->reqs.size();
            if (w->ev)
                (void)hipEventDestroy(w->ev);
            done_cv_.notify_all();
            return true;
        }
        waves_.push_back(std::move(w));
        return true;
    }

    gsn_server_config cfg_;
    void *arena_ = nullptr;
    void *pinned_ = nullptr;
    size_t pin_bytes_ = 0;
    RangeAllocator dev_, pin_;
    std::vector<hipStream_t> streams_;
    std::vector<bool> busy_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::deque<Request *> pending_;
    std::map<uint64_t, std::unique_ptr<Request>> all_;
    std::vector<std::unique_ptr<Wave>> waves_;
    int inflight_ = 0;
    uint64_t next_ticket_ = 0;
    bool stop_ = false;
    std::thread thread_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace gsn

using namespace gsn;

struct gsn_server {
    std::unique_ptr<Server> impl;
};

extern "C" {

int gsn_server_create(gsn_server **server, const gsn_server_config *config)
{
    if (!server)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "server is NULL");
    *server = nullptr;
    gsn_server_config cfg{};
    if (config)
        cfg = *config;
    std::unique_ptr<gsn_server> s(new gsn_server{std::unique_ptr<Server>(new Server(cfg))});
    GSN_TRY(s->impl->start());
    *server = s.release();
    return GSN_OK;
}

int gsn_server_submit(gsn_server *server, int field, unsigned log_n, int inverse, const void *omega,
                      const void *h_in, void *h_out, uint64_t *ticket)
{
    if (!server)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "server is NULL");
    return server->impl->submit(field, log_n, inverse, omega, h_in, h_out, false, ticket);
}

int gsn_server_submit_device(gsn_server *server, int field, unsigned log_n, int inverse, const void *omega,
                             const void *d_in, void *d_out, uint64_t *ticket)
{
    if (!server)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "server is NULL");
    return server->impl->submit(field, log_n, inverse, omega, d_in, d_out, true, ticket);
}

int gsn_server_poll(gsn_server *server, uint64_t ticket, int *done)
{
    if (!server)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "server is NULL");
    return server->impl->poll(ticket, done);
}

int gsn_server_wait(gsn_server *server, uint64_t ticket)
{
    if (!server)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "server is NULL");
    return server->impl->wait(ticket);
}

int gsn_server_ticket_times(gsn_server *server, uint64_t ticket, double *t_submit, double *t_start, double *t_done)
{
    if (!server)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "server is NULL");
    return server->impl->times(ticket, t_submit, t_start, t_done);
}

int gsn_server_destroy(gsn_server *server)
{
    if (!server)
        return GSN_OK;
    server->impl->shutdown();
    delete server;
    return GSN_OK;
}

}  // extern "C"
