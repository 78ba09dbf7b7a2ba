// Host runtime of the gfx950 NTT: the C ABI of include/gpusnarks.h.
//
// Replaces the reference's host wrapper best_fft (cuda/fft_kernel.cu:117-147),
// which allocated device buffers on every call and never freed them, copied
// the vector over, launched one <<<5,256>>> kernel with omega ignored, and
// exited on errors.  Here a plan owns device-resident twiddle tables and a
// workspace, validates omega, picks a pass schedule per (field, size) and
// launches one Stockham pass kernel per pass on the caller's stream.
#include <hip/hip_runtime.h>

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "device_field.h"
#include "gpusnarks.h"
#include "gsn_internal.h"
#include "plan_impl.h"

namespace gsn {

// ------------------------------------------------------------ errors

namespace {
thread_local std::string t_last_error;
std::atomic<int> g_inject_faults{0};
std::mutex g_dev_mutex;
int g_device = -1;
}  // namespace

int set_error(int status, const std::string &msg)
{
    t_last_error = msg;
    return status;
}

int hip_status(hipError_t e, const char *what)
{
    if (e == hipSuccess)
        return GSN_OK;
    const int st = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? GSN_ERR_OUT_OF_MEMORY : GSN_ERR_DEVICE;
    return set_error(st, std::string(what) + ": " + hipGetErrorString(e));
}

int take_injected_fault()
{
    int v = g_inject_faults.load();
    while (v > 0) {
        if (g_inject_faults.compare_exchange_weak(v, v - 1))
            return set_error(GSN_ERR_INJECTED_FAULT, "injected fault");
    }
    return GSN_OK;
}

int ensure_device()
{
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    if (g_device >= 0)
        return hip_status(hipSetDevice(g_device), "hipSetDevice");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return set_error(GSN_ERR_NO_DEVICE, "no HIP device available");
    GSN_TRY(hip_status(hipSetDevice(0), "hipSetDevice"));
    g_device = 0;
    return GSN_OK;
}

// ------------------------------------------------------------ passes

std::vector<PassCfg> default_passes(int field, unsigned L)
{
    std::vector<PassCfg> v;
    if (L == 0)
        return v;
    if (field != GSN_FIELD_FP32) {
        // 753-bit: radix-2 steps in one-wave tiles of 2^7 elements (passes
        // of at most 2^7 rows; spare tile capacity becomes extra columns)
        const int MAX_R = 7;
        const int P = (int)((L + MAX_R - 1) / MAX_R);
        for (int i = 0; i < P; ++i) {
            const int r = (int)L / P + (i < (int)L % P ? 1 : 0);
            int w = MAX_R - r;
            if (w > (int)L - r)
                w = (int)L - r;
            v.push_back({r, w, 1});
        }
        return v;
    }
    // 32-bit: radix-16 register steps; one pass of up to 2^12 elements,
    // otherwise passes of up to 2^10 rows x 16 columns (64 KiB of LDS)
    if (L <= 12) {
        v.push_back({(int)L, 0, (int)(L < 4 ? L : 4)});
        return v;
    }
    const int P = (int)((L + 9) / 10);
    for (int i = 0; i < P; ++i) {
        const int r = (int)L / P + (i < (int)L % P ? 1 : 0);
        v.push_back({r, 4, 4});
    }
    return v;
}

namespace {
std::mutex g_tune_mutex;
std::map<std::pair<int, unsigned>, std::vector<PassCfg>> g_tuned;
std::map<std::pair<int, unsigned>, std::vector<PassCfg>> g_tuned_batch;  // overrides for fused launches
}  // namespace

int load_schedules(const char *path);

// tuning.txt beside the library, or $GSN_TUNING_FILE, read on first use
static void load_default_schedules()
{
    const char *env = std::getenv("GSN_TUNING_FILE");
    if (env && *env) {
        load_schedules(env);
        return;
    }
    Dl_info info;
    if (dladdr((void *)&load_default_schedules, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        const size_t slash = p.rfind('/');
        p = (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/tuning.txt";
        std::ifstream f(p);
        if (f.good())
            load_schedules(p.c_str());
    }
}

std::vector<PassCfg> choose_passes(int field, unsigned log_n, bool batch)
{
    static std::once_flag once;
    std::call_once(once, load_default_schedules);
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    if (batch) {
        auto ib = g_tuned_batch.find({field, log_n});
        if (ib != g_tuned_batch.end())
            return ib->second;
    }
    auto it = g_tuned.find({field, log_n});
    if (it != g_tuned.end())
        return it->second;
    return default_passes(field, log_n);
}

static int validate_schedule(int field, unsigned log_n, const std::vector<PassCfg> &v)
{
    int sum = 0;
    for (const PassCfg &c : v) {
        bool ok = false;
        switch (field) {
        case GSN_FIELD_FP32: ok = pass_supported<F32>(c); break;
        case GSN_FIELD_MNT4753_FR: ok = pass_supported<FMnt4>(c); break;
        case GSN_FIELD_MNT6753_FR: ok = pass_supported<FMnt6>(c); break;
        default: return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id");
        }
        if (!ok)
            return set_error(GSN_ERR_INVALID_ARGUMENT, "unsupported pass shape");
        sum += c.log_r;
    }
    if (sum != (int)log_n)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "pass radices do not multiply to n");
    // every pass needs at least 2^log_w columns: log_n - log_r >= log_w
    for (const PassCfg &c : v)
        if ((int)log_n - c.log_r < c.log_w)
            return set_error(GSN_ERR_INVALID_ARGUMENT, "tile wider than the pass");
    return GSN_OK;
}

int set_schedule(int field, unsigned log_n, const std::vector<PassCfg> &v, bool batch = false)
{
    GSN_TRY(validate_schedule(field, log_n, v));
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    (batch ? g_tuned_batch : g_tuned)[{field, log_n}] = v;
    return GSN_OK;
}

void clear_schedules()
{
    std::lock_guard<std::mutex> lk(g_tune_mutex);
    g_tuned.clear();
    g_tuned_batch.clear();
}

// "field log_n r,w,e r,w,e ..." per line, "batch field log_n ..." for the
// schedule of transforms fused into batches; '#' starts a comment
int load_schedules(const char *path)
{
    std::ifstream f(path);
    if (!f.good())
        return set_error(GSN_ERR_INVALID_ARGUMENT, std::string("cannot read tuning file ") + path);
    std::string line;
    int lineno = 0;
    while (std::getline(f, line)) {
        ++lineno;
        const size_t hash = line.find('#');
        if (hash != std::string::npos)
            line.resize(hash);
        std::istringstream is(line);
        const bool batch = line.find("batch") != std::string::npos;
        if (batch) {
            std::string word;
            is >> word;
            if (word != "batch")
                return set_error(GSN_ERR_INVALID_ARGUMENT,
                                 std::string(path) + ":" + std::to_string(lineno) + ": 'batch' must come first");
        }
        int field;
        unsigned log_n;
        if (!(is >> field >> log_n))
            continue;
        std::vector<PassCfg> v;
        std::string tok;
        while (is >> tok) {
            PassCfg c{};
            if (std::sscanf(tok.c_str(), "%d,%d,%d", &c.log_r, &c.log_w, &c.log_e) != 3)
                return set_error(GSN_ERR_INVALID_ARGUMENT, std::string(path) + ":" + std::to_string(lineno) +
                                                               ": bad pass shape '" + tok + "'");
            v.push_back(c);
        }
        GSN_TRY(set_schedule(field, log_n, v, batch));
    }
    return GSN_OK;
}

DeviceBuffer &DeviceBuffer::operator=(DeviceBuffer &&o) noexcept
{
    if (this != &o) {
        if (ptr)
            (void)hipFree(ptr);
        ptr = o.ptr;
        bytes = o.bytes;
        o.ptr = nullptr;
        o.bytes = 0;
    }
    return *this;
}

DeviceBuffer::~DeviceBuffer()
{
    if (ptr)
        (void)hipFree(ptr);
}
int DeviceBuffer::alloc(size_t n)
{
    if (ptr && bytes >= n)
        return GSN_OK;
    if (ptr) {
        (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    if (n == 0)
        return GSN_OK;
    GSN_TRY(hip_status(hipMalloc(&ptr, n), "hipMalloc"));
    bytes = n;
    return GSN_OK;
}

template <>
F32::tw_t PlanImpl<GSN_FIELD_FP32>::make_tw(const fields::Fp32 &w)
{
    return F32::tw_t{w.im_rep[0], w.im_rep[0] * Fp32Params::PINV};
}

template <>
TW753 PlanImpl<GSN_FIELD_MNT4753_FR>::make_tw(const fields::Scalar &w)
{
    return make_tw753<fields::Scalar, Mnt4753FrParams>(w);
}
template <>
TW753 PlanImpl<GSN_FIELD_MNT6753_FR>::make_tw(const fields::ScalarMNT6 &w)
{
    return make_tw753<fields::ScalarMNT6, Mnt6753FrParams>(w);
}

template <>
int PlanImpl<GSN_FIELD_FP32>::build_tables(const fields::Fp32 &w, size_t n)
{
    using H = fields::Fp32;
    // table of root^x for x < count
    auto powers = [](const H &root, size_t count) {
        std::vector<tw> t(count);
        H acc = H::one();
        for (size_t x = 0; x < count; ++x) {
            t[x] = make_tw(acc);
            acc *= root;
        }
        return t;
    };
    const size_t P = cfg.size();
    col_tabs.resize(P);
    pass_tabs.resize(P);
    pass_los.resize(P);
    pass_as.resize(P);
    pass_lo_bits.assign(P, 0);
    unsigned log_ns = 0;
    for (size_t i = 0; i < P; ++i) {
        const unsigned lr = (unsigned)cfg[i].log_r;
        // column table: w_R^x, w_R = w^(n/R)
        const std::vector<tw> col = powers(w.pow(n >> lr), (size_t)1 << lr);
        GSN_TRY(upload(col_tabs[i], col.data(), col.size() * sizeof(tw)));
        if (log_ns > 0) {
            // pass table in units of w_{Ns R} = w^(n / (Ns R)); exponents r*k < Ns R.
            // Above 2^12 entries a direct table is read at scattered
            // addresses; two tables of ~2^(u/2) entries stay in L1/L2.
            const unsigned u = log_ns + lr;
            const H root = w.pow(n >> u);
            if (u <= 12) {
                const std::vector<tw> t = powers(root, (size_t)1 << u);
                GSN_TRY(upload(pass_tabs[i], t.data(), t.size() * sizeof(tw)));
            } else {
                const unsigned lo = (u + 1) / 2;
                const std::vector<tw> tl = powers(root, (size_t)1 << lo);
                const std::vector<tw> th = powers(root.pow((uint64_t)1 << lo), (size_t)1 << (u - lo));
                GSN_TRY(upload(pass_los[i], tl.data(), tl.size() * sizeof(tw)));
                GSN_TRY(upload(pass_tabs[i], th.data(), th.size() * sizeof(tw)));
                pass_lo_bits[i] = lo;
                // past 2^17 columns, each thread's base twiddle root^(jj k)
                // directly, laid out so adjacent columns are adjacent: one
                // coalesced load instead of two scattered ones and a
                // product (measured 1-3% faster at 2^26 and 2^28; at 2^16
                // columns the extra table traffic made 2^24 6% slower)
                if (log_ns >= 17) {
                    const size_t rows = (size_t)1 << (lr - std::min(lr, (unsigned)cfg[i].log_e));
                    const size_t ns = (size_t)1 << log_ns;
                    std::vector<tw> ta(rows * ns);
                    H step = H::one();
                    for (size_t jj = 0; jj < rows; ++jj, step *= root) {
                        H acc = H::one();
                        for (size_t k = 0; k < ns; ++k, acc *= step)
                            ta[jj * ns + k] = make_tw(acc);
                    }
                    GSN_TRY(upload(pass_as[i], ta.data(), ta.size() * sizeof(tw)));
                }
            }
        }
        log_ns += lr;
    }
    return GSN_OK;
}

template <class Plan, class H, class F, class P>
static int build_tables_753(Plan &plan, const H &w, unsigned log_n, size_t n)
{
    // every record is w_n^e = hi[e >> h] * lo[e & (2^h - 1)] for e < n/2,
    // computed on the device from hi (canonical residues) and lo (records)
    const unsigned h = (log_n + 1) / 2;
    const size_t nlo = (size_t)1 << h, nhi = std::max<size_t>(1, (n / 2) >> h);
    std::vector<TW753> lo(nlo);
    std::vector<E753> hi(nhi);
    H acc = H::one();
    for (size_t x = 0; x < nlo; ++x) {
        lo[x] = make_tw753<H, P>(acc);
        acc *= w;
    }
    const H step = acc;
    acc = H::one();
    for (size_t y = 0; y < nhi; ++y) {
        acc.to_canonical(hi[y].v);  // w^(y 2^h)
        acc *= step;
    }
    DeviceBuffer dlo, dhi;
    GSN_TRY(dlo.alloc(nlo * sizeof(TW753)));
    GSN_TRY(dhi.alloc(nhi * sizeof(E753)));
    GSN_TRY(hip_status(hipMemcpy(dlo.ptr, lo.data(), nlo * sizeof(TW753), hipMemcpyHostToDevice), "hipMemcpy"));
    GSN_TRY(hip_status(hipMemcpy(dhi.ptr, hi.data(), nhi * sizeof(E753), hipMemcpyHostToDevice), "hipMemcpy"));
    const size_t P_ = plan.cfg.size();
    plan.step_tabs.resize(P_);
    plan.step_nrec.assign(P_, 0);
    unsigned log_ns = 0;
    for (size_t i = 0; i < P_; ++i) {
        const unsigned lr = (unsigned)plan.cfg[i].log_r;
        const uint32_t nrec = (uint32_t)((((size_t)1 << lr) - 1) << log_ns);
        // records, then the scale record (n^-1 for an inverse plan)
        GSN_TRY(plan.step_tabs[i].alloc((size_t)(nrec + 1) * sizeof(TW753)));
        plan.step_nrec[i] = nrec + 1;
        // the inverse's n^-1 rides on the last pass's step-0 records when
        // there is more than one pass (PASS_SCALE_IN)
        const bool scale0 = plan.inverse && P_ >= 2 && i == P_ - 1;
        launch_build_steps753<F>((uint4 *)plan.step_tabs[i].ptr, nrec, (const E753 *)dhi.ptr, (const TW753 *)dlo.ptr,
                                 h, log_n, log_ns, lr, plan.scale, scale0, 0);
        GSN_TRY(hip_status(hipGetLastError(), "build_steps753 launch"));
        log_ns += lr;
    }
    return hip_status(hipStreamSynchronize(nullptr), "build_steps753");
}

template <>
int PlanImpl<GSN_FIELD_MNT4753_FR>::build_tables(const fields::Scalar &w, size_t n)
{
    return build_tables_753<PlanImpl, fields::Scalar, FMnt4, Mnt4753FrParams>(*this, w, log_n, n);
}
template <>
int PlanImpl<GSN_FIELD_MNT6753_FR>::build_tables(const fields::ScalarMNT6 &w, size_t n)
{
    return build_tables_753<PlanImpl, fields::ScalarMNT6, FMnt6, Mnt6753FrParams>(*this, w, log_n, n);
}

template <int FID>
static int create_plan_t(PlanBase **out, unsigned log_n, const void *omega, int inverse, bool batch)
{
    using Info = FieldInfo<FID>;
    using H = typename Info::Host;
    if (log_n > Info::TWO_ADICITY)
        return set_error(GSN_ERR_SIZE_EXCEEDS_TWO_ADICITY,
                         "log2(n) = " + std::to_string(log_n) + " exceeds the field's 2-adicity " +
                             std::to_string(Info::TWO_ADICITY));
    H w;
    if (omega) {
        w = H::from_raw((const uint32_t *)omega);
        if (!valid_root(w, log_n))
            return set_error(GSN_ERR_BAD_OMEGA, "omega is not a primitive 2^log_n-th root of unity in canonical form");
    } else {
        w = root_of_unity(Info::root_max(), Info::TWO_ADICITY, log_n);
    }
    GSN_TRY(ensure_device());
    std::unique_ptr<PlanImpl<FID>> p(new PlanImpl<FID>());
    p->field = FID;
    p->log_n = log_n;
    p->inverse = inverse ? 1 : 0;
    p->batch_mode = batch;
    GSN_TRY(p->build(w));
    *out = p.release();
    return GSN_OK;
}

int create_plan(PlanBase **out, int field, unsigned log_n, const void *omega, int inverse, bool batch)
{
    if (!out)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "plan pointer is NULL");
    switch (field) {
    case GSN_FIELD_FP32: return create_plan_t<GSN_FIELD_FP32>(out, log_n, omega, inverse, batch);
    case GSN_FIELD_MNT4753_FR: return create_plan_t<GSN_FIELD_MNT4753_FR>(out, log_n, omega, inverse, batch);
    case GSN_FIELD_MNT6753_FR: return create_plan_t<GSN_FIELD_MNT6753_FR>(out, log_n, omega, inverse, batch);
    default: return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id " + std::to_string(field));
    }
}

size_t element_bytes(int field)
{
    switch (field) {
    case GSN_FIELD_FP32: return 4;
    case GSN_FIELD_MNT4753_FR:
    case GSN_FIELD_MNT6753_FR: return 96;
    default: return 0;
    }
}

int two_adicity(int field)
{
    switch (field) {
    case GSN_FIELD_FP32: return (int)FieldInfo<GSN_FIELD_FP32>::TWO_ADICITY;
    case GSN_FIELD_MNT4753_FR: return (int)FieldInfo<GSN_FIELD_MNT4753_FR>::TWO_ADICITY;
    case GSN_FIELD_MNT6753_FR: return (int)FieldInfo<GSN_FIELD_MNT6753_FR>::TWO_ADICITY;
    default: return -1;
    }
}

int log2_exact(size_t n, unsigned *log_n)
{
    if (n == 0 || (n & (n - 1)))
        return set_error(GSN_ERR_NOT_POWER_OF_TWO, "n = " + std::to_string(n) + " is not a power of two");
    unsigned l = 0;
    while (((size_t)1 << l) < n)
        ++l;
    *log_n = l;
    return GSN_OK;
}

// Plans for the host entry point, cached per (field, size, direction, omega).
namespace {
std::mutex g_cache_mutex;
std::map<std::string, std::unique_ptr<PlanBase>> g_plan_cache;
// plans dropped from the cache stay alive: batches may still point at them
std::vector<std::unique_ptr<PlanBase>> g_retired_plans;
}  // namespace

void clear_plan_cache()
{
    std::lock_guard<std::mutex> lk(g_cache_mutex);
    for (auto &kv : g_plan_cache)
        g_retired_plans.push_back(std::move(kv.second));
    g_plan_cache.clear();
}

int cached_plan(PlanBase **out, int field, unsigned log_n, const void *omega, int inverse, bool batch)
{
    std::string key(16, '\0');
    std::memcpy(&key[0], &field, 4);
    std::memcpy(&key[4], &log_n, 4);
    std::memcpy(&key[8], &inverse, 4);
    key[12] = batch ? 1 : 0;
    if (omega)
        key.append((const char *)omega, element_bytes(field));
    std::lock_guard<std::mutex> lk(g_cache_mutex);
    auto it = g_plan_cache.find(key);
    if (it != g_plan_cache.end()) {
        *out = it->second.get();
        return GSN_OK;
    }
    PlanBase *p = nullptr;
    GSN_TRY(create_plan(&p, field, log_n, omega, inverse, batch));
    g_plan_cache[key].reset(p);
    *out = p;
    return GSN_OK;
}

}  // namespace gsn

// ================================================================ C ABI

using namespace gsn;

extern "C" {

const char *gsn_strerror(int status)
{
    switch (status) {
    case GSN_OK: return "ok";
    case GSN_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GSN_ERR_NOT_POWER_OF_TWO: return "size is not a power of two";
    case GSN_ERR_SIZE_EXCEEDS_TWO_ADICITY: return "size exceeds the field's 2-adicity";
    case GSN_ERR_BAD_OMEGA: return "omega is not a primitive root of unity of the requested order";
    case GSN_ERR_NO_DEVICE: return "no HIP device";
    case GSN_ERR_DEVICE: return "HIP runtime error";
    case GSN_ERR_OUT_OF_MEMORY: return "out of device memory";
    case GSN_ERR_ADMISSION: return "request exceeds the admission budget";
    case GSN_ERR_NOT_READY: return "not ready";
    case GSN_ERR_INJECTED_FAULT: return "injected fault";
    case GSN_ERR_SHUTDOWN: return "server is shutting down";
    default: return "unknown status";
    }
}

const char *gsn_last_error(void) { return t_last_error.c_str(); }

int gsn_version(void) { return 100; }

int gsn_device_init(int device)
{
    std::lock_guard<std::mutex> lk(g_dev_mutex);
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return set_error(GSN_ERR_NO_DEVICE, "no HIP device available");
    if (device < 0 || device >= count)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "device index out of range");
    GSN_TRY(hip_status(hipSetDevice(device), "hipSetDevice"));
    g_device = device;
    return GSN_OK;
}

size_t gsn_element_bytes(int field) { return element_bytes(field); }

int gsn_two_adicity(int field) { return two_adicity(field); }

int gsn_root_of_unity(int field, unsigned log_n, void *omega_out)
{
    if (!omega_out)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "omega_out is NULL");
    const int s = two_adicity(field);
    if (s < 0)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id");
    if ((int)log_n > s)
        return set_error(GSN_ERR_SIZE_EXCEEDS_TWO_ADICITY, "log2(n) exceeds the field's 2-adicity");
    switch (field) {
    case GSN_FIELD_FP32: {
        auto w = root_of_unity(FieldInfo<0>::root_max(), FieldInfo<0>::TWO_ADICITY, log_n);
        std::memcpy(omega_out, w.im_rep, sizeof(w.im_rep));
        break;
    }
    case GSN_FIELD_MNT4753_FR: {
        auto w = root_of_unity(FieldInfo<1>::root_max(), FieldInfo<1>::TWO_ADICITY, log_n);
        std::memcpy(omega_out, w.im_rep, sizeof(w.im_rep));
        break;
    }
    default: {
        auto w = root_of_unity(FieldInfo<2>::root_max(), FieldInfo<2>::TWO_ADICITY, log_n);
        std::memcpy(omega_out, w.im_rep, sizeof(w.im_rep));
        break;
    }
    }
    return GSN_OK;
}

int gsn_debug_inject_faults(int count)
{
    if (count < 0)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "negative fault count");
    g_inject_faults.store(count);
    return GSN_OK;
}

int gsn_plan_create(gsn_plan **plan, int field, unsigned log_n, const void *omega, int inverse)
{
    if (!plan)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "plan is NULL");
    *plan = nullptr;
    PlanBase *impl = nullptr;
    GSN_TRY(create_plan(&impl, field, log_n, omega, inverse));
    *plan = new gsn_plan{impl};
    return GSN_OK;
}

int gsn_plan_destroy(gsn_plan *plan)
{
    if (!plan)
        return GSN_OK;
    delete plan->impl;
    delete plan;
    return GSN_OK;
}

int gsn_plan_execute(gsn_plan *plan, const void *d_in, void *d_out, void *stream)
{
    if (!plan || !d_in || !d_out)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL plan or buffer");
    GSN_TRY(ensure_device());
    return plan->impl->execute(d_in, d_out, (hipStream_t)stream);
}

}  // extern "C"

// The allocation-free entry point's workspace: one array per device, held
// until the process ends.  Its name is a host-side stand-in, not a device
// address: host code uses only the address hipGetSymbolAddress returns for
// the current device (noalloc_workspace).
static __device__ __align__(256) unsigned char noalloc_scratch[GSN_NOALLOC_WORKSPACE_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_noalloc_mutex;
void *g_noalloc_ptr[MAX_DEVICES];  // per device; nullptr until looked up
}  // namespace

// The device address of the current device's noalloc_scratch, looked up
// once per device with that device current.
static int noalloc_workspace(void **out)
{
    int dev = 0;
    GSN_TRY(hip_status(hipGetDevice(&dev), "hipGetDevice"));
    if (dev < 0 || dev >= MAX_DEVICES)
        return set_error(GSN_ERR_DEVICE, "device index out of range for the allocation-free workspace");
    std::lock_guard<std::mutex> lk(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void *p = nullptr;
        GSN_TRY(hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)), "hipGetSymbolAddress"));
        if (!p)
            return set_error(GSN_ERR_DEVICE, "hipGetSymbolAddress returned no address");
        g_noalloc_ptr[dev] = p;
    }
    *out = g_noalloc_ptr[dev];
    return GSN_OK;
}

extern "C" {

int gsn_plan_execute_noalloc(gsn_plan *plan, const void *d_in, void *d_out, void *stream)
{
    if (!plan || !d_in || !d_out)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL plan or buffer");
    GSN_TRY(ensure_device());
    void *ws = nullptr;
    GSN_TRY(noalloc_workspace(&ws));
    return plan->impl->execute_workspace(d_in, d_out, ws, GSN_NOALLOC_WORKSPACE_BYTES, (hipStream_t)stream);
}

size_t gsn_plan_workspace_bytes(const gsn_plan *plan) { return plan ? plan->impl->workspace_bytes() : 0; }

size_t gsn_plan_device_bytes(const gsn_plan *plan) { return plan ? plan->impl->device_bytes : 0; }

int gsn_plan_num_passes(const gsn_plan *plan) { return plan ? plan->impl->num_passes : 0; }

// Every element below p (the kernels' value bounds assume canonical inputs):
// the top limb decides unless it equals p's.
extern "C++" template <int L>
static int check_canonical(const uint32_t *x, size_t n, const uint32_t *mod)
{
    for (size_t i = 0; i < n; ++i) {
        const uint32_t *e = x + i * L;
        if (e[L - 1] < mod[L - 1])
            continue;
        if (e[L - 1] > mod[L - 1] || fields::detail::geq<L>(e, mod))
            return set_error(GSN_ERR_INVALID_ARGUMENT, "element " + std::to_string(i) + " is not below p");
    }
    return GSN_OK;
}

int gsn_fft_host(int field, void *data, size_t n, const void *omega, int inverse)
{
    if (!data)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "data is NULL");
    const size_t eb = element_bytes(field);
    if (!eb)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id");
    unsigned log_n = 0;
    GSN_TRY(log2_exact(n, &log_n));
    switch (field) {
    case GSN_FIELD_FP32: GSN_TRY(check_canonical<1>((const uint32_t *)data, n, &Fp32Params::P)); break;
    case GSN_FIELD_MNT4753_FR:
        GSN_TRY(check_canonical<24>((const uint32_t *)data, n, Mnt4753FrParams::MOD));
        break;
    default: GSN_TRY(check_canonical<24>((const uint32_t *)data, n, Mnt6753FrParams::MOD)); break;
    }
    PlanBase *p = nullptr;
    GSN_TRY(cached_plan(&p, field, log_n, omega, inverse));
    std::lock_guard<std::mutex> lk(p->host_mutex);
    GSN_TRY(p->host_buf.alloc(n * eb));
    GSN_TRY(staged_h2d(p->host_buf.ptr, data, n * eb, nullptr));
    GSN_TRY(p->execute(p->host_buf.ptr, p->host_buf.ptr, nullptr));
    GSN_TRY(staged_d2h(data, p->host_buf.ptr, n * eb, nullptr));
    return GSN_OK;
}

// ---------------------------------------------------------------- tuning

int gsn_tuning_set(int field, unsigned log_n, const int *shapes, int num_passes)
{
    if (num_passes < 0 || (num_passes > 0 && !shapes))
        return set_error(GSN_ERR_INVALID_ARGUMENT, "bad schedule arguments");
    std::vector<PassCfg> v;
    for (int i = 0; i < num_passes; ++i)
        v.push_back({shapes[3 * i], shapes[3 * i + 1], shapes[3 * i + 2]});
    (void)choose_passes(field, log_n);  // load the default file first, so it cannot override this
    GSN_TRY(set_schedule(field, log_n, v));
    clear_plan_cache();
    return GSN_OK;
}

int gsn_tuning_get(int field, unsigned log_n, int *shapes, int max_passes, int *num_passes)
{
    if (!num_passes)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "num_passes is NULL");
    const std::vector<PassCfg> v = choose_passes(field, log_n);
    *num_passes = (int)v.size();
    if ((int)v.size() > max_passes || (!shapes && !v.empty()))
        return set_error(GSN_ERR_INVALID_ARGUMENT, "shapes buffer too small");
    for (size_t i = 0; i < v.size(); ++i) {
        shapes[3 * i] = v[i].log_r;
        shapes[3 * i + 1] = v[i].log_w;
        shapes[3 * i + 2] = v[i].log_e;
    }
    return GSN_OK;
}

int gsn_tuning_clear(void)
{
    (void)choose_passes(GSN_FIELD_FP32, 1);
    clear_schedules();
    clear_plan_cache();
    return GSN_OK;
}

int gsn_tuning_load(const char *path)
{
    if (!path)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "path is NULL");
    (void)choose_passes(GSN_FIELD_FP32, 1);
    GSN_TRY(load_schedules(path));
    clear_plan_cache();
    return GSN_OK;
}

// ---------------------------------------------------------------- plumbing

int gsn_device_malloc(void **ptr, size_t bytes)
{
    if (!ptr)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "ptr is NULL");
    *ptr = nullptr;
    GSN_TRY(ensure_device());
    return hip_status(hipMalloc(ptr, bytes ? bytes : 1), "hipMalloc");
}

int gsn_device_free(void *ptr)
{
    if (!ptr)
        return GSN_OK;
    return hip_status(hipFree(ptr), "hipFree");
}

int gsn_device_memory(size_t *free_bytes, size_t *total_bytes)
{
    if (!free_bytes || !total_bytes)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL output");
    GSN_TRY(ensure_device());
    return hip_status(hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
}

int gsn_device_synchronize(void)
{
    GSN_TRY(ensure_device());
    return hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

static int copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, void *stream)
{
    if ((!dst || !src) && bytes)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL buffer");
    GSN_TRY(ensure_device());
    GSN_TRY(hip_status(hipMemcpyAsync(dst, src, bytes, kind, (hipStream_t)stream), "hipMemcpyAsync"));
    return hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int gsn_memcpy_htod(void *d_dst, const void *h_src, size_t bytes, void *stream)
{
    return copy_sync(d_dst, h_src, bytes, hipMemcpyHostToDevice, stream);
}

int gsn_memcpy_dtoh(void *h_dst, const void *d_src, size_t bytes, void *stream)
{
    return copy_sync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, stream);
}

int gsn_memcpy_dtod(void *d_dst, const void *d_src, size_t bytes, void *stream)
{
    return copy_sync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, stream);
}

int gsn_stream_create(void **stream)
{
    if (!stream)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "stream is NULL");
    GSN_TRY(ensure_device());
    hipStream_t s = nullptr;
    GSN_TRY(hip_status(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreate"));
    *stream = s;
    return GSN_OK;
}

int gsn_stream_destroy(void *stream)
{
    if (!stream)
        return GSN_OK;
    return hip_status(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}

int gsn_stream_synchronize(void *stream)
{
    GSN_TRY(ensure_device());
    return hip_status(hipStreamSynchronize((hipStream_t)stream), "hipStreamSynchronize");
}

int gsn_event_create(void **event)
{
    if (!event)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "event is NULL");
    GSN_TRY(ensure_device());
    hipEvent_t e = nullptr;
    GSN_TRY(hip_status(hipEventCreate(&e), "hipEventCreate"));
    *event = e;
    return GSN_OK;
}

int gsn_event_destroy(void *event)
{
    if (!event)
        return GSN_OK;
    return hip_status(hipEventDestroy((hipEvent_t)event), "hipEventDestroy");
}

int gsn_event_record(void *event, void *stream)
{
    return hip_status(hipEventRecord((hipEvent_t)event, (hipStream_t)stream), "hipEventRecord");
}

int gsn_event_synchronize(void *event)
{
    return hip_status(hipEventSynchronize((hipEvent_t)event), "hipEventSynchronize");
}

int gsn_debug_copy_u32(void *d_out, const void *d_in, size_t count, void *stream)
{
    if (!d_out || !d_in)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL buffer");
    GSN_TRY(ensure_device());
    launch_scale<F32>((const uint32_t *)d_in, (uint32_t *)d_out, F32::tw_t{}, count, 0, (hipStream_t)stream);
    return hip_status(hipGetLastError(), "copy launch");
}

int gsn_debug_clock_ghz(double *ghz)
{
    if (!ghz)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "ghz is NULL");
    GSN_TRY(ensure_device());
    int dev = 0, cus = 0, wall_khz = 0;
    GSN_TRY(hip_status(hipGetDevice(&dev), "hipGetDevice"));
    GSN_TRY(hip_status(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count"));
    GSN_TRY(hip_status(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate"));
    const unsigned blocks = (unsigned)cus * 4;  // 4 waves per SIMD
    DeviceBuffer spans, sink;
    GSN_TRY(spans.alloc((size_t)blocks * 2 * sizeof(unsigned long long)));
    GSN_TRY(sink.alloc(sizeof(E753)));
    for (int rep = 0; rep < 2; ++rep) {  // the first run lets the clock settle
        launch_clock_probe((unsigned long long *)spans.ptr, (E753 *)sink.ptr, blocks, 256, nullptr);
        GSN_TRY(hip_status(hipGetLastError(), "clock probe launch"));
    }
    GSN_TRY(hip_status(hipStreamSynchronize(nullptr), "clock probe"));
    std::vector<unsigned long long> h((size_t)blocks * 2);
    GSN_TRY(hip_status(hipMemcpy(h.data(), spans.ptr, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost), "hipMemcpy"));
    double cyc = 0, wall = 0;
    for (unsigned b = 0; b < blocks; ++b) {
        cyc += (double)h[2 * b];
        wall += (double)h[2 * b + 1];
    }
    if (wall <= 0 || wall_khz <= 0)
        return set_error(GSN_ERR_DEVICE, "clock probe measured nothing");
    *ghz = cyc / wall * (double)wall_khz * 1e-6;
    return GSN_OK;
}

int gsn_debug_mad_rate(double *mads_per_s, double *ghz)
{
    if (!mads_per_s || !ghz)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL argument");
    GSN_TRY(ensure_device());
    int dev = 0, cus = 0, wall_khz = 0;
    GSN_TRY(hip_status(hipGetDevice(&dev), "hipGetDevice"));
    GSN_TRY(hip_status(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count"));
    GSN_TRY(hip_status(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, dev), "wall clock rate"));
    const unsigned blocks = (unsigned)cus * 8;  // 8 waves per SIMD
    const int iters = 2048;
    DeviceBuffer spans, sink;
    GSN_TRY(spans.alloc((size_t)blocks * 2 * sizeof(unsigned long long)));
    GSN_TRY(sink.alloc(sizeof(unsigned long long)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    GSN_TRY(hip_status(hipEventCreate(&e0), "hipEventCreate"));
    int rc = hip_status(hipEventCreate(&e1), "hipEventCreate");
    uint64_t mads = 0;
    if (rc == GSN_OK) {
        // the first launch lets the clock settle; the second is timed
        launch_mad_rate_probe((unsigned long long *)spans.ptr, (unsigned long long *)sink.ptr, blocks, iters, nullptr);
        rc = hip_status(hipEventRecord(e0, nullptr), "hipEventRecord");
        mads = launch_mad_rate_probe((unsigned long long *)spans.ptr, (unsigned long long *)sink.ptr, blocks, iters,
                                     nullptr);
        if (rc == GSN_OK)
            rc = hip_status(hipGetLastError(), "mad rate probe launch");
        if (rc == GSN_OK)
            rc = hip_status(hipEventRecord(e1, nullptr), "hipEventRecord");
    }
    float ms = 0.f;
    if (rc == GSN_OK)
        rc = hip_status(hipEventSynchronize(e1), "mad rate probe");
    if (rc == GSN_OK)
        rc = hip_status(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
    (void)hipEventDestroy(e0);
    if (e1)
        (void)hipEventDestroy(e1);
    GSN_TRY(rc);
    std::vector<unsigned long long> h((size_t)blocks * 2);
    GSN_TRY(hip_status(hipMemcpy(h.data(), spans.ptr, h.size() * sizeof(h[0]), hipMemcpyDeviceToHost), "hipMemcpy"));
    double cyc = 0, wall = 0;
    for (unsigned b = 0; b < blocks; ++b) {
        cyc += (double)h[2 * b];
        wall += (double)h[2 * b + 1];
    }
    if (ms <= 0.f || wall <= 0 || wall_khz <= 0)
        return set_error(GSN_ERR_DEVICE, "mad rate probe measured nothing");
    *mads_per_s = (double)mads / (ms * 1e-3);
    *ghz = cyc / wall * (double)wall_khz * 1e-6;
    return GSN_OK;
}

int gsn_event_elapsed_ms(float *ms, void *start, void *stop)
{
    if (!ms)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "ms is NULL");
    return hip_status(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop), "hipEventElapsedTime");
}

}  // extern "C"
