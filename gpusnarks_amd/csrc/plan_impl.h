// Transform plans: pass schedule, device-resident twiddle tables, workspace.
// Shared by the plain-transform runtime and the batch scheduler.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "device_field.h"
#include "gpusnarks.h"
#include "gsn_internal.h"
#include "gsn_launch.h"
#include "tuning.h"

namespace gsn {

// ------------------------------------------------------------ fields

template <int FID>
struct FieldInfo;
template <>
struct FieldInfo<GSN_FIELD_FP32> {
    using Host = fields::Fp32;
    using Dev = F32;
    static constexpr unsigned TWO_ADICITY = Fp32Params::TWO_ADICITY;
    static Host root_max() { return Host::from_raw(&Fp32Params::ROOT_MONT); }
};
template <>
struct FieldInfo<GSN_FIELD_MNT4753_FR> {
    using Host = fields::Scalar;
    using Dev = FMnt4;
    using Params = Mnt4753FrParams;
    static constexpr unsigned TWO_ADICITY = Params::TWO_ADICITY;
    static Host root_max() { return Host::from_raw(Params::ROOT_MONT); }
};
template <>
struct FieldInfo<GSN_FIELD_MNT6753_FR> {
    using Host = fields::ScalarMNT6;
    using Dev = FMnt6;
    using Params = Mnt6753FrParams;
    static constexpr unsigned TWO_ADICITY = Params::TWO_ADICITY;
    static Host root_max() { return Host::from_raw(Params::ROOT_MONT); }
};

template <class H>
H root_of_unity(H root_max, unsigned two_adicity, unsigned log_n)
{
    H w = root_max;
    for (unsigned i = log_n; i < two_adicity; ++i)
        w = w.square();
    return w;
}

template <class H>
bool is_canonical(const H &x)
{
    return !fields::detail::geq<H::LIMBS>(x.im_rep, H::mod());
}

// omega^n == 1 and omega^(n/2) == -1 (n >= 2), omega canonical
template <class H>
bool valid_root(const H &w, unsigned log_n)
{
    if (!is_canonical(w))
        return false;
    if (log_n == 0)
        return w == H::one();
    H x = w;
    for (unsigned i = 0; i + 1 < log_n; ++i)
        x = x.square();
    return x == -H::one();  // then x^2 = 1 as well
}

// The device twiddle record of the field element w (given in Montgomery
// form): its residue as 29-bit digits and wq = floor(w 2^783 / p).
template <class H, class P>
TW753 make_tw753(const H &w)
{
    uint32_t v[24], rem[24], d[27];
    w.to_canonical(v);
    (w * H::from_canonical(P::C783)).to_canonical(rem);  // w 2^783 mod p
    TW753 t{};
    limbs_to_digits27(d, v);
    for (int i = 0; i < 26; ++i)
        t.w[i] = d[i];
    shoup_wq<P>(t.wq, rem);
    return t;
}

// ------------------------------------------------------------ plans

template <int FID>
struct PlanImpl final : PlanBase {
    using Info = FieldInfo<FID>;
    using H = typename Info::Host;
    using F = typename Info::Dev;
    using elem = typename F::elem_t;
    using tw = typename F::tw_t;
    static constexpr bool IS32 = (FID == GSN_FIELD_FP32);

    std::vector<PassCfg> cfg;
    // 753-bit, per pass: step table of Ns (2^r - 1) records in 14 planes
    std::vector<DeviceBuffer> step_tabs;
    std::vector<uint32_t> step_nrec;
    // 32-bit, per pass: column table w_R^x (x < R) and pass table in units
    // of w_{Ns R} -- direct, or hi/lo halves of a two-level split
    std::vector<DeviceBuffer> col_tabs, pass_tabs, pass_los, pass_as;
    std::vector<uint32_t> pass_lo_bits;
    DeviceBuffer scratch[2];
    tw scale{};

    static tw make_tw(const H &w);

    int build(const H &omega_fwd)
    {
        const H w = inverse ? omega_fwd.mul_inv() : omega_fwd;
        const size_t n = (size_t)1 << log_n;
        cfg = choose_passes(FID, log_n, batch_mode);
        num_passes = (int)cfg.size();
        for (const PassCfg &c : cfg)
            if (!pass_supported<F>(c))
                return set_error(GSN_ERR_INVALID_ARGUMENT, "no kernel for the pass configuration");
        scale = make_tw(H((uint32_t)n).mul_inv());
        if (log_n == 0)
            return GSN_OK;
        GSN_TRY(build_tables(w, n));
        device_bytes = 0;
        for (auto *v : {&col_tabs, &pass_tabs, &pass_los, &pass_as, &step_tabs})
            for (auto &b : *v)
                device_bytes += b.bytes;
        return GSN_OK;
    }

    // workspace buffers an execution needs (0, 1 or 2 of n elements)
    int workspace_buffers() const { return cfg.size() >= 3 ? 2 : (cfg.size() == 2 ? 1 : 0); }

    int ensure_scratch()
    {
        const size_t bytes = ((size_t)1 << log_n) * sizeof(elem);
        for (int i = 0; i < workspace_buffers(); ++i)
            GSN_TRY(scratch[i].alloc(bytes));
        return GSN_OK;
    }

    int upload(DeviceBuffer &buf, const void *src, size_t bytes)
    {
        GSN_TRY(buf.alloc(bytes));
        return hip_status(hipMemcpy(buf.ptr, src, bytes, hipMemcpyHostToDevice), "hipMemcpy(twiddles)");
    }

    int build_tables(const H &w, size_t n);

    // Pass jobs for in -> out with workspace ws[0..1]: buffer i (0 < i < P)
    // alternates ws[0], ws[1], so no pass reads the buffer it writes.
    void make_jobs(const void *in, void *out, void *ws0, void *ws1, std::vector<PassJob<F>> &jobs) const
    {
        jobs.resize(cfg.size());
        make_jobs(in, out, ws0, ws1, jobs.data());
    }
    // The same into jobs[0 .. cfg.size()) (at most MAX_PASSES entries).
    static constexpr int MAX_PASSES = 32;
    void make_jobs(const void *in, void *out, void *ws0, void *ws1, PassJob<F> *jobs) const
    {
        const int P = (int)cfg.size();
        uint32_t log_ns = 0;
        void *ws[2] = {ws0, ws1};
        for (int i = 0; i < P; ++i) {
            const PassCfg &c = cfg[i];
            PassJob<F> jb{};
            PassArgs<F> &a = jb.a;
            a.in = (const elem *)(i == 0 ? in : ws[(i - 1) & 1]);
            a.out = (elem *)(i == P - 1 ? out : ws[i & 1]);
            if (IS32) {
                a.tw_col = (const tw *)col_tabs[i].ptr;
                a.col_shift = 0;
                a.tw_pass = (const tw *)pass_tabs[i].ptr;
                a.tw_pass_lo = (const tw *)pass_los[i].ptr;  // nullptr: direct table
                a.tw_pass_a = (const tw *)pass_as[i].ptr;    // nullptr: from the pass table
                a.lo_bits = pass_lo_bits[i];
                a.pass_shift = 0;
            } else {
                // the 753-bit kernel folds the pass twiddles into its step
                // twiddles, read from the pass's step table
                a.tw_col = (const tw *)step_tabs[i].ptr;
                a.tw_stride = step_nrec[i];
                a.col_shift = 0;
                a.tw_pass = nullptr;
                a.tw_pass_lo = nullptr;
                a.tw_pass_a = nullptr;
                a.lo_bits = 0;
                a.pass_shift = 0;
            }
            a.log_n = log_n;
            a.log_ns = log_ns;
            // the inverse's n^-1: for the 753-bit field with two or more
            // passes, on the last pass's inputs (its step-0 records carry it)
            const uint32_t scale_flag = !IS32 && P >= 2 ? PASS_SCALE_IN : PASS_SCALE;
            a.flags = (i == P - 1) ? (PASS_CANON | (inverse ? scale_flag : 0u)) : 0u;
            a.scale = scale;
            jb.log_r = (uint32_t)c.log_r;
            jb.log_w = (uint32_t)c.log_w;
            jb.num_tiles = (uint32_t)((((size_t)1 << log_n) >> c.log_r) >> c.log_w);
            jb.first_tile = 0;
            jobs[i] = jb;
            log_ns += (uint32_t)c.log_r;
        }
    }

    int execute(const void *in, void *out, hipStream_t s) override
    {
        GSN_TRY(take_injected_fault());
        const size_t n = (size_t)1 << log_n;
        if (cfg.empty()) {  // n == 1: canonical copy
            launch_scale<F>((const elem *)in, (elem *)out, scale, n, 0, s);
            return hip_status(hipGetLastError(), "scale_kernel launch");
        }
        GSN_TRY(ensure_scratch());
        std::vector<PassJob<F>> jobs;
        make_jobs(in, out, scratch[0].ptr, scratch[1].ptr, jobs);
        for (size_t i = 0; i < jobs.size(); ++i) {
            launch_pass<F>(jobs[i], cfg[i], s);
            GSN_TRY(hip_status(hipGetLastError(), "ntt pass launch"));
        }
        return GSN_OK;
    }

    // Device bytes of the workspace an execution over a caller's buffer
    // needs: the second buffer starts at a 256-byte boundary.
    size_t workspace_bytes() const override
    {
        const size_t one = ((size_t)1 << log_n) * sizeof(elem);
        const int nb = workspace_buffers();
        return nb == 0 ? 0 : nb == 1 ? one : ((one + 255) & ~(size_t)255) + one;
    }

    // execute() with the workspace in the caller's device buffer `ws` of
    // ws_bytes: no allocation of any kind (the jobs live on the stack).
    int execute_workspace(const void *in, void *out, void *ws, size_t ws_bytes, hipStream_t s) override
    {
        GSN_TRY(take_injected_fault());
        const size_t n = (size_t)1 << log_n;
        if (cfg.empty()) {  // n == 1: canonical copy
            launch_scale<F>((const elem *)in, (elem *)out, scale, n, 0, s);
            return hip_status(hipGetLastError(), "scale_kernel launch");
        }
        if (cfg.size() > (size_t)MAX_PASSES)
            return set_error(GSN_ERR_INVALID_ARGUMENT, "more passes than the allocation-free path holds");
        const size_t need = workspace_bytes();
        if (need > ws_bytes)
            return set_error(GSN_ERR_INVALID_ARGUMENT,
                             "the transform needs " + std::to_string(need) + " workspace bytes, more than the " +
                                 std::to_string(ws_bytes) + " available");
        char *w0 = (char *)ws;
        char *w1 = w0 + ((n * sizeof(elem) + 255) & ~(size_t)255);
        PassJob<F> jobs[MAX_PASSES];
        make_jobs(in, out, w0, w1, jobs);
        for (size_t i = 0; i < cfg.size(); ++i) {
            launch_pass<F>(jobs[i], cfg[i], s);
            GSN_TRY(hip_status(hipGetLastError(), "ntt pass launch"));
        }
        return GSN_OK;
    }
};

// twiddle records of field elements (defined in gsn_runtime.hip)
template <>
F32::tw_t PlanImpl<GSN_FIELD_FP32>::make_tw(const fields::Fp32 &w);
template <>
TW753 PlanImpl<GSN_FIELD_MNT4753_FR>::make_tw(const fields::Scalar &w);
template <>
TW753 PlanImpl<GSN_FIELD_MNT6753_FR>::make_tw(const fields::ScalarMNT6 &w);

}  // namespace gsn
