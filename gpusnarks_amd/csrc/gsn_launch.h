// Kernel argument types and launch entry points.  The kernels live in one
// translation unit per field (kernels_*.hip) so the large 753-bit kernels
// compile in parallel with the rest.
#pragma once

#include <hip/hip_runtime.h>

#include "field_core.h"

namespace gsn {

enum : uint32_t {
    PASS_SCALE = 1u,  // multiply outputs by `scale` (n^-1 for the inverse)
    PASS_CANON = 2u,  // reduce outputs to [0, p)
    // 753-bit passes after the first: multiply the first step's x0 by
    // `scale`; the pass's step-0 records carry the same factor for x1, so
    // the pass scales its inputs for one product per butterfly column
    // instead of two per column at the end
    PASS_SCALE_IN = 4u,
};

// Everything one Stockham pass of one transform needs.
template <class F>
struct PassArgs {
    const typename F::elem_t *in;
    typename F::elem_t *out;
    const typename F::tw_t *tw_col;
    const typename F::tw_t *tw_pass;
    const typename F::tw_t *tw_pass_lo;
    // 32-bit passes with a two-level pass table: w_{Ns R}^(jj k) at
    // jj Ns + k (jj < R / E, k < Ns), read by adjacent lanes at adjacent k
    const typename F::tw_t *tw_pass_a;
    uint32_t log_n;
    uint32_t log_ns;
    uint32_t col_shift;   // column table index = exponent << col_shift
    uint32_t pass_shift;  // pass table exponent = (r * k) << pass_shift
    uint32_t lo_bits;     // two-level pass table split (0: one table)
    uint32_t tw_stride;   // 753-bit: quads per plane of the pass's step table (records + 1)
    uint32_t flags;
    typename F::tw_t scale;
};

// A pass plus its tile shape and its first tile within a multi-job launch.
template <class F>
struct PassJob {
    PassArgs<F> a;
    uint32_t log_r;
    uint32_t log_w;
    uint32_t first_tile;
    uint32_t num_tiles;
};

// One Stockham pass: radix 2^log_r over 2^log_w columns per workgroup with
// radix-2^log_e register steps.
struct PassCfg {
    int log_r, log_w, log_e;
};

// 753-bit tiles: at most 2^10 elements (99 KiB of LDS, plus up to 56 KiB of
// staged twiddles).  A tile of 2^t elements runs on a workgroup of 2^(t-1)
// threads, t in {7, 8, 9, 10} (tiles under 2^7 use the 64-thread kernel).
constexpr int T753_LOG_TILE = 10;
constexpr int t753_log_tile(const PassCfg &c) { return c.log_r + c.log_w <= 7 ? 7 : c.log_r + c.log_w; }

// Launches one pass.  For the 32-bit field the tile shape must be one of
// the compiled instantiations (pass_supported); 753-bit shapes are runtime.
template <class F>
bool pass_supported(const PassCfg &cfg);
template <class F>
void launch_pass(const PassJob<F> &job, const PassCfg &cfg, hipStream_t stream);
// Launches one pass of many transforms sharing cfg (32-bit) or any shapes
// (753-bit): jobs and tile_job are device arrays; total_tiles blocks.
template <class F>
void launch_pass_multi(const PassJob<F> *d_jobs, const uint32_t *d_tile_job, uint32_t total_tiles, const PassCfg &cfg,
                       hipStream_t stream);

// gsn_debug_clock_ghz's kernel (kernels_mnt6.hip)
void launch_clock_probe(unsigned long long *spans, E753 *sink, unsigned blocks, int iters, hipStream_t stream);
// gsn_debug_mad_rate's kernel (kernels_mnt6.hip); returns the number of
// v_mad_u64_u32 lane operations the launch issues
uint64_t launch_mad_rate_probe(unsigned long long *spans, unsigned long long *sink, unsigned blocks, int iters,
                               hipStream_t stream);

template <class F>
void launch_scale(const typename F::elem_t *in, typename F::elem_t *out, typename F::tw_t s, size_t n,
                  int do_mul, hipStream_t stream);

// Fills the step table of one 753-bit pass (see build_steps753_kernel):
// nrec = Ns (2^log_r - 1) twiddle records and the scale record *scale after
// them, in 14 planes of nrec + 1 quads.  scale0 multiplies the step-0
// records (the first Ns) by *scale.
template <class F>
void launch_build_steps753(uint4 *out, uint32_t nrec, const E753 *hi, const TW753 *lo, uint32_t h, uint32_t log_n,
                           uint32_t log_ns, uint32_t log_r, const TW753 &scale, bool scale0, hipStream_t stream);

}  // namespace gsn
