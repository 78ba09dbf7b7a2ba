// NTT kernels for gfx950: Stockham passes.
//
// A transform of size n = 2^L runs as a sequence of passes.  Pass p has
// radix R = 2^LOG_R and stride Ns = 2^log_ns (the product of the radices of
// the passes before it).  With nR = n / R, for every j in [0, nR):
//
//   x[r]  = in[j + r*nR] * w_n^(r*k*n/(Ns*R)),  k = j mod Ns   (pass twiddle)
//   X     = DFT_R(x)                                         (size-R DFT)
//   out[(j/Ns)*Ns*R + k + r*Ns] = X[r]
//
// which sorts itself: natural-order input gives natural-order output, so the
// bit-reversal permutation of the reference (cuda/fft_kernel.cu:41-45,
// test/fft_host.h:24-29) disappears into the passes' index maps.
//
// A workgroup takes W consecutive j (a tile of R rows by W columns).  The
// size-R DFT of each column is itself a Stockham sequence of radix-E steps:
// each thread holds E elements in registers, does a radix-2 DFT_E on them,
// and exchanges through LDS between steps.  The first step reads global
// memory and the last writes it directly when the output rows are
// W-contiguous (Ns >= W); otherwise (the first pass) the tile is transposed
// through LDS so that the global store is linear.
//
// Two kernel families share that structure:
//   * ntt_pass_kernel<F32, LOG_R, LOG_W, LOG_E>: compile-time tile shape,
//     radix-16 register steps (32-bit field);
//   * ntt753_kernel<F, LOG_TILE>: runtime tile shape (R*W <= 2^LOG_TILE,
//     2^(LOG_TILE-1) threads, radix-2 steps) for the 753-bit fields, whose
//     cost is the multiplier, not the indexing.
// Each has a multi-job form: blockIdx -> (job, tile) through a tile map, so
// one launch can run a pass of many transforms of different sizes.
//
// Twiddles are device-resident tables built once per plan:
//   32-bit field, per pass: a column table (entry x holds w_R^x) and a pass
//     table, either direct (w_{Ns R}^e) or, above 2^12 entries, a two-level
//     split w^e = hi[e >> lo_bits] * lo[e & (2^lo_bits - 1)];
//   753-bit fields, per pass: a step table of the radix-2 step twiddles with
//     the pass twiddles folded in (see tile753_body), n - 1 records over all
//     passes.
#pragma once

#include <hip/hip_runtime.h>

#include "gsn_launch.h"

namespace gsn {

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b mod 8), each with its own L2.  Renumbering blocks so that every XCD
// gets one contiguous range of tiles puts neighbouring tiles -- which share
// cache lines at the edges of their column groups -- under the same L2.
__device__ __forceinline__ uint32_t xcd_tile(uint32_t b, uint32_t grid)
{
    const uint32_t per = grid >> 3, rem = grid & 7u, x = b & 7u, slot = b >> 3;
    return x * per + (x < rem ? x : rem) + slot;
}

__host__ __device__ constexpr int brev_bits(int v, int bits)
{
    int r = 0;
    for (int i = 0; i < bits; ++i)
        r |= ((v >> i) & 1) << (bits - 1 - i);
    return r;
}

// ------------------------------------------------ 32-bit field, compile-time tile shape

// Column-table entries that every lane reads at the same index come through
// the scalar cache: the tables are written before a launch and never during
// it, so they can be read through the constant address space, which costs
// no vector instructions and no VGPRs.
__device__ __forceinline__ F32::tw_t tw32_uniform(const F32::tw_t *tab, uint32_t i)
{
    typedef __attribute__((address_space(4))) const unsigned long long cu64;
    const unsigned long long v = ((cu64 *)tab)[i];
    return F32::tw_t{(uint32_t)v, (uint32_t)(v >> 32)};
}

// Element access with a 32-bit byte offset (n <= 2^30 elements), which maps
// to the scalar-base + 32-bit vector-offset form of the global load/store.
__device__ __forceinline__ uint32_t ld32(const uint32_t *base, uint32_t i)
{
    return *reinterpret_cast<const uint32_t *>(reinterpret_cast<const char *>(base) + (i << 2));
}
__device__ __forceinline__ void st32(uint32_t *base, uint32_t i, uint32_t v)
{
    *reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(base) + (i << 2)) = v;
}
__device__ __forceinline__ F32::tw_t ldtw32(const F32::tw_t *tab, uint32_t i)
{
    const uint2 v = *reinterpret_cast<const uint2 *>(reinterpret_cast<const char *>(tab) + (i << 3));
    return F32::tw_t{v.x, v.y};
}
// the twiddle record of the product of two twiddles (Montgomery forms)
__device__ __forceinline__ F32::tw_t tw32_product(const F32::tw_t &u, const F32::tw_t &v)
{
    const uint32_t w = F32::mul(u.w, v);
    return F32::tw_t{w, w * Fp32Params::PINV};
}

// Natural-order radix-2 DFT of ES = 2^LOG_ES register elements x[OFF..OFF+ES),
// root w_ES = w_R^(R/ES); w_{2m}^jx = w_R^(jx * R / 2m) from the column table.
template <int LOG_R, int LOG_ES, int OFF, int N>
__device__ __forceinline__ void dft32_regs(uint32_t (&x)[N], const F32::tw_t *tw_col)
{
    constexpr int ES = 1 << LOG_ES;
    uint32_t y[ES];
#pragma unroll
    for (int i = 0; i < ES; ++i)
        y[brev_bits(i, LOG_ES)] = x[OFF + i];
#pragma unroll
    for (int s = 1; s <= LOG_ES; ++s) {
        const int m = 1 << (s - 1);
#pragma unroll
        for (int kb = 0; kb < ES; kb += 2 * m) {
#pragma unroll
            for (int jx = 0; jx < m; ++jx) {
                uint32_t t = y[kb + jx + m];
                if (jx)
                    t = F32::mul(t, tw32_uniform(tw_col, (uint32_t)jx << (LOG_R - s)));
                y[kb + jx + m] = F32::sub(y[kb + jx], t);
                y[kb + jx] = F32::add(y[kb + jx], t);
            }
        }
    }
#pragma unroll
    for (int i = 0; i < ES; ++i)
        x[OFF + i] = y[i];
}

template <int LOG_R, int LOG_ES, int N, int G>
__device__ __forceinline__ void dft32_groups(uint32_t (&x)[N], const F32::tw_t *tw_col)
{
    if constexpr (G > 0) {
        dft32_groups<LOG_R, LOG_ES, N, G - 1>(x, tw_col);
        dft32_regs<LOG_R, LOG_ES, (G - 1) << LOG_ES, N>(x, tw_col);
    }
}

// Column steps: step 0 has radix E0 = 2^LOG_E0 (LOG_R mod LOG_E, or LOG_E
// when that is 0) and every later step radix E.  In a step of radix ES each
// thread owns G = E/ES groups; group g is column index jj' = jj + g*(R/E)
// of that step and holds rows jj' + e*(R/ES).  The twiddles between steps
// vary per thread and are read from an LDS copy of the column table.
//
// LDS layout: with W >= 16 columns, row r of the tile sits at words
// r*(W+1) .. r*(W+1)+W-1 (one pad word per row keeps the 16-lane column
// groups of a wave on distinct banks), so every access is a per-thread base
// plus a compile-time offset; single-column tiles use lds_pad.
template <int LOG_R, int LOG_W, int LOG_E>
__device__ __forceinline__ void pass32_body(const PassArgs<F32> &a, uint32_t blk)
{
    using F = F32;
    static_assert(LOG_E <= LOG_R, "register radix larger than the column");
    constexpr int R = 1 << LOG_R, W = 1 << LOG_W, E = 1 << LOG_E;
    constexpr int LOG_E0 = (LOG_R % LOG_E) ? (LOG_R % LOG_E) : LOG_E;
    constexpr int E0 = 1 << LOG_E0, G0 = E / E0;
    constexpr int TILE = R * W, NT = TILE / E;
    constexpr int STEPS = 1 + (LOG_R - LOG_E0) / LOG_E;
    constexpr bool ROWPAD = W >= 16;
    constexpr int LDS_WORDS = ROWPAD ? R * (W + 1) : lds_padded_words(TILE);
    __shared__ uint32_t lds[LDS_WORDS];
    __shared__ F::tw_t twl[STEPS > 1 ? R : 1];
    auto li = [](int row, int col) { return ROWPAD ? row * (W + 1) + col : lds_pad(row * W + col); };

    const int t = threadIdx.x;
    const uint32_t L = a.log_n;
    const uint32_t nR = 1u << (L - LOG_R);
    const uint32_t ns_mask = (1u << a.log_ns) - 1u;
    const uint32_t j0 = blk * W;
    const int c = t & (W - 1);
    const int jj = t >> LOG_W;
    const uint32_t j = j0 + c;
    const uint32_t k = j & ns_mask;

    // column table into LDS; the first barrier of the step loop orders these
    // stores before any read
    if constexpr (STEPS > 1) {
        for (int i = t; i < R; i += NT)
            twl[i] = ldtw32(a.tw_col, (uint32_t)i);
    }

    uint32_t x[E];
    // step 0: group g holds rows jj + g*(R/E) + e*(R/E0), pass twiddle applied
#pragma unroll
    for (int g = 0; g < G0; ++g)
#pragma unroll
        for (int e = 0; e < E0; ++e)
            x[g * E0 + e] = ld32(a.in, j + (uint32_t)(jj + g * (R / E) + e * (R / E0)) * nR);
    if (a.log_ns != 0) {
        // The thread's rows are r = jj + g*(R/E) + e*(R/E0) of one column k,
        // so its twiddles w^(r k) = A B^g C^e with A = w^(jj k),
        // B = w^((R/E) k), C = w^((R/E0) k) (w = w_{Ns R}): three table
        // lookups and E - 1 twiddle products instead of E scattered lookups
        // (each a two-level pair above 2^12 entries), which would keep the
        // texture path busy with one cache line per lane.  An exponent of 0
        // reads the table's exact 1; the table form is uniform.
        const uint32_t sh = a.pass_shift, lb = a.lo_bits, lm = (1u << lb) - 1u;
        auto twx = [&](uint32_t ex) {
            ex <<= sh;
            if (a.tw_pass_lo == nullptr)
                return ldtw32(a.tw_pass, ex);
            return tw32_product(ldtw32(a.tw_pass, ex >> lb), ldtw32(a.tw_pass_lo, ex & lm));
        };
        // B and C depend on the column only: formed once per tile and column
        // (2W lookups instead of two per thread)
        __shared__ F::tw_t colBC[2 * W];
        for (int i = t; i < 2 * W; i += NT) {
            const uint32_t kc = (j0 + (uint32_t)(i & (W - 1))) & ns_mask;
            colBC[i] = twx((i < W ? (uint32_t)(R / E) : (uint32_t)(R / E0)) * kc);
        }
        F::tw_t tg = a.tw_pass_a ? ldtw32(a.tw_pass_a, ((uint32_t)jj << a.log_ns) + k) : twx((uint32_t)jj * k);
        __syncthreads();
        const F::tw_t tB = colBC[c], tC = colBC[W + c];
#pragma unroll
        for (int g = 0; g < G0; ++g) {
            F::tw_t te = tg;
#pragma unroll
            for (int e = 0; e < E0; ++e) {
                x[g * E0 + e] = F::mul(x[g * E0 + e], te);
                if (e + 1 < E0)
                    te = tw32_product(te, tC);
            }
            if (g + 1 < G0)
                tg = tw32_product(tg, tB);
        }
    }
    dft32_groups<LOG_R, LOG_E0, E, G0>(x, a.tw_col);

    if constexpr (STEPS > 1) {
        // step-0 outputs: column jj' lands on rows jj'*E0 + e
#pragma unroll
        for (int g = 0; g < G0; ++g)
#pragma unroll
            for (int e = 0; e < E0; ++e)
                lds[li((jj + g * (R / E)) * E0 + e, c)] = x[g * E0 + e];
#pragma unroll
        for (int s = 1; s < STEPS; ++s) {
            const int log_ns2 = LOG_E0 + (s - 1) * LOG_E;
            const int ns2 = 1 << log_ns2;
            __syncthreads();
            const int kk = jj & (ns2 - 1);
#pragma unroll
            for (int e = 1; e < E; ++e)
                x[e] = F::mul(lds[li(jj + e * (R / E), c)],
                              twl[(e * kk) << (LOG_R - log_ns2 - LOG_E)]);
            x[0] = lds[li(jj, c)];
            dft32_regs<LOG_R, LOG_E, 0, E>(x, a.tw_col);
            if (s < STEPS - 1) {
                __syncthreads();
#pragma unroll
                for (int e = 0; e < E; ++e)
                    lds[li((jj >> log_ns2) * (ns2 * E) + kk + e * ns2, c)] = x[e];
            }
        }
    }
    // x[e] now holds output row jj + e*R/E of column c
    if (a.flags & PASS_SCALE) {
#pragma unroll
        for (int e = 0; e < E; ++e)
            x[e] = F::mul(x[e], a.scale);
    }
    const uint32_t Ns = 1u << a.log_ns;
    if (Ns >= (uint32_t)W) {
        // the W columns share j/Ns: each output row is W contiguous elements
        const uint32_t base = (j0 >> a.log_ns) * Ns * R + k;
#pragma unroll
        for (int e = 0; e < E; ++e)
            st32(a.out, base + (uint32_t)(jj + e * (R / E)) * Ns, x[e]);
    } else {
        // Ns < W: the tile's outputs form one contiguous run of TILE
        // elements; transpose through LDS and store it linearly
        __syncthreads();
#pragma unroll
        for (int e = 0; e < E; ++e)
            lds[li(jj + e * (R / E), c)] = x[e];
        __syncthreads();
        const uint32_t base = j0 * R;  // j0 is a multiple of W, hence of Ns
        const uint32_t log_ns = a.log_ns;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int o = t + e * NT;
            const int c_hi = o >> (log_ns + LOG_R);
            const int rem = o & ((1 << (log_ns + LOG_R)) - 1);
            const int r2 = rem >> log_ns;
            const int c_lo = rem & ((1 << log_ns) - 1);
            const int cc = (c_hi << log_ns) + c_lo;
            st32(a.out, base + o, lds[li(r2, cc)]);
        }
    }
}

template <class F, int LOG_R, int LOG_W, int LOG_E>
__global__ __launch_bounds__(1 << (LOG_R + LOG_W - LOG_E)) void ntt_pass_kernel(PassArgs<F> a)
{
    pass32_body<LOG_R, LOG_W, LOG_E>(a, xcd_tile(blockIdx.x, gridDim.x));
}

template <class F, int LOG_R, int LOG_W, int LOG_E>
__global__ __launch_bounds__(1 << (LOG_R + LOG_W - LOG_E)) void ntt_pass_multi_kernel(
    const PassJob<F> *__restrict__ jobs, const uint32_t *__restrict__ tile_job)
{
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    const PassJob<F> &job = jobs[tile_job[tile]];
    pass32_body<LOG_R, LOG_W, LOG_E>(job.a, tile - job.first_tile);
}

// ------------------------------------------------ 753-bit, runtime tile shape

// One tile of R = 2^log_r rows by W = 2^log_w columns (log_r + log_w <=
// LOG_TILE) with radix-2 steps.  Threads past R*W/2 only join the barriers.
//
// In-place decimation in time over positions i = 0..R-1 of each column:
// thread t = jj W + c starts with positions 2 jj and 2 jj + 1, i.e. input
// rows brev(jj) and brev(jj) + R/2 (brev over log_r - 1 bits), and step s
// pairs positions i and i + 2^s.  Between steps s - 1 and s, position bit s
// (held by thread bit tau = log_w + s - 1) and the slot bit trade places:
// threads t and t ^ 2^tau exchange through LDS, each writing both outputs
// into the slots their next owners read.  For tau < 6 the partner is in the
// same wavefront, whose LDS operations execute in order, so the exchange
// needs no workgroup barrier (only a compiler fence); a partner in another
// wavefront costs a barrier before the writes and one after.  After the
// last step thread t holds positions jj and jj + R/2, which are output rows
// jj and jj + R/2 (the decimation in time reads bit-reversed and writes
// natural order).
//
// The pass twiddle w_{Ns R}^(r k) is not applied to the inputs: folded into
// the steps, step s multiplies x1 by w_{Ns 2^(s+1)}^(k + Ns kk), kk = jj mod
// 2^s (for Ns = 1, k = 0 these are the plain radix-2 twiddles), so a pass of
// radix 2^r costs r products per butterfly column, not r + 1.
//
// Every butterfly is (x0 + t, x0 + 3p - t) with t = x1 * w in [0, 3p); a
// trivial twiddle is the table's exact 1, so t is always a product (only
// step 0 of the first pass, whose x1 is a canonical input below p, skips
// it).  A stage adds at most 3p to the value bound.  The next step's
// twiddle is requested as soon as the current product is done, so its
// latency hides behind the sum, the LDS exchange and the barrier while its
// registers are those the product no longer needs.  (An empty asm statement
// makes the request's index depend on the product's last limb, which keeps
// the compiler from hoisting the load into the product.)  The twiddles come
// from a per-pass step table laid out in the order the lanes read them.
template <class F, int LOG_TILE>
__device__ __forceinline__ void tile753_body(const PassArgs<F> &a, int log_r, int log_w, size_t blk)
{
    using elem = typename F::elem_t;
    using tw = typename F::tw_t;
    constexpr int STRIDE = lds_padded_words(1 << LOG_TILE);  // quads per plane
    __shared__ __align__(16) uint32_t lds[STRIDE * F::LDS_WORDS_PER_ELEM];

    const int t = threadIdx.x;
    const int active = 1 << (log_r + log_w - 1);
    const bool on = t < active;
    const int R = 1 << log_r, W = 1 << log_w, half = R >> 1;
    const uint32_t L = a.log_n, log_ns = a.log_ns;
    const size_t nR = (size_t)1 << (L - log_r);
    const size_t Ns = (size_t)1 << log_ns;
    const size_t j0 = blk * (size_t)W;
    const int c = t & (W - 1);
    const int jj = t >> log_w;  // in [0, R/2) for active threads
    const size_t j = j0 + c;
    const uint32_t k = (uint32_t)(j & (Ns - 1));
    // Every device-memory access goes through global (address space 1)
    // pointers.  Through the generic pointers of a by-value argument the
    // compiler emits flat loads, which also count against the LDS wait
    // counter, so waiting for an LDS exchange would also wait for the next
    // step's twiddle record, the load this loop issues early to hide.
    typedef __attribute__((address_space(1))) const u32x4 gcq4;
    typedef __attribute__((address_space(1))) u32x4 gq4;
    const gcq4 *gin = (const gcq4 *)a.in;  // element i: quads 6 i .. 6 i + 5
    gq4 *gout = (gq4 *)a.out;
    auto gload = [&](size_t i) {
        elem r;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const u32x4 v = gin[6 * i + q];
            r.v[4 * q] = v.x;
            r.v[4 * q + 1] = v.y;
            r.v[4 * q + 2] = v.z;
            r.v[4 * q + 3] = v.w;
        }
        return r;
    };
    auto gstore = [&](size_t i, const elem &x) {
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            u32x4 v;
            v.x = x.v[4 * q];
            v.y = x.v[4 * q + 1];
            v.z = x.v[4 * q + 2];
            v.w = x.v[4 * q + 3];
            gout[6 * i + q] = v;
        }
    };
    // the pass's step table (build_steps753_kernel): step s, column k, kk
    // at (2^s - 1) Ns + k 2^s + kk; step log_r, past the last, reads record
    // 0 (a harmless, branch-free load); the scale record (n^-1) is the last
    // one, read only by the passes that scale.  Reading it from the table
    // keeps 53 uniform digits out of the scalar registers.
    const gcq4 *tab = (const gcq4 *)a.tw_col;
    const uint32_t planes = a.tw_stride;
    // Record i: the wq half (planes 7..13) first, since the product's
    // quotient half needs only it and the w half (planes 0..6) may still be
    // in flight then; the quads that end in padding are loaded as 2 or 3
    // words, so that no load writes a register that nothing reads (the
    // compiler would reuse such a register and wait for the load first).
    typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
    typedef __attribute__((address_space(1))) const u32x2 gcq2;
    typedef __attribute__((address_space(1))) const uint32_t gcq1;
    auto load_record = [&](uint32_t i) {
        tw r;
#pragma unroll
        for (int q = 7; q < 13; ++q) {
            const u32x4 v = tab[(size_t)q * planes + i];
            r.wq[4 * q - 28] = v.x;
            r.wq[4 * q - 27] = v.y;
            r.wq[4 * q - 26] = v.z;
            r.wq[4 * q - 25] = v.w;
        }
        const gcq2 *p13 = (const gcq2 *)(tab + (size_t)13 * planes + i);
        const u32x2 v13 = p13[0];
        r.wq[24] = v13.x;
        r.wq[25] = v13.y;
        r.wq[26] = ((const gcq1 *)p13)[2];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const u32x4 v = tab[(size_t)q * planes + i];
            r.w[4 * q] = v.x;
            r.w[4 * q + 1] = v.y;
            r.w[4 * q + 2] = v.z;
            r.w[4 * q + 3] = v.w;
        }
        const u32x2 v6 = ((const gcq2 *)(tab + (size_t)6 * planes + i))[0];
        r.w[24] = v6.x;
        r.w[25] = v6.y;
        r.pad0[0] = r.pad0[1] = r.pad1 = 0;
        return r;
    };
    auto step_tw = [&](int s, uint32_t after) {
        const uint32_t kk = (uint32_t)jj & ((1u << s) - 1u);
        uint32_t idx = s < log_r ? (((1u << s) - 1u) << log_ns) + (k << s) + kk : 0u;
        asm volatile("" : "+v"(idx) : "v"(after));
        return load_record(idx);
    };

    elem x[2];
    tw w;  // twiddle of the current step
    if (on) {
        // step 0: positions 2 jj and 2 jj + 1 = rows brev(jj) and
        // brev(jj) + R/2 from global memory
        const int jr = log_r > 1 ? (int)(__builtin_bitreverse32((uint32_t)jj) >> (33 - log_r)) : 0;
        x[1] = gload(j + (size_t)(jr + half) * nR);
        x[0] = gload(j + (size_t)jr * nR);
        if (log_ns != 0) {
            x[1] = F::mul(x[1], load_record(k));
            if (a.flags & PASS_SCALE_IN)
                x[0] = F::mul(x[0], load_record(planes - 1u));
        }
        w = step_tw(1, x[1].v[23]);
        F::butterfly(x[0], x[1]);
    }
    for (int s = 1; s < log_r; ++s) {
        // x[e] is position bit s - 1 = e; thread bit tau holds position bit
        // s.  Output e goes to the thread with bit tau = e, into its slot
        // (own bit tau): LDS element (own bit tau) * R W/2 + owner.
        const int tau = log_w + s - 1;
        const bool cross = tau >= 6;  // the partner is in another wavefront
        if (cross)
            __syncthreads();  // every wavefront has read the previous exchange
        if (on) {
            const int m = 1 << tau;
            const int slot = (t & m) ? active : 0;
            F::lds_store(lds, STRIDE, slot + (t & ~m), x[0]);
            F::lds_store(lds, STRIDE, slot + (t | m), x[1]);
        }
        if (cross) {
            __syncthreads();
        } else {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (on) {
            x[1] = F::mul(F::lds_load(lds, STRIDE, active + t), w);
            w = step_tw(s + 1, x[1].v[23]);
            x[0] = F::lds_load(lds, STRIDE, t);
            F::butterfly(x[0], x[1]);
        }
    }
    // x[e] holds output row jj + e*R/2 of column c
    if (on) {
        if (a.flags & PASS_SCALE) {
            const tw sc = load_record(planes - 1u);
            x[0] = F::mul(x[0], sc);
            x[1] = F::mul(x[1], sc);
        }
        if (a.flags & PASS_CANON) {
            x[0] = F::canon(x[0]);
            x[1] = F::canon(x[1]);
        }
    }
    if (a.log_ns == 0) {
        // first pass: the tile's R*W outputs are one contiguous run
        // (out[j*R + r]); stage it in LDS and write it as consecutive
        // 16-byte quads, lane after lane, with non-temporal stores (the next
        // pass reads it from HBM, not from this CU's caches).  The barrier
        // frees LDS: every wavefront has read its last exchange.
        __syncthreads();
        if (on) {
            F::lds_store(lds, STRIDE, jj * W + c, x[0]);
            F::lds_store(lds, STRIDE, (jj + half) * W + c, x[1]);
        }
        __syncthreads();
        const uint4 *q4 = reinterpret_cast<const uint4 *>(lds);
        gq4 *o = (gq4 *)(a.out + j0 * R);
        const int nq = (R * W) * 6;
        for (int m = t; m < nq; m += (int)blockDim.x) {
            const int e = m / 6, q = m - 6 * e;
            const uint4 v = q4[q * STRIDE + lds_pad((e & (R - 1)) * W + (e >> log_r))];
            u32x4 w;
            w.x = v.x;
            w.y = v.y;
            w.z = v.z;
            w.w = v.w;
            __builtin_nontemporal_store(w, o + m);
        }
    } else if (Ns >= (size_t)W) {
        if (on) {
            const size_t base = (j0 / Ns) * Ns * R + k;
            gstore(base + (size_t)jj * Ns, x[0]);
            gstore(base + (size_t)(jj + half) * Ns, x[1]);
        }
    } else {
        // one contiguous run of R*W outputs: transpose through LDS
        __syncthreads();
        if (on) {
            F::lds_store(lds, STRIDE, jj * W + c, x[0]);
            F::lds_store(lds, STRIDE, (jj + half) * W + c, x[1]);
        }
        __syncthreads();
        if (on) {
            const size_t base = j0 * R;
            const uint32_t log_ns = a.log_ns;
            for (int e = 0; e < 2; ++e) {
                const int o = t + e * active;
                const int c_hi = o >> (log_ns + log_r);
                const int rem = o & ((1 << (log_ns + log_r)) - 1);
                const int r2 = rem >> log_ns;
                const int c_lo = rem & ((1 << log_ns) - 1);
                gstore(base + o, F::lds_load(lds, STRIDE, r2 * W + (c_hi << log_ns) + c_lo));
            }
        }
    }
}

// Single job (jobs == nullptr) or many: blockIdx -> tile_job[] -> job.
// LOG_TILE sets the workgroup: 2^(LOG_TILE-1) threads for tiles of up to
// 2^LOG_TILE elements.  A one-wave workgroup (LOG_TILE = 7) never waits on
// another wave at its barriers.  Two waves per SIMD is the register budget
// (256 VGPRs): a single transform of 2^18 has no more threads than that.
template <class F, int LOG_TILE>
__global__ __launch_bounds__(1 << (LOG_TILE - 1)) __attribute__((amdgpu_waves_per_eu(2))) void ntt753_kernel(PassJob<F> single,
                                                                   const PassJob<F> *__restrict__ jobs,
                                                                   const uint32_t *__restrict__ tile_job)
{
    const uint32_t tile = xcd_tile(blockIdx.x, gridDim.x);
    if (jobs == nullptr) {
        tile753_body<F, LOG_TILE>(single.a, (int)single.log_r, (int)single.log_w, tile);
    } else {
        const PassJob<F> &job = jobs[tile_job[tile]];
        tile753_body<F, LOG_TILE>(job.a, (int)job.log_r, (int)job.log_w, tile - job.first_tile);
    }
}

// Step table of one 753-bit pass (log_ns, log_r) of a 2^log_n transform:
// the records tile753_body reads, in the order a wave reads them, stored as
// 14 planes of 16-byte quads (quad q of record i at q * (nrec + 1) + i) so
// that a wave's request for one quad of 64 consecutive records is one
// contiguous kilobyte.  Record nrec is the pass's scale record.  Step s (2^s Ns records from offset (2^s - 1) Ns) holds, at
// k 2^s + kk, w_{Ns 2^(s+1)}^(k + Ns kk) = w_n^e with
// e = (k + Ns kk) << (log_n - log_ns - s - 1) < n/2; each record is
// w_n^e = hi[e >> h] * lo[e & (2^h - 1)] reduced to [0, p).
template <class F>
__global__ __launch_bounds__(256) void build_steps753_kernel(uint4 *out, uint32_t nrec, const E753 *hi, const TW753 *lo,
                                                            uint32_t h, uint32_t log_n, uint32_t log_ns, uint32_t log_r,
                                                            TW753 scale0, int use_scale0)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t stride = nrec + 1u;
    if (i > nrec)
        return;
    if (i == nrec) {
        const uint4 *q = reinterpret_cast<const uint4 *>(&scale0);
#pragma unroll
        for (int p = 0; p < 14; ++p)
            out[(size_t)p * stride + i] = q[p];
        return;
    }
    const uint32_t s = 31u - (uint32_t)__builtin_clz((i >> log_ns) + 1u);
    const uint32_t jx = i - (((1u << s) - 1u) << log_ns);
    const uint32_t k = jx >> s, kk = jx & ((1u << s) - 1u);
    const uint32_t e = (k + (kk << log_ns)) << (log_n - log_ns - s - 1u);
    E753 v = F::canon(F::mul(hi[e >> h], lo[e & ((1u << h) - 1u)]));
    if (use_scale0 && s == 0)
        v = F::canon(F::mul(v, scale0));
    const TW753 r = F::make_tw(v);
    const uint4 *q = reinterpret_cast<const uint4 *>(&r);
#pragma unroll
    for (int p = 0; p < 14; ++p)
        out[(size_t)p * stride + i] = q[p];
    (void)log_r;
}

// Elementwise: out[i] = canon(in[i] * s) (used for n <= 1 and scaling).
template <class F>
__global__ __launch_bounds__(256) void scale_kernel(const typename F::elem_t *in, typename F::elem_t *out,
                                                   typename F::tw_t s, size_t n, int do_mul)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    typename F::elem_t v = F::load(in, i);
    if (do_mul)
        v = F::mul(v, s);
    F::store(out, i, F::canon(v));
}

}  // namespace gsn
