// Domain helpers around the transform: the elementwise steps that sit on
// either side of an NTT in a prover (the reference has none of them; its
// FFT, cuda/fft_kernel.cu:117-147, is the only operation on the path), and
// the field's operators over whole vectors (the reference's Scalar
// operator+ / - / * / ^, cuda/device_field.h:115-158, one element at a time):
//
//   gsn_pointwise_mul    out[i] = a[i] * b[i]            (evaluation-form product)
//   gsn_vec_add          out[i] = a[i] + b[i]
//   gsn_vec_sub          out[i] = a[i] - b[i]
//   gsn_vec_pow          out[i] = in[i]^e                 (e a 64-bit integer)
//   gsn_scale            out[i] = in[i] * s               (e.g. n^-1, a constant)
//   gsn_coset_shift      out[i] = in[i] * g^(+-i)         (coset NTT: shift, transform)
//   gsn_to_montgomery    out[i] = in[i] * R mod p         (plain residues -> API form)
//   gsn_from_montgomery  out[i] = in[i] * R^-1 mod p      (API form -> plain residues)
//
// All work on device buffers of canonical residues (out may equal an input),
// are asynchronous on the given stream and write canonical residues.
// Products of two vectors are general Montgomery products (F753<P>::mont_mul);
// products by a constant are the field's twiddle product (F32::mul,
// F753<P>::mul) by a precomputed record of the constant; a coset's g^i is
// lo[i mod 2^h] * hi[i >> h] from two tables of about sqrt(n) records, cached
// per (field, g, n, direction).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "batch_impl.h"
#include "gpusnarks.h"
#include "gsn_internal.h"
#include "plan_impl.h"

namespace gsn {

// ------------------------------------------------------------ kernels

// out[i] = in[i] * rec
template <class F>
__global__ __launch_bounds__(256) void ew_scale_kernel(typename F::elem_t *out, const typename F::elem_t *in,
                                                      typename F::tw_t rec, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        F::store(out, i, F::canon(F::mul(F::load(in, i), rec)));
}

// out[i] = in[i] * lo[i & (2^h - 1)] * hi[i >> h]
template <class F>
__global__ __launch_bounds__(256) void ew_coset_kernel(typename F::elem_t *out, const typename F::elem_t *in,
                                                      const typename F::tw_t *lo, const typename F::tw_t *hi,
                                                      uint32_t h, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const typename F::elem_t v = F::mul(F::load(in, i), F::load_tw(lo, i & ((1u << h) - 1u)));
    F::store(out, i, F::canon(F::mul(v, F::load_tw(hi, i >> h))));
}

// Elementwise field operations on canonical operands, one adapter per field:
// add/sub/mul give canonical results (the 32-bit product is the Montgomery
// product with the multiplier's companion computed on the fly; the 753-bit
// one is F753<P>::mont_mul).
struct Ops32 {
    using elem = uint32_t;
    __device__ static elem add(elem a, elem b) { return F32::add(a, b); }
    __device__ static elem sub(elem a, elem b) { return F32::sub(a, b); }
    __device__ static elem mul(elem a, elem b) { return F32::mul(a, F32::tw_t{b, b * Fp32Params::PINV}); }
    // product with a constant's precomputed record
    using tw = F32::tw_t;
    __device__ static elem mul_tw(elem a, tw t) { return F32::mul(a, t); }
    __device__ static elem mul_tw2(elem a, tw t, tw u) { return F32::mul(F32::mul(a, t), u); }
    __device__ static tw load_tw(const tw *p, size_t i) { return F32::load_tw(p, i); }
    __device__ static elem one() { return Fp32Params::ONE; }
    __device__ static elem zero() { return 0u; }
    __device__ static elem load(const elem *p, size_t i) { return p[i]; }
    __device__ static void store(elem *p, size_t i, elem v) { p[i] = v; }
};
template <class P>
struct Ops753 {
    using elem = E753;
    using F = F753<P>;
    __device__ static elem add(const elem &a, const elem &b) { return F::add_mod(a, b); }
    __device__ static elem sub(const elem &a, const elem &b) { return F::sub_mod(a, b); }
    __device__ static elem mul(const elem &a, const elem &b) { return F::mont_mul(a, b); }
    // product with a constant's precomputed record (quotient estimate),
    // reduced to [0, p)
    using tw = TW753;
    __device__ static elem mul_tw(const elem &a, const tw &t) { return F::canon(F::mul(a, t)); }
    // by two constants, one reduction: the first product, below 3p, is a
    // valid input of the second
    __device__ static elem mul_tw2(const elem &a, const tw &t, const tw &u) { return F::canon(F::mul(F::mul(a, t), u)); }
    __device__ static tw load_tw(const tw *p, size_t i) { return F::load_tw(p, i); }
    __device__ static elem one()
    {
        elem o;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            o.v[i] = P::ONE[i];
        return o;
    }
    __device__ static elem zero()
    {
        elem o;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            o.v[i] = 0u;
        return o;
    }
    __device__ static elem load(const elem *p, size_t i) { return F::load(p, i); }
    __device__ static void store(elem *p, size_t i, const elem &v) { F::store(p, i, v); }
};

enum { EW_ADD = 0, EW_SUB = 1, EW_MUL = 2 };

// out[i] = a[i] (+ - *) b[i]
template <class O, int OP>
__global__ __launch_bounds__(256) void ew_binop_kernel(typename O::elem *out, const typename O::elem *a,
                                                      const typename O::elem *b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const typename O::elem x = O::load(a, i), y = O::load(b, i);
    O::store(out, i, OP == EW_ADD ? O::add(x, y) : OP == EW_SUB ? O::sub(x, y) : O::mul(x, y));
}

// out[i] = in[i]^e: left-to-right square and multiply over the exponent's
// bits (uniform across lanes), starting from the field's one
template <class O>
__global__ __launch_bounds__(256) void ew_pow_kernel(typename O::elem *out, const typename O::elem *in, size_t n,
                                                    uint64_t e)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const typename O::elem x = O::load(in, i);
    typename O::elem r = O::one();
    for (int bit = 63 - __builtin_clzll(e | 1); bit >= 0; --bit) {
        r = O::mul(r, r);
        if ((e >> bit) & 1u)
            r = O::mul(r, x);
    }
    O::store(out, i, r);
}

// Inner product sum_i a_i b_i (the reference's multiexp over field
// elements, cuda/multi_exp.cu:88-100,104-135 with FieldT = FieldMul =
// Scalar): each block sums a grid-stride slice of products, then a tree in
// LDS; `b` null sums `a` alone (the second level, over the block sums).
constexpr int DOT_THREADS = 256;
template <class O>
__global__ __launch_bounds__(DOT_THREADS) void dot_kernel(typename O::elem *partial, const typename O::elem *a,
                                                         const typename O::elem *b, size_t n)
{
    __shared__ typename O::elem sm[DOT_THREADS];
    const int t = threadIdx.x;
    typename O::elem acc = O::zero();
    for (size_t i = (size_t)blockIdx.x * DOT_THREADS + t; i < n; i += (size_t)gridDim.x * DOT_THREADS)
        acc = O::add(acc, b ? O::mul(O::load(a, i), O::load(b, i)) : O::load(a, i));
    sm[t] = acc;
    __syncthreads();
    for (int s = DOT_THREADS / 2; s > 0; s >>= 1) {
        if (t < s)
            sm[t] = O::add(sm[t], sm[t + s]);
        __syncthreads();
    }
    if (t == 0)
        O::store(partial, blockIdx.x, sm[0]);
}

// Extended radix-2 domain H u g*H, |H| = m (libfqfft's
// extended_radix2_domain construction; the reference has no counterpart).
// Forward, before the two size-m transforms: out[i] = a_i + a_(i+m) and
// out[i+m] = g^i (a_i + G a_(i+m)), G = g^m.
template <class F>
__global__ __launch_bounds__(256) void ext_fwd_kernel(typename F::elem_t *out, const typename F::elem_t *in,
                                                     const typename F::tw_t *lo, const typename F::tw_t *hi,
                                                     uint32_t h, typename F::tw_t G, size_t m)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m)
        return;
    const typename F::elem_t u = F::load(in, i), w = F::load(in, i + m);
    typename F::elem_t t = F::add_mod(u, F::canon(F::mul(w, G)));
    t = F::canon(F::mul(F::canon(F::mul(t, F::load_tw(lo, i & ((1u << h) - 1u)))), F::load_tw(hi, i >> h)));
    F::store(out, i, F::add_mod(u, w));
    F::store(out, i + m, t);
}
// Inverse, after the two size-m inverse transforms (u_i, then g^i v_i):
// c_(i+m) = (u_i - v_i) K with K = 1 / (1 - G), c_i = u_i - c_(i+m).
template <class F>
__global__ __launch_bounds__(256) void ext_inv_kernel(typename F::elem_t *a, const typename F::tw_t *lo,
                                                     const typename F::tw_t *hi, uint32_t h, typename F::tw_t K,
                                                     size_t m)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m)
        return;
    const typename F::elem_t u = F::load(a, i);
    typename F::elem_t v = F::load(a, i + m);
    v = F::canon(F::mul(F::canon(F::mul(v, F::load_tw(lo, i & ((1u << h) - 1u)))), F::load_tw(hi, i >> h)));
    const typename F::elem_t c_hi = F::canon(F::mul(F::sub_mod(u, v), K));
    F::store(a, i, F::sub_mod(u, c_hi));
    F::store(a, i + m, c_hi);
}

// Mixed-radix transforms, n = n1 * n2 with n1 odd (3, 5, 7 factors of p - 1)
// and n2 = 2^a, by the four-step factorisation (j = n2 j1 + j2, k = k1 + n1 k2):
//   Y[k1][j2] = w^(j2 k1) sum_j1 x[n2 j1 + j2] w^(n2 j1 k1)   (mixed_pass_kernel)
//   Z[k1][k2] = sum_j2 Y[k1][j2] w^(n1 j2 k2)                  (radix-2 plans on rows)
//   X[k1 + n1 k2] = Z[k1][k2]                                   (mixed_transpose_kernel)
// The odd DFT of every column runs as radix-3/5/7 Stockham passes, one
// kernel per pass over all columns (thread = column j2 x group j, adjacent
// threads on adjacent columns, so every access is coalesced): a pass of
// radix q over stride L twiddles input r of group j by w_{Lq}^(r (j mod L))
// and writes output s to position (j / L) L q + (j mod L) + s L.  Position
// i of column j2 lives at i n2 + j2 in every buffer.  The last pass also
// multiplies by w^(j2 k1) and so leaves Y row-major by k1.  Every product
// is by a constant, so its tables hold the transform's twiddle records
// (precomputed quotient for the 753-bit fields, reduced to [0, p) after each
// product): T[e] = (w^n2)^e for e < n1, and w^e = lo[e mod 2^h] hi[e >> h].
//
// Each q-point DFT folds its inputs in pairs: with u = w_q, a_m = x_m +
// x_{q-m} and b_m = x_m - x_{q-m} (m = 1..(q-1)/2),
//   X_0 = x_0 + sum a_m,  X_k, X_{q-k} = x_0 + U_k +- V_k,
//   U_k = sum_m a_m C[m k mod q],  V_k = sum_m b_m S[m k mod q],
// C[e] = (u^e + u^-e)/2, S[e] = (u^e - u^-e)/2 (the table CS, per radix):
// (q-1)^2/2 products instead of (q-1)^2.
struct MixRadices {
    uint32_t count;
    uint32_t q[24];
};
// CS layout: radix 3 at 0, 5 at 6, 7 at 16; C[e] at off + e, S[e] at off + q + e
__host__ __device__ constexpr uint32_t mix_cs_offset(uint32_t q) { return q == 3 ? 0u : (q == 5 ? 6u : 16u); }
constexpr uint32_t MIX_CS_ELEMS = 30;

template <class O, int Q, bool LAST>
__device__ __forceinline__ void mixed_pass_body(typename O::elem *dst, const typename O::elem *src, uint32_t n1,
                                                uint32_t log_n2, uint32_t L, const typename O::tw *T,
                                                const typename O::tw *CS, const typename O::tw *lo,
                                                const typename O::tw *hi, uint32_t h)
{
    using E = typename O::elem;
    constexpr int H = (Q - 1) / 2;
    const uint32_t nq = n1 / Q, span = n1 / (L * Q);
    const size_t tid = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (tid >= ((size_t)nq << log_n2))
        return;
    const size_t j2 = tid & (((size_t)1 << log_n2) - 1);
    const uint32_t j = (uint32_t)(tid >> log_n2);
    const uint32_t k = j % L, jb = j / L;
    const typename O::tw *C = CS + mix_cs_offset(Q), *S = C + Q;
    E x[Q];
#pragma unroll
    for (int r = 0; r < Q; ++r)
        x[r] = O::load(src, ((size_t)(j + r * nq) << log_n2) + j2);
    if (k != 0)  // r k span < n1: w_{Lq}^(r k) = T[r k span]
#pragma unroll
        for (int r = 1; r < Q; ++r)
            x[r] = O::mul_tw(x[r], O::load_tw(T, (size_t)r * k * span));
    const uint32_t pos0 = jb * L * Q + k;
    auto put = [&](uint32_t s, const E &v) {
        const uint32_t pos = pos0 + s * L;
        if (LAST) {
            const uint64_t e = (uint64_t)j2 * pos;
            O::store(dst, ((size_t)pos << log_n2) + j2,
                     O::mul_tw2(v, O::load_tw(lo, (size_t)(e & ((1ull << h) - 1))), O::load_tw(hi, (size_t)(e >> h))));
        } else {
            O::store(dst, ((size_t)pos << log_n2) + j2, v);
        }
    };
    E s0 = x[0];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        const E am = O::add(x[m], x[Q - m]);
        x[Q - m] = O::sub(x[m], x[Q - m]);  // b_m
        x[m] = am;
        s0 = O::add(s0, am);
    }
    put(0, s0);
#pragma unroll
    for (int kk = 1; kk <= H; ++kk) {
        E U = O::mul_tw(x[1], O::load_tw(C, (uint32_t)kk % Q));
        E V = O::mul_tw(x[Q - 1], O::load_tw(S, (uint32_t)kk % Q));
#pragma unroll
        for (int m = 2; m <= H; ++m) {
            U = O::add(U, O::mul_tw(x[m], O::load_tw(C, (uint32_t)(m * kk) % Q)));
            V = O::add(V, O::mul_tw(x[Q - m], O::load_tw(S, (uint32_t)(m * kk) % Q)));
        }
        const E base = O::add(x[0], U);
        put((uint32_t)kk, O::add(base, V));
        put((uint32_t)(Q - kk), O::sub(base, V));
    }
}

template <class O, int Q, bool LAST>
__global__ __launch_bounds__(256) void mixed_pass_kernel(typename O::elem *dst, const typename O::elem *src,
                                                         uint32_t n1, uint32_t log_n2, uint32_t L,
                                                         const typename O::tw *T, const typename O::tw *CS,
                                                         const typename O::tw *lo, const typename O::tw *hi,
                                                         uint32_t h)
{
    mixed_pass_body<O, Q, LAST>(dst, src, n1, log_n2, L, T, CS, lo, hi, h);
}
// radix 3 holds three elements: held to 256 registers it runs two
// wavefronts per SIMD (radix 5 and 7 need more registers than that)
template <class O, bool LAST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void mixed_pass3_kernel(
    typename O::elem *dst, const typename O::elem *src, uint32_t n1, uint32_t log_n2, uint32_t L,
    const typename O::tw *T, const typename O::tw *CS, const typename O::tw *lo, const typename O::tw *hi, uint32_t h)
{
    mixed_pass_body<O, 3, LAST>(dst, src, n1, log_n2, L, T, CS, lo, hi, h);
}

template <class O, int Q>
void launch_mixed_pass(bool last, typename O::elem *dst, const typename O::elem *src, uint32_t n1, uint32_t log_n2,
                       uint32_t L, const typename O::tw *T, const typename O::tw *CS, const typename O::tw *lo,
                       const typename O::tw *hi, uint32_t h, hipStream_t st)
{
    const size_t threads = ((size_t)(n1 / Q)) << log_n2;
    const dim3 grid((unsigned)((threads + 255) / 256)), block(256);
    if constexpr (Q == 3) {
        if (last)
            hipLaunchKernelGGL((mixed_pass3_kernel<O, true>), grid, block, 0, st, dst, src, n1, log_n2, L, T, CS, lo,
                               hi, h);
        else
            hipLaunchKernelGGL((mixed_pass3_kernel<O, false>), grid, block, 0, st, dst, src, n1, log_n2, L, T, CS,
                               lo, hi, h);
    } else {
        if (last)
            hipLaunchKernelGGL((mixed_pass_kernel<O, Q, true>), grid, block, 0, st, dst, src, n1, log_n2, L, T, CS,
                               lo, hi, h);
        else
            hipLaunchKernelGGL((mixed_pass_kernel<O, Q, false>), grid, block, 0, st, dst, src, n1, log_n2, L, T, CS,
                               lo, hi, h);
    }
}

// out[k1 + n1 k2] = ws[k1 n2 + k2]
template <class O>
__global__ __launch_bounds__(256) void mixed_transpose_kernel(typename O::elem *out, const typename O::elem *ws,
                                                             uint32_t n1, uint32_t log_n2, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n)
        return;
    const size_t k1 = i >> log_n2, k2 = i & (((size_t)1 << log_n2) - 1);
    O::store(out, k1 + (size_t)n1 * k2, O::load(ws, i));
}

// ------------------------------------------------------------ host side

namespace {

unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

template <int FID>
typename FieldInfo<FID>::Dev::tw_t record(const typename FieldInfo<FID>::Host &v)
{
    return PlanImpl<FID>::make_tw(v);
}

// value R mod p and R^-1 mod p as field elements
template <int FID>
typename FieldInfo<FID>::Host value_r()
{
    using H = typename FieldInfo<FID>::Host;
    return H::from_canonical(H::one().im_rep);  // one() holds R mod p
}

struct CosetTables {
    DeviceBuffer lo, hi;
    uint32_t h = 0;
};

std::mutex g_coset_mutex;
std::map<std::tuple<int, unsigned, int, std::vector<uint32_t>>, std::unique_ptr<CosetTables>> g_coset;

// g^(+-i) = lo[i & (2^h - 1)] * hi[i >> h] for i < 2^log_n
template <int FID>
int coset_tables(CosetTables **out, const typename FieldInfo<FID>::Host &g, unsigned log_n, int inverse)
{
    using H = typename FieldInfo<FID>::Host;
    using tw = typename FieldInfo<FID>::Dev::tw_t;
    std::vector<uint32_t> key(g.im_rep, g.im_rep + H::LIMBS);
    std::lock_guard<std::mutex> lk(g_coset_mutex);
    auto &slot = g_coset[std::make_tuple(FID, log_n, inverse, key)];
    if (!slot) {
        std::unique_ptr<CosetTables> t(new CosetTables());
        const H base = inverse ? g.mul_inv() : g;
        t->h = (log_n + 1) / 2;
        const size_t nlo = (size_t)1 << t->h, nhi = std::max<size_t>(1, ((size_t)1 << log_n) >> t->h);
        std::vector<tw> lo(nlo), hi(nhi);
        H acc = H::one();
        for (size_t x = 0; x < nlo; ++x) {
            lo[x] = record<FID>(acc);
            acc *= base;
        }
        const H step = acc;
        acc = H::one();
        for (size_t y = 0; y < nhi; ++y) {
            hi[y] = record<FID>(acc);
            acc *= step;
        }
        GSN_TRY(t->lo.alloc(nlo * sizeof(tw)));
        GSN_TRY(t->hi.alloc(nhi * sizeof(tw)));
        GSN_TRY(hip_status(hipMemcpy(t->lo.ptr, lo.data(), nlo * sizeof(tw), hipMemcpyHostToDevice), "hipMemcpy"));
        GSN_TRY(hip_status(hipMemcpy(t->hi.ptr, hi.data(), nhi * sizeof(tw), hipMemcpyHostToDevice), "hipMemcpy"));
        slot = std::move(t);
    }
    *out = slot.get();
    return GSN_OK;
}

template <int FID>
int scale_t(void *out, const void *in, size_t n, const typename FieldInfo<FID>::Host &s, hipStream_t st)
{
    using F = typename FieldInfo<FID>::Dev;
    hipLaunchKernelGGL(ew_scale_kernel<F>, dim3(blocks_for(n)), dim3(256), 0, st, (typename F::elem_t *)out,
                       (const typename F::elem_t *)in, record<FID>(s), n);
    return hip_status(hipGetLastError(), "scale launch");
}

template <int FID>
int coset_t(void *out, const void *in, size_t n, const void *g_limbs, int inverse, hipStream_t st)
{
    using H = typename FieldInfo<FID>::Host;
    using F = typename FieldInfo<FID>::Dev;
    unsigned log_n = 0;
    GSN_TRY(log2_exact(n, &log_n));
    const H g = H::from_raw((const uint32_t *)g_limbs);
    if (fields::detail::geq<H::LIMBS>(g.im_rep, H::mod()) || g == H::zero())
        return set_error(GSN_ERR_INVALID_ARGUMENT, "the coset generator must be a nonzero canonical element");
    CosetTables *t = nullptr;
    GSN_TRY(coset_tables<FID>(&t, g, log_n, inverse));
    hipLaunchKernelGGL(ew_coset_kernel<F>, dim3(blocks_for(n)), dim3(256), 0, st, (typename F::elem_t *)out,
                       (const typename F::elem_t *)in, (const typename F::tw_t *)t->lo.ptr,
                       (const typename F::tw_t *)t->hi.ptr, t->h, n);
    return hip_status(hipGetLastError(), "coset launch");
}

int check_args(int field, const void *out, const void *in, size_t n)
{
    if (!element_bytes(field))
        return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id");
    if (!out || !in)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL buffer");
    if (n == 0)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "n is 0");
    if (n > (size_t)0xFFFFFFFFu * 256u)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "n exceeds one launch's grid");
    return GSN_OK;
}

template <int OP>
int binop(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream)
{
    GSN_TRY(check_args(field, d_out, d_a, n));
    if (!d_b)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL buffer");
    GSN_TRY(take_injected_fault());
    GSN_TRY(ensure_device());
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid(blocks_for(n)), block(256);
    switch (field) {
    case GSN_FIELD_FP32:
        hipLaunchKernelGGL((ew_binop_kernel<Ops32, OP>), grid, block, 0, st, (uint32_t *)d_out, (const uint32_t *)d_a,
                           (const uint32_t *)d_b, n);
        break;
    case GSN_FIELD_MNT4753_FR:
        hipLaunchKernelGGL((ew_binop_kernel<Ops753<Mnt4753FrParams>, OP>), grid, block, 0, st, (E753 *)d_out,
                           (const E753 *)d_a, (const E753 *)d_b, n);
        break;
    default:
        hipLaunchKernelGGL((ew_binop_kernel<Ops753<Mnt6753FrParams>, OP>), grid, block, 0, st, (E753 *)d_out,
                           (const E753 *)d_a, (const E753 *)d_b, n);
        break;
    }
    return hip_status(hipGetLastError(), "elementwise launch");
}

// the field's smallest quadratic non-residue, the default shift of an
// extended domain (g^m != 1 for every m dividing the 2-power part of p - 1)
template <int FID>
typename FieldInfo<FID>::Host default_shift();
template <>
fields::Fp32 default_shift<GSN_FIELD_FP32>() { return fields::Fp32(Fp32Params::GENERATOR); }
template <>
fields::Scalar default_shift<GSN_FIELD_MNT4753_FR>() { return fields::Scalar(Mnt4753FrParams::QNR); }
template <>
fields::ScalarMNT6 default_shift<GSN_FIELD_MNT6753_FR>() { return fields::ScalarMNT6(Mnt6753FrParams::QNR); }

// A transform over the extended domain H u g*H of n = 2m points: one
// elementwise kernel and two size-m transforms (forward: combine, then
// transform both halves; inverse: transform both halves, then solve).  The
// two halves run as one batch (each pass level of both in one fused
// launch), built for the buffers of the first call and rebuilt when they
// change.
template <int FID>
struct ExtPlan : PlanBase {
    using H = typename FieldInfo<FID>::Host;
    using F = typename FieldInfo<FID>::Dev;
    PlanBase *half = nullptr;
    CosetTables *tab = nullptr;  // g^i (forward) or g^-i (inverse), i < m
    typename F::tw_t c;          // G (forward) or 1 / (1 - G) (inverse)
    std::unique_ptr<BatchImpl> halves;  // both half transforms, fused
    const void *halves_in = nullptr;
    void *halves_out = nullptr;
    ~ExtPlan() override { delete half; }

    // both size-m transforms of in[0, m), in[m, 2m) into out[0, m), out[m, 2m)
    int run_halves(const void *in, void *out, hipStream_t st)
    {
        if (!halves || halves_in != in || halves_out != out) {
            const size_t hb = ((size_t)1 << (log_n - 1)) * sizeof(typename F::elem_t);
            const gsn_batch_item items[2] = {
                {FID, log_n - 1, inverse, nullptr, in, out},
                {FID, log_n - 1, inverse, nullptr, (const char *)in + hb, (char *)out + hb}};
            std::unique_ptr<BatchImpl> b(new BatchImpl());
            GSN_TRY(build_batch(*b, items, 2, 0));
            halves = std::move(b);
            halves_in = in;
            halves_out = out;
        }
        return halves->enqueue(st);
    }

    int init(unsigned L, const H &g, int inv)
    {
        field = FID;
        log_n = L;
        inverse = inv;
        const unsigned lm = L - 1;
        const H G = g.pow((uint64_t)1 << lm);
        if (G == H::one())
            return set_error(GSN_ERR_INVALID_ARGUMENT, "shift^(n/2) = 1: g*H would meet H");
        GSN_TRY(create_plan(&half, FID, lm, nullptr, inv));
        GSN_TRY(coset_tables<FID>(&tab, g, lm, inv));
        c = inv ? record<FID>((H::one() - G).mul_inv()) : record<FID>(G);
        num_passes = half->num_passes + 1;  // the halves' passes are fused
        device_bytes = half->device_bytes;
        return GSN_OK;
    }

    int execute(const void *in, void *out, hipStream_t st) override
    {
        using E = typename F::elem_t;
        const size_t m = (size_t)1 << (log_n - 1);
        const dim3 grid(blocks_for(m)), block(256);
        const auto *lo = (const typename F::tw_t *)tab->lo.ptr, *hi = (const typename F::tw_t *)tab->hi.ptr;
        if (!inverse) {
            GSN_TRY(take_injected_fault());
            hipLaunchKernelGGL(ext_fwd_kernel<F>, grid, block, 0, st, (E *)out, (const E *)in, lo, hi, tab->h, c, m);
            GSN_TRY(hip_status(hipGetLastError(), "extended-domain launch"));
            return run_halves(out, out, st);
        }
        GSN_TRY(run_halves(in, out, st));
        GSN_TRY(take_injected_fault());
        hipLaunchKernelGGL(ext_inv_kernel<F>, grid, block, 0, st, (E *)out, lo, hi, tab->h, c, m);
        return hip_status(hipGetLastError(), "extended-domain launch");
    }
};

// ---- mixed radix

template <int FID>
struct MixInfo;
template <>
struct MixInfo<GSN_FIELD_FP32> {
    using O = Ops32;
    static constexpr int MAXN1 = 3;
    static constexpr uint32_t ODD_MAX = Fp32Params::ODD_MAX;
    static fields::Fp32 root_odd() { return fields::Fp32::from_raw(&Fp32Params::ROOT_ODD); }
};
template <>
struct MixInfo<GSN_FIELD_MNT4753_FR> {
    using O = Ops753<Mnt4753FrParams>;
    static constexpr int MAXN1 = 525;
    static constexpr uint32_t ODD_MAX = Mnt4753FrParams::ODD_MAX;
    static fields::Scalar root_odd() { return fields::Scalar::from_raw(Mnt4753FrParams::ROOT_ODD); }
};
template <>
struct MixInfo<GSN_FIELD_MNT6753_FR> {
    using O = Ops753<Mnt6753FrParams>;
    static constexpr int MAXN1 = 735;  // 3 5 7^2: the largest odd part the tests cover
    static constexpr uint32_t ODD_MAX = Mnt6753FrParams::ODD_MAX;
    static fields::ScalarMNT6 root_odd() { return fields::ScalarMNT6::from_raw(Mnt6753FrParams::ROOT_ODD); }
};

// A transform of n = n1 2^a points (n1 > 1 odd, dividing the 3/5/7-smooth
// part of p - 1) over the primitive n-th root w = w_{2^a} w_{n1} (the
// standard 2^a-th root times ROOT_ODD^(ODD_MAX / n1)): the odd-DFT kernel
// into the plan's workspace, n1 radix-2 transforms of its rows as one
// batch, and the transpose into the output.
template <int FID>
struct MixedPlan : PlanBase {
    using H = typename FieldInfo<FID>::Host;
    using MI = MixInfo<FID>;
    using O = typename MI::O;
    using E = typename O::elem;
    using TW = typename O::tw;
    uint32_t n1 = 0, a = 0, h = 0;
    MixRadices rad{};
    DeviceBuffer T, lo, hi, ws, ws2, CS;  // T, lo, hi, CS: records of constants
    std::unique_ptr<BatchImpl> rows;

    static void put(std::vector<TW> &v, size_t i, const H &x) { v[i] = record<FID>(x); }

    int init(uint32_t n1_, uint32_t a_, int inv)
    {
        field = FID;
        inverse = inv;
        n1 = n1_;
        a = a_;
        const size_t n2 = (size_t)1 << a, n = (size_t)n1 * n2;
        log_n = 0;  // not a power of two; the size is n1 2^a
        uint32_t m = n1;
        for (uint32_t l : {7u, 5u, 3u})
            for (; m % l == 0; m /= l)
                rad.q[rad.count++] = l;
        const H w2 = root_of_unity(FieldInfo<FID>::root_max(), FieldInfo<FID>::TWO_ADICITY, a);
        const H w = w2 * MI::root_odd().pow(MI::ODD_MAX / n1);
        const H base = inv ? w.mul_inv() : w;
        // T[e] = (base^n2)^e, e < n1
        std::vector<TW> t(n1);
        const H w1 = base.pow(n2);
        H acc = H::one();
        for (uint32_t e = 0; e < n1; ++e, acc = acc * w1)
            put(t, e, acc);
        // symmetric DFT constants per radix q: u = w1^(n1/q), C[e] = (u^e +
        // u^-e)/2, S[e] = (u^e - u^-e)/2
        std::vector<TW> cs(MIX_CS_ELEMS);
        const H half = H(2u).mul_inv();
        for (uint32_t q : {3u, 5u, 7u}) {
            if (n1 % q)
                continue;
            const H u = w1.pow(n1 / q), ui = u.mul_inv();
            H up = H::one(), un = H::one();
            for (uint32_t e = 0; e < q; ++e, up = up * u, un = un * ui) {
                put(cs, mix_cs_offset(q) + e, (up + un) * half);
                put(cs, mix_cs_offset(q) + q + e, (up - un) * half);
            }
        }
        // base^e = lo[e mod 2^h] hi[e >> h] (hi scaled by n1^-1 for the inverse)
        while (((size_t)1 << (2 * h)) < n)
            ++h;
        const size_t nlo = (size_t)1 << h, nhi = (n + nlo - 1) >> h;
        std::vector<TW> vlo(nlo), vhi(nhi);
        acc = H::one();
        for (size_t e = 0; e < nlo; ++e, acc = acc * base)
            put(vlo, e, acc);
        H hacc = inv ? H(n1).mul_inv() : H::one();
        for (size_t y = 0; y < nhi; ++y, hacc = hacc * acc)
            put(vhi, y, hacc);
        GSN_TRY(T.alloc(n1 * sizeof(TW)));
        GSN_TRY(lo.alloc(nlo * sizeof(TW)));
        GSN_TRY(hi.alloc(nhi * sizeof(TW)));
        GSN_TRY(ws.alloc(n * sizeof(E)));
        if (rad.count > 1)
            GSN_TRY(ws2.alloc(n * sizeof(E)));
        GSN_TRY(CS.alloc(MIX_CS_ELEMS * sizeof(TW)));
        GSN_TRY(hip_status(hipMemcpy(CS.ptr, cs.data(), MIX_CS_ELEMS * sizeof(TW), hipMemcpyHostToDevice),
                           "hipMemcpy"));
        GSN_TRY(hip_status(hipMemcpy(T.ptr, t.data(), n1 * sizeof(TW), hipMemcpyHostToDevice), "hipMemcpy"));
        GSN_TRY(hip_status(hipMemcpy(lo.ptr, vlo.data(), nlo * sizeof(TW), hipMemcpyHostToDevice), "hipMemcpy"));
        GSN_TRY(hip_status(hipMemcpy(hi.ptr, vhi.data(), nhi * sizeof(TW), hipMemcpyHostToDevice), "hipMemcpy"));
        num_passes = 2;
        device_bytes = (n1 + nlo + nhi + MIX_CS_ELEMS) * sizeof(TW) + (rad.count > 1 ? 2 : 1) * n * sizeof(E);
        if (a > 0) {
            // the rows: radix-2 transforms over w^n1 (a primitive 2^a-th root)
            const H wr = w.pow(n1);
            std::vector<gsn_batch_item> items(n1);
            for (uint32_t k1 = 0; k1 < n1; ++k1) {
                E *row = (E *)ws.ptr + (size_t)k1 * n2;
                items[k1] = {FID, a, inv, wr.im_rep, row, row};
            }
            rows.reset(new BatchImpl());
            GSN_TRY(build_batch(*rows, items.data(), items.size(), 0));
            num_passes += (int)rows->groups.size() + (int)rows->copies.size();
        }
        return GSN_OK;
    }

    int execute(const void *in, void *out, hipStream_t st) override
    {
        GSN_TRY(take_injected_fault());
        const size_t n2 = (size_t)1 << a, n = (size_t)n1 * n2;
        // passes alternate between the two workspaces so the last lands in
        // ws (the caller's buffers are only read, then written by the
        // transpose)
        const E *src = (const E *)in;
        uint32_t L = 1;
        for (uint32_t pi = 0; pi < rad.count; ++pi) {
            const bool last = pi + 1 == rad.count;
            E *dst = (E *)(((rad.count - 1 - pi) & 1) ? ws2.ptr : ws.ptr);
            const uint32_t q = rad.q[pi];
            if (q == 7)
                launch_mixed_pass<O, 7>(last, dst, src, n1, a, L, (const TW *)T.ptr, (const TW *)CS.ptr,
                                        (const TW *)lo.ptr, (const TW *)hi.ptr, h, st);
            else if (q == 5)
                launch_mixed_pass<O, 5>(last, dst, src, n1, a, L, (const TW *)T.ptr, (const TW *)CS.ptr,
                                        (const TW *)lo.ptr, (const TW *)hi.ptr, h, st);
            else
                launch_mixed_pass<O, 3>(last, dst, src, n1, a, L, (const TW *)T.ptr, (const TW *)CS.ptr,
                                        (const TW *)lo.ptr, (const TW *)hi.ptr, h, st);
            src = dst;
            L *= q;
        }
        GSN_TRY(hip_status(hipGetLastError(), "mixed-radix launch"));
        if (rows)
            GSN_TRY(rows->enqueue(st));
        hipLaunchKernelGGL(mixed_transpose_kernel<O>, dim3(blocks_for(n)), dim3(256), 0, st, (E *)out,
                           (const E *)ws.ptr, n1, a, n);
        return hip_status(hipGetLastError(), "mixed-radix transpose launch");
    }
};

template <int FID>
int create_mixed_plan(PlanBase **out, size_t n, int inverse)
{
    using MI = MixInfo<FID>;
    uint32_t a = 0;
    size_t m = n;
    while (m && (m & 1) == 0) {
        m >>= 1;
        ++a;
    }
    if (n == 0 || m == 1 || m > (size_t)MI::MAXN1 || MI::ODD_MAX % m != 0 || a > FieldInfo<FID>::TWO_ADICITY)
        return set_error(GSN_ERR_INVALID_ARGUMENT,
                         "mixed-radix size " + std::to_string(n) + " must be n1 2^a with 1 < n1 <= " +
                             std::to_string(MI::MAXN1) + " dividing " + std::to_string(MI::ODD_MAX) +
                             " and a <= the 2-adicity");
    std::unique_ptr<MixedPlan<FID>> p(new MixedPlan<FID>());
    GSN_TRY(p->init((uint32_t)m, a, inverse));
    *out = p.release();
    return GSN_OK;
}

template <int FID>
int create_ext_plan(PlanBase **out, unsigned log_n, const void *shift, int inverse)
{
    using H = typename FieldInfo<FID>::Host;
    if (log_n < 1 || log_n - 1 > FieldInfo<FID>::TWO_ADICITY)
        return set_error(GSN_ERR_SIZE_EXCEEDS_TWO_ADICITY, "extended domain of 2^" + std::to_string(log_n) +
                                                       " points needs 1 <= log_n <= 2-adicity + 1");
    H g = default_shift<FID>();
    if (shift) {
        g = H::from_raw((const uint32_t *)shift);
        if (fields::detail::geq<H::LIMBS>(g.im_rep, H::mod()) || g == H::zero())
            return set_error(GSN_ERR_INVALID_ARGUMENT, "the shift must be a nonzero canonical element");
    }
    std::unique_ptr<ExtPlan<FID>> p(new ExtPlan<FID>());
    GSN_TRY(p->init(log_n, g, inverse));
    *out = p.release();
    return GSN_OK;
}

template <class O>
int dot_t(void *d_out, const void *d_a, const void *d_b, size_t n, hipStream_t st)
{
    using E = typename O::elem;
    const unsigned blocks = (unsigned)std::min<size_t>(1024, (n + DOT_THREADS - 1) / DOT_THREADS);
    // block sums in stream-ordered scratch: concurrent calls on different
    // streams get their own
    void *partial = nullptr;
    GSN_TRY(hip_status(hipMallocAsync(&partial, blocks * sizeof(E), st), "hipMallocAsync"));
    hipLaunchKernelGGL(dot_kernel<O>, dim3(blocks), dim3(DOT_THREADS), 0, st, (E *)partial, (const E *)d_a,
                       (const E *)d_b, n);
    hipLaunchKernelGGL(dot_kernel<O>, dim3(1), dim3(DOT_THREADS), 0, st, (E *)d_out, (const E *)partial,
                       (const E *)nullptr, (size_t)blocks);
    const int rc = hip_status(hipGetLastError(), "inner product launch");
    const int rf = hip_status(hipFreeAsync(partial, st), "hipFreeAsync");
    return rc ? rc : rf;
}

}  // namespace

}  // namespace gsn

using namespace gsn;

extern "C" {

int gsn_pointwise_mul(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream)
{
    return binop<EW_MUL>(field, d_out, d_a, d_b, n, stream);
}

int gsn_vec_add(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream)
{
    return binop<EW_ADD>(field, d_out, d_a, d_b, n, stream);
}

int gsn_vec_sub(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream)
{
    return binop<EW_SUB>(field, d_out, d_a, d_b, n, stream);
}

int gsn_vec_pow(int field, void *d_out, const void *d_in, size_t n, uint64_t exponent, void *stream)
{
    GSN_TRY(check_args(field, d_out, d_in, n));
    GSN_TRY(take_injected_fault());
    GSN_TRY(ensure_device());
    hipStream_t st = (hipStream_t)stream;
    const dim3 grid(blocks_for(n)), block(256);
    switch (field) {
    case GSN_FIELD_FP32:
        hipLaunchKernelGGL(ew_pow_kernel<Ops32>, grid, block, 0, st, (uint32_t *)d_out, (const uint32_t *)d_in, n,
                           exponent);
        break;
    case GSN_FIELD_MNT4753_FR:
        hipLaunchKernelGGL(ew_pow_kernel<Ops753<Mnt4753FrParams>>, grid, block, 0, st, (E753 *)d_out,
                           (const E753 *)d_in, n, exponent);
        break;
    default:
        hipLaunchKernelGGL(ew_pow_kernel<Ops753<Mnt6753FrParams>>, grid, block, 0, st, (E753 *)d_out,
                           (const E753 *)d_in, n, exponent);
        break;
    }
    return hip_status(hipGetLastError(), "pow launch");
}

int gsn_inner_product(int field, void *d_out, const void *d_a, const void *d_b, size_t n, void *stream)
{
    GSN_TRY(check_args(field, d_out, d_a, n));
    if (!d_b)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL buffer");
    GSN_TRY(take_injected_fault());
    GSN_TRY(ensure_device());
    hipStream_t st = (hipStream_t)stream;
    switch (field) {
    case GSN_FIELD_FP32: return dot_t<Ops32>(d_out, d_a, d_b, n, st);
    case GSN_FIELD_MNT4753_FR: return dot_t<Ops753<Mnt4753FrParams>>(d_out, d_a, d_b, n, st);
    default: return dot_t<Ops753<Mnt6753FrParams>>(d_out, d_a, d_b, n, st);
    }
}

int gsn_plan_create_mixed(gsn_plan **plan, int field, size_t n, int inverse)
{
    if (!plan)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "plan is NULL");
    *plan = nullptr;
    GSN_TRY(ensure_device());
    PlanBase *impl = nullptr;
    switch (field) {
    case GSN_FIELD_FP32: GSN_TRY(create_mixed_plan<GSN_FIELD_FP32>(&impl, n, inverse)); break;
    case GSN_FIELD_MNT4753_FR: GSN_TRY(create_mixed_plan<GSN_FIELD_MNT4753_FR>(&impl, n, inverse)); break;
    case GSN_FIELD_MNT6753_FR: GSN_TRY(create_mixed_plan<GSN_FIELD_MNT6753_FR>(&impl, n, inverse)); break;
    default: return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id");
    }
    *plan = new gsn_plan{impl};
    return GSN_OK;
}

int gsn_plan_create_ext(gsn_plan **plan, int field, unsigned log_n, const void *shift, int inverse)
{
    if (!plan)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "plan is NULL");
    *plan = nullptr;
    GSN_TRY(ensure_device());
    PlanBase *impl = nullptr;
    switch (field) {
    case GSN_FIELD_FP32: GSN_TRY(create_ext_plan<GSN_FIELD_FP32>(&impl, log_n, shift, inverse)); break;
    case GSN_FIELD_MNT4753_FR: GSN_TRY(create_ext_plan<GSN_FIELD_MNT4753_FR>(&impl, log_n, shift, inverse)); break;
    case GSN_FIELD_MNT6753_FR: GSN_TRY(create_ext_plan<GSN_FIELD_MNT6753_FR>(&impl, log_n, shift, inverse)); break;
    default: return set_error(GSN_ERR_INVALID_ARGUMENT, "unknown field id");
    }
    *plan = new gsn_plan{impl};
    return GSN_OK;
}

int gsn_scale(int field, void *d_out, const void *d_in, size_t n, const void *s, void *stream)
{
    GSN_TRY(check_args(field, d_out, d_in, n));
    if (!s)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL scale");
    GSN_TRY(take_injected_fault());
    GSN_TRY(ensure_device());
    hipStream_t st = (hipStream_t)stream;
    switch (field) {
    case GSN_FIELD_FP32:
        return scale_t<GSN_FIELD_FP32>(d_out, d_in, n, fields::Fp32::from_raw((const uint32_t *)s), st);
    case GSN_FIELD_MNT4753_FR:
        return scale_t<GSN_FIELD_MNT4753_FR>(d_out, d_in, n, fields::Scalar::from_raw((const uint32_t *)s), st);
    default:
        return scale_t<GSN_FIELD_MNT6753_FR>(d_out, d_in, n, fields::ScalarMNT6::from_raw((const uint32_t *)s), st);
    }
}

int gsn_coset_shift(int field, void *d_out, const void *d_in, size_t n, const void *g, int inverse, void *stream)
{
    GSN_TRY(check_args(field, d_out, d_in, n));
    if (!g)
        return set_error(GSN_ERR_INVALID_ARGUMENT, "NULL generator");
    GSN_TRY(take_injected_fault());
    GSN_TRY(ensure_device());
    hipStream_t st = (hipStream_t)stream;
    switch (field) {
    case GSN_FIELD_FP32: return coset_t<GSN_FIELD_FP32>(d_out, d_in, n, g, inverse, st);
    case GSN_FIELD_MNT4753_FR: return coset_t<GSN_FIELD_MNT4753_FR>(d_out, d_in, n, g, inverse, st);
    default: return coset_t<GSN_FIELD_MNT6753_FR>(d_out, d_in, n, g, inverse, st);
    }
}

int gsn_to_montgomery(int field, void *d_out, const void *d_in, size_t n, void *stream)
{
    GSN_TRY(check_args(field, d_out, d_in, n));
    GSN_TRY(take_injected_fault());
    GSN_TRY(ensure_device());
    hipStream_t st = (hipStream_t)stream;
    switch (field) {
    case GSN_FIELD_FP32: return scale_t<GSN_FIELD_FP32>(d_out, d_in, n, value_r<GSN_FIELD_FP32>(), st);
    case GSN_FIELD_MNT4753_FR:
        return scale_t<GSN_FIELD_MNT4753_FR>(d_out, d_in, n, value_r<GSN_FIELD_MNT4753_FR>(), st);
    default: return scale_t<GSN_FIELD_MNT6753_FR>(d_out, d_in, n, value_r<GSN_FIELD_MNT6753_FR>(), st);
    }
}

int gsn_from_montgomery(int field, void *d_out, const void *d_in, size_t n, void *stream)
{
    GSN_TRY(check_args(field, d_out, d_in, n));
    GSN_TRY(take_injected_fault());
    GSN_TRY(ensure_device());
    hipStream_t st = (hipStream_t)stream;
    switch (field) {
    case GSN_FIELD_FP32: return scale_t<GSN_FIELD_FP32>(d_out, d_in, n, value_r<GSN_FIELD_FP32>().mul_inv(), st);
    case GSN_FIELD_MNT4753_FR:
        return scale_t<GSN_FIELD_MNT4753_FR>(d_out, d_in, n, value_r<GSN_FIELD_MNT4753_FR>().mul_inv(), st);
    default:
        return scale_t<GSN_FIELD_MNT6753_FR>(d_out, d_in, n, value_r<GSN_FIELD_MNT6753_FR>().mul_inv(), st);
    }
}

}  // extern "C"
