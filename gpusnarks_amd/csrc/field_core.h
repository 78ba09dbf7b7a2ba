// Device field arithmetic for the NTT kernels (gfx950).
//
// Two field policies share one interface so the transform kernels are
// written once:
//
//   F32      p = 3*2^30 + 1, one u32 per element, Montgomery radix 2^32.
//            Twiddles carry a precomputed companion w' = w * p^-1 mod 2^32,
//            so a product is three 32-bit multiplies and three cheap ops.
//   F753<P>  753-bit MNT fields, 24 x u32 limbs per element in memory.  The
//            product by a twiddle estimates its quotient from a precomputed
//            wq = floor(w 2^783 / p) (Shoup's method carried to 27 digits of
//            29 bits): every column sum fits a 64-bit accumulator, so each
//            partial product is a single v_mad_u64_u32 / v_mad_i64_i32 with
//            no carry handling, 1158 of them against the 1326 of a
//            26-digit Montgomery product.  Sums and differences are not
//            reduced at all; the last pass reduces to [0, p).
//
// Replaces the reference's field core (cuda/device_field.h:67-215,
// cuda/device_field_operators.h:92-258); the reference's 24-limb CIOS
// (device_field_operators.h:151-187) is the shape of the oracle's multiply,
// not of this one.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "gsn_field_params.h"

namespace gsn {

// One spare LDS word after every 32: strides that are multiples of 32
// words (row-to-row steps, transposed reads) then spread over the banks.
__host__ __device__ constexpr int lds_pad(int i) { return i + (i >> 5); }
__host__ __device__ constexpr int lds_padded_words(int n) { return n + (n >> 5) + 1; }

// ------------------------------------------------------------------ F32

struct F32 {
    using elem_t = uint32_t;
    struct tw_t {
        uint32_t w;   // twiddle, Montgomery form (radix 2^32)
        uint32_t wq;  // w * p^-1 mod 2^32
    };
    static constexpr int LIMBS = 1;
    static constexpr uint32_t P = Fp32Params::P;

    // Montgomery product x * w * 2^-32 mod p for x, w in [0, p), result in
    // [0, p), by the subtractive form: with m = x*w*p^-1 mod 2^32, x*w - m*p
    // is divisible by 2^32 and its low words cancel exactly, so
    // (x*w - m*p) / 2^32 = hi(x*w) - hi(m*p), which lies in (-p, p).
    __device__ __forceinline__ static uint32_t mul(uint32_t x, tw_t t)
    {
        const uint32_t m = x * t.wq;
        const uint32_t h = __umulhi(x, t.w);
        const uint32_t q = __umulhi(m, P);
        unsigned bw;
        const uint32_t r = __builtin_subc(h, q, 0u, &bw);  // borrow: h < q
        return bw ? r + P : r;
    }
    __device__ __forceinline__ static uint32_t add(uint32_t a, uint32_t b)
    {
        unsigned bw;
        const uint32_t r = __builtin_subc(a, P - b, 0u, &bw);  // a + b - p; borrow iff a + b < p
        return bw ? r + P : r;
    }
    __device__ __forceinline__ static uint32_t sub(uint32_t a, uint32_t b)
    {
        unsigned bw;
        const uint32_t r = __builtin_subc(a, b, 0u, &bw);
        return bw ? r + P : r;
    }
    __device__ __forceinline__ static uint32_t canon(uint32_t a) { return a; }
    // the canonical sum and difference (add and sub already reduce)
    __device__ __forceinline__ static uint32_t add_mod(uint32_t a, uint32_t b) { return add(a, b); }
    __device__ __forceinline__ static uint32_t sub_mod(uint32_t a, uint32_t b) { return sub(a, b); }

    __device__ __forceinline__ static elem_t load(const elem_t *base, size_t i) { return base[i]; }
    __device__ __forceinline__ static void store(elem_t *base, size_t i, const elem_t &v) { base[i] = v; }
    __device__ __forceinline__ static tw_t load_tw(const tw_t *tab, size_t i)
    {
        const uint2 v = *reinterpret_cast<const uint2 *>(tab + i);
        return tw_t{v.x, v.y};
    }
    // LDS: u32 array, element i at word lds_pad(i).
    static constexpr int LDS_WORDS_PER_ELEM = 1;
    __device__ __forceinline__ static void lds_store(uint32_t *lds, int stride, int i, const elem_t &v)
    {
        (void)stride;
        lds[lds_pad(i)] = v;
    }
    __device__ __forceinline__ static elem_t lds_load(const uint32_t *lds, int stride, int i)
    {
        (void)stride;
        return lds[lds_pad(i)];
    }
};

// ----------------------------------------------------------------- F753

struct E753 {
    uint32_t v[24];
};
// four u32 as a native vector (usable in any address space, unlike uint4)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
// A twiddle of the 753-bit fields: the residue w itself (canonical, not in
// Montgomery form) as 26 digits of 29 bits, and wq = floor(w 2^783 / p) as
// 27 digits; padded to 224 bytes so it moves as 14 aligned quads, w in
// quads 0..6 and wq in quads 7..13 (the product uses them at different
// times, so each half can be loaded on its own).
struct TW753 {
    uint32_t w[26];
    uint32_t pad0[2];
    uint32_t wq[27];
    uint32_t pad1;
};

// Digits [29i, 29i + 29) of a value held in 24 u32 limbs, 27 digits.
__host__ __device__ inline void limbs_to_digits27(uint32_t d[27], const uint32_t y[24])
{
    for (int i = 0; i < 27; ++i) {
        const int bit = 29 * i, w = bit >> 5, sh = bit & 31;
        const uint64_t lo = y[w];
        const uint64_t hi = (w + 1 < 24) ? y[w + 1] : 0u;
        d[i] = (uint32_t)(((hi << 32) | lo) >> sh) & ((1u << 29) - 1u);
    }
}

// wq = floor(w 2^783 / p) from rem = w 2^783 mod p (canonical): w 2^783 -
// rem is an exact multiple of p, so wq = (2^783 - rem) * p^-1 mod 2^783
// (and wq < 2^783 since w < p).  29-bit digits throughout.
template <class P>
__host__ __device__ inline void shoup_wq(uint32_t wq[27], const uint32_t rem[24])
{
    constexpr uint32_t M = (1u << 29) - 1u;
    uint32_t r[27], neg[27];
    limbs_to_digits27(r, rem);
    int64_t b = 0;  // (0 - rem) mod 2^783, digit by digit
    for (int i = 0; i < 27; ++i) {
        const int64_t v = -(int64_t)r[i] - b;
        neg[i] = (uint32_t)v & M;
        b = v < 0 ? 1 : 0;
    }
    uint64_t acc = 0;
    for (int c = 0; c < 27; ++c) {
        for (int i = 0; i <= c; ++i)
            acc += (uint64_t)neg[i] * P::PINV783[c - i];
        wq[c] = (uint32_t)acc & M;
        acc >>= 29;
    }
}

// 753-bit field arithmetic for the NTT kernels.
//
// Elements are 24 x u32 little-endian limbs, the public Montgomery
// representation.  The transform is linear, so multiplying a Montgomery
// representation by the plain twiddle value w keeps it one.  Inside a
// transform values are only partly reduced: every intermediate is below
// 2^767 (a transform of 2^L adds at most 3p per stage to canonical inputs,
// and (1 + 3*30) p < 2^761), and the last pass reduces to [0, p).
//
// The product y*w mod p, for any y < 2^768, is formed from a quotient
// estimate (Shoup's method with a precomputed wq = floor(w 2^783 / p)):
//   q = floor(y * wq / 2^783)  from the top columns of y*wq only,
//   t = y*w - q*p              computed modulo 2^783 (its low 27 digits).
// wq >= w 2^783/p - 1 and y < 2^783 make the exact quotient estimate at most
// one below floor(y w / p); dropping the columns below 25 of y*wq changes
// the sum by less than 25 * 2^(58 + 29*24) < 2^759 = 2^-24 * 2^783, at most
// one more.  So t lies in [0, 3p).  Digits are 29 bits: every column of at
// most 27 + 27 products of 58 bits fits one 64-bit accumulator, signed for
// the q*p terms.
template <class P>
struct F753 {
    using elem_t = E753;
    using tw_t = TW753;
    static constexpr int LIMBS = 24;
    static constexpr uint32_t M29 = (1u << 29) - 1u;

    // bits [29i, 29i + 29) of a 768-bit value (digit 26 has 14 bits)
    __device__ __forceinline__ static void to29(uint32_t a[27], const uint32_t y[24])
    {
#pragma unroll
        for (int i = 0; i < 27; ++i) {
            const int bit = 29 * i, w = bit >> 5, sh = bit & 31;
            const uint32_t lo = y[w];
            const uint32_t hi = (w + 1 < 24) ? y[w + 1] : 0u;
            // 32-bit funnel shift (v_alignbit_b32); a 64-bit shift of the
            // pair would be folded into unaligned loads of a stack copy
            const uint32_t v = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
            a[i] = (i == 26) ? v : (v & M29);
        }
    }
    // inverse of to29 for digits < 2^29 and a value < 2^768
    __device__ __forceinline__ static void from29(uint32_t r[24], const uint32_t d[27])
    {
#pragma unroll
        for (int j = 0; j < 24; ++j) {
            const int bit = 32 * j, i0 = bit / 29, sh = bit % 29;
            uint32_t v = d[i0] >> sh;
            if (i0 + 1 < 27)
                v |= d[i0 + 1] << (29 - sh);
            if (58 - sh < 32 && i0 + 2 < 27)
                v |= d[i0 + 2] << (58 - sh);
            r[j] = v;
        }
    }

    // general product of two elements in Montgomery form (R = 2^768):
    // a b 2^-768 mod p, canonical, for canonical a and b.  REDC in 29-bit
    // digits with radix 2^783 on a' = a 2^15 (a' b 2^-783 = a b 2^-768),
    // column by column: column c < 27 fixes the reduction digit m_c that
    // clears it; columns 27..53 are the result, a' b + m p < 2^783 2p.
    // Each column holds at most 53 products of 58 bits: 64-bit sums.
    __device__ __forceinline__ static E753 mont_mul(const E753 &a, const E753 &b)
    {
        uint32_t x[27], y[27], m[27], r[27];
#pragma unroll
        for (int i = 0; i < 27; ++i) {
            // digit i of a 2^15: bits [29i - 15, 29i + 14) of a
            const int bit = 29 * i - 15;
            if (bit < 0) {
                x[i] = (a.v[0] << 15) & M29;
            } else {
                const int w = bit >> 5, sh = bit & 31;
                const uint32_t lo = a.v[w];
                const uint32_t hi = (w + 1 < 24) ? a.v[w + 1] : 0u;
                x[i] = (sh ? ((lo >> sh) | (hi << (32 - sh))) : lo) & M29;
            }
        }
        to29(y, b.v);
        constexpr uint32_t MP29 = P::MPRIME & M29;  // -p^-1 mod 2^29
        uint64_t acc = 0;
#pragma unroll
        for (int c = 0; c < 27; ++c) {
#pragma unroll
            for (int i = 0; i <= c; ++i)
                acc += (uint64_t)x[i] * y[c - i];
#pragma unroll
            for (int i = (c > 25 ? c - 25 : 0); i < c; ++i)
                acc += (uint64_t)m[i] * P::MOD29[c - i];
            m[c] = ((uint32_t)acc * MP29) & M29;
            acc += (uint64_t)m[c] * P::MOD29[0];
            acc >>= 29;
        }
#pragma unroll
        for (int c = 27; c < 54; ++c) {
#pragma unroll
            for (int i = c - 26; i < 27; ++i)
                acc += (uint64_t)x[i] * y[c - i];
#pragma unroll
            for (int i = c - 25; i < 27; ++i)
                acc += (uint64_t)m[i] * P::MOD29[c - i];
            r[c - 27] = (uint32_t)acc & M29;
            acc >>= 29;
        }
        E753 t;
        from29(t.v, r);
        return sub_p_if_geq(t);
    }
    // x - p if x >= p, for x < 2p
    __device__ __forceinline__ static E753 sub_p_if_geq(const E753 &x)
    {
        uint32_t d[24];
        unsigned bw = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            d[i] = __builtin_subc(x.v[i], P::MOD[i], bw, &bw);
        E753 o;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            o.v[i] = bw ? x.v[i] : d[i];
        return o;
    }
    // (a + b) mod p and (a - b) mod p for canonical a, b
    __device__ __forceinline__ static E753 add_mod(const E753 &a, const E753 &b) { return sub_p_if_geq(add(a, b)); }
    __device__ __forceinline__ static E753 sub_mod(const E753 &a, const E753 &b)
    {
        E753 d;
        unsigned bw = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            d.v[i] = __builtin_subc(a.v[i], b.v[i], bw, &bw);
        if (bw) {
            unsigned c = 0;
#pragma unroll
            for (int i = 0; i < 24; ++i)
                d.v[i] = __builtin_addc(d.v[i], P::MOD[i], c, &c);
        }
        return d;
    }

    // y * w mod p as a value in [0, 3p), for any y < 2^768, in two halves
    // (so that a caller can issue loads between them): mul_q forms the
    // quotient estimate from the digits a of y and wq, mul_r the remainder.
    // q = floor(a * wq / 2^783): columns 25..52, digits from column 27
    __device__ __forceinline__ static void mul_q(uint32_t q[27], const uint32_t a[27], const uint32_t wq[27])
    {
        uint64_t acc = 0;
#pragma unroll
        for (int c = 25; c <= 52; ++c) {
#pragma unroll
            for (int i = (c > 26 ? c - 26 : 0); i <= (c < 26 ? c : 26); ++i)
                acc += (uint64_t)a[i] * wq[c - i];
            if (c >= 27)
                q[c - 27] = (uint32_t)acc & M29;
            acc >>= 29;
        }
        q[26] = (uint32_t)acc;
    }
    // a*w - q*p modulo 2^783: columns 0..26.  The digit -p_0 goes through
    // an SGPR the compiler cannot see into: for p_0 = 1 it would otherwise
    // turn q_c * -1 into a two-instruction 64-bit subtraction.
    __device__ __forceinline__ static E753 mul_r(const uint32_t a[27], const uint32_t w[26], const uint32_t q[27])
    {
        int32_t neg_p0;
        asm("s_mov_b32 %0, %1" : "=s"(neg_p0) : "i"(-(int32_t)P::MOD29[0]));
        uint32_t r[27];
        int64_t s = 0;
#pragma unroll
        for (int c = 0; c <= 26; ++c) {
#pragma unroll
            for (int i = (c > 25 ? c - 25 : 0); i <= (c < 26 ? c : 26); ++i)
                s += (int64_t)((uint64_t)a[i] * w[c - i]);
#pragma unroll
            for (int i = (c > 25 ? c - 25 : 0); i <= (c < 26 ? c : 26); ++i)
                s += (int64_t)(int32_t)q[i] * (int64_t)(c - i == 0 ? neg_p0 : -(int32_t)P::MOD29[c - i]);
            r[c] = (uint32_t)s & M29;
            s >>= 29;
        }
        E753 o;
        from29(o.v, r);
        return o;
    }
    __device__ __forceinline__ static E753 mul(const E753 &y, const TW753 &t)
    {
        uint32_t a[27], q[27];
        to29(a, y.v);
        mul_q(q, a, t.wq);
        return mul_r(a, t.w, q);
    }

    // the butterfly's sums (x0 + t, x0 + 3p - t) for t < 3p, in place (no
    // reduction), with the three carry chains interleaved limb by limb so
    // that no carry-in waits on the instruction right before it
    __device__ __forceinline__ static void butterfly(E753 &x0, E753 &t)
    {
        unsigned c0 = 0, b1 = 0, c2 = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i) {
            const uint32_t a = x0.v[i], b = t.v[i];
            const uint32_t d = __builtin_subc(P::THREE_P[i], b, b1, &b1);
            x0.v[i] = __builtin_addc(a, b, c0, &c0);
            t.v[i] = __builtin_addc(a, d, c2, &c2);
        }
    }
    // a + b (no reduction: callers keep the sum below 2^768)
    __device__ __forceinline__ static E753 add(const E753 &a, const E753 &b)
    {
        E753 r;
        unsigned c = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            r.v[i] = __builtin_addc(a.v[i], b.v[i], c, &c);
        return r;
    }
    // a + 3p - t for t < 3p (no reduction)
    __device__ __forceinline__ static E753 sub(const E753 &a, const E753 &t)
    {
        uint32_t d[24];
        unsigned bw = 0, c = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            d[i] = __builtin_subc(P::THREE_P[i], t.v[i], bw, &bw);
        E753 r;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            r.v[i] = __builtin_addc(a.v[i], d[i], c, &c);
        return r;
    }
    // x mod p for x < 2^767: a quotient estimate from the top 96 bits in
    // double precision, q_est in {q - 1, q}, then one conditional subtraction
    __device__ __forceinline__ static E753 canon(const E753 &x)
    {
        constexpr double TWO32 = 4294967296.0;
        constexpr double P_TOP = ((double)P::MOD[23] * TWO32 + (double)P::MOD[22]) * TWO32 + (double)P::MOD[21];
        const double x_top = ((double)x.v[23] * TWO32 + (double)x.v[22]) * TWO32 + (double)x.v[21];
        // x_top / P_TOP is x/p to within a relative 2^-50; x/p < 2^15
        const uint32_t q = (uint32_t)(x_top * ((1.0 - 0x1p-40) / P_TOP));
        uint32_t r[24];
        uint64_t carry = 0;
        unsigned bw = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i) {
            const uint64_t pr = (uint64_t)q * P::MOD[i] + carry;
            carry = pr >> 32;
            r[i] = __builtin_subc(x.v[i], (uint32_t)pr, bw, &bw);
        }
        uint32_t d[24];
        bw = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            d[i] = __builtin_subc(r[i], P::MOD[i], bw, &bw);
        E753 o;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            o.v[i] = bw ? r[i] : d[i];
        return o;
    }

    // the twiddle record of a canonical residue v: its digits, and wq from
    // rem = v 2^783 mod p, itself the product of v and the constant 2^783 mod p
    __device__ __forceinline__ static TW753 make_tw(const E753 &v)
    {
        TW753 c;
#pragma unroll
        for (int i = 0; i < 26; ++i)
            c.w[i] = P::C783_29[i];
#pragma unroll
        for (int i = 0; i < 27; ++i)
            c.wq[i] = P::C783_WQ[i];
        const E753 rem = canon(mul(v, c));
        TW753 o;
        uint32_t d[27];
        limbs_to_digits27(d, v.v);
#pragma unroll
        for (int i = 0; i < 26; ++i)
            o.w[i] = d[i];
        shoup_wq<P>(o.wq, rem.v);
        o.pad0[0] = o.pad0[1] = o.pad1 = 0;
        return o;
    }

    __device__ __forceinline__ static E753 load(const E753 *base, size_t i)
    {
        const uint4 *p = reinterpret_cast<const uint4 *>(base + i);
        E753 r;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const uint4 v = p[q];
            r.v[4 * q + 0] = v.x;
            r.v[4 * q + 1] = v.y;
            r.v[4 * q + 2] = v.z;
            r.v[4 * q + 3] = v.w;
        }
        return r;
    }
    __device__ __forceinline__ static void store(E753 *base, size_t i, const E753 &x)
    {
        uint4 *p = reinterpret_cast<uint4 *>(base + i);
#pragma unroll
        for (int q = 0; q < 6; ++q)
            p[q] = make_uint4(x.v[4 * q + 0], x.v[4 * q + 1], x.v[4 * q + 2], x.v[4 * q + 3]);
    }
    __device__ __forceinline__ static TW753 load_tw(const TW753 *tab, size_t i)
    {
        const uint4 *p = reinterpret_cast<const uint4 *>(tab + i);
        uint32_t d[56];
#pragma unroll
        for (int q = 0; q < 14; ++q) {
            const uint4 v = p[q];
            d[4 * q] = v.x;
            d[4 * q + 1] = v.y;
            d[4 * q + 2] = v.z;
            d[4 * q + 3] = v.w;
        }
        TW753 r;
#pragma unroll
        for (int k = 0; k < 26; ++k)
            r.w[k] = d[k];
#pragma unroll
        for (int k = 0; k < 27; ++k)
            r.wq[k] = d[28 + k];
        r.pad0[0] = r.pad0[1] = r.pad1 = 0;
        return r;
    }
    // record i of a table of 14 planes of `stride` quads each (Q: uint4, or
    // u32x4 in any address space)
    template <class Q>
    __device__ __forceinline__ static TW753 load_tw_planes(const Q *tab, uint32_t stride, uint32_t i)
    {
        uint32_t d[56];
#pragma unroll
        for (int q = 0; q < 14; ++q) {
            const auto v = tab[(size_t)q * stride + i];
            d[4 * q] = v.x;
            d[4 * q + 1] = v.y;
            d[4 * q + 2] = v.z;
            d[4 * q + 3] = v.w;
        }
        TW753 r;
#pragma unroll
        for (int k = 0; k < 26; ++k)
            r.w[k] = d[k];
#pragma unroll
        for (int k = 0; k < 27; ++k)
            r.wq[k] = d[28 + k];
        r.pad0[0] = r.pad0[1] = r.pad1 = 0;
        return r;
    }
    // LDS: six planes of 16-byte quads; limbs 4q..4q+3 of element i sit in
    // quad q * stride + lds_pad(i), so an element moves with six
    // ds_write_b128 / ds_read_b128 and lanes touching consecutive elements
    // touch consecutive quads.  `stride` counts quads per plane.
    static constexpr int LDS_WORDS_PER_ELEM = 24;
    __device__ __forceinline__ static void lds_store(uint32_t *lds, int stride, int i, const E753 &x)
    {
        uint4 *q4 = reinterpret_cast<uint4 *>(lds);
        const int pi = lds_pad(i);
#pragma unroll
        for (int q = 0; q < 6; ++q)
            q4[q * stride + pi] = make_uint4(x.v[4 * q], x.v[4 * q + 1], x.v[4 * q + 2], x.v[4 * q + 3]);
    }
    __device__ __forceinline__ static E753 lds_load(const uint32_t *lds, int stride, int i)
    {
        const uint4 *q4 = reinterpret_cast<const uint4 *>(lds);
        const int pi = lds_pad(i);
        E753 r;
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            const uint4 v = q4[q * stride + pi];
            r.v[4 * q] = v.x;
            r.v[4 * q + 1] = v.y;
            r.v[4 * q + 2] = v.z;
            r.v[4 * q + 3] = v.w;
        }
        return r;
    }
};

using FMnt4 = F753<Mnt4753FrParams>;
using FMnt6 = F753<Mnt6753FrParams>;

}  // namespace gsn
