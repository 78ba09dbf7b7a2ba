// Device field arithmetic for the NTT kernels (gfx950).
//
// Two field policies share one interface so the transform kernels are
// written once:
//
//   F32      p = 3*2^30 + 1, one u32 per element, Montgomery radix 2^32.
//            Twiddles carry a precomputed companion w' = w * p^-1 mod 2^32,
//            so a product is three 32-bit multiplies and three cheap ops.
//   F753<P>  753-bit MNT fields, 24 x u32 limbs per element in memory.  The
//            product is a finely integrated product-scanning Montgomery
//            multiply over 26 digits of 29 bits: every column sum fits in a
//            64-bit accumulator, so each partial product is a single
//            v_mad_u64_u32 with no carry handling.  Its radix is 2^754.
//
// Replaces the reference's field core (cuda/device_field.h:67-215,
// cuda/device_field_operators.h:92-258); the reference's 24-limb CIOS
// (device_field_operators.h:151-187) is the shape of the oracle's multiply,
// not of this one.
//
// Value ranges (753-bit): data values live in [0, 2p) between operations.
// mul(y, w) needs y < 2p and a canonical twiddle w < p, and returns < 2p
// because (2p * p + 2^754 * p) / 2^754 < 2p when p < 0.45 * 2^754.
// add/sub return values in [0, 2p).  canon() maps [0, 2p) to [0, p).
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
struct TW753 {
    uint32_t d[26];  // 29-bit digits, value w * 2^754 mod p, canonical
};

template <class P>
struct F753 {
    using elem_t = E753;
    using tw_t = TW753;
    static constexpr int LIMBS = 24;
    static constexpr uint32_t M29 = (1u << 29) - 1u;
    static constexpr bool MONT_FRIENDLY = (P::MOD29[0] == 1u);  // p == 1 mod 2^29

    // bits [29i, 29i + 29) of a value held in 24 u32 limbs (value < 2^754)
    __device__ __forceinline__ static void to29(uint32_t a[26], const uint32_t y[24])
    {
#pragma unroll
        for (int i = 0; i < 26; ++i) {
            const int bit = 29 * i, w = bit >> 5, sh = bit & 31;
            const uint32_t lo = y[w];
            const uint32_t hi = (w + 1 < 24) ? y[w + 1] : 0u;
            // 32-bit funnel shift (v_alignbit_b32); a 64-bit shift of the
            // pair would be folded into unaligned loads of a stack copy
            const uint32_t v = sh ? ((lo >> sh) | (hi << (32 - sh))) : lo;
            a[i] = v & M29;
        }
    }
    // inverse of to29 for digits < 2^29 and a value < 2^768
    __device__ __forceinline__ static void from29(uint32_t r[24], const uint32_t d[26])
    {
#pragma unroll
        for (int j = 0; j < 24; ++j) {
            const int bit = 32 * j, i0 = bit / 29, sh = bit % 29;
            uint32_t v = d[i0] >> sh;
            if (i0 + 1 < 26)
                v |= d[i0 + 1] << (29 - sh);
            if (58 - sh < 32 && i0 + 2 < 26)
                v |= d[i0 + 2] << (58 - sh);
            r[j] = v;
        }
    }

    // Montgomery product y * w * 2^-754 mod p, y < 2p, w canonical in 29-bit
    // digits; result < 2p in 24 limbs.  Finely integrated product scanning.
    __device__ __forceinline__ static E753 mul(const E753 &y, const TW753 &w)
    {
        uint32_t a[26], m[26];
        to29(a, y.v);
        uint64_t acc = 0;
#pragma unroll
        for (int k = 0; k < 26; ++k) {
#pragma unroll
            for (int i = 0; i < k; ++i) {
                acc += (uint64_t)a[i] * w.d[k - i];
                acc += (uint64_t)m[i] * P::MOD29[k - i];
            }
            acc += (uint64_t)a[k] * w.d[0];
            if constexpr (MONT_FRIENDLY) {
                // m' = -1 mod 2^29 and p_0 = 1
                m[k] = (0u - (uint32_t)acc) & M29;
                acc += m[k];
            } else {
                m[k] = ((uint32_t)acc * P::MPRIME29) & M29;
                acc += (uint64_t)m[k] * P::MOD29[0];
            }
            acc >>= 29;
        }
        // high columns; digit k-26 of the result replaces m[k-26], which
        // column k no longer reads
#pragma unroll
        for (int k = 26; k < 51; ++k) {
#pragma unroll
            for (int i = k - 25; i < 26; ++i) {
                acc += (uint64_t)a[i] * w.d[k - i];
                acc += (uint64_t)m[i] * P::MOD29[k - i];
            }
            m[k - 26] = (uint32_t)acc & M29;
            acc >>= 29;
        }
        m[25] = (uint32_t)acc;
        E753 r;
        from29(r.v, m);
        return r;
    }

    // (a + b) reduced to [0, 2p) for a, b in [0, 2p)
    __device__ __forceinline__ static E753 add(const E753 &a, const E753 &b)
    {
        uint32_t s[24], d[24];
        unsigned c = 0, bw = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            s[i] = __builtin_addc(a.v[i], b.v[i], c, &c);
#pragma unroll
        for (int i = 0; i < 24; ++i)
            d[i] = __builtin_subc(s[i], P::TWO_P[i], bw, &bw);
        E753 r;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            r.v[i] = bw ? s[i] : d[i];
        return r;
    }
    // (a - b) reduced to [0, 2p) for a, b in [0, 2p)
    __device__ __forceinline__ static E753 sub(const E753 &a, const E753 &b)
    {
        uint32_t d[24];
        unsigned bw = 0, c = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            d[i] = __builtin_subc(a.v[i], b.v[i], bw, &bw);
        const uint32_t mask = 0u - (uint32_t)bw;
        E753 r;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            r.v[i] = __builtin_addc(d[i], P::TWO_P[i] & mask, c, &c);
        return r;
    }
    // [0, 2p) -> [0, p)
    __device__ __forceinline__ static E753 canon(const E753 &a)
    {
        uint32_t d[24];
        unsigned bw = 0;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            d[i] = __builtin_subc(a.v[i], P::MOD[i], bw, &bw);
        E753 r;
#pragma unroll
        for (int i = 0; i < 24; ++i)
            r.v[i] = bw ? a.v[i] : d[i];
        return r;
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
        const uint2 *p = reinterpret_cast<const uint2 *>(tab + i);
        TW753 r;
#pragma unroll
        for (int q = 0; q < 13; ++q) {
            const uint2 v = p[q];
            r.d[2 * q] = v.x;
            r.d[2 * q + 1] = v.y;
        }
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
