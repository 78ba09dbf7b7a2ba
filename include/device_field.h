/*
 * Field types of the NTT API -- drop-in for the reference's fields::Scalar.
 *
 * Replaces cuda/device_field.h:67-215 (type surface) and
 * cuda/device_field_operators.h:92-258 (arithmetic) with correct prime-field
 * arithmetic, and keeps the field-type contract the reference's FFT
 * templates rely on (test/fft_host.h, cuda/fft_kernel.cu:52-147):
 *   trivially copyable, default-constructible, static zero()/one(),
 *   FieldT(uint32_t), FieldT(const uint32_t *limbs), + - * ^ == and
 *   the named operations of fields/field.h:33-223 (is_zero, square,
 *   negate, add, subtract, mul, mul_inv, pow).
 *
 * Representation: little-endian u32 limbs holding a residue in [0, p) in
 * Montgomery form with radix R = 2^(32 * LIMBS) (2^768 for the 753-bit
 * fields, as the reference's 24-limb CIOS implies; 2^32 for Fp32).
 *   FieldT(const uint32_t *limbs) copies raw limbs (a Montgomery
 *     representation), as the reference's constructor does (:100-104);
 *   FieldT(uint32_t v) is the field element v, i.e. v*R mod p, so that
 *     FieldT(1) == one() and the reference's known answers hold
 *     (cuda/device_field_operator_test.cpp:222-320);
 *   one() is R mod p (the reference returns raw 1, device_field.h:86-91).
 *
 * The same header compiles for the host with any C++14 compiler and for the
 * device under hipcc; the NTT kernels use their own field core.
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstdio>

#include "gsn_field_params.h"

#if defined(__HIPCC__)
#define GSN_HD __host__ __device__
#else
#define GSN_HD
#endif

namespace fields {

namespace detail {

template <int L>
GSN_HD inline bool geq(const uint32_t *a, const uint32_t *b)
{
    for (int i = L - 1; i >= 0; --i)
        if (a[i] != b[i])
            return a[i] > b[i];
    return true;
}

template <int L>
GSN_HD inline uint32_t add_n(uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    uint64_t c = 0;
    for (int i = 0; i < L; ++i) {
        c += (uint64_t)a[i] + b[i];
        r[i] = (uint32_t)c;
        c >>= 32;
    }
    return (uint32_t)c;
}

template <int L>
GSN_HD inline uint32_t sub_n(uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    uint32_t borrow = 0;
    for (int i = 0; i < L; ++i) {
        const uint64_t d = (uint64_t)a[i] - b[i] - borrow;
        r[i] = (uint32_t)d;
        borrow = (uint32_t)(d >> 63);
    }
    return borrow;
}

// CIOS Montgomery product (the algorithm of device_field_operators.h:151-187
// with the correct m' and a complete final subtraction); inputs canonical.
template <int L>
GSN_HD inline void mont_mul(uint32_t *r, const uint32_t *a, const uint32_t *b, const uint32_t *mod,
                            uint32_t mprime)
{
    uint32_t t[L + 2] = {0};
    for (int i = 0; i < L; ++i) {
        uint64_t C = 0, x;
        for (int j = 0; j < L; ++j) {
            x = (uint64_t)t[j] + (uint64_t)a[j] * b[i] + C;
            t[j] = (uint32_t)x;
            C = x >> 32;
        }
        x = (uint64_t)t[L] + C;
        t[L] = (uint32_t)x;
        t[L + 1] = (uint32_t)(x >> 32);
        const uint32_t m = t[0] * mprime;
        x = (uint64_t)t[0] + (uint64_t)m * mod[0];
        C = x >> 32;
        for (int j = 1; j < L; ++j) {
            x = (uint64_t)t[j] + (uint64_t)m * mod[j] + C;
            t[j - 1] = (uint32_t)x;
            C = x >> 32;
        }
        x = (uint64_t)t[L] + C;
        t[L - 1] = (uint32_t)x;
        t[L] = t[L + 1] + (uint32_t)(x >> 32);
    }
    if (t[L] || geq<L>(t, mod))
        sub_n<L>(t, t, mod);
    for (int i = 0; i < L; ++i)
        r[i] = t[i];
}

}  // namespace detail

// Prime field with LIMBS x u32 limbs.  Params supplies MOD, MPRIME, ONE, R2.
template <class Params, int LIMBS_>
struct PrimeField {
    static constexpr int LIMBS = LIMBS_;
    // Same member name and zero default as the reference (device_field.h:75).
    uint32_t im_rep[LIMBS] = {};

    GSN_HD static const uint32_t *mod() { return Params::MOD; }

    PrimeField() = default;
    // The field element v (stored as v*R mod p).
    GSN_HD explicit PrimeField(uint32_t v)
    {
        uint32_t raw[LIMBS] = {0};
        raw[0] = v;
        detail::mont_mul<LIMBS>(im_rep, raw, Params::R2, Params::MOD, Params::MPRIME);
    }
    // Raw limbs (a Montgomery representation), device_field.h:100-104.
    GSN_HD explicit PrimeField(const uint32_t *limbs)
    {
        for (int i = 0; i < LIMBS; ++i)
            im_rep[i] = limbs[i];
    }

    GSN_HD static PrimeField zero()
    {
        PrimeField r;
        for (int i = 0; i < LIMBS; ++i)
            r.im_rep[i] = 0;
        return r;
    }
    GSN_HD static PrimeField one() { return PrimeField(Params::ONE); }
    GSN_HD static PrimeField from_raw(const uint32_t *limbs) { return PrimeField(limbs); }
    // x * R mod p for a canonical integer x given as limbs
    GSN_HD static PrimeField from_canonical(const uint32_t *x)
    {
        PrimeField r;
        detail::mont_mul<LIMBS>(r.im_rep, x, Params::R2, Params::MOD, Params::MPRIME);
        return r;
    }
    // canonical integer limbs of this element
    GSN_HD void to_canonical(uint32_t *out) const
    {
        uint32_t one_raw[LIMBS] = {0};
        one_raw[0] = 1;
        detail::mont_mul<LIMBS>(out, im_rep, one_raw, Params::MOD, Params::MPRIME);
    }

    GSN_HD bool is_zero() const
    {
        for (int i = 0; i < LIMBS; ++i)
            if (im_rep[i])
                return false;
        return true;
    }
    GSN_HD bool operator==(const PrimeField &o) const
    {
        for (int i = 0; i < LIMBS; ++i)
            if (im_rep[i] != o.im_rep[i])
                return false;
        return true;
    }
    GSN_HD bool operator!=(const PrimeField &o) const { return !(*this == o); }

    GSN_HD PrimeField operator+(const PrimeField &o) const
    {
        PrimeField r;
        const uint32_t c = detail::add_n<LIMBS>(r.im_rep, im_rep, o.im_rep);
        if (c || detail::geq<LIMBS>(r.im_rep, Params::MOD))
            detail::sub_n<LIMBS>(r.im_rep, r.im_rep, Params::MOD);
        return r;
    }
    GSN_HD PrimeField operator-(const PrimeField &o) const
    {
        PrimeField r;
        if (detail::sub_n<LIMBS>(r.im_rep, im_rep, o.im_rep))
            detail::add_n<LIMBS>(r.im_rep, r.im_rep, Params::MOD);
        return r;
    }
    GSN_HD PrimeField operator-() const { return zero() - *this; }
    GSN_HD PrimeField operator*(const PrimeField &o) const
    {
        PrimeField r;
        detail::mont_mul<LIMBS>(r.im_rep, im_rep, o.im_rep, Params::MOD, Params::MPRIME);
        return r;
    }
    GSN_HD PrimeField &operator+=(const PrimeField &o) { return *this = *this + o; }
    GSN_HD PrimeField &operator-=(const PrimeField &o) { return *this = *this - o; }
    GSN_HD PrimeField &operator*=(const PrimeField &o) { return *this = *this * o; }

    GSN_HD PrimeField square() const { return *this * *this; }
    GSN_HD PrimeField negate() const { return -*this; }
    // square-and-multiply inside the Montgomery domain
    GSN_HD PrimeField pow(uint64_t e) const
    {
        PrimeField acc = one(), base = *this;
        while (e) {
            if (e & 1)
                acc *= base;
            e >>= 1;
            if (e)
                base = base.square();
        }
        return acc;
    }
    GSN_HD PrimeField operator^(uint64_t e) const { return pow(e); }
    // Fermat inversion, x^(p-2); zero maps to zero
    GSN_HD PrimeField mul_inv() const
    {
        uint32_t e[LIMBS];
        uint32_t two[LIMBS] = {0};
        two[0] = 2;
        detail::sub_n<LIMBS>(e, Params::MOD, two);
        PrimeField acc = one();
        for (int i = LIMBS - 1; i >= 0; --i)
            for (int b = 31; b >= 0; --b) {
                acc = acc.square();
                if ((e[i] >> b) & 1)
                    acc *= *this;
            }
        return acc;
    }

    // fields/field.h named-operation contract (:33-223)
    GSN_HD static PrimeField add(const PrimeField &a, const PrimeField &b) { return a + b; }
    GSN_HD static PrimeField subtract(const PrimeField &a, const PrimeField &b) { return a - b; }
    GSN_HD static PrimeField mul(const PrimeField &a, const PrimeField &b) { return a * b; }

    // device_field.h:197-214 test helpers
    static void print(const PrimeField &f)
    {
        for (int i = 0; i < LIMBS; ++i)
            std::printf("%u, ", f.im_rep[i]);
        std::printf("\n");
    }
    static bool testEquality(const PrimeField &a, const PrimeField &b) { return a == b; }
};

// Single-limb parameter adapter for Fp32 (arrays for the generic code).
struct Fp32ArrayParams {
    static constexpr uint32_t MOD[1] = {gsn::Fp32Params::P};
    static constexpr uint32_t MPRIME = gsn::Fp32Params::MPRIME;
    static constexpr uint32_t ONE[1] = {gsn::Fp32Params::ONE};
    static constexpr uint32_t R2[1] = {gsn::Fp32Params::R2};
};

// The reference's type name: MNT4-753 Fr (cuda/device_field.h:42-65).
using Scalar = PrimeField<gsn::Mnt4753FrParams, 24>;
// MNT6-753 Fr (cuda/device_field.h:445-446), 2-adicity 30.
using ScalarMNT6 = PrimeField<gsn::Mnt6753FrParams, 24>;
// p = 3*2^30 + 1, 2-adicity 30.
using Fp32 = PrimeField<Fp32ArrayParams, 1>;

static_assert(sizeof(Scalar) == 96, "Scalar must keep the reference's 96-byte layout");
static_assert(sizeof(ScalarMNT6) == 96, "ScalarMNT6 must be 96 bytes");
static_assert(sizeof(Fp32) == 4, "Fp32 must be 4 bytes");

}  // namespace fields

// Field identifiers of the C ABI (gpusnarks.h) for each type.
template <class FieldT>
struct gsn_field_traits;
template <>
struct gsn_field_traits<fields::Fp32> {
    static constexpr int id = 0;
};
template <>
struct gsn_field_traits<fields::Scalar> {
    static constexpr int id = 1;
};
template <>
struct gsn_field_traits<fields::ScalarMNT6> {
    static constexpr int id = 2;
};
