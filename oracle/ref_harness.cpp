// Harness that compiles the reference's host FFT source, test/fft_host.h, as
// it lies under the reference tree, and instantiates its two templates over
// the oracle's field arithmetic.  TEST INFRASTRUCTURE ONLY.
//
// With correct field arithmetic, the reference's FFT loop structure (bit
// reversal, stage order, running twiddle, C-way split and scatter) can be
// compared bit for bit with the oracle's restatement run with
// literal_butterfly = 1 (the reference's butterfly statements,
// test/fft_host.h:44-48).
#include <cstdint>
#include <cstring>
#include <vector>
#include <utility>

#include "oracle.h"
#include "fft_host.h"  // test/fft_host.h from the reference tree (-I<ref>/test)

namespace {

const or_field *g_field = nullptr;  // set before each call; shared by OpenMP threads

struct OracleField {
    uint32_t v[OR_MAX_LIMBS];
    static OracleField zero() { OracleField r; std::memset(r.v, 0, sizeof(r.v)); return r; }
    static OracleField one() { OracleField r; std::memset(r.v, 0, sizeof(r.v)); std::memcpy(r.v, g_field->one, sizeof(uint32_t) * g_field->nlimbs); return r; }
    OracleField operator+(const OracleField &o) const { OracleField r = zero(); or_add(g_field, r.v, v, o.v); return r; }
    OracleField operator-(const OracleField &o) const { OracleField r = zero(); or_sub(g_field, r.v, v, o.v); return r; }
    OracleField operator*(const OracleField &o) const { OracleField r = zero(); or_mul(g_field, r.v, v, o.v); return r; }
    OracleField operator^(uint64_t e) const { OracleField r = zero(); or_pow(g_field, r.v, v, e); return r; }
};

std::vector<OracleField> load(const uint32_t *a, size_t n, int s)
{
    std::vector<OracleField> v(n, OracleField::zero());
    for (size_t i = 0; i < n; ++i)
        std::memcpy(v[i].v, a + i * s, sizeof(uint32_t) * s);
    return v;
}

void store(uint32_t *a, const std::vector<OracleField> &v, int s)
{
    for (size_t i = 0; i < v.size(); ++i)
        std::memcpy(a + i * s, v[i].v, sizeof(uint32_t) * s);
}

}  // namespace

extern "C" int ref_serial_radix2_fft(int field, uint32_t *a, size_t n, const uint32_t *omega)
{
    g_field = or_field_get(field);
    if (!g_field) return -1;
    const int s = g_field->nlimbs;
    std::vector<OracleField> v = load(a, n, s);
    OracleField w = OracleField::zero();
    std::memcpy(w.v, omega, sizeof(uint32_t) * s);
    _basic_serial_radix2_FFT(v, w, OracleField::one());
    store(a, v, s);
    return 0;
}

extern "C" int ref_parallel_radix2_fft_inner(int field, uint32_t *a, size_t n, const uint32_t *omega,
                                             unsigned log_cpus)
{
    g_field = or_field_get(field);
    if (!g_field) return -1;
    const int s = g_field->nlimbs;
    std::vector<OracleField> v = load(a, n, s);
    OracleField w = OracleField::zero();
    std::memcpy(w.v, omega, sizeof(uint32_t) * s);
    _basic_parallel_radix2_FFT_inner(v, w, log_cpus, OracleField::one());
    store(a, v, s);
    return 0;
}
