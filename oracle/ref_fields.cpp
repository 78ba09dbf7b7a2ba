// Compiles the reference's 753-bit field type (cuda/device_field.h and
// cuda/device_field_operators.h) in its host mode (-DDEBUG, as the
// reference's own operator test does: cuda/device_field_operator_test.cpp:19)
// and exports the operations whose arithmetic is correct, so the oracle's
// add and subtract can be compared with them.  TEST INFRASTRUCTURE ONLY.
#include <cstdint>
#include <cstring>
#include "device_field.h"
#include "device_field_operators.h"

extern "C" void ref_scalar_add(uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    fields::Scalar x(a), y(b);
    fields::Scalar z = x + y;
    std::memcpy(r, z.im_rep, sizeof(z.im_rep));
}

extern "C" void ref_scalar_sub(uint32_t *r, const uint32_t *a, const uint32_t *b)
{
    fields::Scalar x(a), y(b);
    fields::Scalar z = x - y;
    std::memcpy(r, z.im_rep, sizeof(z.im_rep));
}

extern "C" void ref_scalar_modulus(uint32_t *r)
{
    std::memcpy(r, fields::_mod, sizeof(uint32_t) * 24);
}
