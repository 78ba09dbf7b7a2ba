// The reference's multiexp template (cuda/multi_exp.h:24-25) for field
// elements, over the C ABI: both vectors to device memory, the inner
// product on the GPU (gsn_inner_product), one element back.
#include <stdexcept>
#include <string>
#include <vector>

#include "gpusnarks.h"
#include "multi_exp.h"

namespace {

void check(int st, const char *what)
{
    if (st != GSN_OK)
        throw std::runtime_error(std::string("multiexp: ") + what + ": " + gsn_strerror(st) + " (" +
                                 gsn_last_error() + ")");
}

// device allocation released on every path
struct DevBuf {
    void *p = nullptr;
    explicit DevBuf(size_t bytes) { check(gsn_device_malloc(&p, bytes), "gsn_device_malloc"); }
    ~DevBuf() { gsn_device_free(p); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
};

}  // namespace

template <typename FieldT, typename FieldMul>
FieldT multiexp(std::vector<FieldT> &a, std::vector<FieldMul> &mul)
{
    static_assert(sizeof(FieldT) == sizeof(FieldMul), "field-element multiexp: one field");
    if (a.size() != mul.size())
        throw std::runtime_error("multiexp: the vectors differ in size");
    FieldT result = FieldT::zero();
    if (a.empty())
        return result;
    const size_t bytes = a.size() * sizeof(FieldT);
    DevBuf da(bytes), db(bytes), dr(sizeof(FieldT));
    check(gsn_memcpy_htod(da.p, a.data(), bytes, nullptr), "copy");
    check(gsn_memcpy_htod(db.p, mul.data(), bytes, nullptr), "copy");
    check(gsn_inner_product(gsn_field_traits<FieldT>::id, dr.p, da.p, db.p, a.size(), nullptr), "inner product");
    // synchronous on the null stream, after the inner product's launches
    check(gsn_memcpy_dtoh(result.im_rep, dr.p, sizeof(FieldT), nullptr), "copy");
    return result;
}

template fields::Scalar multiexp<fields::Scalar, fields::Scalar>(std::vector<fields::Scalar> &,
                                                                 std::vector<fields::Scalar> &);
template fields::ScalarMNT6 multiexp<fields::ScalarMNT6, fields::ScalarMNT6>(std::vector<fields::ScalarMNT6> &,
                                                                             std::vector<fields::ScalarMNT6> &);
template fields::Fp32 multiexp<fields::Fp32, fields::Fp32>(std::vector<fields::Fp32> &, std::vector<fields::Fp32> &);
