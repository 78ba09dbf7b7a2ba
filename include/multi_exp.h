/*
 * Drop-in for the reference's multiexp entry point, cuda/multi_exp.h:24-25:
 *
 *     template <typename FieldT, typename FieldMul>
 *     FieldT multiexp(std::vector<FieldT> &a, std::vector<FieldMul> &mul);
 *
 * Same signature and meaning for field elements: returns sum_i a[i] * mul[i]
 * (cuda/multi_exp.cu:104-135; its host version test/multiexp.h:3-14).
 * Instantiated (in libgpusnarks.so) for fields::Scalar x fields::Scalar, the
 * reference's first instantiation (cuda/multi_exp.cu:139), and for
 * fields::ScalarMNT6 and fields::Fp32.  Sizes must match (the reference
 * asserts it); errors throw std::runtime_error.  The curve instantiation
 * (fields::mnt4753_G1 x Scalar, :140) is not provided.
 */
#pragma once

#include <vector>

#include "device_field.h"

#ifdef __cplusplus

template <typename FieldT, typename FieldMul>
FieldT multiexp(std::vector<FieldT> &a, std::vector<FieldMul> &mul);

extern template fields::Scalar multiexp<fields::Scalar, fields::Scalar>(std::vector<fields::Scalar> &,
                                                                        std::vector<fields::Scalar> &);
extern template fields::ScalarMNT6 multiexp<fields::ScalarMNT6, fields::ScalarMNT6>(
    std::vector<fields::ScalarMNT6> &, std::vector<fields::ScalarMNT6> &);
extern template fields::Fp32 multiexp<fields::Fp32, fields::Fp32>(std::vector<fields::Fp32> &,
                                                                  std::vector<fields::Fp32> &);

#endif
