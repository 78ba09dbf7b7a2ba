/*
 * Drop-in for the reference's FFT entry point, cuda/fft_kernel.h:24-25:
 *
 *     template <typename FieldT> void best_fft(std::vector<FieldT> &a, const FieldT &omg);
 *
 * Same signature and meaning: transforms `a` in place, natural order in and
 * out, X_k = sum_j a_j omg^(j k).  Unlike the reference
 * (cuda/fft_kernel.cu:117-147), omg is honoured and validated, any
 * power-of-two size up to the field's 2-adicity is accepted, and errors
 * throw std::runtime_error instead of printing and calling exit(-1).
 *
 * Instantiated (in libgpusnarks.so) for fields::Scalar (MNT4-753 Fr, the
 * reference's only instantiation, cuda/fft_kernel.cu:150),
 * fields::ScalarMNT6 (MNT6-753 Fr) and fields::Fp32 (p = 3*2^30 + 1).
 * best_ifft is the inverse (omg^-1, then n^-1 scaling), which the reference
 * does not have.
 */
#pragma once

#include <vector>

#include "device_field.h"

#ifdef __cplusplus

template <typename FieldT>
void best_fft(std::vector<FieldT> &a, const FieldT &omg);

template <typename FieldT>
void best_ifft(std::vector<FieldT> &a, const FieldT &omg);

extern template void best_fft<fields::Scalar>(std::vector<fields::Scalar> &, const fields::Scalar &);
extern template void best_fft<fields::ScalarMNT6>(std::vector<fields::ScalarMNT6> &, const fields::ScalarMNT6 &);
extern template void best_fft<fields::Fp32>(std::vector<fields::Fp32> &, const fields::Fp32 &);
extern template void best_ifft<fields::Scalar>(std::vector<fields::Scalar> &, const fields::Scalar &);
extern template void best_ifft<fields::ScalarMNT6>(std::vector<fields::ScalarMNT6> &, const fields::ScalarMNT6 &);
extern template void best_ifft<fields::Fp32>(std::vector<fields::Fp32> &, const fields::Fp32 &);

#endif
