// The reference's template entry point (cuda/fft_kernel.h:24-25), defined
// over the C ABI with explicit instantiations, as the reference instantiates
// its template in the library (cuda/fft_kernel.cu:150).
#include <stdexcept>
#include <string>
#include <vector>

#include "fft_kernel.h"
#include "gpusnarks.h"

namespace {

template <typename FieldT>
void run(std::vector<FieldT> &a, const FieldT &omg, int inverse, const char *name)
{
    const int st = gsn_fft_host(gsn_field_traits<FieldT>::id, a.data(), a.size(), omg.im_rep, inverse);
    if (st != GSN_OK)
        throw std::runtime_error(std::string(name) + ": " + gsn_strerror(st) + " (" + gsn_last_error() + ")");
}

}  // namespace

template <typename FieldT>
void best_fft(std::vector<FieldT> &a, const FieldT &omg)
{
    run(a, omg, 0, "best_fft");
}

template <typename FieldT>
void best_ifft(std::vector<FieldT> &a, const FieldT &omg)
{
    run(a, omg, 1, "best_ifft");
}

template void best_fft<fields::Scalar>(std::vector<fields::Scalar> &, const fields::Scalar &);
template void best_fft<fields::ScalarMNT6>(std::vector<fields::ScalarMNT6> &, const fields::ScalarMNT6 &);
template void best_fft<fields::Fp32>(std::vector<fields::Fp32> &, const fields::Fp32 &);
template void best_ifft<fields::Scalar>(std::vector<fields::Scalar> &, const fields::Scalar &);
template void best_ifft<fields::ScalarMNT6>(std::vector<fields::ScalarMNT6> &, const fields::ScalarMNT6 &);
template void best_ifft<fields::Fp32>(std::vector<fields::Fp32> &, const fields::Fp32 &);
