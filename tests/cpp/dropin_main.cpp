// A caller of the reference's FFT entry point (cuda/fft_kernel.h:24-25),
// written as test/main.cpp:34-60 calls it, compiled with g++ against
// include/ and linked to libgpusnarks.so -- no HIP headers.
//
//   dropin_main host   argument checks and field identities (no GPU)
//   dropin_main gpu    best_fft / best_ifft against an O(n^2) DFT computed
//                      with the host field type, for all three fields, and
//                      multiexp (cuda/multi_exp.h:24-25) against the host
//                      loop of test/multiexp.h:3-14
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "fft_kernel.h"
#include "gpusnarks.h"
#include "multi_exp.h"

static int failures = 0;
#define CHECK(c)                                                    \
    do {                                                            \
        if (!(c)) {                                                 \
            std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
            ++failures;                                             \
        }                                                           \
    } while (0)

template <class F>
F root(unsigned log_n)
{
    F w;
    if (gsn_root_of_unity(gsn_field_traits<F>::id, log_n, w.im_rep) != GSN_OK)
        throw std::runtime_error("gsn_root_of_unity failed");
    return w;
}

template <class F>
std::vector<F> inputs(size_t n, uint32_t seed)
{
    std::vector<F> v(n);
    uint32_t s = seed;
    for (size_t i = 0; i < n; ++i) {
        s = s * 1664525u + 1013904223u;
        v[i] = F(s) * F(s ^ 0x9E3779B9u) + F((uint32_t)i);  // spread over the field
    }
    return v;
}

template <class F>
bool throws(std::vector<F> v, const F &w)
{
    try {
        best_fft<F>(v, w);
    } catch (const std::runtime_error &) {
        return true;
    }
    return false;
}

template <class F>
void host_checks()
{
    CHECK(F(1) == F::one());
    CHECK(F(2) + F(3) == F(5));
    CHECK(F(7) * F(6) == F(42));
    CHECK(F(5) - F(7) == -F(2));
    CHECK(F(12345).mul_inv() * F(12345) == F::one());
    // n not a power of two, and a non-root omega (test/main.cpp:57 passes
    // Scalar(123)): both rejected before any device work
    CHECK(throws<F>(std::vector<F>(3, F(1)), root<F>(2)));
    CHECK(throws<F>(std::vector<F>(16, F(1)), F(123)));
    // multiexp: mismatched sizes throw (the reference asserts); empty is zero
    std::vector<F> a(3, F(1)), b(4, F(1)), e;
    bool threw = false;
    try {
        multiexp<F, F>(a, b);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    CHECK(threw);
    CHECK((multiexp<F, F>(e, e) == F::zero()));
}

// test/multiexp.h:3-14: result = result + a[i] * b[i]
template <class F>
F host_multi_exp(const std::vector<F> &a, const std::vector<F> &b)
{
    F r = F::zero();
    for (size_t i = 0; i < a.size(); ++i)
        r = r + a[i] * b[i];
    return r;
}

template <class F>
void multiexp_checks(size_t n)
{
    std::vector<F> a = inputs<F>(n, 0xA0A0u), b = inputs<F>(n, 0xB0B0u);
    CHECK((multiexp<F, F>(a, b) == host_multi_exp(a, b)));
}

template <class F>
void gpu_checks(unsigned log_n)
{
    const size_t n = (size_t)1 << log_n;
    const F w = root<F>(log_n);
    const std::vector<F> x = inputs<F>(n, 0xD0D0u + log_n);
    std::vector<F> y = x;
    best_fft<F>(y, w);
    // O(n^2) DFT with the host field type
    std::vector<F> wp(n);
    wp[0] = F::one();
    for (size_t i = 1; i < n; ++i)
        wp[i] = wp[i - 1] * w;
    bool same = true;
    for (size_t k = 0; k < n && same; ++k) {
        F acc = F::zero();
        for (size_t j = 0; j < n; ++j)
            acc += x[j] * wp[(j * k) & (n - 1)];
        same = same && (acc == y[k]);
    }
    CHECK(same);
    best_ifft<F>(y, w);
    bool round = true;
    for (size_t i = 0; i < n; ++i)
        round = round && (y[i] == x[i]);
    CHECK(round);
}

int main(int argc, char **argv)
{
    const bool gpu = argc > 1 && std::strcmp(argv[1], "gpu") == 0;
    host_checks<fields::Scalar>();
    host_checks<fields::ScalarMNT6>();
    host_checks<fields::Fp32>();
    if (gpu) {
        gpu_checks<fields::Scalar>(8);
        gpu_checks<fields::ScalarMNT6>(9);
        gpu_checks<fields::Fp32>(12);
        multiexp_checks<fields::Scalar>(1000);
        multiexp_checks<fields::ScalarMNT6>(333);
        multiexp_checks<fields::Fp32>(100000);
        // the reference harness's shape (test/main.cpp:89-132)
        std::vector<fields::Scalar> v(1 << 18, fields::Scalar(1234));
        CHECK((multiexp<fields::Scalar, fields::Scalar>(v, v) == host_multi_exp(v, v)));
    }
    std::printf("%s: %d failure(s)\n", gpu ? "gpu" : "host", failures);
    return failures ? 1 : 0;
}
