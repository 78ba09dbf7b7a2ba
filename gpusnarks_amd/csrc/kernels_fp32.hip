// Pass kernels of the 32-bit field p = 3*2^30 + 1.
#include "ntt_kernels.h"

namespace gsn {

// Compiled tile shapes (log_r, log_w, log_e): single-pass transforms up to
// 2^12 with one column, 16-column passes of 2^5..2^10 rows, 32-column
// passes (128-byte rows) of 2^5..2^9 rows with 16 elements per thread, and
// 32-column passes of 2^9 and 2^10 rows with 32 elements per thread (a
// 2^10 x 32 tile in 1024 threads, 135 KB of LDS).
#define GSN_F32_SHAPES(X)                                                                 \
    X(1, 0, 1) X(2, 0, 2) X(3, 0, 3) X(4, 0, 4) X(5, 0, 4) X(6, 0, 4) X(7, 0, 4) X(8, 0, 4) \
    X(9, 0, 4) X(10, 0, 4) X(11, 0, 4) X(12, 0, 4) X(5, 4, 4) X(6, 4, 4) X(7, 4, 4)        \
    X(8, 4, 4) X(9, 4, 4) X(10, 4, 4) X(5, 5, 4) X(6, 5, 4) X(7, 5, 4) X(8, 5, 4)          \
    X(9, 5, 4) X(9, 5, 5) X(10, 5, 5)

static inline int shape_key(const PassCfg &c) { return c.log_r * 100 + c.log_w * 10 + c.log_e; }

template <>
bool pass_supported<F32>(const PassCfg &c)
{
    switch (shape_key(c)) {
#define GSN_CASE(R, W, E) case R * 100 + W * 10 + E: return true;
        GSN_F32_SHAPES(GSN_CASE)
#undef GSN_CASE
    default: return false;
    }
}

template <>
void launch_pass<F32>(const PassJob<F32> &job, const PassCfg &c, hipStream_t st)
{
    const size_t grid = ((size_t)1 << (job.a.log_n - c.log_r)) >> c.log_w;
    switch (shape_key(c)) {
#define GSN_CASE(R, W, E)                                                                               \
    case R * 100 + W * 10 + E:                                                                          \
        hipLaunchKernelGGL((ntt_pass_kernel<F32, R, W, E>), dim3((unsigned)grid), dim3(1 << (R + W - E)), 0, st, \
                           job.a);                                                                      \
        break;
        GSN_F32_SHAPES(GSN_CASE)
#undef GSN_CASE
    default: break;
    }
}

template <>
void launch_pass_multi<F32>(const PassJob<F32> *jobs, const uint32_t *tile_job, uint32_t total_tiles,
                            const PassCfg &c, hipStream_t st)
{
    switch (shape_key(c)) {
#define GSN_CASE(R, W, E)                                                                                     \
    case R * 100 + W * 10 + E:                                                                                \
        hipLaunchKernelGGL((ntt_pass_multi_kernel<F32, R, W, E>), dim3(total_tiles), dim3(1 << (R + W - E)), 0, st, \
                           jobs, tile_job);                                                                   \
        break;
        GSN_F32_SHAPES(GSN_CASE)
#undef GSN_CASE
    default: break;
    }
}

template <>
void launch_scale<F32>(const uint32_t *in, uint32_t *out, F32::tw_t s, size_t n, int do_mul, hipStream_t st)
{
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(scale_kernel<F32>, dim3(blocks), dim3(256), 0, st, in, out, s, n, do_mul);
}

}  // namespace gsn
