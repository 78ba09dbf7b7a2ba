// Pass and table kernels of the 753-bit field MNT4-753 Fr.
#include "ntt_kernels.h"

namespace gsn {

template <>
bool pass_supported<FMnt4>(const PassCfg &c)
{
    return c.log_e == 1 && c.log_r >= 1 && c.log_w >= 0 && c.log_r + c.log_w <= T753_LOG_TILE;
}

static void launch753(const PassJob<FMnt4> &single, const PassJob<FMnt4> *jobs, const uint32_t *tile_job,
                      uint32_t grid, int log_tile, hipStream_t st)
{
    switch (log_tile) {
    case 7:
        hipLaunchKernelGGL((ntt753_kernel<FMnt4, 7>), dim3(grid), dim3(64), 0, st, single, jobs, tile_job);
        break;
    case 8:
        hipLaunchKernelGGL((ntt753_kernel<FMnt4, 8>), dim3(grid), dim3(128), 0, st, single, jobs, tile_job);
        break;
    case 9:
        hipLaunchKernelGGL((ntt753_kernel<FMnt4, 9>), dim3(grid), dim3(256), 0, st, single, jobs, tile_job);
        break;
    default:
        hipLaunchKernelGGL((ntt753_kernel<FMnt4, 10>), dim3(grid), dim3(512), 0, st, single, jobs, tile_job);
        break;
    }
}

template <>
void launch_pass<FMnt4>(const PassJob<FMnt4> &job, const PassCfg &c, hipStream_t st)
{
    const size_t grid = ((size_t)1 << (job.a.log_n - c.log_r)) >> c.log_w;
    launch753(job, nullptr, nullptr, (uint32_t)grid, t753_log_tile(c), st);
}

template <>
void launch_pass_multi<FMnt4>(const PassJob<FMnt4> *jobs, const uint32_t *tile_job, uint32_t total_tiles,
                             const PassCfg &c, hipStream_t st)
{
    PassJob<FMnt4> unused{};
    launch753(unused, jobs, tile_job, total_tiles, t753_log_tile(c), st);
}

template <>
void launch_scale<FMnt4>(const E753 *in, E753 *out, TW753 s, size_t n, int do_mul, hipStream_t st)
{
    const unsigned blocks = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(scale_kernel<FMnt4>, dim3(blocks), dim3(256), 0, st, in, out, s, n, do_mul);
}

template <>
void launch_build_steps753<FMnt4>(uint4 *out, uint32_t nrec, const E753 *hi, const TW753 *lo, uint32_t h,
                                 uint32_t log_n, uint32_t log_ns, uint32_t log_r, const TW753 &scale, bool scale0,
                                 hipStream_t st)
{
    const unsigned blocks = (nrec + 1 + 255) / 256;
    hipLaunchKernelGGL(build_steps753_kernel<FMnt4>, dim3(blocks), dim3(256), 0, st, out, nrec, hi, lo, h, log_n, log_ns,
                       log_r, scale, (int)scale0);
}

}  // namespace gsn
