// series.hip - deterministic second pass of the per-step k sums.
//
// The step kernels write one row of nv FP64 partial sums per wave (or per
// work unit) and time step: partials[step][row][nv]. This kernel adds the
// rows of each step in a fixed order (strided per-thread sums, then an
// ordered pass over LDS), so a given k block always produces bit-identical
// series no matter what else shares the GPU. It replaces the MPI_Reduce with
// the custom complex-sum op of antelope.cpp:482-512,598-606 for the k points
// owned by one context.
#include "sbe_internal.hpp"

namespace sbe {
namespace {

constexpr int kReduceThreads = 256;

__global__ __launch_bounds__(kReduceThreads) void series_reduce_kernel(const double *__restrict__ partials,
                                                                       int32_t nrows, int nv,
                                                                       double *__restrict__ series)
{
    __shared__ double acc[kReduceThreads];
    const int64_t step = blockIdx.x;
    const double *p = partials + step * static_cast<int64_t>(nrows) * nv;
    const int vl = nv < 64 ? nv : 64;          // values per block
    const int rl = kReduceThreads / vl;        // row lanes
    const int tid = threadIdx.x;
    const int vlane = tid % vl, rlane = tid / vl;
    const int vb = blockIdx.y * vl;            // this block's values [vb, vb + vl)
    const int v = vb + vlane;
    double s = 0.0;
    if (rlane < rl && v < nv)
        for (int r = rlane; r < nrows; r += rl) s += p[static_cast<int64_t>(r) * nv + v];
    acc[tid] = s;
    __syncthreads();
    if (tid < vl && vb + tid < nv) {
        double t = 0.0;
        for (int r = 0; r < rl; ++r) t += acc[r * vl + tid];
        series[step * nv + vb + tid] = t;
    }
}

}  // namespace

hipError_t launch_series_reduce(const double *partials, int64_t nsteps, int32_t nrows, int nv, double *series,
                                hipStream_t s)
{
    if (nsteps <= 0) return hipSuccess;
    // one block per (step, block of 64 values): the summation order per value
    // (row lanes, then the lanes in order) does not depend on the grid
    const int vl = nv < 64 ? nv : 64;
    const dim3 grid(static_cast<unsigned>(nsteps), static_cast<unsigned>((nv + vl - 1) / vl));
    series_reduce_kernel<<<grid, kReduceThreads, 0, s>>>(partials, nrows, nv, series);
    return hipGetLastError();
}

}  // namespace sbe
