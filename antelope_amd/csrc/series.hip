// series.hip - deterministic second pass of the per-step k sums.
//
// The step kernels write one row of nv FP64 partial sums per wave (or per
// work unit) and time step: partials[step][row][nv]. This kernel adds the
// rows of each step in a fixed order (strided per-thread sums, then an
// ordered pass over LDS), so a given k block always produces bit-identical
// series no matter what else shares the GPU. It replaces the MPI_Reduce with
// the custom complex-sum op of antelope.cpp:482-512,598-606 for the k points
// owned by one context.
//
// Steps with many rows (large grids: 2^18 waves at 4096^2 k points) are
// summed in two passes, so that one step is not left to a single block: the
// rows split into groups (kGroupRows, or a caller's smaller size for wide rows),
// each group's sum is written over the group's first row, and the second pass
// adds the group sums in order. The order depends only on the row count and
// the group size.
#include "sbe_internal.hpp"

namespace sbe {
namespace {

constexpr int kReduceThreads = 256;
constexpr int32_t kGroupRows = 4096;

// Sum over the rows r0 + r rs, r < nr, of one step's block p, for this
// block's values [vb, vb + vl), into out[0 .. vl). Thread (rlane, vlane) sums
// rows rlane, rlane + rl, ... in order; the row lanes are then added in order.
__device__ __forceinline__ void reduce_rows(const double *p, int64_t r0, int64_t rs, int64_t nr, int nv, double *out,
                                            int32_t *flag = nullptr)
{
    __shared__ double acc[kReduceThreads];
    const int vl = nv < 64 ? nv : 64;          // values per block
    const int rl = kReduceThreads / vl;        // row lanes
    const int tid = threadIdx.x;
    const int vlane = tid % vl, rlane = tid / vl;
    const int vb = blockIdx.y * vl;            // this block's values [vb, vb + vl)
    const int v = vb + vlane;
    double s = 0.0;
    if (rlane < rl && v < nv)
        for (int64_t r = rlane; r < nr; r += rl) s += p[(r0 + r * rs) * nv + v];
    acc[tid] = s;
    __syncthreads();
    if (tid < vl && vb + tid < nv) {
        double t = 0.0;
        for (int r = 0; r < rl; ++r) t += acc[r * vl + tid];
        out[tid] = t;
        if (flag && !isfinite(t)) *flag = 1;  // a NaN or inf anywhere in the step reaches its sum
    }
}

// one block per (step, block of 64 values): the rows 0, rs, 2 rs, ... (nr)
__global__ __launch_bounds__(kReduceThreads) void series_reduce_kernel(const double *__restrict__ partials,
                                                                       int32_t nrows, int32_t nr, int32_t rs,
                                                                       int nv, double *__restrict__ series,
                                                                       int32_t *flag)
{
    const int64_t step = blockIdx.x;
    const int vl = nv < 64 ? nv : 64;
    reduce_rows(partials + step * static_cast<int64_t>(nrows) * nv, 0, rs, nr, nv, series + step * nv + blockIdx.y * vl,
                flag);
}

// first pass: group g = blockIdx.z of a step sums its rows [g G, (g + 1) G)
// and writes the sum over row g G. Only this block reads or writes those rows;
// its reads are done (barrier in reduce_rows) before the write.
__global__ __launch_bounds__(kReduceThreads) void series_group_kernel(double *__restrict__ partials, int32_t nrows,
                                                                      int nv, int32_t grows)
{
    const int64_t step = blockIdx.x;
    const int64_t g0 = static_cast<int64_t>(blockIdx.z) * grows;
    const int64_t nr = nrows - g0 < grows ? nrows - g0 : grows;
    double *p = partials + step * static_cast<int64_t>(nrows) * nv;
    __shared__ double sums[64];
    reduce_rows(p, g0, 1, nr, nv, sums);
    __syncthreads();
    const int vl = nv < 64 ? nv : 64, vb = blockIdx.y * vl;
    const int tid = static_cast<int>(threadIdx.x);
    if (tid < vl && vb + tid < nv) p[g0 * nv + vb + tid] = sums[tid];
}

// many k blocks (the solves of a batch) in one launch: block (x, y) works
// for the job whose steps [first, first + nsteps) hold x, and sums that
// step exactly as series_reduce_kernel does for the job alone (one pass)
__global__ __launch_bounds__(kReduceThreads) void series_reduce_multi_kernel(const SeriesJob *__restrict__ jobs,
                                                                             int njobs, int nv)
{
    const int64_t g = blockIdx.x;
    int lo = 0, hi = njobs - 1;  // last job with first <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (jobs[mid].first <= g) lo = mid;
        else hi = mid - 1;
    }
    const SeriesJob &j = jobs[lo];
    const int64_t step = g - j.first;
    if (step >= j.nsteps) return;
    const int vl = nv < 64 ? nv : 64;
    reduce_rows(j.partials + step * static_cast<int64_t>(j.nrows) * nv, 0, 1, j.nrows, nv,
                j.series + step * nv + blockIdx.y * vl, j.flag);
}

}  // namespace

bool series_multi_ok(int32_t nrows) { return nrows <= 2 * kGroupRows; }

hipError_t launch_series_reduce_multi(const SeriesJob *d_jobs, int njobs, int64_t total_steps, int nv, hipStream_t s)
{
    if (njobs <= 0 || total_steps <= 0) return hipSuccess;
    const int vl = nv < 64 ? nv : 64;
    const unsigned vblocks = static_cast<unsigned>((nv + vl - 1) / vl);
    series_reduce_multi_kernel<<<dim3(static_cast<unsigned>(total_steps), vblocks), kReduceThreads, 0, s>>>(
        d_jobs, njobs, nv);
    return hipGetLastError();
}

hipError_t launch_series_reduce(double *partials, int64_t nsteps, int32_t nrows, int nv, double *series,
                                hipStream_t s, int32_t *flag, int32_t group_rows)
{
    const int32_t G = group_rows > 0 ? group_rows : kGroupRows;
    if (nsteps <= 0) return hipSuccess;
    // one block per (step, block of 64 values): the summation order per value
    // (row lanes, then the lanes in order) does not depend on the grid
    const int vl = nv < 64 ? nv : 64;
    const unsigned vblocks = static_cast<unsigned>((nv + vl - 1) / vl);
    if (nrows > 2 * G) {
        const int32_t groups = (nrows + G - 1) / G;
        series_group_kernel<<<dim3(static_cast<unsigned>(nsteps), vblocks, static_cast<unsigned>(groups)),
                              kReduceThreads, 0, s>>>(partials, nrows, nv, G);
        series_reduce_kernel<<<dim3(static_cast<unsigned>(nsteps), vblocks), kReduceThreads, 0, s>>>(
            partials, nrows, groups, G, nv, series, flag);
    } else {
        series_reduce_kernel<<<dim3(static_cast<unsigned>(nsteps), vblocks), kReduceThreads, 0, s>>>(
            partials, nrows, nrows, 1, nv, series, flag);
    }
    return hipGetLastError();
}

}  // namespace sbe
