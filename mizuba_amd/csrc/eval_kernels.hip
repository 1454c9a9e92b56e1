// Batched polyjectory evaluation on the GPU: polyjectory::state_eval and
// state_meval (mizuba/polyjectory.cpp:1094-1366).
//
// One lane per evaluation (row r, time k). The lane finds the trajectory step
// holding its time by binary search over the object's step end times, then
// evaluates the 7 channels by Horner over that step's coefficients. Lanes of
// one wave are consecutive evaluations, usually of one object at nearby times,
// so the searches and the 7 * (order + 1) coefficients come from cache, and
// the 56-byte outputs of a wave are contiguous.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

namespace mzh
{

namespace
{

__global__ __launch_bounds__(256) void k_state_eval(PjView pj, const double *__restrict__ tm, int time_mode,
                                                    const uint64_t *__restrict__ sel, uint64_t n_rows,
                                                    uint64_t n_times, double *__restrict__ out,
                                                    unsigned long long *__restrict__ first_bad)
{
    const uint64_t n_eval = n_rows * n_times;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    const uint32_t op1 = pj.order + 1u;
    for (uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x; e < n_eval; e += stride) {
        const uint64_t r = e / n_times, k = e - r * n_times;
        double t;
        switch (time_mode) {
            case MZH_EVAL_TIME_SCALAR:
                t = tm[0];
                break;
            case MZH_EVAL_TIME_PER_OBJECT:
                t = tm[r];
                break;
            case MZH_EVAL_TIME_SHARED:
                t = tm[k];
                break;
            default:
                t = tm[e];
                break;
        }
        const uint64_t obj = sel ? sel[r] : r;
        double *o = out + e * 7u;
        // a non-finite time (polyjectory.cpp:1097-1100), then an object index
        // out of range (:1009-1014): the host reports the first such element
        if (!is_finite_d(t) || obj >= pj.n_objs) {
            atomicMin(first_bad, (unsigned long long)e);
            continue;
        }
        const uint64_t ns = pj.nsteps[obj];
        const double *T = pj.time + pj.time_off[obj];
        if (ns == 0 || t < T[0] || t >= T[ns]) {
            // no data, or outside [t_0, t_ns) (:1108-1120)
#pragma unroll
            for (int ch = 0; ch < 7; ++ch) o[ch] = __builtin_nan("");
            continue;
        }
        // first end time > t among T[1..ns] (upper_bound, :1136-1138); T[ns] > t
        uint64_t lo = 1, hi = ns;
        while (lo < hi) {
            const uint64_t mid = lo + (hi - lo) / 2u;
            if (T[mid] > t) {
                hi = mid;
            } else {
                lo = mid + 1u;
            }
        }
        const uint64_t step = lo - 1u;
        const double h = t - T[step];  // :1141
        // coefficient-major step record [op1][7]; Horner with stride 7
        // (detail/poly_utils.hpp:41-51)
        const double *c = pj.traj + (pj.step0[obj] + step) * (uint64_t)op1 * 7u;
        double v[7];
#pragma unroll
        for (int ch = 0; ch < 7; ++ch) v[ch] = c[(uint64_t)pj.order * 7u + ch];
        for (uint32_t d = pj.order; d-- > 0;) {
#pragma unroll
            for (int ch = 0; ch < 7; ++ch) v[ch] = c[(uint64_t)d * 7u + ch] + v[ch] * h;
        }
#pragma unroll
        for (int ch = 0; ch < 7; ++ch) o[ch] = v[ch];
    }
}

} // namespace

hipError_t launch_state_eval(const PjView &pj, const double *tm, int time_mode, const uint64_t *sel,
                             uint64_t n_rows, uint64_t n_times, double *out, unsigned long long *first_bad,
                             hipStream_t stream)
{
    const uint64_t n_eval = n_rows * n_times;
    if (n_eval == 0) return hipSuccess;
    // enough workgroups to fill the chip several times over; lanes stride beyond
    const uint64_t blocks = std::min<uint64_t>((n_eval + 255u) / 256u, 256u * 64u);
    hipLaunchKernelGGL(k_state_eval, dim3((unsigned)blocks), dim3(256), 0, stream, pj, tm, time_mode, sel, n_rows,
                       n_times, out, first_bad);
    return hipGetLastError();
}

} // namespace mzh
