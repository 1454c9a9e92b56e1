// Batched polyjectory evaluation on the GPU: polyjectory::state_eval and
// state_meval (mizuba/polyjectory.cpp:1094-1366).
//
// One lane per evaluation (row r, time k). The lane finds the trajectory step
// holding its time by binary search over the object's step end times, then
// evaluates the 7 channels by Horner over that step's coefficients. Lanes of
// one wave are consecutive evaluations, usually of one object at nearby times,
// so the searches and the 7 * (order + 1) coefficients come from cache, and
// the 56-byte outputs of a wave are contiguous.
//
// The per-lane loads and the 56-byte-strided stores make the walk bound by
// the texture address and data units, and the chain of dependent loads (the
// object, its offsets, the binary search, the record) by latency. So a wave
// whose 64 evaluations are all of one object takes a faster path: it copies
// the object's step times into its part of LDS and searches there, loads the
// few consecutive step records its times fall in once, with contiguous
// loads, and writes its 3584 bytes of results through LDS as contiguous
// 16-byte stores. The values and the arithmetic are those of the per-lane
// path, so both give the same bits.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

namespace mzh
{

namespace
{

constexpr uint32_t kEvalRecs = 8;     // step records one wave keeps in LDS
constexpr uint32_t kEvalTimes = 512;  // step times one wave keeps in LDS (the results' tile afterwards)

// Order this wave's LDS accesses: its writes become visible to its lanes and
// its earlier reads complete before later writes.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline uint64_t wave_min_u64(uint64_t v)
{
    for (int m = 32; m >= 1; m >>= 1) v = min(v, (uint64_t)__shfl_xor(v, m));
    return v;
}

__device__ inline uint64_t wave_max_u64(uint64_t v)
{
    for (int m = 32; m >= 1; m >>= 1) v = max(v, (uint64_t)__shfl_xor(v, m));
    return v;
}

__device__ inline double eval_time(const double *tm, int time_mode, uint64_t e, uint64_t r, uint64_t k)
{
    switch (time_mode) {
        case MZH_EVAL_TIME_SCALAR:
            return tm[0];
        case MZH_EVAL_TIME_PER_OBJECT:
            return tm[r];
        case MZH_EVAL_TIME_SHARED:
            return tm[k];
        default:
            return tm[e];
    }
}

// The step of T[0..ns] holding t (T[0] <= t < T[ns]): the first end time > t
// among T[1..ns] (upper_bound, polyjectory.cpp:1136-1138), minus one.
__device__ inline uint64_t find_step(const double *T, uint64_t ns, double t)
{
    uint64_t lo = 1, hi = ns;
    while (lo < hi) {
        const uint64_t mid = lo + (hi - lo) / 2u;
        if (T[mid] > t)
            hi = mid;
        else
            lo = mid + 1u;
    }
    return lo - 1u;
}

// Horner over a coefficient-major step record [op1][7] with stride 7
// (detail/poly_utils.hpp:41-51).
__device__ inline void horner7(const double *c, uint32_t order, double h, double v[7])
{
#pragma unroll
    for (int ch = 0; ch < 7; ++ch) v[ch] = c[(uint64_t)order * 7u + ch];
    for (uint32_t d = order; d-- > 0;) {
#pragma unroll
        for (int ch = 0; ch < 7; ++ch) v[ch] = c[(uint64_t)d * 7u + ch] + v[ch] * h;
    }
}

__global__ __launch_bounds__(256) void k_state_eval(PjView pj, const double *__restrict__ tm, int time_mode,
                                                    const uint64_t *__restrict__ sel, uint64_t n_rows,
                                                    uint64_t n_times, double *__restrict__ out,
                                                    unsigned long long *__restrict__ first_bad)
{
    const uint64_t n_eval = n_rows * n_times;
    const uint64_t stride = (uint64_t)gridDim.x * 256u;  // a multiple of the wave size
    const uint32_t op1 = pj.order + 1u;
    const uint32_t rec_len = op1 * 7u;
    const uint32_t lane = threadIdx.x % 64u;
    // per wave: the step times, then the 64 x 7 results; kEvalRecs step
    // records (at most order 8 fits)
    static_assert(kEvalTimes >= 64 * 7, "the results' tile reuses the times' space");
    __shared__ double lds_times[4][kEvalTimes];
    __shared__ double lds_rec[4][kEvalRecs * 63u];
    double *recs = lds_rec[threadIdx.x / 64u];
    double *outs = lds_times[threadIdx.x / 64u];
    double *Ts = outs;
    const bool out16 = (reinterpret_cast<uintptr_t>(out) % 16u) == 0 && rec_len <= 63u;
    uint64_t e = (uint64_t)blockIdx.x * 256u + threadIdx.x;
    for (; e - lane + 64u <= n_eval && out16; e += stride) {
        // a full wave: the fast path when all 64 evaluations are of one valid
        // object at finite times
        const uint64_t r = e / n_times, k = e - r * n_times;
        const uint64_t obj = sel ? sel[r] : r;
        const double t = eval_time(tm, time_mode, e, r, k);
        const uint64_t obj0 = __shfl(obj, 0);
        if (__ballot(obj != obj0 || !is_finite_d(t)) != 0 || obj0 >= pj.n_objs) break;
        const uint64_t ns = pj.nsteps[obj0];
        const double *T = pj.time + pj.time_off[obj0];
        wave_lds_sync();
        if (ns < kEvalTimes) {
            for (uint64_t i = lane; i <= ns; i += 64u) Ts[i] = T[i];
            wave_lds_sync();
            T = Ts;
        }
        const bool inside = ns != 0 && t >= T[0] && t < T[ns];
        const uint64_t step = inside ? find_step(T, ns, t) : ~0ull;
        const uint64_t s_lo = wave_min_u64(step);
        const uint64_t s_hi = wave_max_u64(inside ? step : 0ull);
        double v[7];
        if (s_lo != ~0ull && s_hi - s_lo < kEvalRecs) {
            // the wave's records, consecutive in memory, through LDS
            const double *src = pj.traj + (pj.step0[obj0] + s_lo) * (uint64_t)rec_len;
            const uint32_t n_d = (uint32_t)(s_hi - s_lo + 1u) * rec_len;
            wave_lds_sync();
            for (uint32_t i = lane; i < n_d; i += 64u) recs[i] = src[i];
            wave_lds_sync();
            if (inside) horner7(recs + (step - s_lo) * rec_len, pj.order, t - T[step], v);
        } else if (inside) {
            horner7(pj.traj + (pj.step0[obj0] + step) * (uint64_t)rec_len, pj.order, t - T[step], v);
        }
        if (!inside) {
            // no data, or outside [t_0, t_ns) (:1108-1120)
#pragma unroll
            for (int ch = 0; ch < 7; ++ch) v[ch] = __builtin_nan("");
        }
        // the wave's 64 x 7 results leave as contiguous 16-byte stores
        wave_lds_sync();
#pragma unroll
        for (int ch = 0; ch < 7; ++ch) outs[lane * 7u + ch] = v[ch];
        wave_lds_sync();
        double2 *dst = reinterpret_cast<double2 *>(out + (e - lane) * 7u);
        const double2 *srcv = reinterpret_cast<const double2 *>(outs);
        for (uint32_t i = lane; i < 224u; i += 64u) dst[i] = srcv[i];
    }
    // the remaining evaluations of this lane, one by one
    for (; e < n_eval; e += stride) {
        const uint64_t r = e / n_times, k = e - r * n_times;
        const double t = eval_time(tm, time_mode, e, r, k);
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
        const uint64_t step = find_step(T, ns, t);
        const double h = t - T[step];  // :1141
        double v[7];
        horner7(pj.traj + (pj.step0[obj] + step) * (uint64_t)op1 * 7u, pj.order, h, v);
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
