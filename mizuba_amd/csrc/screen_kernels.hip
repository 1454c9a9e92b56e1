// Screening kernels for gfx950: AABBs, global box, Morton codes, segmented
// radix sort, gather, BVH build and broad phase.
//
// Batching: every kernel covers all conjunction steps of a problem in one
// launch (grid.y or grid.x = step), so a full-day screen is a fixed handful
// of launches instead of per-step loops.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

namespace mzh
{

namespace
{

constexpr BinomTable kBinom = make_binom_table();
__constant__ BinomTable c_binom = kBinom;

// The problem of a block or work item: the largest p with pre[p] <= x (pre is
// an exclusive prefix of np + 1 entries).
__device__ inline uint32_t find_prob(const uint32_t *__restrict__ pre, uint32_t np, uint32_t x)
{
    uint32_t lo = 0, n = np;
    while (n > 1) {
        const uint32_t h = n / 2;
        if (pre[lo + h] <= x) {
            lo += h;
            n -= h;
        } else {
            n = h;
        }
    }
    return lo;
}

// ---------------------------------------------------------------------------
// upload-time transpose
// ---------------------------------------------------------------------------

__global__ void k_transpose_cf4(const double *__restrict__ traj, double *__restrict__ cf4, float *__restrict__ cf4f,
                                uint64_t n_steps_total, uint32_t op1)
{
    const uint64_t idx = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t per_step = 4ull * op1;
    if (idx >= n_steps_total * per_step) return;
    const uint64_t s = idx / per_step;
    const uint32_t r = (uint32_t)(idx - s * per_step);
    const uint32_t ch = r / op1, k = r - ch * op1;
    const uint32_t src_ch = (ch == 3) ? 6u : ch;
    const double v = traj[(s * op1 + k) * 7u + src_ch];
    cf4[idx] = v;
    cf4f[idx] = (float)v;  // the rounding the FP32 enclosure applies to each coefficient
}

// Bounds of one trajectory step for the mixed-precision pre-screen (one lane
// per step of the concatenated steps), each rounded up to float:
// sbnd = {D3, Jv, Jd, Ja}. Over the step, tau in [0, H],
//   |d^a x / dtau^a| <= B_a = sum_{k >= a} k! / (k - a)! |c_k| H^(k - a)
// per channel (Horner in H on positive terms, so the rounding is at most
// 2 order ulps, covered by the factor 1 + 2^-40), and D_a = |(B_a,x, B_a,y,
// B_a,z)| bounds the a-th derivative of the position; D3 = D_3 (0 for order
// 2). Jv, Jd and Ja: the jumps of position, velocity and acceleration from
// the end of the object's previous step to the start of this one (0 for its
// first step), from FP64 evaluations bounded with 1e-14 of their magnitude
// sums. Two passes for the polyjectory's scalars lam and A (as float bits),
// with D_a <= A lam^a for every step and a = 1..order: the first (AMP false)
// raises lam to each step's own rate max_a (D_a / D_1)^(1 / (a - 1)); the
// second (AMP true) raises A to each step's max_a D_a / lam^a, which holds
// whatever lam is.
template <int ORDER, bool AMP>
__global__ void k_step_bounds(const double *__restrict__ cf4, const double *__restrict__ time,
                              const uint64_t *__restrict__ step0, const uint64_t *__restrict__ nsteps,
                              const uint64_t *__restrict__ time_off, uint32_t n_objs, uint64_t n_steps_total,
                              float4 *__restrict__ sbnd, uint32_t *__restrict__ scal)
{
    constexpr int OP1 = ORDER + 1;
    constexpr double up = 1.0 + 0x1p-40;
    const uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_steps_total) return;
    // the object owning step s: the last o with step0[o] <= s (objects without
    // steps share their successor's step0 and come before it)
    uint32_t lo = 0, n = n_objs;
    while (n > 1) {
        const uint32_t h = n / 2;
        if (step0[lo + h] <= s) {
            lo += h;
            n -= h;
        } else {
            n = h;
        }
    }
    const uint64_t k = s - step0[lo];
    if (k >= nsteps[lo]) return;  // not reached: every step belongs to an object
    const double *t = time + time_off[lo];
    const double H = fabs(t[k + 1] - t[k]);
    double acc[ORDER];
#pragma unroll
    for (int a = 0; a < ORDER; ++a) acc[a] = 0.0;
    double jv2 = 0.0, jd2 = 0.0, ja2 = 0.0;
    const double Hp = k > 0 ? fabs(t[k] - t[k - 1]) : 0.0;
#pragma unroll
    for (int ch = 0; ch < 3; ++ch) {
        const double *c = cf4 + (s * 4u + (uint64_t)ch) * OP1;
        double fa = 1.0;  // a!
#pragma unroll
        for (int a = 1; a <= ORDER; ++a) {
            fa *= (double)a;
            // k! / (k - a)! = C(k, a) a!, for k = ORDER down to a, in Horner order
            double b = 0.0;
#pragma unroll
            for (int kk = ORDER; kk >= a; --kk) b = b * H + (c_binom.v[kk][a] * fa) * fabs(c[kk]);
            acc[a - 1] += b * b;
        }
        if (!AMP && k > 0) {
            // the previous step's position, velocity and half acceleration at its end
            const double *cp = cf4 + ((s - 1u) * 4u + (uint64_t)ch) * OP1;
            double v = cp[ORDER], dv = 0.0, hv = 0.0, mv = fabs(cp[ORDER]), md = 0.0, mh = 0.0;
#pragma unroll
            for (int kk = ORDER - 1; kk >= 0; --kk) {
                hv = hv * Hp + dv;
                mh = mh * Hp + md;
                dv = dv * Hp + v;
                md = md * Hp + mv;
                v = v * Hp + cp[kk];
                mv = mv * Hp + fabs(cp[kk]);
            }
            const double ev = fabs(c[0] - v) + 1e-14 * (mv + fabs(c[0]));
            const double ed = fabs(c[1] - dv) + 1e-14 * (md + fabs(c[1]));
            const double ea = 2.0 * (fabs(c[2] - hv) + 1e-14 * (mh + fabs(c[2])));
            jv2 += ev * ev;
            jd2 += ed * ed;
            ja2 += ea * ea;
        }
    }
    if constexpr (!AMP) {
        const double d1 = sqrt(acc[0]);
        double rate = 0.0;
        if (d1 > 0.0) {
#pragma unroll
            for (int a = 2; a <= ORDER; ++a) {
                const double r = sqrt(acc[a - 1]) / d1;
                if (r > 0.0) rate = fmax(rate, exp(log(r) / (double)(a - 1)));
            }
        }
        // non-negative floats order as their bit patterns
        if (rate > 0.0) atomicMax(&scal[0], __float_as_uint(next_up_f((float)rate)));
        const double d3 = ORDER >= 3 ? sqrt(acc[ORDER >= 3 ? 2 : 0]) : 0.0;
        sbnd[s] = make_float4(next_up_f((float)(d3 * up)), next_up_f((float)(sqrt(jv2) * up)),
                              next_up_f((float)(sqrt(jd2) * up)), next_up_f((float)(sqrt(ja2) * up)));
    } else {
        const double lam = (double)__uint_as_float(scal[0]);
        double amp = 0.0, la = 1.0;
#pragma unroll
        for (int a = 1; a <= ORDER; ++a) {
            la *= lam;
            const double d = sqrt(acc[a - 1]) * up;
            if (d > 0.0) amp = fmax(amp, la > 0.0 ? d / la : __builtin_inf());
        }
        if (amp > 0.0) atomicMax(&scal[1], __float_as_uint(next_up_f((float)(amp * (1.0 + 1e-12)))));
    }
}

template <int ORDER>
hipError_t dispatch_step_bounds(uint32_t order, const double *cf4, const double *time, const uint64_t *step0,
                                const uint64_t *nsteps, const uint64_t *time_off, uint32_t n_objs,
                                uint64_t n_steps_total, float4 *sbnd, uint32_t *scal, hipStream_t stream)
{
    if (order == (uint32_t)ORDER) {
        const dim3 g((unsigned)((n_steps_total + 255) / 256));
        hipLaunchKernelGGL((k_step_bounds<ORDER, false>), g, dim3(256), 0, stream, cf4, time, step0, nsteps, time_off,
                           n_objs, n_steps_total, sbnd, scal);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL((k_step_bounds<ORDER, true>), g, dim3(256), 0, stream, cf4, time, step0, nsteps, time_off,
                           n_objs, n_steps_total, sbnd, scal);
        return hipGetLastError();
    }
    if constexpr (ORDER < kMaxOrder) {
        return dispatch_step_bounds<ORDER + 1>(order, cf4, time, step0, nsteps, time_off, n_objs, n_steps_total, sbnd,
                                               scal, stream);
    } else {
        return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// stage 1: AABBs (detect_conjunctions_aabbs.cpp:64-252, conjunctions_jit.cpp:179-298)
// ---------------------------------------------------------------------------

// Conjunction steps handled by one thread of the AABB kernel: the thread walks
// its object's time array once across the group instead of searching it per
// step, and consecutive steps reuse the trajectory pieces it has just read.
constexpr uint32_t kAabbGroup = kAabbStepGroup;

// Stage 1: per (object, conjunction step) boxes (detect_conjunctions_aabbs.cpp:64-252)
// and the per-step global box (:314-397). One thread per object and group of
// kAabbGroup consecutive steps; the 256 lanes of a block are 256 consecutive
// objects, so each step's boxes are written coalesced and reduced per block.
//
// F32: the mixed-precision problems (the FP64 ones return), whose enclosures
// are computed in FP32 from the float copy of the coefficients (cf4f): half
// the bytes read and half the LDS of the piece cache. The others: the FP64
// problems.
//
// The mixed-precision problems also get each (object, step)'s pre-screen
// model (nprec, read by the narrow phase's np_pair_far, 32 bytes: p0 and v0 in
// float, a0 and R in bfloat16): over the step [cb, ce] with midpoint tm and
// half-width w, the position p0, velocity v0 and acceleration a0 at tm, and
// R >= |r(t) - p0 - v0 s - a0 s^2 / 2| for every t in [cb, ce], s = t - tm:
//  - on the piece holding tm, Taylor: (D3 / 6) |s|^3;
//  - across each piece boundary tb inside the step, the same with D3 the
//    largest of the pieces', plus the jumps there, Jv + Jd |t - tb| +
//    Ja |t - tb|^2 / 2 (|t - tb| <= 2w; each jump of the value or a
//    derivative carries on as a Taylor polynomial past the boundary);
//  - p0, v0 evaluated in FP32 from the float coefficients, within
//    (4 order + 8) 6.1e-8 of their magnitude sums of the FP64 polynomial's
//    values (coefficient rounding, FP32 Horner and the rounding of
//    tm - t_k; stored as computed, the acceleration rounded to bfloat16 with
//    the rounding added, R rounded up).
// Objects whose data do not cover the whole step get R = infinity (never
// screened out).
// Two waves per SIMD for the mixed-precision kernels up to order 8, whose
// registers then just fit (20 bytes spilled at order 8); the others as their
// registers allow.
template <int ORDER, bool F32>
constexpr unsigned kAabbWaves = (F32 && ORDER <= 8) ? 2u : 1u;

template <int ORDER, bool F32>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kAabbWaves<ORDER, F32>))) void k_aabbs(const Prob *__restrict__ P, uint32_t np,
                                               const uint32_t *__restrict__ blk_pre, uint32_t blk_begin)
{
    using CT = typename std::conditional<F32, float, double>::type;
    const uint32_t b = blk_begin + blockIdx.x;
    const uint32_t pi = find_prob(blk_pre, np, b);
    const Prob &q = P[pi];
    if (q.pj.order != (uint32_t)ORDER) return;
    if ((q.mode == MZH_MODE_MIXED) != F32) return;
    const uint32_t lb = b - blk_pre[pi];
    const PjView pj = q.pj;
    const double *cd_begin = q.cd_begin, *cd_end = q.cd_end;
    const uint32_t n_steps = q.C;
    const double radius = q.radius;
    const int mode = q.mode;
    float *__restrict__ aabbs = q.aabbs;
    uint2 *__restrict__ kbke = q.kbke;
    uint32_t *__restrict__ gbox_ord = q.gbox;
    DevError *err = q.err;
    const uint32_t N = pj.n_objs;
    const uint32_t bpo = (N + 255u) / 256u;
    const uint32_t g = lb / bpo;
    const uint32_t o = (lb - g * bpo) * 256u + threadIdx.x;
    const uint32_t c0 = g * kAabbGroup, c1 = min(n_steps, c0 + kAabbGroup);
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    __shared__ uint32_t s_l[4][4], s_u[4][4];
    // The lane's current trajectory piece (4 channels x (ORDER + 1)
    // coefficients) is kept in LDS, lane-minor: a piece spans several
    // conjunction steps and is read from HBM once. Up to order 8 (72 KB).
    constexpr int kCf = 4 * (ORDER + 1);
    constexpr bool kCache = kCf * 256 * sizeof(CT) <= 80 * 1024;
    __shared__ CT s_cf[kCache ? kCf : 1][256];
    uint64_t k_cached = ~0ull;

    uint64_t ns = 0, kb = 0, sbase = 0;
    const double *t = nullptr;
    const CT *cfb = nullptr;
    const double *cfb64 = nullptr;
    float4 sb_cached = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // the step bounds of the cached piece (mixed)
    // error of the FP32 model at tm off the FP64 polynomial, per unit of the
    // magnitude sums (value: sum |c_k| |tau|^k; velocity: sum k |c_k|
    // |tau|^(k-1); half acceleration: sum C(k, 2) |c_k| |tau|^(k-2)), in units
    // of 2^-24: the coefficients' float rounding (1), the FP32 Horner
    // evaluation (2 order) and tau = tm - t_k rounded to float (|p'| |tau|
    // 2^-24 <= order 2^-24 of the value's sum, and likewise for the
    // derivatives); with room to spare
    constexpr double kModelErr = (4.0 * ORDER + 8.0) * 6.1e-8;
    if (o < N) {
        ns = pj.nsteps[o];
        sbase = pj.step0[o];
        if (ns > 0) {
            t = pj.time + pj.time_off[o];
            cfb64 = pj.cf4 + (pj.step0[o] * 4u) * (ORDER + 1);
            if constexpr (F32) {
                cfb = pj.cf4f + (pj.step0[o] * 4u) * (ORDER + 1);
            } else {
                cfb = cfb64;
            }
            // upper_bound(t + 1, t + ns + 1, cd_begin[c0]): first piece ending after the group's start
            const double cb = cd_begin[c0];
            uint64_t lo = 0, n = ns;
            while (n > 0) {
                const uint64_t h = n / 2;
                if (!(cb < t[1 + lo + h])) {
                    lo += h + 1;
                    n -= h + 1;
                } else {
                    n = h;
                }
            }
            kb = lo;
        }
    }
    for (uint32_t c = c0; c < c1; ++c) {
        float lb4[4] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
        float ub4[4] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
        if (o < N) {
            const double cb = cd_begin[c], ce = cd_end[c];
            uint2 range = make_uint2(0u, 0u);
            // the pre-screen model (mixed precision)
            const double tm = cb + 0.5 * (ce - cb), mw = fmax(tm - cb, ce - tm);
            const bool full = F32 && ns > 0 && t[0] <= cb && t[ns] >= ce;
            float d3m = 0.0f, esum = 0.0f, aerr = 0.0f;
            double jsum = 0.0;
            uint32_t az = 0u;  // the acceleration's z, as bfloat16
            float4 *dm = F32 ? q.nprec + ((size_t)c * N + o) * 2u : nullptr;
            if (ns > 0 && !(t[0] >= ce || t[ns] <= cb)) {
                // kb: upper_bound of cb (monotone in c); ke: lower_bound of ce from kb, plus one
                while (kb < ns && !(cb < t[1 + kb])) ++kb;
                uint64_t ke = kb;
                while (ke < ns && t[1 + ke] < ce) ++ke;
                ke += (ke != ns) ? 1 : 0;
                range = make_uint2((uint32_t)kb, (uint32_t)ke);
                for (uint64_t k = kb; k < ke; ++k) {
                    const double ss = t[k];
                    const double ev_lb = cb > ss ? cb : ss;
                    const double ev_ub = ce < t[k + 1] ? ce : t[k + 1];
                    const double h_lb = ev_lb - ss, h_ub = ev_ub - ss;
                    const CT *cf = cfb + k * (uint64_t)kCf;
                    constexpr int kStride = kCache ? 256 : 1;
                    if constexpr (kCache) {
                        if (k != k_cached) {
#pragma unroll
                            for (int j = 0; j < kCf; ++j) s_cf[j][threadIdx.x] = cf[j];
                            if constexpr (F32) sb_cached = pj.sbnd[sbase + k];
                            k_cached = k;
                        }
                        cf = &s_cf[0][threadIdx.x];
                    } else if constexpr (F32) {
                        sb_cached = pj.sbnd[sbase + k];
                    }
#pragma unroll
                    for (int d = 0; d < 4; ++d) {
                        float fl, fu;
                        const CT *cfd = cf + d * (ORDER + 1) * kStride;
                        if constexpr (!F32) {
                            double emin, emax;
                            cs_enclosure_fp64<ORDER, kStride>(cfd, h_lb, h_ub, emin, emax, c_binom);
                            emin = emin - radius;
                            emax = emax + radius;
                            fl = next_down_f((float)emin);
                            fu = next_up_f((float)emax);
                        } else {
                            float emin, emax;
                            cs_enclosure_fp32<ORDER, kStride>(cfd, h_lb, h_ub, emin, emax, c_binom);
                            double dmin = (double)emin, dmax = (double)emax;
                            if (!is_finite_f(emin) || !is_finite_f(emax)) {
                                // coefficients or intermediates outside the float
                                // range (ill-conditioned short steps): use the FP64
                                // enclosure for this piece, from the FP64 copy
                                cs_enclosure_fp64<ORDER, 1>(cfb64 + k * (uint64_t)kCf + d * (ORDER + 1), h_lb, h_ub,
                                                            dmin, dmax, c_binom);
                            }
                            // dmin <= FP64 enclosure minimum (as reals); the same
                            // monotone double subtraction and outward float
                            // conversion then keep this box around the FP64 box.
                            fl = next_down_f((float)(dmin - radius));
                            fu = next_up_f((float)(dmax + radius));
                        }
                        if (!is_finite_f(fl)) {
                            report_error(err, kErrAabbLb, o, 0, (double)fl);
                            fl = -3.0e38f;
                        }
                        if (!is_finite_f(fu)) {
                            report_error(err, kErrAabbUb, o, 0, (double)fu);
                            fu = 3.0e38f;
                        }
                        lb4[d] = lb4[d] < fl ? lb4[d] : fl;
                        ub4[d] = ub4[d] > fu ? ub4[d] : fu;
                    }
                    if constexpr (F32) {
                        if (full) {
                            const float4 sb = sb_cached;
                            d3m = fmaxf(d3m, sb.x);
                            if (k > kb) {
                                const double w2 = 2.0 * mw;  // |t - tb| <= 2 w
                                jsum += (double)sb.y + w2 * (double)sb.z + 0.5 * w2 * w2 * (double)sb.w;
                            }
                            if (ss <= tm && tm < t[k + 1]) {
                                // position, velocity, half acceleration and
                                // their magnitude sums at tm, in FP32 from
                                // the piece's float coefficients
                                const float tau = (float)(tm - ss), at = fabsf(tau), fw = (float)mw;
                                float p0[3], v0[3], a0[3];
#pragma unroll
                                for (int ch = 0; ch < 3; ++ch) {
                                    const CT *cc = cf + ch * (ORDER + 1) * kStride;
                                    float pv = cc[ORDER * kStride], dv = 0.0f, hv = 0.0f;
                                    float mp = fabsf(pv), md = 0.0f, mh = 0.0f;
#pragma unroll
                                    for (int kk = ORDER - 1; kk >= 0; --kk) {
                                        const float x = cc[kk * kStride];
                                        hv = hv * tau + dv;
                                        mh = mh * at + md;
                                        dv = dv * tau + pv;
                                        md = md * at + mp;
                                        pv = pv * tau + x;
                                        mp = mp * at + fabsf(x);
                                    }
                                    p0[ch] = pv;
                                    v0[ch] = dv;
                                    a0[ch] = 2.0f * hv;
                                    // value, velocity x w and acceleration x w^2 / 2
                                    esum += mp + md * fw + mh * fw * fw;
                                }
                                // stored at once (R follows the loop); the
                                // acceleration as bfloat16, its rounding (exact
                                // in float) part of the model's error
                                const uint32_t ax = bf16_rn(a0[0]), ay = bf16_rn(a0[1]);
                                az = bf16_rn(a0[2]);
                                const float ex = a0[0] - bf16_f(ax), ey = a0[1] - bf16_f(ay), ez = a0[2] - bf16_f(az);
                                aerr = fabsf(ex) + fabsf(ey) + fabsf(ez);
                                dm[0] = make_float4(p0[0], p0[1], p0[2], v0[0]);
                                dm[1] = make_float4(v0[1], v0[2], __uint_as_float(ax | (ay << 16)), __uint_as_float(az));
                            }
                        }
                    }
                }
            }
            if constexpr (F32) {
                if (full) {
                    // the model's error: the channels' sum bounds their norm,
                    // and 1e-3 covers the float sums' rounding
                    const double R = (1.0 / 6.0) * (double)d3m * mw * mw * mw + jsum + kModelErr * 1.001 * (double)esum
                                     + 0.5 * mw * mw * 1.001 * (double)aerr;
                    dm[1].w = __uint_as_float(az | (bf16_up(next_up_f((float)(R * (1.0 + 1e-9)))) << 16));
                } else {
                    // never screened out (R infinite)
                    dm[0] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    dm[1] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(0x7f800000u));
                }
            }
            float4 *dst = reinterpret_cast<float4 *>(aabbs + ((size_t)c * (N + 1) + o) * 8u);
            dst[0] = make_float4(lb4[0], lb4[1], lb4[2], lb4[3]);
            dst[1] = make_float4(ub4[0], ub4[1], ub4[2], ub4[3]);
            kbke[(size_t)c * N + o] = range;
        }
        // block-wide min/max of this step's boxes -> one atomic per block and component
        uint32_t ol[4], ou[4];
#pragma unroll
        for (int d = 0; d < 4; ++d) {
            ol[d] = f2ord(lb4[d]);
            ou[d] = f2ord(ub4[d]);
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const uint32_t x = __shfl_xor(ol[d], off, 64), y = __shfl_xor(ou[d], off, 64);
                ol[d] = x < ol[d] ? x : ol[d];
                ou[d] = y > ou[d] ? y : ou[d];
            }
        }
        __syncthreads();
        if (lane == 0) {
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                s_l[w][d] = ol[d];
                s_u[w][d] = ou[d];
            }
        }
        __syncthreads();
        if (threadIdx.x < 4) {
            const int d = threadIdx.x;
            uint32_t l = s_l[0][d], u = s_u[0][d];
            for (int k = 1; k < 4; ++k) {
                l = s_l[k][d] < l ? s_l[k][d] : l;
                u = s_u[k][d] > u ? s_u[k][d] : u;
            }
            atomicMin(&gbox_ord[c * 8u + d], l);
            atomicMax(&gbox_ord[c * 8u + 4u + d], u);
        }
    }
}

// one thread per conjunction step of the batch
__global__ void k_gbox_finalize(const Prob *__restrict__ P, uint32_t np, const uint32_t *__restrict__ step_pre,
                                uint32_t n_all)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_all) return;
    const Prob &q = P[find_prob(step_pre, np, s)];
    const uint32_t c = s - q.step0, N = q.N;
    const uint32_t *gbox_ord = q.gbox;
    float *row = q.aabbs + ((size_t)c * (N + 1) + N) * 8u;
    for (int d = 0; d < 4; ++d) {
        const float l = ord2f(gbox_ord[c * 8u + d]), u = ord2f(gbox_ord[c * 8u + 4u + d]);
        if (is_finite_f(l) && !is_finite_f(u - l)) report_error(q.err, kErrGlobalBox, c, 0, 0.0);
        row[d] = l;
        row[4 + d] = u;
    }
}

// The per-step build words of the walk tree in the BVH scratch
// (lbvh_step_words): the parent of each node whose parent spans blocks, the
// frontier list (nodes finished inside their block whose parent spans
// blocks), its length, and the spanning nodes' list and count. A spanning
// node's arrival counter and split word live in its record's box words until
// the climb writes its box there.
struct LbvhWords {
    uint32_t *par, *nfront, *nspan, *span, *front;
    uint32_t span_cap, front_cap;
};

__device__ inline LbvhWords lbvh_words(const Prob &q, uint32_t c)
{
    uint32_t *h = reinterpret_cast<uint32_t *>(q.bvh_hdr + 0) + (size_t)c * q.lbvh_stride;
    const uint32_t np = 2u * q.N;
    const uint32_t sc = (uint32_t)lbvh_span_cap(q.N);
    return LbvhWords{h, h + np, h + np + 1u, h + np + 2u, h + np + 2u + sc, sc, 2u * sc + 2u};
}

// One thread per conjunction step of the batch: the step's global box starts
// empty; the first step of a problem also clears the problem's counters,
// flags and error word, and thread 0 the batch's work counters.
__global__ void k_batch_init(const Prob *__restrict__ P, uint32_t np, const uint32_t *__restrict__ step_pre,
                             uint32_t n_all, BatchCounters *bc)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) {
        bc->next_sort = 0;
        bc->next_bvh = 0;
        bc->next_bvh_hbm = 0;
    }
    if (s >= n_all) return;
    const Prob &q = P[find_prob(step_pre, np, s)];
    const uint32_t c = s - q.step0;
#pragma unroll
    for (int d = 0; d < 8; ++d) q.gbox[c * 8u + d] = (d < 4) ? f2ord(__builtin_inff()) : f2ord(-__builtin_inff());
    q.tree_size[c] = 0u;  // counted by the walk-tree build
    *lbvh_words(q, c).nfront = 0u;
    *lbvh_words(q, c).nspan = 0u;
    for (uint32_t k = c; k < kReportWords; k += q.C) q.np_report[k] = 0ull;
    for (uint32_t k = c; k < kPoolShards; k += q.C) q.ovf_count[k * kPoolStride] = 0u;
    if (c == 0) {
        q.err->code = 0;
        q.err->i = q.err->j = 0u;
        q.err->value = 0.0;
        *q.flags = 0u;
        for (int k = 0; k < 4; ++k) q.counts[k] = 0ull;
    }
}

// ---------------------------------------------------------------------------
// stage 2: Morton codes (detect_conjunctions_morton.cpp:70-175)
// ---------------------------------------------------------------------------

__device__ inline uint64_t disc_coord(float x, float mn, float mx)
{
    const float isize = mx - mn;
    float rx = (x - mn) / isize;
    rx = rx >= 0.f ? rx : 0.f;
    rx *= 65536.f;
    // values beyond 2^16 are clamped below; keep the conversion in range
    rx = rx < 131072.f ? rx : 131072.f;
    const uint64_t r = (uint64_t)rx;
    return r < 65535ull ? r : 65535ull;
}

// Morton code of a box (detect_conjunctions_morton.cpp:70-175): centre lb/2 +
// ub/2 of each dimension quantized against the global box, bit b of dimension
// d at bit 4b + d; all ones when the box is empty (no data).
__device__ inline uint64_t morton_code(const float4 &blo, const float4 &bhi, const float4 &glo, const float4 &ghi)
{
    if (!is_finite_f(blo.x)) return ~0ull;
    const uint64_t q0 = disc_coord(blo.x / 2 + bhi.x / 2, glo.x, ghi.x);
    const uint64_t q1 = disc_coord(blo.y / 2 + bhi.y / 2, glo.y, ghi.y);
    const uint64_t q2 = disc_coord(blo.z / 2 + bhi.z / 2, glo.z, ghi.z);
    const uint64_t q3 = disc_coord(blo.w / 2 + bhi.w / 2, glo.w, ghi.w);
    return spread4(q0) | (spread4(q1) << 1) | (spread4(q2) << 2) | (spread4(q3) << 3);
}

// ---------------------------------------------------------------------------
// stage 2b: segmented stable LSD radix sort, one workgroup per step
// (replaces tbb::parallel_sort of detect_conjunctions_morton.cpp:179-196).
// Keys are the 64-bit codes, values the object indices. Eight 8-bit passes
// with wave64 ballot matching give a sort that is stable in the object index;
// a final stable partition moves objects without data (code 2^64-1, infinite
// box) behind finite objects that happen to share the all-ones code.
// ---------------------------------------------------------------------------

constexpr int kSortBlock = 1024;
constexpr int kSortItems = 4;
constexpr int kSortWaves = kSortBlock / 64;

__device__ inline uint32_t block_excl_scan_1024(uint32_t v, uint32_t *s_w, uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    if (w == 0) {
        uint32_t x = lane < (blockDim.x >> 6) ? s_w[lane] : 0u;
        const uint32_t xi = wave_inclusive_scan(x);
        if (lane < (blockDim.x >> 6)) s_w[lane] = xi - x;
        if (lane == 63) s_w[64] = xi;
    }
    __syncthreads();
    const uint32_t r = s_w[w] + inc - v;
    total = s_w[64];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kSortBlock) void k_sort(const Prob *__restrict__ P, uint32_t np,
                                                       const uint32_t *__restrict__ step_pre)
{
    const Prob &q = P[find_prob(step_pre, np, blockIdx.x)];
    const uint32_t c = blockIdx.x - q.step0;
    const uint32_t N = q.N;
    const uint64_t *__restrict__ mcodes = q.mcodes;
    const float *__restrict__ aabbs = q.aabbs;
    uint64_t *__restrict__ keys_tmp = q.keys_tmp;
    uint32_t *__restrict__ vals_tmp = q.vals_tmp;
    uint64_t *__restrict__ keys_out = q.srt_mcodes;
    uint32_t *__restrict__ vals_out = q.srt_idx;
    uint32_t *__restrict__ n_eff = q.n_eff;
    if (!q.sort_todo[c]) return;  // sorted by the bucket sort
    const size_t base = (size_t)c * N;
    const uint64_t *kin = mcodes + base;
    uint64_t *kt = keys_tmp + base;
    uint32_t *vt = vals_tmp + base;
    uint64_t *ko = keys_out + base;
    uint32_t *vo = vals_out + base;
    const float *boxes = aabbs + (size_t)c * (N + 1) * 8u;

    __shared__ uint32_t hist[8][256];
    __shared__ uint32_t wc[kSortWaves][256];
    __shared__ uint32_t s_w[65];
    __shared__ uint32_t s_nfin;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;

    for (uint32_t i = tid; i < 8u * 256u; i += kSortBlock) (&hist[0][0])[i] = 0;
    if (tid == 0) s_nfin = 0;
    __syncthreads();
    uint32_t nfin = 0;
    for (uint32_t e = tid; e < N; e += kSortBlock) {
        const uint64_t k = kin[e];
#pragma unroll
        for (int p = 0; p < 8; ++p) atomicAdd(&hist[p][(uint32_t)(k >> (8 * p)) & 255u], 1u);
        nfin += is_finite_f(boxes[(size_t)e * 8u]) ? 1u : 0u;
    }
    atomicAdd(&s_nfin, nfin);
    __syncthreads();
    // exclusive scans of the 8 histograms: wave p scans hist[p] (4 bins per lane)
    if (w < 8) {
        uint32_t v[4], s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = hist[w][lane * 4 + k];
            s += v[k];
        }
        const uint32_t inc = wave_inclusive_scan(s);
        uint32_t run = inc - s;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            hist[w][lane * 4 + k] = run;
            run += v[k];
        }
    }
    __syncthreads();

    const uint64_t lt = lanemask_lt();
    for (int p = 0; p < 8; ++p) {
        const uint64_t *ksrc = (p == 0) ? kin : ((p & 1) ? kt : ko);
        const uint32_t *vsrc = (p & 1) ? vt : vo;
        uint64_t *kdst = (p & 1) ? ko : kt;
        uint32_t *vdst = (p & 1) ? vo : vt;
        const uint32_t shift = 8u * (uint32_t)p;
        for (uint32_t tb = 0; tb < N; tb += kSortBlock * kSortItems) {
            // zero this wave's digit counters
#pragma unroll
            for (int k = 0; k < 4; ++k) wc[w][lane * 4 + k] = 0;
            uint64_t key[kSortItems];
            uint32_t val[kSortItems], dig[kSortItems], rank[kSortItems];
            bool ok[kSortItems];
#pragma unroll
            for (int it = 0; it < kSortItems; ++it) {
                const uint32_t e = tb + w * 64u * kSortItems + it * 64u + lane;
                ok[it] = e < N;
                key[it] = ok[it] ? ksrc[e] : 0ull;
                val[it] = ok[it] ? (p == 0 ? e : vsrc[e]) : 0u;
                dig[it] = (uint32_t)(key[it] >> shift) & 255u;
                const uint64_t vmask = __ballot(ok[it]);
                uint64_t peers = vmask;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const bool bit = (dig[it] >> b) & 1u;
                    const uint64_t bal = __ballot(bit);
                    peers &= bit ? bal : ~bal;
                }
                const uint32_t before = (uint32_t)__popcll(peers & lt);
                const uint32_t old = wc[w][dig[it]];
                rank[it] = old + before;
                if (ok[it] && before == 0) wc[w][dig[it]] = old + (uint32_t)__popcll(peers);
            }
            __syncthreads();
            uint32_t tile_tot = 0;
            if (tid < 256) {
                uint32_t run = 0;
                for (int ww = 0; ww < kSortWaves; ++ww) {
                    const uint32_t x = wc[ww][tid];
                    wc[ww][tid] = run;
                    run += x;
                }
                tile_tot = run;
            }
            __syncthreads();
#pragma unroll
            for (int it = 0; it < kSortItems; ++it) {
                if (ok[it]) {
                    const uint32_t pos = hist[p][dig[it]] + wc[w][dig[it]] + rank[it];
                    kdst[pos] = key[it];
                    vdst[pos] = val[it];
                }
            }
            __syncthreads();
            if (tid < 256) hist[p][tid] += tile_tot;
        }
        __syncthreads();
    }
    // after 8 passes the result is in (ko, vo)
    const uint32_t nf = s_nfin;
    if (tid == 0) n_eff[c] = nf;
    if (nf == N) return;
    // first position whose code is all ones
    uint32_t lo = 0, n = N;
    while (n > 0) {
        const uint32_t h = n / 2;
        if (ko[lo + h] != ~0ull) {
            lo += h + 1;
            n -= h + 1;
        } else {
            n = h;
        }
    }
    const uint32_t t0 = lo;
    if (nf == t0) return;  // the tail holds only objects without data
    // stable partition of the tail [t0, N): finite boxes first (into vt), then copy back
    uint32_t fin_base = t0, inf_base = nf;
    for (uint32_t cb = t0; cb < N; cb += kSortBlock) {
        const uint32_t e = cb + tid;
        uint32_t v = 0, isf = 0;
        if (e < N) {
            v = vo[e];
            isf = is_finite_f(boxes[(size_t)v * 8u]) ? 1u : 0u;
        }
        uint32_t tot;
        const uint32_t r = block_excl_scan_1024(isf, s_w, tot);
        if (e < N) {
            if (isf) {
                vt[fin_base + r] = v;
            } else {
                vt[inf_base + (e - cb - r)] = v;
            }
        }
        fin_base += tot;
        inf_base += (N - cb < kSortBlock ? N - cb : kSortBlock) - tot;
    }
    __syncthreads();
    for (uint32_t e = t0 + tid; e < N; e += kSortBlock) vo[e] = vt[e];
}

__global__ __launch_bounds__(256) void k_gather(const Prob *__restrict__ P, uint32_t np,
                                                const uint32_t *__restrict__ blk_pre)
{
    const uint32_t pi = find_prob(blk_pre, np, blockIdx.x);
    const Prob &q = P[pi];
    const uint32_t lb = blockIdx.x - blk_pre[pi];
    const uint32_t N = q.N;
    const float *__restrict__ aabbs = q.aabbs;
    const uint32_t *__restrict__ srt_idx = q.srt_idx;
    float *__restrict__ srt_aabbs = q.srt_aabbs;
    const uint32_t bps = (N + 1u + 255u) / 256u;
    const uint32_t c = lb / bps;
    if (!q.sort_todo[c]) return;  // gathered by the bucket sort
    const uint32_t p = (lb - c * bps) * 256u + threadIdx.x;
    if (p > N) return;
    const size_t sb = (size_t)c * (N + 1) * 8u;
    const uint32_t src = (p == N) ? N : srt_idx[(size_t)c * N + p];
    const float4 *s = reinterpret_cast<const float4 *>(aabbs + sb + (size_t)src * 8u);
    float4 *d = reinterpret_cast<float4 *>(srt_aabbs + sb + (size_t)p * 8u);
    d[0] = s[0];
    d[1] = s[1];
}

// ---------------------------------------------------------------------------
// stage 2: Morton codes, sort and gather in one workgroup per step
// (detect_conjunctions_morton.cpp:70-223). The sort key is (box without
// data, code, object index): the reference's comparator (:179-196) with its
// unspecified tie order fixed to the object index. Objects are distributed
// into 2^12 buckets by the top code bits (LDS counters) and scattered to them
// by one workgroup per step; then each block of 512 sorted positions, widened
// to whole buckets, is staged in LDS by a workgroup of its own
// (k_bucket_rank) and every object finds its place by counting the smaller
// keys of its bucket. Objects
// without data go after the finite ones in index order. A step with a bucket
// larger than kBktMax (a tight cluster of objects) is left to the radix sort
// (k_sort + k_gather), flagged in todo.
// ---------------------------------------------------------------------------

constexpr uint32_t kBktMax = 1024;

// LDS of the bucket sort of one step.
template <uint32_t BITS>
struct BucketLds {
    uint32_t off[1u << BITS];
    uint32_t w[65];
    uint32_t big;
};

// The bucket sort of step c of problem q by one 1024-thread workgroup.
template <uint32_t BITS>
__device__ void bucket_sort_step(const Prob &q, uint32_t c, BucketLds<BITS> &L)
{
    constexpr uint32_t kBkts = 1u << BITS;
    constexpr uint32_t kBktBits = BITS;
    const uint32_t N = q.N;
    const uint32_t tid = threadIdx.x;
    uint32_t *__restrict__ todo = q.sort_todo;
    uint32_t *__restrict__ n_eff = q.n_eff;
    const float *boxes = q.aabbs + (size_t)c * (N + 1) * 8u;
    float *sboxes = q.srt_aabbs + (size_t)c * (N + 1) * 8u;
    uint64_t *codes = q.mcodes + (size_t)c * N;
    uint64_t *kt = q.keys_tmp + (size_t)c * N;
    uint32_t *vt = q.vals_tmp + (size_t)c * N;
    uint64_t *ko = q.srt_mcodes + (size_t)c * N;
    uint32_t *vo = q.srt_idx + (size_t)c * N;
    uint32_t *s_off = L.off;
    uint32_t *s_w = L.w;
    uint32_t &s_big = L.big;
    const float4 glo = reinterpret_cast<const float4 *>(boxes + (size_t)N * 8u)[0];
    const float4 ghi = reinterpret_cast<const float4 *>(boxes + (size_t)N * 8u)[1];

    for (uint32_t k = tid; k < kBkts; k += 1024u) s_off[k] = 0;
    if (tid == 0) s_big = 0;
    __syncthreads();
    // codes and bucket sizes; the next object's box is read while the
    // current one's code is made
    {
        const float4 *bx = reinterpret_cast<const float4 *>(boxes);
        float4 blo = make_float4(0.f, 0.f, 0.f, 0.f), bhi = blo;
        if (tid < N) {
            blo = bx[(size_t)tid * 2u];
            bhi = bx[(size_t)tid * 2u + 1u];
        }
        for (uint32_t o = tid; o < N; o += 1024u) {
            // (the last object's box again past the end)
            const uint32_t on = min(o + 1024u, N - 1u);
            const float4 nlo = bx[(size_t)on * 2u], nhi = bx[(size_t)on * 2u + 1u];
            const uint64_t code = morton_code(blo, bhi, glo, ghi);
            codes[o] = code;
            if (is_finite_f(blo.x)) atomicAdd(&s_off[(uint32_t)(code >> (64u - kBktBits))], 1u);
            blo = nlo;
            bhi = nhi;
        }
    }
    __syncthreads();
    // exclusive offsets of the buckets (kBkts / 1024 per thread), the largest bucket
    constexpr uint32_t kPer = kBkts / 1024u;
    uint32_t cnt[kPer], sum = 0, mx = 0;
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        cnt[k] = s_off[tid * kPer + k];
        sum += cnt[k];
        mx = cnt[k] > mx ? cnt[k] : mx;
    }
    uint32_t nfin;
    uint32_t run = block_excl_scan_1024(sum, s_w, nfin);
#pragma unroll
    for (uint32_t k = 0; k < kPer; ++k) {
        s_off[tid * kPer + k] = run;
        run += cnt[k];
    }
    if (mx > kBktMax) atomicMax(&s_big, mx);
    __syncthreads();
    if (s_big > kBktMax) {
        if (tid == 0) todo[c] = 1u;
        return;
    }
    if (tid == 0) {
        todo[c] = 0u;
        n_eff[c] = nfin;
        // row N: the global box
        reinterpret_cast<float4 *>(sboxes + (size_t)N * 8u)[0] = glo;
        reinterpret_cast<float4 *>(sboxes + (size_t)N * 8u)[1] = ghi;
    }
    // finite objects into their buckets (any order inside a bucket); objects
    // without data straight to their final place, in index order
    if (nfin == N && N > 0u) {
        // every object has data: each thread moves the codes it made, four
        // at a time, the next four read while the current ones are placed
        // (the last object's code again past the end)
        uint64_t cd[4];
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) cd[k] = codes[min(tid + k * 1024u, N - 1u)];
        for (uint32_t o = tid; o < N; o += 4u * 1024u) {
            uint64_t nx[4];
#pragma unroll
            for (uint32_t k = 0; k < 4u; ++k) nx[k] = codes[min(o + (4u + k) * 1024u, N - 1u)];
#pragma unroll
            for (uint32_t k = 0; k < 4u; ++k) {
                if (o + k * 1024u < N) {
                    const uint32_t pos = atomicAdd(&s_off[(uint32_t)(cd[k] >> (64u - kBktBits))], 1u);
                    kt[pos] = cd[k];
                    vt[pos] = o + k * 1024u;
                }
                cd[k] = nx[k];
            }
        }
    } else if (nfin < N) {
        uint32_t inf_carry = 0;
        for (uint32_t base = 0; base < N; base += 1024u) {
            const uint32_t o = base + tid;
            bool fin = false, inf = false;
            float4 blo = make_float4(0.f, 0.f, 0.f, 0.f), bhi = blo;
            if (o < N) {
                blo = reinterpret_cast<const float4 *>(boxes + (size_t)o * 8u)[0];
                bhi = reinterpret_cast<const float4 *>(boxes + (size_t)o * 8u)[1];
                fin = is_finite_f(blo.x);
                inf = !fin;
            }
            uint32_t tot;
            const uint32_t ex = block_excl_scan_1024(inf ? 1u : 0u, s_w, tot);
            if (fin) {
                const uint64_t code = codes[o];
                const uint32_t pos = atomicAdd(&s_off[(uint32_t)(code >> (64u - kBktBits))], 1u);
                kt[pos] = code;
                vt[pos] = o;
            } else if (inf) {
                const uint32_t p = nfin + inf_carry + ex;
                ko[p] = ~0ull;
                vo[p] = o;
                reinterpret_cast<float4 *>(sboxes + (size_t)p * 8u)[0] = blo;
                reinterpret_cast<float4 *>(sboxes + (size_t)p * 8u)[1] = bhi;
            }
            inf_carry += tot;
        }
    }
    __syncthreads();
    // s_off[b] is now the end of bucket b: for the ranking (k_bucket_rank)
    uint32_t *ends = q.bkt_end + (size_t)c * kBktEndStride;
    for (uint32_t k = tid; k < kBkts; k += 1024u) ends[k] = s_off[k];
}

// The ranking of the bucket sort, one workgroup per block of kLbvhBlock
// sorted positions of a step (one thread each): the block's positions,
// widened to the whole buckets they cut, are staged in LDS, and each object's
// place is its bucket's start plus the number of smaller codes in its bucket
// (object indices compared only among equal codes). The buckets of the
// block's first and last positions are those of the keys scattered there; a
// thread's own key, bucket range and box are read while the block is staged.
// The final write gathers the box, so the sorted boxes, codes and indices
// leave in one pass.
template <uint32_t BITS>
__global__ __launch_bounds__(kLbvhBlock) void k_bucket_rank(const Prob *__restrict__ P, uint32_t np,
                                                           const uint32_t *__restrict__ blk_pre)
{
    constexpr uint32_t kWin = kLbvhBlock + 2u * kBktMax;
    __shared__ uint64_t s_key[kWin];
    __shared__ uint32_t s_val[kWin];
    const uint32_t gblk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t pi = find_prob(blk_pre, np, gblk);
    const Prob &q = P[pi];
    const uint32_t lblk = gblk - blk_pre[pi];
    const uint32_t N = q.N;
    const uint32_t bps = (N + kLbvhBlock - 1u) / kLbvhBlock;
    const uint32_t c = lblk / bps;
    const uint32_t B0 = (lblk - c * bps) * kLbvhBlock;
    if (q.sort_todo[c]) return;  // left to the radix sort
    const uint32_t nfin = q.n_eff[c];
    if (B0 >= nfin) return;  // the whole workgroup
    const uint32_t tid = threadIdx.x;
    const uint32_t *ends = q.bkt_end + (size_t)c * kBktEndStride;
    const uint64_t *kt = q.keys_tmp + (size_t)c * N;
    const uint32_t *vt = q.vals_tmp + (size_t)c * N;
    const uint32_t e = B0 + tid;
    const bool live = e < nfin;
    // the thread's key (in scatter order: ranked among its bucket's keys)
    const uint64_t key = kt[live ? e : B0];
    const uint32_t val = vt[live ? e : B0];
    const uint32_t b = (uint32_t)(key >> (64u - BITS));
    const uint32_t beg = b ? ends[b - 1u] : 0u, end = ends[b];
    const float *boxes = q.aabbs + (size_t)c * (N + 1) * 8u;
    const float4 glo = reinterpret_cast<const float4 *>(boxes + (size_t)val * 8u)[0];
    const float4 ghi = reinterpret_cast<const float4 *>(boxes + (size_t)val * 8u)[1];
    // the window: the buckets of the block's first and last positions
    const uint32_t b0 = (uint32_t)(kt[B0] >> (64u - BITS));
    const uint32_t b1 = (uint32_t)(kt[min(B0 + kLbvhBlock - 1u, nfin - 1u)] >> (64u - BITS));
    const uint32_t wb = b0 ? ends[b0 - 1u] : 0u, we = ends[b1];
    if (we < wb || we - wb > kWin) {  // (the whole workgroup: the window's bounds are its own)
        report_error(q.err, kErrSortWindow, c, B0, 0.0);
        return;
    }
    for (uint32_t x = wb + tid; x < we; x += kLbvhBlock) {
        s_key[x - wb] = kt[x];
        s_val[x - wb] = vt[x];
    }
    __syncthreads();
    if (!live) return;
    if (beg < wb || end > we) {
        report_error(q.err, kErrSortWindow, c, e, 0.0);
        return;
    }
    uint32_t rank = 0, eq = 0;
    for (uint32_t m = beg; m < end; ++m) {
        const uint64_t km = s_key[m - wb];
        rank += km < key ? 1u : 0u;
        eq += km == key ? 1u : 0u;
    }
    if (eq > 1u) {
        // equal codes: the lower object index first
        for (uint32_t m = beg; m < end; ++m) rank += (s_key[m - wb] == key && s_val[m - wb] < val) ? 1u : 0u;
    }
    const uint32_t p = beg + rank;
    q.srt_mcodes[(size_t)c * N + p] = key;
    q.srt_idx[(size_t)c * N + p] = val;
    float *sboxes = q.srt_aabbs + (size_t)c * (N + 1) * 8u;
    reinterpret_cast<float4 *>(sboxes + (size_t)p * 8u)[0] = glo;
    reinterpret_cast<float4 *>(sboxes + (size_t)p * 8u)[1] = ghi;
}

// Next work item of a persistent workgroup: one atomic per workgroup on the
// batch's counter. Every workgroup leaves once the counter passes n_items,
// which it does after a finite number of fetches: no workgroup ever waits.
__device__ inline uint32_t next_item(uint32_t *counter, uint32_t *s_item)
{
    __syncthreads();  // the previous item's LDS is no longer read
    if (threadIdx.x == 0) *s_item = atomicAdd(counter, 1u);
    __syncthreads();
    return *s_item;
}

// Stage 2, persistent: each workgroup takes (problem, step) items in the
// order of the list (largest problems first, so that the long items start
// early and the short ones fill the end).
template <uint32_t BITS>
__global__ __launch_bounds__(1024) void k_bucket_sort(const Prob *__restrict__ P, const uint2 *__restrict__ items,
                                                      uint32_t n_items, uint32_t *counter)
{
    __shared__ BucketLds<BITS> L;
    __shared__ uint32_t s_item;
    for (uint32_t it = next_item(counter, &s_item); it < n_items; it = next_item(counter, &s_item)) {
        const uint2 w = items[it];
        bucket_sort_step<BITS>(P[w.x], w.y, L);
    }
}

// ---------------------------------------------------------------------------
// stage 3: BVH (detect_conjunctions_bvh.cpp:267-699), one workgroup per step.
// Level-synchronous top-down radix split; the split bit of a node is the
// first bit in which its first and last codes differ, which is where the
// reference's ++split_idx search stops. Children of a level are numbered by
// an inclusive scan of the child counts (BFS order, adjacent siblings).
// ---------------------------------------------------------------------------

constexpr int kBvhBlock = 1024;
constexpr size_t kBvhLdsMax = 64u * 1024u;

__device__ inline void box_union(float l[4], float u[4], const float *bx)
{
    const float4 bl = reinterpret_cast<const float4 *>(bx)[0];
    const float4 bu = reinterpret_cast<const float4 *>(bx)[1];
    l[0] = l[0] < bl.x ? l[0] : bl.x;
    l[1] = l[1] < bl.y ? l[1] : bl.y;
    l[2] = l[2] < bl.z ? l[2] : bl.z;
    l[3] = l[3] < bl.w ? l[3] : bl.w;
    u[0] = u[0] > bu.x ? u[0] : bu.x;
    u[1] = u[1] > bu.y ? u[1] : bu.y;
    u[2] = u[2] > bu.z ? u[2] : bu.z;
    u[3] = u[3] > bu.w ? u[3] : bu.w;
}

// Per-step scratch of the BVH build, in LDS when it fits, else in HBM.
//   D[q]     highest differing bit of codes q-1 and q (-1 if equal), q in [1, n)
//   dmax[k]  maximum of D over the 64 positions of chunk k
struct BvhScratch {
    int8_t *D;
    int8_t *dmax;
};

__host__ __device__ inline size_t bvh_scratch_bytes(uint32_t n)
{
    const size_t nch = ((size_t)n + 63u) / 64u;
    return nch * 64u + (nch + 15u) / 16u * 16u;
}

__device__ inline BvhScratch bvh_scratch_at(uint8_t *base, uint32_t n)
{
    const size_t nch = ((size_t)n + 63u) / 64u;
    BvhScratch sc;
    sc.D = reinterpret_cast<int8_t *>(base);
    sc.dmax = reinterpret_cast<int8_t *>(base + nch * 64u);
    return sc;
}

// Maximum of the signed bytes A[lo..hi] and the first position holding it,
// folded into (best, at); reads whole aligned 16-byte groups.
__device__ inline void max_bytes(const int8_t *A, uint32_t lo, uint32_t hi, int &best, uint32_t &at)
{
    for (uint32_t g = lo & ~15u; g <= hi; g += 16u) {
        const uint4 w = *reinterpret_cast<const uint4 *>(A + g);
        const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
        for (uint32_t j = 0; j < 16u; ++j) {
            const uint32_t p = g + j;
            const int v = (int)(int8_t)(ws[j >> 2] >> (8u * (j & 3u)));
            if (p >= lo && p <= hi && v > best) {
                best = v;
                at = p;
            }
        }
    }
}

// Position of the strongest boundary in [lo, hi] (lo <= hi), and its strength
// (-1 if the codes over the range are all equal). Inside a BVH node the
// strongest boundary is unique: its highest differing bit flips once.
__device__ inline int bvh_split(const BvhScratch &sc, uint32_t lo, uint32_t hi, uint32_t &arg)
{
    int best = -1;
    uint32_t at = lo;
    const uint32_t c0 = (lo + 63u) >> 6, c1 = (hi + 1u) >> 6;  // full chunks [c0, c1)
    if (c0 >= c1) {
        max_bytes(sc.D, lo, hi, best, at);
    } else {
        if (lo < c0 * 64u) max_bytes(sc.D, lo, c0 * 64u - 1u, best, at);
        int cb = -1;
        uint32_t ck = c0;
        max_bytes(sc.dmax, c0, c1 - 1u, cb, ck);
        if (cb > best) {
            int b2 = -1;
            max_bytes(sc.D, ck * 64u, ck * 64u + 63u, b2, at);
            best = cb;
        }
        if (hi >= c1 * 64u) max_bytes(sc.D, c1 * 64u, hi, best, at);
    }
    arg = at;
    return best;
}

// BVH build over one conjunction step (detect_conjunctions_bvh.cpp:267-699).
//
// The tree is the binary radix tree of the sorted Morton codes: a node over
// [b, e) whose end codes differ splits at the one position where its highest
// differing bit flips, which is the strongest boundary (highest differing bit
// of neighbouring codes) inside the range; a run of equal codes is one
// multi-object leaf. Nodes are laid out level by level in the reference's
// breadth-first order (children of a level appended in order, left before
// right, detect_conjunctions_bvh.cpp:521-580) and refitted bottom-up
// (:629-691). Boundary strengths and their per-chunk maxima live in LDS, so
// the split search touches no global memory, and consecutive lanes read and
// write consecutive node records. The escape index comes top-down: a left
// child escapes to its right sibling, a right child to its parent's escape.
//
// The headers (escape, begin, end or split, left) live in a scratch array hdr
// while the tree is built: a node's header is made by its parent and completed
// at its own level, where an internal node keeps its split in place of its
// end. Every node record is then written once, whole: a leaf's at its level,
// an internal node's in the refit, so that consecutive lanes write full cache
// lines instead of revisiting parts of records.
template <bool LDS>
__device__ void bvh_step(const Prob &pb, uint32_t c, uint8_t *s_dyn, uint32_t *lvl, uint32_t *s_w)
{
    const uint32_t N = pb.N;
    const uint64_t node_cap = pb.node_cap;
    uint32_t *__restrict__ tree_size = pb.tree_size;
    const uint32_t n = pb.n_eff[c];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    if (n == 0) {
        if (tid == 0) tree_size[c] = 0;
        return;
    }
    const uint64_t *codes = pb.srt_mcodes + (size_t)c * N;
    const float *boxes = pb.srt_aabbs + (size_t)c * (N + 1) * 8u;
    TravNode *T = pb.trav + (size_t)c * node_cap;
    uint4 *H = pb.bvh_hdr + (size_t)c * node_cap;
    const BvhScratch sc = bvh_scratch_at(LDS ? s_dyn : pb.bvh_scr + (size_t)c * pb.bvh_scr_stride, n);
    const uint32_t nch = (n + 63u) / 64u;

    // ---- boundary strengths and chunk maxima (one wave per chunk) ----------
    for (uint32_t q = tid; q < nch * 64u; q += kBvhBlock) {
        int d = -1;
        if (q > 0 && q < n) {
            const uint64_t x = codes[q - 1] ^ codes[q];
            d = x ? (int)(63u - (uint32_t)__clzll(x)) : -1;
        }
        sc.D[q] = (int8_t)d;
        int m = d;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const int o = __shfl_xor(m, off, 64);
            m = m > o ? m : o;
        }
        if (lane == 0) sc.dmax[q >> 6] = (int8_t)m;
    }
    if (tid == 0) H[0] = make_uint4(kTravEnd, 0u, n, 0u);
    __syncthreads();

    // ---- top-down, one level at a time ------------------------------------
    uint32_t lb = 0, le = 1, nlev = 0;
    for (;;) {
        if (tid == 0) lvl[nlev] = lb;
        const uint32_t cnt = le - lb;
        uint32_t carry = 0;
        for (uint32_t cb = 0; cb < cnt; cb += kBvhBlock) {
            const uint32_t k = cb + tid;
            uint32_t esc = 0, b = 0, e = 0, q = 0;
            bool inner = false;
            if (k < cnt) {
                const uint4 h = H[lb + k];
                esc = h.x;
                b = h.y;
                e = h.z;
                if (e - b > 1u) inner = bvh_split(sc, b + 1u, e - 1u, q) >= 0;
            }
            uint32_t tot;
            const uint32_t ex = block_excl_scan_1024(inner ? 2u : 0u, s_w, tot);
            if ((uint64_t)le + carry + tot > node_cap) {
                if (tid == 0) tree_size[c] = 0xffffffffu;
                return;
            }
            if (k < cnt) {
                if (inner) {
                    const uint32_t lc = le + carry + ex;
                    H[lc] = make_uint4(lc + 1u, b, q, 0u);
                    H[lc + 1u] = make_uint4(esc, q, e, 0u);
                    H[lb + k] = make_uint4(esc, b, q, lc);
                } else {
                    float l[4] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
                    float u[4] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
                    for (uint32_t o = b; o != e; ++o) box_union(l, u, boxes + (size_t)o * 8u);
                    uint4 *nd = reinterpret_cast<uint4 *>(&T[lb + k]);
                    nd[0] = make_uint4(__float_as_uint(l[0]), __float_as_uint(l[1]), __float_as_uint(l[2]),
                                       __float_as_uint(l[3]));
                    nd[1] = make_uint4(__float_as_uint(u[0]), __float_as_uint(u[1]), __float_as_uint(u[2]),
                                       __float_as_uint(u[3]));
                    nd[2] = make_uint4(esc, b, e, 0u);
                }
            }
            carry += tot;
        }
        __syncthreads();
        ++nlev;
        lb = le;
        le += carry;
        if (carry == 0 || nlev > kBvhMaxLevels) break;
    }
    if (tid == 0) lvl[nlev] = le;
    __syncthreads();

    // ---- bottom-up refit of the internal nodes ----------------------------
    for (int j = (int)nlev - 2; j >= 0; --j) {
        const uint32_t a = lvl[j], z = lvl[j + 1];
        for (uint32_t ni = a + tid; ni < z; ni += kBvhBlock) {
            const uint4 h = H[ni];
            if (h.w != 0u) {
                const uint32_t left = h.w;
                float l[4] = {__builtin_inff(), __builtin_inff(), __builtin_inff(), __builtin_inff()};
                float u[4] = {-__builtin_inff(), -__builtin_inff(), -__builtin_inff(), -__builtin_inff()};
                box_union(l, u, T[left].lb);
                box_union(l, u, T[left + 1u].lb);
                uint4 *nd = reinterpret_cast<uint4 *>(&T[ni]);
                nd[0] = make_uint4(__float_as_uint(l[0]), __float_as_uint(l[1]), __float_as_uint(l[2]),
                                   __float_as_uint(l[3]));
                nd[1] = make_uint4(__float_as_uint(u[0]), __float_as_uint(u[1]), __float_as_uint(u[2]),
                                   __float_as_uint(u[3]));
                nd[2] = h;
            }
        }
        __syncthreads();
    }
    if (tid == 0) tree_size[c] = le;
}

// Stage 3, persistent: each workgroup takes (problem, step) items in list
// order (largest problems first). LDS: the build scratch of the largest step
// of the list (dynamic) when LDS, else the scratch is in HBM. Eight waves per
// SIMD (at most 64 VGPRs) let two workgroups share a CU: a build is a chain of
// dependent level passes, and the second workgroup's passes fill the first's
// waits.
template <bool LDS>
__global__ __launch_bounds__(kBvhBlock) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_bvh(const Prob *__restrict__ P, const uint2 *__restrict__ items,
                                                     uint32_t n_items, uint32_t *counter)
{
    extern __shared__ __align__(16) uint8_t s_dyn[];
    __shared__ uint32_t lvl[kBvhMaxLevels + 2];
    __shared__ uint32_t s_w[65];
    __shared__ uint32_t s_item;
    for (uint32_t it = next_item(counter, &s_item); it < n_items; it = next_item(counter, &s_item)) {
        const uint2 w = items[it];
        bvh_step<LDS>(P[w.x], w.y, s_dyn, lvl, s_w);
    }
}

// The reference's node records (conjunctions.hpp:65-72) from the device
// layout, for the steps' trees as exposed by get_bvh_tree.
__global__ __launch_bounds__(256) void k_bvh_export(const TravNode *__restrict__ trav,
                                                    const uint32_t *__restrict__ tree_size,
                                                    mzh_bvh_node *__restrict__ nodes, uint64_t node_cap,
                                                    uint32_t blocks_per_step)
{
    const uint32_t c = blockIdx.x / blocks_per_step;
    const uint32_t i = (blockIdx.x - c * blocks_per_step) * 256u + threadIdx.x;
    const uint32_t nn = tree_size[c];
    if (i >= nn) return;
    const TravNode *T = trav + (size_t)c * node_cap;
    const TravNode t = T[i];
    uint32_t end = t.end;
    if (t.left != 0u) {
        // an internal node ends where the last leaf of its right spine ends
        uint32_t x = t.left + 1u;
        while (T[x].left != 0u) x = T[x].left + 1u;
        end = T[x].end;
    }
    mzh_bvh_node r;
    r.begin = t.begin;
    r.end = end;
    r.left = t.left ? (int32_t)t.left : -1;
    r.right = t.left ? (int32_t)(t.left + 1u) : -1;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
        r.lb[d] = t.lb[d];
        r.ub[d] = t.ub[d];
    }
    nodes[(size_t)c * node_cap + i] = r;
}

// ---------------------------------------------------------------------------
// stage 3, the broad phase's walk tree: the radix tree of mzh_lbvh.hpp (the
// reference's tree over runs of equal codes, numbered internal node k at 2k,
// leaf p at 2p + 1), built for every step at once. The reference's
// breadth-first records (k_bvh) are made only when a caller asks for a tree.
// ---------------------------------------------------------------------------

__device__ inline void box_min_max(float4 &l, float4 &u, const float4 &bl, const float4 &bu)
{
    l.x = l.x < bl.x ? l.x : bl.x;
    l.y = l.y < bl.y ? l.y : bl.y;
    l.z = l.z < bl.z ? l.z : bl.z;
    l.w = l.w < bl.w ? l.w : bl.w;
    u.x = u.x > bu.x ? u.x : bu.x;
    u.y = u.y > bu.y ? u.y : bu.y;
    u.z = u.z > bu.z ? u.z : bu.z;
    u.w = u.w > bu.w ? u.w : bu.w;
}

// Split and child flags of an internal node: split << 2 | left_leaf | right_leaf << 1.
__device__ inline uint32_t lbvh_info(const LbvhNode &nd)
{
    return (nd.split << 2) | (nd.left_leaf ? 1u : 0u) | (nd.right_leaf ? 2u : 0u);
}

__device__ inline uint32_t lbvh_sibling(uint32_t info, uint32_t self)
{
    const uint32_t split = info >> 2;
    const uint32_t lc = lbvh_left(split, (info & 1u) != 0u), rc = lbvh_right(split, (info & 2u) != 0u);
    return self == lc ? rc : lc;
}

// The walk's fields of an internal node: escape, then the right and left
// children (the right child's index halved is the split); a run's root is
// the walk's leaf for the run: escape, begin, end, 0.
__device__ inline uint4 lbvh_meta(const LbvhNode &nd, uint32_t esc)
{
    return nd.run ? make_uint4(esc, nd.lo, nd.hi + 1u, 0u)
                  : make_uint4(esc, lbvh_right(nd.split, nd.right_leaf), lbvh_left(nd.split, nd.left_leaf), 1u);
}

// The codes of a block's positions [B0 - 1, B1 + 2) staged in LDS.
struct LdsCodes {
    const uint64_t *s;
    uint32_t b0;
    __device__ uint64_t operator()(uint32_t i) const { return s[i + 1u - b0]; }
};

// One workgroup per block of BS consecutive sorted positions of one step, one
// position per thread, with the block's codes and a halo of one position
// before and two after in LDS. Boxes are on the step's quantization grid
// (qbox). The thread of position p writes leaf p (the object's box) and counts the runs of equal codes (the step's node total,
// 2 runs - 1, is the reference's tree size); for internal node k = p it finds
// whether the node's range leaves the block (one code beyond the block,
// lbvh_leaves), and if not, the node from the block's codes alone
// (lbvh_node); nodes that leave go on the step's spanning list for
// k_lbvh_top. Internal boxes are then made bottom-up inside the block: from
// each leaf a thread climbs to the parent, and of the two children's threads
// the second to arrive (an LDS counter per node) makes the parent's box from
// both children's boxes in LDS, writes the parent's whole record (walk fields
// kept in LDS until then, so that each record is written once) and climbs on.
// A node whose parent spans goes
// on the step's frontier list instead. Boxes are unions (min / max), so the
// order does not matter.
template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_lbvh_build(const Prob *__restrict__ P, uint32_t np,
                                                   const uint32_t *__restrict__ blk_pre)
{
    __shared__ uint4 s_box[2u * BS];      // node grid boxes by local index 2 (k - B0) + leaf
    __shared__ uint64_t s_code[BS + 3u];  // codes of the positions [B0 - 1, B1 + 2)
    __shared__ uint16_t s_par[2u * BS];   // parent's slot when the parent lies inside the block
    __shared__ uint4 s_meta[BS];          // walk fields of internal node B0 + slot
    __shared__ uint32_t s_info[BS];       // split and child flags of internal node B0 + slot
    __shared__ uint32_t s_flag[BS];       // arrivals at internal node B0 + slot
    constexpr uint16_t kNoParent = 0xffffu;
    static_assert(2u * BS <= kNoParent, "LDS slots");
    const uint32_t gblk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t pi = find_prob(blk_pre, np, gblk);
    const Prob &q = P[pi];
    const uint32_t lblk = gblk - blk_pre[pi];
    const uint32_t bps = (q.N + BS - 1u) / BS;
    const uint32_t c = lblk / bps;
    const uint32_t B0 = (lblk - c * bps) * BS;
    const uint32_t n = q.n_eff[c];
    if (B0 >= n) return;  // the whole workgroup
    const uint32_t B1 = n - B0 < BS ? n : B0 + BS;
    const uint32_t t = threadIdx.x, p = B0 + t;
    const bool live = p < B1;
    const uint64_t *codes = q.srt_mcodes + (size_t)c * q.N;
    const float *boxes = q.srt_aabbs + (size_t)c * (q.N + 1) * 8u;
    WalkNode *T = q.wtree + (size_t)c * q.node_cap;
    const LbvhWords W = lbvh_words(q, c);
    const LdsCodes code{s_code, B0};
    auto local = [&](uint32_t g) { return 2u * ((g >> 1) - B0) + (g & 1u); };
    for (uint32_t x = t; x < BS + 3u; x += BS) {
        const int64_t pos = (int64_t)B0 - 1 + x;
        s_code[x] = pos >= 0 && pos < (int64_t)n ? codes[pos] : 0ull;
    }
    s_par[2u * t] = kNoParent;
    s_par[2u * t + 1u] = kNoParent;
    s_flag[t] = 0u;
    __syncthreads();
    uint4 qb = make_uint4(0u, 0u, 0u, 0u);
    uint32_t runs = 0;
    if (live) {
        const QGrid g = qgrid(reinterpret_cast<const float4 *>(boxes + (size_t)q.N * 8u)[0],
                              reinterpret_cast<const float4 *>(boxes + (size_t)q.N * 8u)[1]);
        qb = qbox(g, reinterpret_cast<const float4 *>(boxes + (size_t)p * 8u)[0],
                  reinterpret_cast<const float4 *>(boxes + (size_t)p * 8u)[1]);
        s_box[2u * t + 1u] = qb;
        // the leaf's whole record (the walk reads a node's box and fields together)
        uint4 *leaf = reinterpret_cast<uint4 *>(&T[2u * p + 1u]);
        leaf[0] = qb;
        leaf[1] = make_uint4(lbvh_escape(code, n, p), p, p + 1u, 0u);
        runs = p == 0u ? 1u : (code(p) != code(p - 1u) ? 2u : 0u);
        if (p + 1u < n) {
            int dmin;
            const int64_t dir = lbvh_dir(code, n, p, dmin);
            if (lbvh_leaves(code, n, p, B0, B1, dir, dmin)) {
                const uint32_t at = atomicAdd(W.nspan, 1u);
                if (at < W.span_cap) {
                    W.span[at] = p;
                } else {
                    report_error(q.err, kErrBvhBuild, c, 2u * p, 0.0);
                }
            } else {
                const LbvhNode nd = lbvh_node(code, n, p, B0, B1, dir, dmin);
                s_meta[t] = lbvh_meta(nd, lbvh_escape(code, n, nd.hi));
                s_info[t] = lbvh_info(nd);
                // inside the block, so are its children
                s_par[local(lbvh_left(nd.split, nd.left_leaf))] = (uint16_t)t;
                s_par[local(lbvh_right(nd.split, nd.right_leaf))] = (uint16_t)t;
            }
        }
    }
    {
        uint32_t r = runs;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
        if ((t & 63u) == 0u && r) atomicAdd(&q.tree_size[c], r);
    }
    __syncthreads();
    if (!live) return;
    const uint32_t root = lbvh_root(n);
    uint32_t self = 2u * p + 1u;
    // a climb is at most as long as the tree is deep (96 key bits)
    for (uint32_t guard = 0; guard < 128u; ++guard) {
        const uint32_t pa = s_par[local(self)];
        if (pa == kNoParent) {
            if (self != root) {
                const uint32_t at = atomicAdd(W.nfront, 1u);
                if (at < W.front_cap) {
                    W.front[at] = self;
                } else {
                    report_error(q.err, kErrBvhBuild, c, self, 0.0);
                }
            }
            break;
        }
        // the first child to arrive leaves; its box is in LDS for the second
        // (the climb exchanges data through LDS only: its fences order LDS
        // accesses and leave the node records' global stores in flight)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        const uint32_t old = __hip_atomic_fetch_add(&s_flag[pa], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (old == 0u) break;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
        const uint32_t sib = local(lbvh_sibling(s_info[pa], self));
        qb = qbox_union(qb, s_box[sib]);
        s_box[2u * pa] = qb;
        // the whole record at once
        uint4 *nb = reinterpret_cast<uint4 *>(&T[2u * (B0 + pa)]);
        nb[0] = qb;
        nb[1] = s_meta[pa];
        self = 2u * (B0 + pa);
    }
}

// Threads per workgroup of k_lbvh_top (256 measured slower: 0.75 against
// 0.71 ms for the stage).
constexpr uint32_t kLbvhTopBlock = 1024;

// The nodes that span blocks, one workgroup per step: first the spanning
// nodes themselves (from the step's codes in global memory, each with its own
// thread), their children's parent links and counters; then the same
// bottom-up climb as inside the blocks, from the step's frontier list, with
// the counters and split words in the spanning nodes' box words (the second
// child to arrive reads the split word, then writes the box there).
__global__ __launch_bounds__(kLbvhTopBlock) void k_lbvh_top(const Prob *__restrict__ P, uint32_t np,
                                                   const uint32_t *__restrict__ step_pre)
{
    const uint32_t s = blockIdx.x;
    const Prob &q = P[find_prob(step_pre, np, s)];
    const uint32_t c = s - q.step0;
    const uint32_t n = q.n_eff[c];
    if (n < 2u) return;
    const LbvhWords W = lbvh_words(q, c);
    const uint64_t *codes = q.srt_mcodes + (size_t)c * q.N;
    WalkNode *T = q.wtree + (size_t)c * q.node_cap;
    const uint32_t ns = min(min(*W.nspan, n - 1u), W.span_cap);
    for (uint32_t x = threadIdx.x; x < ns; x += blockDim.x) {
        const uint32_t k = W.span[x];
        if (k + 1u >= n) {
            report_error(q.err, kErrBvhBuild, c, 2u * k, 0.0);
            continue;
        }
        const LbvhNode nd = lbvh_node(codes, n, k);
        *reinterpret_cast<uint4 *>(&T[2u * k].escape) = lbvh_meta(nd, lbvh_escape(codes, n, nd.hi));
        // the arrival counter and split word, in the box words until the box
        *reinterpret_cast<uint4 *>(T[2u * k].qb) = make_uint4(0u, lbvh_info(nd), 0u, 0u);
        W.par[lbvh_left(nd.split, nd.left_leaf)] = 2u * k;
        W.par[lbvh_right(nd.split, nd.right_leaf)] = 2u * k;
    }
    __syncthreads();
    const uint32_t nf = min(*W.nfront, W.front_cap);
    for (uint32_t f = threadIdx.x; f < nf; f += blockDim.x) {
        uint32_t self = W.front[f];
        if (self >= 2u * n) {
            report_error(q.err, kErrBvhBuild, c, self, 0.0);
            break;
        }
        uint4 qb = *reinterpret_cast<const uint4 *>(T[self].qb);
        for (uint32_t guard = 0; guard < 128u && self != 0u; ++guard) {
            const uint32_t pa = W.par[self];
            if (pa >= 2u * n || (pa & 1u)) {
                report_error(q.err, kErrBvhBuild, c, self, 0.0);
                break;
            }
            uint32_t *cw = T[pa].qb;
            const uint32_t old = __hip_atomic_fetch_add(&cw[0], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (old == 0u) break;
            const uint32_t sib = lbvh_sibling(cw[1], self);
            if (sib >= 2u * n) {
                report_error(q.err, kErrBvhBuild, c, pa, 0.0);
                break;
            }
            qb = qbox_union(qb, *reinterpret_cast<const uint4 *>(T[sib].qb));
            *reinterpret_cast<uint4 *>(T[pa].qb) = qb;
            self = pa;
        }
    }
}

// ---------------------------------------------------------------------------
// stage 4: broad phase (detect_conjunctions_broad_phase.cpp:153-339)
// One lane per (step, Morton-sorted query object); stack traversal with
// closed-interval 4-D overlap; a pair (i, j) is emitted from the query of the
// lower original index i. Counts go to cnt[step][i]; an exclusive scan over i
// then places every query's pairs in (i, j) order without a global sort.
// ---------------------------------------------------------------------------

constexpr uint32_t kBpSlots = 4;  // partners kept inline per object and step

__device__ inline bool box_overlap(const float4 &ql, const float4 &qu, const float4 &nl, const float4 &nu)
{
    return (qu.x >= nl.x && ql.x <= nu.x) && (qu.y >= nl.y && ql.y <= nu.y) && (qu.z >= nl.z && ql.z <= nu.z)
           && (qu.w >= nl.w && ql.w <= nu.w);
}

// Record the broad-phase pair (i, j), i < j: count it for i, keep the first
// kBpSlots partners inline, and push the others on i's list in the overflow
// pool (entry = partner, next entry; head[step][i] = latest entry).
//
// Called by the lanes of a wave that record a pair (any subset of the wave):
// the pool entries of the lanes that need one are taken with one atomic on the
// counter of the block's part of the pool.
__device__ inline void bp_emit(bool has, uint32_t c, uint32_t i, uint32_t j, uint32_t N, uint32_t *cnt,
                               uint32_t *slots, uint2 *ovf, uint32_t *head, uint32_t *ovf_count, uint32_t ovf_cap)
{
    const size_t g = (size_t)c * N + i;
    bool need = false;
    if (has) {
        const uint32_t k = atomicAdd(&cnt[g], 1u);
        if (k < kBpSlots) {
            slots[g * kBpSlots + k] = j;
        } else {
            need = true;
        }
    }
    const uint64_t m = __ballot(need);
    if (m == 0) return;
    const uint32_t part = blockIdx.x & (kPoolShards - 1u);
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if ((int)__lane_id() == leader) base = atomicAdd(&ovf_count[part * kPoolStride], (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (need) {
        const uint32_t loc = base + (uint32_t)__popcll(m & ((1ull << __lane_id()) - 1ull));
        const uint32_t part_cap = ovf_cap / kPoolShards;
        if (loc < part_cap) {
            const uint32_t e = part * part_cap + loc;
            const uint32_t next = atomicExch(&head[g], e);
            ovf[e] = make_uint2(j, next);
        }
    }
}

// Broad phase (detect_conjunctions_broad_phase.cpp:153-339). The reference
// walks the tree from every object i and emits (i, j), i < j, for every member
// j of every leaf whose box overlaps box_i (closed intervals, :269-271), with
// otype_i + otype_j <= 3 (:294-310). Equivalently: (i, j) is a pair iff box_i
// overlaps the box of the leaf holding j.
//
// Here every pair is found once. The query at sorted position o walks only
// the part of the tree to the right of its own run of equal codes [rb, re):
// it starts at the escape of its run's leaf, the node that follows the leaf
// in depth-first order, so it visits the subtrees hanging to the right of the
// path from the root to its leaf and never the path itself (whose nodes all
// hold the query's leaf, so a walk from the root would enter each). Its query box
// is the box of its own leaf, which contains box_o, so every pair between o
// and a later leaf is reached from o whichever of the two has the lower
// original index; the reference's condition is then checked exactly for the
// pair (min, max) of the two original indices. Pairs inside o's own run are
// emitted by o for the members after it. Each pair is counted for its lower
// index i (bp_emit); the order within i's partners is fixed when they are
// placed.
template <uint32_t BS>
__global__ __launch_bounds__(BS) void k_bp(const Prob *__restrict__ P, uint32_t np,
                                           const uint32_t *__restrict__ blk_pre)
{
    // logical block: consecutive logical blocks (one step of one problem) on one XCD
    const uint32_t gblk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t pi = find_prob(blk_pre, np, gblk);
    const Prob &q = P[pi];
    const uint32_t lblk = gblk - blk_pre[pi];
    const uint32_t N = q.N;
    const uint64_t node_cap = q.node_cap;
    const WalkNode *__restrict__ wtree = q.wtree;
    const float *__restrict__ srt_aabbs = q.srt_aabbs;
    const uint64_t *__restrict__ srt_mcodes = q.srt_mcodes;
    const uint32_t *__restrict__ srt_idx = q.srt_idx;
    const uint32_t *__restrict__ n_eff = q.n_eff;
    const int32_t *__restrict__ otypes = q.otypes;
    uint32_t *__restrict__ cnt = q.cnt;
    uint32_t *__restrict__ slots = q.slots;
    uint2 *__restrict__ ovf = q.ovf;
    uint32_t *__restrict__ head = q.head;
    uint32_t *__restrict__ ovf_count = q.ovf_count;
    const uint32_t ovf_cap = q.ovf_cap;
    const uint32_t bps = (N + BS - 1u) / BS;
    const uint32_t c = lblk / bps;
    const uint32_t o = (lblk - c * bps) * BS + threadIdx.x;
    const uint32_t n = n_eff[c];
    if (o >= n) return;
    const WalkNode *T = wtree + (size_t)c * node_cap;
    const uint32_t *vidx = srt_idx + (size_t)c * N;
    const float *boxes = srt_aabbs + (size_t)c * (N + 1) * 8u;
    const float4 bl = reinterpret_cast<const float4 *>(boxes + (size_t)o * 8u)[0];
    const float4 bu = reinterpret_cast<const float4 *>(boxes + (size_t)o * 8u)[1];
    const uint32_t orig = vidx[o];
    const int32_t ot = otypes[orig];
    if (ot + 1 > 3) return;  // no object type can pair with this one
    // own run of equal codes [rb, re) and its leaf box
    const uint64_t *codes = srt_mcodes + (size_t)c * N;
    const uint64_t code = codes[o];
    uint32_t rb = o, re = o + 1u;
    while (rb > 0 && codes[rb - 1u] == code) --rb;
    while (re < n && codes[re] == code) ++re;
    const bool multi = re - rb > 1u;
    // partners found during the walk wait in registers and are recorded
    // together (all lanes at once) when the buffer fills and at the end
    constexpr uint32_t kBuf = 4;
    uint32_t buf[kBuf] = {};
    uint32_t nbuf = 0;
    auto flush = [&]() {
#pragma unroll
        for (uint32_t k = 0; k < kBuf; ++k) {
            if (__ballot(k < nbuf) == 0) break;  // the calling lanes' buffers are done
            const uint32_t oi = buf[k];
            bp_emit(k < nbuf, c, orig < oi ? orig : oi, orig < oi ? oi : orig, N, cnt, slots, ovf, head, ovf_count,
                    ovf_cap);
        }
        nbuf = 0;
    };
    auto add = [&](uint32_t oi) {
#pragma unroll
        for (uint32_t k = 0; k < kBuf; ++k)
            if (k == nbuf) buf[k] = oi;
        if (++nbuf == kBuf) flush();
    };
    float4 ql = bl, qu = bu;
    if (multi) {
        float l[4] = {bl.x, bl.y, bl.z, bl.w}, u[4] = {bu.x, bu.y, bu.z, bu.w};
        for (uint32_t m = rb; m != re; ++m) box_union(l, u, boxes + (size_t)m * 8u);
        ql = make_float4(l[0], l[1], l[2], l[3]);
        qu = make_float4(u[0], u[1], u[2], u[3]);
        // pairs inside the run: every member overlaps the leaf box
        for (uint32_t m = o + 1u; m < re; ++m) {
            const uint32_t oi = vidx[m];
            if (ot + otypes[oi] <= 3) add(oi);
        }
    }
    // the walk tree (mzh_lbvh.hpp): a depth-first walk visits each of its
    // 2n - 1 reachable nodes at most once. Its boxes are on the step's grid
    // (qbox), which holds every box that meets the query box on the grid too.
    const uint4 qq = qbox(qgrid(reinterpret_cast<const float4 *>(boxes + (size_t)N * 8u)[0],
                                reinterpret_cast<const float4 *>(boxes + (size_t)N * 8u)[1]),
                          ql, qu);
    const uint32_t lim = 2u * n;
    uint32_t i = lbvh_escape(LbvhArray{codes}, n, re - 1u), visits = 0;
    while (i < lim) {
        if (++visits > lim) {
            report_error(q.err, kErrBvhWalk, c, o, 0.0);
            break;
        }
        const uint4 nq = *reinterpret_cast<const uint4 *>(T[i].qb);
        const uint4 meta = *reinterpret_cast<const uint4 *>(&T[i].escape);
        // an internal node {escape, right, left, 1} is entered, to its left
        // child (every node the walk reaches lies after the run); a leaf
        // {escape, begin, end, 0}, or a node missed, is left by its escape.
        // The next node is chosen from the record's words on every path, so
        // that the record is read with one load.
        const bool hit_node = qbox_overlap(qq, nq);
        const bool enter = hit_node && meta.w != 0u;
        const uint32_t next = enter ? meta.z : meta.x;
        if (hit_node && meta.w == 0u) {
            const bool single = meta.z - meta.y == 1u;
            // the leaf's float box: its object's, or the union of its run's
            float4 nl = reinterpret_cast<const float4 *>(boxes + (size_t)meta.y * 8u)[0];
            float4 nu = reinterpret_cast<const float4 *>(boxes + (size_t)meta.y * 8u)[1];
#pragma unroll 1
            for (uint32_t m = meta.y + 1u; m < meta.z; ++m)
                box_min_max(nl, nu, reinterpret_cast<const float4 *>(boxes + (size_t)m * 8u)[0],
                            reinterpret_cast<const float4 *>(boxes + (size_t)m * 8u)[1]);
            // box_o against the leaf box, needed when the query box is larger than box_o
            const bool o_hits_leaf = !multi || box_overlap(bl, bu, nl, nu);
            const uint32_t m_end = box_overlap(ql, qu, nl, nu) ? meta.z : meta.y;
            for (uint32_t m = meta.y; m != m_end; ++m) {
                const uint32_t oi = vidx[m];
                if (ot + otypes[oi] > 3) continue;
                bool hit;
                if (orig < oi) {
                    hit = o_hits_leaf;  // box_o against the leaf of m
                } else if (!multi && single) {
                    hit = true;  // box_m is the leaf box and the query box is box_o
                } else {
                    // box_m against the leaf of o
                    const float4 ml = reinterpret_cast<const float4 *>(boxes + (size_t)m * 8u)[0];
                    const float4 mu = reinterpret_cast<const float4 *>(boxes + (size_t)m * 8u)[1];
                    hit = box_overlap(ml, mu, ql, qu);
                }
                if (hit) add(oi);
            }
        }
        i = next;
    }
    flush();
}

// Place every object's pairs at its offset in (i, j) order: the inline
// partners and, past kBpSlots, those on its overflow list, sorted by j.
__global__ __launch_bounds__(256) void k_bp_place(const Prob *__restrict__ P, uint32_t np,
                                                  const uint32_t *__restrict__ blk_pre)
{
    const uint32_t pi = find_prob(blk_pre, np, blockIdx.x);
    const Prob &q = P[pi];
    const uint32_t N = q.N, C = q.C;
    const uint32_t *__restrict__ cnt = q.cnt;
    const uint32_t *__restrict__ slots = q.slots;
    const uint2 *__restrict__ ovf = q.ovf;
    const uint32_t *__restrict__ head = q.head;
    const uint32_t ovf_cap = q.ovf_cap;
    const uint64_t *__restrict__ off = q.off;
    const uint64_t *__restrict__ step_off = q.step_off;
    const uint64_t g = (uint64_t)(blockIdx.x - blk_pre[pi]) * 256u + threadIdx.x;
    if (g >= (uint64_t)N * C) return;
    const uint32_t n = cnt[g];
    if (n == 0) return;
    const uint32_t c = (uint32_t)(g / N), orig = (uint32_t)(g - (uint64_t)c * N);
    const uint64_t at = step_off[c] + off[g];
    if (at + n > q.pair_cap) return;  // too small a pair buffer: flagged, the screening runs again
    mzh_aabb_collision *out = q.pairs + at;
    // the step of every pair, for the narrow phase
    uint32_t *out_c = q.pair_c + at;
    for (uint32_t k = 0; k < n; ++k) out_c[k] = c;
    if (n <= kBpSlots) {
        uint32_t v[kBpSlots];
#pragma unroll
        for (uint32_t k = 0; k < kBpSlots; ++k) v[k] = (k < n) ? slots[g * kBpSlots + k] : 0xffffffffu;
        // sorting network on kBpSlots = 4 values (padding sorts last)
        auto cs = [](uint32_t &x, uint32_t &y) {
            const uint32_t a = x < y ? x : y, b = x < y ? y : x;
            x = a;
            y = b;
        };
        cs(v[0], v[1]);
        cs(v[2], v[3]);
        cs(v[0], v[2]);
        cs(v[1], v[3]);
        cs(v[1], v[2]);
#pragma unroll
        for (uint32_t k = 0; k < kBpSlots; ++k)
            if (k < n) out[k] = mzh_aabb_collision{orig, v[k]};
        return;
    }
    // gather, then insertion sort in place (objects with many partners are rare)
    uint32_t k = 0;
    for (; k < kBpSlots; ++k) out[k] = mzh_aabb_collision{orig, slots[g * kBpSlots + k]};
    for (uint32_t e = head[g]; e < ovf_cap && k < n; e = ovf[e].y) out[k++] = mzh_aabb_collision{orig, ovf[e].x};
    for (uint32_t a = 1; a < k; ++a) {
        const mzh_aabb_collision x = out[a];
        uint32_t b = a;
        while (b > 0 && out[b - 1].j > x.j) {
            out[b] = out[b - 1];
            --b;
        }
        out[b] = x;
    }
}

// per-step exclusive scan of cnt[step][0..N) -> off, step totals (one
// workgroup per conjunction step of the batch)
__global__ __launch_bounds__(1024) void k_scan_counts(const Prob *__restrict__ P, uint32_t np,
                                                      const uint32_t *__restrict__ step_pre)
{
    const Prob &q = P[find_prob(step_pre, np, blockIdx.x)];
    const uint32_t c = blockIdx.x - q.step0;
    const uint32_t N = q.N;
    const uint32_t *__restrict__ cnt = q.cnt;
    uint64_t *__restrict__ off = q.off;
    uint64_t *__restrict__ step_total = q.step_total;
    __shared__ uint32_t s_w[65];
    uint64_t carry = 0;
    for (uint32_t cb = 0; cb < N; cb += 1024) {
        const uint32_t e = cb + threadIdx.x;
        const uint32_t v = e < N ? cnt[(size_t)c * N + e] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan_1024(v, s_w, tot);
        if (e < N) off[(size_t)c * N + e] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) step_total[c] = carry;
}

// exclusive scan over the step totals -> step_off[0..C], step_off[C] = the
// problem's pair count (one workgroup per problem)
__global__ __launch_bounds__(1024) void k_scan_steps(const Prob *__restrict__ P)
{
    const Prob &q = P[blockIdx.x];
    const uint64_t *__restrict__ step_total = q.step_total;
    uint64_t *__restrict__ step_off = q.step_off;
    const uint32_t C = q.C;
    __shared__ uint64_t s_w[17];
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint64_t carry = 0;
    for (uint32_t cb = 0; cb < C; cb += 1024) {
        const uint32_t e = cb + threadIdx.x;
        const uint64_t v = e < C ? step_total[e] : 0ull;
        uint64_t inc = v;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint64_t t = __shfl_up(inc, d, 64);
            if (lane >= (uint32_t)d) inc += t;
        }
        if (lane == 63) s_w[w] = inc;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint64_t run = 0;
            for (int k = 0; k < 16; ++k) {
                const uint64_t x = s_w[k];
                s_w[k] = run;
                run += x;
            }
            s_w[16] = run;
        }
        __syncthreads();
        if (e < C) step_off[e] = carry + s_w[w] + inc - v;
        carry += s_w[16];
        __syncthreads();
    }
    if (threadIdx.x == 0) step_off[C] = carry;
}

// Pair and pool counts of every problem, their overflow flags, and the
// exclusive prefix of the placed pairs over the batch (one wave: the pool
// count a problem needs is kPoolShards times the count of its most used part).
__global__ void k_pair_prefix(const Prob *__restrict__ P, uint32_t np, uint64_t *__restrict__ pair_pre)
{
    if (blockIdx.x != 0 || threadIdx.x >= 64) return;
    uint64_t run = 0;
    for (uint32_t p = 0; p < np; ++p) {
        const Prob &q = P[p];
        uint32_t mp = threadIdx.x < kPoolShards ? q.ovf_count[threadIdx.x * kPoolStride] : 0u;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = __shfl_xor(mp, off, 64);
            mp = mp > o ? mp : o;
        }
        if (threadIdx.x != 0) continue;
        const uint64_t n = q.step_off[q.C];
        const uint64_t nov = (uint64_t)mp * kPoolShards;
        q.counts[0] = n;
        q.counts[1] = nov;
        uint32_t f = 0;
        if (n > q.pair_cap) f |= kFlagPairs;
        if (nov > q.ovf_cap) f |= kFlagOvf;
        if (f) atomicOr(q.flags, f);
        pair_pre[p] = run;
        // an incomplete pair list is not screened
        run += (f == 0) ? n : 0;
    }
    if (threadIdx.x == 0) pair_pre[np] = run;
}

// Broad-phase counts and list heads start empty (one thread per (step, object)).
__global__ __launch_bounds__(256) void k_bp_reset(const Prob *__restrict__ P, uint32_t np,
                                                  const uint32_t *__restrict__ blk_pre)
{
    const uint32_t pi = find_prob(blk_pre, np, blockIdx.x);
    const Prob &q = P[pi];
    const uint64_t g = (uint64_t)(blockIdx.x - blk_pre[pi]) * 256u + threadIdx.x;
    if (g >= (uint64_t)q.N * q.C) return;
    q.cnt[g] = 0u;
    q.head[g] = 0xffffffffu;
}

template <int ORDER>
hipError_t dispatch_aabbs(uint32_t order, int mode, const Prob *P, uint32_t np, const BatchTables &t,
                          uint32_t blk_begin, uint32_t n_blocks, hipStream_t stream)
{
    if (order == (uint32_t)ORDER) {
        if (mode == MZH_MODE_MIXED) {
            hipLaunchKernelGGL((k_aabbs<ORDER, true>), dim3(n_blocks), dim3(256), 0, stream, P, np, t.aabb_blk,
                               blk_begin);
        } else {
            hipLaunchKernelGGL((k_aabbs<ORDER, false>), dim3(n_blocks), dim3(256), 0, stream, P, np, t.aabb_blk,
                               blk_begin);
        }
        return hipGetLastError();
    }
    if constexpr (ORDER < kMaxOrder) {
        return dispatch_aabbs<ORDER + 1>(order, mode, P, np, t, blk_begin, n_blocks, stream);
    } else {
        return hipErrorInvalidValue;
    }
}

} // namespace

hipError_t launch_transpose_cf4(const double *traj, double *cf4, float *cf4f, uint64_t n_steps_total, uint32_t order,
                                hipStream_t stream)
{
    const uint64_t n = n_steps_total * 4ull * (order + 1u);
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_transpose_cf4, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, traj, cf4, cf4f,
                       n_steps_total, order + 1u);
    return hipGetLastError();
}

hipError_t launch_step_bounds(const double *cf4, const double *time, const uint64_t *step0, const uint64_t *nsteps,
                              const uint64_t *time_off, uint32_t n_objs, uint64_t n_steps_total, uint32_t order,
                              float4 *sbnd, uint32_t *scal, hipStream_t stream)
{
    hipError_t e = hipMemsetAsync(scal, 0, 2 * sizeof(uint32_t), stream);
    if (e != hipSuccess || n_steps_total == 0) return e;
    return dispatch_step_bounds<kMinOrder>(order, cf4, time, step0, nsteps, time_off, n_objs, n_steps_total, sbnd,
                                           scal, stream);
}

hipError_t launch_batch_init(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                             BatchCounters *bc, hipStream_t stream)
{
    hipLaunchKernelGGL(k_batch_init, dim3((n_steps_all + 255u) / 256u), dim3(256), 0, stream, P, np, t.steps,
                       n_steps_all, bc);
    return hipGetLastError();
}

hipError_t launch_aabbs(uint32_t order, int mode, const Prob *P, uint32_t np, const BatchTables &t,
                        uint32_t blk_begin, uint32_t n_blocks, hipStream_t stream)
{
    if (n_blocks == 0) return hipSuccess;
    return dispatch_aabbs<kMinOrder>(order, mode, P, np, t, blk_begin, n_blocks, stream);
}

hipError_t launch_gbox_finalize(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                                hipStream_t stream)
{
    hipLaunchKernelGGL(k_gbox_finalize, dim3((n_steps_all + 255u) / 256u), dim3(256), 0, stream, P, np, t.steps,
                       n_steps_all);
    return hipGetLastError();
}

hipError_t launch_morton_sort(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                              uint32_t n_gather_blocks, uint32_t n_rank_blocks, BatchCounters *bc,
                              const LaunchShape &ls, hipStream_t stream)
{
    const dim3 g(std::max(1u, std::min(ls.sort_wgs, t.n_sort_items)));
    switch (ls.bucket_bits) {
        case 10:
            hipLaunchKernelGGL(k_bucket_sort<10>, g, dim3(1024), 0, stream, P, t.sort_items, t.n_sort_items,
                               &bc->next_sort);
            break;
        case 11:
            hipLaunchKernelGGL(k_bucket_sort<11>, g, dim3(1024), 0, stream, P, t.sort_items, t.n_sort_items,
                               &bc->next_sort);
            break;
        default:
            hipLaunchKernelGGL(k_bucket_sort<12>, g, dim3(1024), 0, stream, P, t.sort_items, t.n_sort_items,
                               &bc->next_sort);
            break;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (n_rank_blocks) {
        switch (ls.bucket_bits) {
            case 10:
                hipLaunchKernelGGL(k_bucket_rank<10>, dim3(n_rank_blocks), dim3(kLbvhBlock), 0, stream, P, np,
                                   t.lbvh_blk);
                break;
            case 11:
                hipLaunchKernelGGL(k_bucket_rank<11>, dim3(n_rank_blocks), dim3(kLbvhBlock), 0, stream, P, np,
                                   t.lbvh_blk);
                break;
            default:
                hipLaunchKernelGGL(k_bucket_rank<12>, dim3(n_rank_blocks), dim3(kLbvhBlock), 0, stream, P, np,
                                   t.lbvh_blk);
                break;
        }
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    // steps with an oversized bucket: stable radix sort and gather
    hipLaunchKernelGGL(k_sort, dim3(n_steps_all), dim3(kSortBlock), 0, stream, P, np, t.steps);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gather, dim3(n_gather_blocks), dim3(256), 0, stream, P, np, t.gather_blk);
    return hipGetLastError();
}

size_t bvh_scratch_stride(uint32_t n_objs) { return (bvh_scratch_bytes(n_objs) + 255u) / 256u * 256u; }
size_t bvh_scratch_bytes_host(uint32_t n_objs) { return bvh_scratch_bytes(n_objs); }
static_assert(kBvhLdsMaxHost == kBvhLdsMax, "BVH LDS bound");

hipError_t launch_lbvh(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_blocks, uint32_t n_steps_all,
                       hipStream_t stream)
{
    if (n_blocks) {
        hipLaunchKernelGGL(k_lbvh_build<kLbvhBlock>, dim3(n_blocks), dim3(kLbvhBlock), 0, stream, P, np, t.lbvh_blk);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (n_steps_all == 0) return hipSuccess;
    hipLaunchKernelGGL(k_lbvh_top, dim3(n_steps_all), dim3(kLbvhTopBlock), 0, stream, P, np, t.steps);
    return hipGetLastError();
}

hipError_t launch_bvh(const Prob *P, uint32_t np, const BatchTables &t, BatchCounters *bc, const LaunchShape &ls,
                      hipStream_t stream)
{
    (void)np;
    // LDS when the scratch of the largest step of the list fits; otherwise in HBM
    if (t.n_bvh_items) {
        hipLaunchKernelGGL(k_bvh<true>, dim3(std::max(1u, std::min(ls.bvh_wgs, t.n_bvh_items))), dim3(kBvhBlock),
                           t.bvh_lds_bytes, stream, P, t.bvh_items, t.n_bvh_items, &bc->next_bvh);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (t.n_bvh_hbm_items) {
        hipLaunchKernelGGL(k_bvh<false>, dim3(std::max(1u, std::min(ls.bvh_wgs, t.n_bvh_hbm_items))),
                           dim3(kBvhBlock), 0, stream, P, t.bvh_hbm_items, t.n_bvh_hbm_items, &bc->next_bvh_hbm);
        return hipGetLastError();
    }
    return hipSuccess;
}

hipError_t launch_bvh_export(const TravNode *trav, const uint32_t *tree_size, mzh_bvh_node *nodes, uint64_t node_cap,
                             uint32_t n_steps, hipStream_t stream)
{
    const uint32_t bps = (uint32_t)((node_cap + 255u) / 256u);
    if (n_steps == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bvh_export, dim3(bps * n_steps), dim3(256), 0, stream, trav, tree_size, nodes, node_cap, bps);
    return hipGetLastError();
}

hipError_t launch_bp_reset(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_nc_blocks,
                           hipStream_t stream)
{
    if (n_nc_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bp_reset, dim3(n_nc_blocks), dim3(256), 0, stream, P, np, t.nc_blk);
    return hipGetLastError();
}

hipError_t launch_bp_count(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_bp_blocks,
                           const LaunchShape &ls, hipStream_t stream)
{
    if (n_bp_blocks == 0) return hipSuccess;
    switch (ls.bp_block) {
        case 128:
            hipLaunchKernelGGL(k_bp<128>, dim3(n_bp_blocks), dim3(128), 0, stream, P, np, t.bp_blk);
            break;
        case 512:
            hipLaunchKernelGGL(k_bp<512>, dim3(n_bp_blocks), dim3(512), 0, stream, P, np, t.bp_blk);
            break;
        default:
            hipLaunchKernelGGL(k_bp<256>, dim3(n_bp_blocks), dim3(256), 0, stream, P, np, t.bp_blk);
            break;
    }
    return hipGetLastError();
}

hipError_t launch_scan_counts(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                              hipStream_t stream)
{
    hipLaunchKernelGGL(k_scan_counts, dim3(n_steps_all), dim3(1024), 0, stream, P, np, t.steps);
    return hipGetLastError();
}

hipError_t launch_scan_steps(const Prob *P, uint32_t np, hipStream_t stream)
{
    hipLaunchKernelGGL(k_scan_steps, dim3(np), dim3(1024), 0, stream, P);
    return hipGetLastError();
}

hipError_t launch_pair_prefix(const Prob *P, uint32_t np, uint64_t *pair_pre, hipStream_t stream)
{
    hipLaunchKernelGGL(k_pair_prefix, dim3(1), dim3(64), 0, stream, P, np, pair_pre);
    return hipGetLastError();
}

hipError_t launch_bp_place(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_nc_blocks,
                           hipStream_t stream)
{
    if (n_nc_blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_bp_place, dim3(n_nc_blocks), dim3(256), 0, stream, P, np, t.nc_blk);
    return hipGetLastError();
}

} // namespace mzh
