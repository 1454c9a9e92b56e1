// Device-side helpers shared by the gfx950 kernels of mizuba_amd.
//
// All kernels are compiled with -ffp-contract=off: every FP64/FP32 multiply
// and add is rounded separately, so the FP64 screen reproduces the CPU
// restatement of the reference bit for bit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/mizuba_amd.h"
#include "mzh_enclosure.hpp"
#include "mzh_lbvh.hpp"
#include "mzh_qbox.hpp"

namespace mzh
{

constexpr uint32_t kBvhMaxLevels = 72;

// Per-problem parameters passed by value to the kernels.
struct StepGrid {
    const double *cd_begin; // [C]
    const double *cd_end;   // [C]
    uint32_t n_steps;       // C
};

// Device buffers reached through a pointer held in device memory (a Prob
// record, a PjView inside one) are declared global: a pointer loaded from
// memory is otherwise generic (flat) to the compiler, and a flat access also
// counts against the LDS wait counter, so kernels that mix LDS work with such
// loads would wait on both. gl() converts a host-side device pointer for
// these fields (a no-op in the host compilation).
#if defined(__HIP_DEVICE_COMPILE__)
#define MZH_GLOBAL __attribute__((address_space(1)))
#else
#define MZH_GLOBAL
#endif
template <class T>
using gbuf = T MZH_GLOBAL *;
template <class T>
__host__ __device__ inline gbuf<T> gl(T *p)
{
    return (gbuf<T>)p;
}

// Polyjectory as laid out in HBM.
struct PjView {
    gbuf<const double> traj;      // AoS [S][op1][7]
    gbuf<const double> time;      // per object nsteps + 1 values
    gbuf<const double> cf4;       // step-major transposed [S][4][op1], channels x y z r
    gbuf<const float> cf4f;       // the same rounded to float (the mixed-precision AABBs)
    // [S]: {D3, Jv, Jd, Ja} of each step (launch_step_bounds): a bound on
    // |d^3 (x, y, z) / dt^3| over the step and the jumps of position,
    // velocity and acceleration from the object's previous step; scal: the
    // polyjectory's lam and A (float bits), |d^a (x, y, z) / dt^a| <= A lam^a
    gbuf<const float4> sbnd;
    gbuf<const uint32_t> scal;
    gbuf<const uint64_t> step0;   // global index of the object's first step
    gbuf<const uint64_t> nsteps;  // steps per object
    gbuf<const uint64_t> time_off;
    uint32_t n_objs;
    uint32_t order;
};

// BVH node of the reference's breadth-first order
// (detect_conjunctions_bvh.cpp:521-580), made on request for the tree
// accessors: {escape, begin, end or split, left}, where left is the index of
// the left child (the right child is left + 1) and 0 for a leaf, since the
// root is nobody's child; an internal node holds its split (the first
// position of its right child) in place of its end. The escape index is the
// node that follows the subtree in a depth-first walk; that of the root's
// right spine is kTravEnd.
constexpr uint32_t kTravEnd = 0xffffffffu;
struct alignas(16) TravNode {
    float lb[4];
    float ub[4];
    uint32_t escape;
    uint32_t begin;
    uint32_t end;   // leaf: end; internal node: split
    uint32_t left;
};
static_assert(sizeof(TravNode) == 48, "TravNode layout");

// Node of the broad phase's walk tree (mzh_lbvh.hpp: internal node k at 2k,
// leaf p at 2p + 1), read by the walk with two 16-byte loads: the node's box
// on the step's grid (mzh_qbox.hpp) and {escape, right child, left
// child, 1} for an internal node, {escape, begin, end, 0} for a leaf or a
// run's root.
struct alignas(16) WalkNode {
    uint32_t qb[4];
    uint32_t escape, a, b, internal;
};
static_assert(sizeof(WalkNode) == 32, "WalkNode layout");


// Error word written by kernels; the host turns it into the reference's
// exception message. Only the first error (lowest code wins ties) is kept.
struct DevError {
    int32_t code;      // 0 = none
    uint32_t i, j;     // object indices involved
    double value;      // offending value
};

enum DevErrCode : int32_t {
    kErrNone = 0,
    kErrAabbLb = 1,        // non-finite lower bound in an AABB
    kErrAabbUb = 2,        // non-finite upper bound in an AABB
    kErrGlobalBox = 3,     // non-finite global box size
    kErrNpInterval = 10,   // invalid root-finding interval
    kErrNpDist2 = 11,      // non-finite dist2 enclosure
    kErrNpRoot = 12,       // non-finite root
    kErrNpRootDer = 13,    // non-finite derivative at a root
    kErrNpIsolation = 14,  // isolation limits exceeded (value = wlist size, i/j = sizes)
    kErrNpIters = 15,      // TOMS 748 iteration limit
    kErrNpErrno = 16,      // TOMS 748 domain error
    kErrNpConjDist2 = 17,  // non-finite conjunction distance square
    kErrNpBndDerUb = 18,
    kErrNpBndDist2Ub = 19,
    kErrNpBndDerLb = 20,
    kErrNpBndDist2Lb = 21,
    kErrRfInterval = 23,   // invalid root-finding interval (value, i = polynomial)
    kErrBvhWalk = 24,      // a broad-phase walk longer than its tree (i = step, j = sorted position)
    kErrBvhBuild = 25,     // a walk-tree node with a parent or sibling outside the tree (i = step, j = node)
    kErrSortWindow = 26,   // a bucket-sort ranking block whose buckets exceed its window (i = step, j = position)
};

__device__ inline void report_error(DevError *e, int32_t code, uint32_t i, uint32_t j, double v)
{
    if (atomicCAS(&e->code, 0, code) == 0) {
        e->i = i;
        e->j = j;
        e->value = v;
    }
}

// ---- IEEE helpers --------------------------------------------------------

// bfloat16 (the top 16 bits of a float): rounded to nearest (ties to even),
// rounded up (for x >= 0), and back to float.
__device__ inline uint32_t bf16_rn(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (b + 0x7fffu + ((b >> 16) & 1u)) >> 16;
}
__device__ inline uint32_t bf16_up(float x)
{
    const uint32_t b = __float_as_uint(x);
    return (b >> 16) + ((b & 0xffffu) ? 1u : 0u);
}
__device__ inline float bf16_f(uint32_t h) { return __uint_as_float(h << 16); }

// nextafter(x, -inf) / nextafter(x, +inf) for float, exact bit stepping.
__device__ inline float next_down_f(float x)
{
    if (x != x || x == -__builtin_inff()) return x;
    if (x == 0.0f) return __uint_as_float(0x80000001u);
    uint32_t u = __float_as_uint(x);
    u = (u >> 31) ? u + 1u : u - 1u;
    return __uint_as_float(u);
}

__device__ inline float next_up_f(float x)
{
    if (x != x || x == __builtin_inff()) return x;
    if (x == 0.0f) return __uint_as_float(0x00000001u);
    uint32_t u = __float_as_uint(x);
    u = (u >> 31) ? u - 1u : u + 1u;
    return __uint_as_float(u);
}

// nextafter(x, toward) for double (finite x).
__device__ inline double next_toward_d(double x, double toward)
{
    if (x != x || toward != toward) return x + toward;
    if (x == toward) return toward;
    if (x == 0.0) {
        const double tiny = __longlong_as_double(1ll);
        return toward > 0 ? tiny : -tiny;
    }
    long long u = __double_as_longlong(x);
    if ((x < toward) == (x > 0)) {
        ++u;
    } else {
        --u;
    }
    return __longlong_as_double(u);
}

__device__ inline bool is_finite_f(float x) { return __builtin_isfinite(x); }
__device__ inline bool is_finite_d(double x) { return __builtin_isfinite(x); }

// Order-preserving float <-> uint32 map for atomic min/max on floats
// (replaces the CAS loops of detail/atomic_minmax.hpp:29-57).
__device__ inline uint32_t f2ord(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline float ord2f(uint32_t u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// Spread the low 16 bits of x so that bit b lands on bit 4b.
__device__ inline uint64_t spread4(uint64_t x)
{
    x &= 0xffffull;
    x = (x | (x << 24)) & 0x000000ff000000ffull;
    x = (x | (x << 12)) & 0x000f000f000f000full;
    x = (x | (x << 6)) & 0x0303030303030303ull;
    x = (x | (x << 3)) & 0x1111111111111111ull;
    return x;
}

// XCD-aware block order: the dispatcher hands block b to XCD b % 8, so the
// logical block id x * q + b / 8 keeps consecutive logical blocks (those of
// one conjunction step) on one XCD and its L2. The tail beyond 8 * q keeps
// its own index, which makes the map a bijection.
constexpr uint32_t kNumXcd = 8;
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nblocks)
{
    const uint32_t q = nblocks / kNumXcd;
    if (b >= q * kNumXcd) return b;
    return (b % kNumXcd) * q + b / kNumXcd;
}

// ---- wave64 primitives ----------------------------------------------------

__device__ inline uint64_t lanemask_lt()
{
    const uint32_t lane = __lane_id();
    return lane == 0 ? 0ull : (~0ull >> (64u - lane));
}

// Inclusive prefix sum across the 64 lanes of a wave.
__device__ inline uint32_t wave_inclusive_scan(uint32_t v)
{
    const uint32_t lane = __lane_id();
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(v, off, 64);
        if (lane >= (uint32_t)off) v += t;
    }
    return v;
}

__device__ inline uint64_t wave_reduce_add_u64(uint64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum over a workgroup of blockDim.x <= 1024 threads; every thread gets the
// total. s_w needs one slot per wave. Contains barriers: call from all threads.
__device__ inline unsigned long long block_reduce_add_u64(unsigned long long v, unsigned long long *s_w)
{
    v = wave_reduce_add_u64(v);
    const uint32_t w = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    __syncthreads();
    if ((threadIdx.x & 63u) == 0) s_w[w] = v;
    __syncthreads();
    unsigned long long t = 0;
    for (uint32_t k = 0; k < nw; ++k) t += s_w[k];
    return t;
}

// Exclusive prefix over a workgroup (blockDim.x <= 1024) of per-thread
// counts, and the block total. s_w needs one slot per wave + 1.
__device__ inline uint32_t block_excl_scan_u32(uint32_t v, uint32_t *s_w, uint32_t &total)
{
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6, nw = (blockDim.x + 63u) >> 6;
    const uint32_t inc = wave_inclusive_scan(v);
    __syncthreads();
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
    for (uint32_t k = 0; k < nw; ++k) {
        const uint32_t x = s_w[k];
        before += (k < w) ? x : 0u;
        tot += x;
    }
    total = tot;
    return before + inc - v;
}

} // namespace mzh
