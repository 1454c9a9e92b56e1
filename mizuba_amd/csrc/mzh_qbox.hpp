// Boxes of the broad phase's walk tree on a grid over the step's global box
// (host and device code: the kernels include it through mzh_device.hpp, and a
// CPU test compiles it on its own).
//
// The grid has 65535 cells per dimension: dimension d of a box is the word
// lo | hi << 16, lo and hi being floor and ceil of
// f(x) = clamp((x - g0_d) * s_d, 0, 65535), s_d = 65535 / (g1_d - g0_d) (0 for
// an empty or non-finite extent). f is non-decreasing, so closed intervals
// that meet keep meeting on the grid (a <= b gives floor f(a) <= ceil f(b)),
// and the union of two grid boxes (min of the lows, max of the highs) is the
// grid box of the union: a node's grid box contains the grid box of every
// object under it, and a walk that tests grid boxes misses no pair. A leaf the
// grid says meets the query is checked again with the objects' float boxes.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MZH_QBOX_HD __host__ __device__
#else
#define MZH_QBOX_HD
struct float4 {
    float x, y, z, w;
};
struct uint4 {
    uint32_t x, y, z, w;
};
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline uint4 make_uint4(uint32_t x, uint32_t y, uint32_t z, uint32_t w) { return uint4{x, y, z, w}; }
#endif

namespace mzh
{

// x - y and x * y rounded to nearest, never fused
MZH_QBOX_HD inline float qbox_sub(float x, float y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __fsub_rn(x, y);
#else
    return x - y;
#endif
}

MZH_QBOX_HD inline float qbox_mul(float x, float y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __fmul_rn(x, y);
#else
    return x * y;
#endif
}

struct QGrid {
    float4 g0, s;
};

MZH_QBOX_HD inline float qgrid_scale(float lo, float hi)
{
    const float s = 65535.0f / qbox_sub(hi, lo);
    return (std::isfinite(s) && s > 0.0f) ? s : 0.0f;
}

MZH_QBOX_HD inline QGrid qgrid(const float4 &glo, const float4 &ghi)
{
    QGrid g;
    g.g0 = glo;
    g.s = make_float4(qgrid_scale(glo.x, ghi.x), qgrid_scale(glo.y, ghi.y), qgrid_scale(glo.z, ghi.z),
                      qgrid_scale(glo.w, ghi.w));
    return g;
}

MZH_QBOX_HD inline uint32_t qdim(float lo, float hi, float g0, float s)
{
    const float fl = fminf(fmaxf(qbox_mul(qbox_sub(lo, g0), s), 0.0f), 65535.0f);
    const float fh = fminf(fmaxf(qbox_mul(qbox_sub(hi, g0), s), 0.0f), 65535.0f);
    return (uint32_t)floorf(fl) | (uint32_t)ceilf(fh) << 16;
}

MZH_QBOX_HD inline uint4 qbox(const QGrid &g, const float4 &lo, const float4 &hi)
{
    return make_uint4(qdim(lo.x, hi.x, g.g0.x, g.s.x), qdim(lo.y, hi.y, g.g0.y, g.s.y),
                      qdim(lo.z, hi.z, g.g0.z, g.s.z), qdim(lo.w, hi.w, g.g0.w, g.s.w));
}

MZH_QBOX_HD inline uint32_t qdim_union(uint32_t a, uint32_t b)
{
    const uint32_t lo = (a & 0xffffu) < (b & 0xffffu) ? (a & 0xffffu) : (b & 0xffffu);
    const uint32_t hi = (a >> 16) > (b >> 16) ? (a >> 16) : (b >> 16);
    return lo | hi << 16;
}

MZH_QBOX_HD inline uint4 qbox_union(const uint4 &a, const uint4 &b)
{
    return make_uint4(qdim_union(a.x, b.x), qdim_union(a.y, b.y), qdim_union(a.z, b.z), qdim_union(a.w, b.w));
}

// (bitwise ands: every word is read on every path, so that a box is loaded
// with one instruction)
MZH_QBOX_HD inline uint32_t qdim_meet(uint32_t a, uint32_t b)
{
    return (uint32_t)((a & 0xffffu) <= (b >> 16)) & (uint32_t)((b & 0xffffu) <= (a >> 16));
}

MZH_QBOX_HD inline bool qbox_overlap(const uint4 &a, const uint4 &b)
{
    return (qdim_meet(a.x, b.x) & qdim_meet(a.y, b.y) & qdim_meet(a.z, b.z) & qdim_meet(a.w, b.w)) != 0u;
}

} // namespace mzh
