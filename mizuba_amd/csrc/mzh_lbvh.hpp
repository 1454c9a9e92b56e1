// The broad phase's walk tree of one conjunction step, node by node (host and
// device code: the kernels include it through mzh_device.hpp, and a CPU test
// compiles it on its own to compare it with the reference's tree).
//
// The tree is the binary radix tree of the step's sorted Morton codes
// (detect_conjunctions_bvh.cpp:267-699): a node whose range holds different
// codes splits where its highest differing bit flips, and a run of equal codes
// is one leaf. Here it is built with Karras' construction (T. Karras, "Maximizing
// parallelism in the construction of BVHs, octrees, and k-d trees", HPG 2012),
// which finds every internal node independently: over the n keys
// (code, position), all distinct, internal node k (0 <= k <= n - 2) covers a
// range with one end at k, and leaf p is position p. A run of equal codes is a
// subtree of this tree whose root covers exactly the run; the walk stops there
// and treats it as the run's leaf, so the nodes the walk can reach are the
// reference's nodes, numbered differently: internal node k at index 2k, leaf p
// at index 2p + 1. The root is internal node 0 (leaf 0 when n == 1).
//
// The escape index of a node (the node that follows its subtree in a
// depth-first walk) depends only on the position right after its range.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define MZH_LBVH_HD __host__ __device__
#else
#define MZH_LBVH_HD
#endif

namespace mzh
{

constexpr uint32_t kLbvhEnd = 0xffffffffu;  // escape of the root's right spine

MZH_LBVH_HD inline int lbvh_clz64(uint64_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll((long long)x);
#else
    return __builtin_clzll(x);
#endif
}

MZH_LBVH_HD inline int lbvh_clz32(uint32_t x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __clz((int)x);
#else
    return __builtin_clz(x);
#endif
}

// The codes as a function of the position: a plain array, or a window of
// them staged elsewhere (the build keeps a block's codes in LDS).
struct LbvhArray {
    const uint64_t *c;
    MZH_LBVH_HD uint64_t operator()(uint32_t i) const { return c[i]; }
};

// Length of the common prefix of the keys (codes[i], i) and (codes[j], j):
// the Morton code's 64 bits, then the position's 32 bits; -1 when j is
// outside [s0, s1).
template <class Code>
MZH_LBVH_HD inline int lbvh_delta(const Code &code, int64_t s0, int64_t s1, uint32_t i, int64_t j)
{
    if (j < s0 || j >= s1) return -1;
    const uint64_t x = code(i) ^ code((uint32_t)j);
    if (x) return lbvh_clz64(x);
    return 64 + lbvh_clz32(i ^ (uint32_t)j);
}

// Internal node k of the radix tree over the n keys: its range [lo, hi]
// (inclusive), its split (first position of the right child) and whether
// each child is a leaf. run: the range holds one code only (the node is, or
// lies inside, a run's leaf).
struct LbvhNode {
    uint32_t lo, hi, split;
    bool left_leaf, right_leaf, run;
};

// The direction of internal node k (+1: its range starts at k) and the
// common prefix of key k with its neighbour on the other side.
template <class Code>
MZH_LBVH_HD inline int64_t lbvh_dir(const Code &code, uint32_t n, uint32_t k, int &dmin)
{
    const int dr = lbvh_delta(code, 0, n, k, (int64_t)k + 1), dl = lbvh_delta(code, 0, n, k, (int64_t)k - 1);
    dmin = dr > dl ? dl : dr;
    return dr > dl ? 1 : -1;  // the range extends towards the longer common prefix
}

// True when the range of internal node k leaves the positions [s0, s1)
// (which hold k): the first key beyond them on the node's side shares more
// than dmin bits with key k. Reads the codes at k - 1, k, k + 1 and s0 - 1 or
// s1.
template <class Code>
MZH_LBVH_HD inline bool lbvh_leaves(const Code &code, uint32_t n, uint32_t k, int64_t s0, int64_t s1, int64_t d,
                                    int dmin)
{
    return lbvh_delta(code, 0, n, k, d > 0 ? s1 : s0 - 1) > dmin;
}

template <class Code>
MZH_LBVH_HD inline bool lbvh_leaves(const Code &code, uint32_t n, uint32_t k, int64_t s0, int64_t s1)
{
    int dmin;
    const int64_t d = lbvh_dir(code, n, k, dmin);
    return lbvh_leaves(code, n, k, s0, s1, d, dmin);
}

// Internal node k, its other end and split searched over [s0, s1) only:
// exact when the node's range lies inside [s0, s1) (lbvh_leaves false), since
// the common prefix with key k only shrinks away from k. Reads the codes at
// k - 1, k + 1 and inside [s0, s1).
// (the same with k's direction d and prefix dmin known, from lbvh_dir)
template <class Code>
MZH_LBVH_HD inline LbvhNode lbvh_node(const Code &code, uint32_t n, uint32_t k, int64_t s0, int64_t s1, int64_t d,
                                      int dmin)
{
    // the other end: exponential then binary search for the farthest key
    // sharing more than dmin bits with key k
    uint64_t lmax = 2;
    while (lbvh_delta(code, s0, s1, k, (int64_t)k + (int64_t)lmax * d) > dmin) lmax <<= 1;
    uint64_t l = 0;
    for (uint64_t t = lmax >> 1; t >= 1; t >>= 1)
        if (lbvh_delta(code, s0, s1, k, (int64_t)k + (int64_t)(l + t) * d) > dmin) l += t;
    const int64_t j = (int64_t)k + (int64_t)l * d;
    const int dnode = lbvh_delta(code, s0, s1, k, j);
    // the split: the last key, from k, sharing more than dnode bits with key k
    uint64_t s = 0;
    for (uint64_t div = 2;; div <<= 1) {
        const uint64_t t = (l + div - 1) / div;
        if (lbvh_delta(code, s0, s1, k, (int64_t)k + (int64_t)(s + t) * d) > dnode) s += t;
        if (t <= 1) break;
    }
    const int64_t gamma = (int64_t)k + (int64_t)s * d + (d < 0 ? -1 : 0);
    LbvhNode r;
    r.lo = (uint32_t)(d > 0 ? (int64_t)k : j);
    r.hi = (uint32_t)(d > 0 ? j : (int64_t)k);
    r.split = (uint32_t)(gamma + 1);
    r.left_leaf = (int64_t)r.lo == gamma;
    r.right_leaf = (int64_t)r.hi == gamma + 1;
    r.run = code(r.lo) == code(r.hi);
    return r;
}

template <class Code>
MZH_LBVH_HD inline LbvhNode lbvh_node(const Code &code, uint32_t n, uint32_t k, int64_t s0, int64_t s1)
{
    int dmin;
    const int64_t d = lbvh_dir(code, n, k, dmin);
    return lbvh_node(code, n, k, s0, s1, d, dmin);
}

MZH_LBVH_HD inline LbvhNode lbvh_node(const uint64_t *codes, uint32_t n, uint32_t k)
{
    return lbvh_node(LbvhArray{codes}, n, k, 0, n);
}

// Escape index of a node whose range ends at position hi (inclusive): the
// right child of the node that splits between hi and hi + 1, which is
// internal node hi + 1 when that node's range starts at hi + 1, else leaf
// hi + 1. Reads the codes at hi, hi + 1 and hi + 2.
template <class Code>
MZH_LBVH_HD inline uint32_t lbvh_escape(const Code &code, uint32_t n, uint32_t hi)
{
    if ((uint64_t)hi + 1u >= n) return kLbvhEnd;
    const uint32_t k = hi + 1u;
    if ((uint64_t)k + 1u >= n) return 2u * k + 1u;
    int dmin;
    return lbvh_dir(code, n, k, dmin) > 0 ? 2u * k : 2u * k + 1u;
}

MZH_LBVH_HD inline uint32_t lbvh_escape(const uint64_t *codes, uint32_t n, uint32_t hi)
{
    return lbvh_escape(LbvhArray{codes}, n, hi);
}

MZH_LBVH_HD inline uint32_t lbvh_root(uint32_t n) { return n >= 2u ? 0u : 1u; }

// Children of an internal node in the walk's numbering.
MZH_LBVH_HD inline uint32_t lbvh_left(uint32_t split, bool left_leaf) { return 2u * (split - 1u) + (left_leaf ? 1u : 0u); }
MZH_LBVH_HD inline uint32_t lbvh_right(uint32_t split, bool right_leaf) { return 2u * split + (right_leaf ? 1u : 0u); }

} // namespace mzh
