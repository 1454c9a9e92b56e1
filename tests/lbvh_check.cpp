// Host build of the broad phase's walk tree (mizuba_amd/csrc/mzh_lbvh.hpp:
// Karras' construction over (code, position) keys, runs of equal codes as
// leaves, escape indices) and of its grid boxes (mzh_qbox.hpp), driven by
// tests/test_lbvh_host.py.
#include <cstdint>
#include <vector>

#include "../mizuba_amd/csrc/mzh_lbvh.hpp"
#include "../mizuba_amd/csrc/mzh_qbox.hpp"

namespace
{

// The node records as the build writes them: internal node k at 2k, leaf p
// at 2p + 1; a leaf (or a run's root) holds its range, an internal node its
// children.
struct Rec {
    uint32_t escape, a, b, internal;  // internal: a = right child, b = left child; else a = begin, b = end
};

std::vector<Rec> records(uint32_t n, const uint64_t *codes)
{
    std::vector<Rec> T(2ull * n);
    for (uint32_t p = 0; p < n; ++p) T[2ull * p + 1] = Rec{mzh::lbvh_escape(codes, n, p), p, p + 1, 0u};
    for (uint32_t k = 0; k + 1 < n; ++k) {
        const mzh::LbvhNode nd = mzh::lbvh_node(codes, n, k);
        const uint32_t esc = mzh::lbvh_escape(codes, n, nd.hi);
        T[2ull * k] = nd.run ? Rec{esc, nd.lo, nd.hi + 1, 0u}
                             : Rec{esc, mzh::lbvh_right(nd.split, nd.right_leaf), mzh::lbvh_left(nd.split, nd.left_leaf), 1u};
    }
    return T;
}

// The codes of a window [w0, w1) of positions only, as the build stages them
// in LDS: a read outside the window is counted.
struct Window {
    const uint64_t *c;
    int64_t w0, w1;
    mutable uint64_t outside = 0;
    uint64_t operator()(uint32_t i) const
    {
        if ((int64_t)i < w0 || (int64_t)i >= w1) ++outside;
        return c[i];
    }
};

} // namespace

// Walks the tree of the n sorted codes as the broad phase does, descending
// into every internal node (to its left child, whose begin and the right
// child's give the split), and records each node it reaches in walk order:
// begin, end (exclusive), split (0 for a leaf). Returns the number of nodes,
// or -1 when a walk step leaves the node array or the walk does not end
// within 2n steps (a broken escape chain).
extern "C" int64_t lbvh_walk(uint32_t n, const uint64_t *codes, uint32_t *begin, uint32_t *end, uint32_t *split)
{
    if (n == 0) return 0;
    const std::vector<Rec> T = records(n, codes);
    // ranges of every record, for the internal nodes' begin and end
    std::vector<uint32_t> lo(2ull * n), hi(2ull * n);
    for (uint32_t p = 0; p < n; ++p) lo[2ull * p + 1] = p, hi[2ull * p + 1] = p + 1;
    for (uint32_t k = 0; k + 1 < n; ++k) {
        const mzh::LbvhNode nd = mzh::lbvh_node(codes, n, k);
        lo[2ull * k] = nd.lo;
        hi[2ull * k] = nd.hi + 1;
    }
    int64_t m = 0;
    uint32_t i = mzh::lbvh_root(n);
    for (uint64_t guard = 0; i != mzh::kLbvhEnd; ++guard) {
        if (guard > 2ull * n || i >= 2ull * n) return -1;
        const Rec &r = T[i];
        begin[m] = lo[i];
        end[m] = hi[i];
        split[m] = r.internal ? r.a >> 1 : 0u;
        if (r.internal && (lo[r.b] != lo[i] || hi[r.b] != r.a >> 1 || lo[r.a] != r.a >> 1 || hi[r.a] != hi[i]))
            return -1;  // children that do not tile the node
        ++m;
        i = r.internal ? r.b : r.escape;
    }
    return m;
}

// The escape index and range of every internal node, for the checks that the
// escape of a node is the node after its subtree: lo, hi (inclusive), escape,
// split, run flag and the indices of both children.
extern "C" void lbvh_nodes(uint32_t n, const uint64_t *codes, uint32_t *lo, uint32_t *hi, uint32_t *esc,
                           uint32_t *split, uint8_t *run, uint32_t *left, uint32_t *right)
{
    for (uint32_t k = 0; k + 1 < n; ++k) {
        const mzh::LbvhNode nd = mzh::lbvh_node(codes, n, k);
        lo[k] = nd.lo;
        hi[k] = nd.hi;
        esc[k] = mzh::lbvh_escape(codes, n, nd.hi);
        split[k] = nd.split;
        run[k] = nd.run ? 1 : 0;
        left[k] = mzh::lbvh_left(nd.split, nd.left_leaf);
        right[k] = mzh::lbvh_right(nd.split, nd.right_leaf);
    }
}

// The build's block form: for every block [B0, B1) of bs positions and every
// internal node k in it, lbvh_leaves must say whether k's range leaves the
// block, and for a node that stays inside, lbvh_node searched over the block
// (and the node's and its leaves' escapes) must equal the full computation
// while reading codes only in [B0 - 1, B1 + 2). Returns the number of
// mismatches.
extern "C" int64_t lbvh_block_check(uint32_t n, const uint64_t *codes, uint32_t bs)
{
    int64_t bad = 0;
    for (uint32_t B0 = 0; B0 < n; B0 += bs) {
        const uint32_t B1 = n - B0 < bs ? n : B0 + bs;
        Window w{codes, (int64_t)B0 - 1, (int64_t)B1 + 2};
        for (uint32_t p = B0; p < B1; ++p) {
            if (mzh::lbvh_escape(w, n, p) != mzh::lbvh_escape(codes, n, p)) ++bad;
            if (p + 1 >= n) continue;
            const mzh::LbvhNode f = mzh::lbvh_node(codes, n, p);
            const bool inside = f.lo >= B0 && f.hi < B1;
            if (mzh::lbvh_leaves(w, n, p, B0, B1) == inside) ++bad;
            if (!inside) continue;
            const mzh::LbvhNode r = mzh::lbvh_node(w, n, p, B0, B1);
            if (r.lo != f.lo || r.hi != f.hi || r.split != f.split || r.left_leaf != f.left_leaf
                || r.right_leaf != f.right_leaf || r.run != f.run)
                ++bad;
            if (mzh::lbvh_escape(w, n, r.hi) != mzh::lbvh_escape(codes, n, f.hi)) ++bad;
        }
        if (w.outside) bad += (int64_t)w.outside;
    }
    return bad;
}

// The grid boxes of n float boxes (lo[4], hi[4] each, inside or on the global
// box glo, ghi) against the float tests, over every pair (a, b): a pair whose
// closed boxes meet must meet on the grid, and the grid box of the union of
// two boxes must be the union of their grid boxes. Returns the number of
// violations; meets_grid counts the pairs that meet on the grid only.
extern "C" int64_t qbox_check(uint32_t n, const float *boxes, const float *g, int64_t *meets_float, int64_t *meets_grid)
{
    const float4 glo = make_float4(g[0], g[1], g[2], g[3]), ghi = make_float4(g[4], g[5], g[6], g[7]);
    const mzh::QGrid grid = mzh::qgrid(glo, ghi);
    auto lo = [&](uint32_t i) { return make_float4(boxes[8 * i], boxes[8 * i + 1], boxes[8 * i + 2], boxes[8 * i + 3]); };
    auto hi = [&](uint32_t i) {
        return make_float4(boxes[8 * i + 4], boxes[8 * i + 5], boxes[8 * i + 6], boxes[8 * i + 7]);
    };
    int64_t bad = 0, mf = 0, mg = 0;
    for (uint32_t a = 0; a < n; ++a) {
        const uint4 qa = mzh::qbox(grid, lo(a), hi(a));
        for (uint32_t b = 0; b < n; ++b) {
            const uint4 qb = mzh::qbox(grid, lo(b), hi(b));
            bool meet = true;
            for (int d = 0; d < 4; ++d) meet = meet && boxes[8 * a + 4 + d] >= boxes[8 * b + d] && boxes[8 * a + d] <= boxes[8 * b + 4 + d];
            const bool gm = mzh::qbox_overlap(qa, qb);
            mf += meet;
            mg += gm && !meet;
            if (meet && !gm) ++bad;
            float4 ul = lo(a), uh = hi(a);
            const float4 bl = lo(b), bh = hi(b);
            ul = make_float4(fminf(ul.x, bl.x), fminf(ul.y, bl.y), fminf(ul.z, bl.z), fminf(ul.w, bl.w));
            uh = make_float4(fmaxf(uh.x, bh.x), fmaxf(uh.y, bh.y), fmaxf(uh.z, bh.z), fmaxf(uh.w, bh.w));
            const uint4 qu = mzh::qbox(grid, ul, uh), qv = mzh::qbox_union(qa, qb);
            if (qu.x != qv.x || qu.y != qv.y || qu.z != qv.z || qu.w != qv.w) ++bad;
        }
    }
    *meets_float = mf;
    *meets_grid = mg;
    return bad;
}
