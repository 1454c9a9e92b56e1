// Host-side launchers of the gfx950 kernels (one declaration per stage).
// Every launcher enqueues on `stream` and returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mzh_device.hpp"

namespace mzh
{

// Upload-time layout transform: AoS [S][op1][7] -> step-major [S][4][op1]
// holding channels x, y, z, r, in FP64 (cf4) and rounded to FP32 (cf4f).
hipError_t launch_transpose_cf4(const double *traj, double *cf4, float *cf4f, uint64_t n_steps_total, uint32_t order,
                                hipStream_t stream);

// Upload-time bounds of every trajectory step for the mixed-precision
// pre-screen (PjView::sbnd, PjView::scal), from the coefficients of cf4 and
// the step lengths, rounded up: D3 >= max over the step of |d^3 (x, y, z) /
// dt^3|; Jv, Jd and Ja >= the jumps of position, velocity and acceleration
// from the object's previous step; and for the polyjectory (two floats, as
// their bits) lam and A with |d^a (x, y, z) / dt^a| <= A lam^a over every
// step, a = 1..order.
hipError_t launch_step_bounds(const double *cf4, const double *time, const uint64_t *step0, const uint64_t *nsteps,
                              const uint64_t *time_off, uint32_t n_objs, uint64_t n_steps_total, uint32_t order,
                              float4 *sbnd, uint32_t *scal, hipStream_t stream);

// ---- a batch of screening problems ------------------------------------------
// Conjunction steps handled by one lane of the AABB kernel.
constexpr uint32_t kAabbStepGroup = 16;

// Every stage runs as one launch over all problems of a batch (a single
// screening is a batch of one). Each problem is described to the kernels by a
// Prob record in device memory; a launch table (the exclusive prefix of the
// blocks, or work items, each problem contributes) maps a block to its
// problem and its block within the problem.

// Overflow flags of a problem (Prob::flags), set on the device when a
// capacity chosen before the launch is too small; the screening is then run
// again with the sizes the kernels counted.
constexpr uint32_t kFlagPairs = 1u;  // more pairs than pair_cap
constexpr uint32_t kFlagOvf = 2u;    // broad-phase overflow pool full
constexpr uint32_t kFlagConj = 4u;   // more conjunctions than conj_out_cap

// The broad-phase overflow pool is split into kPoolShards equal parts
// (ovf_cap / kPoolShards entries each) with a counter each, one 128-byte line
// apart: a block takes entries from the part its index picks, one atomic per
// wave, so that a step with many pairs per object does not queue every entry
// on one address. The pool a screening needs is kPoolShards times its most
// used part.
constexpr uint32_t kPoolShards = 64;
constexpr uint32_t kPoolStride = 32;

// Narrow-phase report counters of a problem (Prob::np_report): kReportSlots
// copies of the MZH_NP_COUNT counters, one 128-byte line each. A wave adds its
// sums to the copy its index picks, so that the additions of a launch spread
// over memory channels instead of queueing on one address; the host adds the
// copies up.
constexpr uint32_t kReportSlots = 64;
constexpr uint32_t kReportStride = 16;
constexpr uint32_t kReportWords = kReportSlots * kReportStride;
static_assert(MZH_NP_COUNT <= kReportStride, "report counters per copy");

struct Prob {
    PjView pj;
    gbuf<const double> cd_begin, cd_end;  // [C]
    uint32_t C, N;
    uint32_t step0;     // index of its first conjunction step among the batch's
    int32_t mode;       // MZH_MODE_FP64 / MZH_MODE_MIXED
    double radius;      // conj_thresh / 2
    double thresh;
    uint64_t node_cap;  // BVH nodes per step
    uint64_t bvh_scr_stride;
    uint64_t lbvh_stride;  // 32-bit words per step of the walk-tree build's scratch (lbvh_step_words)
    // the problem's buffers (its context's)
    gbuf<float> aabbs, srt_aabbs;  // [C][N + 1][2][4]
    gbuf<uint2> kbke;  // [C][N]
    gbuf<float4> nprec;  // [C][N][2] mixed-precision pre-screen model of each (step, object), mixed problems
    double np_e;         // mixed problems: the bound on the FP64 path's dist^2 interpolation error of a pair
    gbuf<uint32_t> gbox;  // [C][8] order-preserving images of the global box
    gbuf<uint64_t> mcodes, srt_mcodes, keys_tmp;
    gbuf<uint32_t> srt_idx, vals_tmp, n_eff, sort_todo;
    gbuf<uint32_t> bkt_end;  // [C][kBktEndStride]: the bucket sort's bucket ends per step
    gbuf<TravNode> trav;  // [C][node_cap], the breadth-first trees, made on request
    gbuf<WalkNode> wtree;  // [C][node_cap], the broad phase's walk trees
    gbuf<uint4> bvh_hdr;  // walk-tree build: [C][lbvh_stride] 32-bit words; breadth-first build: [C][node_cap]
    gbuf<uint8_t> bvh_scr;  // [C][bvh_scr_stride], when the build scratch is not in LDS
    gbuf<uint32_t> tree_size;  // [C]
    gbuf<const int32_t> otypes;  // [N]
    gbuf<uint32_t> cnt, head, slots;  // [C][N], [C][N], [C][N][4]
    gbuf<uint2> ovf;  // overflow pool {partner, next}
    gbuf<uint32_t> ovf_count;  // [kPoolShards][kPoolStride]: entries taken from each part of the pool
    uint32_t ovf_cap;
    gbuf<uint64_t> off;  // [C][N] offset of each object's pairs in its step
    gbuf<uint64_t> step_total;  // [C]
    gbuf<uint64_t> step_off;  // [C + 1]: exclusive prefix of the steps' pairs; [C] = all pairs
    gbuf<mzh_aabb_collision> pairs;  // pair_cap entries
    gbuf<uint32_t> pair_c;           // pair_cap entries: the conjunction step of each pair
    uint64_t pair_cap;
    gbuf<mzh_conj> conj_out;  // the problem's conjunctions, sorted (conj_out_cap entries)
    uint64_t conj_out_cap;
    gbuf<unsigned long long> np_report;  // [kReportSlots][kReportStride], MZH_NP_COUNT used of each
    gbuf<DevError> err;
    gbuf<uint32_t> flags;  // kFlag*
    gbuf<uint64_t> counts;  // [4]: pairs, pool entries, conjunctions, first conjunction in the batch's list
};

// Work counters shared by the launches of one batch run (device memory),
// cleared by launch_batch_init.
struct BatchCounters {
    uint32_t next_sort, next_bvh, next_bvh_hbm, pad;
};

// Exclusive prefix arrays (np + 1 entries, device memory) of one batch.
struct BatchTables {
    const uint32_t *steps;       // conjunction steps (Prob::step0)
    const uint32_t *aabb_blk;    // blocks of the AABB kernel
    const uint32_t *nc_blk;      // blocks of 256 (step, object) entries
    const uint32_t *bp_blk;      // blocks of the broad-phase walk
    const uint32_t *gather_blk;  // blocks of 256 (step, sorted position) entries of the radix-sort gather
    const uint32_t *lbvh_blk;    // blocks of kLbvhBlock sorted positions of one step (walk-tree build)
    const uint2 *sort_items;     // (problem, step) work items of the sort, largest problems first
    const uint2 *bvh_items;      // same, problems whose BVH scratch fits in LDS
    const uint2 *bvh_hbm_items;  // same, the others
    uint32_t n_sort_items, n_bvh_items, n_bvh_hbm_items;
    uint32_t bvh_lds_bytes;      // dynamic LDS of the LDS build (largest problem of its list)
};

// Launch shapes that leave the results unchanged (chosen by the tuner).
struct LaunchShape {
    uint32_t bp_block = 128;       // broad-phase walk workgroup: 128, 256 or 512
    uint32_t sort_wgs = 512;       // persistent sort workgroups
    uint32_t bvh_wgs = 512;        // persistent BVH workgroups
    uint32_t roots_wgs = 512;      // root-finding workgroups
    uint32_t bucket_bits = 12;     // bucket-sort width: 10, 11 or 12 bits
};

// Clear the per-problem counters, flags and error words, the global boxes
// and the batch's work counters. One thread per conjunction step.
hipError_t launch_batch_init(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                             BatchCounters *bc, hipStream_t stream);
// Stage 1: per-(step, object) AABBs plus the per-step global box; also
// stores, per (step, object), the range [kb, ke) of the object's trajectory
// steps overlapping the conjunction step, for the narrow phase. One launch
// per polynomial order present in the batch: problems of another order
// return at once.
// mode: the problems of that mode (MZH_MODE_FP64 or MZH_MODE_MIXED) are
// handled, the others return at once.
hipError_t launch_aabbs(uint32_t order, int mode, const Prob *P, uint32_t np, const BatchTables &t,
                        uint32_t blk_begin, uint32_t n_blocks, hipStream_t stream);
hipError_t launch_gbox_finalize(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                                hipStream_t stream);
// Stage 2: Morton codes, sort by (no data, code, index), gather of the boxes
// in sorted order (persistent workgroups over the sort items); steps with a
// bucket too large for the bucket sort go to the radix sort and gather.
hipError_t launch_morton_sort(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                              uint32_t n_gather_blocks, uint32_t n_rank_blocks, BatchCounters *bc,
                              const LaunchShape &ls, hipStream_t stream);
// Bytes of per-step BVH scratch (LDS when it fits kBvhLdsMax, else HBM).
size_t bvh_scratch_bytes_host(uint32_t n_objs);
size_t bvh_scratch_stride(uint32_t n_objs);
constexpr size_t kBvhLdsMaxHost = 64u * 1024u;
// Stage 3: the broad phase's walk tree (radix tree over runs, Karras'
// construction, mzh_lbvh.hpp) of every step: n_blocks blocks of kLbvhBlock
// positions (the lbvh_blk table), then one workgroup per step for the nodes
// that span blocks. Also counts each step's tree size.
constexpr uint32_t kLbvhBlock = 512;
constexpr uint32_t kBktEndStride = 4096;  // bucket ends per step (2^12, the widest bucket sort)
// 32-bit words per step of the walk-tree build's scratch (in the BVH scratch
// buffer) for steps of at most n_objs objects: the parent of every node, two
// counts, and the lists of spanning and frontier nodes sized by their bound.
// A node spans the boundary between two blocks only if it holds both
// positions next to it, so each boundary has at most 96 of them (one per
// prefix length of the 96-bit keys); a frontier node is a child of a spanning
// node.
__host__ __device__ inline uint64_t lbvh_span_cap(uint32_t n_objs) { return 96ull * ((n_objs + kLbvhBlock - 1u) / kLbvhBlock) + 1u; }
__host__ __device__ inline uint64_t lbvh_step_words(uint32_t n_objs)
{
    return 2ull * n_objs + 2u + lbvh_span_cap(n_objs) + (2u * lbvh_span_cap(n_objs) + 2u);
}
hipError_t launch_lbvh(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_blocks, uint32_t n_steps_all,
                       hipStream_t stream);
// The reference's breadth-first trees (persistent workgroups over the BVH
// items), made when a caller asks for a tree.
hipError_t launch_bvh(const Prob *P, uint32_t np, const BatchTables &t, BatchCounters *bc, const LaunchShape &ls,
                      hipStream_t stream);
// The reference's bvh_node records of every step's tree, from the device
// layout (one problem).
hipError_t launch_bvh_export(const TravNode *trav, const uint32_t *tree_size, mzh_bvh_node *nodes, uint64_t node_cap,
                             uint32_t n_steps, hipStream_t stream);
// Stage 4: broad phase. launch_bp_count finds every pair once (each query
// walks the tree to the right of its own code run) and counts it for the
// pair's lower original index i, keeping i's first 4 partners inline and
// pushing the others on i's list in the overflow pool. After the scans
// (per-step offsets, then the step prefix whose last entry is the problem's
// pair count), launch_bp_place writes every object's pairs sorted by partner.
// launch_bp_reset clears the counts and list heads first.
hipError_t launch_bp_reset(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_nc_blocks,
                           hipStream_t stream);
hipError_t launch_bp_count(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_bp_blocks,
                           const LaunchShape &ls, hipStream_t stream);
hipError_t launch_scan_counts(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_steps_all,
                              hipStream_t stream);
hipError_t launch_scan_steps(const Prob *P, uint32_t np, hipStream_t stream);
// Per problem: pair and pool counts into counts[0..1], kFlagPairs /
// kFlagOvf when they exceed the capacities; pair_pre (np + 1 entries) =
// exclusive prefix of the problems' placed pairs, the narrow phase's global
// pair index.
hipError_t launch_pair_prefix(const Prob *P, uint32_t np, uint64_t *pair_pre, hipStream_t stream);
hipError_t launch_bp_place(const Prob *P, uint32_t np, const BatchTables &t, uint32_t n_nc_blocks,
                           hipStream_t stream);

// Stage 5: narrow phase over the placed pairs of problems [p_first, p_end)
// (one order group: all of polynomial order `order`), pairs [pair_begin,
// pair_end) counted from the group's first, of the batch's global pair index
// (problem p's pairs are [pair_pre[p], pair_pre[p + 1]); the range is clamped
// on the device to the group's pairs). Conjunction records of
// all problems go to one list, with the batch-wide index of their step.
struct NarrowArgs {
    const Prob *P;
    uint32_t np;
    uint32_t order;
    const uint64_t *pair_pre;     // [np + 1]
    uint32_t p_first, p_end;      // the problems of the pass's order group
    uint64_t pair_begin, pair_end;  // relative to the group's first pair
    uint32_t roots_wgs;           // root-finding workgroups
    double cheby[kMaxOrder + 1];  // (cos((2i+1)pi/(2 op1)) + 1) / 2, i = 0..order
    mzh_conj *conj;               // output records (capacity conj_cap)
    uint32_t *conj_step;          // batch-wide step index of each record (Prob::step0 + step)
    uint64_t conj_cap;
    unsigned long long *conj_count;
    // overlapping trajectory-step pairs of every broad-phase pair
    void *steps;                  // StepPair records
    uint64_t steps_cap;
    unsigned long long *steps_count;
    // queue of step-pairs whose dist^2 enclosure reaches below thresh^2
    void *cand;                   // records of narrow_cand_bytes(order) bytes
    uint64_t cand_cap;
    unsigned long long *cand_count;
    // root isolation: entries of the private work list (<= 48), and the
    // candidates that outgrow it (indices into cand), isolated again with
    // kNarrowSlowLanes work lists of 251 entries in wl_slow
    int wl_fast_cap;
    uint64_t *slow;
    uint64_t slow_cap;
    unsigned long long *slow_count;
    void *wl_slow;                // narrow_wl_slow_bytes(order)
    // largest queue lengths of the passes (steps, cand, slow), for sizing
    unsigned long long *peak;
};
// One overlapping pair of trajectory steps of a broad-phase pair of problem p.
struct StepPair {
    uint32_t p, i, j, c, ki, kj;
};
// Narrow phase of the pairs [pair_begin, pair_end): enumeration, screen, root
// finding (and its slow pass), then the pass's queue lengths folded into
// peak and the queue counters cleared for the next pass.
// launch_narrow_reset clears the problems' counters, the conjunction count,
// the queue counters and peak before the first pass.
hipError_t launch_narrow_reset(const NarrowArgs &a, hipStream_t stream);
// Bytes of one queued root-finding candidate for a polynomial order.
size_t narrow_cand_bytes(uint32_t order);
// Root isolation work lists: an entry holds an interval and a polynomial; the
// slow pass runs kNarrowSlowLanes lanes, each with a list of 251 entries.
constexpr size_t narrow_wl_entry_bytes(uint32_t order) { return 8u * (order + 3u); }
constexpr uint32_t kNarrowSlowLanes = 256;
inline size_t narrow_wl_slow_bytes(uint32_t order) { return kNarrowSlowLanes * 251u * narrow_wl_entry_bytes(order); }
constexpr int kNarrowWlFastCap = 48;

// Root finding of free-standing polynomials (run_poly_root_finding with
// direction +1, poly_utils.cpp:272-460), as the narrow phase does it.
//   polys [n][order + 1], rf_int [n] (device)
//   roots [n][2 * order + 2]: the roots found; n_roots[k]: their number
//   (0xfffffffe on error: the first error is in err, with i = the
//   polynomial's index); flags[k] (may be null): bit 0 the fast exclusion
//   check fired, bit 1 the isolation ran on the slow path.
//   Scratch: slow (n entries), slow_count and err (cleared by the launch),
//   wl_slow (narrow_wl_slow_bytes(order)).
struct RootFindArgs {
    const double *polys;
    const double *rf_int;
    uint64_t n;
    double *roots;
    uint32_t *n_roots;
    uint32_t *flags;
    uint64_t *slow;
    unsigned long long *slow_count;
    void *wl_slow;
    int wl_fast_cap;
    DevError *err;
};
hipError_t launch_rootfind(uint32_t order, const RootFindArgs &a, hipStream_t stream);
template <int ORDER>
hipError_t launch_rootfind_order(const RootFindArgs &a, hipStream_t stream);
hipError_t launch_narrow(const NarrowArgs &a, hipStream_t stream);
template <int ORDER>
hipError_t launch_narrow_order(const NarrowArgs &a, hipStream_t stream);

// Group conjunction records by step and sort each step by (tca, then every
// field lexicographically): a merge sort of 32-byte keys, ping-ponging between
// keys0 and keys1 (n entries each). n_dev (may be null) holds the number of
// records on the device, at most n (the launches are sized by n). Each
// problem then takes its range of the sorted list (its steps are contiguous
// in the batch's step index) into Prob::conj_out, with counts[2..3] = its
// number of conjunctions and its first index, and kFlagConj where conj_out
// is too small.
struct ConjKey {
    uint32_t step, idx;
    uint64_t tca;  // bits of tca + 0.0
    uint64_t ij;   // i << 32 | j
    uint64_t pad;
};
hipError_t launch_conj_sort(const Prob *P, uint32_t np, const uint32_t *step_pre, const mzh_conj *in,
                            const uint32_t *in_step, uint64_t n, const unsigned long long *n_dev, ConjKey *keys0,
                            ConjKey *keys1, hipStream_t stream);

// ---- SGP4 (near-earth model) ------------------------------------------------
// Deep-space (SDP4) terms of an element set with a period of 225 minutes or
// more: the Greenwich sidereal angle at epoch, the lunar and solar periodic
// coefficients, the secular rates and the geopotential resonance terms.
struct Sdp4Terms {
    double gsto;
    double e3, ee2, peo, pgho, pho, pinco, plo, se2, se3, sgh2, sgh3, sgh4, sh2, sh3, si2, si3, sl2, sl3, sl4;
    double xgh2, xgh3, xgh4, xh2, xh3, xi2, xi3, xl2, xl3, xl4, zmol, zmos;
    double dedt, didt, dmdt, dnodt, domdt;
    double d2201, d2211, d3210, d3222, d4410, d4422, d5220, d5232, d5421, d5433;
    double del1, del2, del3, xfact, xlamo;
    int irez;  // 0 none, 1 one-day (synchronous), 2 half-day resonance
};
// The initialised model of one element set.
struct Sgp4Sat {
    double bstar, ecco, argpo, inclo, mo, nodeo, no_unkozai;
    double aycof, con41, cc1, cc4, cc5, d2, d3, d4, delmo, eta, argpdot, omgcof, sinmao, t2cof, t3cof, t4cof, t5cof,
        x1mth2, x7thm1, mdot, nodedot, xlcof, xmcof, nodecf;
    int isimp, error, deep;
    Sdp4Terms ds;  // deep != 0 only
};
hipError_t launch_sgp4_propagate(const mzh_gpe *gpes, uint64_t n_gpes, const double *tsince, uint64_t n_times,
                                 double *out, int32_t *err, hipStream_t stream);
// The generator of polyjectories from GPEs. A satellite's trajectory is a
// sequence of segments, one per GPE (from the later of the range's beginning
// and the GPE's epoch to the next GPE's epoch or the range's end); a segment
// is cut into nominal steps of a twelfth of the orbital period, each of which
// is interpolated independently (bisected where the interpolation is not
// accurate enough), one lane per nominal step.
struct Sgp4Seg {
    double t_b, t_e;    // the segment, days from the polyjectory's epoch
    double eoff;        // the GPE's epoch, days from the polyjectory's epoch
    double h0;          // nominal step, days (k_sgp4_seg_init)
    uint64_t step0;     // index of its first nominal step (host)
    uint32_t n_steps;   // nominal steps (k_sgp4_seg_init; 1 if status != 0)
    uint32_t sat, gpe;  // satellite and GPE indices
    int32_t status;     // initialisation status (0 or an SGP4 error code)
    int32_t deep;       // deep-space model (k_sgp4_seg_init)
};
struct Sgp4Gen {
    const mzh_gpe *gpes;
    Sgp4Seg *segs;
    Sgp4Sat *sats;                 // [n_segs] the initialised model of each segment
    uint32_t n_segs, n_sats;
    const uint32_t *sat_seg0;      // [n_sats + 1] segments of each satellite
    const uint64_t *seg_step0;     // [n_segs + 1] first nominal step of each segment
    uint64_t n_steps;              // nominal steps in total
    bool any_deep;                 // a segment uses the deep-space model
    double reentry_r2, exit_r2;    // squared radii (km^2)
    int32_t *step_res;             // [n_steps] polynomials of the step (> 0), or -status
    uint32_t *step_out;            // [n_steps] index of its first polynomial within the satellite, or ~0u
    uint64_t *sat_npolys;          // [n_sats]
    int32_t *sat_status;           // [n_sats]
    double *sat_tlast;             // [n_sats] end of the last polynomial
    // staging of single-piece steps (null: none): the count pass keeps
    // coefficient k of the step's polynomial in stage[k * n_steps + step]
    // (coefficient-major, so that a wave's stores are contiguous) and its end
    // time in stage_t[step + 1]; k_sgp4_copy moves them into place
    double *stage, *stage_t;
    // the kept steps to interpolate again (multi-piece ones, or all when not
    // staging), listed by k_sgp4_scan
    uint32_t *redo;
    unsigned long long *redo_count;
    // output (k_sgp4_copy, k_sgp4_write)
    double *traj, *time;
    const uint64_t *sat_poly0, *sat_time0;
};
// Initialise every segment's model, its nominal step and step count.
hipError_t launch_sgp4_seg_init(const Sgp4Gen &g, hipStream_t stream);
// Interpolate every nominal step, counting its polynomials (or its status).
hipError_t launch_sgp4_count(const Sgp4Gen &g, uint32_t order, hipStream_t stream);
// Per satellite: the first failing step ends the trajectory; offsets of the
// kept steps, polynomial counts, statuses and end times.
hipError_t launch_sgp4_scan(const Sgp4Gen &g, hipStream_t stream);
// Move the staged single-piece steps into place; interpolate the n_redo
// listed steps again and write their polynomials and times.
hipError_t launch_sgp4_write(const Sgp4Gen &g, uint32_t order, uint64_t n_redo, hipStream_t stream);

// Polyjectory evaluation: out[r][k][7] for n_rows x n_times evaluations,
// times and rows as in mzh_state_eval (time_mode = MZH_EVAL_TIME_*). The
// smallest flat index r * n_times + k of an evaluation with a non-finite time
// or an object index >= n_objs is atomicMin-ed into first_bad.
hipError_t launch_state_eval(const PjView &pj, const double *tm, int time_mode, const uint64_t *sel,
                             uint64_t n_rows, uint64_t n_times, double *out, unsigned long long *first_bad,
                             hipStream_t stream);

} // namespace mzh
