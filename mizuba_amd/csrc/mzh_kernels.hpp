// Host-side launchers of the gfx950 kernels (one declaration per stage).
// Every launcher enqueues on `stream` and returns the hipError_t of the launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "mzh_device.hpp"

namespace mzh
{

// Upload-time layout transform: AoS [S][op1][7] -> step-major [S][4][op1]
// holding channels x, y, z, r.
hipError_t launch_transpose_cf4(const double *traj, double *cf4, uint64_t n_steps_total, uint32_t order,
                                hipStream_t stream);

// Stage 1: per-(step, object) AABBs plus the per-step global box.
// mode: MZH_MODE_FP64 or MZH_MODE_MIXED.
// Also stores, per (step, object), the range [kb, ke) of the object's
// trajectory steps overlapping the conjunction step, for the narrow phase.
hipError_t launch_aabbs(const PjView &pj, const StepGrid &sg, double conj_thresh, int mode, float *aabbs,
                        uint2 *kbke, uint32_t *gbox_ord, DevError *err, hipStream_t stream);
hipError_t launch_gbox_finalize(const uint32_t *gbox_ord, float *aabbs, uint32_t n_objs, uint32_t n_steps,
                                DevError *err, hipStream_t stream);

// Stage 2: Morton codes, sort by (no data, code, index), gather of the boxes
// in sorted order; n_eff = number of objects with data per step. todo
// (one word per step) marks the steps left to the radix sort fallback.
hipError_t launch_morton_sort(const float *aabbs, uint64_t *mcodes, uint64_t *keys_tmp, uint32_t *vals_tmp,
                              uint64_t *srt_mcodes, uint32_t *srt_idx, float *srt_aabbs, uint32_t *n_eff,
                              uint32_t *todo, uint32_t n_objs, uint32_t n_steps, hipStream_t stream);

// Stage 3: BVH, one workgroup per step. nodes has node_cap entries per step.
// Also writes the depth-first traversal copy `trav` (node_cap per step);
// nolc_buf, ps_buf and pos_buf are node_cap-sized scratch per step.
// Bytes of per-step BVH scratch in HBM (used when it does not fit in LDS).
size_t bvh_scratch_stride(uint32_t n_objs);
// hdr: node_cap uint4 headers per step, scratch of the build.
hipError_t launch_bvh(const uint64_t *srt_mcodes, const float *srt_aabbs, const uint32_t *n_eff, TravNode *trav,
                      uint4 *hdr, uint8_t *scr_buf, uint32_t *tree_size, uint32_t n_objs, uint64_t node_cap,
                      uint32_t n_steps, hipStream_t stream);
// The reference's bvh_node records of every step's tree, from the device layout.
hipError_t launch_bvh_export(const TravNode *trav, const uint32_t *tree_size, mzh_bvh_node *nodes, uint64_t node_cap,
                             uint32_t n_steps, hipStream_t stream);

// Stage 4: broad phase. launch_bp_count finds every pair once (each query
// walks the tree to the right of its own code run) and counts it for the
// pair's lower original index i, keeping i's first 4 partners inline (slots: 4
// per object and step) and pushing the others on i's list in the overflow
// pool (ovf entries {partner, next}, head per object and step). After the
// scans, launch_bp_place writes every object's pairs sorted by partner.
// launch_bp_reset clears the counts, list heads and pool counter first.
hipError_t launch_bp_reset(uint32_t *cnt, uint32_t *head, uint32_t *ovf_count, uint32_t n_objs, uint32_t n_steps,
                           hipStream_t stream);
hipError_t launch_bp_count(const TravNode *nodes, const uint32_t *tree_size, const float *srt_aabbs,
                           const uint64_t *srt_mcodes, const uint32_t *srt_idx, const uint32_t *n_eff,
                           const int32_t *otypes, uint32_t *cnt, uint32_t *slots, uint2 *ovf, uint32_t *head,
                           uint32_t *ovf_count, uint32_t ovf_cap, uint32_t n_objs, uint64_t node_cap, uint32_t n_steps,
                           hipStream_t stream);
hipError_t launch_scan_counts(const uint32_t *cnt, uint64_t *off, uint64_t *step_total, uint32_t n_objs,
                              uint32_t n_steps, hipStream_t stream);
hipError_t launch_scan_steps(const uint64_t *step_total, uint64_t *step_off, uint64_t *grand_total, uint32_t n_steps,
                             hipStream_t stream);
hipError_t launch_bp_place(const uint32_t *cnt, const uint32_t *slots, const uint2 *ovf, const uint32_t *head,
                           uint32_t ovf_cap, const uint64_t *off, const uint64_t *step_off, mzh_aabb_collision *pairs,
                           uint32_t n_objs, uint32_t n_steps, hipStream_t stream);

// Stage 5: narrow phase over the broad-phase pairs [pair_begin, pair_end) of
// the concatenated pair list (all steps).
struct NarrowArgs {
    PjView pj;
    StepGrid sg;
    const uint2 *kbke;  // [C][N] overlapping trajectory-step ranges from the AABB stage
    const mzh_aabb_collision *pairs;
    const uint64_t *step_off;  // [C + 1], exclusive prefix of the per-step pair counts
    uint64_t pair_begin, pair_end;
    double conj_thresh;
    double cheby[kMaxOrder + 1];  // (cos((2i+1)pi/(2 op1)) + 1) / 2, i = 0..order
    mzh_conj *conj;               // output records (capacity conj_cap)
    uint32_t *conj_step;          // step index of each record
    uint64_t conj_cap;
    unsigned long long *conj_count;
    unsigned long long *np_report;  // [MZH_NP_COUNT]
    DevError *err;
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
};
// One overlapping pair of trajectory steps of a broad-phase pair.
struct StepPair {
    uint32_t i, j, c, ki, kj;
};
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
// keys0 and keys1 (n entries each); the sorted records go to out.
struct ConjKey {
    uint32_t step, idx;
    uint64_t tca;  // bits of tca + 0.0
    uint64_t ij;   // i << 32 | j
    uint64_t pad;
};
hipError_t launch_conj_sort(const mzh_conj *in, const uint32_t *in_step, uint64_t n, ConjKey *keys0, ConjKey *keys1,
                            mzh_conj *out, hipStream_t stream);

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
