/*
 * mzoracle -- CPU restatement of mizuba's conjunction-detection path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the correctness oracle that the
 * HIP implementation under mizuba_amd/ is checked against. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker (or as the timed CPU baseline), never as a fallback for
 * the product path.
 *
 * The reference (bluescarni/mizuba 0.0.12) cannot be built here (it needs
 * heyoka/LLVM, oneTBB, Boost, fmt, spdlog), so the oracle restates its
 * algorithms statement by statement in plain C11 + OpenMP. Each function in
 * mzoracle.c cites the reference file:line it follows. Where the reference
 * calls into a third-party library (heyoka 7.3 JIT expressions, Boost.Math
 * toms748_solve, tbb::parallel_sort) the published algorithm is restated and
 * the choice of operation order / tie rule is fixed here (see mzoracle.c).
 *
 * Layouts follow the reference exactly:
 *   trajectory step  : double[op1][7]  (coefficient-major, channels x y z vx vy vz r)
 *   aabbs            : float[C][N+1][2][4]   (row N = global box of the step)
 *   mcodes, srt_mcodes: uint64[C][N];  srt_idx: uint32[C][N]
 *   bvh_node         : 48 bytes, aabb_collision: 8 bytes, conj: 120 bytes
 */
#ifndef MZORACLE_H
#define MZORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    uint32_t begin, end;
    int32_t left, right;
    float lb[4], ub[4];
} mzo_bvh_node;

typedef struct {
    uint32_t i, j;
} mzo_aabb_collision;

typedef struct {
    uint32_t i, j;
    double tca, dca;
    double ri[3], vi[3], rj[3], vj[3];
} mzo_conj;

/* In-memory polyjectory view (polyjectory.cpp:1006-1036). traj_off is in
 * doubles into traj, time_off in doubles into time; an object with nsteps == 0
 * has no time data. */
typedef struct {
    uint32_t n_objs;
    uint32_t order;
    const double *traj;
    const double *time;
    const uint64_t *traj_off;
    const uint64_t *nsteps;
    const uint64_t *time_off;
    double maxT;
} mzo_pj;

/* Error kinds, mirroring the exception types the reference throws. */
enum { MZO_OK = 0, MZO_EINVAL = 1, MZO_ERANGE = 2, MZO_ERUNTIME = 3, MZO_ENOMEM = 4 };

/* Narrow-phase counters, same meaning and order as np_report
 * (conjunctions.hpp:150-175). */
enum {
    MZO_NP_CANDIDATES = 0,
    MZO_NP_DIST2_CHECK,
    MZO_NP_POLY_ROOTS,
    MZO_NP_FEX_CHECK,
    MZO_NP_POLY_NO_ROOTS,
    MZO_NP_DIST_MINIMA,
    MZO_NP_DISCARDED_MINIMA,
    MZO_NP_COUNT
};

typedef struct {
    int status;
    char err[512];
    uint64_t n_objs;
    uint64_t n_cd_steps;     /* total steps of the problem */
    uint64_t step_begin;     /* steps [step_begin, step_end) were computed */
    uint64_t step_end;
    int skip_cd;
    float *aabbs;            /* [ns][N+1][2][4], ns = step_end - step_begin */
    float *srt_aabbs;
    uint64_t *mcodes;        /* [ns][N] */
    uint64_t *srt_mcodes;
    uint32_t *srt_idx;
    double *cd_end_times;    /* [ns] */
    mzo_bvh_node *trees;     /* concatenated */
    uint64_t *tree_offsets;  /* [ns][2] (offset, size) */
    mzo_aabb_collision *bp;  /* concatenated */
    uint64_t *bp_offsets;    /* [ns][2] */
    mzo_conj *conj;
    uint64_t n_conj;
    uint64_t np_report[MZO_NP_COUNT];
    double stage_seconds[5]; /* aabbs, morton, bvh, broad, narrow (summed over threads) */
} mzo_result;

/* Whole pipeline over conjunction steps [step_begin, step_end) (all steps if
 * step_end == 0). Mirrors conjunctions::conjunctions (conjunctions.cpp:263-360)
 * and detect_conjunctions (detect_conjunctions.cpp:55-537). n_threads <= 0
 * uses the OpenMP default. Returns the status (also stored in out->status). */
int mzo_conjunctions(const mzo_pj *pj, double conj_thresh, double conj_det_interval, const int32_t *otypes,
                     uint64_t step_begin, uint64_t step_end, int n_threads, mzo_result *out);
void mzo_free_result(mzo_result *res);

/* ---- unit-level entry points (used by the oracle's own tests) ---- */

/* Step time range (conjunctions.cpp:373-392). Returns MZO_OK or MZO_EINVAL. */
int mzo_cd_begin_end(double maxT, uint64_t cd_idx, double conj_det_interval, uint64_t n_cd_steps, double *cbegin,
                     double *cend);
/* Taylor shift p(x) -> p(x + a) of one polynomial with stride
 * (conjunctions_jit.cpp:68-115). */
void mzo_poly_translate(const double *cfs, uint32_t order, uint32_t stride, double a, double *out, uint32_t out_stride);
/* Cargo-Shisha enclosure over [lb, ub] (conjunctions_jit.cpp:179-267). */
void mzo_cs_enclosure(const double *cfs, uint32_t order, uint32_t stride, double lb, double ub, double *out_min,
                      double *out_max);
/* Per-object AABB of one step (detect_conjunctions_aabbs.cpp:64-252);
 * returns MZO_OK or MZO_EINVAL (non-finite bound). */
int mzo_object_aabb(const mzo_pj *pj, uint32_t obj, double cbegin, double cend, double conj_thresh, float lb[4],
                    float ub[4]);
/* 16-bit quantisation + 4-D interleave (detect_conjunctions_morton.cpp:70-102). */
uint64_t mzo_disc_single_coord(float x, float min, float max);
uint64_t mzo_morton_encode4(uint64_t n0, uint64_t n1, uint64_t n2, uint64_t n3);
/* Chebyshev nodes on [0, h], ascending (conjunctions_jit.cpp:345-361). */
void mzo_cheby_nodes(uint32_t order, double h, double *nodes);
/* Bjorck-Pereyra interpolation (poly_utils.cpp:467-535). */
void mzo_vm_interp(uint32_t order, const double *alpha, const double *f, double *out);
/* Real-root isolation + refinement of poly on [0, rf_int) with direction +1
 * (poly_utils.cpp:272-460). roots receives up to `order` roots. Returns
 * MZO_OK or MZO_EINVAL; *fex is set to 1 when the fast exclusion check fired. */
int mzo_root_finding(const double *poly, uint32_t order, double rf_int, int direction, double *roots,
                     uint32_t *n_roots, int *fex);
/* Bracketed refinement (TOMS 748) of the single root of poly in [lb, ub);
 * returns 0 on success (poly_utils.cpp:213-268). */
int mzo_bracketed_root(const double *poly, uint32_t order, double lb, double ub, double *root, uint32_t *n_iter);

/* ---- polyjectory evaluation (polyjectory.cpp:1094-1366) ---- */

/* State of object `obj` at time tm (pj_eval_obj_state, polyjectory.cpp:1094-1149):
 * the 7 channels by Horner over the trajectory step holding tm (the first step
 * whose end time is > tm). NaN if the object has no data or tm is outside
 * [t_0, t_nsteps). Returns MZO_EINVAL for a non-finite tm, MZO_ERANGE for
 * obj >= n_objs (the out_of_range of polyjectory::operator[], :1009-1014). */
int mzo_state_eval(const mzo_pj *pj, uint64_t obj, double tm, double out[7]);

/* Time arguments of mzo_state_meval. */
enum { MZO_TIME_SCALAR = 0, MZO_TIME_PER_OBJECT = 1, MZO_TIME_SHARED = 2, MZO_TIME_PER_EVAL = 3 };

/* Batched evaluation (state_eval_impl / state_meval_impl, polyjectory.cpp:1151-1350):
 * out[r][k][7] = state of object sel[r] (r if sel is NULL) at the time
 * tm[0] (SCALAR), tm[r] (PER_OBJECT), tm[k] (SHARED) or tm[r * n_times + k]
 * (PER_EVAL). On error returns the status of the first failing element in
 * (r, k) order and stores its flat index r * n_times + k in *bad. */
int mzo_state_meval(const mzo_pj *pj, const double *tm, int time_mode, const uint64_t *sel, uint64_t n_rows,
                    uint64_t n_times, double *out, uint64_t *bad);

#ifdef __cplusplus
}
#endif

#endif
