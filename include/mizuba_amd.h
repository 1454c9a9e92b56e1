/*
 * mizuba_amd C ABI -- MI355X (gfx950) conjunction screening over a polyjectory.
 *
 * This is the drop-in boundary for the path
 *     conjunctions(pj, conj_thresh, conj_det_interval, otypes)
 * of bluescarni/mizuba 0.0.12. The reference has no plugin registry: the path
 * sits behind the C++ class `mizuba::conjunctions` (mizuba/conjunctions.hpp:57-237),
 * whose constructor runs `detect_conjunctions` (mizuba/detect_conjunctions.cpp:55-537)
 * and whose accessors expose the per-step arrays. Each entry point below names
 * the reference interface it replaces.
 *
 * Conventions
 *  - Plain C types only; every entry point returns an int status (MZH_OK on
 *    success). The message of the last failure of a context is read with
 *    mzh_last_error(); errors raised before a context exists are read with
 *    mzh_global_error(). No C++ exception crosses this boundary.
 *  - Status kinds mirror the exception types of the reference:
 *    MZH_EINVAL  <-> std::invalid_argument (ValueError in Python),
 *    MZH_ERANGE  <-> std::out_of_range     (IndexError),
 *    MZH_ERUNTIME<-> std::runtime_error / HIP failures (RuntimeError).
 *  - A context owns one HIP stream on one device and all device buffers of
 *    one problem. Calls on one context must not overlap; different contexts
 *    may be used concurrently from different threads.
 *  - Layouts and record sizes are the reference's:
 *      aabbs, srt_aabbs : float   [C][N+1][2][4]   (row N = global box)
 *      mcodes, srt_mcodes: uint64 [C][N];  srt_idx: uint32 [C][N]
 *      cd_end_times     : double  [C]
 *      mzh_bvh_node 48 B, mzh_aabb_collision 8 B, mzh_conj 120 B.
 */
#ifndef MIZUBA_AMD_H
#define MIZUBA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MZH_OK 0
#define MZH_EINVAL 1
#define MZH_ERANGE 2
#define MZH_ERUNTIME 3
#define MZH_ENOMEM 4
/* A screening too large for one batch (objects x conjunction steps above
 * 2^32 - 1): std::invalid_argument (ValueError); mzh_screen_steps screens
 * such a problem in ranges of steps. */
#define MZH_ESIZE 5

/* Precision modes of the screen (AABB -> broad phase).
 * MZH_MODE_FP64 : drop-in arithmetic of the reference -- FP64 Cargo-Shisha
 *                 enclosures converted outward to FP32 boxes; bit-exact with
 *                 the CPU restatement.
 * MZH_MODE_MIXED: FP32 enclosures with a rigorous a-priori error margin; every
 *                 box contains the FP64-mode box, so the candidate set is a
 *                 superset and the FP64 narrow phase returns the same
 *                 conjunctions. */
#define MZH_MODE_FP64 0
#define MZH_MODE_MIXED 1

/* conjunctions::bvh_node (mizuba/conjunctions.hpp:65-72). */
typedef struct {
    uint32_t begin, end;
    int32_t left, right;
    float lb[4], ub[4];
} mzh_bvh_node;

/* conjunctions::aabb_collision (mizuba/conjunctions.hpp:75-79). */
typedef struct {
    uint32_t i, j;
} mzh_aabb_collision;

/* conjunctions::conj (mizuba/conjunctions.hpp:82-94). */
typedef struct {
    uint32_t i, j;
    double tca, dca;
    double ri[3], vi[3], rj[3], vj[3];
} mzh_conj;

/* Narrow-phase counters in the order of conjunctions::np_report
 * (mizuba/conjunctions.hpp:150-175). */
#define MZH_NP_COUNT 7

/* Stage timings reported by mzh_get_stage_times (seconds, device clock):
 * aabbs, morton (codes + sort + gather), bvh, broad, narrow, total, then the
 * AABB kernel alone, the narrow-phase kernels alone and the broad-phase
 * traversal kernel alone. */
#define MZH_STAGE_COUNT 9

typedef struct mzh_ctx mzh_ctx;

/* ---- context ---------------------------------------------------------- */

/* Create a context on HIP device `device` with its own stream. No reference
 * counterpart (the reference runs on the host); it stands for the per-call
 * state that conjunctions::conjunctions builds (mizuba/conjunctions.cpp:263-360). */
int mzh_create(int device, mzh_ctx **out);
int mzh_destroy(mzh_ctx *ctx);
/* Copy the message of the last failure on ctx into buf (NUL-terminated). */
int mzh_last_error(const mzh_ctx *ctx, char *buf, size_t len);
/* Message of the last failure of mzh_create (thread-local). */
int mzh_global_error(char *buf, size_t len);
/* Number of HIP devices visible (0 when none). */
int mzh_device_count(int *count);

/* ---- input: the polyjectory -------------------------------------------- */

/* Upload a polyjectory to HBM. Replaces the reads through polyjectory::operator[]
 * (mizuba/polyjectory.cpp:1006-1036) that every stage of the reference does:
 *   traj      : double[S_tot][order+1][7], all objects' steps concatenated
 *   time      : double[T_tot], per object nsteps+1 values (none if nsteps == 0)
 *   traj_off  : per object, offset in doubles of its first step in traj
 *   nsteps    : per object, number of trajectory steps
 *   time_off  : per object, offset in doubles of its first time in time
 *   maxT      : largest final time over all objects (polyjectory::get_maxT)
 * The data must already satisfy the polyjectory invariants (finite values,
 * strictly increasing non-negative times, order >= 2); the host front end
 * checks them with the reference's messages. The upload also builds the
 * step-major transposed coefficient copy used by the AABB kernels. */
int mzh_upload_polyjectory(mzh_ctx *ctx, uint32_t n_objs, uint32_t order, const double *traj, uint64_t traj_len,
                           const double *time, uint64_t time_len, const uint64_t *traj_off, const uint64_t *nsteps,
                           const uint64_t *time_off, double maxT);

/* ---- polyjectories from SGP4 element sets ------------------------------- */

/* A general perturbations element set (mizuba/make_sgp4_polyjectory.hpp:35-47):
 * UTC Julian epoch as two parts, mean motion (rad/min), eccentricity,
 * inclination, right ascension of the ascending node, argument of perigee,
 * mean anomaly (rad) and the B* drag term (1 / Earth radii). */
typedef struct mzh_gpe {
    uint64_t norad_id;
    double epoch_jd1, epoch_jd2;
    double n0, e0, i0, node0, omega0, m0, bstar;
} mzh_gpe;

/* Satellite status codes of a generated polyjectory, beside the SGP4 error
 * codes 1-6 (make_sgp4_polyjectory.cpp:478-913): */
#define MZH_SGP4_NONFINITE 10  /* non-finite values in the interpolation */
#define MZH_SGP4_REENTRY 11    /* below the reentry radius */
#define MZH_SGP4_EXIT 12       /* at or beyond the exit radius */
#define MZH_SGP4_INTERPOLATION 14 /* more than 20 pieces needed for one nominal step */

/* SGP4 / SDP4 (WGS-72, AFSPC mode) states of n_gpes element sets at n_times
 * times each: out[g][k] = {x, y, z, vx, vy, vz} (km, km/s, TEME) at tsince[g][k]
 * minutes from the epoch of gpes[g], err[g][k] = the SGP4 error code (0: none;
 * an error of the propagation to the epoch, which the published
 * initialisation performs, is reported for every time). The arithmetic is that of the published SGP4 code, including its
 * stopping rule for Kepler's equation. Host pointers. No reference
 * counterpart as a function: it is the propagation step of
 * make_sgp4_polyjectory (make_sgp4_polyjectory.cpp:722-758, 822-835). */
int mzh_sgp4_propagate(mzh_ctx *ctx, const mzh_gpe *gpes, uint64_t n_gpes, const double *tsince, uint64_t n_times,
                       double *out, int32_t *err);

/* Build, on the GPU, the polyjectory of the satellites of gpes (sorted by
 * NORAD ID, then epoch) over the UTC Julian dates [jd_begin, jd_end), and make
 * it the context's polyjectory, as mzh_upload_polyjectory would. Replaces
 * make_sgp4_polyjectory (mizuba/make_sgp4_polyjectory.cpp:1460-1605) with the
 * reference's checks, GPE selection and status codes. Each GPE's stretch of
 * a trajectory is cut into nominal steps of a twentieth of the orbital period
 * (shortened for eccentric orbits by the perigee's angular rate), each
 * interpolated at the Chebyshev nodes and halved until the positions are
 * within 1e-7 km of SGP4 between the nodes (pieces shorter than 5 s are kept
 * as they are; more than 20 pieces is MZH_SGP4_INTERPOLATION). A satellite's
 * trajectory ends at its first failing step, whose code becomes its status.
 * Deep-space element sets use SDP4's lunar-solar and resonance terms. Time in days
 * from jd_begin in TAI; the interpolated model solves Kepler's equation to
 * convergence. order: the polynomial order (8 in this build). */
int mzh_make_sgp4_polyjectory(mzh_ctx *ctx, const mzh_gpe *gpes, uint64_t n_gpes, double jd_begin, double jd_end,
                              double reentry_radius, double exit_radius, uint32_t order);

/* Device times (seconds, HIP events on the context's stream) of the kernels
 * of the last mzh_make_sgp4_polyjectory: segment set-up, count, scan, write
 * and the transposed copy for screening. No reference counterpart. */
int mzh_get_sgp4_times(const mzh_ctx *ctx, double seconds[5]);

/* The context's polyjectory: sizes, then the arrays (host pointers) in the
 * layout of mzh_upload_polyjectory, the per-object status (0 for uploaded
 * polyjectories) and the epoch of a generated one (TAI Julian date, two parts).
 * A NULL destination skips that array. */
int mzh_get_polyjectory_info(const mzh_ctx *ctx, uint64_t *n_objs, uint32_t *order, uint64_t *traj_len,
                             uint64_t *time_len, double *maxT, double epoch[2]);
int mzh_copy_polyjectory(mzh_ctx *ctx, double *traj, double *time, uint64_t *traj_off, uint64_t *nsteps,
                         uint64_t *time_off, int32_t *status);

/* ---- evaluation of the uploaded polyjectory ----------------------------- */

/* Where the evaluation time of row r (object sel[r], or r) and column k comes
 * from in mzh_state_eval: */
#define MZH_EVAL_TIME_SCALAR 0     /* tm[0] for every evaluation: state_eval(t) */
#define MZH_EVAL_TIME_PER_OBJECT 1 /* tm[r]: state_eval(t[n_rows]) */
#define MZH_EVAL_TIME_SHARED 2     /* tm[k], the same times for every row: state_meval(t[n_times]) */
#define MZH_EVAL_TIME_PER_EVAL 3   /* tm[r * n_times + k]: state_meval(t[n_rows][n_times]) */

/* Evaluate the uploaded polyjectory: out[r][k][0..6] = (x, y, z, vx, vy, vz, r)
 * of object sel[r] (object r when sel is NULL, and then n_rows must be n_objs)
 * at the time given by time_mode. Replaces polyjectory::state_eval and
 * polyjectory::state_meval (mizuba/polyjectory.hpp:184-204,
 * mizuba/polyjectory.cpp:1094-1366): Horner over the trajectory step whose end
 * time is the first one after t; NaN where the object has no data or t lies
 * outside [t_begin, t_end). n_times must be 1 for the SCALAR and PER_OBJECT
 * modes. Errors (the reference's messages, reported for the first failing
 * evaluation in (r, k) order): a non-finite time -> MZH_EINVAL, an object
 * index >= n_objs -> MZH_ERANGE.
 * mzh_state_eval takes host pointers (copied through HBM);
 * mzh_state_eval_device takes device pointers and leaves out in HBM. Both
 * return once the results are in out. */
int mzh_state_eval(mzh_ctx *ctx, const double *tm, int time_mode, const uint64_t *sel, uint64_t n_rows,
                   uint64_t n_times, double *out);
int mzh_state_eval_device(mzh_ctx *ctx, const double *tm, int time_mode, const uint64_t *sel, uint64_t n_rows,
                          uint64_t n_times, double *out);
/* Device memory on the context's device, for callers of the device-pointer
 * entry points that have no HIP allocator of their own. No reference
 * counterpart. mzh_copy_to_device / mzh_copy_from_device copy between host
 * and such memory and return once the copy is complete. */
int mzh_device_alloc(mzh_ctx *ctx, uint64_t bytes, void **ptr);
int mzh_device_free(mzh_ctx *ctx, void *ptr);
int mzh_copy_to_device(mzh_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int mzh_copy_from_device(mzh_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* Device time (seconds, HIP events on the context's stream) of the evaluation
 * kernel of the last mzh_state_eval / mzh_state_eval_device call. No
 * reference counterpart: measurement only. */
int mzh_get_eval_time(const mzh_ctx *ctx, double *seconds);

/* ---- the hot path ------------------------------------------------------ */

/* Run conjunction screening over all conjunction steps. Replaces
 * conjunctions::conjunctions (mizuba/conjunctions.cpp:263-360: validation,
 * n_cd_steps = ceil(maxT / conj_det_interval)) followed by
 * conjunctions::detect_conjunctions (mizuba/detect_conjunctions.cpp:55-537)
 * and its five stages:
 *   detect_conjunctions_aabbs        (mizuba/detect_conjunctions_aabbs.cpp:300-398)
 *   detect_conjunctions_morton       (mizuba/detect_conjunctions_morton.cpp:113-224)
 *   detect_conjunctions_bvh          (mizuba/detect_conjunctions_bvh.cpp:267-699)
 *   detect_conjunctions_broad_phase  (mizuba/detect_conjunctions_broad_phase.cpp:153-339)
 *   detect_conjunctions_narrow_phase (mizuba/detect_conjunctions_narrow_phase.cpp:167-769)
 * otypes may be NULL (all primaries); otherwise it holds n_objs values in
 * {1, 2, 4} (otype PRIMARY / SECONDARY / MASKED). Results stay in HBM and are
 * read with the accessors below. Error messages are the reference's. */
int mzh_screen(mzh_ctx *ctx, double conj_thresh, double conj_det_interval, const int32_t *otypes, int mode);

/* The validation of conjunctions::conjunctions (mizuba/conjunctions.cpp:263-360)
 * without a context or a device: the threshold, the interval, the object
 * types (otypes may be NULL) and the number of conjunction steps
 * ceil(maxT / conj_det_interval), returned in *n_cd_steps. Also checks this
 * build's one-batch size limit: n_objs * n_cd_steps must not exceed
 * 2^32 - 1 (MZH_ESIZE, with *n_cd_steps set). Same status and message
 * (mzh_global_error) as mzh_screen for the same arguments. */
int mzh_check_screen(uint64_t n_objs, double maxT, double conj_thresh, double conj_det_interval,
                     const int32_t *otypes, int mode, uint64_t *n_cd_steps);

/* mzh_screen restricted to the conjunction steps [step_begin, step_end) of
 * the screening (0 <= step_begin < step_end <= n_cd_steps): the same step
 * times (get_cd_begin_end with the whole screening's step count), and the
 * results of those steps only, as if their rows had been cut out of a whole
 * screening: the per-step arrays hold step_end - step_begin steps, the
 * conjunctions are those found in these steps. Screening a problem range by
 * range and concatenating gives the whole screening's results bit for bit;
 * the one-batch size limit applies to the range. Replaces the chunked step
 * loop of detect_conjunctions (mizuba/detect_conjunctions.cpp:61-79, the
 * chunks of cd_chunk_size steps; :184-242, :313-424, each chunk computed and
 * then written out before the next), so that a screening needs device memory
 * for one range of steps only. A range outside the screening is MZH_ERANGE. */
int mzh_screen_steps(mzh_ctx *ctx, double conj_thresh, double conj_det_interval, const int32_t *otypes, int mode,
                     uint64_t step_begin, uint64_t step_end);

/* Free and total memory of the context's device, in bytes. */
int mzh_device_memory(mzh_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes);

/* Conjunction steps per range for a screening of n_objs objects over
 * n_cd_steps steps whose device buffers (the per-step arrays, trees, pair and
 * sort scratch: about MZH_SCREEN_BYTES_PER_OBJECT_STEP bytes per object and
 * step) must fit in device_bytes: at least 1, at most n_cd_steps, and small
 * enough for the one-batch size limit. n_cd_steps itself means one batch. The
 * reference's counterpart is its fixed chunk of 128 steps
 * (mizuba/detect_conjunctions.cpp:79). */
#define MZH_SCREEN_BYTES_PER_OBJECT_STEP 400
int mzh_chunk_steps(uint64_t n_objs, uint64_t n_cd_steps, uint64_t device_bytes, uint64_t *chunk_steps);

/* Screen n independent problems together: ctxs[k] (each with its own uploaded
 * polyjectory, all on one device) is screened with conj_thresh[k],
 * conj_det_interval[k] and otypes[k] (otypes may be NULL, as may any
 * otypes[k]), exactly as mzh_screen would. Every stage runs as one launch for
 * all problems, on the first context's stream, so that many small problems
 * fill the GPU together. status[k] receives each problem's return code (its
 * message through mzh_last_error on ctxs[k]); the function returns the first
 * non-zero code, MZH_OK if none, or MZH_EINVAL if the arguments are malformed
 * (a null or repeated context, contexts on different devices).
 * The reference has no batched entry: it is conjunctions::conjunctions
 * (mizuba/conjunctions.cpp:263-360) called once per problem. */
int mzh_screen_batch(mzh_ctx *const *ctxs, uint64_t n, const double *conj_thresh, const double *conj_det_interval,
                     const int32_t *const *otypes, int mode, int *status);

/* The narrow phase's root finder applied to n free-standing polynomials:
 * run_poly_root_finding with direction +1 (mizuba/detail/poly_utils.cpp:272-460),
 * the minima of the narrow phase. polys[k][0..order] (ascending powers) on
 * [0, rf_int[k]); roots[k][0 .. 2*order+1] receives the n_roots[k] roots
 * found, flags[k] (may be NULL) bit 0 = the fast exclusion check fired, bit 1
 * = the isolation outgrew the private work list and ran with the 250-entry
 * list. Host pointers. An isolation failure is MZH_EINVAL with the
 * reference's message. */
int mzh_root_finding(mzh_ctx *ctx, uint32_t order, uint64_t n, const double *polys, const double *rf_int,
                     double *roots, uint32_t *n_roots, uint32_t *flags);

/* mzh_root_finding with device pointers (polys, rf_int, roots, n_roots and
 * flags in device memory on the context's device; flags may be NULL) that
 * allocates no memory: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No
 * reference counterpart. Its scratch (the slow path's counter, queue and
 * 251-entry work lists, and the error word) is a fixed 16 MiB array of the
 * library on each device, held until the process ends; n may not exceed what
 * it holds ((16 MiB - 256 B - the work lists) / 8 polynomials: about 1.39
 * million at order 8, 0.62 million at order 20), else MZH_EINVAL with a
 * message saying so. The call clears the scratch it uses, waits for its
 * kernels and returns MZH_EINVAL with the reference's message on an
 * isolation failure. Calls on one device share that scratch: calls that
 * overlap on one device, from any contexts, race. Outputs as in
 * mzh_root_finding: n_roots[k] = 0xfffffffe for a polynomial whose root
 * finding failed. */
int mzh_root_finding_noalloc(mzh_ctx *ctx, uint32_t order, uint64_t n, const double *polys, const double *rf_int,
                             double *roots, uint32_t *n_roots, uint32_t *flags);

/* ---- a batch handle: the same problems screened again and again ---------- */

/* A batch of n problems (as mzh_screen_batch takes them), kept for repeated
 * screening: the serving loop of a catalogue whose polyjectories are
 * refreshed in place, or the benchmark's repeated steps. The reference has no
 * counterpart: conjunctions::conjunctions (mizuba/conjunctions.cpp:263-360)
 * is called once per problem. Every stage runs as one launch for all
 * problems; the sort and the BVH build are persistent kernels over
 * (problem, conjunction step) work items. flags:
 *   MZH_BATCH_GRAPH  after the first (eager) run, the whole launch sequence
 *                    is captured into a HIP graph and later runs replay it,
 *                    with the buffer sizes of the previous run plus a quarter;
 *                    a replay whose sizes are too small is redone eagerly
 *                    and captured again;
 *   MZH_BATCH_TUNE   the first run first times the launch shapes (bucket-sort
 *                    width, persistent grid sizes, broad-phase workgroup,
 *                    root-finding grid) on this batch and keeps the fastest,
 *                    also for later batches of the same size class in the
 *                    process; the results do not depend on them.
 * The contexts (on one device, each with a polyjectory) must outlive the
 * batch; the arguments are validated as mzh_screen would (MZH_EINVAL,
 * message in mzh_global_error). A context changed between runs (a new
 * upload, a release, a screening by another call) is described again on the
 * next run. Results are read through the contexts' accessors, as after
 * mzh_screen. mzh_batch_run returns the first non-zero status; status[k]
 * receives each problem's. */
#define MZH_BATCH_GRAPH 1
#define MZH_BATCH_TUNE 2
typedef struct mzh_batch mzh_batch;
int mzh_batch_create(mzh_ctx *const *ctxs, uint64_t n, const double *conj_thresh, const double *conj_det_interval,
                     const int32_t *const *otypes, int mode, int flags, mzh_batch **out);
int mzh_batch_run(mzh_batch *batch, int *status);
/* Graph replays so far, device seconds of the last run, and the launch shape
 * in use: bucket-sort bits, sort and BVH workgroups, broad-phase workgroup
 * size, root-finding workgroups (any pointer may be NULL). */
int mzh_batch_info(const mzh_batch *batch, uint64_t *replays, double *seconds, uint32_t shape[5]);
int mzh_batch_destroy(mzh_batch *batch);

/* ---- accessors (mizuba/conjunctions.cpp:394-509) ------------------------ */

/* conjunctions::get_n_cd_steps (conjunctions.cpp:457-460). */
int mzh_get_n_cd_steps(const mzh_ctx *ctx, uint64_t *out);
/* Number of objects of the uploaded polyjectory. */
int mzh_get_n_objs(const mzh_ctx *ctx, uint64_t *out);
/* conjunctions::get_cd_end_times (conjunctions.cpp:399-402): double[C]. */
int mzh_copy_cd_end_times(mzh_ctx *ctx, double *dst);
/* conjunctions::get_aabbs / get_srt_aabbs (conjunctions.cpp:394-397, 404-407):
 * float[C][N+1][2][4] into dst. */
int mzh_copy_aabbs(mzh_ctx *ctx, float *dst);
int mzh_copy_srt_aabbs(mzh_ctx *ctx, float *dst);
/* conjunctions::get_mcodes / get_srt_mcodes / get_srt_idx (conjunctions.cpp:409-422). */
int mzh_copy_mcodes(mzh_ctx *ctx, uint64_t *dst);
int mzh_copy_srt_mcodes(mzh_ctx *ctx, uint64_t *dst);
int mzh_copy_srt_idx(mzh_ctx *ctx, uint32_t *dst);
/* conjunctions::get_bvh_tree (conjunctions.cpp:424-439): size, then the nodes
 * of step `step`. MZH_ERANGE with the reference's message if step >= C. */
int mzh_get_bvh_tree_size(mzh_ctx *ctx, uint64_t step, uint64_t *size);
int mzh_copy_bvh_tree(mzh_ctx *ctx, uint64_t step, mzh_bvh_node *dst);
/* conjunctions::get_aabb_collisions (conjunctions.cpp:441-455): per-step
 * broad-phase pairs, sorted by (i, j). */
int mzh_get_aabb_collisions_size(mzh_ctx *ctx, uint64_t step, uint64_t *size);
int mzh_copy_aabb_collisions(mzh_ctx *ctx, uint64_t step, mzh_aabb_collision *dst);
/* (offset, size) of every step's pairs in the concatenated pair array, as
 * the reference's bp_offsets (detect_conjunctions.cpp:141-152): uint64[C][2]. */
int mzh_copy_bp_offsets(mzh_ctx *ctx, uint64_t *dst);
/* conjunctions::get_conjunctions (conjunctions.cpp:478-491): all conjunctions,
 * step by step, each step sorted by (tca, then lexicographically). */
int mzh_get_n_conjunctions(mzh_ctx *ctx, uint64_t *out);
int mzh_copy_conjunctions(mzh_ctx *ctx, mzh_conj *dst);
/* np_report counters of the last screen (detect_conjunctions.cpp:490-514). */
int mzh_get_np_report(mzh_ctx *ctx, uint64_t out[MZH_NP_COUNT]);
/* Per-stage device times of the last screen (the trace breakdown of
 * detect_conjunctions.cpp:444-454), seconds: see MZH_STAGE_COUNT. */
int mzh_get_stage_times(mzh_ctx *ctx, double out[MZH_STAGE_COUNT]);

/* Device-resident views (no reference counterpart; the reference's arrays are
 * memory-mapped files): device pointer and byte size of an output array.
 * which: 0 aabbs, 1 srt_aabbs, 2 mcodes, 3 srt_mcodes, 4 srt_idx,
 *        5 bvh nodes (fixed capacity per step), 6 pairs, 7 conjunctions. */
int mzh_device_view(mzh_ctx *ctx, int which, void **ptr, uint64_t *bytes);

/* Make dst use the polyjectory of src (same device), without a copy: the
 * device arrays are shared, read-only, and stay allocated while any context
 * uses them (a later upload into either context gets arrays of its own). Lets
 * several screenings of one polyjectory (conjunctions() called again with
 * other thresholds, intervals or object types) skip the upload. No reference
 * counterpart: the reference's polyjectory is shared memory-mapped files
 * (mizuba/polyjectory.cpp:91-236). */
int mzh_share_polyjectory(mzh_ctx *dst, const mzh_ctx *src);

/* Free device memory of a context (the reference's hint_release, which
 * madvises its memory-mapped results away, mizuba/conjunctions.cpp:466-477):
 *   MZH_RELEASE_SCRATCH      buffers that no accessor reads (queues, counts,
 *                            sort and build scratch); a later mzh_screen
 *                            allocates them again;
 *   MZH_RELEASE_RESULTS      the device copies of the results: the copy
 *                            accessors and mzh_device_view then fail, while
 *                            the sizes, counters, step times and n_cd_steps
 *                            stay available;
 *   MZH_RELEASE_POLYJECTORY  the context's reference to its polyjectory.
 * Waits for the context's stream first. */
#define MZH_RELEASE_SCRATCH 1
#define MZH_RELEASE_RESULTS 2
#define MZH_RELEASE_POLYJECTORY 4
int mzh_release(mzh_ctx *ctx, int what);

/* Wait for all work queued on the context's stream. */
int mzh_synchronize(mzh_ctx *ctx);

/* Measurement helper, no counterpart in the reference: `launches` launches of
 * a plain streaming-copy kernel that reads `bytes` (a multiple of 16) and
 * writes `bytes` on the context's device, into two scratch buffers allocated
 * and freed here; *seconds receives the average launch duration. Its hardware
 * counters (FETCH_SIZE, WRITE_SIZE) calibrate the HBM traffic figures. */
int mzh_stream_copy(mzh_ctx *ctx, uint64_t bytes, int launches, double *seconds);

#ifdef __cplusplus
}
#endif

#endif
