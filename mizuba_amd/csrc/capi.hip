// C ABI of mizuba_amd and the host-side orchestrator of one screening problem.
//
// The orchestrator replaces conjunctions::conjunctions + detect_conjunctions
// (mizuba/conjunctions.cpp:263-360, mizuba/detect_conjunctions.cpp:55-537):
// instead of a TBB loop over conjunction steps in chunks of 128 with a writer
// thread, every stage runs once over all steps as a batched gfx950 launch,
// and the outputs stay resident in HBM until an accessor copies them out.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <limits>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/mizuba_amd.h"
#include "host_ctx.hpp"
#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

using namespace mzh;
using namespace mzh::host;

namespace
{

thread_local std::string g_global_error;

// ---- scratch of the allocation-free root finder ---------------------------
// mzh_root_finding_noalloc keeps its scratch (slow-path counter, error word,
// the slow pass's work lists and its queue) in this array; every device the
// library runs on holds its own copy until the process ends. Host code must
// not pass `scratch` itself to a kernel: it stands for a device address only
// on the device. Its per-device address comes from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = 16u << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_m;
void *g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on `device`, which must be current;
// resolved once per device.
hipError_t scratch_address(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_m);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[device] = p;
    }
    *ptr = g_scratch_dev[device];
    return hipSuccess;
}

// Layout of the scratch for a root-finding call at `order`: slow-path
// counter at 0, error word at 64, the slow pass's work lists from 256, then
// the slow queue (8 bytes per polynomial).
constexpr size_t kScrCount = 0, kScrErr = 64, kScrLists = 256;
size_t scr_queue_offset(uint32_t order) { return kScrLists + (narrow_wl_slow_bytes(order) + 255u) / 256u * 256u; }
uint64_t noalloc_max_polys(uint32_t order) { return (kNoAllocScratchBytes - scr_queue_offset(order)) / 8u; }

int fail(mzh_ctx *ctx, int code, const std::string &msg)
{
    if (ctx) ctx->last_error = msg;
    return code;
}

template <typename F>
int guarded(mzh_ctx *ctx, F &&f)
{
    if (!ctx) return MZH_EINVAL;
    try {
        MZH_HIP(hipSetDevice(ctx->device));
        f();
        return MZH_OK;
    } catch (const StatusError &e) {
        return fail(ctx, e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(ctx, MZH_ENOMEM, "host memory allocation failed");
    } catch (...) {
        return fail(ctx, MZH_ERUNTIME, "unknown error");
    }
}

} // namespace

namespace mzh::host
{

std::string dev_error_message(const DevError &e)
{
    const std::string i = std::to_string(e.i), j = std::to_string(e.j), v = fmt_double(e.value);
    switch (e.code) {
        case kErrAabbLb:
            return "The computation of the bounding box for the object at index " + i
                   + " produced the non-finite lower bound " + v;
        case kErrAabbUb:
            return "The computation of the bounding box for the object at index " + i
                   + " produced the non-finite upper bound " + v;
        case kErrGlobalBox:
            return "A global bounding box with non-finite size was generated";
        case kErrNpInterval:
            return "During the narrow phase collision detection of objects " + i + " and " + j
                   + ", an invalid time interval for polynomial root finding was generated";
        case kErrNpDist2:
            return "Non-finite value(s) detected during conjunction tracking for objects " + i + " and " + j;
        case kErrNpRoot:
            return "Polynomial root finding produced a non-finite root of " + v + " involving objects " + i + " and "
                   + j;
        case kErrNpRootDer:
            return "Polynomial root finding produced the root " + v + " with nonfinite derivative involving objects "
                   + i + " and " + j;
        case kErrNpIsolation:
            return "The polynomial root isolation algorithm failed: the working list size is " + i
                   + " and the number of isolating intervals is " + j;
        case kErrRfInterval:
            return "Invalid root-finding interval " + v + " for the polynomial at index " + i
                   + ": it must be finite and positive";
        case kErrNpIters:
            return "Polynomial root finding failed due to too many iterations";
        case kErrNpErrno:
            return "Polynomial root finding returned a nonzero errno with code '" + std::to_string((int)e.value) + "'";
        case kErrNpConjDist2:
            return "A non-finite conjunction distance square of " + v + " was computed for the objects at indices " + i
                   + " and " + j;
        case kErrNpBndDerUb:
            return "An invalid value of " + v
                   + " was computed for the derivative of the distance square function at the upper boundary of a "
                     "trajectory";
        case kErrNpBndDist2Ub:
            return "An invalid value of " + v
                   + " was computed for a candidate minimum of the distance square function at the upper boundary "
                     "of a trajectory";
        case kErrNpBndDerLb:
            return "An invalid value of " + v
                   + " was computed for the derivative of the distance square function at the lower boundary of a "
                     "trajectory";
        case kErrNpBndDist2Lb:
            return "An invalid value of " + v
                   + " was computed for a candidate minimum of the distance square function at the lower boundary "
                     "of a trajectory";
        case kErrBvhWalk:
            return "Internal error: the broad-phase walk of the object at sorted position " + j
                   + " in the conjunction step " + i + " did not end within its tree";
        case kErrBvhBuild:
            return "Internal error: the broad-phase tree of the conjunction step " + i
                   + " links its node " + j + " outside the tree";
        case kErrSortWindow:
            return "Internal error: the Morton sort of the conjunction step " + i + " found the buckets at sorted position "
                   + j + " outside its window";
        default:
            return "device error code " + std::to_string(e.code);
    }
}

} // namespace mzh::host

namespace
{

// Device arrays for a new polyjectory: the context's own, or fresh ones when
// other contexts share the current arrays (they keep those).
void own_pj_bufs(mzh_ctx *ctx)
{
    if (!ctx->pjb || ctx->pjb.use_count() > 1) ctx->pjb = std::make_shared<PjBufs>();
}

} // namespace

namespace mzh::host
{

PjView pj_view(const mzh_ctx *ctx)
{
    PjView v;
    v.traj = gl(ctx->pjb->traj.as<const double>());
    v.time = gl(ctx->pjb->time.as<const double>());
    v.cf4 = gl(ctx->pjb->cf4.as<const double>());
    v.cf4f = gl(ctx->pjb->cf4f.as<const float>());
    v.sbnd = gl(ctx->pjb->sbnd.as<const float4>());
    v.scal = gl(ctx->pjb->scal.as<const uint32_t>());
    v.step0 = gl(ctx->pjb->step0.as<const uint64_t>());
    v.nsteps = gl(ctx->pjb->nsteps.as<const uint64_t>());
    v.time_off = gl(ctx->pjb->time_off.as<const uint64_t>());
    v.n_objs = ctx->n_objs;
    v.order = ctx->order;
    return v;
}


} // namespace mzh::host

namespace
{

} // namespace

namespace mzh::host
{

// This build's size limit of one batch: the per-step stages index (step,
// object) with 32-bit counts, so N * C must stay below 2^32 (MZH_ESIZE).
void check_batch_size(uint64_t N, uint64_t C)
{
    if (N * C > 0xffffffffull)
        throw StatusError{MZH_ESIZE, "A screening of " + std::to_string(N) + " objects over " + std::to_string(C)
                                         + " conjunction steps is too large: the number of objects times the number "
                                           "of conjunction steps must not exceed 4294967295"};
}

// conjunctions::conjunctions validation (conjunctions.cpp:279-335) with the
// reference's messages, the step ranges of get_cd_begin_end
// (conjunctions.cpp:373-392), and (one_batch) the size limit of one batch.
// Everything is checked before any device memory is touched.
ScreenPlan plan_screen(uint32_t N, double maxT, double conj_thresh, double conj_det_interval,
                       const int32_t *otypes_in, int mode, bool one_batch, uint64_t k0, uint64_t k1)
{
    if (mode != MZH_MODE_FP64 && mode != MZH_MODE_MIXED)
        throw StatusError{MZH_EINVAL, "Invalid screening mode " + std::to_string(mode)};
    // conjunctions.cpp:279-296
    if (!std::isfinite(conj_thresh) || conj_thresh <= 0)
        throw StatusError{MZH_EINVAL, "The conjunction threshold must be finite and positive, but instead a value of "
                                          + fmt_double(conj_thresh) + " was provided"};
    if (!std::isfinite(conj_thresh * conj_thresh))
        throw StatusError{MZH_EINVAL, "A conjunction threshold of " + fmt_double(conj_thresh)
                                          + " is too large and results in an overflow error"};
    if (!std::isfinite(conj_det_interval) || conj_det_interval <= 0)
        throw StatusError{MZH_EINVAL,
                          "The conjunction detection interval must be finite and positive, but instead a value of "
                              + fmt_double(conj_det_interval) + " was provided"};
    ScreenPlan pl;
    // otypes (conjunctions.cpp:305-332)
    pl.ot.assign(N, 1);
    if (otypes_in) {
        for (uint32_t k = 0; k < N; ++k) {
            const int32_t v = otypes_in[k];
            if (v != 1 && v != 2 && v != 4)
                throw StatusError{MZH_EINVAL, "The value of an object type must be one of [1, 2, 4], but a value of "
                                                  + std::to_string(v) + " was detected instead"};
            pl.skip_cd = pl.skip_cd && (v != 1);
            pl.ot[k] = v;
        }
    } else {
        pl.skip_cd = false;
    }
    // n_cd_steps (conjunctions.cpp:335)
    const double cs = std::ceil(maxT / conj_det_interval);
    if (!(cs >= 1) || cs > 4.0e9)
        throw StatusError{MZH_EINVAL, "The number of conjunction steps " + fmt_double(cs) + " is not representable"};
    pl.C = (uint64_t)cs;
    if (one_batch) check_batch_size(N, pl.C);
    // get_cd_begin_end (conjunctions.cpp:373-392) for the steps [k0, k1)
    k1 = std::min(k1, pl.C);
    k0 = std::min(k0, k1);
    pl.cd_begin.resize(k1 - k0);
    pl.cd_end.resize(k1 - k0);
    for (uint64_t k = k0; k < k1; ++k) {
        const double b = conj_det_interval * static_cast<double>(k);
        const double e = (k == pl.C - 1) ? maxT : conj_det_interval * static_cast<double>(k + 1);
        if (!std::isfinite(b) || !std::isfinite(e) || !(e > b) || b < 0 || e > maxT)
            throw StatusError{MZH_EINVAL, "Invalid conjunction step time range: [" + fmt_double(b) + ", "
                                              + fmt_double(e) + ")"};
        pl.cd_begin[k - k0] = b;
        pl.cd_end[k - k0] = e;
    }
    return pl;
}

} // namespace mzh::host

namespace
{

// conjunctions::conjunctions validation + detect_conjunctions: a batch of one.
void run_screen(mzh_ctx *ctx, double conj_thresh, double conj_det_interval, const int32_t *otypes_in, int mode)
{
    if (!ctx->has_pj) throw StatusError{MZH_EINVAL, "No polyjectory has been uploaded to this context"};
    set_plan(ctx, plan_screen(ctx->n_objs, ctx->maxT, conj_thresh, conj_det_interval, otypes_in, mode), conj_thresh,
             conj_det_interval, mode);
    if (!ctx->solo) ctx->solo.reset(batch_new(ctx->stream));
    int status = MZH_OK;
    batch_run(ctx->solo.get(), {ctx}, false, false, &status);
    if (status != MZH_OK) throw StatusError{status, ctx->last_error};
}

// The conjunction steps [k0, k1) of a screening: the whole screening's plan
// cut to the range (its step times and object types), a batch of one.
void run_screen_steps(mzh_ctx *ctx, double conj_thresh, double conj_det_interval, const int32_t *otypes_in, int mode,
                      uint64_t k0, uint64_t k1)
{
    if (!ctx->has_pj) throw StatusError{MZH_EINVAL, "No polyjectory has been uploaded to this context"};
    // the range's step times only (a screening beyond one batch may have more
    // steps than host memory holds step times for)
    ScreenPlan pl =
        plan_screen(ctx->n_objs, ctx->maxT, conj_thresh, conj_det_interval, otypes_in, mode, false, k0, k1);
    if (!(k0 < k1 && k1 <= pl.C))
        throw StatusError{MZH_ERANGE, "Invalid range of conjunction steps [" + std::to_string(k0) + ", "
                                          + std::to_string(k1) + ") for a screening of " + std::to_string(pl.C)
                                          + " conjunction steps"};
    check_batch_size(ctx->n_objs, k1 - k0);
    pl.C = k1 - k0;
    set_plan(ctx, std::move(pl), conj_thresh, conj_det_interval, mode);
    if (!ctx->solo) ctx->solo.reset(batch_new(ctx->stream));
    int status = MZH_OK;
    batch_run(ctx->solo.get(), {ctx}, false, false, &status);
    if (status != MZH_OK) throw StatusError{status, ctx->last_error};
}

void require_screened(const mzh_ctx *ctx)
{
    if (!ctx->screened) throw StatusError{MZH_EINVAL, "No screening results are available in this context"};
}

// For the accessors that read the device copies of the results.
void require_results(const mzh_ctx *ctx)
{
    require_screened(ctx);
    if (ctx->results_released)
        throw StatusError{MZH_EINVAL, "The device copies of the results of this screening have been released"};
}

// The screening walks a tree of its own numbering; the reference's
// breadth-first trees and their 48-byte bvh_node records are made on the
// first request after a screen.
void export_nodes(mzh_ctx *ctx)
{
    if (ctx->nodes_ready) return;
    build_bfs_trees(ctx);
    MZH_HIP(ctx->nodes.ensure(sizeof(mzh_bvh_node) * ctx->node_cap * ctx->C));
    MZH_HIP(launch_bvh_export(ctx->trav.as<TravNode>(), ctx->tree_size.as<uint32_t>(), ctx->nodes.as<mzh_bvh_node>(),
                              ctx->node_cap, (uint32_t)ctx->C, ctx->stream));
    MZH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nodes_ready = true;
}

// Streaming copy of n 16-byte words (the traffic calibration of
// mzh_stream_copy): every word read once and written once.
__global__ __launch_bounds__(256) void k_stream_copy(const uint4 *__restrict__ src, uint4 *__restrict__ dst, uint64_t n)
{
    const uint64_t stride = (uint64_t)gridDim.x * 256u;
    for (uint64_t k = (uint64_t)blockIdx.x * 256u + threadIdx.x; k < n; k += stride) dst[k] = src[k];
}

} // namespace

extern "C" {

int mzh_device_count(int *count)
{
    if (!count) return MZH_EINVAL;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return MZH_OK;
}

int mzh_global_error(char *buf, size_t len)
{
    if (!buf || len == 0) return MZH_EINVAL;
    std::snprintf(buf, len, "%s", g_global_error.c_str());
    return MZH_OK;
}

int mzh_create(int device, mzh_ctx **out)
{
    if (!out) return MZH_EINVAL;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        g_global_error = "No HIP device is available";
        return MZH_ERUNTIME;
    }
    if (device < 0 || device >= n) {
        g_global_error = "Invalid HIP device index " + std::to_string(device);
        return MZH_EINVAL;
    }
    mzh_ctx *ctx = nullptr;
    try {
        ctx = new mzh_ctx();
    } catch (...) {
        g_global_error = "host memory allocation failed";
        return MZH_ENOMEM;
    }
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    // set-up of the device's scratch address (once per device); a failure is
    // reported by mzh_root_finding_noalloc, which tries again
    if (e == hipSuccess) {
        void *scr = nullptr;
        (void)scratch_address(device, &scr);
    }
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipEventCreate(&ctx->ev_eval[k]);
    for (int k = 0; k < 8 && e == hipSuccess; ++k) e = hipEventCreate(&ctx->ev_sgp4[k]);
    if (e != hipSuccess) {
        g_global_error = std::string("HIP error while creating a context: ") + hipGetErrorString(e);
        delete ctx;
        return MZH_ERUNTIME;
    }
    *out = ctx;
    return MZH_OK;
}

int mzh_destroy(mzh_ctx *ctx)
{
    if (!ctx) return MZH_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &e : ctx->ev_eval)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : ctx->ev_sgp4)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return MZH_OK;
}

int mzh_last_error(const mzh_ctx *ctx, char *buf, size_t len)
{
    if (!ctx || !buf || len == 0) return MZH_EINVAL;
    std::snprintf(buf, len, "%s", ctx->last_error.c_str());
    return MZH_OK;
}

int mzh_upload_polyjectory(mzh_ctx *ctx, uint32_t n_objs, uint32_t order, const double *traj, uint64_t traj_len,
                           const double *time, uint64_t time_len, const uint64_t *traj_off, const uint64_t *nsteps,
                           const uint64_t *time_off, double maxT)
{
    return guarded(ctx, [&] {
        if (n_objs == 0) throw StatusError{MZH_EINVAL, "Cannot upload a polyjectory with zero objects"};
        if (order < (uint32_t)kMinOrder || order > (uint32_t)kMaxOrder)
            throw StatusError{MZH_EINVAL, "Polynomial order " + std::to_string(order)
                                              + " is not supported (the supported range is [2, "
                                              + std::to_string(kMaxOrder) + "])"};
        if (!traj_off || !nsteps || !time_off || (!traj && traj_len) || (!time && time_len))
            throw StatusError{MZH_EINVAL, "Null pointer passed to mzh_upload_polyjectory"};
        if (!std::isfinite(maxT) || maxT <= 0) throw StatusError{MZH_EINVAL, "Invalid maxT"};
        const uint64_t step_sz = 7ull * (order + 1);
        std::vector<uint64_t> step0(n_objs);
        uint64_t steps_total = traj_len / step_sz;
        if (traj_len % step_sz) throw StatusError{MZH_EINVAL, "Trajectory data size is not a whole number of steps"};
        for (uint32_t o = 0; o < n_objs; ++o) {
            if (traj_off[o] % step_sz) throw StatusError{MZH_EINVAL, "Misaligned trajectory offset"};
            step0[o] = traj_off[o] / step_sz;
            if (step0[o] + nsteps[o] > steps_total) throw StatusError{MZH_EINVAL, "Trajectory offsets out of range"};
            if (nsteps[o] >= 0xffffffffull)
                throw StatusError{MZH_EINVAL, "Too many trajectory steps for one object"};
            if (nsteps[o] && time_off[o] + nsteps[o] + 1 > time_len)
                throw StatusError{MZH_EINVAL, "Time offsets out of range"};
        }
        hipStream_t st = ctx->stream;
        ctx->has_pj = false;
        ctx->screened = false;
        own_pj_bufs(ctx);
        ++ctx->generation;
        MZH_HIP(ctx->pjb->traj.ensure(sizeof(double) * (traj_len ? traj_len : 1)));
        MZH_HIP(ctx->pjb->time.ensure(sizeof(double) * (time_len ? time_len : 1)));
        MZH_HIP(ctx->pjb->cf4.ensure(sizeof(double) * 4u * (order + 1) * (steps_total ? steps_total : 1)));
        MZH_HIP(ctx->pjb->cf4f.ensure(sizeof(float) * 4u * (order + 1) * (steps_total ? steps_total : 1)));
        MZH_HIP(ctx->pjb->sbnd.ensure(sizeof(float4) * (steps_total ? steps_total : 1)));
        MZH_HIP(ctx->pjb->scal.ensure(2 * sizeof(uint32_t)));
        MZH_HIP(ctx->pjb->step0.ensure(sizeof(uint64_t) * n_objs));
        MZH_HIP(ctx->pjb->nsteps.ensure(sizeof(uint64_t) * n_objs));
        MZH_HIP(ctx->pjb->time_off.ensure(sizeof(uint64_t) * n_objs));
        if (traj_len) MZH_HIP(hipMemcpyAsync(ctx->pjb->traj.p, traj, sizeof(double) * traj_len, hipMemcpyHostToDevice, st));
        if (time_len) MZH_HIP(hipMemcpyAsync(ctx->pjb->time.p, time, sizeof(double) * time_len, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->pjb->step0.p, step0.data(), sizeof(uint64_t) * n_objs, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->pjb->nsteps.p, nsteps, sizeof(uint64_t) * n_objs, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->pjb->time_off.p, time_off, sizeof(uint64_t) * n_objs, hipMemcpyHostToDevice, st));
        MZH_HIP(launch_transpose_cf4(ctx->pjb->traj.as<double>(), ctx->pjb->cf4.as<double>(), ctx->pjb->cf4f.as<float>(),
                                 steps_total, order, st));
        MZH_HIP(launch_step_bounds(ctx->pjb->cf4.as<double>(), ctx->pjb->time.as<double>(),
                                   ctx->pjb->step0.as<uint64_t>(), ctx->pjb->nsteps.as<uint64_t>(),
                                   ctx->pjb->time_off.as<uint64_t>(), n_objs, steps_total, order,
                                   ctx->pjb->sbnd.as<float4>(), ctx->pjb->scal.as<uint32_t>(), st));
        MZH_HIP(hipMemcpyAsync(ctx->pjb->h_scal, ctx->pjb->scal.p, sizeof(ctx->pjb->h_scal), hipMemcpyDeviceToHost, st));
        MZH_HIP(hipStreamSynchronize(st));
        ctx->n_objs = n_objs;
        ctx->order = order;
        ctx->maxT = maxT;
        ctx->n_steps_total = steps_total;
        ctx->h_traj_off.assign(traj_off, traj_off + n_objs);
        ctx->h_nsteps.assign(nsteps, nsteps + n_objs);
        ctx->h_time_off.assign(time_off, time_off + n_objs);
        ctx->h_status.assign(n_objs, 0);
        ctx->traj_len = traj_len;
        ctx->time_len = time_len;
        ctx->pj_epoch[0] = ctx->pj_epoch[1] = 0;
        ctx->has_pj = true;
    });
}

namespace
{

// Double-length sum (Knuth's two-sum), normalised.
struct DL {
    double hi, lo;
};
DL dl_norm(double a, double b)
{
    const double s = a + b, bp = s - a, ap = s - bp;
    return DL{s, (a - ap) + (b - bp)};
}
bool dl_less(const DL &a, const DL &b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }

// TAI - UTC (leap seconds) from 1972 on, as published by the IERS: (UTC
// Julian date from which it applies, seconds).
constexpr double kLeap[][2] = {
    {2441317.5, 10}, {2441499.5, 11}, {2441683.5, 12}, {2442048.5, 13}, {2442413.5, 14}, {2442778.5, 15},
    {2443144.5, 16}, {2443509.5, 17}, {2443874.5, 18}, {2444239.5, 19}, {2444786.5, 20}, {2445151.5, 21},
    {2445516.5, 22}, {2446247.5, 23}, {2447161.5, 24}, {2447892.5, 25}, {2448257.5, 26}, {2448804.5, 27},
    {2449169.5, 28}, {2449534.5, 29}, {2450083.5, 30}, {2450630.5, 31}, {2451179.5, 32}, {2453736.5, 33},
    {2454832.5, 34}, {2456109.5, 35}, {2457204.5, 36}, {2457754.5, 37}};

// TAI - UTC in seconds at the UTC Julian date jd (>= 1972).
double tai_minus_utc(double jd)
{
    double dat = kLeap[0][1];
    for (const auto &e : kLeap)
        if (jd >= e[0]) dat = e[1];
    return dat;
}

// UTC Julian date (two parts) -> TAI Julian date, double-length, as the IAU
// SOFA routine utctai states it (make_sgp4_polyjectory.cpp:115-121 converts
// through heyoka's jd_utc_to_tai): the fraction of the day is scaled by
// (86400 + leap) / 86400 on a day that ends with a leap second, then TAI - UTC
// at the day's start is added.
DL utc_to_tai(double jd1, double jd2)
{
    const DL a = dl_norm(jd1, jd2);
    if (!(a.hi >= kLeap[0][0]))
        throw StatusError{MZH_EINVAL, "UTC dates before 1972 are not supported by make_sgp4_polyjectory()"};
    double day0 = std::floor(a.hi - 0.5) + 0.5;
    double fd = (a.hi - day0) + a.lo;
    if (fd < 0.0) {
        day0 -= 1.0;
        fd += 1.0;
    } else if (fd >= 1.0) {
        day0 += 1.0;
        fd -= 1.0;
    }
    const double dat0 = tai_minus_utc(day0), leap = tai_minus_utc(day0 + 1.0) - dat0;
    return dl_norm(day0, fd * ((86400.0 + leap) / 86400.0) + dat0 / 86400.0);
}

// b - a in days, for nearby double-length dates.
double dl_diff(const DL &b, const DL &a) { return (b.hi - a.hi) + (b.lo - a.lo); }

DL gpe_epoch(const mzh_gpe &g) { return dl_norm(g.epoch_jd1, g.epoch_jd2); }

std::string fmt_u64(uint64_t v) { return std::to_string(v); }

// make_sgp4_polyjectory (make_sgp4_polyjectory.cpp:1460-1605).
void make_sgp4_impl(mzh_ctx *ctx, const mzh_gpe *gpes, uint64_t n_gpes, double jd_begin, double jd_end,
                    double reentry_radius, double exit_radius, uint32_t order)
{
    if (!std::isfinite(jd_begin) || !std::isfinite(jd_end) || !(jd_begin < jd_end))
        throw StatusError{MZH_EINVAL, "Invalid Julian date interval [" + fmt_double(jd_begin) + ", " + fmt_double(jd_end)
                                          + ") supplied to make_sgp4_polyjectory(): the begin/end dates must be "
                                            "finite and the end date must be strictly after the begin date"};
    if (std::isnan(reentry_radius) || std::isnan(exit_radius))
        throw StatusError{MZH_EINVAL, "The reentry/exit radiuses in make_sgp4_polyjectory() cannot be NaN"};
    if (!(reentry_radius < exit_radius))
        throw StatusError{MZH_EINVAL, "The reentry radius (" + fmt_double(reentry_radius)
                                          + ") must be less than the exit radius (" + fmt_double(exit_radius)
                                          + ") in make_sgp4_polyjectory()"};
    if (n_gpes == 0 || !gpes)
        throw StatusError{MZH_EINVAL, "make_sgp4_polyjectory() requires a non-empty array of GPEs in input"};
    if (order != 8)
        throw StatusError{MZH_EINVAL, "make_sgp4_polyjectory() builds polyjectories of order 8 in this build, but "
                                      "the order "
                                          + std::to_string(order) + " was requested"};
    // finiteness (make_sgp4_polyjectory.cpp:123-147), then the order (:150-186)
    for (uint64_t i = 0; i < n_gpes; ++i) {
        const mzh_gpe &g = gpes[i];
        const std::pair<const char *, double> f[] = {{"epoch_jd1", g.epoch_jd1}, {"epoch_jd2", g.epoch_jd2},
                                                     {"n0", g.n0},               {"e0", g.e0},
                                                     {"i0", g.i0},               {"node0", g.node0},
                                                     {"omega0", g.omega0},       {"m0", g.m0},
                                                     {"bstar", g.bstar}};
        for (const auto &x : f)
            if (!std::isfinite(x.second))
                throw StatusError{MZH_EINVAL, std::string("Invalid GPE detected by make_sgp4_polyjectory(): the '")
                                                  + x.first + "' of the GPE at index " + std::to_string(i)
                                                  + " is not finite"};
    }
    for (uint64_t i = 1; i < n_gpes; ++i) {
        const mzh_gpe &a = gpes[i - 1], &b = gpes[i];
        if (a.norad_id < b.norad_id) continue;
        if (a.norad_id == b.norad_id) {
            const DL ea = gpe_epoch(a), eb = gpe_epoch(b);
            if (dl_less(ea, eb)) continue;
            if (!dl_less(eb, ea))
                throw StatusError{MZH_EINVAL, "Two GPEs with identical NORAD ID " + fmt_u64(a.norad_id) + " and epoch ("
                                                  + fmt_double(a.epoch_jd1) + ", " + fmt_double(a.epoch_jd2)
                                                  + ") were identified by make_sgp4_polyjectory() - this is not "
                                                    "allowed"};
        }
        throw StatusError{MZH_EINVAL, "The set of GPEs passed to make_sgp4_polyjectory() is not sorted correctly: it "
                                      "must be ordered first by NORAD ID and then by epoch"};
    }
    // satellites, the GPE each starts with (the latest at or before jd_begin,
    // else the first: make_sgp4_polyjectory.cpp:1308-1335) and the segments:
    // each GPE from the later of jd_begin and its epoch to the next GPE's
    // epoch, or to jd_end when there is no next GPE before it (:1340-1375).
    // Times in days from the polyjectory's epoch (jd_begin in TAI).
    const DL pj_epoch = utc_to_tai(jd_begin, 0.0);
    const double t_end = dl_diff(utc_to_tai(jd_end, 0.0), pj_epoch);
    const DL begin_utc = dl_norm(jd_begin, 0.0), end_utc = dl_norm(jd_end, 0.0);
    if (n_gpes > 0xfffffff0ull) throw StatusError{MZH_EINVAL, "Too many GPEs for make_sgp4_polyjectory()"};
    std::vector<Sgp4Seg> segs;
    std::vector<uint32_t> sat_seg0;
    for (uint64_t i = 0; i < n_gpes;) {
        uint64_t j = i;
        while (j < n_gpes && gpes[j].norad_id == gpes[i].norad_id) ++j;
        uint64_t start = i;
        for (uint64_t k = i; k < j; ++k)
            if (!dl_less(begin_utc, gpe_epoch(gpes[k]))) start = k;
        const uint32_t sat = (uint32_t)sat_seg0.size();
        sat_seg0.push_back((uint32_t)segs.size());
        double t_b = 0.0;
        for (uint64_t g = start; g < j; ++g) {
            const bool last = g + 1u == j || !dl_less(gpe_epoch(gpes[g + 1u]), end_utc);
            Sgp4Seg sg{};
            sg.eoff = dl_diff(utc_to_tai(gpes[g].epoch_jd1, gpes[g].epoch_jd2), pj_epoch);
            sg.t_b = t_b;
            sg.t_e = last ? t_end : dl_diff(utc_to_tai(gpes[g + 1u].epoch_jd1, gpes[g + 1u].epoch_jd2), pj_epoch);
            sg.sat = sat;
            sg.gpe = (uint32_t)g;
            segs.push_back(sg);
            if (last) break;
            t_b = sg.t_e;
        }
        i = j;
    }
    const uint32_t n_sats = (uint32_t)sat_seg0.size();
    const uint32_t n_segs = (uint32_t)segs.size();
    sat_seg0.push_back(n_segs);
    const uint32_t op1 = order + 1u;
    hipStream_t st = ctx->stream;
    DevBuf d_gpes, d_segs, d_sats, d_satseg, d_segstep, d_res, d_out, d_np, d_status, d_tlast;
    MZH_HIP(d_gpes.ensure(sizeof(mzh_gpe) * n_gpes));
    MZH_HIP(d_segs.ensure(sizeof(Sgp4Seg) * n_segs));
    MZH_HIP(d_sats.ensure(sizeof(Sgp4Sat) * n_segs));
    MZH_HIP(d_satseg.ensure(sizeof(uint32_t) * (n_sats + 1u)));
    MZH_HIP(d_segstep.ensure(sizeof(uint64_t) * (n_segs + 1u)));
    MZH_HIP(d_np.ensure(sizeof(uint64_t) * n_sats));
    MZH_HIP(d_status.ensure(sizeof(int32_t) * n_sats));
    MZH_HIP(d_tlast.ensure(sizeof(double) * n_sats));
    MZH_HIP(hipMemcpyAsync(d_gpes.p, gpes, sizeof(mzh_gpe) * n_gpes, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemcpyAsync(d_segs.p, segs.data(), sizeof(Sgp4Seg) * n_segs, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemcpyAsync(d_satseg.p, sat_seg0.data(), sizeof(uint32_t) * (n_sats + 1u), hipMemcpyHostToDevice, st));
    Sgp4Gen g{};
    g.gpes = d_gpes.as<const mzh_gpe>();
    g.segs = d_segs.as<Sgp4Seg>();
    g.sats = d_sats.as<Sgp4Sat>();
    g.n_segs = n_segs;
    g.n_sats = n_sats;
    g.sat_seg0 = d_satseg.as<const uint32_t>();
    g.reentry_r2 = reentry_radius * reentry_radius;
    g.exit_r2 = std::isinf(exit_radius) ? std::numeric_limits<double>::infinity() : exit_radius * exit_radius;
    hipEvent_t *E = ctx->ev_sgp4;
    MZH_HIP(hipEventRecord(E[0], st));
    MZH_HIP(launch_sgp4_seg_init(g, st));
    MZH_HIP(hipEventRecord(E[1], st));
    MZH_HIP(hipMemcpyAsync(segs.data(), d_segs.p, sizeof(Sgp4Seg) * n_segs, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipStreamSynchronize(st));
    // nominal steps of each segment (none for an empty stretch of a valid GPE)
    std::vector<uint64_t> seg_step0(n_segs + 1u);
    uint64_t n_steps = 0;
    g.any_deep = false;
    for (uint32_t k = 0; k < n_segs; ++k) {
        if (segs[k].status == 0 && segs[k].deep) g.any_deep = true;
        if (segs[k].status == 0 && !(segs[k].t_e > segs[k].t_b)) segs[k].n_steps = 0;
        segs[k].step0 = n_steps;
        seg_step0[k] = n_steps;
        n_steps += segs[k].n_steps;
    }
    seg_step0[n_segs] = n_steps;
    MZH_HIP(hipMemcpyAsync(d_segs.p, segs.data(), sizeof(Sgp4Seg) * n_segs, hipMemcpyHostToDevice, st));
    MZH_HIP(d_segstep.ensure(sizeof(uint64_t) * (n_segs + 1u)));
    MZH_HIP(hipMemcpyAsync(d_segstep.p, seg_step0.data(), sizeof(uint64_t) * (n_segs + 1u), hipMemcpyHostToDevice,
                           st));
    // step indices are 32-bit in the generator's lists
    if (n_steps > 0xffffffffull) throw StatusError{MZH_EINVAL, "Too many trajectory steps for make_sgp4_polyjectory()"};
    MZH_HIP(d_res.ensure(sizeof(int32_t) * (n_steps ? n_steps : 1)));
    MZH_HIP(d_out.ensure(sizeof(uint32_t) * (n_steps ? n_steps : 1)));
    // stage single-piece steps when that takes at most a quarter of the free
    // device memory (else the write pass interpolates every step again)
    DevBuf d_stage, d_stage_t, d_redo, d_redo_n;
    g.stage = g.stage_t = nullptr;
    MZH_HIP(d_redo.ensure(sizeof(uint32_t) * (n_steps ? n_steps : 1)));
    MZH_HIP(d_redo_n.ensure(sizeof(unsigned long long)));
    MZH_HIP(hipMemsetAsync(d_redo_n.p, 0, sizeof(unsigned long long), st));
    g.redo = d_redo.as<uint32_t>();
    g.redo_count = d_redo_n.as<unsigned long long>();
    {
        size_t free_b = 0, total_b = 0;
        const uint64_t stage_b = n_steps * sizeof(double) * 7u * (order + 1u);
        if (n_steps && hipMemGetInfo(&free_b, &total_b) == hipSuccess && stage_b <= free_b / 4u) {
            MZH_HIP(d_stage.ensure(stage_b));
            MZH_HIP(d_stage_t.ensure(sizeof(double) * (n_steps + 1u)));
            g.stage = d_stage.as<double>();
            g.stage_t = d_stage_t.as<double>();
        }
    }
    g.seg_step0 = d_segstep.as<const uint64_t>();
    g.n_steps = n_steps;
    g.step_res = d_res.as<int32_t>();
    g.step_out = d_out.as<uint32_t>();
    g.sat_npolys = d_np.as<uint64_t>();
    g.sat_status = d_status.as<int32_t>();
    g.sat_tlast = d_tlast.as<double>();
    MZH_HIP(hipEventRecord(E[2], st));
    MZH_HIP(launch_sgp4_count(g, order, st));
    MZH_HIP(hipEventRecord(E[3], st));
    MZH_HIP(launch_sgp4_scan(g, st));
    MZH_HIP(hipEventRecord(E[4], st));
    std::vector<uint64_t> nsteps(n_sats);
    std::vector<int32_t> status(n_sats);
    std::vector<double> t_last(n_sats);
    unsigned long long n_redo = 0;
    MZH_HIP(hipMemcpyAsync(&n_redo, d_redo_n.p, sizeof(n_redo), hipMemcpyDeviceToHost, st));
    MZH_HIP(hipMemcpyAsync(nsteps.data(), d_np.p, sizeof(uint64_t) * n_sats, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipMemcpyAsync(status.data(), d_status.p, sizeof(int32_t) * n_sats, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipMemcpyAsync(t_last.data(), d_tlast.p, sizeof(double) * n_sats, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipStreamSynchronize(st));
    // layout of the polyjectory
    std::vector<uint64_t> step0(n_sats), traj_off(n_sats), time_off(n_sats);
    uint64_t steps_total = 0, time_len = 0;
    double maxT = 0;
    for (uint32_t k = 0; k < n_sats; ++k) {
        if (nsteps[k] >= 0xffffffffull) throw StatusError{MZH_EINVAL, "Too many trajectory steps for one object"};
        step0[k] = steps_total;
        traj_off[k] = steps_total * op1 * 7u;
        time_off[k] = time_len;
        steps_total += nsteps[k];
        time_len += nsteps[k] ? nsteps[k] + 1u : 0u;
        if (nsteps[k]) maxT = std::max(maxT, t_last[k]);
    }
    if (steps_total == 0)
        throw StatusError{MZH_EINVAL,
                          "All the trajectories in a polyjectory have a number of steps equal to zero: this is not "
                          "allowed"};
    ctx->has_pj = false;
    ctx->screened = false;
    own_pj_bufs(ctx);
    ++ctx->generation;
    MZH_HIP(ctx->pjb->traj.ensure(sizeof(double) * steps_total * op1 * 7u));
    MZH_HIP(ctx->pjb->time.ensure(sizeof(double) * time_len));
    MZH_HIP(ctx->pjb->cf4.ensure(sizeof(double) * 4u * op1 * steps_total));
    MZH_HIP(ctx->pjb->cf4f.ensure(sizeof(float) * 4u * op1 * steps_total));
    MZH_HIP(ctx->pjb->sbnd.ensure(sizeof(float4) * steps_total));
    MZH_HIP(ctx->pjb->scal.ensure(2 * sizeof(uint32_t)));
    MZH_HIP(ctx->pjb->step0.ensure(sizeof(uint64_t) * n_sats));
    MZH_HIP(ctx->pjb->nsteps.ensure(sizeof(uint64_t) * n_sats));
    MZH_HIP(ctx->pjb->time_off.ensure(sizeof(uint64_t) * n_sats));
    MZH_HIP(hipMemcpyAsync(ctx->pjb->step0.p, step0.data(), sizeof(uint64_t) * n_sats, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemcpyAsync(ctx->pjb->nsteps.p, nsteps.data(), sizeof(uint64_t) * n_sats, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemcpyAsync(ctx->pjb->time_off.p, time_off.data(), sizeof(uint64_t) * n_sats, hipMemcpyHostToDevice, st));
    g.traj = ctx->pjb->traj.as<double>();
    g.time = ctx->pjb->time.as<double>();
    g.sat_poly0 = ctx->pjb->step0.as<const uint64_t>();
    g.sat_time0 = ctx->pjb->time_off.as<const uint64_t>();
    MZH_HIP(hipEventRecord(E[5], st));
    MZH_HIP(launch_sgp4_write(g, order, n_redo, st));
    MZH_HIP(hipEventRecord(E[6], st));
    MZH_HIP(launch_transpose_cf4(ctx->pjb->traj.as<double>(), ctx->pjb->cf4.as<double>(), ctx->pjb->cf4f.as<float>(),
                                 steps_total, order, st));
    MZH_HIP(launch_step_bounds(ctx->pjb->cf4.as<double>(), ctx->pjb->time.as<double>(), ctx->pjb->step0.as<uint64_t>(),
                               ctx->pjb->nsteps.as<uint64_t>(), ctx->pjb->time_off.as<uint64_t>(), n_sats, steps_total,
                               order, ctx->pjb->sbnd.as<float4>(), ctx->pjb->scal.as<uint32_t>(), st));
    MZH_HIP(hipEventRecord(E[7], st));
    MZH_HIP(hipMemcpyAsync(ctx->pjb->h_scal, ctx->pjb->scal.p, sizeof(ctx->pjb->h_scal), hipMemcpyDeviceToHost, st));
    MZH_HIP(hipStreamSynchronize(st));
    {
        const int pairs[5][2] = {{0, 1}, {2, 3}, {3, 4}, {5, 6}, {6, 7}};
        for (int k = 0; k < 5; ++k) {
            float ms = 0;
            MZH_HIP(hipEventElapsedTime(&ms, E[pairs[k][0]], E[pairs[k][1]]));
            ctx->sgp4_s[k] = ms * 1e-3;
        }
    }
    ctx->n_objs = n_sats;
    ctx->order = order;
    ctx->maxT = maxT;
    ctx->n_steps_total = steps_total;
    ctx->h_traj_off = std::move(traj_off);
    ctx->h_nsteps = std::move(nsteps);
    ctx->h_time_off = std::move(time_off);
    ctx->h_status = std::move(status);
    ctx->traj_len = steps_total * op1 * 7u;
    ctx->time_len = time_len;
    ctx->pj_epoch[0] = pj_epoch.hi;
    ctx->pj_epoch[1] = pj_epoch.lo;
    ctx->has_pj = true;
}

} // namespace

int mzh_sgp4_propagate(mzh_ctx *ctx, const mzh_gpe *gpes, uint64_t n_gpes, const double *tsince, uint64_t n_times,
                       double *out, int32_t *err)
{
    return guarded(ctx, [&] {
        const uint64_t n = n_gpes * n_times;
        if (n == 0) return;
        if (!gpes || !tsince || !out || !err) throw StatusError{MZH_EINVAL, "Null pointer passed to mzh_sgp4_propagate"};
        if (n_times && n_gpes > (uint64_t)1 << 40) throw StatusError{MZH_EINVAL, "Too many propagations"};
        hipStream_t st = ctx->stream;
        DevBuf d_g, d_t, d_o, d_e;
        MZH_HIP(d_g.ensure(sizeof(mzh_gpe) * n_gpes));
        MZH_HIP(d_t.ensure(sizeof(double) * n));
        MZH_HIP(d_o.ensure(sizeof(double) * 6u * n));
        MZH_HIP(d_e.ensure(sizeof(int32_t) * n));
        MZH_HIP(hipMemcpyAsync(d_g.p, gpes, sizeof(mzh_gpe) * n_gpes, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(d_t.p, tsince, sizeof(double) * n, hipMemcpyHostToDevice, st));
        MZH_HIP(launch_sgp4_propagate(d_g.as<const mzh_gpe>(), n_gpes, d_t.as<const double>(), n_times,
                                      d_o.as<double>(), d_e.as<int32_t>(), st));
        MZH_HIP(hipMemcpyAsync(out, d_o.p, sizeof(double) * 6u * n, hipMemcpyDeviceToHost, st));
        MZH_HIP(hipMemcpyAsync(err, d_e.p, sizeof(int32_t) * n, hipMemcpyDeviceToHost, st));
        MZH_HIP(hipStreamSynchronize(st));
    });
}

int mzh_make_sgp4_polyjectory(mzh_ctx *ctx, const mzh_gpe *gpes, uint64_t n_gpes, double jd_begin, double jd_end,
                              double reentry_radius, double exit_radius, uint32_t order)
{
    return guarded(ctx,
                   [&] { make_sgp4_impl(ctx, gpes, n_gpes, jd_begin, jd_end, reentry_radius, exit_radius, order); });
}

int mzh_get_polyjectory_info(const mzh_ctx *ctx, uint64_t *n_objs, uint32_t *order, uint64_t *traj_len,
                             uint64_t *time_len, double *maxT, double epoch[2])
{
    if (!ctx || !ctx->has_pj) return MZH_EINVAL;
    if (n_objs) *n_objs = ctx->n_objs;
    if (order) *order = ctx->order;
    if (traj_len) *traj_len = ctx->traj_len;
    if (time_len) *time_len = ctx->time_len;
    if (maxT) *maxT = ctx->maxT;
    if (epoch) {
        epoch[0] = ctx->pj_epoch[0];
        epoch[1] = ctx->pj_epoch[1];
    }
    return MZH_OK;
}

int mzh_copy_polyjectory(mzh_ctx *ctx, double *traj, double *time, uint64_t *traj_off, uint64_t *nsteps,
                         uint64_t *time_off, int32_t *status)
{
    return guarded(ctx, [&] {
        if (!ctx->has_pj) throw StatusError{MZH_EINVAL, "No polyjectory has been uploaded to this context"};
        if (traj && ctx->traj_len)
            MZH_HIP(hipMemcpy(traj, ctx->pjb->traj.p, sizeof(double) * ctx->traj_len, hipMemcpyDeviceToHost));
        if (time && ctx->time_len)
            MZH_HIP(hipMemcpy(time, ctx->pjb->time.p, sizeof(double) * ctx->time_len, hipMemcpyDeviceToHost));
        const size_t n = ctx->n_objs;
        if (traj_off) std::copy(ctx->h_traj_off.begin(), ctx->h_traj_off.end(), traj_off);
        if (nsteps) std::copy(ctx->h_nsteps.begin(), ctx->h_nsteps.end(), nsteps);
        if (time_off) std::copy(ctx->h_time_off.begin(), ctx->h_time_off.end(), time_off);
        if (status) std::copy(ctx->h_status.begin(), ctx->h_status.begin() + n, status);
    });
}

namespace
{

// polyjectory::state_eval / state_meval (polyjectory.cpp:1094-1366).
void state_eval_impl(mzh_ctx *ctx, const double *tm, int time_mode, const uint64_t *sel, uint64_t n_rows,
                     uint64_t n_times, double *out, bool on_device)
{
    if (!ctx->has_pj) throw StatusError{MZH_EINVAL, "No polyjectory has been uploaded to this context"};
    if (time_mode < MZH_EVAL_TIME_SCALAR || time_mode > MZH_EVAL_TIME_PER_EVAL)
        throw StatusError{MZH_EINVAL, "Invalid time mode " + std::to_string(time_mode) + " passed to state_eval()"};
    if ((time_mode == MZH_EVAL_TIME_SCALAR || time_mode == MZH_EVAL_TIME_PER_OBJECT) && n_times != 1)
        throw StatusError{MZH_EINVAL, "state_eval() evaluates one time per object, but " + std::to_string(n_times)
                                          + " time evaluations per object were requested"};
    if (!sel && n_rows != ctx->n_objs)
        throw StatusError{MZH_EINVAL, "Invalid output array passed to state_eval(): the number of objects is "
                                          + std::to_string(ctx->n_objs)
                                          + " but the size of the first dimension of the array is "
                                          + std::to_string(n_rows) + " (the two numbers must be equal)"};
    if (n_times && n_rows > std::numeric_limits<uint64_t>::max() / 7u / n_times)
        throw StatusError{MZH_EINVAL, "Too many evaluations requested from state_eval()"};
    const uint64_t n_eval = n_rows * n_times;
    if (n_eval == 0) return;
    if (!tm || !out) throw StatusError{MZH_EINVAL, "Null pointer passed to state_eval()"};
    const uint64_t n_tm = time_mode == MZH_EVAL_TIME_SCALAR       ? 1u
                          : time_mode == MZH_EVAL_TIME_PER_OBJECT ? n_rows
                          : time_mode == MZH_EVAL_TIME_SHARED     ? n_times
                                                                  : n_eval;
    hipStream_t st = ctx->stream;
    const double *d_tm = tm;
    const uint64_t *d_sel = sel;
    double *d_out = out;
    if (!on_device) {
        MZH_HIP(ctx->ev_tm.ensure(sizeof(double) * n_tm));
        MZH_HIP(ctx->ev_out.ensure(sizeof(double) * 7u * n_eval));
        MZH_HIP(hipMemcpyAsync(ctx->ev_tm.p, tm, sizeof(double) * n_tm, hipMemcpyHostToDevice, st));
        d_tm = ctx->ev_tm.as<const double>();
        d_out = ctx->ev_out.as<double>();
        if (sel) {
            MZH_HIP(ctx->ev_sel.ensure(sizeof(uint64_t) * n_rows));
            MZH_HIP(hipMemcpyAsync(ctx->ev_sel.p, sel, sizeof(uint64_t) * n_rows, hipMemcpyHostToDevice, st));
            d_sel = ctx->ev_sel.as<const uint64_t>();
        }
    }
    MZH_HIP(ctx->ev_bad.ensure(sizeof(unsigned long long)));
    MZH_HIP(hipMemsetAsync(ctx->ev_bad.p, 0xff, sizeof(unsigned long long), st));
    MZH_HIP(hipEventRecord(ctx->ev_eval[0], st));
    MZH_HIP(launch_state_eval(pj_view(ctx), d_tm, time_mode, d_sel, n_rows, n_times, d_out,
                              ctx->ev_bad.as<unsigned long long>(), st));
    MZH_HIP(hipEventRecord(ctx->ev_eval[1], st));
    unsigned long long bad = 0;
    MZH_HIP(hipMemcpyAsync(&bad, ctx->ev_bad.p, sizeof(bad), hipMemcpyDeviceToHost, st));
    MZH_HIP(hipStreamSynchronize(st));
    float ms = 0;
    MZH_HIP(hipEventElapsedTime(&ms, ctx->ev_eval[0], ctx->ev_eval[1]));
    ctx->eval_s = ms * 1e-3;
    if (bad != ~0ull) {
        // the first failing evaluation: its time is checked before its index
        const uint64_t r = bad / n_times, k = bad - r * n_times;
        const uint64_t ti = time_mode == MZH_EVAL_TIME_SCALAR       ? 0u
                            : time_mode == MZH_EVAL_TIME_PER_OBJECT ? r
                            : time_mode == MZH_EVAL_TIME_SHARED     ? k
                                                                    : bad;
        double t = 0;
        uint64_t obj = r;
        if (on_device) {
            MZH_HIP(hipMemcpy(&t, tm + ti, sizeof(double), hipMemcpyDeviceToHost));
            if (sel) MZH_HIP(hipMemcpy(&obj, sel + r, sizeof(uint64_t), hipMemcpyDeviceToHost));
        } else {
            t = tm[ti];
            if (sel) obj = sel[r];
        }
        if (!std::isfinite(t))
            throw StatusError{MZH_EINVAL, "An non-finite evaluation time of " + fmt_double(t)
                                              + " was detected during the evaluation of a polyjectory"};
        throw StatusError{MZH_ERANGE, "Invalid object index " + std::to_string(obj)
                                          + " specified - the total number of objects in the polyjectory is only "
                                          + std::to_string(ctx->n_objs)};
    }
    if (!on_device)
        MZH_HIP(hipMemcpy(out, ctx->ev_out.p, sizeof(double) * 7u * n_eval, hipMemcpyDeviceToHost));
}

} // namespace

int mzh_state_eval(mzh_ctx *ctx, const double *tm, int time_mode, const uint64_t *sel, uint64_t n_rows,
                   uint64_t n_times, double *out)
{
    return guarded(ctx, [&] { state_eval_impl(ctx, tm, time_mode, sel, n_rows, n_times, out, false); });
}

int mzh_state_eval_device(mzh_ctx *ctx, const double *tm, int time_mode, const uint64_t *sel, uint64_t n_rows,
                          uint64_t n_times, double *out)
{
    return guarded(ctx, [&] { state_eval_impl(ctx, tm, time_mode, sel, n_rows, n_times, out, true); });
}

int mzh_device_alloc(mzh_ctx *ctx, uint64_t bytes, void **ptr)
{
    if (!ptr) return MZH_EINVAL;
    *ptr = nullptr;
    return guarded(ctx, [&] { MZH_HIP(hipMalloc(ptr, bytes ? bytes : 1)); });
}

int mzh_device_free(mzh_ctx *ctx, void *ptr)
{
    return guarded(ctx, [&] {
        if (ptr) {
            MZH_HIP(hipStreamSynchronize(ctx->stream));
            MZH_HIP(hipFree(ptr));
        }
    });
}

int mzh_copy_to_device(mzh_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    return guarded(ctx, [&] {
        if (!bytes) return;
        if (!dst || !src) throw StatusError{MZH_EINVAL, "Null pointer passed to mzh_copy_to_device"};
        MZH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        MZH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int mzh_copy_from_device(mzh_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    return guarded(ctx, [&] {
        if (!bytes) return;
        if (!dst || !src) throw StatusError{MZH_EINVAL, "Null pointer passed to mzh_copy_from_device"};
        MZH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        MZH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int mzh_get_sgp4_times(const mzh_ctx *ctx, double seconds[5])
{
    if (!ctx || !seconds) return MZH_EINVAL;
    for (int k = 0; k < 5; ++k) seconds[k] = ctx->sgp4_s[k];
    return MZH_OK;
}

int mzh_get_eval_time(const mzh_ctx *ctx, double *seconds)
{
    if (!ctx || !seconds) return MZH_EINVAL;
    *seconds = ctx->eval_s;
    return MZH_OK;
}

int mzh_check_screen(uint64_t n_objs, double maxT, double conj_thresh, double conj_det_interval,
                     const int32_t *otypes, int mode, uint64_t *n_cd_steps)
{
    try {
        if (n_objs > 0xffffffffull)
            throw StatusError{MZH_EINVAL, "Too many objects for one screening: " + std::to_string(n_objs)};
        if (!std::isfinite(maxT) || maxT <= 0) throw StatusError{MZH_EINVAL, "Invalid maxT " + fmt_double(maxT)};
        // the arguments and the step count, then the limit of one batch, then
        // (within it) every step's time range
        const ScreenPlan pl =
            plan_screen((uint32_t)n_objs, maxT, conj_thresh, conj_det_interval, otypes, mode, false, 0, 0);
        if (n_cd_steps) *n_cd_steps = pl.C;
        check_batch_size(n_objs, pl.C);
        (void)plan_screen((uint32_t)n_objs, maxT, conj_thresh, conj_det_interval, otypes, mode);
        return MZH_OK;
    } catch (const StatusError &e) {
        g_global_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_global_error = "host memory allocation failed";
        return MZH_ENOMEM;
    }
}

int mzh_root_finding(mzh_ctx *ctx, uint32_t order, uint64_t n, const double *polys, const double *rf_int,
                     double *roots, uint32_t *n_roots, uint32_t *flags)
{
    return guarded(ctx, [&] {
        if (order < (uint32_t)kMinOrder || order > (uint32_t)kMaxOrder)
            throw StatusError{MZH_EINVAL, "Polynomial order " + std::to_string(order) + " is not supported"};
        if (n == 0) return;
        if (!polys || !rf_int || !roots || !n_roots) throw StatusError{MZH_EINVAL, "Null pointer passed to mzh_root_finding"};
        const uint32_t op1 = order + 1, nmax = 2u * order + 2u;
        hipStream_t st = ctx->stream;
        DevBuf d_p, d_rf, d_roots, d_cnt_r, d_flags, d_slow, d_cnt, d_wl, d_err;
        MZH_HIP(d_p.ensure(sizeof(double) * op1 * n));
        MZH_HIP(d_rf.ensure(sizeof(double) * n));
        MZH_HIP(d_roots.ensure(sizeof(double) * nmax * n));
        MZH_HIP(d_cnt_r.ensure(sizeof(uint32_t) * n));
        MZH_HIP(d_flags.ensure(sizeof(uint32_t) * n));
        MZH_HIP(d_slow.ensure(sizeof(uint64_t) * n));
        MZH_HIP(d_cnt.ensure(sizeof(unsigned long long)));
        MZH_HIP(d_wl.ensure(narrow_wl_slow_bytes(order)));
        MZH_HIP(d_err.ensure(sizeof(DevError)));
        MZH_HIP(hipMemcpyAsync(d_p.p, polys, sizeof(double) * op1 * n, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(d_rf.p, rf_int, sizeof(double) * n, hipMemcpyHostToDevice, st));
        RootFindArgs a;
        a.polys = d_p.as<const double>();
        a.rf_int = d_rf.as<const double>();
        a.n = n;
        a.roots = d_roots.as<double>();
        a.n_roots = d_cnt_r.as<uint32_t>();
        a.flags = d_flags.as<uint32_t>();
        a.slow = d_slow.as<uint64_t>();
        a.slow_count = d_cnt.as<unsigned long long>();
        a.wl_slow = d_wl.p;
        a.wl_fast_cap = (int)std::min<uint64_t>(env_u64("MIZUBA_AMD_WL_FAST_CAP", kNarrowWlFastCap), kNarrowWlFastCap);
        a.err = d_err.as<DevError>();
        MZH_HIP(launch_rootfind(order, a, st));
        DevError herr;
        MZH_HIP(hipMemcpyAsync(&herr, d_err.p, sizeof(DevError), hipMemcpyDeviceToHost, st));
        MZH_HIP(hipStreamSynchronize(st));
        if (herr.code != kErrNone) throw StatusError{MZH_EINVAL, dev_error_message(herr)};
        MZH_HIP(hipMemcpy(n_roots, d_cnt_r.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        MZH_HIP(hipMemcpy(roots, d_roots.p, sizeof(double) * nmax * n, hipMemcpyDeviceToHost));
        if (flags) MZH_HIP(hipMemcpy(flags, d_flags.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    });
}

int mzh_root_finding_noalloc(mzh_ctx *ctx, uint32_t order, uint64_t n, const double *polys, const double *rf_int,
                             double *roots, uint32_t *n_roots, uint32_t *flags)
{
    return guarded(ctx, [&] {
        if (order < (uint32_t)kMinOrder || order > (uint32_t)kMaxOrder)
            throw StatusError{MZH_EINVAL, "Polynomial order " + std::to_string(order) + " is not supported"};
        if (n == 0) return;
        if (!polys || !rf_int || !roots || !n_roots)
            throw StatusError{MZH_EINVAL, "Null pointer passed to mzh_root_finding_noalloc"};
        if (n > noalloc_max_polys(order))
            throw StatusError{MZH_EINVAL, "mzh_root_finding_noalloc: " + std::to_string(n)
                                              + " polynomials of order " + std::to_string(order)
                                              + " exceed the " + std::to_string(noalloc_max_polys(order))
                                              + " that its fixed scratch of "
                                              + std::to_string(kNoAllocScratchBytes >> 20) + " MiB holds"};
        uint8_t *scr = nullptr;
        MZH_HIP(scratch_address(ctx->device, reinterpret_cast<void **>(&scr)));
        hipStream_t st = ctx->stream;
        RootFindArgs a;
        a.polys = polys;
        a.rf_int = rf_int;
        a.n = n;
        a.roots = roots;
        a.n_roots = n_roots;
        a.flags = flags;
        a.slow_count = reinterpret_cast<unsigned long long *>(scr + kScrCount);
        a.err = reinterpret_cast<DevError *>(scr + kScrErr);
        a.wl_slow = scr + kScrLists;
        a.slow = reinterpret_cast<uint64_t *>(scr + scr_queue_offset(order));
        a.wl_fast_cap = (int)std::min<uint64_t>(env_u64("MIZUBA_AMD_WL_FAST_CAP", kNarrowWlFastCap), kNarrowWlFastCap);
        // the launch clears the counter and the error word first
        MZH_HIP(launch_rootfind(order, a, st));
        DevError herr;
        MZH_HIP(hipMemcpyAsync(&herr, a.err, sizeof(DevError), hipMemcpyDeviceToHost, st));
        MZH_HIP(hipStreamSynchronize(st));
        if (herr.code != kErrNone) throw StatusError{MZH_EINVAL, dev_error_message(herr)};
    });
}

int mzh_screen(mzh_ctx *ctx, double conj_thresh, double conj_det_interval, const int32_t *otypes, int mode)
{
    return guarded(ctx, [&] { run_screen(ctx, conj_thresh, conj_det_interval, otypes, mode); });
}

int mzh_screen_steps(mzh_ctx *ctx, double conj_thresh, double conj_det_interval, const int32_t *otypes, int mode,
                     uint64_t step_begin, uint64_t step_end)
{
    return guarded(ctx, [&] {
        run_screen_steps(ctx, conj_thresh, conj_det_interval, otypes, mode, step_begin, step_end);
    });
}

int mzh_device_memory(mzh_ctx *ctx, uint64_t *free_bytes, uint64_t *total_bytes)
{
    return guarded(ctx, [&] {
        if (!free_bytes || !total_bytes) throw StatusError{MZH_EINVAL, "Null pointer"};
        MZH_HIP(hipSetDevice(ctx->device));
        size_t f = 0, t = 0;
        MZH_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = f;
        *total_bytes = t;
    });
}

int mzh_chunk_steps(uint64_t n_objs, uint64_t n_cd_steps, uint64_t device_bytes, uint64_t *chunk_steps)
{
    if (!chunk_steps || n_cd_steps == 0) return MZH_EINVAL;
    const uint64_t n = n_objs ? n_objs : 1;
    uint64_t k = n_cd_steps;
    k = std::min<uint64_t>(k, 0xffffffffull / n);
    k = std::min<uint64_t>(k, device_bytes / ((uint64_t)MZH_SCREEN_BYTES_PER_OBJECT_STEP * n));
    *chunk_steps = std::max<uint64_t>(k, 1);
    return MZH_OK;
}

} // extern "C"

// A batch handle: the problems, their screening arguments and the batch that
// screens them (with its captured graph and launch shapes).
struct mzh_batch {
    std::vector<mzh_ctx *> ctxs;
    std::vector<double> thresh, interval;
    std::vector<std::vector<int32_t>> ot;
    std::vector<char> has_ot;
    std::vector<uint64_t> gen_after_run;
    int mode = MZH_MODE_FP64;
    int flags = 0;
    bool tuned = false;
    std::unique_ptr<ScreenBatch, BatchDeleter> b;
};

namespace
{

// The arguments of a batch: contexts distinct, non-null, on one device, each
// with a polyjectory. Returns MZH_OK or MZH_EINVAL with g_global_error set.
int check_batch_args(mzh_ctx *const *ctxs, uint64_t n, const double *conj_thresh, const double *conj_det_interval)
{
    if (!ctxs || !conj_thresh || !conj_det_interval) {
        g_global_error = "Null pointer passed to a batched screening";
        return MZH_EINVAL;
    }
    std::vector<const mzh_ctx *> seen(ctxs, ctxs + n);
    std::sort(seen.begin(), seen.end());
    if (!seen.front() || std::adjacent_find(seen.begin(), seen.end()) != seen.end()) {
        g_global_error = "The contexts of a batched screening must be distinct and non-null";
        return MZH_EINVAL;
    }
    for (uint64_t k = 0; k < n; ++k)
        if (ctxs[k]->device != ctxs[0]->device) {
            g_global_error = "The contexts of a batched screening must be on one device";
            return MZH_EINVAL;
        }
    return MZH_OK;
}

// Validate and set the plan of problem k; its status (message in the context).
int plan_one(mzh_ctx *ctx, double thresh, double interval, const int32_t *ot, int mode)
{
    try {
        if (!ctx->has_pj) throw StatusError{MZH_EINVAL, "No polyjectory has been uploaded to this context"};
        set_plan(ctx, plan_screen(ctx->n_objs, ctx->maxT, thresh, interval, ot, mode), thresh, interval, mode);
        return MZH_OK;
    } catch (const StatusError &e) {
        return fail(ctx, e.code, e.msg);
    } catch (const std::bad_alloc &) {
        return fail(ctx, MZH_ENOMEM, "host memory allocation failed");
    }
}

// Screen the planned problems among ctxs (status[k] == MZH_OK) with batch b.
void run_planned(ScreenBatch *b, mzh_ctx *const *ctxs, uint64_t n, bool graph, bool tune, int *status)
{
    std::vector<mzh_ctx *> todo;
    std::vector<uint64_t> where;
    for (uint64_t k = 0; k < n; ++k)
        if (status[k] == MZH_OK) {
            todo.push_back(ctxs[k]);
            where.push_back(k);
        }
    if (todo.empty()) return;
    std::vector<int> st(todo.size(), MZH_OK);
    try {
        MZH_HIP(hipSetDevice(todo[0]->device));
        batch_run(b, todo, graph, tune, st.data());
    } catch (const StatusError &e) {
        for (size_t k = 0; k < todo.size(); ++k) st[k] = fail(todo[k], e.code, e.msg);
    } catch (const std::bad_alloc &) {
        for (size_t k = 0; k < todo.size(); ++k) st[k] = fail(todo[k], MZH_ENOMEM, "host memory allocation failed");
    }
    for (size_t k = 0; k < todo.size(); ++k) status[where[k]] = st[k];
}

} // namespace

extern "C" {

int mzh_screen_batch(mzh_ctx *const *ctxs, uint64_t n, const double *conj_thresh, const double *conj_det_interval,
                     const int32_t *const *otypes, int mode, int *status)
{
    if (n == 0) return MZH_OK;
    if (!status) return MZH_EINVAL;
    const int chk = check_batch_args(ctxs, n, conj_thresh, conj_det_interval);
    if (chk != MZH_OK) return chk;
    for (uint64_t k = 0; k < n; ++k)
        status[k] = plan_one(ctxs[k], conj_thresh[k], conj_det_interval[k], otypes ? otypes[k] : nullptr, mode);
    // every valid problem in one batch on the first context's stream
    std::unique_ptr<ScreenBatch, BatchDeleter> b;
    try {
        MZH_HIP(hipSetDevice(ctxs[0]->device));
        b.reset(batch_new(ctxs[0]->stream));
    } catch (const StatusError &e) {
        for (uint64_t k = 0; k < n; ++k)
            if (status[k] == MZH_OK) status[k] = fail(ctxs[k], e.code, e.msg);
    }
    if (b) run_planned(b.get(), ctxs, n, false, false, status);
    for (uint64_t k = 0; k < n; ++k)
        if (status[k] != MZH_OK) return status[k];
    return MZH_OK;
}

int mzh_batch_create(mzh_ctx *const *ctxs, uint64_t n, const double *conj_thresh, const double *conj_det_interval,
                     const int32_t *const *otypes, int mode, int flags, mzh_batch **out)
{
    if (!out) return MZH_EINVAL;
    *out = nullptr;
    if (n == 0) {
        g_global_error = "A batch needs at least one problem";
        return MZH_EINVAL;
    }
    const int chk = check_batch_args(ctxs, n, conj_thresh, conj_det_interval);
    if (chk != MZH_OK) return chk;
    if (mode != MZH_MODE_FP64 && mode != MZH_MODE_MIXED) {
        g_global_error = "Invalid screening mode " + std::to_string(mode);
        return MZH_EINVAL;
    }
    try {
        std::unique_ptr<mzh_batch> bt(new mzh_batch());
        bt->ctxs.assign(ctxs, ctxs + n);
        bt->thresh.assign(conj_thresh, conj_thresh + n);
        bt->interval.assign(conj_det_interval, conj_det_interval + n);
        bt->ot.resize(n);
        bt->has_ot.assign(n, 0);
        bt->gen_after_run.assign(n, ~0ull);
        for (uint64_t k = 0; k < n; ++k) {
            if (!ctxs[k]->has_pj) {
                g_global_error = "No polyjectory has been uploaded to the context of problem " + std::to_string(k);
                return MZH_EINVAL;
            }
            const int32_t *ot = otypes ? otypes[k] : nullptr;
            if (ot) {
                bt->ot[k].assign(ot, ot + ctxs[k]->n_objs);
                bt->has_ot[k] = 1;
            }
            // the arguments are checked now, as conjunctions() would
            (void)plan_screen(ctxs[k]->n_objs, ctxs[k]->maxT, conj_thresh[k], conj_det_interval[k], ot, mode);
        }
        bt->mode = mode;
        bt->flags = flags;
        MZH_HIP(hipSetDevice(ctxs[0]->device));
        bt->b.reset(batch_new(ctxs[0]->stream));
        *out = bt.release();
        return MZH_OK;
    } catch (const StatusError &e) {
        g_global_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_global_error = "host memory allocation failed";
        return MZH_ENOMEM;
    }
}

int mzh_batch_run(mzh_batch *bt, int *status)
{
    if (!bt || !status) return MZH_EINVAL;
    const uint64_t n = bt->ctxs.size();
    for (uint64_t k = 0; k < n; ++k) {
        mzh_ctx *ctx = bt->ctxs[k];
        status[k] = MZH_OK;
        // plan again only when the context was touched since this batch's run
        if (ctx->generation != bt->gen_after_run[k] || ctx->last_batch != bt->b.get())
            status[k] = plan_one(ctx, bt->thresh[k], bt->interval[k], bt->has_ot[k] ? bt->ot[k].data() : nullptr,
                                 bt->mode);
    }
    const bool tune = (bt->flags & MZH_BATCH_TUNE) && !bt->tuned;
    run_planned(bt->b.get(), bt->ctxs.data(), n, (bt->flags & MZH_BATCH_GRAPH) != 0, tune, status);
    bt->tuned = bt->tuned || tune;
    for (uint64_t k = 0; k < n; ++k) bt->gen_after_run[k] = bt->ctxs[k]->generation;
    for (uint64_t k = 0; k < n; ++k)
        if (status[k] != MZH_OK) return status[k];
    return MZH_OK;
}

int mzh_batch_info(const mzh_batch *bt, uint64_t *replays, double *seconds, uint32_t shape[5])
{
    if (!bt) return MZH_EINVAL;
    if (replays) *replays = batch_replays(bt->b.get());
    if (seconds) *seconds = batch_seconds(bt->b.get());
    if (shape) {
        const LaunchShape ls = batch_shape(bt->b.get());
        shape[0] = ls.bucket_bits;
        shape[1] = ls.sort_wgs;
        shape[2] = ls.bvh_wgs;
        shape[3] = ls.bp_block;
        shape[4] = ls.roots_wgs;
    }
    return MZH_OK;
}

int mzh_batch_destroy(mzh_batch *bt)
{
    if (!bt) return MZH_OK;
    if (!bt->ctxs.empty()) (void)hipSetDevice(bt->ctxs[0]->device);
    if (!bt->ctxs.empty()) (void)hipStreamSynchronize(bt->ctxs[0]->stream);
    delete bt;
    return MZH_OK;
}

int mzh_get_n_cd_steps(const mzh_ctx *ctx, uint64_t *out)
{
    if (!ctx || !out) return MZH_EINVAL;
    if (!ctx->screened) return fail(const_cast<mzh_ctx *>(ctx), MZH_EINVAL, "No screening results are available");
    *out = ctx->C;
    return MZH_OK;
}

int mzh_get_n_objs(const mzh_ctx *ctx, uint64_t *out)
{
    if (!ctx || !out) return MZH_EINVAL;
    *out = ctx->n_objs;
    return MZH_OK;
}

int mzh_copy_cd_end_times(mzh_ctx *ctx, double *dst)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        std::memcpy(dst, ctx->h_cd_end.data(), sizeof(double) * ctx->C);
    });
}

namespace
{
int copy_dev(mzh_ctx *ctx, void *dst, const DevBuf &src, size_t bytes)
{
    return guarded(ctx, [&] {
        require_results(ctx);
        if (!dst && bytes) throw StatusError{MZH_EINVAL, "Null destination"};
        if (bytes) MZH_HIP(hipMemcpy(dst, src.p, bytes, hipMemcpyDeviceToHost));
    });
}
} // namespace

int mzh_copy_aabbs(mzh_ctx *ctx, float *dst)
{
    if (!ctx) return MZH_EINVAL;
    return copy_dev(ctx, dst, ctx->aabbs, sizeof(float) * 8u * (size_t)(ctx->n_objs + 1) * ctx->C);
}

int mzh_copy_srt_aabbs(mzh_ctx *ctx, float *dst)
{
    if (!ctx) return MZH_EINVAL;
    return copy_dev(ctx, dst, ctx->srt_aabbs, sizeof(float) * 8u * (size_t)(ctx->n_objs + 1) * ctx->C);
}

int mzh_copy_mcodes(mzh_ctx *ctx, uint64_t *dst)
{
    if (!ctx) return MZH_EINVAL;
    return copy_dev(ctx, dst, ctx->mcodes, sizeof(uint64_t) * (size_t)ctx->n_objs * ctx->C);
}

int mzh_copy_srt_mcodes(mzh_ctx *ctx, uint64_t *dst)
{
    if (!ctx) return MZH_EINVAL;
    return copy_dev(ctx, dst, ctx->srt_mcodes, sizeof(uint64_t) * (size_t)ctx->n_objs * ctx->C);
}

int mzh_copy_srt_idx(mzh_ctx *ctx, uint32_t *dst)
{
    if (!ctx) return MZH_EINVAL;
    return copy_dev(ctx, dst, ctx->srt_idx, sizeof(uint32_t) * (size_t)ctx->n_objs * ctx->C);
}

int mzh_get_bvh_tree_size(mzh_ctx *ctx, uint64_t step, uint64_t *size)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        if (step >= ctx->C)
            throw StatusError{MZH_ERANGE, "Cannot fetch the BVH tree for the conjunction timestep at index "
                                              + std::to_string(step) + ": the total number of conjunction steps is only "
                                              + std::to_string(ctx->C)};
        *size = ctx->h_tree_size[step];
    });
}

int mzh_copy_bvh_tree(mzh_ctx *ctx, uint64_t step, mzh_bvh_node *dst)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        if (step >= ctx->C)
            throw StatusError{MZH_ERANGE, "Cannot fetch the BVH tree for the conjunction timestep at index "
                                              + std::to_string(step) + ": the total number of conjunction steps is only "
                                              + std::to_string(ctx->C)};
        const size_t n = ctx->h_tree_size[step];
        if (n) {
            require_results(ctx);
            export_nodes(ctx);
            MZH_HIP(hipMemcpy(dst, ctx->nodes.as<mzh_bvh_node>() + step * ctx->node_cap, sizeof(mzh_bvh_node) * n,
                              hipMemcpyDeviceToHost));
        }
    });
}

int mzh_get_aabb_collisions_size(mzh_ctx *ctx, uint64_t step, uint64_t *size)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        if (step >= ctx->C)
            throw StatusError{MZH_ERANGE,
                              "Cannot fetch the list of AABB collisions for the conjunction timestep at index "
                                  + std::to_string(step) + ": the total number of conjunction steps is only "
                                  + std::to_string(ctx->C)};
        *size = ctx->h_step_off[step + 1] - ctx->h_step_off[step];
    });
}

int mzh_copy_aabb_collisions(mzh_ctx *ctx, uint64_t step, mzh_aabb_collision *dst)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        if (step >= ctx->C)
            throw StatusError{MZH_ERANGE,
                              "Cannot fetch the list of AABB collisions for the conjunction timestep at index "
                                  + std::to_string(step) + ": the total number of conjunction steps is only "
                                  + std::to_string(ctx->C)};
        const uint64_t a = ctx->h_step_off[step], n = ctx->h_step_off[step + 1] - a;
        if (n) require_results(ctx);
        if (n)
            MZH_HIP(hipMemcpy(dst, ctx->pairs.as<mzh_aabb_collision>() + a, sizeof(mzh_aabb_collision) * n,
                              hipMemcpyDeviceToHost));
    });
}

int mzh_copy_bp_offsets(mzh_ctx *ctx, uint64_t *dst)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        for (uint64_t k = 0; k < ctx->C; ++k) {
            dst[2 * k] = ctx->h_step_off[k];
            dst[2 * k + 1] = ctx->h_step_off[k + 1] - ctx->h_step_off[k];
        }
    });
}

int mzh_get_n_conjunctions(mzh_ctx *ctx, uint64_t *out)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        *out = ctx->n_conj;
    });
}

int mzh_copy_conjunctions(mzh_ctx *ctx, mzh_conj *dst)
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        if (ctx->n_conj) require_results(ctx);
        if (ctx->n_conj)
            MZH_HIP(hipMemcpy(dst, ctx->conj_sorted.p, sizeof(mzh_conj) * ctx->n_conj, hipMemcpyDeviceToHost));
    });
}

int mzh_get_np_report(mzh_ctx *ctx, uint64_t out[MZH_NP_COUNT])
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        for (int k = 0; k < MZH_NP_COUNT; ++k) out[k] = ctx->h_np[k];
    });
}

int mzh_get_stage_times(mzh_ctx *ctx, double out[MZH_STAGE_COUNT])
{
    return guarded(ctx, [&] {
        require_screened(ctx);
        for (int k = 0; k < MZH_STAGE_COUNT; ++k) out[k] = ctx->stage_s[k];
    });
}

int mzh_device_view(mzh_ctx *ctx, int which, void **ptr, uint64_t *bytes)
{
    return guarded(ctx, [&] {
        require_results(ctx);
        const size_t NC = (size_t)ctx->n_objs * ctx->C;
        switch (which) {
            case 0:
                *ptr = ctx->aabbs.p;
                *bytes = sizeof(float) * 8u * (size_t)(ctx->n_objs + 1) * ctx->C;
                break;
            case 1:
                *ptr = ctx->srt_aabbs.p;
                *bytes = sizeof(float) * 8u * (size_t)(ctx->n_objs + 1) * ctx->C;
                break;
            case 2:
                *ptr = ctx->mcodes.p;
                *bytes = sizeof(uint64_t) * NC;
                break;
            case 3:
                *ptr = ctx->srt_mcodes.p;
                *bytes = sizeof(uint64_t) * NC;
                break;
            case 4:
                *ptr = ctx->srt_idx.p;
                *bytes = sizeof(uint32_t) * NC;
                break;
            case 5:
                export_nodes(ctx);
                *ptr = ctx->nodes.p;
                *bytes = sizeof(mzh_bvh_node) * ctx->node_cap * ctx->C;
                break;
            case 6:
                *ptr = ctx->pairs.p;
                *bytes = sizeof(mzh_aabb_collision) * ctx->n_pairs;
                break;
            case 7:
                *ptr = ctx->conj_sorted.p;
                *bytes = sizeof(mzh_conj) * ctx->n_conj;
                break;
            default:
                throw StatusError{MZH_EINVAL, "Invalid device view index " + std::to_string(which)};
        }
    });
}

int mzh_share_polyjectory(mzh_ctx *dst, const mzh_ctx *src)
{
    return guarded(dst, [&] {
        if (!src || !src->has_pj) throw StatusError{MZH_EINVAL, "The source context holds no polyjectory"};
        if (src == dst) return;
        if (src->device != dst->device)
            throw StatusError{MZH_EINVAL, "Contexts on different devices cannot share a polyjectory"};
        dst->screened = false;
        ++dst->generation;
        dst->pjb = src->pjb;
        dst->n_objs = src->n_objs;
        dst->order = src->order;
        dst->maxT = src->maxT;
        dst->n_steps_total = src->n_steps_total;
        dst->h_traj_off = src->h_traj_off;
        dst->h_nsteps = src->h_nsteps;
        dst->h_time_off = src->h_time_off;
        dst->h_status = src->h_status;
        dst->traj_len = src->traj_len;
        dst->time_len = src->time_len;
        dst->pj_epoch[0] = src->pj_epoch[0];
        dst->pj_epoch[1] = src->pj_epoch[1];
        dst->has_pj = true;
    });
}

int mzh_release(mzh_ctx *ctx, int what)
{
    return guarded(ctx, [&] {
        MZH_HIP(hipStreamSynchronize(ctx->stream));
        ++ctx->generation;
        if (what & MZH_RELEASE_SCRATCH) {
            for (DevBuf *b : {&ctx->kbke, &ctx->nprec, &ctx->keys_tmp, &ctx->vals_tmp, &ctx->sort_todo, &ctx->gbox,
                              &ctx->wtree, &ctx->bkt_end, &ctx->bvh_hdr, &ctx->bvh_scr, &ctx->otypes, &ctx->cnt, &ctx->head, &ctx->off,
                              &ctx->step_total, &ctx->slots, &ctx->ovf, &ctx->ovf_count, &ctx->pair_c, &ctx->ev_tm,
                              &ctx->ev_sel,
                              &ctx->ev_out, &ctx->ev_bad})
                b->release();
            // the narrow phase's queues of the context's own screenings
            ctx->solo.reset();
        }
        if (what & MZH_RELEASE_RESULTS) {
            for (DevBuf *b : {&ctx->aabbs, &ctx->srt_aabbs, &ctx->mcodes, &ctx->srt_mcodes, &ctx->srt_idx, &ctx->nodes,
                              &ctx->n_eff,
                              &ctx->trav, &ctx->tree_size, &ctx->pairs, &ctx->conj_sorted, &ctx->step_off,
                              &ctx->cd_begin, &ctx->cd_end})
                b->release();
            ctx->nodes_ready = false;
            if (ctx->screened) ctx->results_released = true;
        }
        if (what & MZH_RELEASE_POLYJECTORY) {
            ctx->pjb.reset();
            ctx->has_pj = false;
        }
    });
}

int mzh_synchronize(mzh_ctx *ctx)
{
    return guarded(ctx, [&] { MZH_HIP(hipStreamSynchronize(ctx->stream)); });
}

int mzh_stream_copy(mzh_ctx *ctx, uint64_t bytes, int launches, double *seconds)
{
    if (!ctx || bytes == 0 || bytes % 16u != 0 || launches < 1) return MZH_EINVAL;
    return guarded(ctx, [&] {
        DevBuf src, dst;
        MZH_HIP(src.ensure(bytes));
        MZH_HIP(dst.ensure(bytes));
        MZH_HIP(hipMemsetAsync(src.p, 1, bytes, ctx->stream));
        hipEvent_t e0, e1;
        MZH_HIP(hipEventCreate(&e0));
        MZH_HIP(hipEventCreate(&e1));
        const uint64_t n = bytes / 16u;
        const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255u) / 256u, 8192u);
        MZH_HIP(hipEventRecord(e0, ctx->stream));
        for (int k = 0; k < launches; ++k)
            hipLaunchKernelGGL(k_stream_copy, dim3(blocks), dim3(256), 0, ctx->stream, src.as<const uint4>(),
                               dst.as<uint4>(), n);
        MZH_HIP(hipGetLastError());
        MZH_HIP(hipEventRecord(e1, ctx->stream));
        MZH_HIP(hipEventSynchronize(e1));
        float ms = 0;
        MZH_HIP(hipEventElapsedTime(&ms, e0, e1));
        MZH_HIP(hipEventDestroy(e0));
        MZH_HIP(hipEventDestroy(e1));
        if (seconds) *seconds = 1e-3 * ms / launches;
    });
}

} // extern "C"
