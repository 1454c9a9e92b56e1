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
#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

namespace
{

using namespace mzh;

// Shortest decimal text that round-trips a double (the {} formatting of the
// reference's fmt messages).
std::string fmt_double(double v)
{
    if (std::isnan(v)) return "nan";
    if (std::isinf(v)) return v > 0 ? "inf" : "-inf";
    char buf[64];
    for (int prec = 1; prec <= 17; ++prec) {
        std::snprintf(buf, sizeof(buf), "%.*g", prec, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string s(buf);
    // fmt prints integral doubles without exponent and without a trailing '.0'
    return s;
}

// Positive integer from the environment, or dflt. Test hooks: the initial
// sizes of the growable screening buffers can be made small so that the
// resize-and-rerun paths run on small problems.
uint64_t env_u64(const char *name, uint64_t dflt)
{
    if (const char *v = std::getenv(name)) {
        const unsigned long long x = std::strtoull(v, nullptr, 10);
        if (x > 0) return x;
    }
    return dflt;
}

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t ensure(size_t b)
    {
        if (b <= bytes && p) return hipSuccess;
        if (p) {
            (void)hipFree(p);
            p = nullptr;
            bytes = 0;
        }
        const size_t alloc = b ? b : 256;
        hipError_t e = hipMalloc(&p, alloc);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        bytes = alloc;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    // Grow to at least b bytes, keeping the first `keep` bytes.
    hipError_t grow(size_t b, size_t keep, hipStream_t st)
    {
        if (b <= bytes && p) return hipSuccess;
        void *q = nullptr;
        hipError_t e = hipMalloc(&q, b);
        if (e != hipSuccess) return e;
        if (p && keep) {
            e = hipMemcpyAsync(q, p, std::min(keep, bytes), hipMemcpyDeviceToDevice, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) {
                (void)hipFree(q);
                return e;
            }
        }
        if (p) (void)hipFree(p);
        p = q;
        bytes = b;
        return hipSuccess;
    }
    template <typename T>
    T *as() const
    {
        return static_cast<T *>(p);
    }
};

struct StatusError {
    int code;
    std::string msg;
};

// Device arrays of a polyjectory: the AoS coefficients, the times, the
// step-major transposed positions and the per-object layout.
struct PjBufs {
    DevBuf traj, time, cf4, step0, nsteps, time_off;
};

} // namespace

struct mzh_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string last_error;

    // polyjectory: the device arrays may be shared by several contexts
    // (mzh_share_polyjectory); they are read-only once uploaded
    bool has_pj = false;
    uint32_t n_objs = 0, order = 0;
    double maxT = 0;
    uint64_t n_steps_total = 0;
    std::shared_ptr<PjBufs> pjb;
    // host copies of its layout, statuses and epoch (TAI Julian date) for copy-out
    std::vector<uint64_t> h_traj_off, h_nsteps, h_time_off;
    std::vector<int32_t> h_status;
    uint64_t traj_len = 0, time_len = 0;
    double pj_epoch[2] = {0, 0};

    // problem state
    bool screened = false;
    bool results_released = false;  // mzh_release(MZH_RELEASE_RESULTS) after the screen
    bool skip_cd = false;
    uint64_t C = 0;
    double conj_thresh = 0, conj_det_interval = 0;
    int mode = MZH_MODE_FP64;
    std::vector<double> h_cd_begin, h_cd_end;
    DevBuf cd_begin, cd_end;
    DevBuf aabbs, kbke, srt_aabbs, mcodes, srt_mcodes, srt_idx, keys_tmp, vals_tmp, n_eff, sort_todo, gbox;
    uint64_t node_cap = 0;
    DevBuf nodes, trav, bvh_hdr, bvh_scr, tree_size;
    bool nodes_ready = false;  // nodes holds the reference-layout export of trav
    DevBuf otypes, cnt, head, off, step_total, step_off, grand, slots, ovf, ovf_count;
    uint32_t ovf_cap = 0;
    DevBuf pairs;
    uint64_t n_pairs = 0;
    DevBuf conj_raw, conj_step, conj_sorted, conj_count, np_report, conj_keys;
    uint64_t conj_cap = 0, n_conj = 0;
    DevBuf cand, cand_count, steps, steps_count, slow, slow_count, wl_slow;
    uint64_t cand_cap = 0, steps_cap = 0, slow_cap = 0;
    DevBuf err;
    // polyjectory evaluation: staged times, selector and results, first bad element
    DevBuf ev_tm, ev_sel, ev_out, ev_bad;
    hipEvent_t ev_eval[2] = {};
    double eval_s = 0;  // device time of the last evaluation kernel
    // SGP4 generation: device times of its kernels (segment set-up, count,
    // scan, write, transpose) in the last mzh_make_sgp4_polyjectory
    hipEvent_t ev_sgp4[8] = {};
    double sgp4_s[5] = {0, 0, 0, 0, 0};
    hipEvent_t ev[12] = {};
    double stage_s[MZH_STAGE_COUNT] = {};
    uint64_t h_np[MZH_NP_COUNT] = {0, 0, 0, 0, 0, 0, 0};

    // host caches of per-step sizes (filled after a screen)
    std::vector<uint32_t> h_tree_size;
    std::vector<uint64_t> h_step_off;  // [C + 1]
};

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

#define MZH_HIP(call)                                                                                                  \
    do {                                                                                                               \
        const hipError_t e_ = (call);                                                                                  \
        if (e_ != hipSuccess)                                                                                          \
            throw StatusError{e_ == hipErrorOutOfMemory ? MZH_ENOMEM : MZH_ERUNTIME,                                   \
                              std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)};                    \
    } while (0)

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
        default:
            return "device error code " + std::to_string(e.code);
    }
}

// Device arrays for a new polyjectory: the context's own, or fresh ones when
// other contexts share the current arrays (they keep those).
void own_pj_bufs(mzh_ctx *ctx)
{
    if (!ctx->pjb || ctx->pjb.use_count() > 1) ctx->pjb = std::make_shared<PjBufs>();
}

PjView pj_view(const mzh_ctx *ctx)
{
    PjView v;
    v.traj = ctx->pjb->traj.as<const double>();
    v.time = ctx->pjb->time.as<const double>();
    v.cf4 = ctx->pjb->cf4.as<const double>();
    v.step0 = ctx->pjb->step0.as<const uint64_t>();
    v.nsteps = ctx->pjb->nsteps.as<const uint64_t>();
    v.time_off = ctx->pjb->time_off.as<const uint64_t>();
    v.n_objs = ctx->n_objs;
    v.order = ctx->order;
    return v;
}

// The arguments of one screening after validation.
struct ScreenPlan {
    uint64_t C = 0;
    bool skip_cd = true;
    std::vector<int32_t> ot;
    std::vector<double> cd_begin, cd_end;
};

// conjunctions::conjunctions validation (conjunctions.cpp:279-335) with the
// reference's messages, the step ranges of get_cd_begin_end
// (conjunctions.cpp:373-392), and this build's size limit: the per-step
// stages index (step, object) with 32-bit counts, so N * C must stay below
// 2^32. Everything is checked before any device memory is touched.
ScreenPlan plan_screen(uint32_t N, double maxT, double conj_thresh, double conj_det_interval,
                       const int32_t *otypes_in, int mode)
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
    if ((uint64_t)N * pl.C > 0xffffffffull)
        throw StatusError{MZH_EINVAL, "A screening of " + std::to_string(N) + " objects over " + std::to_string(pl.C)
                                          + " conjunction steps is too large: the number of objects times the number "
                                            "of conjunction steps must not exceed 4294967295"};
    // get_cd_begin_end for every step (conjunctions.cpp:373-392)
    pl.cd_begin.resize(pl.C);
    pl.cd_end.resize(pl.C);
    for (uint64_t k = 0; k < pl.C; ++k) {
        const double b = conj_det_interval * static_cast<double>(k);
        const double e = (k == pl.C - 1) ? maxT : conj_det_interval * static_cast<double>(k + 1);
        if (!std::isfinite(b) || !std::isfinite(e) || !(e > b) || b < 0 || e > maxT)
            throw StatusError{MZH_EINVAL, "Invalid conjunction step time range: [" + fmt_double(b) + ", "
                                              + fmt_double(e) + ")"};
        pl.cd_begin[k] = b;
        pl.cd_end[k] = e;
    }
    return pl;
}

// conjunctions::conjunctions validation + detect_conjunctions, batched.
void run_screen(mzh_ctx *ctx, double conj_thresh, double conj_det_interval, const int32_t *otypes_in, int mode)
{
    if (!ctx->has_pj) throw StatusError{MZH_EINVAL, "No polyjectory has been uploaded to this context"};
    const uint32_t N = ctx->n_objs;
    ScreenPlan pl = plan_screen(N, ctx->maxT, conj_thresh, conj_det_interval, otypes_in, mode);
    const uint64_t C = pl.C;
    const bool skip_cd = pl.skip_cd;
    const std::vector<int32_t> &ot = pl.ot;
    ctx->h_cd_begin = std::move(pl.cd_begin);
    ctx->h_cd_end = std::move(pl.cd_end);
    const size_t boxes_bytes = sizeof(float) * 8u * (size_t)(N + 1) * C;
    const size_t NC = (size_t)N * C;
    hipStream_t st = ctx->stream;
    ctx->screened = false;
    ctx->results_released = false;
    ctx->C = C;
    ctx->conj_thresh = conj_thresh;
    ctx->conj_det_interval = conj_det_interval;
    ctx->mode = mode;
    ctx->skip_cd = skip_cd;
    ctx->node_cap = 2ull * (N ? N : 1);

    MZH_HIP(ctx->cd_begin.ensure(sizeof(double) * C));
    MZH_HIP(ctx->cd_end.ensure(sizeof(double) * C));
    MZH_HIP(ctx->aabbs.ensure(boxes_bytes));
    MZH_HIP(ctx->kbke.ensure(sizeof(uint2) * NC));
    MZH_HIP(ctx->srt_aabbs.ensure(boxes_bytes));
    MZH_HIP(ctx->mcodes.ensure(sizeof(uint64_t) * NC));
    MZH_HIP(ctx->srt_mcodes.ensure(sizeof(uint64_t) * NC));
    MZH_HIP(ctx->keys_tmp.ensure(sizeof(uint64_t) * NC));
    MZH_HIP(ctx->srt_idx.ensure(sizeof(uint32_t) * NC));
    MZH_HIP(ctx->vals_tmp.ensure(sizeof(uint32_t) * NC));
    MZH_HIP(ctx->n_eff.ensure(sizeof(uint32_t) * C));
    MZH_HIP(ctx->sort_todo.ensure(sizeof(uint32_t) * C));
    MZH_HIP(ctx->gbox.ensure(sizeof(uint32_t) * 8u * C));
    MZH_HIP(ctx->trav.ensure(sizeof(TravNode) * ctx->node_cap * C));
    MZH_HIP(ctx->bvh_hdr.ensure(sizeof(uint4) * ctx->node_cap * C));
    MZH_HIP(ctx->bvh_scr.ensure(bvh_scratch_stride(N) * C));
    ctx->nodes_ready = false;
    MZH_HIP(ctx->tree_size.ensure(sizeof(uint32_t) * C));
    MZH_HIP(ctx->otypes.ensure(sizeof(int32_t) * (N ? N : 1)));
    MZH_HIP(ctx->cnt.ensure(sizeof(uint32_t) * NC));
    MZH_HIP(ctx->head.ensure(sizeof(uint32_t) * NC));
    MZH_HIP(ctx->slots.ensure(sizeof(uint32_t) * 4u * NC));
    if (ctx->ovf_cap == 0)
        ctx->ovf_cap = (uint32_t)std::min<uint64_t>(env_u64("MIZUBA_AMD_INIT_OVF_CAP", NC / 8 + 4096), 0x7fffffffu);
    MZH_HIP(ctx->ovf.ensure(sizeof(uint2) * ctx->ovf_cap));
    MZH_HIP(ctx->ovf_count.ensure(sizeof(uint32_t)));
    MZH_HIP(ctx->off.ensure(sizeof(uint64_t) * NC));
    MZH_HIP(ctx->step_total.ensure(sizeof(uint64_t) * C));
    MZH_HIP(ctx->step_off.ensure(sizeof(uint64_t) * (C + 1)));
    MZH_HIP(ctx->grand.ensure(sizeof(uint64_t)));
    MZH_HIP(ctx->err.ensure(sizeof(DevError)));
    MZH_HIP(ctx->conj_count.ensure(sizeof(unsigned long long)));
    MZH_HIP(ctx->np_report.ensure(sizeof(unsigned long long) * MZH_NP_COUNT));

    MZH_HIP(hipMemcpyAsync(ctx->cd_begin.p, ctx->h_cd_begin.data(), sizeof(double) * C, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemcpyAsync(ctx->cd_end.p, ctx->h_cd_end.data(), sizeof(double) * C, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemcpyAsync(ctx->otypes.p, ot.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemsetAsync(ctx->err.p, 0, sizeof(DevError), st));
    MZH_HIP(hipMemsetAsync(ctx->np_report.p, 0, sizeof(unsigned long long) * MZH_NP_COUNT, st));

    const PjView pj = pj_view(ctx);
    StepGrid sg;
    sg.cd_begin = ctx->cd_begin.as<const double>();
    sg.cd_end = ctx->cd_end.as<const double>();
    sg.n_steps = (uint32_t)C;
    DevError *derr = ctx->err.as<DevError>();

    MZH_HIP(hipEventRecord(ctx->ev[0], st));
    // stage 1: AABBs
    MZH_HIP(hipEventRecord(ctx->ev[6], st));
    MZH_HIP(launch_aabbs(pj, sg, conj_thresh, mode, ctx->aabbs.as<float>(), ctx->kbke.as<uint2>(),
                         ctx->gbox.as<uint32_t>(), derr, st));
    MZH_HIP(hipEventRecord(ctx->ev[7], st));
    MZH_HIP(launch_gbox_finalize(ctx->gbox.as<uint32_t>(), ctx->aabbs.as<float>(), N, (uint32_t)C, derr, st));
    MZH_HIP(hipEventRecord(ctx->ev[1], st));
    // stage 2: Morton codes, sort, gather
    MZH_HIP(launch_morton_sort(ctx->aabbs.as<float>(), ctx->mcodes.as<uint64_t>(), ctx->keys_tmp.as<uint64_t>(),
                               ctx->vals_tmp.as<uint32_t>(), ctx->srt_mcodes.as<uint64_t>(),
                               ctx->srt_idx.as<uint32_t>(), ctx->srt_aabbs.as<float>(), ctx->n_eff.as<uint32_t>(),
                               ctx->sort_todo.as<uint32_t>(), N, (uint32_t)C, st));
    MZH_HIP(hipEventRecord(ctx->ev[2], st));
    // stage 3: BVH
    MZH_HIP(launch_bvh(ctx->srt_mcodes.as<uint64_t>(), ctx->srt_aabbs.as<float>(), ctx->n_eff.as<uint32_t>(),
                       ctx->trav.as<TravNode>(), ctx->bvh_hdr.as<uint4>(), ctx->bvh_scr.as<uint8_t>(),
                       ctx->tree_size.as<uint32_t>(), N, ctx->node_cap, (uint32_t)C, st));
    MZH_HIP(hipEventRecord(ctx->ev[3], st));
    ctx->n_pairs = 0;
    ctx->n_conj = 0;
    if (!skip_cd) {
        // stage 4: broad phase
        uint32_t n_ovf = 0;
        for (int attempt = 0; attempt < 3; ++attempt) {
            MZH_HIP(launch_bp_reset(ctx->cnt.as<uint32_t>(), ctx->head.as<uint32_t>(), ctx->ovf_count.as<uint32_t>(), N,
                                    (uint32_t)C, st));
            MZH_HIP(hipEventRecord(ctx->ev[10], st));
            MZH_HIP(launch_bp_count(ctx->trav.as<TravNode>(), ctx->tree_size.as<uint32_t>(),
                                    ctx->srt_aabbs.as<float>(), ctx->srt_mcodes.as<uint64_t>(),
                                    ctx->srt_idx.as<uint32_t>(), ctx->n_eff.as<uint32_t>(), ctx->otypes.as<int32_t>(),
                                    ctx->cnt.as<uint32_t>(), ctx->slots.as<uint32_t>(), ctx->ovf.as<uint2>(),
                                    ctx->head.as<uint32_t>(), ctx->ovf_count.as<uint32_t>(), ctx->ovf_cap, N,
                                    ctx->node_cap, (uint32_t)C, st));
            MZH_HIP(hipEventRecord(ctx->ev[11], st));
            MZH_HIP(launch_scan_counts(ctx->cnt.as<uint32_t>(), ctx->off.as<uint64_t>(),
                                       ctx->step_total.as<uint64_t>(), N, (uint32_t)C, st));
            MZH_HIP(launch_scan_steps(ctx->step_total.as<uint64_t>(), ctx->step_off.as<uint64_t>(),
                                      ctx->grand.as<uint64_t>(), (uint32_t)C, st));
            MZH_HIP(hipMemcpyAsync(&n_ovf, ctx->ovf_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
            MZH_HIP(hipStreamSynchronize(st));
            if (n_ovf <= ctx->ovf_cap) break;
            ctx->ovf_cap = n_ovf + n_ovf / 4 + 4096;
            MZH_HIP(ctx->ovf.ensure(sizeof(uint2) * ctx->ovf_cap));
        }
        if (n_ovf > ctx->ovf_cap) throw StatusError{MZH_ERUNTIME, "broad-phase overflow list capacity exceeded"};
        uint64_t total = 0;
        MZH_HIP(hipMemcpy(&total, ctx->grand.p, sizeof(uint64_t), hipMemcpyDeviceToHost));
        ctx->n_pairs = total;
        MZH_HIP(ctx->pairs.ensure(sizeof(mzh_aabb_collision) * (total ? total : 1)));
        MZH_HIP(launch_bp_place(ctx->cnt.as<uint32_t>(), ctx->slots.as<uint32_t>(), ctx->ovf.as<uint2>(),
                                ctx->head.as<uint32_t>(), ctx->ovf_cap, ctx->off.as<uint64_t>(),
                                ctx->step_off.as<uint64_t>(), ctx->pairs.as<mzh_aabb_collision>(), N, (uint32_t)C, st));
        MZH_HIP(hipEventRecord(ctx->ev[4], st));
        // stage 5: narrow phase, in passes over at most pass_pairs broad-phase
        // pairs (64 Mi, or MIZUBA_AMD_NARROW_PASS_PAIRS) so that the step-pair
        // and candidate queues stay bounded; a pass whose queues or output
        // overflow is re-run with larger buffers
        NarrowArgs na;
        na.pj = pj;
        na.sg = sg;
        na.kbke = ctx->kbke.as<const uint2>();
        na.pairs = ctx->pairs.as<const mzh_aabb_collision>();
        na.step_off = ctx->step_off.as<const uint64_t>();
        na.conj_thresh = conj_thresh;
        const uint32_t op1 = ctx->order + 1;
        const double pi = 3.141592653589793238462643383279502884;
        for (int m = 0; m <= kMaxOrder; ++m) na.cheby[m] = 0;
        for (uint32_t m = 0; m < op1; ++m) na.cheby[m] = (std::cos((2. * m + 1.) / (2. * op1) * pi) + 1.) / 2.;
        na.np_report = ctx->np_report.as<unsigned long long>();
        na.conj_count = ctx->conj_count.as<unsigned long long>();
        na.err = derr;
        // bounds the queues to a few GB
        const uint64_t pass_pairs = env_u64("MIZUBA_AMD_NARROW_PASS_PAIRS", 64ull << 20);
        const uint64_t chunk = std::min<uint64_t>(total, pass_pairs);
        if (ctx->conj_cap == 0) ctx->conj_cap = env_u64("MIZUBA_AMD_INIT_CONJ_CAP", 1u << 16);
        // queued root-finding candidates: about a fifth of the pairs on LEO catalogs
        const size_t cand_bytes = narrow_cand_bytes(ctx->order);
        ctx->cand_cap = std::max<uint64_t>(ctx->cand_cap, env_u64("MIZUBA_AMD_INIT_CAND_CAP", chunk / 2 + 65536));
        // overlapping trajectory-step pairs: about 1.5 per pair for T/8 steps, 2.2 for the T/20 steps
        // of SGP4 polyjectories, with 3-minute conjunction steps (a larger need re-runs the pass)
        ctx->steps_cap = std::max<uint64_t>(ctx->steps_cap, env_u64("MIZUBA_AMD_INIT_STEPS_CAP", 3 * chunk + 65536));
        // candidates whose root isolation outgrows the private work list (rare)
        ctx->slow_cap = std::max<uint64_t>(ctx->slow_cap, env_u64("MIZUBA_AMD_INIT_SLOW_CAP", 4096));
        na.wl_fast_cap = (int)std::min<uint64_t>(env_u64("MIZUBA_AMD_WL_FAST_CAP", kNarrowWlFastCap), kNarrowWlFastCap);
        MZH_HIP(ctx->cand_count.ensure(sizeof(unsigned long long)));
        MZH_HIP(ctx->steps_count.ensure(sizeof(unsigned long long)));
        MZH_HIP(ctx->slow_count.ensure(sizeof(unsigned long long)));
        MZH_HIP(ctx->wl_slow.ensure(narrow_wl_slow_bytes(ctx->order)));
        na.cand_count = ctx->cand_count.as<unsigned long long>();
        na.steps_count = ctx->steps_count.as<unsigned long long>();
        na.slow_count = ctx->slow_count.as<unsigned long long>();
        na.wl_slow = ctx->wl_slow.p;
        MZH_HIP(ctx->conj_raw.grow(sizeof(mzh_conj) * ctx->conj_cap, 0, st));
        MZH_HIP(ctx->conj_step.grow(sizeof(uint32_t) * ctx->conj_cap, 0, st));
        MZH_HIP(hipMemsetAsync(ctx->conj_count.p, 0, sizeof(unsigned long long), st));
        MZH_HIP(hipMemsetAsync(ctx->np_report.p, 0, sizeof(unsigned long long) * MZH_NP_COUNT, st));
        // state after the last completed pass: conjunction count, counters
        unsigned long long done_conj = 0, done_rep[MZH_NP_COUNT] = {0, 0, 0, 0, 0, 0, 0};
        MZH_HIP(hipEventRecord(ctx->ev[8], st));
        for (uint64_t pb = 0; pb < total; pb += chunk) {
            na.pair_begin = pb;
            na.pair_end = std::min<uint64_t>(total, pb + chunk);
            bool ok = false;
            for (int attempt = 0; attempt < 8 && !ok; ++attempt) {
                MZH_HIP(ctx->conj_raw.grow(sizeof(mzh_conj) * ctx->conj_cap, sizeof(mzh_conj) * done_conj, st));
                MZH_HIP(ctx->conj_step.grow(sizeof(uint32_t) * ctx->conj_cap, sizeof(uint32_t) * done_conj, st));
                MZH_HIP(ctx->cand.ensure(cand_bytes * ctx->cand_cap));
                MZH_HIP(ctx->steps.ensure(sizeof(StepPair) * ctx->steps_cap));
                MZH_HIP(ctx->slow.ensure(sizeof(uint64_t) * ctx->slow_cap));
                if (attempt > 0) {
                    // undo the overflowed pass
                    MZH_HIP(hipMemcpyAsync(ctx->conj_count.p, &done_conj, sizeof(done_conj), hipMemcpyHostToDevice,
                                           st));
                    MZH_HIP(hipMemcpyAsync(ctx->np_report.p, done_rep, sizeof(done_rep), hipMemcpyHostToDevice, st));
                }
                MZH_HIP(hipMemsetAsync(ctx->cand_count.p, 0, sizeof(unsigned long long), st));
                MZH_HIP(hipMemsetAsync(ctx->steps_count.p, 0, sizeof(unsigned long long), st));
                MZH_HIP(hipMemsetAsync(ctx->slow_count.p, 0, sizeof(unsigned long long), st));
                na.slow = ctx->slow.as<uint64_t>();
                na.slow_cap = ctx->slow_cap;
                na.steps = ctx->steps.p;
                na.steps_cap = ctx->steps_cap;
                na.conj = ctx->conj_raw.as<mzh_conj>();
                na.conj_step = ctx->conj_step.as<uint32_t>();
                na.conj_cap = ctx->conj_cap;
                na.cand = ctx->cand.p;
                na.cand_cap = ctx->cand_cap;
                MZH_HIP(launch_narrow(na, st));
                unsigned long long cnts[4] = {0, 0, 0, 0}, rep[MZH_NP_COUNT];
                MZH_HIP(hipMemcpyAsync(&cnts[0], ctx->conj_count.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                       st));
                MZH_HIP(hipMemcpyAsync(&cnts[1], ctx->cand_count.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                       st));
                MZH_HIP(hipMemcpyAsync(&cnts[2], ctx->steps_count.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                       st));
                MZH_HIP(hipMemcpyAsync(&cnts[3], ctx->slow_count.p, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                       st));
                MZH_HIP(hipMemcpyAsync(rep, ctx->np_report.p, sizeof(rep), hipMemcpyDeviceToHost, st));
                MZH_HIP(hipStreamSynchronize(st));
                const bool conj_ok = cnts[0] <= ctx->conj_cap, cand_ok = cnts[1] <= ctx->cand_cap,
                           steps_ok = cnts[2] <= ctx->steps_cap, slow_ok = cnts[3] <= ctx->slow_cap;
                ok = conj_ok && cand_ok && steps_ok && slow_ok;
                if (ok) {
                    done_conj = cnts[0];
                    std::memcpy(done_rep, rep, sizeof(rep));
                    break;
                }
                if (!conj_ok) ctx->conj_cap = cnts[0] + cnts[0] / 4 + 1024;
                if (!cand_ok) ctx->cand_cap = cnts[1] + cnts[1] / 4 + 1024;
                if (!steps_ok) ctx->steps_cap = cnts[2] + cnts[2] / 4 + 1024;
                if (!slow_ok) ctx->slow_cap = cnts[3] + cnts[3] / 4 + 1024;
            }
            if (!ok) throw StatusError{MZH_ERUNTIME, "narrow-phase buffers overflowed after resizing"};
        }
        MZH_HIP(hipEventRecord(ctx->ev[9], st));
        ctx->n_conj = done_conj;
        const uint64_t M = ctx->n_conj;
        if (M >= 0xffffffffull) throw StatusError{MZH_ERUNTIME, "Too many conjunctions for one screening"};
        MZH_HIP(ctx->conj_sorted.ensure(sizeof(mzh_conj) * (M ? M : 1)));
        MZH_HIP(ctx->conj_keys.ensure(2u * sizeof(ConjKey) * (M ? M : 1)));
        MZH_HIP(launch_conj_sort(ctx->conj_raw.as<mzh_conj>(), ctx->conj_step.as<uint32_t>(), M,
                                 ctx->conj_keys.as<ConjKey>(), ctx->conj_keys.as<ConjKey>() + M,
                                 ctx->conj_sorted.as<mzh_conj>(), st));
    } else {
        MZH_HIP(hipMemsetAsync(ctx->step_off.p, 0, sizeof(uint64_t) * (C + 1), st));
        MZH_HIP(hipEventRecord(ctx->ev[4], st));
        MZH_HIP(hipEventRecord(ctx->ev[8], st));
        MZH_HIP(hipEventRecord(ctx->ev[9], st));
    }
    MZH_HIP(hipEventRecord(ctx->ev[5], st));
    MZH_HIP(hipStreamSynchronize(st));

    DevError herr;
    MZH_HIP(hipMemcpy(&herr, ctx->err.p, sizeof(DevError), hipMemcpyDeviceToHost));
    if (herr.code != kErrNone) throw StatusError{MZH_EINVAL, dev_error_message(herr)};

    float ms[5];
    for (int k = 0; k < 5; ++k) MZH_HIP(hipEventElapsedTime(&ms[k], ctx->ev[k], ctx->ev[k + 1]));
    for (int k = 0; k < 5; ++k) ctx->stage_s[k] = ms[k] * 1e-3;
    float tot = 0;
    MZH_HIP(hipEventElapsedTime(&tot, ctx->ev[0], ctx->ev[5]));
    ctx->stage_s[5] = tot * 1e-3;
    float ka = 0, kn = 0;
    MZH_HIP(hipEventElapsedTime(&ka, ctx->ev[6], ctx->ev[7]));
    MZH_HIP(hipEventElapsedTime(&kn, ctx->ev[8], ctx->ev[9]));
    ctx->stage_s[6] = ka * 1e-3;
    ctx->stage_s[7] = kn * 1e-3;
    float kb = 0;
    if (!skip_cd) MZH_HIP(hipEventElapsedTime(&kb, ctx->ev[10], ctx->ev[11]));
    ctx->stage_s[8] = kb * 1e-3;
    unsigned long long rep[MZH_NP_COUNT];
    MZH_HIP(hipMemcpy(rep, ctx->np_report.p, sizeof(rep), hipMemcpyDeviceToHost));
    for (int k = 0; k < MZH_NP_COUNT; ++k) ctx->h_np[k] = rep[k];

    ctx->h_tree_size.resize(C);
    MZH_HIP(hipMemcpy(ctx->h_tree_size.data(), ctx->tree_size.p, sizeof(uint32_t) * C, hipMemcpyDeviceToHost));
    for (uint64_t k = 0; k < C; ++k) {
        if (ctx->h_tree_size[k] > ctx->node_cap) throw StatusError{MZH_ERUNTIME, "BVH node capacity exceeded"};
    }
    ctx->h_step_off.resize(C + 1);
    MZH_HIP(hipMemcpy(ctx->h_step_off.data(), ctx->step_off.p, sizeof(uint64_t) * (C + 1), hipMemcpyDeviceToHost));
    ctx->screened = true;
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

// The trees are kept in the broad phase's layout; the reference's 48-byte
// bvh_node records are produced on the first request after a screen.
void export_nodes(mzh_ctx *ctx)
{
    if (ctx->nodes_ready) return;
    MZH_HIP(ctx->nodes.ensure(sizeof(mzh_bvh_node) * ctx->node_cap * ctx->C));
    MZH_HIP(launch_bvh_export(ctx->trav.as<TravNode>(), ctx->tree_size.as<uint32_t>(), ctx->nodes.as<mzh_bvh_node>(),
                              ctx->node_cap, (uint32_t)ctx->C, ctx->stream));
    MZH_HIP(hipStreamSynchronize(ctx->stream));
    ctx->nodes_ready = true;
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
    for (int k = 0; k < 12 && e == hipSuccess; ++k) e = hipEventCreate(&ctx->ev[k]);
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
    for (auto &e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
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
        MZH_HIP(ctx->pjb->traj.ensure(sizeof(double) * (traj_len ? traj_len : 1)));
        MZH_HIP(ctx->pjb->time.ensure(sizeof(double) * (time_len ? time_len : 1)));
        MZH_HIP(ctx->pjb->cf4.ensure(sizeof(double) * 4u * (order + 1) * (steps_total ? steps_total : 1)));
        MZH_HIP(ctx->pjb->step0.ensure(sizeof(uint64_t) * n_objs));
        MZH_HIP(ctx->pjb->nsteps.ensure(sizeof(uint64_t) * n_objs));
        MZH_HIP(ctx->pjb->time_off.ensure(sizeof(uint64_t) * n_objs));
        if (traj_len) MZH_HIP(hipMemcpyAsync(ctx->pjb->traj.p, traj, sizeof(double) * traj_len, hipMemcpyHostToDevice, st));
        if (time_len) MZH_HIP(hipMemcpyAsync(ctx->pjb->time.p, time, sizeof(double) * time_len, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->pjb->step0.p, step0.data(), sizeof(uint64_t) * n_objs, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->pjb->nsteps.p, nsteps, sizeof(uint64_t) * n_objs, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->pjb->time_off.p, time_off, sizeof(uint64_t) * n_objs, hipMemcpyHostToDevice, st));
        MZH_HIP(launch_transpose_cf4(ctx->pjb->traj.as<double>(), ctx->pjb->cf4.as<double>(), steps_total, order, st));
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
    MZH_HIP(ctx->pjb->traj.ensure(sizeof(double) * steps_total * op1 * 7u));
    MZH_HIP(ctx->pjb->time.ensure(sizeof(double) * time_len));
    MZH_HIP(ctx->pjb->cf4.ensure(sizeof(double) * 4u * op1 * steps_total));
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
    MZH_HIP(launch_transpose_cf4(ctx->pjb->traj.as<double>(), ctx->pjb->cf4.as<double>(), steps_total, order, st));
    MZH_HIP(hipEventRecord(E[7], st));
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
        const ScreenPlan pl = plan_screen((uint32_t)n_objs, maxT, conj_thresh, conj_det_interval, otypes, mode);
        if (n_cd_steps) *n_cd_steps = pl.C;
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

int mzh_screen_batch(mzh_ctx *const *ctxs, uint64_t n, const double *conj_thresh, const double *conj_det_interval,
                     const int32_t *const *otypes, int mode, int *status)
{
    if (n == 0) return MZH_OK;
    if (!ctxs || !conj_thresh || !conj_det_interval || !status) return MZH_EINVAL;
    {
        std::vector<const mzh_ctx *> seen(ctxs, ctxs + n);
        std::sort(seen.begin(), seen.end());
        if (!seen.front() || std::adjacent_find(seen.begin(), seen.end()) != seen.end()) return MZH_EINVAL;
    }
    // One host thread per problem in flight: each drives its own context's
    // stream, including the two synchronisations that size the pair and
    // conjunction buffers, so the problems overlap on the device.
    std::atomic<uint64_t> next{0};
    auto worker = [&] {
        for (uint64_t k = next.fetch_add(1); k < n; k = next.fetch_add(1))
            status[k] = mzh_screen(ctxs[k], conj_thresh[k], conj_det_interval[k], otypes ? otypes[k] : nullptr, mode);
    };
    const uint64_t nthreads = std::min<uint64_t>(n, 8);
    std::vector<std::thread> pool;
    try {
        for (uint64_t t = 1; t < nthreads; ++t) pool.emplace_back(worker);
    } catch (...) {
        // fewer threads: the remaining problems are taken by those that started
    }
    worker();
    for (auto &th : pool) th.join();
    for (uint64_t k = 0; k < n; ++k)
        if (status[k] != MZH_OK) return status[k];
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
        if (what & MZH_RELEASE_SCRATCH) {
            for (DevBuf *b : {&ctx->kbke, &ctx->keys_tmp, &ctx->vals_tmp, &ctx->n_eff, &ctx->sort_todo, &ctx->gbox,
                              &ctx->bvh_hdr, &ctx->bvh_scr, &ctx->otypes, &ctx->cnt, &ctx->head, &ctx->off,
                              &ctx->step_total, &ctx->grand, &ctx->slots, &ctx->ovf, &ctx->ovf_count, &ctx->conj_raw,
                              &ctx->conj_step, &ctx->conj_keys, &ctx->conj_count, &ctx->cand, &ctx->cand_count,
                              &ctx->steps, &ctx->steps_count, &ctx->slow, &ctx->slow_count, &ctx->wl_slow,
                              &ctx->ev_tm, &ctx->ev_sel, &ctx->ev_out, &ctx->ev_bad})
                b->release();
        }
        if (what & MZH_RELEASE_RESULTS) {
            for (DevBuf *b : {&ctx->aabbs, &ctx->srt_aabbs, &ctx->mcodes, &ctx->srt_mcodes, &ctx->srt_idx, &ctx->nodes,
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

} // extern "C"
