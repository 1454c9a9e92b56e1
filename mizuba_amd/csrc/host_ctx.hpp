// Host-side state of a context and the helpers shared by the C ABI
// (capi.hip) and the batch orchestrator (batch.hip). Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/mizuba_amd.h"
#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

namespace mzh::host
{

// Shortest decimal text that round-trips a double (the {} formatting of the
// reference's fmt messages).
inline std::string fmt_double(double v)
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
inline uint64_t env_u64(const char *name, uint64_t dflt)
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
    mzh::host::DevBuf &operator=(const DevBuf &) = delete;
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
    mzh::host::DevBuf traj, time, cf4, cf4f, sbnd, scal, step0, nsteps, time_off;
    uint32_t h_scal[2] = {0u, 0u};  // the polyjectory's lam and A (float bits), read back at upload
};

} // namespace mzh::host

namespace mzh::host
{
class ScreenBatch;
struct BatchDeleter {
    void operator()(ScreenBatch *b) const;
};
} // namespace mzh::host

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
    std::shared_ptr<mzh::host::PjBufs> pjb;
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
    mzh::host::DevBuf cd_begin, cd_end;
    mzh::host::DevBuf aabbs, kbke, srt_aabbs, mcodes, srt_mcodes, srt_idx, keys_tmp, vals_tmp, n_eff, sort_todo, gbox;
    mzh::host::DevBuf nprec;  // mixed-precision pre-screen models, [C][N][2] float4
    uint64_t node_cap = 0;
    mzh::host::DevBuf nodes, trav, wtree, bvh_hdr, bvh_scr, tree_size, bkt_end;
    bool nodes_ready = false;  // nodes holds the reference-layout export of trav
    mzh::host::DevBuf otypes, cnt, head, off, step_total, step_off, slots, ovf, ovf_count;
    uint32_t ovf_cap = 0;
    mzh::host::DevBuf pairs, pair_c;
    uint64_t n_pairs = 0, pair_cap = 0;
    mzh::host::DevBuf conj_sorted;
    uint64_t n_conj = 0, conj_out_cap = 0;
    std::vector<int32_t> h_otypes;  // the object types of the screening
    // bumped by every operation that may move the context's device buffers
    // (uploads, releases, screenings outside a batch's graph replay)
    uint64_t generation = 0;
    const void *last_batch = nullptr;  // the batch that ran the last screening
    // the batch that screens this context alone (mzh_screen), with the
    // narrow phase's queues, kept for the next screening
    std::unique_ptr<mzh::host::ScreenBatch, mzh::host::BatchDeleter> solo;
    // polyjectory evaluation: staged times, selector and results, first bad element
    mzh::host::DevBuf ev_tm, ev_sel, ev_out, ev_bad;
    hipEvent_t ev_eval[2] = {};
    double eval_s = 0;  // device time of the last evaluation kernel
    // SGP4 generation: device times of its kernels (segment set-up, count,
    // scan, write, transpose) in the last mzh_make_sgp4_polyjectory
    hipEvent_t ev_sgp4[8] = {};
    double sgp4_s[5] = {0, 0, 0, 0, 0};
    double stage_s[MZH_STAGE_COUNT] = {};
    uint64_t h_np[MZH_NP_COUNT] = {0, 0, 0, 0, 0, 0, 0};

    // host caches of per-step sizes (filled after a screen)
    std::vector<uint32_t> h_tree_size;
    std::vector<uint64_t> h_step_off;  // [C + 1]
};


namespace mzh::host
{

#define MZH_HIP(call)                                                                                                  \
    do {                                                                                                               \
        const hipError_t e_ = (call);                                                                                  \
        if (e_ != hipSuccess)                                                                                          \
            throw mzh::host::StatusError{e_ == hipErrorOutOfMemory ? MZH_ENOMEM : MZH_ERUNTIME,                        \
                                         std::string("HIP error in ") + #call + ": " + hipGetErrorString(e_)};         \
    } while (0)

// The reference's message for a device error word.
std::string dev_error_message(const DevError &e);
// The polyjectory of a context as the kernels see it.
PjView pj_view(const mzh_ctx *ctx);

// The arguments of one screening after validation.
struct ScreenPlan {
    uint64_t C = 0;
    bool skip_cd = true;
    std::vector<int32_t> ot;
    std::vector<double> cd_begin, cd_end;
};
// conjunctions::conjunctions validation (conjunctions.cpp:279-335) with the
// reference's messages, the step ranges of get_cd_begin_end
// (conjunctions.cpp:373-392), and this build's size limit.
ScreenPlan plan_screen(uint32_t N, double maxT, double conj_thresh, double conj_det_interval,
                       const int32_t *otypes_in, int mode, bool one_batch = true, uint64_t k0 = 0,
                       uint64_t k1 = ~0ull);
// The size limit of one batch (MZH_ESIZE): N * C below 2^32.
void check_batch_size(uint64_t N, uint64_t C);

// Store a validated plan in its context.
void set_plan(mzh_ctx *ctx, ScreenPlan &&pl, double conj_thresh, double conj_det_interval, int mode);

// A batch on `stream` (see batch.hip). batch_run screens the contexts (each
// with its plan set) together, one launch per stage for all of them:
// replaying its captured graph when `graph` (captured after the first run),
// choosing its launch shapes by timing first when `tune`. status[k]
// receives each problem's outcome, its message in ctxs[k]->last_error.
ScreenBatch *batch_new(hipStream_t stream);
void batch_run(ScreenBatch *b, const std::vector<mzh_ctx *> &ctxs, bool graph, bool tune, int *status);
// The reference's breadth-first trees (detect_conjunctions_bvh.cpp:267-699)
// of the context's last screening, into ctx->trav and ctx->tree_size, from
// its sorted codes and boxes; the screening itself walks a tree of its own
// numbering. Synchronous on ctx->stream.
void build_bfs_trees(mzh_ctx *ctx);
// Graph replays so far, device time of the last run (seconds), launch shape.
uint64_t batch_replays(const ScreenBatch *b);
double batch_seconds(const ScreenBatch *b);
LaunchShape batch_shape(const ScreenBatch *b);

} // namespace mzh::host
