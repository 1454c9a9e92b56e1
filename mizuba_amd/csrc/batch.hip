// The screening orchestrator: any number of independent problems (contexts,
// each with its own polyjectory, threshold, interval and object types)
// screened together, one launch per stage for all of them. A single
// screening is a batch of one.
//
// Replaces conjunctions::conjunctions + detect_conjunctions
// (mizuba/conjunctions.cpp:263-360, mizuba/detect_conjunctions.cpp:55-537):
// instead of a TBB loop over conjunction steps in chunks of 128 with a writer
// thread (detect_conjunctions.cpp:61-79, 313-424), every stage runs once over
// all steps of all problems, and the outputs stay in HBM.
//
// Work distribution:
// - the AABB, broad-phase and placement kernels map blocks to (problem,
//   block of the problem) through prefix tables;
// - the sort and the BVH build are persistent: a fixed grid of workgroups
//   takes (problem, step) items from a device counter, largest problems first;
// - the narrow phase runs over one global index of all problems' pairs, and
//   its queues (step pairs, candidates, slow-path candidates, conjunction
//   records) are shared by the batch.
//
// Two ways to run:
// - eager: the host reads the pair, queue and conjunction counts between the
//   phases (three synchronisations) and sizes the buffers exactly;
// - graph: the whole launch sequence, captured once into a HIP graph after an
//   eager run, is replayed with the capacities that run measured (plus a
//   quarter); overflows are flagged on the device, and a replay that flags
//   one is redone eagerly. One synchronisation per run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mizuba_amd.h"
#include "host_ctx.hpp"
#include "mzh_device.hpp"
#include "mzh_kernels.hpp"

namespace mzh::host
{

void set_plan(mzh_ctx *ctx, ScreenPlan &&pl, double conj_thresh, double conj_det_interval, int mode)
{
    ctx->screened = false;
    ctx->results_released = false;
    ctx->C = pl.C;
    ctx->skip_cd = pl.skip_cd;
    ctx->conj_thresh = conj_thresh;
    ctx->conj_det_interval = conj_det_interval;
    ctx->mode = mode;
    ctx->h_cd_begin = std::move(pl.cd_begin);
    ctx->h_cd_end = std::move(pl.cd_end);
    ctx->h_otypes = std::move(pl.ot);
    ctx->node_cap = 2ull * (ctx->n_objs ? ctx->n_objs : 1);
    ++ctx->generation;  // new step arrays and object types to describe
}

namespace
{

uint64_t grown(uint64_t need) { return need + need / 4 + 1024; }

// Launch shapes chosen by the tuner, per (order, size class) of a batch.
std::mutex g_tune_m;
std::map<std::tuple<uint32_t, uint32_t, uint32_t, int>, LaunchShape> g_tuned;

uint32_t log2_class(double x) { return x >= 1 ? (uint32_t)std::lround(std::log2(x)) : 0u; }

} // namespace

class ScreenBatch
{
public:
    explicit ScreenBatch(hipStream_t stream) : st(stream)
    {
        for (auto &e : ev) MZH_HIP(hipEventCreate(&e));
    }
    ~ScreenBatch()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
    ScreenBatch(const ScreenBatch &) = delete;
    ScreenBatch &operator=(const ScreenBatch &) = delete;

    // The problems to screen (plans already set), in the caller's order.
    void set_problems(const std::vector<mzh_ctx *> &ctxs)
    {
        if (ctxs != caller) {
            caller = ctxs;
            invalidate();
        }
    }
    void set_graph(bool g) { use_graph = g; }
    void set_shape(const LaunchShape &s)
    {
        ls = s;
        invalidate();
    }
    const LaunchShape &shape() const { return ls; }
    uint64_t replays() const { return n_replays; }
    double last_run_seconds() const { return run_s; }

    // Screen all problems; status[k] = the outcome of caller[k].
    void run(int *status);
    // Choose the launch shapes by timing the candidates on this batch.
    void tune();

private:
    hipStream_t st;
    std::vector<mzh_ctx *> caller;  // the caller's order
    std::vector<mzh_ctx *> probs;   // batch order: by polynomial order
    std::vector<uint32_t> back;     // probs[k] == caller[back[k]]
    LaunchShape ls;
    bool use_graph = false;
    bool prepared = false;
    std::vector<uint64_t> gens;     // contexts' generations when prepared

    // device tables and batch-wide state
    DevBuf d_P, d_tab, d_items, d_bc, d_pair_pre, d_ctr, d_counts, d_flags, d_np, d_err;
    DevBuf steps, cand, slow, wl_slow, conj_raw, conj_step, conj_keys;
    uint64_t steps_cap = 0, cand_cap = 0, slow_cap = 0, conj_cap = 0;
    std::vector<Prob> hP;
    BatchTables tabs{};
    uint32_t n_steps_all = 0, n_nc_blocks = 0, n_bp_blocks = 0, n_gather_blocks = 0, n_lbvh_blocks = 0;
    struct Group {
        uint32_t order, p_first, p_end, blk_begin, n_blk;
        bool any_fp64, any_mixed;  // the modes of its problems (one AABB launch each)
    };
    std::vector<Group> groups;
    bool any_bp = false;
    uint32_t max_order = 0;
    size_t max_cand_bytes = 0;

    // graph
    hipGraphExec_t exec = nullptr;
    bool exec_ok = false;
    uint64_t n_replays = 0;
    std::vector<uint64_t> group_pair_ub;  // pair bound per order group in the graph
    hipEvent_t ev[12] = {};
    double run_s = 0;

    void invalidate()
    {
        prepared = false;
        exec_ok = false;
    }
    bool gens_changed() const
    {
        if (gens.size() != probs.size()) return true;
        for (size_t k = 0; k < probs.size(); ++k)
            if (probs[k]->generation != gens[k]) return true;
        return false;
    }
    uint64_t pass_pairs() const { return env_u64("MIZUBA_AMD_NARROW_PASS_PAIRS", 64ull << 20); }
    void prepare();
    void upload_probs() { MZH_HIP(hipMemcpyAsync(d_P.p, hP.data(), sizeof(Prob) * hP.size(), hipMemcpyHostToDevice, st)); }
    void ensure_queues();
    void enqueue_stages(bool timed);
    void enqueue_place() { MZH_HIP(launch_bp_place(d_P.as<Prob>(), np(), tabs, n_nc_blocks, st)); }
    NarrowArgs narrow_args(const Group &g) const;
    void enqueue_narrow(const std::vector<uint64_t> &group_pairs);
    void enqueue_conj_sort(uint64_t n, bool dev_count);
    void run_eager(bool timed);
    void capture();
    bool replay();
    void finalize(int *status_batch_order, bool replayed);
    uint32_t np() const { return (uint32_t)probs.size(); }
    void read_back(std::vector<uint64_t> &counts, std::vector<uint32_t> &flags);
};

// The mixed-precision pre-screen's bound on the FP64 narrow phase's dist^2
// interpolation error for any step pair of a problem: with n1 = order + 1
// Chebyshev nodes on an interval of h <= conj_det_interval, at most
// 2 (h/4)^n1 / n1! max |F^(n1)|, F^(n1) = sum_{a=1..order} C(n1, a) d^(a) .
// d^(n1-a), |d^(a)| <= 2 A lam^a (the polyjectory's A and lam): at most
// 8 (2^n1 - 2) / n1! (h lam / 4)^n1 A^2, rounded up.
double mixed_interp_bound(const mzh_ctx *ctx)
{
    float lam = 0.0f, amp = 0.0f;
    std::memcpy(&lam, &ctx->pjb->h_scal[0], sizeof(float));
    std::memcpy(&amp, &ctx->pjb->h_scal[1], sizeof(float));
    const uint32_t n1 = ctx->order + 1u;
    const double hl = ctx->conj_det_interval * (1.0 + 1e-12) * 0.25 * (double)lam;
    double e = 8.0 * (std::ldexp(1.0, (int)n1) - 2.0);
    for (uint32_t k = 1; k <= n1; ++k) e *= hl / (double)k;
    e *= (double)amp * (double)amp * (1.0 + 1e-9);
    return std::isfinite(e) ? e : HUGE_VAL;
}

void build_bfs_trees(mzh_ctx *ctx)
{
    const uint32_t N = ctx->n_objs;
    const uint64_t C = ctx->C;
    if (C == 0) return;
    hipStream_t st = ctx->stream;
    MZH_HIP(ctx->trav.ensure(sizeof(TravNode) * ctx->node_cap * C));
    MZH_HIP(ctx->bvh_hdr.ensure(sizeof(uint4) * ctx->node_cap * C));
    const bool in_lds = bvh_scratch_bytes_host(N) <= kBvhLdsMaxHost;
    if (!in_lds) MZH_HIP(ctx->bvh_scr.ensure(bvh_scratch_stride(N) * C));
    Prob q{};
    q.C = (uint32_t)C;
    q.N = N;
    q.node_cap = ctx->node_cap;
    q.bvh_scr_stride = bvh_scratch_stride(N);
    q.srt_aabbs = gl(ctx->srt_aabbs.as<float>());
    q.srt_mcodes = gl(ctx->srt_mcodes.as<uint64_t>());
    q.n_eff = gl(ctx->n_eff.as<uint32_t>());
    q.trav = gl(ctx->trav.as<TravNode>());
    q.bvh_hdr = gl(ctx->bvh_hdr.as<uint4>());
    q.bvh_scr = gl(ctx->bvh_scr.as<uint8_t>());
    q.tree_size = gl(ctx->tree_size.as<uint32_t>());
    std::vector<uint2> items(C);
    for (uint64_t c = 0; c < C; ++c) items[c] = make_uint2(0u, (uint32_t)c);
    DevBuf dP, dI, dC;
    MZH_HIP(dP.ensure(sizeof(Prob)));
    MZH_HIP(dI.ensure(sizeof(uint2) * C));
    MZH_HIP(dC.ensure(sizeof(BatchCounters)));
    MZH_HIP(hipMemcpyAsync(dP.p, &q, sizeof(Prob), hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemcpyAsync(dI.p, items.data(), sizeof(uint2) * C, hipMemcpyHostToDevice, st));
    MZH_HIP(hipMemsetAsync(dC.p, 0, sizeof(BatchCounters), st));
    BatchTables t{};
    t.bvh_items = in_lds ? dI.as<uint2>() : nullptr;
    t.bvh_hbm_items = in_lds ? nullptr : dI.as<uint2>();
    t.n_bvh_items = in_lds ? (uint32_t)C : 0u;
    t.n_bvh_hbm_items = in_lds ? 0u : (uint32_t)C;
    t.bvh_lds_bytes = in_lds ? (uint32_t)bvh_scratch_bytes_host(N) : 0u;
    MZH_HIP(launch_bvh(dP.as<Prob>(), 1u, t, dC.as<BatchCounters>(), LaunchShape{}, st));
    // the temporaries are freed on return
    MZH_HIP(hipStreamSynchronize(st));
}

void ScreenBatch::prepare()
{
    // only the compiled variants (the block tables below follow bp_block)
    if (ls.bp_block != 128 && ls.bp_block != 512) ls.bp_block = 256;
    if (ls.bucket_bits != 10 && ls.bucket_bits != 11) ls.bucket_bits = 12;
    ls.sort_wgs = std::max(1u, ls.sort_wgs);
    ls.bvh_wgs = std::max(1u, ls.bvh_wgs);
    ls.roots_wgs = std::max(1u, ls.roots_wgs);
    // batch order: by polynomial order (the AABB and narrow-phase kernels are
    // compiled per order), stable in the caller's order
    back.resize(caller.size());
    std::iota(back.begin(), back.end(), 0u);
    std::stable_sort(back.begin(), back.end(), [&](uint32_t a, uint32_t b) { return caller[a]->order < caller[b]->order; });
    probs.resize(caller.size());
    for (size_t k = 0; k < back.size(); ++k) probs[k] = caller[back[k]];
    const uint32_t n = np();
    hP.assign(n, Prob{});
    MZH_HIP(d_counts.ensure(sizeof(uint64_t) * 4u * n));
    MZH_HIP(d_flags.ensure(sizeof(uint32_t) * n));
    MZH_HIP(d_np.ensure(sizeof(unsigned long long) * kReportWords * n));
    MZH_HIP(d_err.ensure(sizeof(DevError) * n));
    MZH_HIP(d_bc.ensure(sizeof(BatchCounters)));
    MZH_HIP(d_pair_pre.ensure(sizeof(uint64_t) * (n + 1)));
    MZH_HIP(d_ctr.ensure(sizeof(unsigned long long) * 8u));
    std::vector<uint32_t> step_pre(n + 1), aabb_pre(n + 1), nc_pre(n + 1), bp_pre(n + 1), gather_pre(n + 1),
        lbvh_pre(n + 1);
    max_order = 0;
    max_cand_bytes = 0;
    any_bp = false;
    groups.clear();
    for (uint32_t k = 0; k < n; ++k) {
        mzh_ctx *ctx = probs[k];
        const uint32_t N = ctx->n_objs;
        const uint64_t C = ctx->C;
        const size_t NC = (size_t)N * C;
        const size_t boxes_bytes = sizeof(float) * 8u * (size_t)(N + 1) * C;
        MZH_HIP(ctx->cd_begin.ensure(sizeof(double) * C));
        MZH_HIP(ctx->cd_end.ensure(sizeof(double) * C));
        MZH_HIP(ctx->aabbs.ensure(boxes_bytes));
        MZH_HIP(ctx->kbke.ensure(sizeof(uint2) * NC));
        if (ctx->mode == MZH_MODE_MIXED) MZH_HIP(ctx->nprec.ensure(sizeof(float4) * 2u * NC));
        MZH_HIP(ctx->srt_aabbs.ensure(boxes_bytes));
        MZH_HIP(ctx->mcodes.ensure(sizeof(uint64_t) * NC));
        MZH_HIP(ctx->srt_mcodes.ensure(sizeof(uint64_t) * NC));
        MZH_HIP(ctx->keys_tmp.ensure(sizeof(uint64_t) * NC));
        MZH_HIP(ctx->srt_idx.ensure(sizeof(uint32_t) * NC));
        MZH_HIP(ctx->vals_tmp.ensure(sizeof(uint32_t) * NC));
        MZH_HIP(ctx->n_eff.ensure(sizeof(uint32_t) * C));
        MZH_HIP(ctx->sort_todo.ensure(sizeof(uint32_t) * C));
        MZH_HIP(ctx->gbox.ensure(sizeof(uint32_t) * 8u * C));
        MZH_HIP(ctx->bkt_end.ensure(sizeof(uint32_t) * kBktEndStride * C));
        MZH_HIP(ctx->wtree.ensure(sizeof(WalkNode) * ctx->node_cap * C));
        MZH_HIP(ctx->bvh_hdr.ensure(sizeof(uint32_t) * lbvh_step_words(N) * C));
        const size_t scr_b = bvh_scratch_bytes_host(N);
        if (scr_b > kBvhLdsMaxHost) MZH_HIP(ctx->bvh_scr.ensure(bvh_scratch_stride(N) * C));
        MZH_HIP(ctx->tree_size.ensure(sizeof(uint32_t) * C));
        MZH_HIP(ctx->otypes.ensure(sizeof(int32_t) * (N ? N : 1)));
        MZH_HIP(ctx->cnt.ensure(sizeof(uint32_t) * NC));
        MZH_HIP(ctx->head.ensure(sizeof(uint32_t) * NC));
        MZH_HIP(ctx->slots.ensure(sizeof(uint32_t) * 4u * NC));
        if (ctx->ovf_cap == 0)
            ctx->ovf_cap = (uint32_t)std::min<uint64_t>(env_u64("MIZUBA_AMD_INIT_OVF_CAP", NC / 8 + 4096), 0x7fffffffu);
        MZH_HIP(ctx->ovf.ensure(sizeof(uint2) * ctx->ovf_cap));
        MZH_HIP(ctx->ovf_count.ensure(sizeof(uint32_t) * kPoolShards * kPoolStride));
        MZH_HIP(ctx->off.ensure(sizeof(uint64_t) * NC));
        MZH_HIP(ctx->step_total.ensure(sizeof(uint64_t) * C));
        MZH_HIP(ctx->step_off.ensure(sizeof(uint64_t) * (C + 1)));
        MZH_HIP(ctx->pairs.ensure(sizeof(mzh_aabb_collision) * (ctx->pair_cap ? ctx->pair_cap : 1)));
        MZH_HIP(ctx->pair_c.ensure(sizeof(uint32_t) * (ctx->pair_cap ? ctx->pair_cap : 1)));
        MZH_HIP(ctx->conj_sorted.ensure(sizeof(mzh_conj) * (ctx->conj_out_cap ? ctx->conj_out_cap : 1)));
        ctx->nodes_ready = false;
        MZH_HIP(hipMemcpyAsync(ctx->cd_begin.p, ctx->h_cd_begin.data(), sizeof(double) * C, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->cd_end.p, ctx->h_cd_end.data(), sizeof(double) * C, hipMemcpyHostToDevice, st));
        MZH_HIP(hipMemcpyAsync(ctx->otypes.p, ctx->h_otypes.data(), sizeof(int32_t) * N, hipMemcpyHostToDevice, st));
        Prob &q = hP[k];
        q.pj = pj_view(ctx);
        q.cd_begin = gl(ctx->cd_begin.as<const double>());
        q.cd_end = gl(ctx->cd_end.as<const double>());
        q.C = (uint32_t)C;
        q.N = N;
        q.step0 = step_pre[k];
        q.mode = ctx->mode;
        q.radius = ctx->conj_thresh / 2;
        q.thresh = ctx->conj_thresh;
        q.node_cap = ctx->node_cap;
        q.bvh_scr_stride = bvh_scratch_stride(N);
        q.lbvh_stride = lbvh_step_words(N);
        q.aabbs = gl(ctx->aabbs.as<float>());
        q.srt_aabbs = gl(ctx->srt_aabbs.as<float>());
        q.kbke = gl(ctx->kbke.as<uint2>());
        q.nprec = gl(ctx->nprec.as<float4>());
        q.np_e = mixed_interp_bound(ctx);
        q.gbox = gl(ctx->gbox.as<uint32_t>());
        q.mcodes = gl(ctx->mcodes.as<uint64_t>());
        q.srt_mcodes = gl(ctx->srt_mcodes.as<uint64_t>());
        q.keys_tmp = gl(ctx->keys_tmp.as<uint64_t>());
        q.srt_idx = gl(ctx->srt_idx.as<uint32_t>());
        q.vals_tmp = gl(ctx->vals_tmp.as<uint32_t>());
        q.bkt_end = gl(ctx->bkt_end.as<uint32_t>());
        q.n_eff = gl(ctx->n_eff.as<uint32_t>());
        q.sort_todo = gl(ctx->sort_todo.as<uint32_t>());
        q.wtree = gl(ctx->wtree.as<WalkNode>());
        q.bvh_hdr = gl(ctx->bvh_hdr.as<uint4>());
        q.bvh_scr = gl(ctx->bvh_scr.as<uint8_t>());
        q.tree_size = gl(ctx->tree_size.as<uint32_t>());
        q.otypes = gl(ctx->otypes.as<const int32_t>());
        q.cnt = gl(ctx->cnt.as<uint32_t>());
        q.head = gl(ctx->head.as<uint32_t>());
        q.slots = gl(ctx->slots.as<uint32_t>());
        q.ovf = gl(ctx->ovf.as<uint2>());
        q.ovf_count = gl(ctx->ovf_count.as<uint32_t>());
        q.ovf_cap = ctx->ovf_cap;
        q.off = gl(ctx->off.as<uint64_t>());
        q.step_total = gl(ctx->step_total.as<uint64_t>());
        q.step_off = gl(ctx->step_off.as<uint64_t>());
        q.pairs = gl(ctx->pairs.as<mzh_aabb_collision>());
        q.pair_c = gl(ctx->pair_c.as<uint32_t>());
        q.pair_cap = ctx->pair_cap;
        q.conj_out = gl(ctx->conj_sorted.as<mzh_conj>());
        q.conj_out_cap = ctx->conj_out_cap;
        q.np_report = gl(d_np.as<unsigned long long>() + (size_t)k * kReportWords);
        q.err = gl(d_err.as<DevError>() + k);
        q.flags = gl(d_flags.as<uint32_t>() + k);
        q.counts = gl(d_counts.as<uint64_t>() + 4u * k);
        step_pre[k + 1] = step_pre[k] + (uint32_t)C;
        aabb_pre[k + 1] = aabb_pre[k] + ((N + 255u) / 256u) * (uint32_t)((C + kAabbStepGroup - 1) / kAabbStepGroup);
        nc_pre[k + 1] = nc_pre[k] + (uint32_t)((NC + 255u) / 256u);
        bp_pre[k + 1] = bp_pre[k] + (ctx->skip_cd ? 0u : ((N + ls.bp_block - 1u) / ls.bp_block) * (uint32_t)C);
        gather_pre[k + 1] = gather_pre[k] + ((N + 1u + 255u) / 256u) * (uint32_t)C;
        lbvh_pre[k + 1] = lbvh_pre[k] + ((N + kLbvhBlock - 1u) / kLbvhBlock) * (uint32_t)C;
        any_bp = any_bp || !ctx->skip_cd;
        max_order = std::max(max_order, ctx->order);
        max_cand_bytes = std::max(max_cand_bytes, narrow_cand_bytes(ctx->order));
        if (groups.empty() || groups.back().order != ctx->order)
            groups.push_back(Group{ctx->order, k, k + 1, aabb_pre[k], 0, false, false});
        groups.back().p_end = k + 1;
        (ctx->mode == MZH_MODE_MIXED ? groups.back().any_mixed : groups.back().any_fp64) = true;
        groups.back().n_blk = aabb_pre[k + 1] - groups.back().blk_begin;
    }
    n_steps_all = step_pre[n];
    n_nc_blocks = nc_pre[n];
    n_bp_blocks = bp_pre[n];
    n_gather_blocks = gather_pre[n];
    n_lbvh_blocks = lbvh_pre[n];
    // work items of the persistent sort and build, largest problems first
    std::vector<uint32_t> by_size(n);
    std::iota(by_size.begin(), by_size.end(), 0u);
    std::stable_sort(by_size.begin(), by_size.end(), [&](uint32_t a, uint32_t b) { return hP[a].N > hP[b].N; });
    std::vector<uint2> items, bvh_lds, bvh_hbm;
    items.reserve(n_steps_all);
    uint32_t lds_bytes = 0;
    for (uint32_t k : by_size) {
        const bool in_lds = bvh_scratch_bytes_host(hP[k].N) <= kBvhLdsMaxHost;
        if (in_lds) lds_bytes = std::max<uint32_t>(lds_bytes, (uint32_t)bvh_scratch_bytes_host(hP[k].N));
        for (uint32_t c = 0; c < hP[k].C; ++c) {
            items.push_back(make_uint2(k, c));
            (in_lds ? bvh_lds : bvh_hbm).push_back(make_uint2(k, c));
        }
    }
    const size_t n_tab = 6u * (n + 1u);
    MZH_HIP(d_tab.ensure(sizeof(uint32_t) * n_tab));
    MZH_HIP(d_items.ensure(sizeof(uint2) * (items.size() + bvh_lds.size() + bvh_hbm.size() + 1u)));
    std::vector<uint32_t> tab;
    tab.reserve(n_tab);
    for (const auto *v : {&step_pre, &aabb_pre, &nc_pre, &bp_pre, &gather_pre, &lbvh_pre})
        tab.insert(tab.end(), v->begin(), v->end());
    MZH_HIP(hipMemcpyAsync(d_tab.p, tab.data(), sizeof(uint32_t) * n_tab, hipMemcpyHostToDevice, st));
    std::vector<uint2> all(items);
    all.insert(all.end(), bvh_lds.begin(), bvh_lds.end());
    all.insert(all.end(), bvh_hbm.begin(), bvh_hbm.end());
    if (!all.empty()) MZH_HIP(hipMemcpyAsync(d_items.p, all.data(), sizeof(uint2) * all.size(), hipMemcpyHostToDevice, st));
    const uint32_t *T = d_tab.as<uint32_t>();
    tabs.steps = T;
    tabs.aabb_blk = T + (n + 1);
    tabs.nc_blk = T + 2 * (n + 1);
    tabs.bp_blk = T + 3 * (n + 1);
    tabs.gather_blk = T + 4 * (n + 1);
    tabs.lbvh_blk = T + 5 * (n + 1);
    const uint2 *I = d_items.as<uint2>();
    tabs.sort_items = I;
    tabs.bvh_items = I + items.size();
    tabs.bvh_hbm_items = I + items.size() + bvh_lds.size();
    tabs.n_sort_items = (uint32_t)items.size();
    tabs.n_bvh_items = (uint32_t)bvh_lds.size();
    tabs.n_bvh_hbm_items = (uint32_t)bvh_hbm.size();
    tabs.bvh_lds_bytes = lds_bytes;
    // batch-wide narrow-phase queues (sizes from earlier runs, else a start)
    MZH_HIP(d_P.ensure(sizeof(Prob) * n));
    upload_probs();
    MZH_HIP(wl_slow.ensure(narrow_wl_slow_bytes(max_order)));
    gens.resize(n);
    for (uint32_t k = 0; k < n; ++k) gens[k] = probs[k]->generation;
    prepared = true;
    exec_ok = false;
}

void ScreenBatch::enqueue_stages(bool timed)
{
    Prob *P = d_P.as<Prob>();
    const uint32_t n = np();
    BatchCounters *bc = d_bc.as<BatchCounters>();
    if (timed) MZH_HIP(hipEventRecord(ev[0], st));
    MZH_HIP(launch_batch_init(P, n, tabs, n_steps_all, bc, st));
    // stage 1: AABBs
    if (timed) MZH_HIP(hipEventRecord(ev[6], st));
    for (const Group &g : groups) {
        if (g.any_fp64) MZH_HIP(launch_aabbs(g.order, MZH_MODE_FP64, P, n, tabs, g.blk_begin, g.n_blk, st));
        if (g.any_mixed) MZH_HIP(launch_aabbs(g.order, MZH_MODE_MIXED, P, n, tabs, g.blk_begin, g.n_blk, st));
    }
    if (timed) MZH_HIP(hipEventRecord(ev[7], st));
    MZH_HIP(launch_gbox_finalize(P, n, tabs, n_steps_all, st));
    if (timed) MZH_HIP(hipEventRecord(ev[1], st));
    // stage 2: Morton codes, sort, gather
    MZH_HIP(launch_morton_sort(P, n, tabs, n_steps_all, n_gather_blocks, n_lbvh_blocks, bc, ls, st));
    if (timed) MZH_HIP(hipEventRecord(ev[2], st));
    // stage 3: the walk tree
    MZH_HIP(launch_lbvh(P, n, tabs, n_lbvh_blocks, n_steps_all, st));
    if (timed) MZH_HIP(hipEventRecord(ev[3], st));
    // stage 4: broad phase (counts, scans and the pair prefix)
    MZH_HIP(launch_bp_reset(P, n, tabs, n_nc_blocks, st));
    if (timed) MZH_HIP(hipEventRecord(ev[10], st));
    MZH_HIP(launch_bp_count(P, n, tabs, n_bp_blocks, ls, st));
    if (timed) MZH_HIP(hipEventRecord(ev[11], st));
    MZH_HIP(launch_scan_counts(P, n, tabs, n_steps_all, st));
    MZH_HIP(launch_scan_steps(P, n, st));
    MZH_HIP(launch_pair_prefix(P, n, d_pair_pre.as<uint64_t>(), st));
}

NarrowArgs ScreenBatch::narrow_args(const Group &g) const
{
    NarrowArgs na{};
    na.P = d_P.as<const Prob>();
    na.np = np();
    na.order = g.order;
    na.pair_pre = d_pair_pre.as<const uint64_t>();
    const uint32_t op1 = g.order + 1;
    const double pi = 3.141592653589793238462643383279502884;
    for (uint32_t m = 0; m < op1; ++m) na.cheby[m] = (std::cos((2. * m + 1.) / (2. * op1) * pi) + 1.) / 2.;
    unsigned long long *ctr = d_ctr.as<unsigned long long>();
    na.conj = conj_raw.as<mzh_conj>();
    na.conj_step = conj_step.as<uint32_t>();
    na.conj_cap = conj_cap;
    na.conj_count = ctr + 0;
    na.steps = steps.p;
    na.steps_cap = steps_cap;
    na.steps_count = ctr + 1;
    na.cand = cand.p;
    na.cand_cap = cand_cap;
    na.cand_count = ctr + 2;
    na.wl_fast_cap = (int)std::min<uint64_t>(env_u64("MIZUBA_AMD_WL_FAST_CAP", kNarrowWlFastCap), kNarrowWlFastCap);
    na.slow = slow.as<uint64_t>();
    na.slow_cap = slow_cap;
    na.slow_count = ctr + 3;
    na.wl_slow = wl_slow.p;
    na.peak = ctr + 4;
    na.roots_wgs = ls.roots_wgs;
    return na;
}

// The narrow phase over every order group's pairs, in passes of at most
// pass_pairs() pairs; group_pairs[g] bounds the group's pairs.
void ScreenBatch::enqueue_narrow(const std::vector<uint64_t> &group_pairs)
{
    const uint64_t pp = pass_pairs();
    ensure_queues();
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        NarrowArgs na = narrow_args(groups[gi]);
        na.p_first = groups[gi].p_first;
        na.p_end = groups[gi].p_end;
        if (gi == 0) MZH_HIP(launch_narrow_reset(na, st));
        for (uint64_t pb = 0; pb < group_pairs[gi]; pb += pp) {
            na.pair_begin = pb;
            na.pair_end = std::min<uint64_t>(group_pairs[gi], pb + pp);
            MZH_HIP(launch_narrow(na, st));
        }
    }
}

void ScreenBatch::ensure_queues()
{
    MZH_HIP(steps.ensure(sizeof(StepPair) * (steps_cap ? steps_cap : 1)));
    MZH_HIP(cand.ensure(max_cand_bytes * (cand_cap ? cand_cap : 1)));
    MZH_HIP(slow.ensure(sizeof(uint64_t) * (slow_cap ? slow_cap : 1)));
    MZH_HIP(conj_raw.grow(sizeof(mzh_conj) * (conj_cap ? conj_cap : 1), 0, st));
    MZH_HIP(conj_step.grow(sizeof(uint32_t) * (conj_cap ? conj_cap : 1), 0, st));
    MZH_HIP(conj_keys.ensure(2u * sizeof(ConjKey) * (conj_cap ? conj_cap : 1)));
}

// Order the batch's conjunction records (n of them, or at most n with the
// count read on the device) into the problems' outputs.
void ScreenBatch::enqueue_conj_sort(uint64_t n, bool dev_count)
{
    const unsigned long long *n_dev = dev_count ? d_ctr.as<unsigned long long>() : nullptr;
    MZH_HIP(launch_conj_sort(d_P.as<Prob>(), np(), tabs.steps, conj_raw.as<mzh_conj>(), conj_step.as<uint32_t>(), n,
                             n_dev, conj_keys.as<ConjKey>(), conj_keys.as<ConjKey>() + (n ? n : 1), st));
}

void ScreenBatch::read_back(std::vector<uint64_t> &counts, std::vector<uint32_t> &flags)
{
    const uint32_t n = np();
    counts.resize(4u * n);
    flags.resize(n);
    MZH_HIP(hipMemcpyAsync(counts.data(), d_counts.p, sizeof(uint64_t) * 4u * n, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipMemcpyAsync(flags.data(), d_flags.p, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipStreamSynchronize(st));
}

// Eager run: the host sizes the pair buffers, the narrow-phase queues and the
// conjunction outputs from the counts of the run itself.
void ScreenBatch::run_eager(bool timed)
{
    const uint32_t n = np();
    std::vector<uint64_t> counts;
    std::vector<uint32_t> flags;
    std::vector<DevError> errs(n);
    enqueue_stages(timed);
    // pair buffers of exactly the problems' pair counts; a full overflow pool
    // runs the broad phase again with a larger one
    for (int attempt = 0;; ++attempt) {
        read_back(counts, flags);
        bool redo = false, resized = false;
        for (uint32_t k = 0; k < n; ++k) {
            mzh_ctx *ctx = probs[k];
            if (flags[k] & kFlagOvf) {
                ctx->ovf_cap = (uint32_t)std::min<uint64_t>(grown(counts[4u * k + 1]), 0x7fffffffu);
                MZH_HIP(ctx->ovf.ensure(sizeof(uint2) * ctx->ovf_cap));
                hP[k].ovf = gl(ctx->ovf.as<uint2>());
                hP[k].ovf_cap = ctx->ovf_cap;
                redo = true;
            }
            if (counts[4u * k] > ctx->pair_cap) {
                // exact, or with the headroom that later graph replays need
                ctx->pair_cap = use_graph ? grown(counts[4u * k]) : counts[4u * k];
                MZH_HIP(ctx->pairs.ensure(sizeof(mzh_aabb_collision) * ctx->pair_cap));
                MZH_HIP(ctx->pair_c.ensure(sizeof(uint32_t) * ctx->pair_cap));
                hP[k].pairs = gl(ctx->pairs.as<mzh_aabb_collision>());
                hP[k].pair_c = gl(ctx->pair_c.as<uint32_t>());
                hP[k].pair_cap = ctx->pair_cap;
                resized = true;
            }
        }
        if (!redo && !resized) break;
        if (attempt == 3) throw StatusError{MZH_ERUNTIME, "broad-phase overflow pool capacity exceeded"};
        upload_probs();
        if (redo) {
            // the whole broad phase again (flags and pool counters cleared first)
            enqueue_stages(timed);
        } else {
            MZH_HIP(hipMemsetAsync(d_flags.p, 0, sizeof(uint32_t) * n, st));
            MZH_HIP(launch_pair_prefix(d_P.as<Prob>(), n, d_pair_pre.as<uint64_t>(), st));
        }
    }
    enqueue_place();
    if (timed) MZH_HIP(hipEventRecord(ev[4], st));
    // narrow phase: queues sized from the pair count, grown and run again when
    // a pass overflows them
    std::vector<uint64_t> gp(groups.size(), 0);
    uint64_t total = 0;
    for (size_t gi = 0; gi < groups.size(); ++gi) {
        for (uint32_t k = groups[gi].p_first; k < groups[gi].p_end; ++k)
            if (!(flags[k] & (kFlagPairs | kFlagOvf))) gp[gi] += counts[4u * k];
        total += gp[gi];
    }
    const uint64_t chunk = std::min<uint64_t>(std::max<uint64_t>(total, 1), pass_pairs());
    // overlapping trajectory-step pairs: about 1.5 per pair for T/8 steps, 2.2
    // for the T/20 steps of SGP4 polyjectories; candidates: about a fifth of
    // the pairs on LEO catalogs
    steps_cap = std::max<uint64_t>(steps_cap, env_u64("MIZUBA_AMD_INIT_STEPS_CAP", 3 * chunk + 65536));
    cand_cap = std::max<uint64_t>(cand_cap, env_u64("MIZUBA_AMD_INIT_CAND_CAP", chunk / 2 + 65536));
    slow_cap = std::max<uint64_t>(slow_cap, env_u64("MIZUBA_AMD_INIT_SLOW_CAP", 4096));
    conj_cap = std::max<uint64_t>(conj_cap, env_u64("MIZUBA_AMD_INIT_CONJ_CAP", 1u << 16));
    unsigned long long ctr[8];
    for (int attempt = 0; attempt < 8; ++attempt) {
        if (timed) MZH_HIP(hipEventRecord(ev[8], st));
        enqueue_narrow(gp);
        if (timed) MZH_HIP(hipEventRecord(ev[9], st));
        MZH_HIP(hipMemcpyAsync(ctr, d_ctr.p, sizeof(ctr), hipMemcpyDeviceToHost, st));
        MZH_HIP(hipStreamSynchronize(st));
        const bool ok = ctr[0] <= conj_cap && ctr[4] <= steps_cap && ctr[5] <= cand_cap && ctr[6] <= slow_cap;
        if (ok) break;
        if (attempt == 7) throw StatusError{MZH_ERUNTIME, "narrow-phase buffers overflowed after resizing"};
        if (ctr[0] > conj_cap) conj_cap = grown(ctr[0]);
        if (ctr[4] > steps_cap) steps_cap = grown(ctr[4]);
        if (ctr[5] > cand_cap) cand_cap = grown(ctr[5]);
        if (ctr[6] > slow_cap) slow_cap = grown(ctr[6]);
    }
    const uint64_t M = ctr[0];
    if (M >= 0xffffffffull) throw StatusError{MZH_ERUNTIME, "Too many conjunctions for one screening"};
    // conjunctions: each problem's output sized by its own count
    for (int attempt = 0; attempt < 2; ++attempt) {
        enqueue_conj_sort(M, false);
        read_back(counts, flags);
        bool resized = false;
        for (uint32_t k = 0; k < n; ++k) {
            mzh_ctx *ctx = probs[k];
            if (counts[4u * k + 2] > ctx->conj_out_cap) {
                ctx->conj_out_cap = use_graph ? grown(counts[4u * k + 2]) : counts[4u * k + 2];
                MZH_HIP(ctx->conj_sorted.ensure(sizeof(mzh_conj) * ctx->conj_out_cap));
                hP[k].conj_out = gl(ctx->conj_sorted.as<mzh_conj>());
                hP[k].conj_out_cap = ctx->conj_out_cap;
                resized = true;
            }
        }
        if (!resized) break;
        upload_probs();
        MZH_HIP(hipMemsetAsync(d_flags.p, 0, sizeof(uint32_t) * n, st));
    }
    if (timed) MZH_HIP(hipEventRecord(ev[5], st));
}

// Capture the whole run into a graph, with capacities that hold the last
// run's counts plus a quarter. Called after an eager run.
void ScreenBatch::capture()
{
    const uint32_t n = np();
    std::vector<uint64_t> counts;
    std::vector<uint32_t> flags;
    read_back(counts, flags);
    unsigned long long ctr[8];
    MZH_HIP(hipMemcpy(ctr, d_ctr.p, sizeof(ctr), hipMemcpyDeviceToHost));
    // the results of the eager run stay where they are: the buffers that hold
    // them already have the headroom (run_eager sized them so); only the
    // overflow pool, scratch, may grow here
    group_pair_ub.assign(groups.size(), 0);
    for (uint32_t k = 0; k < n; ++k) {
        mzh_ctx *ctx = probs[k];
        const uint32_t oc = (uint32_t)std::min<uint64_t>(grown(counts[4u * k + 1]), 0x7fffffffu);
        if (oc > ctx->ovf_cap) {
            ctx->ovf_cap = oc;
            MZH_HIP(ctx->ovf.ensure(sizeof(uint2) * oc));
        }
        hP[k].ovf = gl(ctx->ovf.as<uint2>());
        hP[k].ovf_cap = ctx->ovf_cap;
    }
    for (size_t gi = 0; gi < groups.size(); ++gi)
        for (uint32_t k = groups[gi].p_first; k < groups[gi].p_end; ++k) group_pair_ub[gi] += hP[k].pair_cap;
    steps_cap = std::max<uint64_t>(steps_cap, grown(ctr[4]));
    cand_cap = std::max<uint64_t>(cand_cap, grown(ctr[5]));
    slow_cap = std::max<uint64_t>(slow_cap, grown(ctr[6]));
    conj_cap = std::max<uint64_t>(conj_cap, grown(ctr[0]));
    ensure_queues();
    upload_probs();
    MZH_HIP(hipStreamSynchronize(st));
    if (exec) {
        (void)hipGraphExecDestroy(exec);
        exec = nullptr;
    }
    exec_ok = false;
    hipGraph_t graph = nullptr;
    MZH_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    try {
        enqueue_stages(false);
        enqueue_place();
        enqueue_narrow(group_pair_ub);
        enqueue_conj_sort(conj_cap, true);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    MZH_HIP(hipStreamEndCapture(st, &graph));
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    MZH_HIP(e);
    for (uint32_t k = 0; k < n; ++k) gens[k] = probs[k]->generation;
    exec_ok = true;
}

// One replay of the captured graph. False when a capacity overflowed (the
// results are then incomplete and the caller runs eagerly).
bool ScreenBatch::replay()
{
    MZH_HIP(hipEventRecord(ev[0], st));
    MZH_HIP(hipGraphLaunch(exec, st));
    MZH_HIP(hipEventRecord(ev[5], st));
    std::vector<uint64_t> counts;
    std::vector<uint32_t> flags;
    read_back(counts, flags);
    ++n_replays;
    unsigned long long ctr[8];
    MZH_HIP(hipMemcpy(ctr, d_ctr.p, sizeof(ctr), hipMemcpyDeviceToHost));
    if (ctr[0] > conj_cap || ctr[4] > steps_cap || ctr[5] > cand_cap || ctr[6] > slow_cap) return false;
    for (uint32_t f : flags)
        if (f) return false;
    return true;
}

// Host state of every problem after a run; status in batch order.
void ScreenBatch::finalize(int *status, bool replayed)
{
    const uint32_t n = np();
    std::vector<uint64_t> counts(4u * n);
    std::vector<uint32_t> flags(n);
    std::vector<DevError> errs(n);
    std::vector<unsigned long long> rep(kReportWords * (size_t)n);
    MZH_HIP(hipMemcpyAsync(counts.data(), d_counts.p, sizeof(uint64_t) * 4u * n, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipMemcpyAsync(errs.data(), d_err.p, sizeof(DevError) * n, hipMemcpyDeviceToHost, st));
    MZH_HIP(hipMemcpyAsync(rep.data(), d_np.p, sizeof(unsigned long long) * rep.size(), hipMemcpyDeviceToHost, st));
    for (uint32_t k = 0; k < n; ++k) {
        mzh_ctx *ctx = probs[k];
        ctx->h_tree_size.resize(ctx->C);
        ctx->h_step_off.resize(ctx->C + 1);
        MZH_HIP(hipMemcpyAsync(ctx->h_tree_size.data(), ctx->tree_size.p, sizeof(uint32_t) * ctx->C,
                               hipMemcpyDeviceToHost, st));
        MZH_HIP(hipMemcpyAsync(ctx->h_step_off.data(), ctx->step_off.p, sizeof(uint64_t) * (ctx->C + 1),
                               hipMemcpyDeviceToHost, st));
    }
    MZH_HIP(hipStreamSynchronize(st));
    float tot = 0;
    MZH_HIP(hipEventElapsedTime(&tot, ev[0], ev[5]));
    run_s = tot * 1e-3;
    double stage[MZH_STAGE_COUNT] = {};
    if (!replayed) {
        float ms = 0;
        const int pairs[8][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {4, 5}, {0, 5}, {6, 7}, {8, 9}};
        for (int k = 0; k < 8; ++k) {
            if (hipEventElapsedTime(&ms, ev[pairs[k][0]], ev[pairs[k][1]]) == hipSuccess) stage[k] = ms * 1e-3;
        }
        if (any_bp && hipEventElapsedTime(&ms, ev[10], ev[11]) == hipSuccess) stage[8] = ms * 1e-3;
    } else {
        stage[5] = run_s;
    }
    for (uint32_t k = 0; k < n; ++k) {
        mzh_ctx *ctx = probs[k];
        int st_k = MZH_OK;
        std::string msg;
        if (errs[k].code != kErrNone) {
            st_k = MZH_EINVAL;
            msg = dev_error_message(errs[k]);
        } else {
            for (uint64_t c = 0; c < ctx->C; ++c)
                if (ctx->h_tree_size[c] > ctx->node_cap) {
                    st_k = MZH_ERUNTIME;
                    msg = "BVH node capacity exceeded";
                    break;
                }
        }
        status[k] = st_k;
        if (st_k != MZH_OK) {
            ctx->last_error = msg;
            ctx->screened = false;
            continue;
        }
        ctx->n_pairs = ctx->skip_cd ? 0 : counts[4u * k];
        ctx->n_conj = counts[4u * k + 2];
        for (int m = 0; m < MZH_NP_COUNT; ++m) {
            uint64_t sum = 0;
            for (uint32_t sl = 0; sl < kReportSlots; ++sl) sum += rep[(size_t)k * kReportWords + sl * kReportStride + m];
            ctx->h_np[m] = sum;
        }
        for (int m = 0; m < MZH_STAGE_COUNT; ++m) ctx->stage_s[m] = stage[m];
        ctx->nodes_ready = false;
        ctx->screened = true;
        ctx->results_released = false;
    }
}

void ScreenBatch::run(int *status)
{
    std::vector<int> st_b(caller.size(), MZH_OK);
    // a context screened by another batch since (or changed: uploads,
    // releases, new plans) may have moved its buffers: describe it again
    for (mzh_ctx *c : caller)
        if (c->last_batch != this) {
            ++c->generation;
            c->last_batch = this;
        }
    if (!prepared || gens_changed()) prepare();
    if (use_graph && exec_ok) {
        if (replay()) {
            finalize(st_b.data(), true);
            for (size_t k = 0; k < back.size(); ++k) status[back[k]] = st_b[k];
            return;
        }
        // an overflow: run eagerly, capture again with the new sizes
        exec_ok = false;
    }
    run_eager(true);
    finalize(st_b.data(), false);
    if (use_graph) capture();
    for (size_t k = 0; k < back.size(); ++k) status[back[k]] = st_b[k];
}

// Coordinate search over the launch shapes, timing the kernels they govern
// with HIP events on this batch (every shape gives the same results).
void ScreenBatch::tune()
{
    double n_avg = 0, c_avg = 0;
    for (mzh_ctx *c : caller) {
        n_avg += c->n_objs;
        c_avg += (double)c->C;
    }
    n_avg /= caller.size();
    c_avg /= caller.size();
    const auto key = std::make_tuple(caller.front()->order, log2_class(n_avg), log2_class(c_avg), caller.front()->mode);
    {
        std::lock_guard<std::mutex> lk(g_tune_m);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            set_shape(it->second);
            return;
        }
    }
    std::vector<int> status(caller.size());
    const bool g = use_graph;
    use_graph = false;
    auto time_of = [&](int stage) {
        // device time of the stage in a fresh eager run (median of 3)
        double t[3];
        for (double &x : t) {
            run(status.data());
            x = caller.front()->stage_s[stage];
        }
        std::sort(t, t + 3);
        return t[1];
    };
    LaunchShape best = ls;
    struct Knob {
        uint32_t LaunchShape::*field;
        std::vector<uint32_t> values;
        int stage;  // stage time that the knob governs
    };
    const std::vector<Knob> knobs = {
        {&LaunchShape::bucket_bits, {10, 11, 12}, 1},
        {&LaunchShape::sort_wgs, {256, 512, 1024}, 1},
        {&LaunchShape::bp_block, {128, 256, 512}, 8},
        {&LaunchShape::roots_wgs, {256, 512, 1024}, 7},
    };
    for (const Knob &kb : knobs) {
        uint32_t bv = best.*kb.field;
        double bt = 1e30;
        for (uint32_t v : kb.values) {
            LaunchShape s = best;
            s.*kb.field = v;
            set_shape(s);
            const double t = time_of(kb.stage);
            if (t < bt) {
                bt = t;
                bv = v;
            }
        }
        best.*kb.field = bv;
    }
    set_shape(best);
    use_graph = g;
    std::lock_guard<std::mutex> lk(g_tune_m);
    g_tuned[key] = best;
}

void BatchDeleter::operator()(ScreenBatch *b) const { delete b; }

ScreenBatch *batch_new(hipStream_t stream) { return new ScreenBatch(stream); }

void batch_run(ScreenBatch *b, const std::vector<mzh_ctx *> &ctxs, bool graph, bool tune, int *status)
{
    b->set_problems(ctxs);
    if (tune) b->tune();
    b->set_graph(graph);
    b->run(status);
}

uint64_t batch_replays(const ScreenBatch *b) { return b->replays(); }
double batch_seconds(const ScreenBatch *b) { return b->last_run_seconds(); }
LaunchShape batch_shape(const ScreenBatch *b) { return b->shape(); }

} // namespace mzh::host
