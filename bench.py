"""Benchmark: screened object-days/s on one MI355X.

Workload: 30 000 synthetic objects (80 % LEO perigee 400-1200 km e < 0.01,
20 % perigee 1200-2000 km e < 0.05, i in [0, 100] deg, seeded) over 1 day,
order-8 FP64 polyjectory, conj_thresh 5 km, conj_det_interval 3 min. The
objects are SGP4 mean elements made into a polyjectory by this library's GPU
make_sgp4_polyjectory (`--catalog sgp4`, the default; `--catalog kepler`
takes J2-secular Keplerian orbits fitted by make_synthetic_polyjectory).
One step = one full conjunction screening of that polyjectory (480
conjunction steps: AABBs -> Morton/sort -> BVH -> broad -> narrow phase), with
the polyjectory already resident in HBM. `value` is the mixed-precision
pipeline; the all-FP64 drop-in is timed beside it on the same GPU and their
conjunction lists are compared. The CPU oracle (the C restatement of the
reference path, OpenMP over all host cores) is timed on a bounded sample of
the same workload.

Prints ONE JSON line.
"""

from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "screened object-days/s at 30k objects; mixed vs FP64 GPU at identical conj. list"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6


def aabb_bytes_per_launch(pj, interval, coef_bytes=8):
    """Algorithmic bytes of one AABB launch: per (object, conjunction step)
    304 B per overlapping trajectory step (4 channels x 9 FP64 coefficients +
    2 times; 160 B with the mixed-precision kernel's float coefficients,
    coef_bytes=4) + 32 B box write."""
    traj, tm, toff, ns, tmo = pj._flat()
    op1 = pj.poly_order + 1
    C = int(np.ceil(pj.maxT / interval))
    cb = interval * np.arange(C, dtype=np.float64)
    ce = np.minimum(interval * np.arange(1, C + 1, dtype=np.float64), pj.maxT)
    ce[-1] = pj.maxT
    pieces = 0
    for o in range(pj.n_objs):
        n = int(ns[o])
        if n == 0:
            continue
        t = tm[int(tmo[o]) : int(tmo[o]) + n + 1]
        ov = (t[0] < ce) & (t[n] > cb)
        kb = np.searchsorted(t[1:], cb, side="right")
        ke = np.searchsorted(t[1:], ce, side="left")
        ke = ke + (ke != n)
        pieces += int(np.sum(np.where(ov, ke - kb, 0)))
    per_piece = 4 * op1 * coef_bytes + 16
    return pieces * per_piece + C * pj.n_objs * 32, pieces / (C * pj.n_objs)


def broad_bytes_per_launch(n_objs, n_steps, n_pairs):
    """Algorithmic bytes of one broad-phase traversal launch: per (object,
    conjunction step) the query box (32 B), its sorted-to-original index and
    object type (8 B), its pair count (4 B) and its share of the step's tree
    read once (2 nodes of the reference's 48 B layout; the walk tree's own
    records are 32 B), plus one 4 B partner index per pair."""
    return n_steps * n_objs * (32 + 8 + 4 + 96) + 4 * n_pairs


def bvh_bytes_per_launch(n_objs, n_steps):
    """Algorithmic bytes of one build of the broad phase's walk tree
    (k_lbvh_build, then k_lbvh_top for the nodes that span blocks): per
    (object, conjunction step) the sorted code (8 B) and box (32 B) read once
    and two nodes (a leaf and an internal node) of the reference's 48 B layout
    written: 136 B per object-step (the walk tree's own records are 32 B)."""
    return n_steps * n_objs * (8 + 32 + 96)


def hbm_traffic(kernel_prefix):
    """Measured HBM bytes per dispatch of a kernel, from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel calibration
    recorded in the same file), or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "screen_30k_1day_hbm_traffic.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and (cells[0] == kernel_prefix or cells[0].startswith(kernel_prefix + "<")):
            try:
                return int(float(cells[4]))
            except ValueError:
                return None
    return None


def copy_calibration(mib):
    """Plain streaming kernels of known size for the PMC calibration: this
    library's own HIP copy kernel (k_stream_copy through mzh_stream_copy, in
    this process, 3 launches) and a torch elementwise kernel (3 launches), each
    reading and writing `mib` MiB per launch."""
    import ctypes

    import torch

    nbytes = mib * 1024 * 1024
    lib = ctypes.CDLL(os.path.join(ROOT, "mizuba_amd", "libmizuba_amd.so"))
    ctx = ctypes.c_void_p()
    if lib.mzh_create(0, ctypes.byref(ctx)) != 0:
        raise RuntimeError("mzh_create failed")
    sec = ctypes.c_double()
    rc = lib.mzh_stream_copy(ctx, ctypes.c_uint64(nbytes), 3, ctypes.byref(sec))
    lib.mzh_destroy(ctx)
    if rc != 0:
        raise RuntimeError(f"mzh_stream_copy failed ({rc})")
    n = nbytes // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": nbytes, "copy_calibration_bytes_written": nbytes, "launches": 3,
                      "k_stream_copy_ms": round(sec.value * 1e3, 4),
                      "k_stream_copy_gbs": round(2 * nbytes / sec.value / 1e9, 1)}))


def batch_workload(args):
    """The batching configuration: 64 independent problems, object counts
    log-uniform in [500, 20000], conj_thresh U[1, 10] km, horizons U[0.5, 3]
    days, 3-minute conjunction steps, seeded SGP4 mean elements (same mixture)
    made into polyjectories by make_sgp4_polyjectory, mixed precision. One
    step = one screening of all 64 problems through a batch handle: one launch
    per stage for all problems, replayed as a captured HIP graph, launch
    shapes tuned on the first run. Timed beside it: the same batch run eagerly
    (host sizing between phases) and the problems screened one by one."""
    import mizuba_amd as mz

    rng = np.random.default_rng(args.seed)
    n_prob = 64
    ns = np.exp(rng.uniform(np.log(500), np.log(20000), n_prob)).astype(int)
    threshs = [float(x) for x in rng.uniform(1.0, 10.0, n_prob)]
    days = [float(x) for x in rng.uniform(0.5, 3.0, n_prob)]
    interval = 3.0 / 1440.0
    jd0 = 2460496.5
    t0 = time.perf_counter()
    pjs = []
    for k in range(n_prob):
        g = mz.make_synthetic_gpes(int(ns[k]), seed=args.seed + 1000 + k, frac_high=0.2, epoch_jd=jd0)
        pjs.append(mz.make_sgp4_polyjectory(g, jd0, jd0 + days[k])[0])
    print(f"[bench] {n_prob} SGP4 problems ({int(ns.sum())} objects) in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    obj_days = float(sum(pj.n_objs * pj.maxT for pj in pjs))
    scrs = [mz.Screener(0) for _ in pjs]
    for scr, pj in zip(scrs, pjs):
        scr.upload(pj)
    intervals = [interval] * n_prob
    # graph replays (tuned shapes, captured after the first run)
    bg = mz.Batch(scrs, threshs, intervals, mode="mixed", graph=True, tune=True)
    t = time.perf_counter()
    bg.run()
    tune_s = time.perf_counter() - t
    for _ in range(max(args.warmup, 1)):
        bg.run()
    dev = []
    t = time.perf_counter()
    for _ in range(args.steps):
        bg.run()
        dev.append(bg.seconds)
    dt_graph = (time.perf_counter() - t) / args.steps
    graph_res = [scr.conjunctions() for scr in scrs]
    tuned_shape = bg.shape
    # the 64 polyjectories and contexts fill most of the device: each batch's
    # narrow-phase queues are freed before the next run allocates its own
    del bg
    gc.collect()
    # the same batch run eagerly, with the same shapes
    be = mz.Batch(scrs, threshs, intervals, mode="mixed", graph=False)
    be.run()
    t = time.perf_counter()
    for _ in range(args.steps):
        be.run()
    dt_eager = (time.perf_counter() - t) / args.steps
    stages = scrs[0].stage_times()
    eager_res = [scr.conjunctions() for scr in scrs]
    n_pairs = sum(scr.n_pairs() for scr in scrs)
    del be
    gc.collect()
    # one by one
    t = time.perf_counter()
    for _ in range(args.steps):
        for scr, th in zip(scrs, threshs):
            scr.screen(th, interval, mode="mixed")
    dt_single = (time.perf_counter() - t) / args.steps
    single = [scr.conjunctions() for scr in scrs]
    same = all(np.array_equal(a.view(np.uint8), b.view(np.uint8)) and np.array_equal(a.view(np.uint8),
                                                                                        c.view(np.uint8))
               for a, b, c in zip(graph_res, eager_res, single))
    # roofline: the broad-phase walk of the eager batch run, the batch's
    # dominant kernel (algorithmic bytes summed over the problems)
    kb = stages["broad_kernel"]
    C_all = [int(np.ceil(pj.maxT / interval)) for pj in pjs]
    bp_bytes = sum(broad_bytes_per_launch(pj.n_objs, C, 0) for pj, C in zip(pjs, C_all)) + 4 * n_pairs
    out = {
        "metric": "screened object-days/s, 64 independent problems screened together",
        "value": round(obj_days / dt_graph, 1),
        "unit": "object-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_graph * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 screen + f64 narrow phase",
        "data": "synthetic (seeded SGP4 mean elements, polyjectories from this library's GPU make_sgp4_polyjectory, "
                "order 8)",
        "config": {
            "workload": "64 problems, N log-uniform in [500, 20000], conj_thresh U[1, 10] km, horizon U[0.5, 3] d, "
            "conj_det_interval 3 min, mixed precision",
            "n_problems": n_prob,
            "n_objs_total": int(ns.sum()),
            "object_days": round(obj_days, 1),
            "seed": args.seed,
        },
        "graph_device_ms": round(float(np.median(dev)) * 1e3, 3),
        "eager_batch": {"value": round(obj_days / dt_eager, 1), "ms_per_step": round(dt_eager * 1e3, 3),
                        "stages_ms": {k: round(v * 1e3, 3) for k, v in stages.items()}},
        "one_by_one": {"value": round(obj_days / dt_single, 1), "ms_per_step": round(dt_single * 1e3, 3)},
        "graph_over_one_by_one": round(dt_single / dt_graph, 3),
        "graph_over_eager": round(dt_eager / dt_graph, 3),
        "tuned_shape": tuned_shape,
        "first_run_with_tuning_s": round(tune_s, 3),
        "identical_conj_lists": bool(same),
        "n_conj": int(sum(len(b) for b in graph_res)),
        "n_pairs": int(n_pairs),
        "roofline": {
            "kernel": "k_bp",
            "bound": "hbm",
            "achieved": round(bp_bytes / kb / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(bp_bytes / kb / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes": int(bp_bytes),
            "avg_launch_ms": round(kb * 1e3, 4),
        },
    }
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po

        # the oracle on every problem's first steps (a bounded sample), and
        # the GPU list against it there
        threads = po.n_threads_default()
        nsamp = args.cpu_sample_steps if args.cpu_sample_steps > 0 else 24
        t = time.perf_counter()
        refs = []
        samp_days = 0.0
        for pj, th in zip(pjs, threshs):
            traj, tm, toff, nst, tmo = pj._flat()
            flat = po.Flat.from_arrays(pj.poly_order, traj, tm, nst, toff, tmo, pj.maxT)
            C = int(np.ceil(pj.maxT / interval))
            k1 = min(nsamp, C)
            refs.append(po.run(flat, th, interval, step_end=k1, n_threads=threads))
            samp_days += pj.n_objs * min(k1 * interval, pj.maxT)
        dt_cpu = time.perf_counter() - t
        ok = True
        for g, r in zip(graph_res, refs):
            t_end = r["cd_end_times"][-1]
            gg = g[g["tca"] < t_end]
            ok = ok and np.array_equal(gg.view(np.uint8), r["conjunctions"].view(np.uint8))
        out["cpu_baseline"] = {"value": round(samp_days / dt_cpu, 2), "unit": "object-days/s", "cores": threads,
                               "kind": "port", "cpu": cpu_model(),
                               "sample": f"first {nsamp} conjunction steps of each of the 64 problems "
                                         f"({samp_days:.0f} object-days), oracle parallel over steps only, "
                                         f"wall {dt_cpu:.2f}s"}
        out["oracle_check"] = {"steps_per_problem": nsamp, "identical_bits": bool(ok),
                               "n_conj": int(sum(len(r["conjunctions"]) for r in refs))}
    print(json.dumps(out))


LARGE_CONFIGS = {
    # name: (objects, of them in the dense shells, days, conj_thresh km, interval min, seed, oracle windows of steps)
    "config2": (30000, 0, 7.0, 5.0, 3.0, 2026, 48),
    "config4": (100000, 70000, 1.0, 10.0, 2.0, 2029, 8),
}


def large_config_workload(args):
    """The precision and dense-shell configurations on their own data: seeded
    SGP4 mean elements made into order-8 polyjectories by
    make_sgp4_polyjectory. config2: 30 000 objects (80% LEO, 20% higher) over
    7 days, 5 km, 3-minute steps (3 360 steps); config4: 100 000 objects, 70 000
    of them in five near-circular shells (540-580 km, 72 planes, 53 / 97.6 deg)
    and 30 000 LEO debris, 1 day, 10 km, 2-minute steps. The FP64 drop-in and
    the mixed pipeline are timed on the resident polyjectory; the oracle
    screens three windows of steps (start, middle, end), against which both
    lists are compared: identical (i, j) and the largest TCA / DCA error."""
    import mizuba_amd as mz

    n, n_shell, days, thresh, minutes, seed, win = LARGE_CONFIGS[args.workload]
    interval = minutes / 1440.0
    jd0 = 2460496.5
    t0 = time.perf_counter()
    gpes = mz.make_synthetic_gpes(n, seed=seed, frac_high=0.0 if n_shell else 0.2, n_shell=n_shell, epoch_jd=jd0)
    pj = mz.make_sgp4_polyjectory(gpes, jd0, jd0 + days)[0]
    print(f"[bench] {args.workload}: {pj.n_objs} objects x {days:g} d in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    scr = mz.Screener(0)
    scr.upload(pj)
    res = {}
    for mode in ("fp64", "mixed"):
        for _ in range(max(args.warmup, 1)):
            scr.screen(thresh, interval, mode=mode)
        t = time.perf_counter()
        stages = []
        for _ in range(args.steps):
            scr.screen(thresh, interval, mode=mode)
            stages.append(scr.stage_times())
        dt = (time.perf_counter() - t) / args.steps
        res[mode] = {
            "dt": dt,
            "conj": scr.conjunctions(),
            "n_pairs": int(scr.n_pairs()),
            "stages_ms": {k: round(float(np.mean([s_[k] for s_ in stages])) * 1e3, 3) for k in stages[0]},
        }
        print(f"[bench] {mode} {dt * 1e3:.2f} ms/step", file=sys.stderr, flush=True)
    C = int(np.ceil(pj.maxT / interval))
    obj_days = pj.n_objs * pj.maxT
    c64, cmx = res["fp64"]["conj"], res["mixed"]["conj"]
    same_ij = len(c64) == len(cmx) and bool(np.array_equal(c64["i"], cmx["i"]) and np.array_equal(c64["j"], cmx["j"]))
    out = {
        "metric": "screened object-days/s; mixed vs FP64 GPU at identical conj. list",
        "value": round(obj_days / res["mixed"]["dt"], 1),
        "unit": "object-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(res["mixed"]["dt"] * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 screen + f64 narrow phase",
        "data": "synthetic (seeded SGP4 mean elements, polyjectory from this library's GPU make_sgp4_polyjectory, "
                "order 8)",
        "config": {
            "workload": f"{args.workload}: {pj.n_objs} objects ({n_shell} in dense shells) x {days:g} d, "
                        f"conj_thresh {thresh:g} km, conj_det_interval {minutes:g} min",
            "n_objs": pj.n_objs,
            "n_cd_steps": C,
            "seed": seed,
        },
        "fp64_dropin": {"value": round(obj_days / res["fp64"]["dt"], 1),
                        "ms_per_step": round(res["fp64"]["dt"] * 1e3, 3),
                        "stages_ms": res["fp64"]["stages_ms"], "n_pairs": res["fp64"]["n_pairs"]},
        "mixed": {"stages_ms": res["mixed"]["stages_ms"], "n_pairs": res["mixed"]["n_pairs"]},
        "mixed_over_fp64": round(res["fp64"]["dt"] / res["mixed"]["dt"], 3),
        "identical_conj_list": same_ij,
        "identical_conj_bits": bool(same_ij and np.array_equal(c64.view(np.uint8), cmx.view(np.uint8))),
        "n_conj": int(len(cmx)),
    }
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po

        traj, tm, toff, ns, tmo = pj._flat()
        flat = po.Flat.from_arrays(pj.poly_order, traj, tm, ns, toff, tmo, pj.maxT)
        threads = po.n_threads_default()
        cd_end = np.minimum(interval * np.arange(1, C + 1, dtype=np.float64), pj.maxT)
        cd_end[-1] = pj.maxT
        windows = [(0, win), (C // 2 - win // 2, C // 2 - win // 2 + win), (C - win, C)]
        err = {"fp64": [0.0, 0.0, True, True], "mixed": [0.0, 0.0, True, True]}
        n_ref = 0
        t_cpu = 0.0
        for k0, k1 in windows:
            t = time.perf_counter()
            ref = po.run(flat, thresh, interval, step_begin=k0, step_end=k1, n_threads=threads)
            t_cpu += time.perf_counter() - t
            r = ref["conjunctions"]
            n_ref += len(r)
            t0w = 0.0 if k0 == 0 else cd_end[k0 - 1]
            t1w = cd_end[k1 - 1]
            for mode, cj in (("fp64", c64), ("mixed", cmx)):
                upper = (cj["tca"] <= t1w) if k1 == C else (cj["tca"] < t1w)
                g = cj[(cj["tca"] >= t0w) & upper]
                ok = len(g) == len(r) and np.array_equal(g["i"], r["i"]) and np.array_equal(g["j"], r["j"])
                e = err[mode]
                e[2] = e[2] and bool(ok)
                e[3] = e[3] and bool(ok and np.array_equal(g.view(np.uint8), r.view(np.uint8)))
                if ok and len(r):
                    e[0] = max(e[0], float(np.max(np.abs(g["tca"] - r["tca"]))) * 86400.0)
                    e[1] = max(e[1], float(np.max(np.abs(g["dca"] - r["dca"]))))
        samp_days = sum(min(k1 * interval, pj.maxT) - k0 * interval for k0, k1 in windows)
        out["oracle_check"] = {
            "windows": windows,
            "n_conj": int(n_ref),
            **{f"{m}_identical_ij": err[m][2] for m in err},
            **{f"{m}_bitwise": err[m][3] for m in err},
            **{f"{m}_max_dtca_s": err[m][0] for m in err},
            **{f"{m}_max_ddca_km": err[m][1] for m in err},
        }
        out["cpu_baseline"] = {"value": round(pj.n_objs * samp_days / t_cpu, 2), "unit": "object-days/s",
                               "cores": threads, "kind": "port", "cpu": cpu_model(),
                               "sample": f"3 windows of {win} conjunction steps ({samp_days:.4f} days) of the same "
                                         f"polyjectory, wall {t_cpu:.2f}s; the oracle runs its steps in parallel "
                                         "(OpenMP over steps)"}
    print(json.dumps(out))


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the CPU baseline."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def eval_bytes_per_launch(pj, tgrid):
    """Algorithmic bytes of one k_state_eval launch over a shared time grid:
    the 7-channel state written per evaluation (56 B), the grid read once
    (8 B per time) and the coefficient record of every (object, trajectory
    step) the grid touches read once (7 x (order + 1) x 8 B); the time arrays
    searched are counted once per object (8 B per step end)."""
    traj, time_, traj_off, nsteps, time_off = pj._flat()
    op1 = pj.poly_order + 1
    touched = 0
    searched = 0
    for o in range(pj.n_objs):
        ns = int(nsteps[o])
        if ns == 0:
            continue
        t = time_[int(time_off[o]) : int(time_off[o]) + ns + 1]
        inside = tgrid[(tgrid >= t[0]) & (tgrid < t[-1])]
        touched += len(np.unique(np.searchsorted(t[1:], inside, side="right")))
        searched += ns + 1
    n_eval = pj.n_objs * len(tgrid)
    return 56 * n_eval + 8 * len(tgrid) + touched * 7 * op1 * 8 + 8 * searched


def eval_hbm_traffic():
    path = os.path.join(ROOT, "profiles", "state_eval_30k_2048_times.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0] == "k_state_eval":
            try:
                return int(float(cells[4]))
            except ValueError:
                return None
    return None


def state_eval_workload(args):
    """Batched polyjectory evaluation (polyjectory::state_meval with one time
    grid for all objects): the 30 000-object 1-day polyjectory at 2048 times
    spread over the day, 61 M evaluations of the 7-channel state, times and
    results resident in HBM (mzh_state_eval_device). The oracle's evaluation
    is timed on a sample of the objects."""
    import mizuba_amd as mz

    n_t = 2048
    pj = mz.make_synthetic_polyjectory(args.n_objs, args.days, seed=args.seed, frac_high=0.2, order=8)
    N = pj.n_objs
    tgrid = np.arange(n_t, dtype=np.float64) * (pj.maxT / n_t)
    scr = mz.Screener(0)
    scr.upload(pj)
    # times and results in HBM, in buffers of the context's device
    d_tm = scr.device_buffer(tgrid.nbytes)
    d_tm.write(tgrid)
    d_out = scr.device_buffer(N * n_t * 7 * 8)

    def run():
        scr.state_eval_device(d_tm.ptr, 2, 0, N, n_t, d_out.ptr)

    for _ in range(args.warmup):
        run()
    steps = max(args.steps, 1) * 4
    kt = []
    t = time.perf_counter()
    for _ in range(steps):
        run()
        kt.append(scr.eval_time())
    dt = (time.perf_counter() - t) / steps
    kavg = sum(kt) / len(kt)
    n_eval = N * n_t
    # a sample of rows again through a selector, and a spot check of the full
    # run's output for the same rows
    rng = np.random.default_rng(args.seed)
    rows = np.sort(rng.choice(N, 64, replace=False)).astype(np.uint64)
    d_sel = scr.device_buffer(rows.nbytes)
    d_sel.write(rows)
    d_smp = scr.device_buffer(len(rows) * n_t * 7 * 8)
    scr.state_eval_device(d_tm.ptr, 2, d_sel.ptr, len(rows), n_t, d_smp.ptr)
    got = np.empty((len(rows), n_t, 7))
    d_smp.read(got)
    full = np.empty((N, n_t, 7))
    d_out.read(full)
    same_rows = np.array_equal(full[rows.astype(np.int64)].view(np.uint64), got.view(np.uint64))
    del full
    alg = eval_bytes_per_launch(pj, tgrid)
    traffic = eval_hbm_traffic()
    out = {
        "metric": "polyjectory state evaluations/s (state_meval, one time grid for all objects)",
        "value": round(n_eval / dt, 1),
        "unit": "evaluations/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded J2-secular Keplerian catalog, order-8 Chebyshev-fitted polyjectory)",
        "config": {"workload": f"{N} objects x {args.days:g} day, {n_t} shared times", "n_objs": N, "n_times": n_t,
                   "seed": args.seed},
        "selector_rows_match_full_run": bool(same_rows),
        "roofline": {
            "kernel": "k_state_eval",
            "bound": "hbm",
            "achieved": round(alg / kavg / 1e9, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(alg / kavg / 1e9 / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "algorithmic_bytes": int(alg),
            "avg_launch_ms": round(kavg * 1e3, 4),
        },
    }
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po

        traj, time_, traj_off, nsteps, time_off = pj._flat()
        flat = po.Flat.from_arrays(pj.poly_order, traj, time_, nsteps, traj_off, time_off, pj.maxT)
        # the checker: the sampled rows against the oracle, bit for bit (NaN included)
        st, ref, _ = po.state_meval(flat, tgrid, po.TIME_SHARED, sel=rows, n_times=n_t)
        out["identical_to_oracle_sample"] = bool(st == 0 and np.array_equal(got.view(np.uint64),
                                                                            ref.view(np.uint64)))
        n_s = N  # the whole workload: about 2 s of one core
        t = time.perf_counter()
        po.state_meval(flat, tgrid, po.TIME_SHARED, sel=np.arange(n_s, dtype=np.uint64), n_times=n_t)
        dt_cpu = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(n_s * n_t / dt_cpu, 1), "unit": "evaluations/s", "cores": 1,
                               "kind": "port",
                               "sample": f"first {n_s} objects x {n_t} times, wall {dt_cpu:.2f}s"}
    print(json.dumps(out))


def synthetic_gpes(n, jd_begin, seed):
    """A seeded near-earth catalog: mean motions of 11-16 rev/day (most above
    14), small eccentricities (a few up to 0.2, perigee above 200 km
    altitude), inclinations up to 108 deg, uniform angles, B* around 1e-4 and
    epochs up to 2 days before jd_begin; one GPE per satellite."""
    import mizuba_amd as mz

    rng = np.random.default_rng(seed)
    g = np.zeros(n, dtype=mz.gpe_dtype)
    g["norad_id"] = np.arange(1, n + 1)
    rev = np.where(rng.uniform(size=n) < 0.7, rng.uniform(14.0, 16.0, n), rng.uniform(11.0, 14.0, n))
    n0 = rev * 2.0 * np.pi / 1440.0
    a = (398600.8 / (n0 / 60.0) ** 2) ** (1.0 / 3.0)
    e_max = np.clip(1.0 - 6578.0 / a, 1e-4, 0.2)
    e = np.where(rng.uniform(size=n) < 0.95, rng.exponential(0.003, n), rng.uniform(0.0, 0.2, n))
    g["e0"] = np.clip(e, 1e-5, e_max)
    g["n0"] = n0
    g["i0"] = rng.uniform(0.0, 0.6 * np.pi, n)
    for f in ("node0", "omega0", "m0"):
        g[f] = rng.uniform(0.0, 2.0 * np.pi, n)
    g["bstar"] = 1e-4 * rng.lognormal(0.0, 1.0, n)
    ep = jd_begin - rng.uniform(0.0, 2.0, n)
    g["epoch_jd1"] = np.floor(ep - 0.5) + 0.5
    g["epoch_jd2"] = ep - g["epoch_jd1"]
    return g


def ref_sgp4_lib():
    """The reference's vendored SGP4 built from its sources into oracle/_ref
    (None where it was not built)."""
    import ctypes

    path = os.path.join(ROOT, "oracle", "_ref", "libref_sgp4.so")
    if not os.path.exists(path):
        return None
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.ref_sgp4_propagate.argtypes = [dp, dp, ctypes.c_int64, dp, ctypes.POINTER(ctypes.c_int)]
    lib.ref_sgp4_propagate.restype = ctypes.c_int
    return lib


def ref_sgp4(lib, gpe, tsince):
    import ctypes

    dp = ctypes.POINTER(ctypes.c_double)
    row = np.ascontiguousarray([float(gpe[f]) for f in gpe.dtype.names], dtype=np.float64)
    ts = np.ascontiguousarray(tsince, dtype=np.float64)
    out = np.zeros((len(ts), 6))
    err = np.zeros(len(ts), dtype=np.int32)
    rc = lib.ref_sgp4_propagate(row.ctypes.data_as(dp), ts.ctypes.data_as(dp), len(ts), out.ctypes.data_as(dp),
                                err.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    return rc, out, err


def sgp4_flops_per_launch():
    # FP64 operations of one k_sgp4_count launch on this workload, from the
    # PMC profile (the counters of the VALU's FP64 instructions x 64 lanes)
    path = os.path.join(ROOT, "profiles", "sgp4_30k_1day.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 2 and cells[0] == "fp64_ops_per_launch":
            try:
                return float(cells[1])
            except ValueError:
                return None
    return None


def sgp4_workload(args):
    """Polyjectory generation from GPEs (make_sgp4_polyjectory) on the GPU:
    a seeded near-earth catalog of --n-objs satellites over --days days, the
    polyjectory left in HBM for screening. A sample of the generated
    trajectories is checked against the reference's vendored SGP4."""
    import mizuba_amd as mz

    jd0 = 2460496.5
    N = args.n_objs
    gpes = synthetic_gpes(N, jd0, args.seed)
    scr = mz.Screener(0)
    for _ in range(args.warmup):
        scr.make_sgp4(gpes, jd0, jd0 + args.days)
    steps = max(args.steps, 1)
    wall, kt = [], []
    for _ in range(steps):
        t = time.perf_counter()
        status = scr.make_sgp4(gpes, jd0, jd0 + args.days)
        wall.append(time.perf_counter() - t)
        kt.append(scr.sgp4_times())
    dt = sum(wall) / steps
    info = scr.pj_info()
    stage = {k: sum(d[k] for d in kt) / steps for k in kt[0]}
    n_steps = int(info["n_steps"])
    # a sample of 32 satellites, copied out and evaluated at 16 random times,
    # against SGP4 (the GPU propagator here; the reference's in the CPU leg)
    rng = np.random.default_rng(args.seed)
    smp = np.sort(rng.choice(N, 32, replace=False))
    pj = mz.make_sgp4_polyjectory(gpes[smp], jd0, jd0 + args.days)[0]
    tm = np.sort(rng.uniform(0.0, args.days * 0.999, 16))
    ev = pj.state_meval(tm)

    def sample_err(prop):
        worst_p = worst_v = 0.0
        for k, i in enumerate(smp):
            if pj.status[k] != 0:
                continue
            e_off = (gpes["epoch_jd1"][i] - jd0) + gpes["epoch_jd2"][i]
            ref, err = prop(i, (tm - e_off) * 1440.0)
            ok = err == 0
            worst_p = max(worst_p, float(np.abs(ev[k, ok, :3] - ref[ok, :3]).max()))
            worst_v = max(worst_v, float(np.abs(ev[k, ok, 3:6] - ref[ok, 3:]).max()))
        return {"position_km": worst_p, "velocity_km_s": worst_v}

    def gpu_prop(i, ts):
        ref, err = mz.sgp4_propagate(gpes[i : i + 1], ts)
        return ref[0], err[0]

    flops = sgp4_flops_per_launch()
    kc = stage["count"]
    out = {
        "metric": "SGP4 polyjectory generation throughput (make_sgp4_polyjectory)",
        "value": round(N * args.days / dt, 1),
        "unit": "satellite-days/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded near-earth GPE catalog, one GPE per satellite)",
        "config": {"workload": f"{N} near-earth satellites x {args.days:g} day, order 8, data left in HBM",
                   "n_sats": N, "days": args.days, "seed": args.seed},
        "trajectory_steps": n_steps,
        "satellites_with_data": int(np.sum(status == 0)),
        "stage_ms": {k: round(v * 1e3, 3) for k, v in stage.items()},
        "sample_max_err_vs_gpu_sgp4": sample_err(gpu_prop),
        "roofline": {
            "kernel": "k_sgp4_count",
            "bound": "valu-fp64",
            "achieved": None if flops is None else round(flops / kc / 1e12, 2),
            "peak": 78.6,
            "unit": "TFLOP/s",
            "frac": None if flops is None else round(flops / kc / 1e12 / 78.6, 4),
            "traffic": None,
            "avg_launch_ms": round(kc * 1e3, 3),
        },
    }
    lib = None if args.no_cpu_baseline else ref_sgp4_lib()
    if lib is not None:
        # the checker: the sample against the reference's SGP4
        out["sample_max_err_vs_reference_sgp4"] = sample_err(lambda i, ts: ref_sgp4(lib, gpes[i], ts)[1:])
        # the reference's SGP4 on one core, the evaluations a CPU generator with
        # the same steps would make (19 per step: the step's beginning, 9 nodes
        # and 9 error points), on a sample of satellites
        per_sat = n_steps / max(int(np.sum(status == 0)), 1)
        n_s = 1000
        n_eval = int(per_sat * 19)
        t = time.perf_counter()
        for i in range(n_s):
            ref_sgp4(lib, gpes[i], np.linspace(0.0, args.days * 1440.0, n_eval))
        dt_cpu = time.perf_counter() - t
        out["cpu_baseline"] = {"value": round(n_s * args.days / dt_cpu, 1), "unit": "satellite-days/s", "cores": 1,
                               "kind": "reference",
                               "sample": f"{n_s} satellites x {n_eval} SGP4 evaluations of the reference's vendored "
                                         f"SGP4 (19 per step), interpolation not counted, wall {dt_cpu:.2f}s"}
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n-objs", type=int, default=30000)
    ap.add_argument("--days", type=float, default=1.0)
    ap.add_argument("--seed", type=int, default=2025)
    ap.add_argument("--cpu-sample-steps", type=int, default=0, help="conjunction steps timed on the CPU (0: all)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration-mib", type=int, default=0, help="only run the PMC copy calibration")
    ap.add_argument("--workload", choices=("headline", "batch64", "state_eval", "sgp4", "config2", "config4"),
                    default="headline")
    ap.add_argument("--catalog", choices=("kepler", "sgp4"), default="sgp4",
                    help="headline catalog: J2-secular Keplerian fits, or SGP4 mean elements through make_sgp4_polyjectory")
    args = ap.parse_args()

    if args.gpus != 1:
        print(
            f"bench.py: --gpus {args.gpus} is not supported: the screening runs on one GPU "
            "(conjunction steps are independent, so several GPUs would each take a contiguous range of them)",
            file=sys.stderr,
        )
        sys.exit(2)

    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
        return
    if args.workload == "batch64":
        batch_workload(args)
        return
    if args.workload == "sgp4":
        sgp4_workload(args)
        return
    if args.workload == "state_eval":
        state_eval_workload(args)
        return
    if args.workload in LARGE_CONFIGS:
        large_config_workload(args)
        return

    import mizuba_amd as mz

    thresh, interval = 5.0, 3.0 / 1440.0
    t0 = time.perf_counter()
    if args.catalog == "sgp4":
        jd0 = 2460496.5
        gpes = mz.make_synthetic_gpes(args.n_objs, seed=args.seed, frac_high=0.2, epoch_jd=jd0)
        pj = mz.make_sgp4_polyjectory(gpes, jd0, jd0 + args.days)[0]
    else:
        pj = mz.make_synthetic_polyjectory(args.n_objs, args.days, seed=args.seed, frac_high=0.2, order=8)
    print(f"[bench] synthetic polyjectory: {pj.n_objs} objects in {time.perf_counter() - t0:.1f}s", file=sys.stderr)

    scr = mz.Screener(0)
    scr.upload(pj)
    for _ in range(args.warmup):
        scr.screen(thresh, interval, mode="mixed")
        scr.screen(thresh, interval, mode="fp64")
    print("[bench] warmup done", file=sys.stderr, flush=True)

    def timed(mode):
        stages = []
        scr.screen(thresh, interval, mode=mode)  # prime buffers for this mode
        t = time.perf_counter()
        for _ in range(args.steps):
            scr.screen(thresh, interval, mode=mode)
            stages.append(scr.stage_times())
        dt = (time.perf_counter() - t) / args.steps
        return dt, stages

    dt_fp64, st_fp64 = timed("fp64")
    conj_fp64 = scr.conjunctions()
    pairs_fp64 = scr.n_pairs()
    np_fp64 = scr.np_report()
    dt_mixed, st_mixed = timed("mixed")
    conj_mixed = scr.conjunctions()
    pairs_mixed = scr.n_pairs()
    np_mixed = scr.np_report()
    print(f"[bench] fp64 {dt_fp64*1e3:.2f} ms/step, mixed {dt_mixed*1e3:.2f} ms/step", file=sys.stderr, flush=True)

    same_ij = len(conj_fp64) == len(conj_mixed) and bool(
        np.array_equal(conj_fp64["i"], conj_mixed["i"]) and np.array_equal(conj_fp64["j"], conj_mixed["j"])
    )
    same_bits = same_ij and bool(np.array_equal(conj_fp64.view(np.uint8), conj_mixed.view(np.uint8)))

    obj_days = pj.n_objs * pj.maxT
    value = obj_days / dt_mixed
    value_fp64 = obj_days / dt_fp64

    def avg(stages, key):
        return float(np.mean([s[key] for s in stages]))

    # per-kernel device times (HIP events on the library's stream) and rooflines
    stage_keys = ("aabbs", "morton", "bvh", "broad", "narrow")
    kern_keys = ("aabb_kernel", "narrow_kernel", "broad_kernel")
    st_avg = {k: avg(st_mixed, k) for k in stage_keys + ("total",) + kern_keys}
    st_avg_fp64 = {k: avg(st_fp64, k) for k in stage_keys + ("total",) + kern_keys}
    C = int(np.ceil(pj.maxT / interval))
    aabb_bytes, kbar = aabb_bytes_per_launch(pj, interval)
    aabb_bytes_f32, _ = aabb_bytes_per_launch(pj, interval, coef_bytes=4)
    kernels = {
        "k_aabbs<8, false>": (st_avg_fp64["aabb_kernel"], aabb_bytes),
        "k_aabbs<8, true>": (st_avg["aabb_kernel"], aabb_bytes_f32),
        "k_lbvh_build+k_lbvh_top": (st_avg_fp64["bvh"], bvh_bytes_per_launch(pj.n_objs, C)),
        "k_bp": (st_avg_fp64["broad_kernel"], broad_bytes_per_launch(pj.n_objs, C, pairs_fp64)),
    }
    per_kernel = {
        name: {
            "avg_launch_ms": round(t * 1e3, 4),
            "algorithmic_bytes": b,
            "achieved_gbs": round(b / t / 1e9, 1),
            "frac_hbm": round(b / t / 1e9 / HBM_PEAK_GBS, 4),
        }
        for name, (t, b) in kernels.items()
    }
    dom = max(kernels, key=lambda k: kernels[k][0])
    t_dom, b_dom = kernels[dom]
    gbs = b_dom / t_dom / 1e9
    roofline = {
        "kernel": dom,
        "bound": "hbm",
        "achieved": round(gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(gbs / HBM_PEAK_GBS, 4),
        "traffic": hbm_traffic(dom),
        "algorithmic_bytes": b_dom,
        "avg_launch_ms": round(t_dom * 1e3, 4),
        "kbar": round(kbar, 4),
        "kernels": per_kernel,
    }

    # End to end through the reference's interface: a conjunctions object
    # made from the polyjectory (its own context: the polyjectory copied into
    # it, the mixed screening) and every exposed array read to host memory.
    # Not the headline value, which keeps inputs and outputs in HBM.
    del scr
    t = time.perf_counter()
    cj = mz.conjunctions(pj, thresh, interval, mode="mixed")
    t_made = time.perf_counter()
    nbytes = 0
    for name in ("aabbs", "srt_aabbs", "mcodes", "srt_mcodes", "srt_idx", "cd_end_times", "conjunctions"):
        nbytes += getattr(cj, name).nbytes
    for k in range(cj.n_cd_steps):
        nbytes += cj.get_bvh_tree(k).nbytes + cj.get_aabb_collisions(k).nbytes
    dt_e2e = time.perf_counter() - t
    same_e2e = bool(np.array_equal(cj.conjunctions.view(np.uint8), conj_mixed.view(np.uint8)))
    del cj
    with_d2h = {
        "value": round(obj_days / dt_e2e, 1),
        "unit": "object-days/s",
        "ms": round(dt_e2e * 1e3, 3),
        "construct_ms": round((t_made - t) * 1e3, 3),
        "bytes_to_host": int(nbytes),
        "identical_conj_bits": same_e2e,
        "includes": "conjunctions(pj, ...) in mixed mode: polyjectory copy into the object's context, screening, "
        "and reading every exposed array (boxes, codes, sort order, all steps' BVH trees and pairs, conjunctions)",
    }

    cpu = None
    oracle_check = None
    if not args.no_cpu_baseline:
        from oracle import pyoracle as po

        traj, tm, toff, ns, tmo = pj._flat()
        flat = po.Flat.from_arrays(pj.poly_order, traj, tm, ns, toff, tmo, pj.maxT)
        threads = po.n_threads_default()
        nsamp = C if args.cpu_sample_steps <= 0 else min(args.cpu_sample_steps, C)
        t = time.perf_counter()
        ref = po.run(flat, thresh, interval, step_end=nsamp, n_threads=threads)
        dt_cpu = time.perf_counter() - t
        sample_days = min(nsamp * interval, pj.maxT)
        cpu = {
            "value": round(pj.n_objs * sample_days / dt_cpu, 2),
            "unit": "object-days/s",
            "cores": threads,
            "kind": "port",
            "cpu": cpu_model(),
            "sample": f"first {nsamp} of {C} conjunction steps ({sample_days:.4f} days) of the same "
            f"{pj.n_objs}-object polyjectory, wall {dt_cpu:.2f}s; the oracle runs its steps in parallel "
            "(OpenMP over steps, each step's objects and pairs in one thread)",
        }
        # TCA/DCA error of the mixed GPU pipeline against the FP64 oracle on the sample
        t_end = ref["cd_end_times"][-1]
        g = conj_mixed[conj_mixed["tca"] < t_end]
        r = ref["conjunctions"]
        ok = len(g) == len(r) and np.array_equal(g["i"], r["i"]) and np.array_equal(g["j"], r["j"])
        oracle_check = {
            "steps": nsamp,
            "identical_ij": bool(ok),
            "n_conj": int(len(r)),
            "max_dtca_s": float(np.max(np.abs(g["tca"] - r["tca"])) * 86400.0) if ok and len(r) else None,
            "max_ddca_km": float(np.max(np.abs(g["dca"] - r["dca"]))) if ok and len(r) else None,
            "bitwise": bool(ok and np.array_equal(g.view(np.uint8), r.view(np.uint8))),
        }

    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "object-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt_mixed * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 screen + f64 narrow phase",
        "data": ("synthetic (seeded SGP4 mean elements, polyjectory from this library's GPU make_sgp4_polyjectory, "
                 "order 8)" if args.catalog == "sgp4" else
                 "synthetic (seeded J2-secular Keplerian catalog, order-8 Chebyshev-fitted polyjectory)"),
        "config": {
            "workload": f"{pj.n_objs} objects x {args.days:g} day, order 8, conj_thresh 5 km, conj_det_interval 3 min",
            "n_objs": pj.n_objs,
            "days": args.days,
            "n_cd_steps": int(np.ceil(pj.maxT / interval)),
            "mix": ("80% LEO (perigee 400-1200 km, e < 0.01) / 20% perigee 1200-2000 km, e < 0.05; "
                    "inclination 0-100 deg; bstar 1e-5-1e-4"),
            "catalog": args.catalog,
            "trajectory_steps": int(pj._flat()[3].sum()),
            "seed": args.seed,
        },
        "fp64_dropin": {
            "value": round(value_fp64, 1),
            "ms_per_step": round(dt_fp64 * 1e3, 3),
            "stages_ms": {k: round(v * 1e3, 3) for k, v in st_avg_fp64.items()},
            "n_pairs": int(pairs_fp64),
            "np_report": np_fp64,
        },
        "mixed": {
            "stages_ms": {k: round(v * 1e3, 3) for k, v in st_avg.items()},
            "n_pairs": int(pairs_mixed),
            "np_report": np_mixed,
        },
        "mixed_over_fp64": round(dt_fp64 / dt_mixed, 3),
        "identical_conj_list": same_ij,
        "identical_conj_bits": same_bits,
        "n_conj": int(len(conj_mixed)),
        "roofline": roofline,
        "with_d2h": with_d2h,
        "cpu_baseline": cpu,
        "oracle_check": oracle_check,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
