#!/usr/bin/env python3
"""Step-throughput bench of the device step library (one MI355X).

Workload: the plain method on GPU, one world of 16,384 bodies (1/3 spheres, 1/3 boxes, 1/3
capsules) spawned over a 256x256, 0.5 m fBm heightfield; dt = 1/60, 10 velocity and 20 position
iterations, mu = 0.5, e = 0.  A step is one tick of the whole step (broadphase, narrowphase,
contact solve, integration, projection, sleep) through the C ABI; the scene is resident in HBM
before the timed region starts.

Prints one JSON line: body-steps/s over the timed ticks, plus
  roofline     - the dominant kernel's algorithmic bytes per launch over its mean launch time,
                 both measured live with HIP events on the library's stream during the timed ticks;
  cpu_baseline - the oracle (single-threaded C restatement of the reference step) timed on this
                 host on a bounded, stratified sample of the same ticks: a few oracle ticks
                 continued from the device state at the start of every 100-tick chunk of the timed
                 region, each chunk's CPU time extrapolated from its sample.
Other workloads (--workload): batch (1,024 worlds, rollback windows), learned (the learned warm
start against the previous-frame and zero ones, batched and plain, to residual 1e-4), towers (8
worlds of 16,384 bodies with towers and ball-jointed capsule chains), boxes512 (the box lattice).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "riftphys-rust_amd"))

import rphys  # noqa: E402
import scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure

# Algorithmic bytes per unit of each single-kernel stage: every input read once and every output
# written once (f32 SoA constants).
#   velocity solve : per island contact, Iv row-iterations x 260 B (constraint 140, two bodies'
#                    velocities read+written 96, accumulators 24) + one shock row 230 B
#   projection     : per island contact, Ip x 230 B (PosContact 60, two poses read+written 112,
#                    inverse mass/inertia 56); its shock pass runs in a launch of its own
#   body-pair narrowphase : per pair, 88 B (two poses + shapes)
#   terrain narrowphase   : per (collider, triangle) candidate, 16 B (height samples)
#   ball joints (an island entry each, in the small islands at these sizes): per joint, 3 rows x
#                    (Iv x 96 B velocity, Ip x 120 B position)


def stage_bytes(stage, ctr, iv, ip):
    c_large = ctr["island_contacts_global"]  # islands of 4,097..131,072 contacts
    j = ctr.get("joint_entries", 0)
    c_small = ctr["island_contacts"] - c_large - j  # (islands above 131,072 contacts: none at these sizes)
    table = {
        "vel_solve_small": (iv * 260 + 230) * c_small + iv * 96 * 3 * j,
        "vel_solve_large": (iv * 260 + 230) * c_large,
        "projection_small": ip * 230 * c_small + ip * 120 * 3 * j,
        "projection_large": ip * 230 * c_large,
        "narrow_pairs": 88 * ctr["pairs"],
        "narrow_terrain": 16 * ctr["terrain_candidates"],
    }
    return table.get(stage)


# The kernels of each stage.  The small islands' velocity solve and projection run as concurrent
# launches (the lane tier on one stream, tier 0's largest islands and tier 3 on another), and a
# big island's as one launch sequence per tick (a setup kernel, kl wide-level launches and a tail
# sweep per iteration, and for the velocity solve a shock kernel), so the roofline unit of these
# stages is the stage, per tick.
KERNEL_OF_STAGE = {
    "vel_solve_small": "k_lane_vel + k_sweep_vel<0> + k_sweep_vel<3>",
    "vel_solve_large": "k_sweep_vel<1> + k_level_vel + k_sweep_vel_tail<1> + k_sweep_vel_post<1>",
    "projection_small": "k_lane_proj + k_sweep_proj<0> + k_sweep_proj<3>",
    "projection_large": "k_sweep_proj<1> + k_level_proj + k_sweep_proj_tail<1>",
    "narrow_pairs": "k_narrow_pairs",
    "narrow_terrain": "k_narrow_terrain",
}


def measured_traffic(stage, workload):
    """HBM bytes per tick of a stage from the committed rocprofv3 PMC summary, when one exists."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) >= 3 and f[0] == workload and f[1] == stage:
            try:
                return float(f[2])
            except ValueError:
                return None
    return None


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, strata, ticks_per_stratum, dt):
    """strata: (device state, chunk length) per chunk of the timed region.  Runs the oracle for a
    few ticks from each state and extrapolates each chunk's CPU time from its mean tick time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py

    n = len(strata[0][0]["pos"])
    est, sampled, spent = 0.0, 0, 0.0
    for state, length in strata:
        o = oracle_py.OracleWorld(scene)
        o.load_state(state)
        t0 = time.perf_counter()
        for _ in range(ticks_per_stratum):
            o.step(dt)
        el = time.perf_counter() - t0
        est += el / ticks_per_stratum * length
        sampled += ticks_per_stratum
        spent += el
        del o
    total_ticks = sum(length for _, length in strata)
    return {
        "value": n * total_ticks / est,
        "unit": "body-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": "%d oracle ticks: %d continued from the device state at the start of each of the %d chunks of the timed ticks, "
        "each chunk's time extrapolated from its sample (%.1f s measured, %.0f s estimated for all %d ticks; %s, %d host cores, 1 used)"
        % (sampled, ticks_per_stratum, len(strata), spent, est, total_ticks, cpu_model(), os.cpu_count() or 0),
    }


def merge_profile(stages, ctr, st, c):
    for key, v in st.items():
        stages[key] = stages.get(key, 0.0) + v
    for key, v in c.items():
        if isinstance(v, dict):
            acc = ctr.setdefault(key, {})
            for k2, v2 in v.items():
                acc[k2] = acc.get(k2, 0.0) + v2
        else:
            ctr[key] = ctr.get(key, 0) + v


def roofline_of(stages, ctr, iv, ip, workload):
    cand = {k: v for k, v in stages.items() if k in KERNEL_OF_STAGE}
    dom = max(cand, key=cand.get)
    avg_ms = cand[dom] / max(ctr["ticks"], 1)
    bytes_launch = stage_bytes(dom, ctr, iv, ip) / max(ctr["ticks"], 1)
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dom, workload), "kernel": KERNEL_OF_STAGE[dom], "stage": dom,
            "algorithmic_bytes_per_launch": bytes_launch, "mean_launch_ms": avg_ms,
            "launch": "one tick of the stage" if "+" in KERNEL_OF_STAGE[dom] else "one kernel launch"}


def print_stages(stages, ctr):
    tot = sum(stages.values())
    print("stage split over %d ticks (ms/tick):" % ctr["ticks"], file=sys.stderr)
    for k, v in stages.items():
        print("  %-18s %8.3f  %5.1f%%" % (k, v / ctr["ticks"], 100.0 * v / tot), file=sys.stderr)
    print("counters per tick: " + ", ".join("%s %.0f" % (k, v / ctr["ticks"]) for k, v in ctr.items()
                                             if k not in ("ticks", "sweep_phase")), file=sys.stderr)
    print("large-island sweep phases per tick (ms; levels): " + ", ".join(
        "%s %.3f" % (k, v / ctr["ticks"]) for k, v in ctr["sweep_phase"].items()), file=sys.stderr)


def batch_bench(a):
    """Batched workload: many independent worlds in one launch sequence, each replaying 8-tick
    rollback windows from a device snapshot (snapshot, 8 ticks, restore, the same 8 ticks)."""
    t0 = time.perf_counter()
    worlds = [scenes.batch_world(w) for w in range(a.worlds)]
    handles = [rphys.World(s) for s in worlds]
    b = rphys.Batch(handles)
    del handles
    nb = b.total_bodies()
    dt = float(worlds[0].get("dt", 1.0 / 60.0))
    iv, ip = worlds[0].get("iters", (10, 20))
    wl = {"workload": "batch%d_rollback8" % a.worlds, "worlds": a.worlds, "bodies": nb,
          "bodies_per_world": "round(exp(U(ln 32, ln 2048))), mixed shapes over per-world fBm terrain",
          "rollback_window": 8, "dt": "1/60", "vel_iters": iv, "pos_iters": ip, "graphs": not a.no_graph}
    print("bench: %s, %d bodies, built in %.1f s; warmup %d ticks" % (wl["workload"], nb, time.perf_counter() - t0, a.warmup),
          file=sys.stderr, flush=True)
    if a.warmup > 0:
        b.step_n(a.warmup, dt, graph=not a.no_graph)
    b.hashes()
    # launch shapes of the island sweeps: cached per (workload, size), else a grid search on the
    # device from this state (untimed; the batch is left where it was)
    tuning = None
    if not a.no_graph:
        t1 = time.perf_counter()
        entry, cached = b.autotune(wl["workload"], ticks=2, dt=dt)
        tuning = dict(entry, cached=cached, search_s=time.perf_counter() - t1)
        print("bench: tuning %s (%s, %.1f s)" % (b.get_tuning(), "cached" if cached else "searched", tuning["search_s"]), file=sys.stderr, flush=True)
    wl["tuning"] = tuning
    b.snapshot(1)  # start of the timed windows (both passes below start here)
    windows = max(1, a.steps // 16)
    # timed: graph-captured ticks, each window snapshot -> 8 ticks -> restore -> the same 8 ticks
    el = 0.0
    for k in range(windows):
        t1 = time.perf_counter()
        b.snapshot(0)
        b.step_n(8, dt, graph=not a.no_graph)
        b.restore(0)
        b.step_n(8, dt, graph=not a.no_graph)
        el += time.perf_counter() - t1
        if (k + 1) % 5 == 0 or k + 1 == windows:
            print("bench: %d/%d windows, %.2f s" % (k + 1, windows, el), file=sys.stderr, flush=True)
    ticks = 16 * windows
    value = nb * ticks / el
    # the same windows again with HIP events at every stage boundary (roofline, stage split)
    b.restore(1)
    stages, ctr = {}, {}
    el_prof = 0.0
    for k in range(windows):
        t1 = time.perf_counter()
        b.snapshot(0)
        merge_profile(stages, ctr, *b.step_profiled(8, dt))
        b.restore(0)
        merge_profile(stages, ctr, *b.step_profiled(8, dt))
        el_prof += time.perf_counter() - t1
    print("bench: profiled pass %.2f s" % el_prof, file=sys.stderr, flush=True)
    # rollback check (untimed hashes)
    b.snapshot(0)
    b.step_n(8, dt, graph=not a.no_graph)
    h1 = b.hashes()
    b.restore(0)
    b.step_n(8, dt, graph=not a.no_graph)
    replay_equal = h1 == b.hashes()
    if a.stages:
        print_stages(stages, ctr)
    base = None
    if not a.no_cpu:
        print("bench: oracle sample over %d worlds" % len(range(0, a.worlds, max(1, a.worlds // 8))), file=sys.stderr, flush=True)
        base = batch_cpu_baseline(worlds, a.warmup, ticks, dt)
    wl.update({"contacts_per_tick": ctr["contacts"] / ctr["ticks"], "ticks": "%d..%d, each window stepped twice" % (a.warmup, a.warmup + ticks // 2),
               "profiled_pass_body_steps_per_s": nb * ticks / el_prof, "replayed_hashes_equal": replay_equal})
    return {"metric": "body-steps/s on 1 GPU (plain method, batched worlds with rollback windows)", "value": value,
            "unit": "body-steps/s", "n_gpus": 1, "steps": ticks, "warmup": a.warmup, "ms_per_step": 1000.0 * el / ticks,
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
            "data": "synthetic (seeded XorShift64 scenes)", "config": wl, "roofline": roofline_of(stages, ctr, iv, ip, wl["workload"]),
            "cpu_baseline": base}


def batch_cpu_baseline(worlds, warmup, ticks, dt):
    """The oracle on every (n/8)-th world: its trajectory from spawn through the timed ticks (a
    rollback window's replay repeats the same kind of tick), timing the ticks after warmup (the
    distinct ticks of the timed windows, capped at 40 per world to bound the sample)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py

    sample = worlds[:: max(1, len(worlds) // 8)]
    span = min(40, ticks // 2)
    bodies, t_spent = 0, 0.0
    for s in sample:
        o = oracle_py.OracleWorld(s)
        for _ in range(warmup):
            o.step(dt)
        t0 = time.perf_counter()
        for _ in range(span):
            o.step(dt)
        t_spent += time.perf_counter() - t0
        bodies += len(s["shape"]) * span
        del o
    return {"value": bodies / t_spent, "unit": "body-steps/s", "cores": 1, "kind": "port",
            "sample": "%d of %d worlds (every %d-th), %d ticks each after %d warmup ticks (%.1f s; %s, %d host cores, 1 used)"
            % (len(sample), len(worlds), max(1, len(worlds) // 8), span, warmup, t_spent, cpu_model(), os.cpu_count() or 0)}


def towers_bench(a):
    """One GPU's share of the at-scale workload: `--worlds` (default 8 of 64) worlds of 16,384 bodies,
    669 towers of 20 boxes, 4 loose boxes and 200 chains of 15 capsules joined by 14 ball joints
    each, over fBm terrain, velocity iterations 10, position iterations 2.  Timed: ticks with the
    fixed iterations, launched on the library's streams (`--graph`: graph-captured and autotuned); then, from one snapshot, the PGS run to residual 1e-4 per island
    with the previous-frame, the zero and (with --learned) the learned warm start, the model trained
    on the solver's converged impulses logged from 4 other worlds of the same generator; the oracle
    on one world for the CPU baseline."""
    t0 = time.perf_counter()
    worlds = [scenes.towers_ragdolls(seed=scenes.seed_for(5, w)) for w in range(a.worlds)]
    b = rphys.Batch([rphys.World(s_) for s_ in worlds])
    nb = b.total_bodies()
    dt = float(worlds[0].get("dt", 1.0 / 60.0))
    iv, ip = worlds[0].get("iters", (10, 2))
    nj = sum(len(s_["joints"]) for s_ in worlds)
    wl = {"workload": "towers_chains%d_x16k" % a.worlds, "worlds": a.worlds, "bodies": nb, "ball_joints": nj,
          "world": "669 towers x 20 boxes (he 0.25), 4 loose boxes, 200 chains x 15 capsules (r 0.1, hh 0.15), 256x256 fBm",
          "dt": "1/60", "vel_iters": iv, "pos_iters": ip, "graphs": not a.no_graph}
    print("bench: %s, %d bodies, %d joints, built in %.1f s; warmup %d ticks" % (wl["workload"], nb, nj, time.perf_counter() - t0, a.warmup),
          file=sys.stderr, flush=True)
    for k in range(0, a.warmup, 30):
        b.step_n(min(30, a.warmup - k), dt, graph=not a.no_graph)
        b.hashes()
        print("bench: warmup %d/%d ticks" % (min(k + 30, a.warmup), a.warmup), file=sys.stderr, flush=True)
    tuning = None
    if not a.no_graph:  # launch shapes of the island sweeps, as for the batched workload
        t1 = time.perf_counter()
        entry, cached = b.autotune(wl["workload"], ticks=2, dt=dt)
        tuning = dict(entry, cached=cached, search_s=time.perf_counter() - t1)
        print("bench: tuning %s (%s, %.1f s)" % (b.get_tuning(), "cached" if cached else "searched", tuning["search_s"]), file=sys.stderr, flush=True)
    wl["tuning"] = tuning
    b.snapshot(1)
    el, done = 0.0, 0
    while done < a.steps:
        k = min(20, a.steps - done)
        t1 = time.perf_counter()
        b.step_n(k, dt, graph=not a.no_graph)
        b.hashes()
        el += time.perf_counter() - t1
        done += k
        print("bench: %d/%d ticks, %.2f s" % (done, a.steps, el), file=sys.stderr, flush=True)
    value = nb * a.steps / el
    b.restore(1)
    stages, ctr = {}, {}
    for k in range(0, min(a.steps, 40), 10):
        merge_profile(stages, ctr, *b.step_profiled(10, dt))
    ctr["joint_entries"] = nj * ctr["ticks"]  # every joint of these worlds is active and has its 3 rows
    if a.stages:
        print_stages(stages, ctr)
    # PGS to residual 1e-4 from the same state: previous-frame vs zero (vs learned) warm start
    tol, max_it = 1.0e-4, a.max_sweeps
    params, training = None, None
    if a.learned:
        t1 = time.perf_counter()
        tw = [scenes.towers_ragdolls(seed=scenes.seed_for(5, 100 + w)) for w in range(4)]
        tb = rphys.Batch([rphys.World(s_) for s_ in tw])
        for k in range(0, a.warmup, 30):
            tb.step_n(min(30, a.warmup - k), dt, graph=not a.no_graph)
            tb.hashes()
        tb.set_residual_mode(True, tol, max_it)
        tb.set_contact_log(a.train_rows + 400000)
        rows = 0
        while rows < a.train_rows:
            tb.step_n(2, dt, graph=not a.no_graph)
            rows = tb.contact_log_rows()[0]
            print("bench: training data %d rows, %.1f s" % (rows, time.perf_counter() - t1), file=sys.stderr, flush=True)
        t2 = time.perf_counter()
        params, losses = tb.train_warm_start(epochs=a.epochs, lr=1.0e-3, seed=20260101)
        training = {"worlds": "4 towers-and-chains worlds (seeds of worlds 100..103), ticks %d on" % a.warmup, "rows": rows,
                    "epochs": a.epochs, "lr": 1e-3, "loss_first_last": [float(losses[0]), float(losses[-1])],
                    "data_s": t2 - t1, "train_s": time.perf_counter() - t2}
        print("bench: trained %d epochs in %.1f s, loss %.4f -> %.4f" % (a.epochs, training["train_s"], losses[0], losses[-1]), file=sys.stderr, flush=True)
        del tb
    b.set_residual_mode(True, tol, max_it)
    resid = {}
    for name, mode in (("previous_frame", 0), ("zero", 1)) + ((("learned", 2),) if params is not None else ()):
        b.restore(1)
        b.set_warm_start(mode, params if mode == 2 else None)
        el2, isum, ic, conv, cts, lrn = 0.0, 0.0, 0.0, 0.0, 0.0, 0.0
        for k in range(0, 20, 10):
            t1 = time.perf_counter()
            _, c = b.step_profiled(10, dt)
            el2 += time.perf_counter() - t1
            isum += c["pgs_iters_sum"]; ic += c["island_contacts"]; conv += c["island_contacts_converged"]; cts += c["contacts"]
            lrn += c["warm_start_learned"]
        resid[name] = {"body_steps_per_s": nb * 20 / el2, "mean_sweeps_per_entry": isum / max(ic, 1), "converged_entry_frac": conv / max(ic, 1),
                       "contacts_per_tick": cts / 20, "learned_contact_frac": lrn / max(cts, 1)}
        print("bench: residual 1e-4, %s warm start: %.0f body-steps/s, %.2f sweeps per island entry, %.3f converged, %.3f learned" % (
            name, resid[name]["body_steps_per_s"], resid[name]["mean_sweeps_per_entry"], resid[name]["converged_entry_frac"],
            resid[name]["learned_contact_frac"]), file=sys.stderr, flush=True)
    b.set_residual_mode(False)
    b.set_warm_start(0)
    wl.update({"contacts_per_tick": ctr["contacts"] / ctr["ticks"], "ticks": "%d..%d" % (a.warmup, a.warmup + a.steps),
               "residual_1e-4": resid, "max_sweeps": max_it, "training": training})
    base = None
    if not a.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracle_py
        o = oracle_py.OracleWorld(worlds[0])
        t1 = time.perf_counter()
        for k in range(a.warmup):
            o.step(dt)
            if k % 30 == 29:
                print("bench: oracle warmup %d/%d ticks, %.1f s" % (k + 1, a.warmup, time.perf_counter() - t1), file=sys.stderr, flush=True)
        span = max(1, a.cpu_ticks)
        t1 = time.perf_counter()
        for _ in range(span):
            o.step(dt)
        t_spent = time.perf_counter() - t1
        base = {"value": len(worlds[0]["shape"]) * span / t_spent, "unit": "body-steps/s", "cores": 1, "kind": "port",
                "sample": "world 0 (16,384 bodies), ticks %d..%d after %d warmup ticks (%.1f s; %s, %d host cores, 1 used)"
                % (a.warmup, a.warmup + span, a.warmup, t_spent, cpu_model(), os.cpu_count() or 0)}
    return {"metric": "body-steps/s on 1 GPU (at-scale share: towers and jointed chains)", "value": value, "unit": "body-steps/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": 1000.0 * el / a.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded XorShift64 scenes)", "config": wl,
            "roofline": roofline_of(stages, ctr, iv, ip, wl["workload"]), "cpu_baseline": base}


def residual_pass(step_profiled, nb, ticks, chunk=10):
    """Runs `ticks` residual-mode ticks through a profiled stepper; throughput and sweep counts."""
    el, acc, done = 0.0, {}, 0
    stages = {}
    while done < ticks:
        k = min(chunk, ticks - done)
        t1 = time.perf_counter()
        st, c = step_profiled(k)
        el += time.perf_counter() - t1
        merge_profile(stages, acc, st, c)
        done += k
    ic = max(acc["island_contacts"], 1.0)
    icc = acc["island_contacts_converged"]
    return {"body_steps_per_s": nb * ticks / el, "ms_per_tick": 1000.0 * el / ticks, "mean_sweeps_per_entry": acc["pgs_iters_sum"] / ic,
            "converged_entry_frac": icc / ic, "mean_sweeps_converged_islands": acc["pgs_iters_sum_converged"] / icc if icc else None,
            "islands_converged_frac": acc["islands_converged"] / max(acc["islands"], 1.0), "contacts_per_tick": acc["contacts"] / ticks,
            "learned_contact_frac": acc["warm_start_learned"] / max(acc["contacts"], 1.0)}, stages, acc


def atscale_bench(a):
    """The headline comparison: learned + batched vs plain, PGS to residual 1e-4.

    `--worlds` (default 64) worlds of 16,384 bodies each: 669 towers of 20 boxes (he 0.25), 4 loose
    boxes and 200 chains of 15 capsules joined by 14 ball joints, over 256x256 fBm terrain; dt 1/60,
    velocity iterations 10 (the sweep limit is `--max-sweeps` in residual mode), position
    iterations 2.  Stepped untimed to tick 120 (settled towers, chains at rest) with the fixed
    iterations, then from one snapshot:
      learned + batched : all worlds in one launch sequence per tick, every island iterated to
                          residual 1e-4, the learned warm start (trained first, on the device, on
                          the converged impulses the solver logs for 4 other worlds of the same
                          generator, islands that reached the tolerance only);
      batched           : the same with the previous-frame (the reference's) and zero warm starts;
      plain             : the same worlds one world at a time (one launch sequence per world per
                          tick), previous-frame warm start, to the same residual.
    value = learned + batched body-steps/s.  The plain leg's per-world hashes must equal the
    batched previous-frame leg's."""
    tol, max_it = 1.0e-4, a.max_sweeps
    t0 = time.perf_counter()
    worlds = [scenes.towers_ragdolls(seed=scenes.seed_for(5, w)) for w in range(a.worlds)]
    b = rphys.Batch([rphys.World(s_) for s_ in worlds])
    nb = b.total_bodies()
    dt = float(worlds[0].get("dt", 1.0 / 60.0))
    iv, ip = worlds[0].get("iters", (10, 2))
    nj = sum(len(s_["joints"]) for s_ in worlds)
    settle = max(a.warmup, 120)
    print("bench: atscale, %d worlds, %d bodies, %d ball joints, built in %.1f s; stepping to tick %d" % (
        a.worlds, nb, nj, time.perf_counter() - t0, settle), file=sys.stderr, flush=True)
    for k in range(0, settle, 30):
        b.step_n(min(30, settle - k), dt)
        b.hashes()
        print("bench: settle %d/%d ticks, %.1f s" % (min(k + 30, settle), settle, time.perf_counter() - t0), file=sys.stderr, flush=True)
    b.snapshot(1)

    # the warm-start model: converged impulses of 4 other worlds of the generator
    t1 = time.perf_counter()
    tw = [scenes.towers_ragdolls(seed=scenes.seed_for(5, 100 + w)) for w in range(4)]
    tb = rphys.Batch([rphys.World(s_) for s_ in tw])
    for k in range(0, settle, 30):
        tb.step_n(min(30, settle - k), dt)
        tb.hashes()
    tb.set_residual_mode(True, tol, max_it)
    tb.set_contact_log(a.train_rows + 400000)
    rows, logged = 0, 0
    while rows < a.train_rows and logged < 400:
        tb.step_n(5, dt)
        logged += 5
        rows = tb.contact_log_rows()[0]
        if logged % 25 == 0:
            print("bench: training data %d rows after %d ticks, %.1f s" % (rows, logged, time.perf_counter() - t1), file=sys.stderr, flush=True)
    t2 = time.perf_counter()
    params, losses = tb.train_warm_start(epochs=a.epochs, lr=1.0e-3, seed=20260101)
    training = {"worlds": "4 worlds of the generator (seeds of worlds 100..103), residual 1e-4, from tick %d, %d ticks logged" % (settle, logged),
                "rows": rows, "rows_from": "islands that reached the tolerance", "epochs": a.epochs, "lr": 1e-3,
                "loss_first_last": [float(losses[0]), float(losses[-1])], "data_s": t2 - t1, "train_s": time.perf_counter() - t2}
    print("bench: trained on %d rows, %d epochs in %.1f s, loss %.4f -> %.4f" % (rows, a.epochs, training["train_s"], losses[0], losses[-1]),
          file=sys.stderr, flush=True)
    del tb

    # learned + batched: timed without events (the value), then the same ticks profiled
    b.set_residual_mode(True, tol, max_it)
    b.restore(1)
    b.set_warm_start(2, params)
    done, el = 0, 0.0
    while done < a.steps:
        k = min(10, a.steps - done)
        t1 = time.perf_counter()
        b.step_n(k, dt)  # returns once the ticks have run (the library reads their counters back)
        el += time.perf_counter() - t1
        done += k
        print("bench: learned + batched %d/%d ticks, %.2f s" % (done, a.steps, el), file=sys.stderr, flush=True)
    value = nb * a.steps / el
    modes, stages_l, ctr_l = {}, None, None
    for name, mode in (("learned", 2), ("previous_frame", 0), ("zero", 1)):
        b.restore(1)
        b.set_warm_start(mode, params if mode == 2 else None)
        if mode == 2 and a.mark_window:  # a copy-kernel dispatch on each side marks the roofline's window in a kernel trace
            rphys.copy_probe(1 << 20, 1)
        modes[name], st_, c_ = residual_pass(lambda k: b.step_profiled(k, dt), nb, a.steps)
        if mode == 2 and a.mark_window:
            rphys.copy_probe(1 << 20, 1)
        if mode == 2:
            stages_l, ctr_l = st_, c_
        if mode == 0:
            batched_hashes = b.hashes()
        print("bench: batched, %s warm start: %.0f body-steps/s, %.2f sweeps per island entry, %.3f of entries converged" % (
            name, modes[name]["body_steps_per_s"], modes[name]["mean_sweeps_per_entry"], modes[name]["converged_entry_frac"]),
            file=sys.stderr, flush=True)
    ctr_l["joint_entries"] = nj * ctr_l["ticks"]

    # plain: the same worlds one at a time from the same tick, previous-frame warm start
    b.restore(1)
    b.set_residual_mode(False)
    b.set_warm_start(0)
    pbodies, pel, same = 0, 0.0, 0
    acc = {}
    for w in range(a.worlds):
        pw = b.world(w)
        pw.set_residual_mode(True, tol, max_it)
        pw.set_warm_start(0)
        pw.hash()  # device scene built, off the clock
        t1 = time.perf_counter()
        pw.step_n(a.steps, dt)
        pel += time.perf_counter() - t1
        h = pw.hash()
        pbodies += pw.num_bodies()
        same += int(h == batched_hashes[w])
        pw.close()
        if (w + 1) % 8 == 0 or w + 1 == a.worlds:
            print("bench: plain %d/%d worlds, %.2f s" % (w + 1, a.worlds, pel), file=sys.stderr, flush=True)
    plain = {"body_steps_per_s": pbodies * a.steps / pel, "ms_per_tick_all_worlds": 1000.0 * pel / a.steps,
             "hashes_equal_batched_previous_frame": "%d/%d" % (same, a.worlds)}
    if a.stages:
        print_stages(stages_l, ctr_l)
    base = None
    if not a.no_cpu:
        base = atscale_cpu_baseline(b, worlds[0], a.cpu_ticks, dt)
    wl = {"workload": "atscale%d_towers_chains_x16k" % a.worlds, "worlds": a.worlds, "bodies": nb, "ball_joints": nj,
          "world": "669 towers x 20 boxes (he 0.25), 4 loose boxes, 200 chains x 15 capsules (r 0.1, hh 0.15) with 14 ball joints, 256x256 fBm",
          "dt": "1/60", "vel_iters": iv, "pos_iters": ip, "residual": tol, "max_sweeps": max_it, "ticks": "%d..%d" % (settle, settle + a.steps),
          "launches": "stream", "learned_batched": modes["learned"], "batched_previous_frame": modes["previous_frame"],
          "batched_zero": modes["zero"], "plain": plain, "learned_batched_over_plain": value / plain["body_steps_per_s"],
          "training": training}
    return {"metric": "body-steps/s on 1 GPU, learned+batched vs plain; PGS iters to residual 1e-4", "value": value, "unit": "body-steps/s",
            "n_gpus": 1, "steps": a.steps, "warmup": settle, "ms_per_step": 1000.0 * el / a.steps, "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32 (warm-start model: bf16 weights and activations, f32 accumulation)",
            "data": "synthetic (seeded XorShift64 scenes; the model is trained in the run)", "config": wl,
            "roofline": roofline_of(stages_l, ctr_l, ctr_l["pgs_iters_sum"] / max(ctr_l["island_contacts"], 1.0), ip, wl["workload"]),
            "cpu_baseline": base}


def atscale_cpu_baseline(b, scene, ticks, dt):
    """The oracle on world 0, continued from the batch's state at the snapshot (bodies, sleep,
    warm cache), for `ticks` ticks with the fixed iterations (the reference's step)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_py

    b.restore(1)
    st = b.world(0).export_state()
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    spent, hashes = {}, {}
    for nt in (1, threads):
        oracle_py.lib().or_set_threads(nt)
        o = oracle_py.OracleWorld(scene)
        o.load_state(st)
        t1 = time.perf_counter()
        for _ in range(ticks):
            o.step(dt)
        spent[nt] = time.perf_counter() - t1
        hashes[nt] = o.hash()
        del o
        print("bench: oracle, %d thread(s): %.1f s for %d ticks" % (nt, spent[nt], ticks), file=sys.stderr, flush=True)
    oracle_py.lib().or_set_threads(1)
    n = len(scene["shape"])
    return {"value": n * ticks / spent[threads], "unit": "body-steps/s", "cores": threads, "kind": "port",
            "single_thread_value": n * ticks / spent[1], "threads_bit_identical": hashes[1] == hashes[threads],
            "sample": "world 0 (16,384 bodies, 2,800 ball joints), %d ticks from the timed window's first tick, the reference's fixed "
            "iterations; OpenMP where the reference uses rayon (narrowphase pairs, terrain colliders, PGS level batches >= 64), %d threads "
            "%.1f s, 1 thread %.1f s (%s, %d host cores)" % (ticks, threads, spent[threads], spent[1], cpu_model(), os.cpu_count() or 0)}


def learned_bench(a):
    """Learned + batched vs plain, PGS to residual 1e-4 (per island, at most `--max-sweeps`).

    1. Training data: the solver's own converged impulses (residual mode, less the previous-frame
       warm start) on 16 seeds of the 512-box lattice and 64 mixed-shape worlds of the batched
       workload's generator with their own seeds, ticks 120 on, about 10M contacts; 20 epochs of
       Adam at lr 1e-3 on the device.
    2. Held-out lattices and the batched worlds at tick `warmup`, the same ticks stepped from one
       snapshot with the previous-frame (reference), zero and learned warm start.
    3. Plain: one 16,384-body world at tick `warmup`, previous-frame warm start.
    """
    tol, max_it = 1.0e-4, a.max_sweeps
    t0 = time.perf_counter()
    lat = [scenes.box_lattice(seed=scenes.seed_for(4, w)) for w in range(16)]
    lat += [scenes.batch_world(w, seed_base_k=5) for w in range(64)]
    tb = rphys.Batch([rphys.World(s_) for s_ in lat])
    tb.step_n(120, graph=not a.no_graph)  # the lattices collapse, the mixed shapes land (fixed iterations)
    tb.set_residual_mode(True, tol, max_it)
    tb.set_contact_log(a.train_rows + 200000)
    rows = 0
    while rows < a.train_rows:
        tb.step_n(10, graph=not a.no_graph)
        rows = tb.contact_log_rows()[0]
        print("bench: training data %d rows, %.1f s" % (rows, time.perf_counter() - t0), file=sys.stderr, flush=True)
    t_data = time.perf_counter() - t0
    t1 = time.perf_counter()
    params, losses = tb.train_warm_start(epochs=a.epochs, lr=1.0e-3, seed=20260101)
    t_train = time.perf_counter() - t1
    print("bench: trained %d epochs in %.1f s, loss %.4f -> %.4f" % (a.epochs, t_train, losses[0], losses[-1]), file=sys.stderr, flush=True)
    del tb

    def timed(stepper, nb_):
        el, isum, ic, cts, done, conv, lrn = 0.0, 0.0, 0.0, 0.0, 0, 0.0, 0.0
        while done < a.steps:
            k = min(10, a.steps - done)
            t2 = time.perf_counter()
            st, c = stepper(k)
            el += time.perf_counter() - t2
            isum += c["pgs_iters_sum"]; ic += c["island_contacts"]; cts += c["contacts"]
            conv += c["island_contacts_converged"]; lrn += c["warm_start_learned"]
            done += k
            print("bench:   %d/%d ticks, %.1f s, %.0f contacts/tick, %.1f sweeps/contact" % (done, a.steps, el, c["contacts"] / k,
                  c["pgs_iters_sum"] / max(c["island_contacts"], 1)), file=sys.stderr, flush=True)
        return {"body_steps_per_s": nb_ * a.steps / el, "ms_per_tick": 1000.0 * el / a.steps, "mean_sweeps_per_contact": isum / max(ic, 1),
                "contacts_per_tick": cts / a.steps, "converged_contact_frac": conv / max(ic, 1), "learned_contact_frac": lrn / max(cts, 1)}

    # in distribution: 16 held-out seeds of the lattice
    lb = rphys.Batch([rphys.World(scenes.box_lattice(seed=scenes.seed_for(4, 100 + w))) for w in range(16)])
    lb.step_n(90, graph=not a.no_graph)
    lb.snapshot(1)
    lb.set_residual_mode(True, tol, max_it)
    lattice = {}
    for name, mode in (("previous_frame", 0), ("zero", 1), ("learned", 2)):
        lb.restore(1)
        lb.set_warm_start(mode, params if mode == 2 else None)
        lattice[name] = timed(lambda k: lb.step_profiled(k, 1.0 / 60.0), lb.total_bodies())
        print("bench: lattice, %s warm start: %.0f body-steps/s, %.2f sweeps per contact, %.3f converged, %.3f learned" % (
            name, lattice[name]["body_steps_per_s"], lattice[name]["mean_sweeps_per_contact"], lattice[name]["converged_contact_frac"],
            lattice[name]["learned_contact_frac"]), file=sys.stderr, flush=True)
    del lb

    worlds = [scenes.batch_world(w) for w in range(a.worlds)]
    b = rphys.Batch([rphys.World(s_) for s_ in worlds])
    nb = b.total_bodies()
    dt = float(worlds[0].get("dt", 1.0 / 60.0))
    for k in range(0, a.warmup, 50):  # warm up with the fixed iterations
        b.step_n(min(50, a.warmup - k), dt, graph=not a.no_graph)
        b.hashes()
        print("bench: batched warmup %d/%d ticks" % (min(k + 50, a.warmup), a.warmup), file=sys.stderr, flush=True)
    b.snapshot(1)
    b.set_residual_mode(True, tol, max_it)
    modes = {}
    for name, mode in (("previous_frame", 0), ("zero", 1), ("learned", 2)):
        b.restore(1)
        b.set_warm_start(mode, params if mode == 2 else None)
        modes[name] = timed(lambda k: b.step_profiled(k, dt), nb)
        print("bench: batched, %s warm start: %.0f body-steps/s, %.2f sweeps per contact, %.3f converged, %.3f learned" % (
            name, modes[name]["body_steps_per_s"], modes[name]["mean_sweeps_per_contact"], modes[name]["converged_contact_frac"],
            modes[name]["learned_contact_frac"]), file=sys.stderr, flush=True)
    del b

    w = rphys.World(scenes.mixed_shapes())
    for k in range(0, a.warmup, 50):
        if a.no_graph:
            for _ in range(min(50, a.warmup - k)):
                w.step()
        else:
            w.step_n(min(50, a.warmup - k))
        w.hash()
        print("bench: plain warmup %d/%d ticks" % (min(k + 50, a.warmup), a.warmup), file=sys.stderr, flush=True)
    w.set_residual_mode(True, tol, max_it)
    plain = timed(lambda k: w.step_profiled(k), w.num_bodies())
    print("bench: plain single world: %.0f body-steps/s, %.2f sweeps per contact" % (plain["body_steps_per_s"], plain["mean_sweeps_per_contact"]),
          file=sys.stderr, flush=True)
    value = modes["learned"]["body_steps_per_s"]
    ms = modes["learned"]["ms_per_tick"]
    return {"metric": "body-steps/s on 1 GPU, learned+batched vs plain; PGS iters to residual 1e-4", "value": value, "unit": "body-steps/s",
            "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32 (model: bf16 weights and activations, f32 accumulation)", "data": "synthetic (seeded XorShift64 scenes)",
            "config": {"workload": "learned_batch%d_vs_plain16k" % a.worlds, "residual": tol, "max_sweeps": max_it, "worlds": a.worlds, "bodies": nb,
                       "training": {"scene": "16 seeds of the 512-box lattice + 64 mixed-shape worlds", "rows": rows, "epochs": a.epochs, "lr": 1e-3,
                                    "loss_first_last": [float(losses[0]), float(losses[-1])], "data_s": t_data, "train_s": t_train},
                       "lattice_held_out": lattice, "batched": modes, "plain": plain, "learned_batched_over_plain": value / plain["body_steps_per_s"]},
            "roofline": None, "cpu_baseline": None}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["atscale", "mixed16k", "boxes512", "batch", "learned", "towers"], default="atscale")
    ap.add_argument("--train-rows", type=int, default=10_000_000, help="learned workload: contacts logged for training")
    ap.add_argument("--epochs", type=int, default=20, help="learned workload: training epochs")
    ap.add_argument("--max-sweeps", type=int, default=100, help="learned workload: PGS sweep limit per island in residual mode")
    ap.add_argument("--worlds", type=int, default=1024, help="batch workload: number of worlds")
    ap.add_argument("--cpu-ticks", type=int, default=3, help="oracle ticks sampled per 100-tick chunk")
    ap.add_argument("--no-cpu", action="store_true", help="skip the oracle sample (profiling runs)")
    ap.add_argument("--no-graph", action="store_true",
                    help="step on the stream instead of graph-captured ticks (profiling runs: rocprofv3 kernel tracing of graph launches crashed the profiled process)")
    ap.add_argument("--stages", action="store_true", help="also print the per-stage split to stderr")
    ap.add_argument("--mark-window", action="store_true",
                    help="atscale: launch the copy kernel just before and after the profiled learned + batched pass (the roofline's "
                         "window), so riftphys-rust_amd/kernel_window.py can cut that window out of a rocprofv3 kernel trace")
    ap.add_argument("--learned", action="store_true", help="towers workload: also train and compare the learned warm start")
    ap.add_argument("--graph", action="store_true",
                    help="towers workload: graph-captured, autotuned ticks (default: stream launches, measured faster on this workload)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB MiB (bandwidth reference, counter calibration)")
    a = ap.parse_args()
    if a.copy_probe:
        nbytes = a.copy_probe << 20
        print(json.dumps({"copy_probe_bytes": nbytes, "reps": 20, "gbps_read_plus_write": rphys.copy_probe(nbytes, 20)}), flush=True)
        return
    if a.gpus != 1:
        print("bench.py measures one GPU: the step library has no multi-GPU path (independent worlds would shard by "
              "world across GPUs with no per-tick exchange). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if a.workload == "atscale":
        if a.steps == 600:
            a.steps = 20
        if a.worlds == 1024:
            a.worlds = 64
        if a.train_rows == 10_000_000:
            a.train_rows = 2_000_000
        if a.cpu_ticks == 3:
            a.cpu_ticks = 60
        print(json.dumps(atscale_bench(a)), flush=True)
        return
    if a.workload == "learned":
        if a.steps == 600:
            a.steps = 40
        if a.warmup == 10:
            a.warmup = 150
        print(json.dumps(learned_bench(a)), flush=True)
        return
    if a.workload == "towers":
        if a.steps == 600:
            a.steps = 60
        if a.warmup == 10:
            a.warmup = 120
        if a.worlds == 1024:
            a.worlds = 8
        if a.train_rows == 10_000_000:
            a.train_rows = 3_000_000
        a.no_graph = not a.graph
        print(json.dumps(towers_bench(a)), flush=True)
        return
    if a.workload == "batch":
        if a.steps == 600:
            a.steps = 160
        if a.warmup == 10:
            a.warmup = 150
        print(json.dumps(batch_bench(a)), flush=True)
        return
    if a.workload == "mixed16k":
        scene = scenes.mixed_shapes()
        wl = {"workload": "mixed16k_fbm256", "bodies": 16384, "shapes": "1/3 sphere, 1/3 box, 1/3 capsule",
              "heightfield": "256x256 fBm, 0.5 m", "dt": "1/60", "vel_iters": 10, "pos_iters": 20, "worlds": 1}
    else:
        scene = scenes.box_lattice()
        wl = {"workload": "boxes512_flat64", "bodies": 512, "heightfield": "64x64 flat, 1 m", "dt": "1/60",
              "vel_iters": 10, "pos_iters": 20, "worlds": 1}
    dt = float(scene.get("dt", 1.0 / 60.0))
    iv, ip = scene.get("iters", (10, 20))

    w = rphys.World(scene)
    nb = w.num_bodies()
    print("bench: %s, %d bodies, warmup %d ticks" % (wl["workload"], nb, a.warmup), file=sys.stderr, flush=True)
    if a.warmup > 0:
        if a.no_graph:
            for k in range(a.warmup):
                w.step(dt)
                if (k + 1) % 50 == 0:
                    print("bench: warmup %d/%d ticks" % (k + 1, a.warmup), file=sys.stderr, flush=True)
        else:
            w.step_n(a.warmup, dt)
    w.hash()  # forces completion of the warmup ticks before the clock starts

    # timed region: the profiled step in chunks of up to 100 ticks, a progress line after each
    # (the CPU sample's device states are exported between chunks, off the clock)
    stages, ctr, strata = {}, {}, []
    done, el = 0, 0.0
    while done < a.steps:
        k = min(100, a.steps - done)
        if not a.no_cpu:
            strata.append((w.export_state(), k))
        t0 = time.perf_counter()
        st, c = w.step_profiled(k, dt)
        el += time.perf_counter() - t0
        merge_profile(stages, ctr, st, c)
        done += k
        print("bench: %d/%d ticks, %.2f s" % (done, a.steps, el), file=sys.stderr, flush=True)
    value = nb * a.steps / el

    roof = roofline_of(stages, ctr, iv, ip, wl["workload"])
    if a.stages:
        print_stages(stages, ctr)

    if not a.no_cpu:
        print("bench: oracle sample, %d ticks from each of %d states" % (a.cpu_ticks, len(strata)), file=sys.stderr, flush=True)
    base = None if a.no_cpu else cpu_baseline(scene, strata, a.cpu_ticks, dt)
    wl.update({"pairs_per_tick": ctr["pairs"] / ctr["ticks"], "contacts_per_tick": ctr["contacts"] / ctr["ticks"],
               "terrain_candidates_per_tick": ctr["terrain_candidates"] / ctr["ticks"], "ticks": "%d..%d" % (a.warmup, a.warmup + a.steps)})
    out = {
        "metric": "body-steps/s on 1 GPU (plain method, single world)",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1000.0 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded XorShift64 scene)",
        "config": wl,
        "roofline": roof,
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
