#!/usr/bin/env python3
"""WCSPH step throughput on one MI355X (particle-steps/s).

Workload: the 1,048,576-particle dam break (128 x 128 x 64 lattice at
spacing 0.02 m, jitter +-0.1 r, 3 wall shells, CFL-adaptive dt with
lambda = 0.4), stepped through libwcsph's C ABI.  Inputs are resident on
the GPU before the timed region; every timed step runs the whole step
(dt reduction, Morton sort, cell ranges, neighbour lists, density, forces +
integration).

Prints one JSON line.  `roofline` is for the dominant kernel, timed with
HIP events on the library's stream inside the timed region;
`cpu_baseline` is the OpenMP FP32 CPU step of the same algorithm (the
oracle's float instantiation) on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fluid-simulation_amd"))

import numpy as np  # noqa: E402

import wcsph  # noqa: E402
from wcsph import abi, scenes  # noqa: E402

HBM_PEAK_GBS = 8000.0


def kernel_bytes(name, n, ncell, nf_pairs, nb_pairs, fused_density):
    """Algorithmic HBM bytes of one launch: each per-particle array read or
    written once, list entries written (search) or read (later passes) once,
    neighbour positions re-read as cache hits."""
    pairs = nf_pairs + nb_pairs
    if name == "cell_count":
        return n * (16 + 8)  # pos in; cell, slot out
    if name == "scan":
        return ncell * (4 + 4 + 8 + 4)  # rank_to_lin, count in; start, end out; count cleared
    if name == "scatter":
        return n * (8 + 4 + 4)  # cell, slot, start in; tmp out
    if name == "permute":
        # cell, start/end, tmp; pos, vel, id in; pos, vel, id, perm, x/y/z out
        return n * (4 + 8 + 4 + 36 + 40 + 12)
    if name == "neighbors":
        # pos in; fluid + boundary counts out; list entries out; with the
        # density pass fused: rho, p, xp, vel.w out
        return n * (16 + 8 + (28 if fused_density else 0)) + 4 * pairs
    if name == "density":
        return n * (16 + 8 + 16 + 4 + 8) + 4 * pairs  # pos, counts in; xp, rho, rho, p out; list in
    if name == "normals":
        return n * (16 + 4 + 16) + 4 * nf_pairs
    if name == "force_integrate":
        # xp, vel, counts, p, id in; pos, vel, id out; list entries in; and
        # the next sort's cell and slot out (the list forms count there)
        return n * (16 + 16 + 8 + 4 + 4 + 16 + 16 + 4 + 8) + 4 * pairs
    return 0


def profile_summary(kernel, steps):
    """What the committed rocprofv3 summaries say of `kernel`: HBM bytes per
    launch (FETCH_SIZE + WRITE_SIZE scaled by the copy calibration) over a
    profiled run of the same step count from the lattice (None when no such
    run was profiled), and the counter-derived busy fractions, or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
        out = dict(t["kernels"][kernel])
    except (OSError, KeyError, ValueError):
        return None
    out["hbm_bytes_per_launch"] = t.get("hbm_bytes_by_steps", {}).get(str(steps), {}).get(kernel)
    return out


def cpu_baseline(desc, prm, hint, n_steps):
    sys.path.insert(0, ROOT)
    from oracle import oracle as O  # CPU baseline: the oracle's OpenMP FP32 step

    fl, bd, box = scenes.dam_break(desc)
    threads = O.lib().orc_max_threads()
    solver = O.Solver(prm, fl, bd, box, precision=32, parallel=True)
    solver.step(hint, 1)  # first step outside the timing (page-in)
    t0 = time.perf_counter()
    solver.step(hint, n_steps)
    t = time.perf_counter() - t0
    solver.close()
    return {
        "value": fl.shape[0] * n_steps / t,
        "unit": "particle-steps/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n_steps} steps (after 1 untimed) of the same {fl.shape[0]}-particle dam break, "
                  f"OpenMP FP32 oracle step, {threads} threads",
    }


def calibrate_copy():
    """A plain copy kernel of known size: 256 MiB read + 256 MiB written per launch."""
    import torch

    src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_per_launch": src.numel() * 4, "launches": 3}))


def batched_rate(steps, warmup, hint):
    """The batched form: 256 dam breaks, fluid counts log-uniform in
    [2k, 200k], CFL dt per scene, all stepped together (graph replay), with
    the list builder of each size bucket picked by the size-bucket tuner
    (untimed, before the timed steps)."""
    from wcsph import autotune

    descs = scenes.batch_dam_breaks()
    prm = scenes.preset_params(abi.PRESET_BENCHMARK, adaptive_dt=1, cfl_factor=0.4)
    sc = [scenes.dam_break(d) for d in descs]
    t0 = time.perf_counter()
    sim = wcsph.BatchSimulator(prm, sc)
    t_create = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, report = autotune.tune_scene_buckets(prm, sc, hint, batch=sim, steps_per_trial=3)
    t_tune = time.perf_counter() - t0
    names = {abi.NEIGHBORS_CELL: "cell", abi.NEIGHBORS_PARTICLE: "particle"}
    buckets = [{"bucket": r["bucket"], "scenes": r["scenes"], "fluid_particles": r["fluid_particles"],
                "builder": names[r["builder"]],
                "trials_ms_per_step": {f"{names[d['neighbor_mode']]}/{d['neighbor_groups']}": ms
                                       for d, ms in r["trials"]}} if "bucket" in r else r for r in report]
    modes = sim.get_scene_neighbor_modes()
    buckets.append({"applied": {names[m]: int((modes == m).sum()) for m in (abi.NEIGHBORS_CELL,
                                                                             abi.NEIGHBORS_PARTICLE)}})
    n = sim.n
    if warmup:
        sim.step(hint, warmup)
    sim.reset()
    sim.step(hint, 1)  # graph capture outside the timing
    sim.reset()
    sim.sync()
    t0 = time.perf_counter()
    sim.step_async(hint, steps)
    st = sim.sync()
    t = time.perf_counter() - t0
    out = {"value": n * steps / t, "unit": "particle-steps/s", "scenes": len(sc), "fluid_particles": n,
           "boundary_particles": sim.nb, "steps": steps, "ms_per_step": t * 1e3 / steps,
           "min_scene": min(int(np.prod(d.dims)) for d in descs), "max_scene": max(int(np.prod(d.dims)) for d in descs),
           "create_s": t_create, "tune_s": t_tune, "size_buckets": buckets, "flags": st.flags}
    sim.close()
    return out


def obstacle_rate(steps, hint=5e-4):
    """The obstacle-collapse workload: 8.3M fluid particles (2 x 1 x 1 m
    block, 10% in clusters at spacing 1.8 r) falling onto a voxel-sampled
    sphere + torus, CFL dt.  The autotuner picks the variant on the initial
    state (5 steps per candidate, state restored), then `steps` steps are
    timed as graph replays."""
    desc = scenes.OBSTACLE_COLLAPSE_8M
    prm = scenes.preset_params(abi.PRESET_BENCHMARK, adaptive_dt=1, cfl_factor=0.4, particle_radius=desc.r)
    t0 = time.perf_counter()
    fl, bd, box = scenes.obstacle_collapse(desc)
    sim = wcsph.FluidSimulator(prm, fl, bd, box)
    t_create = time.perf_counter() - t0
    from wcsph import autotune
    best, trials = autotune.tune(sim, hint, steps_per_trial=5)
    sim.step(hint, 1)  # graph capture outside the timing
    sim.reset()
    sim.sync()
    t0 = time.perf_counter()
    sim.step_async(hint, steps)
    st = sim.sync()
    t = time.perf_counter() - t0
    n = sim.n
    out = {"workload": "obstacle_collapse_8M: 2x1x1 m block, 10% clustered at 1.8r, sphere R 0.5 + torus (0.8, 0.2) "
                       "sampled at 2r, CFL dt", "value": n * steps / t, "unit": "particle-steps/s",
           "fluid_particles": n, "boundary_particles": sim.nb, "steps": steps, "ms_per_step": t * 1e3 / steps,
           "create_s": t_create, "variant": autotune.variant_to_dict(best),
           "trials_ms_per_step": [ms for _, ms in trials], "sim_time": st.sim_time, "flags": st.flags,
           "max_fluid_neighbors": st.max_fluid_neighbors}
    sim.close()
    return out


def adjoint_ratio(steps=200, every=20):
    """The differentiable form: a 65,536-particle dam break (64 x 32 x 32
    lattice), fixed dt = 2e-4.  Targets: final positions of a run with
    alpha = 0.02 and seeded Gaussian initial velocities (sigma = 0.1 m/s);
    loss: mean squared distance to them from zero initial velocities at
    alpha = 0.01.  Times the plain forward (graph), the checkpointed forward
    and the backward (recompute + adjoint)."""
    desc = scenes.DamBreak(dims=(64, 32, 32), box_max=(3.2, 1.0, 1.0), seed=77)
    prm = scenes.preset_params(abi.PRESET_BENCHMARK)
    fl, bd, box = scenes.dam_break(desc)
    n = fl.shape[0]
    dt = 2e-4
    sim = wcsph.FluidSimulator(prm, fl, bd, box)
    rng = np.random.default_rng(2025)
    vt = np.zeros((n, 4), np.float32)
    vt[:, :3] = rng.normal(0.0, 0.1, (n, 3)).astype(np.float32)
    sim.set_viscosity(0.02)
    target = sim.grad_forward(vt, dt, steps, every)
    sim.set_viscosity(0.01)
    v0 = np.zeros((n, 4), np.float32)
    sim.grad_forward(v0, dt, steps, every)  # warm-up
    t0 = time.perf_counter()
    xt = sim.grad_forward(v0, dt, steps, every)
    t_fwd = time.perf_counter() - t0
    g = np.zeros((n, 4), np.float32)
    g[:, :3] = 2.0 * (xt[:, :3] - target[:, :3]) / n
    t0 = time.perf_counter()
    gv, ga = sim.grad_backward(g)
    t_bwd = time.perf_counter() - t0
    loss = float(np.sum((xt[:, :3].astype(np.float64) - target[:, :3]) ** 2) / n)
    sim.reset()
    sim.step(dt, 1)
    sim.reset()
    t0 = time.perf_counter()
    sim.step(dt, steps)
    t_plain = time.perf_counter() - t0
    sim.close()
    return {"workload": "dam_break_65k: 64x32x32 lattice, fixed dt 2e-4, checkpoint every %d" % every,
            "fluid_particles": n, "steps": steps, "plain_forward_s": t_plain, "grad_forward_s": t_fwd,
            "grad_backward_s": t_bwd, "fwd_plus_adjoint_over_forward": (t_fwd + t_bwd) / t_plain,
            "loss": loss, "grad_alpha": ga, "grad_v0_norm": float(np.linalg.norm(gv[:, :3]))}


def set_mode(sim, mode):
    if mode:
        v = sim.get_variant()
        v.neighbor_mode = {"cell": abi.NEIGHBORS_CELL, "particle": abi.NEIGHBORS_PARTICLE,
                           "cell_loop": abi.NEIGHBORS_CELL_LOOP, "brick": abi.NEIGHBORS_BRICK}[mode]
        sim.set_variant(v)


def grid_cells(prm, bd):
    """Cells of the scene's fluid grid (cell size h; for the scan's bytes)."""
    cs = 4 * prm.particle_radius
    lo = bd[:, :3].min(axis=0)
    hi = bd[:, :3].max(axis=0)
    return int(np.prod(np.floor(hi / cs) - np.floor(lo / cs) + 5))


def plain_run(desc, prm, hint, steps, warmup, mode=None, split=None, pairs=True):
    """One scene, `steps` steps from its initial state, three times over:
    (1) eager launches with HIP events between the kernels on the library's
    stream (per-kernel launch times), (2) graph replays timed by the host
    clock (the throughput), (3) when `pairs`, one step at a time reading the
    list totals after each (the run is deterministic, so these are the
    totals of the steps timed in (1) and (2)).  Warm-up: the first `warmup`
    steps of the same run, then back to the initial state.  `split` also
    times the first `split` graph-replayed steps on their own."""
    fl, bd, box = scenes.dam_break(desc)
    n = fl.shape[0]
    sim = wcsph.FluidSimulator(prm, fl, bd, box)
    lib = sim.lib
    set_mode(sim, mode)
    fused = sim.get_variant().neighbor_mode == abi.NEIGHBORS_PARTICLE
    if warmup:
        sim.step(hint, warmup)
    sim.reset()

    lib.wcsph_set_profiling(sim.handle, 1)
    lib.wcsph_kernel_times(sim.handle, None, None, 1)
    sim.sync()
    t0 = time.perf_counter()
    sim.step_async(hint, steps)
    st = sim.sync()
    t_eager = time.perf_counter() - t0
    ms = (ctypes.c_double * abi.NUM_KERNELS)()
    cnt = (ctypes.c_uint64 * abi.NUM_KERNELS)()
    lib.wcsph_kernel_times(sim.handle, ms, cnt, 1)
    lib.wcsph_set_profiling(sim.handle, 0)
    kern = {abi.KERNEL_NAMES[k]: ms[k] / cnt[k] for k in range(abi.NUM_KERNELS) if cnt[k]}

    sim.reset()
    sim.step(hint, 1)  # graph capture outside the timing
    sim.reset()
    sim.sync()
    t_split = None
    t0 = time.perf_counter()
    if split and split < steps:
        sim.step_async(hint, split)
        sim.sync()
        t_split = time.perf_counter() - t0
        sim.step_async(hint, steps - split)
    else:
        sim.step_async(hint, steps)
    st2 = sim.sync()
    t_graph = time.perf_counter() - t0

    per_step = []
    if pairs:
        sim.reset()
        for _ in range(steps):
            sim.step(hint, 1)
            per_step.append((sim.neighbor_total(0), sim.neighbor_total(1)))
    out = {"n": n, "steps": steps, "nb": int(bd.shape[0]), "ncell": grid_cells(prm, bd), "fused_density": fused,
           "t_eager": t_eager, "t_graph": t_graph, "t_split": t_split, "kern": kern, "per_step": per_step,
           "st": st, "st2": st2, "variant": sim.get_variant()}
    sim.close()
    return out


def roofline(run):
    """The dominant kernel of a plain run against the HBM roofline: its mean
    algorithmic bytes over the timed steps over its mean launch time over
    the same steps."""
    kern = run["kern"]
    dominant = max((k for k in kern if k != "time_step"), key=lambda k: kern[k])
    per = [kernel_bytes(dominant, run["n"], run["ncell"], nf, nb, run["fused_density"]) for nf, nb in run["per_step"]]
    dom_bytes = sum(per) / len(per)
    achieved = dom_bytes / (kern[dominant] * 1e-3) / 1e9
    prof = profile_summary(dominant, run["steps"]) or {}
    out = {"bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": prof.get("hbm_bytes_per_launch"),
           "bytes_per_launch": dom_bytes, "avg_launch_ms": kern[dominant]}
    if "limiter" in prof:
        # what keeps the kernel below the HBM roof, from the SQ / TA counters
        out["limiter"] = prof["limiter"]
    return out


def summarize(prof_dir):
    """Writes the profiles/ summaries of the rocprofv3 runs under prof_dir
    (subfolders: stats = --kernel-trace --stats; fetch, write = --pmc
    FETCH_SIZE / WRITE_SIZE; cfetch, cwrite = the same over --calibrate-copy;
    sq, ta = --pmc counter passes), each run over `bench_cmd` below; every
    summary names the command it comes from."""
    import csv
    import collections

    bench_cmd = "python3 bench.py --only plain --steps 20 --warmup 5"

    def short(name):
        name = name.split("(")[0].replace("void ", "").replace("wcsph_dev::", "")
        return name.strip()

    def read_counters(sub):
        path = os.path.join(prof_dir, sub, "run_counter_collection.csv")
        rows = collections.defaultdict(lambda: collections.defaultdict(list))
        with open(path) as f:
            for r in csv.DictReader(f):
                rows[short(r["Kernel_Name"])][r["Counter_Name"]].append(
                    (int(r["Dispatch_Id"]), float(r["Counter_Value"])))
        return rows

    def mean_skip2(v):
        v = [x for _, x in sorted(v)]
        v = v[2:] if len(v) > 4 else v
        return sum(v) / len(v)

    def write_stats(sub, out_name, title, cmd, workload):
        path = os.path.join(prof_dir, sub, "run_kernel_stats.csv")
        if not os.path.exists(path):
            return
        lines = [f"# {title}: rocprofv3 kernel statistics", "", "Command (repo root, one MI355X, TMPDIR=/tmp):", "",
                 "```", f"rocprofv3 --kernel-trace --stats --output-format csv -d rocprof_out/{sub} -o run -- "
                 f"{cmd}", "```", "", f"Workload: {workload}", "",
                 "| Kernel | Calls | Average (us) | Min (us) | Max (us) | Share of GPU time |",
                 "|---|---:|---:|---:|---:|---:|"]
        stats = []
        with open(path) as f:
            for r in csv.DictReader(f):
                stats.append((short(r["Name"]), int(r["Calls"]), float(r["AverageNs"]) / 1e3,
                              float(r["MinNs"]) / 1e3, float(r["MaxNs"]) / 1e3, float(r["Percentage"])))
        stats.sort(key=lambda x: -x[5])
        for k, c, a, mn, mx, pct in stats:
            lines.append(f"| `{k}` | {c} | {a:.1f} | {mn:.1f} | {mx:.1f} | {pct:.1f}% |")
        with open(os.path.join(ROOT, "profiles", out_name), "w") as f:
            f.write("\n".join(lines) + "\n")

    write_stats("stats", "wcsph_step_dam_break_1M_kernel_stats.md",
                "WCSPH step, 1,048,576-particle dam break", bench_cmd,
                "128x128x64 fluid lattice, 369,696 wall particles, CFL dt (lambda 0.4), default variant "
                "(per-particle list builder with the density pass fused; the force pass counts the next sort's "
                "cells): steps 1-5 (warm-up), then steps 1-20 three times (eager with HIP events, graph replay "
                "after one capture step, one step at a time for the list totals).")
    write_stats("stats_full", "wcsph_step_dam_break_1M_500_steps_kernel_stats.md",
                "WCSPH step, 1,048,576-particle dam break, the whole 500-step run",
                "python3 bench.py --only plain --steps 500 --warmup 5",
                "as the 20-step summary, over steps 1-500 of the run (the state compresses as it goes: 28 to "
                "34 fluid neighbours per particle).")
    write_stats("stats_batched", "wcsph_batched_256_scenes_kernel_stats.md",
                "Batched WCSPH step, 256 dam-break scenes", "python3 bench.py --only batched --batched-steps 200",
                "256 scenes of 2,028 to 195,112 fluid particles (9,907,267 in all, log-uniform counts), "
                "17,129,940 wall particles, per-scene CFL dt: the size-bucket tuner's trials (each bucket's "
                "scenes as a batch of their own), 5 warm-up steps, a reset, 1 + 200 graph-replayed steps.")
    write_stats("stats_obstacle", "wcsph_step_obstacle_collapse_8M_kernel_stats.md",
                "WCSPH step, 8.3M-particle obstacle collapse", "python3 bench.py --only obstacle --obstacle-steps 200",
                "2x1x1 m block of 8,313,167 particles (10% in clusters at 1.8r), sphere R 0.5 + torus (0.8, 0.2) "
                "sampled at 2r (354,618 particles), CFL dt: the autotuner's trials over its default space (5 "
                "steps each, state restored), then 1 + 200 graph-replayed steps of the chosen variant.")
    write_stats("stats_adjoint", "wcsph_adjoint_dam_break_65k_kernel_stats.md",
                "Differentiable WCSPH step, 65,536-particle dam break", "python3 bench.py --only adjoint",
                "64x32x32 lattice, fixed dt 2e-4, 200 steps: target run (alpha 0.02), a warm-up forward, the "
                "timed checkpointed forward (K = 20), the backward (recompute + adjoint), and 1 + 200 plain steps.")

    # HBM traffic, scaled by the copy calibration
    cal = {}
    for sub, counter in (("cfetch", "FETCH_SIZE"), ("cwrite", "WRITE_SIZE")):
        # KiB of the 256 MiB copy: the largest dispatch of the run (the
        # fill before it writes as much and reads nothing)
        rows = read_counters(sub)
        cal[counter] = max(v for d in rows.values() for _, v in d.get(counter, []))
    copy_bytes = 64 * 1024 * 1024 * 4
    fscale = copy_bytes / (cal["FETCH_SIZE"] * 1024)
    wscale = copy_bytes / (cal["WRITE_SIZE"] * 1024)
    full_cmd = "python3 bench.py --only plain --steps 500 --warmup 5"
    fetch, write = read_counters("fetch"), read_counters("write")
    have_full = os.path.isdir(os.path.join(prof_dir, "fetch_full")) and os.path.isdir(
        os.path.join(prof_dir, "write_full"))
    fetch_full = read_counters("fetch_full") if have_full else {}
    write_full = read_counters("write_full") if have_full else {}
    sq = read_counters("sq") if os.path.isdir(os.path.join(prof_dir, "sq")) else {}
    ta = read_counters("ta") if os.path.isdir(os.path.join(prof_dir, "ta")) else {}
    td = read_counters("td") if os.path.isdir(os.path.join(prof_dir, "td")) else {}
    groups = {}
    prefixes = (("cell_count", "k_cell_count"), ("scan", "k_scan_"), ("scatter", "k_scatter"),
                ("permute", "k_permute"))
    for k in sorted(fetch):
        g = next((g for g, pre in prefixes if k.startswith(pre)), None)
        if g:
            groups.setdefault(g, []).append(k)
        elif k.startswith("k_neighbors"):
            groups.setdefault("neighbors", []).append(k)
        elif k.startswith("k_density"):
            groups.setdefault("density", []).append(k)
        elif k.startswith("k_force_integrate"):
            groups.setdefault("force_integrate", []).append(k)
        elif k.startswith("k_normals"):
            groups.setdefault("normals", []).append(k)
    kernels = {}
    tlines = ["# WCSPH step, 1,048,576-particle dam break: HBM traffic per launch (rocprofv3 FETCH_SIZE / "
              "WRITE_SIZE)", "", "Commands (repo root, one MI355X, TMPDIR=/tmp; one counter per run):", "", "```",
              f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d rocprof_out/fetch -o run -- {bench_cmd}",
              f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d rocprof_out/write -o run -- {bench_cmd}",
              f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d rocprof_out/fetch_full -o run -- {full_cmd}",
              f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d rocprof_out/write_full -o run -- {full_cmd}",
              "rocprofv3 --pmc FETCH_SIZE --output-format csv -d rocprof_out/cfetch -o run -- "
              "python3 bench.py --calibrate-copy",
              "rocprofv3 --pmc WRITE_SIZE --output-format csv -d rocprof_out/cwrite -o run -- "
              "python3 bench.py --calibrate-copy", "```", "",
              f"Calibration: a plain 256 MiB device copy reads FETCH_SIZE {cal['FETCH_SIZE']:.0f} KiB and "
              f"WRITE_SIZE {cal['WRITE_SIZE']:.0f} KiB per launch, so FETCH_SIZE is scaled by {fscale:.3f} and "
              f"WRITE_SIZE by {wscale:.3f}.  Averages skip the first two launches of each kernel.", ""]
    by_steps = {}
    for steps, fe, wr in ((20, fetch, write), (500, fetch_full, write_full)):
        if not fe:
            continue
        tlines += [f"Steps 1-{steps} of the run from the lattice (`--steps {steps} --warmup 5`):", "",
                   "| Kernel | Launches | FETCH (MB, scaled) | WRITE (MB, scaled) | HBM MB / launch |",
                   "|---|---:|---:|---:|---:|"]
        by_steps[str(steps)] = {}
        for g, names in groups.items():
            fb = sum(mean_skip2(fe[k]["FETCH_SIZE"]) for k in names if k in fe) * 1024 * fscale
            wb = sum(mean_skip2(wr[k]["WRITE_SIZE"]) for k in names if k in wr) * 1024 * wscale
            launches = max((len(fe[k]["FETCH_SIZE"]) for k in names if k in fe), default=0)
            if not launches:
                continue
            by_steps[str(steps)][g] = fb + wb
            if steps == 20:
                kernels[g] = {"kernels": names, "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb,
                              "hbm_bytes_per_launch": fb + wb}
            tlines.append(f"| `{'`, `'.join(names)}` | {launches} | {fb / 1e6:.1f} | {wb / 1e6:.1f} | "
                          f"{(fb + wb) / 1e6:.1f} |")
        tlines.append("")
    with open(os.path.join(ROOT, "profiles", "wcsph_step_dam_break_1M_hbm_traffic.md"), "w") as f:
        f.write("\n".join(tlines) + "\n")

    # SQ / TA / TD counters: VALU issue, texture-address and texture-data
    # (the L1-to-register return path) busy fractions
    if sq and ta:
        clines = ["# WCSPH step, 1,048,576-particle dam break: SQ / TA / TD counters per launch", "",
                  "Commands (repo root, one MI355X, TMPDIR=/tmp; one pass per run):", "", "```",
                  "rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR "
                  "SQ_INSTS_LDS SQ_WAVE_CYCLES SQ_BUSY_CYCLES --output-format csv -d rocprof_out/sq -o run -- "
                  f"{bench_cmd}",
                  "rocprofv3 --pmc TA_TA_BUSY_sum TA_BUSY_avr GRBM_GUI_ACTIVE GRBM_COUNT --output-format csv "
                  f"-d rocprof_out/ta -o run -- {bench_cmd}",
                  "rocprofv3 --pmc TD_TD_BUSY_sum TD_TC_STALL_sum GRBM_GUI_ACTIVE --output-format csv "
                  f"-d rocprof_out/td -o run -- {bench_cmd}", "```", "",
                  "cycles = GRBM_GUI_ACTIVE / 8 (the counter sums the 8 XCDs; active us = cycles / 2.4 GHz, which "
                  "includes dispatch overhead, so short kernels read longer than their kernel-trace time); "
                  "VALU busy = 4 x SQ_INSTS_VALU / "
                  "(1024 SIMDs x cycles) (a wave64 VALU instruction occupies its SIMD 4 cycles); TA busy = "
                  "TA_TA_BUSY_sum / (256 TAs x cycles); TD busy = TD_TD_BUSY_sum / (256 TDs x cycles) (the "
                  "texture-data unit returns loaded data from L1 to the registers); VMEM / wave = vector memory "
                  "instructions per wavefront.", "",
                  "| Kernel | active us (counter run) | VALU busy | TA busy | TD busy | VALU instr / wave | "
                  "VMEM rd / wave | VMEM wr / wave | LDS instr / wave |", "|---|---:|---:|---:|---:|---:|---:|---:|---:|"]
        for g, names in groups.items():
            for k in names:
                if k not in sq or k not in ta or "GRBM_GUI_ACTIVE" not in ta[k]:
                    continue
                cyc = mean_skip2(ta[k]["GRBM_GUI_ACTIVE"]) / 8
                waves = mean_skip2(sq[k]["SQ_WAVES"])
                valu = 4 * mean_skip2(sq[k]["SQ_INSTS_VALU"]) / (1024 * cyc)
                tab = mean_skip2(ta[k]["TA_TA_BUSY_sum"]) / (256 * cyc)
                tdb = None
                if k in td and "GRBM_GUI_ACTIVE" in td[k]:
                    tdb = mean_skip2(td[k]["TD_TD_BUSY_sum"]) / (256 * mean_skip2(td[k]["GRBM_GUI_ACTIVE"]) / 8)
                per_wave = {c: mean_skip2(sq[k][c]) / waves for c in
                            ("SQ_INSTS_VALU", "SQ_INSTS_VMEM_RD", "SQ_INSTS_VMEM_WR", "SQ_INSTS_LDS")}
                us = cyc / 2.4e3
                tds = f"{tdb:.0%}" if tdb is not None else "-"
                clines.append(f"| `{k}` | {us:.1f} | {valu:.0%} | {tab:.0%} | {tds} | "
                              f"{per_wave['SQ_INSTS_VALU']:.0f} | "
                              f"{per_wave['SQ_INSTS_VMEM_RD']:.1f} | {per_wave['SQ_INSTS_VMEM_WR']:.1f} | "
                              f"{per_wave['SQ_INSTS_LDS']:.1f} |")
                if k == names[0]:
                    busy = {"vector ALU issue": valu, "texture-address (TA) unit": tab}
                    if tdb is not None:
                        busy["texture-data (TD) return path"] = tdb
                    top = max(busy, key=busy.get)
                    kernels[g]["limiter"] = {"valu_busy": valu, "ta_busy": tab, "td_busy": tdb,
                                             "what": f"{top} ({busy[top]:.0%} busy; rocprofv3 SQ / TA / TD "
                                                     "counters)" if busy[top] > 0.5
                                             else "latency (no unit above 50% busy)"}
        with open(os.path.join(ROOT, "profiles", "wcsph_step_dam_break_1M_counters.md"), "w") as f:
            f.write("\n".join(clines) + "\n")
    out = {"command_fetch": f"rocprofv3 --pmc FETCH_SIZE --output-format csv -d rocprof_out/fetch -o run -- "
                            f"{bench_cmd}",
           "command_write": f"rocprofv3 --pmc WRITE_SIZE --output-format csv -d rocprof_out/write -o run -- "
                            f"{bench_cmd}",
           "command_full": f"the same two over {full_cmd}",
           "calibration": {"copy_bytes": copy_bytes, "fetch_size_kib": cal["FETCH_SIZE"],
                           "write_size_kib": cal["WRITE_SIZE"], "fetch_scale": fscale, "write_scale": wscale},
           "kernels": kernels, "hbm_bytes_by_steps": by_steps}
    with open(os.path.join(ROOT, "profiles", "traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({g: round(v["hbm_bytes_per_launch"] / 1e6, 1) for g, v in kernels.items()}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", choices=["1m", "32k"], default="1m")
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--full-steps", type=int, default=500,
                    help="steps of the whole 1M dam-break run reported beside the timed K (0 skips it)")
    ap.add_argument("--c1-steps", type=int, default=1000,
                    help="steps of the 32,768-particle dam break at fixed dt 1e-4 (0 skips it)")
    ap.add_argument("--batched-steps", type=int, default=200,
                    help="steps of the 256-scene batched form (0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=200,
                    help="steps of the differentiable 65k-particle run (0 skips it)")
    ap.add_argument("--neighbor-mode", choices=["cell", "particle", "cell_loop", "brick"], default=None,
                    help="neighbour search form of the timed runs (default: the library's)")
    ap.add_argument("--obstacle-steps", type=int, default=200,
                    help="steps of the 8.3M-particle obstacle collapse (0 skips it)")
    ap.add_argument("--only", choices=["plain", "batched", "adjoint", "obstacle"], default=None,
                    help="run just that workload and print its JSON (for profiling one workload)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 256 MiB device copy (for scaling rocprofv3 byte counters)")
    ap.add_argument("--summarize", metavar="DIR", default=None,
                    help="write the profiles/ summaries of the rocprofv3 CSV runs under DIR (no GPU)")
    args = ap.parse_args()
    if args.summarize:
        summarize(args.summarize)
        return
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.only in ("batched", "adjoint", "obstacle"):
        hint = 5e-4
        if args.only == "batched":
            print(json.dumps(batched_rate(args.batched_steps, min(args.warmup, 5), hint)))
        elif args.only == "adjoint":
            print(json.dumps(adjoint_ratio(args.adjoint_steps)))
        else:
            print(json.dumps(obstacle_rate(args.obstacle_steps, hint)))
        return

    if args.config == "1m":
        desc = scenes.DAM_BREAK_1M
        prm = scenes.preset_params(abi.PRESET_BENCHMARK, adaptive_dt=1, cfl_factor=0.4)
        hint = 5e-4
        workload = "dam_break_1M: 128x128x64 fluid lattice, 369,696 wall particles, CFL dt (lambda 0.4)"
    else:
        desc = scenes.DAM_BREAK_32K
        prm = scenes.preset_params(abi.PRESET_BENCHMARK)
        hint = 1e-4
        workload = "dam_break_32k: 32^3 fluid lattice, fixed dt 1e-4"

    run = plain_run(desc, prm, hint, args.steps, args.warmup, args.neighbor_mode)
    n = run["n"]
    t_eager, t_graph, st, st2 = run["t_eager"], run["t_graph"], run["st"], run["st2"]
    best_t = min(t_eager, t_graph)
    value = n * args.steps / best_t
    nf, nb = run["per_step"][-1]
    out = {
        "metric": "particle-steps/s at 1M particles (plain WCSPH step)",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best_t * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": workload, "fluid_particles": n, "boundary_particles": run["nb"],
                   "preset": "benchmark", "timed_steps": f"steps 1-{args.steps} of the run from the lattice",
                   "mode": "graph" if t_graph <= t_eager else "eager+events"},
        "roofline": roofline(run),
        "kernels_ms": run["kern"],
        "eager_events_ms_per_step": t_eager * 1e3 / args.steps,
        "graph_ms_per_step": t_graph * 1e3 / args.steps,
        "sum_kernel_ms_per_step": sum(run["kern"].values()),
        "mean_fluid_neighbors_timed": sum(x for x, _ in run["per_step"]) / len(run["per_step"]) / n,
        "mean_fluid_neighbors_last": nf / n,
        "mean_boundary_neighbors_last": nb / n,
        "last_dt": st2.last_dt,
        "max_fluid_neighbors": max(st.max_fluid_neighbors, st2.max_fluid_neighbors),
        "max_boundary_neighbors": max(st.max_boundary_neighbors, st2.max_boundary_neighbors),
        "flags": st.flags | st2.flags,
    }
    if args.only == "plain":
        print(json.dumps(out))
        return
    if args.config == "1m" and args.full_steps:
        # the whole run the metric is quoted on (500 steps from the lattice),
        # timed the same way, and its first batched-steps steps on their own
        full = plain_run(desc, prm, hint, args.full_steps, 0, args.neighbor_mode,
                         split=args.batched_steps or None)
        ft = min(full["t_eager"], full["t_graph"])
        out["full_run"] = {"steps": args.full_steps, "value": n * args.full_steps / ft,
                           "ms_per_step": ft * 1e3 / args.full_steps, "roofline": roofline(full),
                           "kernels_ms": full["kern"],
                           "mean_fluid_neighbors": sum(x for x, _ in full["per_step"]) / len(full["per_step"]) / n,
                           "sim_time": full["st2"].sim_time, "flags": full["st"].flags | full["st2"].flags}
        if full["t_split"]:
            out["full_run"]["first_steps"] = {"steps": args.batched_steps,
                                              "value": n * args.batched_steps / full["t_split"]}
    if args.config == "1m" and args.c1_steps:
        c1 = plain_run(scenes.DAM_BREAK_32K, scenes.preset_params(abi.PRESET_BENCHMARK), 1e-4, args.c1_steps, 20,
                       args.neighbor_mode, pairs=False)
        out["c1_32k"] = {"workload": "dam_break_32k: 32^3 fluid lattice, fixed dt 1e-4", "steps": args.c1_steps,
                         "fluid_particles": c1["n"], "value": c1["n"] * args.c1_steps / c1["t_graph"],
                         "ms_per_step": c1["t_graph"] * 1e3 / args.c1_steps,
                         "sum_kernel_ms_per_step": sum(c1["kern"].values()), "kernels_ms": c1["kern"],
                         "variant_neighbor_mode": c1["variant"].neighbor_mode,
                         "flags": c1["st"].flags | c1["st2"].flags}
    if args.batched_steps and args.config == "1m":
        b = batched_rate(args.batched_steps, min(args.warmup, 5), hint)
        b["ratio_to_plain_timed"] = b["value"] / value
        first = out.get("full_run", {}).get("first_steps")
        if first:
            # the same step count from the same start: batched vs the 1M scene
            b["ratio_to_plain_same_steps"] = b["value"] / first["value"]
        out["batched"] = b
    if args.adjoint_steps and args.config == "1m":
        out["adjoint"] = adjoint_ratio(args.adjoint_steps)
    if args.obstacle_steps and args.config == "1m":
        out["obstacle_collapse"] = obstacle_rate(args.obstacle_steps, hint)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(desc, prm, hint, args.cpu_steps)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
