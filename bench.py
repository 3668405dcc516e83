#!/usr/bin/env python3
"""Benchmark of the fluid-solid RK4 stepper on one MI355X.

A bench "step" is one RK4 time step (4 stages + interface filter + receiver
record) over the whole grid of the workload.  Default workload: the 2001 x 1001
body-fitted seafloor model (z = 500 + 40 sin(2 pi x / 800) m, 5 % random solid,
20 Hz, dt = 0.25 ms, 30-cell CFS-PML), K = 6000 steps = one full shot record.

Prints one JSON line: value = Gcell-updates/s (nx * nz * K / wall time of the
hipGraph-replayed run, inputs resident in HBM), the roofline of the dominant
kernel (the RK stage-2 kernel over both media, its mean duration over the K
steps of the same run with every launch timed by start / stop events its own
dispatch stamps), the FP64 CPU baseline on bounded samples of the same
workload (all host cores and one core), and the shots/s of the batched
forward+adjoint gradient and of the batched forward, each beside the same shots
run one at a time.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "marineseis-fdm-2d_amd"))

HBM_PEAK_GBPS = 8000.0


def _oracle_rate(model, nsteps: int, threads: int):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(model, threads=threads)
    o.reset(0)
    o.run(2)   # first-touch / page-in outside the timed sample
    t0 = time.perf_counter()
    o.run(nsteps)
    dt = time.perf_counter() - t0
    return model.nx * model.nz * nsteps / dt / 1e9, dt


def cpu_baseline(model, nsteps: int, nsteps_1core: int):
    """FP64 oracle (OpenMP over rows) timed on `nsteps` steps of the same model
    with every host core the run may use, and on `nsteps_1core` steps with one."""
    threads = os.cpu_count() or 1
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    if cap > 0:
        threads = min(threads, cap)
    v, dt = _oracle_rate(model, nsteps, threads)
    out = {"value": round(v, 5), "unit": "Gcell-updates/s", "cores": threads, "kind": "port",
           "sample": f"{nsteps} FP64 RK4 steps of the same {model.nx}x{model.nz} model "
                     f"(after 2 untimed steps), {dt:.2f} s wall"}
    if nsteps_1core > 0:
        v1, dt1 = _oracle_rate(model, nsteps_1core, 1)
        out["single_core"] = {"value": round(v1, 5), "unit": "Gcell-updates/s", "cores": 1,
                              "sample": f"{nsteps_1core} FP64 RK4 steps of the same model on one thread, "
                                        f"{dt1:.2f} s wall"}
    return out


def stage2_kernel_names(solid: bool, tile_rows: int, occupancy: int = 0):
    """rocprofv3 names of the RK stage-2 launch (F direction), vertical-column
    metrics first.  Template arguments: direction, stage, tile rows, metric form
    (1 = vertical grid columns), then the top kind (free surface: 1) of the
    solid-only kernel, or the CPML flag (0) and the FIN flag (0) of the
    fluid-solid one.  Occupancy policy 1 runs the _w5 copies (vertical columns
    only)."""
    w5 = "_w5" if occupancy == 1 else ""
    if solid:
        return [f"free_stage_kernel{w5}<0, 2, {tile_rows}, 1, 1>"] + ([] if w5 else [f"free_stage_kernel<0, 2, {tile_rows}, 0, 1>"])
    return [f"stage_kernel{w5}<0, 2, {tile_rows}, 1, 0, 0>"] + ([] if w5 else [f"stage_kernel<0, 2, {tile_rows}, 0, 0, 0>"])


def measured_traffic(workload: str, kernels):
    """(kernel name, HBM bytes per launch) of the first of `kernels` found in the
    committed rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE passes scaled
    by a plain device copy), or (None, None)."""
    path = os.path.join(ROOT, "profiles", f"stage_kernel_{workload}.json")
    try:
        with open(path) as f:
            table = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None, None
    for name in kernels:
        if name in table:
            return name, round(table[name]["traffic_bytes"])
        # the occupancy copies (_w5) run the same tile code: same accesses
        base = name.replace("_w5<", "<")
        if base in table:
            return name, round(table[base]["traffic_bytes"])
    return None, None


def bench_batch(args, nsteps=None):
    """64 shot windows of uneven size: one batched run vs the same shots one at a
    time with the single-shot solver, both on `args.batch_steps` steps per shot
    (shots/s scaled to the full 4000-step record).  Each side runs the tile
    height its own autotuner picks: the batch's over the whole batch, the
    single-shot solver's on the first window (applied to every window)."""
    import msfdm
    from msfdm import modelgen
    models = modelgen.batch_shots(n=args.batch_shots, nt=4000)
    cells = sum(m.nx * m.nz for m in models)
    nsteps = nsteps or args.batch_steps
    b = msfdm.Batch(models)
    best = b.autotune(64)
    occ = b.occupancy
    b.reset()
    b.run(64)
    b.sync()
    t0 = time.perf_counter()
    b.run(nsteps)
    b.sync()
    t_batch = time.perf_counter() - t0
    b.close()
    probe = msfdm.Solver(models[0])
    best_seq = probe.autotune(64)
    occ_seq = probe.occupancy
    probe.close()
    t_seq = 0.0
    for m in models:
        s = msfdm.Solver(m)
        s.tile_rows = best_seq
        s.occupancy = occ_seq
        s.reset(0)
        s.run(64)
        s.sync()
        t0 = time.perf_counter()
        s.run(nsteps)
        s.sync()
        t_seq += time.perf_counter() - t0
        s.close()
    scale = 4000.0 / nsteps
    out = {"metric": "batched forward shots/s (64 uneven shot windows, 4000-step records)",
           "value": round(len(models) / (t_batch * scale), 4), "unit": "shots/s",
           "sequential_value": round(len(models) / (t_seq * scale), 4),
           "speedup_vs_sequential": round(t_seq / t_batch, 3),
           "batched_Gcell_per_s": round(cells * nsteps / t_batch / 1e9, 3),
           "sequential_Gcell_per_s": round(cells * nsteps / t_seq / 1e9, 3),
           "tile_rows": best, "occupancy_policy": occ, "tile_rows_sequential": best_seq,
           "occupancy_policy_sequential": occ_seq, "steps_timed": nsteps, "shots": len(models), "total_cells": cells,
           "note": f"timed {nsteps} of 4000 steps per shot after 64 warm-up steps; shots/s scaled x{scale:g}"}
    return out


def bench_gradient(args):
    """Adjoint-gradient workload: the 401x201 flat-seafloor model with 3000 steps;
    observed records from the true model, gradient of 0.5||rec - obs||^2 w.r.t.
    vp, vs, rho at the true model smoothed by a 100 m Gaussian.  Checked by central
    differences along 16 seeded relative directions (eps 1e-3) and by the FP64 host
    adjoint on the 201x101 cut."""
    import numpy as np
    import msfdm
    from msfdm import autograd as ag
    from msfdm import fwi, modelgen
    true = modelgen.preset("flat401", nt=3000)
    init = fwi.smoothed_model(true, 100.0, 5.0)
    s = msfdm.Solver(true)
    s.reset(0); s.run(true.nt); s.sync()
    obs = s.records()
    solver = msfdm.Solver(init)
    fwi.misfit_and_gradient(solver, init, obs, nsteps=init.nt)   # warm-up (allocation, graph)
    t0 = time.perf_counter()
    loss, g = fwi.misfit_and_gradient(solver, init, obs, nsteps=init.nt)
    t_grad = time.perf_counter() - t0
    t0 = time.perf_counter()
    s.reset(0); s.run(true.nt); s.sync()
    t_fwd = time.perf_counter() - t0
    vel0 = ag.velocity_from_model(init)
    names = ("vp_f", "rho_f", "vp", "vs", "rho")
    rng = np.random.default_rng(4)
    def loss_at(vel):
        media = ag.media_from_velocity(*vel)
        arrays = dict(init.arrays)
        for k, v in media.items():
            arrays[k] = np.ascontiguousarray(v, dtype="<f4")
        solver.update_media(**{k: arrays[k] for k in media})
        solver.reset(0); solver.run(init.nt); solver.sync()
        return sum(0.5 * float(np.sum((a.astype(np.float64) - b.astype(np.float64)) ** 2))
                   for a, b in zip(solver.records(), obs))
    errs = []
    eps = 1e-3
    for _ in range(args.fd_dirs):
        d = [rng.standard_normal(v.shape) * v for v in vel0]
        fd = (loss_at([v + eps * dv for v, dv in zip(vel0, d)]) - loss_at([v - eps * dv for v, dv in zip(vel0, d)])) / (2 * eps)
        an = sum(float(np.sum(g[n] * dv)) for n, dv in zip(names, d))
        errs.append(abs(an - fd) / abs(fd))
    # FP64 host adjoint on the 201x101 cut
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    cut_true = modelgen.preset("cut201")
    cut = fwi.smoothed_model(cut_true, 100.0, 5.0)
    o = pyoracle.Oracle(cut_true); o.reset(0); o.run(cut.nt)
    obs_c = o.records()
    oi = pyoracle.Oracle(cut); oi.reset(0); oi.run(cut.nt)
    res = [a - b for a, b in zip(oi.records(), obs_c)]
    go = oi.gradient(0, cut.nt, res, every=0 or 25)
    sc = msfdm.Solver(cut)
    gs = sc.gradient(0, cut.nt, [r.astype(np.float32) for r in res])
    par = {n: float(np.linalg.norm(gs[n] - go[n]) / np.linalg.norm(go[n])) for n in msfdm._lib.GRADS}
    out = {"metric": "adjoint gradient (vp, vs, rho) of the L2 record misfit, 401x201 flat seafloor, 3000 steps",
           "value": round(1.0 / t_grad, 4), "unit": "gradients/s (1 shot: checkpointed forward + adjoint)",
           "gradient_s": round(t_grad, 3), "forward_s": round(t_fwd, 3),
           "gradient_over_forward": round(t_grad / t_fwd, 2), "misfit": loss,
           "fd_check": {"directions": args.fd_dirs, "eps_rel": eps, "max_rel_err": max(errs),
                        "median_rel_err": float(np.median(errs))},
           "fp64_oracle_cut201": {"steps": cut.nt, "rel_l2_per_array": par}}
    print(json.dumps(out), flush=True)


def bench_shots(args, workload=None, steps=None, nseq=None):
    """Multi-shot gradient: n shots of one model (sources spread evenly over the
    middle 80 % of the line), all shots of a wave stepped together
    (msfdm_gradient_shots) against the same shots one at a time (msfdm_gradient),
    seeded synthetic residuals.  When --shots-steps is below the model's record
    length the timed run covers that many steps and shots/s is scaled to the
    full record."""
    import numpy as np
    import msfdm
    from msfdm import modelgen
    n = args.shots
    workload = workload or args.shots_workload
    nx = modelgen.preset(workload, nt=1).nx
    src = [int(round(x)) for x in np.linspace(0.1 * (nx - 1), 0.9 * (nx - 1), n)]
    model = modelgen.preset(workload, src_x=src, **({"nt": args.shots_record} if args.shots_record else {}))
    full = model.nt
    steps = steps or args.shots_steps or full
    s = msfdm.Solver(model)
    s.reset(0); s.run(steps); s.sync()
    nrec = s.num_records()
    rng = np.random.default_rng(5)
    res = [[rng.standard_normal((nrec, model.nx)).astype(np.float32) for _ in range(3)] for _ in range(n)]
    per_shot, every = s.gradient_shots_plan(steps)
    per_shot_full, every_full = s.gradient_shots_plan(full)
    k = min(n, nseq or args.shots_seq)
    idx = [int(i) for i in np.linspace(0, n - 1, k)]
    s.gradient_shots([0], steps, res[:1])                      # warm-ups (allocation, code objects,
    s.gradient(idx[0], steps, res[idx[0]])                     # checkpoint and segment buffers)
    # the short one-at-a-time sample first, so that it does not inherit the clock
    # state of the long batched run
    t_seq, ref = 0.0, None
    for i in idx:
        t0 = time.perf_counter()
        gi = s.gradient(i, steps, res[i])
        t_seq += time.perf_counter() - t0
        print(f"[shots] {workload}: shot {i} alone in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
        ref = gi if ref is None else {a: ref[a] + gi[a] for a in ref}
    t_seq *= n / k
    t0 = time.perf_counter()
    g, wave = s.gradient_shots(list(range(n)), steps, res)
    t_batch = time.perf_counter() - t0
    print(f"[shots] {workload}: {n} shots x {steps} steps batched in {t_batch:.1f} s "
          f"({wave} shots per wave)", file=sys.stderr, flush=True)
    gsub, _ = s.gradient_shots(idx, steps, [res[i] for i in idx])
    par = {a: float(np.linalg.norm(gsub[a] - ref[a]) / np.linalg.norm(ref[a])) for a in msfdm._lib.GRADS}
    scale = full / steps
    hbm = 288e9
    s.close()
    out = {"metric": f"multi-shot fwd+adjoint gradient shots/s ({workload}, {full}-step records)",
           "value": round(n / (t_batch * scale), 4), "unit": "shots/s",
           "sequential_value": round(n / (t_seq * scale), 4),
           "speedup_vs_sequential": round(t_seq / t_batch, 3),
           "shots": n, "wave_shots": wave, "steps_timed": steps, "every": every,
           "grid": [model.nx, model.nz], "batched_s": round(t_batch, 3), "sequential_s_est": round(t_seq, 3),
           "sequential_shots_timed": k,
           "bytes_per_shot_full_record": per_shot_full, "every_full_record": every_full,
           "shots_per_wave_in_288GB": int(0.9 * hbm // per_shot_full),
           "rel_l2_batched_vs_sum_of_single": par,
           "data": "synthetic residuals (seeded normal), generated model",
           "note": (f"timed {steps} of {full} steps; shots/s scaled x{scale:g}" if scale != 1 else "full records")}
    return out


def bench_solid(args):
    """Solid-only stepper with a traction-free top (grid.calc_type = 0): the
    2001 x 1001 body-fitted grid under a 20 m / 600 m sinusoidal topography, 5 %
    random solid (vp/vs/rho 3200/1800/2300), 20 Hz, dt = 0.25 ms, 30-cell CFS-PML,
    6000 steps; free-surface vz recorded every 4 steps.  Same JSON fields as the
    default solver line, with the FP64 oracle's solid-only stepper as the CPU
    baseline."""
    import msfdm
    from msfdm import modelgen
    model = modelgen.preset(args.solid_workload)
    s = msfdm.Solver(model)
    if args.tile_rows:
        s.tile_rows = args.tile_rows
        s.occupancy = args.occupancy
    else:
        s.autotune(64)   # tile height and occupancy policy from the library's autotuner
    s.reset(0)
    s.run(args.warmup)
    s.sync()
    t0 = time.perf_counter()
    s.run(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    cells = model.nx * model.nz
    bytes_step = s.bytes_per_step()
    reps = args.kernel_reps
    t_stage = s.time_kernel(0, reps)
    t_stage4 = s.time_kernel(1, reps)
    b_stage = s.kernel_bytes(0)
    achieved = b_stage / (t_stage * 1e-3) / 1e9
    kname, traffic = measured_traffic(args.solid_workload, stage2_kernel_names(True, s.tile_rows, s.occupancy))
    cpu = cpu_baseline(model, args.cpu_steps, args.cpu_steps_1core) if args.cpu_steps > 0 else None
    out = {
        "metric": "Gcell-updates/s (plain FP32 solid-only solver, traction-free surface, 1 GPU)",
        "value": round(cells * args.steps / wall / 1e9, 4), "unit": "Gcell-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (generated body-fitted model, seeded random solid)",
        "config": {
            "workload": f"{args.solid_workload}: {model.nx}x{model.nz} solid grid, free surface on row 3, "
                        f"npd={model.npd}, dt={model.dt}, fpeak={model.fpeak}",
            "nx": model.nx, "nz": model.nz, "time_steps": args.steps, "tile_rows": s.tile_rows,
            "occupancy_policy": s.occupancy,
            "hbm_GBps_per_step_alg": round(bytes_step / (wall / args.steps) / 1e9, 1),
            "hbm_frac_per_step_alg": round(bytes_step / (wall / args.steps) / 1e9 / HBM_PEAK_GBPS, 4),
            "bytes_per_step_alg": bytes_step,
            "kernel_ms": {"free_stage_kernel_stage2": round(t_stage, 5),
                          "free_stage_kernel_stage4": round(t_stage4, 5)},
            "stage4_GBps_alg": round(s.kernel_bytes(1) / (t_stage4 * 1e-3) / 1e9, 1),
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic,
                     "kernel": f"{kname or 'free_stage_kernel'} (RK stage 2, F direction)",
                     "alg_bytes_per_launch": b_stage},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def bench_mas(args):
    """Phase-shift dispersion image (tools/MAS.py:50-98) of a record-sized gather:
    1500 samples at 1 kHz x 1941 traces (the 2001-column record without its
    30-column PML strips), 2.5 m spacing, test velocities 100..3000 m/s every
    5 m/s, on the GPU in FP64, against the numpy restatement timed on a bounded
    sample (the first 128 samples, the cost scaled by the frequency count)."""
    import numpy as np
    from msfdm import records
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import mas
    rng = np.random.default_rng(11)
    nt, ntr, fs = 1500, 1941, 1000.0
    g = rng.standard_normal((nt, ntr)).astype(np.float32)
    x = (np.arange(ntr) - 970) * 2.5
    records.dispersion_image(g[:64], x, fs, 100.0, 3000.0, 5.0)      # warm-up (code objects)
    t0 = time.perf_counter()
    f, vel, amp = records.dispersion_image(g, x, fs, 100.0, 3000.0, 5.0)
    t_gpu = time.perf_counter() - t0
    nf = f.size
    ns = 128
    t0 = time.perf_counter()
    mas.dispersion_image(g[:ns].astype(np.float64), x, fs, vel)
    t_cpu = (time.perf_counter() - t0) * nf / (ns // 2 + 1)
    out = {"metric": "dispersion image (phase-shift, FP64) of a 1500 x 1941 gather, 581 test velocities",
           "value": round(t_gpu * 1e3, 2), "unit": "ms (GPU, host copies included)", "higher_is_better": False,
           "stack_terms_per_s": round(nf * vel.size * ntr / t_gpu / 1e9, 2),
           "cpu_baseline": {"value": round(t_cpu * 1e3, 1), "unit": "ms", "cores": 1, "kind": "port",
                            "sample": f"numpy restatement on the first {ns} samples ({ns // 2 + 1} of {nf} "
                                      f"frequencies), scaled by the frequency count"},
           "speedup_vs_cpu": round(t_cpu / t_gpu, 1)}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="seafloor2001")
    ap.add_argument("--cpu-steps", type=int, default=150, help="steps of the CPU-baseline sample (0 = skip)")
    ap.add_argument("--cpu-steps-1core", type=int, default=10, help="steps of the one-thread CPU sample (0 = skip)")
    ap.add_argument("--no-shots", action="store_true",
                    help="default mode: skip the batched-gradient and batched-forward shots/s legs")
    ap.add_argument("--kernel-reps", type=int, default=100)
    ap.add_argument("--tile-rows", type=int, default=0, choices=[0, 8, 16],
                    help="tile height of the stage kernels (0 = the library's autotuner)")
    ap.add_argument("--occupancy", type=int, default=0, choices=[0, 1],
                    help="occupancy policy of the stage kernels when --tile-rows is given")
    ap.add_argument("--mode", default="solver", choices=["solver", "batch", "gradient", "shots", "solid", "mas"])
    ap.add_argument("--solid-workload", default="solid_topo2001")
    ap.add_argument("--shots", type=int, default=32)
    ap.add_argument("--shots-workload", default="flat401")
    ap.add_argument("--shots-steps", type=int, default=0, help="steps per shot timed (0 = the full record)")
    ap.add_argument("--shots-record", type=int, default=0, help="record length in steps (0 = the workload's)")
    ap.add_argument("--shots-seq", type=int, default=4, help="shots timed one at a time for the comparison")
    ap.add_argument("--fd-dirs", type=int, default=16)
    ap.add_argument("--batch-shots", type=int, default=64)
    ap.add_argument("--batch-steps", type=int, default=400)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (x5) to calibrate HBM byte counters")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(shots shard across GPUs with one gradient all-reduce; not built here)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4}))
        return
    if args.mode == "batch":
        print(json.dumps(bench_batch(args)), flush=True)
        return
    if args.mode == "gradient":
        bench_gradient(args)
        return
    if args.mode == "shots":
        print(json.dumps(bench_shots(args)), flush=True)
        return
    if args.mode == "solid":
        bench_solid(args)
        return
    if args.mode == "mas":
        bench_mas(args)
        return

    import msfdm
    from msfdm import modelgen

    model = modelgen.preset(args.workload)
    s = msfdm.Solver(model)
    if args.tile_rows:
        s.tile_rows = args.tile_rows
        s.occupancy = args.occupancy
    else:
        s.autotune(64)   # tile height (8 or 16 rows) and occupancy policy from the library's autotuner
    s.reset(0)
    s.run(args.warmup)
    s.sync()
    # the hipGraph of the time loop is instantiated by the warm-up steps
    t0 = time.perf_counter()
    s.run(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    cells = model.nx * model.nz
    value = cells * args.steps / wall / 1e9
    bytes_step = s.bytes_per_step()

    # dominant kernel: the RK stage kernel over both media.  Its duration is
    # taken inside the step sequence: the same shot again (reset, warm-up), the
    # same K steps with direct launches, every launch timed by start / stop
    # events its own dispatch stamps (no launch gaps, no repeat of one launch
    # over cache-resident inputs).
    # Pass 1 times the stage-2 launches only (every other launch as in a plain
    # run); pass 2 times every launch for the per-stage breakdown.
    s.reset(0)
    s.run(args.warmup)
    s.sync()
    p2 = s.profile_run(args.steps, classes=[1])
    t_stage = p2["kernel_ms"][1] / max(p2["launches"][1], 1)
    s.reset(0)
    s.run(args.warmup)
    s.sync()
    prof = s.profile_run(args.steps)
    kms, nl = prof["kernel_ms"], prof["launches"]
    stage_ms = [kms[i] / max(nl[i], 1) for i in range(4)]
    b_stage, b_stage4 = s.kernel_bytes(0), s.kernel_bytes(1)
    tile_rows, occupancy = s.tile_rows, s.occupancy
    achieved = b_stage / (t_stage * 1e-3) / 1e9
    # the same launch repeated back to back (inputs partly cache resident)
    reps = args.kernel_reps
    t_rep2 = s.time_kernel(0, reps)
    t_rep4 = s.time_kernel(1, reps)

    kname, traffic = measured_traffic(args.workload, stage2_kernel_names(False, tile_rows, occupancy))
    # what a plain streaming kernel reaches on this part: 1 GB streams (HBM),
    # and streams the size of one stage-2 launch's bytes (partly cache resident)
    stream = {"copy_GBps": round(1e3 * msfdm.stream_probe(1000, 1), 1),
              "read3_write1_GBps": round(1e3 * msfdm.stream_probe(1000, 3), 1),
              "read3_write1_stage2_sized_GBps": round(1e3 * msfdm.stream_probe(b_stage / 4e6, 3, reps=200), 1),
              "stream_MB": 1000, "stage2_sized_stream_MB": round(b_stage / 4e6, 1)}

    cpu = None
    if args.cpu_steps > 0:
        cpu = cpu_baseline(model, args.cpu_steps, args.cpu_steps_1core)
    s.close()
    shots = None
    if not args.no_shots:
        shots = {
            "gradient_flat401_32": bench_shots(args, workload="flat401", steps=2000, nseq=4),
            "gradient_twoscale4001_32": bench_shots(args, workload="twoscale4001", steps=200, nseq=2),
            "forward_batch64": bench_batch(args, nsteps=200),
        }

    out = {
        "metric": "Gcell-updates/s (plain FP32 fluid-solid solver, 1 GPU)",
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (generated body-fitted model, seeded random solid)",
        "config": {
            "workload": f"{args.workload}: {model.nx}x{model.nz} grid (fluid {model.nz_fluid} + solid "
                        f"{model.nz_solid} rows), npd={model.npd}, dt={model.dt}, fpeak={model.fpeak}",
            "nx": model.nx, "nz": model.nz, "time_steps": args.steps, "tile_rows": tile_rows,
            "occupancy_policy": occupancy,
            "hbm_GBps_per_step_alg": round(bytes_step / (wall / args.steps) / 1e9, 1),
            "hbm_frac_per_step_alg": round(bytes_step / (wall / args.steps) / 1e9 / HBM_PEAK_GBPS, 4),
            "bytes_per_step_alg": bytes_step,
            "kernel_ms_in_sequence": {"stage1": round(stage_ms[0], 5), "stage2": round(stage_ms[1], 5),
                                      "stage3": round(stage_ms[2], 5), "stage4": round(stage_ms[3], 5),
                                      "iface_final": round(kms[4] / max(nl[4], 1), 5)},
            "launches_in_sequence": nl,
            "ms_per_step_profiled": round(prof["wall_ms"] / args.steps, 5),
            "kernel_ms_isolated_repeat": {"stage2": round(t_rep2, 5), "stage4": round(t_rep4, 5)},
            "stage4_GBps_alg": round(b_stage4 / (stage_ms[3] * 1e-3) / 1e9, 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic,
            "kernel": f"{kname or 'stage_kernel'} (RK stage 2, both directions, fluid + solid tiles)",
            "alg_bytes_per_launch": b_stage,
            "timing": f"mean of the {p2['launches'][1]} stage-2 launches of a {args.steps}-step run "
                      f"(only those launches timed, by events their own dispatch stamps)",
            "frac_of_streaming_kernel": round(achieved / stream["read3_write1_GBps"], 4),
        },
        "hbm_stream_probe": stream,
        "cpu_baseline": cpu,
        "shots": shots,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
