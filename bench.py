"""Throughput of the 3D open-Hockney space-charge step on one MI355X.

Main line (the plain FP64 GPU configuration): a proton bunch at 0.8 GeV
kinetic energy, 4e6 macroparticles, population 1e11, 6D Gaussian matched to
an O-F-O-D cell (0.5 m drifts, 0.2 m quads k1 = +-0.6 m^-2, rms emittance
1 mm mrad), sigma_z = 10 mm, sigma_dp/p = 1e-4, seed 1; 4-cell channel with
10 space-charge kicks per cell (Split_operator_stepper, 40 steps per turn);
grid 64x64x128 (doubled 128x128x256), adaptive domain, linear Green's
function, FP64 throughout.

A step = one space-charge step (the hot path: bunch extent, cell sort and
reorder of the bunch every sort_period steps, deposit, Green's spectrum, FFT
convolution, E field, kick) of a bunch that is tracked through the channel:
the half-step drift/quadrupole maps run between the kicks as their own
kernels.  value = macroparticles x steps / (sum of the
HIP-event times of the timed steps, first kernel of the step to last, on the
stream they run on) = particle-kicks/s.  The wall time of the whole timed
loop, maps included, is reported beside it as the tracking rate.

Also reported: the roofline of the dominant kernel (algorithmic bytes per
launch over its HIP-event-timed duration, MI355X HBM 8 TB/s) and the CPU
restatement of the reference operator (oracle/, OpenMP on every host core
granted to the process) timed on a bounded sample of the same workload.

Sub-records of the same line (default run, no --config):
  precision_form  1.6e7 particles, 128x128x256, FP32 deposit / FFT / field,
                  FP64 particles: throughput, roofline, first-kick field
                  rel-L2 against the oracle and 100-step emittance drift
                  against the FP64 form;
  served_form     the 256-job Poisson stream: saturated throughput, its ratio
                  to the plain FP64 form on the same jobs, step and job
                  p50 / p99 latency of the real-time stream;
  drift_case      1e5 particles through a 1 m drift in 10 kicks on the GPU
                  and on the CPU restatement beside it.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBPS = 8000.0

CONFIGS = {
    # FP64 form: 4e6 particles matched to the cell (rms emittance 1 mm mrad),
    # sigma_z 10 mm, population 1e11, seed 1, grid 64x64x128
    "fodo_fp64_4M_64x64x128": dict(n=4_000_000, grid=(64, 64, 128), pop=1e11, seed=1,
                                   sigma_z=0.010, sigma_dpop=1e-4, emit=1e-6, precision=0),
    # precision form: 1.6e7 particles, sigma_x = sigma_y = 2 mm at the cell
    # entrance (matched), sigma_z 20 mm, population 2e11, seed 2, grid
    # 128x128x256 with FP32 deposit / FFT / field
    "fodo_fp32grid_16M_128x128x256": dict(n=16_000_000, grid=(128, 128, 256), pop=2e11, seed=2,
                                          sigma_z=0.020, sigma_dpop=1e-4, sigma_xy=0.002,
                                          precision=1),
}
KINETIC_ENERGY = 0.8  # GeV, proton


def make_workload(sc, S, config, precision=None, sort_period=None):
    """Bunch + operator + propagator of a named workload (FODO channel of 4
    cells, 10 kicks per cell); precision overrides the config's form."""
    import numpy as np
    cfg = CONFIGS[config]
    gamma = (sc.PROTON_MASS + KINETIC_ENERGY) / sc.PROTON_MASS
    beta, _ = sc.kinematics(sc.PROTON_MASS, gamma)
    cell = S.fodo_channel(1)
    if "sigma_xy" in cfg:
        chol = S.matched_chol_sigma(cell, cfg["sigma_xy"], cfg["sigma_xy"],
                                    cfg["sigma_z"] / beta, cfg["sigma_dpop"])
    else:
        chol = S.matched_chol(cell, cfg["emit"], cfg["sigma_z"] / beta, cfg["sigma_dpop"])
    bunch = sc.Bunch(cfg["n"], cfg["pop"], gamma=gamma)
    bunch.populate_gaussian(np.zeros(6), chol, seed=cfg["seed"])
    opts = sc.default_options(*cfg["grid"])
    opts.precision = cfg["precision"] if precision is None else precision
    if sort_period is not None:
        opts.sort_period = sort_period
    op = sc.SpaceCharge3dOpenHockney(opts)
    prop = S.Propagator(S.fodo_channel(4), 40, bunch, op)
    return bunch, opts, op, prop


def algorithmic_bytes(phase, n, g, w):
    """Algorithmic HBM bytes of one launch (unit: one particle or one node):
    kick     73 B/particle (x,y,cdt,xp,yp,dpop read, xp,yp,dpop written, mask)
             + 3 field words per node read once;
    deposit  25 B/particle (x,y,cdt, mask) + 1 density word per node written."""
    G = g[0] * g[1] * g[2]
    if phase == "sc3d_kick":
        return 73 * n + 3 * w * G
    if phase == "sc3d_local_rho":
        return 25 * n + w * G
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48,
                    help="timed steps (a multiple of the default sort period, 24)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default=None, choices=sorted(CONFIGS),
                    help="time only this workload (default: the FP64 workload as the main line, "
                         "with the precision, served and drift-case sub-records)")
    ap.add_argument("--sort-period", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-seconds", type=float, default=10.0,
                    help="CPU baseline: time space-charge steps until this much CPU time")
    ap.add_argument("--accuracy", action="store_true",
                    help="also state field rel-L2 and rms-emittance drift of the FP32-grid form "
                         "against the FP64 form and the oracle on this workload")
    ap.add_argument("--served", action="store_true",
                    help="measure the served form on the Poisson job stream instead (see "
                         "served_bench)")
    ap.add_argument("--jobs", type=int, default=256, help="served: jobs in the stream")
    ap.add_argument("--branches", type=int, default=8,
                    help="served: streams the per-job chains are spread over")
    ap.add_argument("--admit-interval", type=int, default=8,
                    help="served: iterations between admissions (graph rebuilds)")
    ap.add_argument("--serve-sort-period", type=int, default=24,
                    help="served: iterations between reorders of the resident bunches")
    ap.add_argument("--no-graphs", action="store_true",
                    help="served: launch every iteration's kernels directly")
    ap.add_argument("--form", choices=("served", "plain64", "plain32"), default=None,
                    help="run only one form on the job stream, all jobs at once (served) or "
                         "one after another through the operator API (plain64 / plain32), and "
                         "print its wall time: the unit of the per-form kernel-time split")
    ap.add_argument("--long", action="store_true",
                    help="long tracking of the precision form against the FP64 form on the "
                         "FODO ring (see long_tracking)")
    ap.add_argument("--turns", type=int, default=200, help="long: turns")
    ap.add_argument("--long-particles", type=int, default=10_000_000, help="long: particles")
    ap.add_argument("--no-precision", action="store_true",
                    help="default run: skip the precision-form sub-record")
    ap.add_argument("--no-served", action="store_true",
                    help="default run: skip the served-form sub-record")
    ap.add_argument("--no-c1", action="store_true",
                    help="default run: skip the drift-case (1e5 particles, CPU beside GPU) record")
    ap.add_argument("--copy-probe", action="store_true",
                    help="first time a 1 GiB device-to-device copy (counter calibration and "
                         "achievable HBM bandwidth)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: there is no multi-GPU path (the step would "
              "shard by particles, with one all-reduce of the density octant per step)",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import synergia2_amd as sc
    from synergia2_amd import stepper as S

    if args.form:
        form_run(args, sc, S)
        return
    if args.served:
        served_bench(args, sc, S)
        return
    if args.long:
        long_tracking(args, sc, S)
        return

    copy_gbps = copy_probe(torch) if args.copy_probe else None

    main_cfg = args.config or "fodo_fp64_4M_64x64x128"
    rec, bunch, opts, prop = timed_sc_run(args, sc, S, main_cfg, args.steps, args.warmup)
    rec["roofline"]["copy_probe_GBps"] = copy_gbps
    cfg = CONFIGS[main_cfg]

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(bunch, prop, opts, args.cpu_seconds, sc, S)
    del bunch, prop
    torch.cuda.empty_cache()

    acc = None
    if args.accuracy:
        acc = accuracy_report(sc, S, main_cfg, 100)

    out = {
        "metric": "particle-kicks/s (3D Hockney SC step)",
        "value": rec["value"],
        "unit": "particle-kicks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": rec["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if cfg["precision"] == 0 else "f32-grid/f64-particles",
        "data": "synthetic (seeded 6D Gaussian matched to the FODO cell)",
        "config": rec["config"],
        "roofline": rec["roofline"],
        "cpu_baseline": cpu,
    }
    if acc is not None:
        out["accuracy"] = acc
    # the other forms of the headline metric, measured in the same run when
    # the default workload is the main line: the precision form on its own
    # workload (with its accuracy against the oracle), the served form on the
    # job stream, and the small drift case on the CPU restatement beside the GPU
    if args.config is None:
        if not args.no_precision:
            out["precision_form"] = precision_record(args, sc, S)
            torch.cuda.empty_cache()
        if not args.no_served:
            out["served_form"] = served_record(args, sc, S)
            torch.cuda.empty_cache()
        if not args.no_c1:
            out["drift_case"] = drift_case_record(sc, S)
            torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def timed_sc_run(args, sc, S, config, steps, warmup):
    """Warm up, then time `steps` split-operator steps of a named workload:
    the space-charge applies by one HIP event pair per apply on the apply
    stream (the value), the whole loop (maps included) by the wall clock;
    then the same steps of a fresh copy of the workload with events at every
    phase boundary (the per-phase split and the per-kernel rooflines; those
    events add a few us each).  Returns the record, and the bunch, options
    and propagator for the CPU leg."""
    import torch
    cfg = CONFIGS[config]
    n, grid, precision = cfg["n"], cfg["grid"], cfg["precision"]
    bunch, opts, op, prop = make_workload(sc, S, config, sort_period=args.sort_period)
    print(f"# {config}: n={n} grid={grid} warmup={warmup} steps={steps} "
          f"sort_period={opts.sort_period}", file=sys.stderr, flush=True)
    prop.run(warmup, sync=True)
    torch.cuda.synchronize()
    op.set_timing(sc.TIMING_APPLY)
    op.apply_timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prop.run(steps, fuse=False)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    op.sync()
    wall = t1 - t0
    apply_ms, napplies = op.apply_timing(reset=True)
    assert napplies == steps
    op.set_timing(sc.TIMING_OFF)
    del op, prop, bunch
    torch.cuda.empty_cache()
    sc_s = apply_ms / 1e3
    ms_per_step = 1e3 * sc_s / steps
    value = n * steps / sc_s
    print(f"# SC steps {sc_s:.4f}s -> {ms_per_step:.4f} ms/step, {value:.4e} particle-kicks/s; "
          f"tracking wall {wall:.4f}s incl. maps -> {n * steps / wall:.4e} /s",
          file=sys.stderr, flush=True)
    # the same steps again, phases timed
    bunch, opts, op, prop = make_workload(sc, S, config, sort_period=args.sort_period)
    prop.run(warmup, sync=True)
    torch.cuda.synchronize()
    op.set_timing(sc.TIMING_PHASES)
    op.timing(reset=True)
    prop.run(steps, fuse=False)
    op.sync()
    phases, napplies = op.timing(reset=True)
    assert napplies == steps
    phase_ms = sum(phases.values()) / steps
    print(f"# phases (events at every boundary, {phase_ms:.4f} ms/step in all):",
          file=sys.stderr, flush=True)
    for k, v in phases.items():
        print(f"#   {k:16s} {v / steps * 1e3:9.2f} us/step", file=sys.stderr)
    w = 8 if precision == 0 else 4
    traffic = measured_traffic(config)
    lines = []
    for phase, kname in (("sc3d_kick", "k_kick_tile"), ("sc3d_local_rho", "k_deposit_tile")):
        per_launch_s = phases[phase] / steps / 1e3
        B = algorithmic_bytes(phase, n, grid, w)
        achieved = B / per_launch_s / 1e9
        lines.append({
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": traffic.get(kname),
            "algorithmic_bytes_per_launch": B,
            "launch_us": round(per_launch_s * 1e6, 2),
        })
    # the dominant kernel: the longer of the two particle passes
    roofline = dict(max(lines, key=lambda r: r["launch_us"]))
    roofline["particle_kernels"] = lines
    rec = {
        "value": value,
        "ms_per_step": ms_per_step,
        "config": {
            "workload": config.replace("_", " ") + ", FODO 4 cells, 10 kicks/cell",
            "particles": n, "grid": list(grid), "doubled_grid": [2 * x for x in grid],
            "green_fn": "linear", "domain": "adaptive", "sort_period": opts.sort_period,
            "phase_us_per_step": {k: round(v / steps * 1e3, 2) for k, v in phases.items()},
            "phase_sum_ms_per_step": round(phase_ms, 4),
            "tracking_ms_per_step_incl_maps": round(1e3 * wall / steps, 4),
            "tracking_particle_steps_per_s": n * steps / wall,
        },
        "roofline": roofline,
    }
    op.set_timing(sc.TIMING_OFF)
    # the same number of steps again with the maps after each kick fused into
    # the kick kernel (one particle pass for kick + maps, as the served form
    # runs them): wall time of the tracking loop
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prop.run(steps, fuse=True)
    torch.cuda.synchronize()
    wall_f = time.perf_counter() - t0
    rec["config"]["tracking_ms_per_step_fused_maps"] = round(1e3 * wall_f / steps, 4)
    rec["config"]["tracking_particle_steps_per_s_fused_maps"] = n * steps / wall_f
    print(f"# tracking with the maps fused into the kick: {1e3 * wall_f / steps:.4f} ms/step -> "
          f"{n * steps / wall_f:.4e} /s", file=sys.stderr, flush=True)
    del op
    return rec, bunch, opts, prop


def precision_record(args, sc, S):
    """The precision form (FP32 deposit / FFT / field, FP64 particles and
    maps) on its own workload: 1.6e7 particles, 128x128x256, seed 2 --
    throughput and roofline as the main line, then its accuracy: field
    rel-L2 of the first kick against the oracle and 100-step rms-emittance
    drift against the FP64 form."""
    import torch
    config = "fodo_fp32grid_16M_128x128x256"
    rec, bunch, opts, prop = timed_sc_run(args, sc, S, config, args.steps, args.warmup)
    del bunch, prop
    torch.cuda.empty_cache()
    rec["accuracy"] = accuracy_report(sc, S, config, 100)
    rec["dtype"] = "f32-grid/f64-particles"
    return rec

def long_tracking(args, sc, S):
    """Long tracking: 1e7 macroparticles, 90 % 6D Gaussian matched to the
    cell (sigma_x = sigma_y = 2 mm, sigma_z = 20 mm, sigma_dp/p = 1e-4) + 10 %
    uniform in the transverse 4D ellipsoid at 3 sigma (seed 3; population
    2e11), through a ring of 32 O-F-O-D cells with one kick per element (128
    kicks per turn, Split_operator_stepper_elements(1)), grid 128x128x256.
    The precision form (FP32 grid) and the FP64 form track the same bunch;
    after every turn both report rms emittances and the fraction of
    particles beyond 4 initial sigma (x, y).  The FP64 form stands in for
    the oracle here (it matches the oracle within 1e-12 per step; 25,600
    oracle steps of 1e7 particles would take hours on the host)."""
    import numpy as np
    import torch

    n, turns = args.long_particles, args.turns
    gamma = (sc.PROTON_MASS + KINETIC_ENERGY) / sc.PROTON_MASS
    beta, _ = sc.kinematics(sc.PROTON_MASS, gamma)
    ring = S.fodo_channel(32)
    steps = S.split_operator_steps_elements(ring, 1)
    chol = S.matched_chol_sigma(S.fodo_channel(1), 2e-3, 2e-3, 0.020 / beta, 1e-4)
    parts = S.halo_mix_parts(n, chol, seed=3)
    forms = {}
    for prec in (sc.FP32_GRID, sc.FP64):
        b = sc.Bunch(n, 2e11, gamma=gamma)
        b.set_host(parts)
        opts = sc.default_options(128, 128, 256)
        opts.precision = prec
        op = sc.SpaceCharge3dOpenHockney(opts)
        forms[prec] = (b, op, S.Propagator(ring, len(steps), b, op, steps=steps))
    del parts
    mean0, mom0 = forms[sc.FP64][1].moments(forms[sc.FP64][0])
    width0 = np.sqrt(np.diag(mom0))
    e0 = sc.emittances(mom0)
    hist = {sc.FP32_GRID: [], sc.FP64: []}
    tsec = {sc.FP32_GRID: 0.0, sc.FP64: 0.0}
    print(f"# long: {n} particles, {turns} turns x {len(steps)} kicks, emit0 {e0}",
          file=sys.stderr, flush=True)
    for t in range(1, turns + 1):
        for prec, (b, op, prop) in forms.items():
            t0 = time.perf_counter()
            prop.run(len(steps), sync=True)
            tsec[prec] += time.perf_counter() - t0
            mean, mom2 = op.moments(b)
            tail = op.tail_fraction(b, mean, width0, 4.0)
            hist[prec].append((sc.emittances(mom2), [float(tail[0]), float(tail[2])]))
        if t % 10 == 0 or t == turns:
            e32, tl32 = hist[sc.FP32_GRID][-1]
            e64, tl64 = hist[sc.FP64][-1]
            print(f"# turn {t}: emit32/emit64-1 = "
                  f"{[f'{e32[k] / e64[k] - 1:.2e}' for k in range(3)]}, "
                  f"tail x,y fp32 {tl32} fp64 {tl64}, "
                  f"s/turn fp32 {tsec[sc.FP32_GRID] / t:.3f} fp64 {tsec[sc.FP64] / t:.3f}",
                  file=sys.stderr, flush=True)
    drift = [max(abs(hist[sc.FP32_GRID][i][0][k] / hist[sc.FP64][i][0][k] - 1)
                 for i in range(turns)) for k in range(3)]
    tail_diff = [max(abs(hist[sc.FP32_GRID][i][1][k] - hist[sc.FP64][i][1][k])
                     for i in range(turns)) for k in range(2)]
    out = {
        "metric": "long tracking: precision form vs FP64 form, per turn",
        "config": {"workload": "32-cell FODO ring, 1 kick per element, 90% Gaussian + 10% "
                               "3-sigma 4D ellipsoid halo, seed 3",
                   "particles": n, "turns": turns, "kicks_per_turn": len(steps),
                   "grid": [128, 128, 256]},
        "emittance_rel_drift_max": drift,
        "tail4sigma_abs_diff_max": tail_diff,
        "final": {"emit_fp32": hist[sc.FP32_GRID][-1][0], "emit_fp64": hist[sc.FP64][-1][0],
                  "tail_fp32": hist[sc.FP32_GRID][-1][1], "tail_fp64": hist[sc.FP64][-1][1]},
        "emit0": e0,
        "ms_per_step": {"fp32_grid": 1e3 * tsec[sc.FP32_GRID] / (turns * len(steps)),
                        "fp64": 1e3 * tsec[sc.FP64] / (turns * len(steps))},
        "per_turn": [{"turn": i + 1, "emit_fp32": hist[sc.FP32_GRID][i][0],
                      "emit_fp64": hist[sc.FP64][i][0], "tail_fp32": hist[sc.FP32_GRID][i][1],
                      "tail_fp64": hist[sc.FP64][i][1]} for i in range(turns)],
    }
    print(json.dumps(out), flush=True)


def served_record(args, sc, S):
    """Served form on the job stream: Poisson arrivals (mean 20 jobs/s) of
    `--jobs` FODO-channel jobs (seed 4; particles log-uniform in [1e4, 4e6];
    grid from {32x32x64, 64x64x128, 128x128x256}; steps uniform in [20, 200];
    sigma_x, sigma_y uniform in [0.5, 3] mm (matched), sigma_z in [5, 30] mm),
    every job FP32 grid / FP64 particles.
      stream:    jobs submitted at their arrival times (real time), server
                 iterating; per-step (iteration) and per-job (submit to done)
                 latency percentiles; throughput = particle-steps / makespan
      saturated: the same jobs all submitted at t = 0; throughput =
                 particle-steps / wall time (value)
      fp64_plain: the same jobs one after another through the operator API
                 in the plain FP64 form (Propagator, fused maps); throughput
                 = particle-steps / wall time."""
    import numpy as np
    import torch
    from synergia2_amd import serve as SV

    stream = SV.job_stream(count=args.jobs, seed=4)
    total = sum(kw["n"] * kw["steps"] for _, kw in stream)
    print(f"# served: {len(stream)} jobs, {total:.4e} particle-steps, last arrival "
          f"{stream[-1][0]:.2f} s", file=sys.stderr, flush=True)

    def drive(real_time):
        srv = SV.Server(server_options(SV, args))
        pending = list(stream)
        ids = []
        t0 = time.perf_counter()
        while True:
            now = time.perf_counter() - t0
            while pending and (not real_time or pending[0][0] <= now):
                ids.append(srv.submit(SV.Job(**pending.pop(0)[1])))
            res, wait = srv.iterate()
            if res == 0 and wait == 0:
                if not pending:
                    break
                time.sleep(max(0.0, min(0.005, pending[0][0] - (time.perf_counter() - t0))))
        wall = time.perf_counter() - t0
        st = srv.stats()
        lat = [srv.result(i).t_done - srv.result(i).t_submit for i in ids]
        srv.close()
        torch.cuda.empty_cache()
        return wall, st, lat

    wall_s, st_s, lat_s = drive(True)
    print(f"# stream: wall {wall_s:.2f} s, {st_s['iterations']} iterations, "
          f"{st_s['graph_builds']} graph builds, step p50/p99 {st_s['iter_ms_p50']:.3f}/"
          f"{st_s['iter_ms_p99']:.3f} ms, job p50/p99 {np.percentile(lat_s, 50):.3f}/"
          f"{np.percentile(lat_s, 99):.3f} s", file=sys.stderr, flush=True)
    wall_x, st_x, lat_x = drive(False)
    print(f"# saturated: wall {wall_x:.2f} s -> {total / wall_x:.4e} particle-steps/s, "
          f"{st_x['iterations']} iterations, {st_x['graph_builds']} graph builds",
          file=sys.stderr, flush=True)

    # the plain FP64 form on the same jobs, one at a time
    t0 = time.perf_counter()
    for _, kw in stream:
        job = SV.Job(**kw)
        b = sc.Bunch(job.n, job.real_num, gamma=job.gamma)
        b.populate_gaussian(np.zeros(6), job.chol, seed=job.seed)
        op = sc.SpaceCharge3dOpenHockney(sc.default_options(*job.shape))
        S.Propagator(job.lattice, job.nsteps, b, op).run(job.steps, sync=True)
        op.close()
        del b
    torch.cuda.synchronize()
    wall_64 = time.perf_counter() - t0
    print(f"# fp64 plain: wall {wall_64:.2f} s -> {total / wall_64:.4e} particle-steps/s",
          file=sys.stderr, flush=True)
    return {
        "metric": "particle-kicks/s (served FP32-grid form, Poisson job mix)",
        "value": total / wall_x,
        "unit": "particle-kicks/s",
        "dtype": "f32-grid/f64-particles",
        "workload": f"{len(stream)} FODO jobs, Poisson 20/s, seed 4",
        "particle_steps": total,
        "vs_fp64_plain": (total / wall_x) / (total / wall_64),
        "saturated": {"wall_s": wall_x, "iterations": st_x["iterations"],
                      "graph_builds": st_x["graph_builds"],
                      "step_ms_p50": st_x["iter_ms_p50"], "step_ms_p99": st_x["iter_ms_p99"]},
        "stream": {"wall_s": wall_s, "throughput": total / wall_s,
                   "iterations": st_s["iterations"], "graph_builds": st_s["graph_builds"],
                   "step_ms_p50": st_s["iter_ms_p50"], "step_ms_p99": st_s["iter_ms_p99"],
                   "job_s_p50": float(np.percentile(lat_s, 50)),
                   "job_s_p99": float(np.percentile(lat_s, 99))},
        "fp64_plain": {"wall_s": wall_64, "throughput": total / wall_64},
    }


def server_options(SV, args):
    return SV.options(branches=args.branches, admit_interval=args.admit_interval,
                      use_graphs=not args.no_graphs, sort_period=args.serve_sort_period)


def form_run(args, sc, S):
    """One form alone on the first `--jobs` jobs of the stream (seed 4): the
    served form with every job submitted at t = 0, or the plain operator
    form job after job (FP64 or FP32 grid).  Under rocprofv3 --kernel-trace
    this gives each form's GPU time per kernel on the same job mix."""
    import numpy as np
    import torch
    from synergia2_amd import serve as SV
    stream = SV.job_stream(count=args.jobs, seed=4)
    total = sum(kw["n"] * kw["steps"] for _, kw in stream)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if args.form == "served":
        srv = SV.Server(server_options(SV, args))
        for _, kw in stream:
            srv.submit(SV.Job(**kw))
        while True:
            res, wait = srv.iterate()
            if res == 0 and wait == 0:
                break
        srv.close()
    else:
        prec = sc.FP64 if args.form == "plain64" else sc.FP32_GRID
        for _, kw in stream:
            job = SV.Job(**kw)
            b = sc.Bunch(job.n, job.real_num, gamma=job.gamma)
            b.populate_gaussian(np.zeros(6), job.chol, seed=job.seed)
            o = sc.default_options(*job.shape)
            o.precision = prec
            op = sc.SpaceCharge3dOpenHockney(o)
            S.Propagator(job.lattice, job.nsteps, b, op).run(job.steps, sync=True)
            op.close()
            del b
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    print(json.dumps({"form": args.form, "jobs": len(stream), "particle_steps": total,
                      "wall_s": wall, "particle_steps_per_s": total / wall}), flush=True)


def served_bench(args, sc, S):
    """The served form alone, as one JSON line of its own."""
    rec = served_record(args, sc, S)
    out = {"metric": rec["metric"], "value": rec["value"], "unit": rec["unit"], "n_gpus": 1,
           "higher_is_better": True, "dtype": rec["dtype"],
           "data": "synthetic (seeded job stream)",
           "config": {"workload": rec["workload"], "particle_steps": rec["particle_steps"]}}
    out.update({k: rec[k] for k in ("saturated", "stream", "fp64_plain", "vs_fp64_plain")})
    print(json.dumps(out), flush=True)


def drift_case_record(sc, S):
    """The small drift case: 1e5 particles (sigma 1 mm, 1e-4, 1 mm, 1e-4,
    10 mm / beta, 1e-4; population 5e10, seed 0) through a 1 m drift in 10
    split-operator kicks on 32x32x64, on the GPU (FP64) and on the CPU
    restatement of the reference operator beside it (all granted cores),
    with the largest coordinate difference after the 10 kicks."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    n = 100000
    gamma = (sc.PROTON_MASS + KINETIC_ENERGY) / sc.PROTON_MASS
    beta, _ = sc.kinematics(sc.PROTON_MASS, gamma)
    sig = (1e-3, 1e-4, 1e-3, 1e-4, 0.010 / beta, 1e-4)
    b = sc.Bunch(n, 5e10, gamma=gamma)
    b.populate_gaussian(np.zeros(6), sc.diag_chol(sig), seed=0)
    p0, m0 = b.host()
    op = sc.SpaceCharge3dOpenHockney(sc.default_options(32, 32, 64))
    prop = S.Propagator([(sc.MAP_DRIFT, 1.0, 0.0)], 10, b, op)
    prop.run(10, sync=True)  # warm-up pass on a copy of the state below
    b.set_host(p0, m0)
    prop.i = 0
    op.close()
    op = sc.SpaceCharge3dOpenHockney(sc.default_options(32, 32, 64))
    prop = S.Propagator([(sc.MAP_DRIFT, 1.0, 0.0)], 10, b, op)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    prop.run(10, sync=True)
    gpu_s = time.perf_counter() - t0
    pg, _ = b.host()
    po = np.ascontiguousarray(p0.copy())
    mo = np.ascontiguousarray(m0.copy())
    ob = O.make_bunch(po, mo, n, 5e10, 1, b.mass, b.gamma, b.beta, b.p)
    cores = host_cores()
    cpu_s = 0.0
    for k in range(10):
        first, second, length = prop.steps[k]
        t0 = time.perf_counter()
        O.apply_maps(ob, [(kd, L, k1, 4, O.reference_cdt(kd, L, k1, 4, b.p, b.mass))
                          for kd, L, k1 in first])
        rc, _ = O.apply(O.make_options((32, 32, 64), reorder=(k == 0)), ob,
                        length / (b.beta * sc.C_LIGHT), nthreads=cores["threads"])
        assert rc == 0, O.last_error()
        O.apply_maps(ob, [(kd, L, k1, 4, O.reference_cdt(kd, L, k1, 4, b.p, b.mass))
                          for kd, L, k1 in second])
        cpu_s += time.perf_counter() - t0
    g = pg[:, np.argsort(pg[6], kind="stable")]
    r = po[:, np.argsort(po[6], kind="stable")]
    diff = max(float(np.abs(g[c] - r[c]).max() / np.abs(r[c]).max()) for c in range(6))
    op.close()
    print(f"# drift case: GPU {gpu_s * 1e3:.2f} ms, CPU {cpu_s * 1e3:.1f} ms for 10 steps "
          f"(maps included), max coordinate difference {diff:.2e}", file=sys.stderr, flush=True)
    return {"workload": "1e5 particles, 1 m drift in 10 split-operator kicks, 32x32x64, FP64, "
                        "seed 0",
            "gpu_ms_per_step": 1e3 * gpu_s / 10, "cpu_ms_per_step": 1e3 * cpu_s / 10,
            "gpu_particle_kicks_per_s": n * 10 / gpu_s, "cpu_particle_kicks_per_s": n * 10 / cpu_s,
            "cpu": cores, "max_rel_coordinate_diff_gpu_vs_cpu": diff}


def host_cores():
    """The host threads granted to this process: its CPU affinity set, bounded
    by the cgroup CPU quota when one is set; with the CPU model."""
    aff = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = min(aff, quota) if quota else aff
    return {"threads": threads, "affinity": aff, "cgroup_quota_cpus": quota,
            "nproc": os.cpu_count(), "model": model}


def rel_l2(a, b):
    import numpy as np
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def accuracy_report(sc, S, config, steps):
    """The FP32-grid form against the FP64 form on the same seeded bunch:
    field rel-L2 (Ex, Ey, Ez octants) at the first kick and relative rms
    emittance drift (x, y, z) after `steps` steps; and the first kick's field
    of both forms against the oracle (CPU restatement, FP64) on the bunch
    state right before that kick."""
    import numpy as np
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    taps = (sc.DBG_EX, sc.DBG_EY, sc.DBG_EZ)
    field, emit = {}, {}
    for prec in (0, 1):
        bunch, opts, op, prop = make_workload(sc, S, config, precision=prec)
        if prec == 0:
            # oracle: first-step maps then one space-charge apply on the host
            parts, masks = bunch.host()
            parts = np.ascontiguousarray(parts)
            masks = np.ascontiguousarray(masks)
            ob = O.make_bunch(parts, masks, bunch.total_num, bunch.real_num, bunch.charge,
                              bunch.mass, bunch.gamma, bunch.beta, bunch.p)
            first, _, length = prop.steps[0]
            O.apply_maps(ob, [(kd, L, k1, 4, O.reference_cdt(kd, L, k1, 4, bunch.p, bunch.mass))
                              for kd, L, k1 in first])
            t0 = time.perf_counter()
            rc, out = O.apply(O.make_options(tuple(opts.shape), n_sigma=opts.n_sigma), ob,
                              length / (bunch.beta * sc.C_LIGHT), nthreads=host_cores()["threads"],
                              want=("ex", "ey", "ez"))
            assert rc == 0, O.last_error()
            oracle_field = [out["ex"], out["ey"], out["ez"]]
            print(f"# oracle apply {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
            del parts, masks, ob, out
        e0 = sc.emittances(op.moments(bunch)[1])
        prop.run(1, sync=True, fuse=False)
        field[prec] = [op.debug(t) for t in taps]
        prop.run(steps - 1, sync=True)
        emit[prec] = (e0, sc.emittances(op.moments(bunch)[1]))
        del bunch, op, prop
        torch.cuda.empty_cache()
    rep = {
        "steps": steps,
        "field_rel_l2_fp32_vs_fp64": max(rel_l2(field[1][c], field[0][c]) for c in range(3)),
        "field_rel_l2_fp64_vs_oracle": max(rel_l2(field[0][c], oracle_field[c]) for c in range(3)),
        "field_rel_l2_fp32_vs_oracle": max(rel_l2(field[1][c], oracle_field[c]) for c in range(3)),
        "emittance_rel_drift_fp32_vs_fp64": [abs(emit[1][1][k] / emit[0][1][k] - 1)
                                             for k in range(3)],
        "emittance_growth_fp64": [emit[0][1][k] / emit[0][0][k] for k in range(3)],
    }
    print(f"# accuracy: {json.dumps(rep)}", file=sys.stderr, flush=True)
    return rep


def measured_traffic(config):
    """HBM bytes per launch of the kick (the map-free k_kick_tile of the
    applies that do not reorder) and of
    k_deposit_tile from the committed rocprofv3 FETCH_SIZE + WRITE_SIZE runs
    of this workload (calibrated against a copy of known size), or {} when
    none was taken for it."""
    path = os.path.join(ROOT, "profiles", f"pmc_traffic_sc_step_{config}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        doc = json.load(f)
    out = {}
    for name, rec in doc["kernels"].items():
        if "k_kick_tile<" in name and ", false>" in name:  # not the reordering apply's
            out["k_kick_tile"] = rec["traffic_bytes"]
        elif "k_deposit_tile<" in name:
            out["k_deposit_tile"] = rec["traffic_bytes"]
    return out


def copy_probe(torch):
    """A 1 GiB device-to-device copy (512 MiB read + 512 MiB written), timed
    with events over 5 repeats: the achievable HBM rate of a plain stream,
    and the kernel whose FETCH_SIZE / WRITE_SIZE counters calibrate those of
    the space-charge kernels."""
    nbytes = 512 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(5):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    gbps = 5 * 2 * nbytes / (e0.elapsed_time(e1) / 1e3) / 1e9
    print(f"# copy probe: {2 * nbytes >> 20} MiB per copy, {gbps:.1f} GB/s", file=sys.stderr,
          flush=True)
    del src, dst
    torch.cuda.empty_cache()
    return round(gbps, 1)


def cpu_baseline(bunch, prop, opts, seconds, sc, S):
    """The oracle (CPU restatement of the reference operator) on the host
    cores, continuing the same bunch through the same channel in FP64: the
    maps between kicks run untimed, the space-charge applies are timed, until
    they have taken `seconds` (at least 2 steps)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    host = host_cores()
    cores = host["threads"]
    parts, masks = bunch.host()
    parts = np.ascontiguousarray(parts)
    masks = np.ascontiguousarray(masks)
    b = O.make_bunch(parts, masks, bunch.total_num, bunch.real_num, bunch.charge, bunch.mass,
                     bunch.gamma, bunch.beta, bunch.p)
    o = O.make_options(tuple(opts.shape), n_sigma=opts.n_sigma)
    k0 = prop.i
    sc_time = 0.0
    steps = 0
    while steps < 2 or (sc_time < seconds and steps < 400):
        k = (k0 + steps) % len(prop.steps)
        first, second, length = prop.steps[k]
        O.apply_maps(b, [(kd, L, k1, 4, O.reference_cdt(kd, L, k1, 4, bunch.p, bunch.mass))
                         for kd, L, k1 in first])
        t0 = time.perf_counter()
        rc, _ = O.apply(o, b, length / (bunch.beta * sc.C_LIGHT), nthreads=cores)
        sc_time += time.perf_counter() - t0
        assert rc == 0, O.last_error()
        O.apply_maps(b, [(kd, L, k1, 4, O.reference_cdt(kd, L, k1, 4, bunch.p, bunch.mass))
                         for kd, L, k1 in second])
        steps += 1
    return {
        "value": bunch.n * steps / sc_time,
        "unit": "particle-kicks/s",
        "cores": cores,
        "host": host,
        "kind": "port",
        "sample": f"{steps} space-charge steps of the same workload ({bunch.n} particles, "
                  f"grid {tuple(opts.shape)}, FODO channel; maps between kicks untimed), "
                  f"oracle/ CPU restatement of the reference operator with OpenMP ({cores} "
                  f"threads on {host['model']}; the reference binary cannot be built: no Kokkos, "
                  f"FFTW absent, so its own FFT); {sc_time:.2f} s of space-charge time",
    }


if __name__ == "__main__":
    main()
