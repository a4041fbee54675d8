#!/usr/bin/env python3
"""Benchmark of the coupled FDTD + Rayleigh-Plesset step on one MI355X.

Workload (one "step" = one coupled mixture time step of the whole problem):
  2048 x 2048 cells, dx = dy = 0.1 mm, 32-cell PML on all sides, water
  (c = 1500 m/s, rho = 1000 kg/m^3, mu = 1e-3 Pa s, sigma = 0.0725 N/m,
  P0 = 101325 Pa, gamma = 1.4), focused 1 MHz continuous wave of 1.5 MPa from
  the y0 edge (focal length 15 mm, focus centred in x), 2^20 bubbles with
  log-normal R0 (median 3 um, sigma_log 0.4) uniform in an 8 x 8 mm box around
  the focus (seed 2), initial mean void fraction 1e-6 in the box, CFL 0.5.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

W untimed steps bring the wavefront to the bubbles (the focus is ~300 steps
from the source), then K steps are timed with the mixed-precision path
(FP32 fields, FP64 RP and void fraction), then the same W + K steps run with
the all-FP64 path for the speed and error comparison.  One JSON line on
stdout.  Progress lines go to stderr.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "cell-updates/s & bubble-steps/s: mixed vs FP64 at stated err; adjoint/fwd time"
# Algorithmic FP64 operations of one RP sub-step (add, sub, mul, div, sqrt,
# tanh, tan each counted once; the oracle's operation counter gives 470.6 on a
# driven log-normal population) and of one bubble's gather + scatter.
RP_FLOPS_PER_SUBSTEP = 470.0
BUBBLE_FLOPS_PER_CALL = 60.0
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def focused_2048_workload(abi, precision, n=2048, nb=1 << 20, seed=2):
    cfg = abi.water_config(n, npml=32, cfl=0.5, amp_pa=1.5e6, freq=1.0e6, focal_dist=15e-3,
                           bubbles=True, fg0=1e-6, precision=precision)
    r = np.random.default_rng(seed)
    fpx, fpy = cfg.ext_focus_x, 15e-3 * 0.5 * math.sqrt(3.0)
    box = 8e-3
    x = fpx + (r.random(nb) - 0.5) * box
    y = fpy + (r.random(nb) - 0.5) * box
    R0 = np.exp(math.log(3e-6) + 0.4 * r.standard_normal(nb))
    fg_in = 1e-6 * (box / 1e-4) ** 2 / nb
    return cfg, (x, y, R0, None, fg_in)


def clustered_4096_workload(abi, precision, n=4096, nb=1 << 22, seed=3):
    """4096^2, 32-cell PML, 2 MHz 3 MPa focused CW, 2^22 log-normal bubbles
    (median 3 um, sigma_log 0.4) clustered in 16 Gaussian clouds (centres
    uniform over an 8 x 8 mm box at the focus, sigma 0.25 mm, clipped to the
    box), mean void fraction 1e-6 over the box."""
    cfg = abi.water_config(n, npml=32, cfl=0.5, amp_pa=3.0e6, freq=2.0e6, focal_dist=15e-3,
                           bubbles=True, fg0=1e-6, precision=precision)
    r = np.random.default_rng(seed)
    fpx, fpy = cfg.ext_focus_x, 15e-3 * 0.5 * math.sqrt(3.0)
    box = 8e-3
    cx = fpx + (r.random(16) - 0.5) * box
    cy = fpy + (r.random(16) - 0.5) * box
    k = r.integers(0, 16, nb)
    x = np.clip(cx[k] + 0.25e-3 * r.standard_normal(nb), fpx - 0.5 * box, fpx + 0.5 * box)
    y = np.clip(cy[k] + 0.25e-3 * r.standard_normal(nb), fpy - 0.5 * box, fpy + 0.5 * box)
    R0 = np.exp(math.log(3e-6) + 0.4 * r.standard_normal(nb))
    fg_in = 1e-6 * (box / 1e-4) ** 2 / nb
    return cfg, (x, y, R0, None, fg_in)


def adjoint_1024_workload(abi, precision, n=1024, nb=65536, n_elem=64, seed=4):
    """The adjoint workload: 1024^2, 65536 bubbles with R0 ~ U(1, 5) um in a
    4 x 4 mm box at the focus (mean void fraction 1e-6 there), a 64-element
    aperture with amplitudes ~ U(0.3, 0.7) MPa and phases ~ U(0, 2 pi)."""
    cfg = abi.water_config(n, npml=32, cfl=0.5, amp_pa=0.5e6, freq=1.0e6, focal_dist=15e-3,
                           bubbles=True, fg0=1e-6, precision=precision)
    r = np.random.default_rng(seed)
    fpx, fpy = cfg.ext_focus_x, 15e-3 * 0.5 * math.sqrt(3.0)
    box = 4e-3
    x = fpx + (r.random(nb) - 0.5) * box
    y = fpy + (r.random(nb) - 0.5) * box
    R0 = 1e-6 + 4e-6 * r.random(nb)
    fg_in = 1e-6 * (box / 1e-4) ** 2 / nb
    amp = 0.3e6 + 0.4e6 * r.random(n_elem)
    phase = 2.0 * math.pi * r.random(n_elem)
    dx = cfg.grid_lx / n
    ic, jc = int(round(fpx / dx + 0.5)), int(round(fpy / dx + 0.5))
    focal_box = (ic - 10, ic + 9, jc - 10, jc + 9)  # 2 x 2 mm
    return cfg, (x, y, R0, None, fg_in), (amp, phase), focal_box


def adjoint_leg(brp, abi, steps, window, every):
    """brp_adjoint on the adjoint workload next to a plain forward of the
    same steps: adjoint/forward time."""
    cfg, bub, src, box = adjoint_1024_workload(abi, abi.PREC_MIXED)
    f = brp.Simulation(cfg)
    f.set_bubbles(*bub)
    f.set_source(*src)
    f.prepare()
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        m = min(500, steps - done)
        f.step(m)
        done += m
        log(f"  adjoint workload forward step {done}: {(time.perf_counter() - t0) / done * 1e3:.2f} ms/step")
    t_fwd = time.perf_counter() - t0
    f.close()
    a = brp.Simulation(cfg)
    a.set_bubbles(*bub)
    a.set_source(*src)
    a.prepare()
    log(f"  adjoint: {steps} steps, window {window}, checkpoint every {every}")
    t0 = time.perf_counter()
    r = a.adjoint(box, steps, window, every)
    t_adj = time.perf_counter() - t0
    a.close()
    g = np.concatenate([r["grad_amp"], r["grad_phase"]])
    log(f"  adjoint: {t_adj:.2f} s vs forward {t_fwd:.2f} s")
    return {"workload": "1024x1024 grid, 65536 bubbles R0~U(1,5) um in a 4x4 mm box, "
                        "64 elements amp~U(0.3,0.7) MPa phase~U(0,2pi), mixed precision",
            "steps": steps, "window": window, "checkpoint_every": every,
            "J": r["J"], "forward_s": t_fwd, "adjoint_s": t_adj,
            "adjoint_over_forward": t_adj / t_fwd,
            "adjoint_forward_sweep_s": r["forward_ms"] / 1e3,
            "adjoint_reverse_sweep_s": r["reverse_ms"] / 1e3,
            "memory_gb": r["memory_bytes"] / 1e9, "n_checkpoints": r["n_checkpoints"],
            "grad_norm": float(np.linalg.norm(g)),
            "grad_fg0_abs_max": float(np.abs(r["grad_fg0"]).max())}


def stepped(sim, k, tag, chunk=25):
    """sim.step(k) in chunks, printing one progress line per chunk."""
    out = []
    phases = None
    done = 0
    while done < k:
        m = min(chunk, k - done)
        t0 = time.perf_counter()
        st = sim.step(m)
        dt = time.perf_counter() - t0
        ph = sim.phase_times()
        phases = ph if phases is None else {a: phases[a] + ph[a] for a in ph}
        done += m
        out += st
        log(f"  {tag} step {st[-1]['nstep']}: {dt / m * 1e3:.2f} ms/step, Picard "
            f"{sum(s['n_pic'] for s in st) / m:.2f}, RP sub-steps mean "
            f"{np.mean([s['rp_mean_substeps'] for s in st]):.1f} max "
            f"{max(s['rp_max_substeps'] for s in st)}")
    return out, phases


def run_gpu(brp, cfg, bub, warmup, steps, first):
    sim = brp.Simulation(cfg)
    sim.set_bubbles(*bub)
    sim.prepare()
    t0 = time.perf_counter()
    st_first = sim.step(first) if first else []
    t_first = time.perf_counter() - t0
    if warmup > first:
        stepped(sim, warmup - first, "warmup", chunk=50)
    at_warmup = (sim.field("p0"), sim.bubbles("R_t"))
    log(f"  warmup done ({warmup} steps), timing {steps} steps")
    t0 = time.perf_counter()
    st, phases = stepped(sim, steps, "timed")
    wall = time.perf_counter() - t0
    return sim, st, wall, phases, t_first, at_warmup


def cpu_baseline(abi, cfg, bub, sample_steps, threads):
    import oracle_lib as ol
    o = ol.OracleSim(cfg, threads=threads)
    o.set_bubbles(*bub)
    o.prepare()
    t0 = time.perf_counter()
    st = o.step(sample_steps)
    wall = time.perf_counter() - t0
    o.close()
    return wall, st


def measured_traffic():
    """HBM bytes per RP launch (FETCH_SIZE x 2 + WRITE_SIZE, copy-calibrated)
    from the committed counter summary, or None."""
    path = os.path.join(ROOT, "profiles", "k_bubble_persistent_2048_counters.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def grid_roofline(ph, n, steps, mixed):
    """The grid kernels (velocity, first pressure evaluation and their ghost
    kernels) against HBM: 20 + 32 algorithmic bytes per cell-update in mixed
    mode (vx, vy, p0 read, vx, vy written; vx, vy, pn.x, pn.y, p0 read, p.x,
    p.y, p0 written), twice that in FP64; the Picard re-evaluations cover the
    bubble window only and are left out of the byte count."""
    per_cell = 52.0 if mixed else 104.0
    ms = (ph["velocity"] + ph["pressure"]) / steps
    achieved = per_cell * n * n / (ms * 1e-3) / 1e9
    return {"kernels": "k_velocity + k_pressure_first + ghosts (+ window Picard evaluations)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "ms_per_step": ms,
            "bytes_per_cell_update": per_cell}


def thermal_leg(brp, abi, n, nb, bub, warmup, steps, other_without_ms):
    """The same workload with the one-way temperature solve and the focal
    search on (bubble heat scatter, conductivity/flux/increment pass, the add
    pass, ghosts, argmax of the work field, one 16-byte read-back), timed over
    `steps` steps after `warmup`.  Its device time is the "other" phase minus
    the same phase without the thermal solve.  Algorithmic bytes per cell:
    increment T, f_g, f_g(n), heat (8 each), p.x, p.y, pn.x, pn.y (4 each in
    mixed mode), dT written (8); add T, dT read, T written (24); argmax dT
    read (8): 88 bytes."""
    cfg, _ = focused_2048_workload(abi, abi.PREC_MIXED, n, nb)
    cfg.thermal = 1
    sim, st, wall, ph, _, _ = run_gpu(brp, cfg, bub, warmup, steps, 0)
    sim.close()
    th_ms = max(ph["other"] / steps - other_without_ms, 1e-9)
    per_cell = 88.0
    achieved = per_cell * n * n / (th_ms * 1e-3) / 1e9
    return {"ms_per_step": wall / steps * 1e3, "thermal_ms_per_step": th_ms,
            "steps_window": f"{warmup + 1}-{warmup + steps}",
            "focal_y_last": st[-1]["focal_y"],
            "roofline": {"kernels": "k_bubble_heat + k_temperature_increment + k_temperature_add "
                                    "+ k_temperature_ghosts + k_focal_blocks/final",
                         "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "bytes_per_cell": per_cell}}


def clustered_leg(brp, abi, warmup, steps):
    """The 4096^2 workload with 2^22 clustered bubbles at 2 MHz / 3 MPa
    (clustered_4096_workload), mixed precision, timed over `steps` steps after `warmup`
    (the wavefront reaches the clouds near step 200)."""
    cfg, bub = clustered_4096_workload(abi, abi.PREC_MIXED)
    n, nb = cfg.grid_x, len(bub[0])
    sim, st, wall, ph, _, _ = run_gpu(brp, cfg, bub, warmup, steps, 0)
    sim.close()
    launches = sum(s["n_pic"] for s in st)
    return {"workload": f"{n}x{n} grid, 32-cell PML, 2 MHz 3 MPa focused CW, {nb} log-normal "
                        "bubbles in 16 Gaussian clouds (sigma 0.25 mm) in an 8x8 mm box, mixed "
                        "precision",
            "steps_window": f"{warmup + 1}-{warmup + steps}",
            "ms_per_step": wall / steps * 1e3, "value": n * n * steps / wall,
            "unit": "cell-updates/s", "bubble_steps_per_s": nb * steps / wall,
            "picard_per_step": launches / steps,
            "rp_launch_ms": ph["bubble_rp"] / max(launches, 1),
            "rp_max_substeps": max(s["rp_max_substeps"] for s in st),
            "phase_ms_per_step": {k: v / steps for k, v in ph.items()}}


def calibrate_copy(mib=1024, reps=5):
    """A plain device copy of known size (reads and writes `mib` MiB per
    launch), run under `rocprofv3 --pmc FETCH_SIZE` / `WRITE_SIZE` to scale
    those counters to bytes."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        torch.mul(a, 1.0, out=b)
    torch.cuda.synchronize()
    print(json.dumps({"copy_kernel_bytes_read": n * 4, "copy_kernel_bytes_written": n * 4,
                      "launches": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain copy kernel of 1 GiB (counter calibration)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--cpu-steps", type=int, default=20)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--bubbles", type=int, default=1 << 20)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-thermal", action="store_true")
    ap.add_argument("--no-clustered", action="store_true")
    ap.add_argument("--clustered-steps", type=int, default=100)
    ap.add_argument("--clustered-warmup", type=int, default=300)
    ap.add_argument("--thermal-steps", type=int, default=100)
    ap.add_argument("--adjoint-steps", type=int, default=4000)
    ap.add_argument("--adjoint-window", type=int, default=1000)
    ap.add_argument("--adjoint-checkpoint", type=int, default=200)
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint leg (printed as one JSON line)")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the coupled step runs on one MI355X)",
              file=sys.stderr)
        sys.exit(2)

    import importlib
    brp = importlib.import_module("cuda-rp_amd")
    abi = brp.abi
    if args.adjoint_only:
        print(json.dumps(adjoint_leg(brp, abi, args.adjoint_steps, args.adjoint_window,
                                     args.adjoint_checkpoint)), flush=True)
        return
    n, nb = args.grid, args.bubbles
    first = min(args.cpu_steps, args.warmup)

    cfg_m, bub = focused_2048_workload(abi, abi.PREC_MIXED, n, nb)
    log(f"mixed: {n}^2 cells, {nb} bubbles, warmup {args.warmup}, steps {args.steps}")
    sim_m, st_m, wall_m, ph_m, tfirst_m, warm_m = run_gpu(brp, cfg_m, bub, args.warmup,
                                                          args.steps, first)
    cells = n * n
    cups = cells * args.steps / wall_m
    bsps = nb * args.steps / wall_m
    launches = sum(s["n_pic"] for s in st_m)
    substeps = sum(s["rp_mean_substeps"] * s["n_pic"] * nb for s in st_m)
    rp_ms = ph_m["bubble_rp"]
    avg_ms = rp_ms / max(launches, 1)
    flops_per_launch = (RP_FLOPS_PER_SUBSTEP * substeps + BUBBLE_FLOPS_PER_CALL * nb * launches) \
        / max(launches, 1)
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    log(f"  mixed: {wall_m / args.steps * 1e3:.3f} ms/step, {cups:.3e} cell-updates/s, "
        f"phases {ph_m}, Picard {launches / args.steps:.2f}/step")

    out = {
        "metric": METRIC,
        "value": cups,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall_m / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields + f64 RP / int64 void fraction (mixed)",
        "data": "synthetic (seeded bubble population, analytic source)",
        "config": {"workload": f"{n}x{n} grid, 32-cell PML, 1 MHz 1.5 MPa focused CW, "
                               f"{nb} log-normal bubbles in an 8x8 mm box, CFL 0.5",
                   "grid": n, "bubbles": nb, "precision": "mixed",
                   "steps_window": f"{args.warmup + 1}-{args.warmup + args.steps}"},
        "bubble_steps_per_s": bsps,
        "picard_per_step": launches / args.steps,
        "rp_substeps_per_s": substeps / wall_m,
        "phase_ms_per_step": {k: v / args.steps for k, v in ph_m.items()},
        "roofline": {"kernel": "k_bubble_persistent (gather + Rayleigh-Plesset + scatter)",
                     "bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic(), "avg_launch_ms": avg_ms,
                     "flops_per_launch": flops_per_launch,
                     "note": "FP64 vector ALU; MI355X FP64 vector and matrix peaks are both "
                             "78.6 TFLOP/s.  A launch is bound by the longest bubbles' "
                             "sequential sub-step chains, each on a wavefront of its own "
                             "that issues at most one instruction per 4 cycles whatever its "
                             "active lanes (about 700 instructions per sub-step), not by FP64 "
                             "throughput; traffic = HBM bytes per launch from rocprofv3 "
                             "FETCH_SIZE / WRITE_SIZE (profiles/)"},
        "grid_roofline": grid_roofline(ph_m, n, args.steps, mixed=True),
    }

    if not args.no_fp64:
        cfg_d, _ = focused_2048_workload(abi, abi.PREC_FP64, n, nb)
        log("fp64 run")
        sim_d, st_d, wall_d, ph_d, _, warm_d = run_gpu(brp, cfg_d, bub, args.warmup, args.steps,
                                                       0)
        cups_d = cells * args.steps / wall_d
        from cases import rel_l2
        err_p0 = rel_l2(sim_m.field("p0"), sim_d.field("p0"))
        err_rt = rel_l2(sim_m.bubbles("R_t"), sim_d.bubbles("R_t"))
        out["fp64"] = {"value": cups_d, "ms_per_step": wall_d / args.steps * 1e3,
                       "bubble_steps_per_s": nb * args.steps / wall_d,
                       "phase_ms_per_step": {k: v / args.steps for k, v in ph_d.items()},
                       "picard_per_step": sum(s["n_pic"] for s in st_d) / args.steps,
                       "grid_roofline": grid_roofline(ph_d, n, args.steps, mixed=False)}
        out["mixed_vs_fp64"] = {"speedup": cups / cups_d, "p0_rel_l2": err_p0,
                                "R_t_rel_l2": err_rt,
                                "at_step": args.warmup + args.steps,
                                "p0_rel_l2_at_warmup_end": rel_l2(warm_m[0], warm_d[0]),
                                "R_t_rel_l2_at_warmup_end": rel_l2(warm_m[1], warm_d[1]),
                                "warmup_end_step": args.warmup,
                                "picard_count_mismatch_steps":
                                    int(sum(a["n_pic"] != b["n_pic"] for a, b in zip(st_m, st_d)))}
        log(f"  fp64: {wall_d / args.steps * 1e3:.3f} ms/step; p0 err {err_p0:.2e}")
        sim_d.close()
    sim_m.close()

    if not args.no_clustered:
        log("clustered 4096^2 leg")
        out["clustered_4096"] = clustered_leg(brp, abi, args.clustered_warmup, args.clustered_steps)
        log(f"  clustered: {out['clustered_4096']['ms_per_step']:.2f} ms/step")

    if not args.no_thermal:
        log("thermal leg")
        out["thermal"] = thermal_leg(brp, abi, n, nb, bub, args.warmup, args.thermal_steps,
                                     ph_m["other"] / args.steps)
        log(f"  thermal: {out['thermal']['thermal_ms_per_step']:.3f} ms/step")

    if not args.no_adjoint:
        log("adjoint leg")
        out["adjoint"] = adjoint_leg(brp, abi, args.adjoint_steps, args.adjoint_window,
                                     args.adjoint_checkpoint)

    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        cfg_c, _ = focused_2048_workload(abi, abi.PREC_FP64, n, nb)
        log(f"cpu baseline: oracle, {threads} threads, steps 1-{first}")
        wall_c, st_c = cpu_baseline(abi, cfg_c, bub, first, threads)
        out["cpu_baseline"] = {
            "value": cells * first / wall_c, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"the FP64 oracle on this workload, steps 1-{first} from rest "
                      f"({wall_c:.1f} s; the bubbles are still at equilibrium)",
            "bubble_steps_per_s": nb * first / wall_c,
            "gpu_same_sample_value": cells * first / tfirst_m,
        }
        if threads > 1:  # the same oracle on one core, on a shorter sample
            k1 = max(1, min(2, first))
            log(f"cpu baseline: oracle, 1 thread, steps 1-{k1}")
            wall_1, _ = cpu_baseline(abi, cfg_c, bub, k1, 1)
            out["cpu_baseline"]["single_thread"] = {
                "value": cells * k1 / wall_1, "unit": "cell-updates/s", "cores": 1,
                "sample": f"steps 1-{k1} ({wall_1:.1f} s)"}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
