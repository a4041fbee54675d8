#!/usr/bin/env python3
"""Benchmark of the coupled FDTD + Rayleigh-Plesset step on one MI355X.

Headline workload (one "step" = one coupled mixture time step of the whole
problem), the clustered 4096^2 configuration:
  4096 x 4096 cells, dx = dy = 0.1 mm, 32-cell PML on all sides, water
  (c = 1500 m/s, rho = 1000 kg/m^3, mu = 1e-3 Pa s, sigma = 0.0725 N/m,
  P0 = 101325 Pa, gamma = 1.4), a focused 2 MHz continuous wave of 3 MPa from
  the y0 edge (focal length 15 mm, focus centred in x), 2^22 bubbles with
  log-normal R0 (median 3 um, sigma_log 0.4) in 16 seeded Gaussian clouds
  (sigma 0.25 mm) in an 8 x 8 mm box at the focus, mean void fraction 1e-6
  over the box, CFL 0.5, mixed precision (FP32 fields, FP64 RP and
  void-fraction accumulation).

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Every run first takes PREROLL = 300 untimed steps, which bring the wavefront
onto the bubble clouds (the focus is ~260 steps from the source) so that the
timed window is inside the cavitating regime the benchmark is about; then W
further untimed warm-up steps, then K timed steps.  Beside the headline: the
all-FP64 run of the same window, mixed-vs-FP64 window statistics, the
void-fraction accumulation schemes on the FP64 run's radii, the error of the
mixed run against the FP64 oracle seeded with the GPU's state at the window
start, the CPU baseline, the ASCII writer, the 2048^2 workload, the
temperature solve and the 1024^2 adjoint workload.  --at-scale runs only the
4096^2 adjoint workload (8000 steps, a few minutes).  One JSON line on
stdout; progress on stderr.
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

import cases  # noqa: E402  (workload definitions shared with the tests)

METRIC = "cell-updates/s & bubble-steps/s: mixed vs FP64 at stated err; adjoint/fwd time"
PREROLL = 300
# Algorithmic FP64 operations of one RP sub-step (add, sub, mul, div, sqrt,
# tanh, tan each counted once; the oracle's operation counter gives 470.6 on a
# driven log-normal population) and of one bubble's gather + scatter.
RP_FLOPS_PER_SUBSTEP = 470.0
BUBBLE_FLOPS_PER_CALL = 60.0
FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes of one step: per cell-update 116 + 56 n_pic (mixed) or
# 176 + 88 n_pic (FP64); per bubble-step 96 + 288 n_pic.
CELL_BYTES = {"mixed": (116.0, 56.0), "fp64": (176.0, 88.0)}
BUBBLE_BYTES = (96.0, 288.0)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stepped(sim, k, tag, chunk=25):
    """sim.step(k) in chunks, one progress line per chunk; returns the step
    records and the summed device phase times (ms)."""
    out, phases, done = [], None, 0
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


def timed_window(brp, cfg, bub, warmup, steps, tag, keep_state=False):
    """PREROLL + warmup untimed steps, then `steps` timed ones (synchronous:
    each brp_step call returns after its last kernel).  With keep_state the
    state at the window start is saved in slot 0 and exported."""
    sim = brp.Simulation(cfg)
    sim.set_bubbles(*bub)
    sim.prepare()
    stepped(sim, PREROLL + warmup, f"{tag} pre-roll", chunk=50)
    state = None
    if keep_state:
        sim.save_state(0)
        state = sim.export_state()
    t0 = time.perf_counter()
    st, ph = stepped(sim, steps, f"{tag} timed")
    wall = time.perf_counter() - t0
    return sim, st, wall, ph, state


def summarize(st, wall, ph, n, nb, prec):
    steps = len(st)
    launches = sum(s["n_pic"] for s in st)
    substeps = sum(s["rp_mean_substeps"] * s["n_pic"] * nb for s in st)
    a, b = CELL_BYTES[prec]
    step_bytes = sum(n * n * (a + b * s["n_pic"]) + nb * (BUBBLE_BYTES[0] + BUBBLE_BYTES[1] *
                                                          s["n_pic"]) for s in st)
    return {"ms_per_step": wall / steps * 1e3, "cell_updates_per_s": n * n * steps / wall,
            "bubble_steps_per_s": nb * steps / wall, "rp_substeps_per_s": substeps / wall,
            "picard_per_step": launches / steps, "launches": launches, "substeps": substeps,
            "rp_launch_ms": ph["bubble_rp"] / max(launches, 1),
            "rp_max_substeps": max(s["rp_max_substeps"] for s in st),
            "rp_lean_launches": sum(s["rp_lean_launches"] for s in st),
            "rp_kept_fraction": sum(s["rp_reused"] for s in st) / max(launches * nb, 1),
            "phase_ms_per_step": {k: v / steps for k, v in ph.items()},
            "device_ms_per_step": sum(ph.values()) / steps,
            "step_bytes": step_bytes / steps}


def rp_roofline(sm, nb, traffic):
    """The dominant kernel, k_bubble_persistent (RP + scatter of every bubble,
    one launch per Picard iteration, after k_queue_keys gathered the
    pressures): algorithmic FP64 operations of the sub-steps it integrated
    (results kept from the previous iteration count none) per launch over its
    mean launch time from HIP events on the library's stream."""
    flops = (RP_FLOPS_PER_SUBSTEP * sm["substeps"] + BUBBLE_FLOPS_PER_CALL * nb * sm["launches"]) \
        / max(sm["launches"], 1)
    achieved = flops / (sm["rp_launch_ms"] * 1e-3) / 1e12
    return {"kernel": "k_bubble_persistent (Rayleigh-Plesset + void-fraction scatter, one "
                      "launch per Picard iteration)",
            "bound": "fp64-valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
            "avg_launch_ms": sm["rp_launch_ms"], "flops_per_launch": flops,
            "substeps_per_launch": sm["substeps"] / max(sm["launches"], 1),
            "note": "FP64 vector ALU (no MFMA: the RP sub-step is a scalar recurrence per "
                    "bubble); peak = MI355X FP64 vector 78.6 TFLOP/s.  traffic = HBM bytes "
                    "per launch, FETCH_SIZE x 2 + WRITE_SIZE from rocprofv3 (copy-calibrated)"}


def step_roofline(sm, prec):
    achieved = sm["step_bytes"] / (sm["device_ms_per_step"] * 1e-3) / 1e9
    a, b = CELL_BYTES[prec]
    return {"scope": "whole coupled step (all kernels), device time",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "bytes_per_step": sm["step_bytes"],
            "model": f"N^2 ({a:.0f} + {b:.0f} n_pic) + Nb (96 + 288 n_pic) bytes per step"}


def grid_roofline(ph, n, steps, mixed):
    """The grid kernels (velocity, first pressure evaluation and their ghost
    kernels) against HBM: 20 + 32 algorithmic bytes per cell-update in mixed
    mode, twice that in FP64; the window-only Picard re-evaluations are left
    out of the byte count."""
    per_cell = 52.0 if mixed else 104.0
    ms = (ph["velocity"] + ph["pressure"]) / steps
    achieved = per_cell * n * n / (ms * 1e-3) / 1e9
    return {"kernels": "k_velocity + k_pressure_first + ghosts (+ window Picard evaluations)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "ms_per_step": ms,
            "bytes_per_cell_update": per_cell}


def measured_traffic(name):
    """HBM bytes per RP launch (FETCH_SIZE x 2 + WRITE_SIZE, copy-calibrated)
    from a committed counter summary, or None."""
    try:
        for line in open(os.path.join(ROOT, "profiles", name)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def focal_cells(cfg, half=10):
    return cases.focal_box(cfg, half)


def window_statistics(sim_m, sim_d, cfg, steps):
    """Mixed against FP64 over `steps` more steps of both runs: pointwise
    errors (which grow as collapse times drift apart) and the statistics a
    user of the simulation reads -- the time-averaged focal pressure squared
    (the heat-deposition proxy), the R_t / R0 percentiles and the fraction of
    bubbles expanded beyond 2 R0."""
    i0, i1, j0, j1 = focal_cells(cfg)
    acc = {"m": 0.0, "d": 0.0}
    for _ in range(steps):
        for tag, s in (("m", sim_m), ("d", sim_d)):
            s.step(1)
            p = s.field("p0")[j0:j1 + 1, i0:i1 + 1]
            acc[tag] += float(np.mean(p * p)) / steps
    R0 = sim_m.bubbles("R0")
    rm, rd = sim_m.bubbles("R_t") / R0, sim_d.bubbles("R_t") / R0
    pct = (50, 90, 99, 99.9)
    qm, qd = np.percentile(rm, pct), np.percentile(rd, pct)
    return {"steps": steps,
            "focal_p2_mean_mixed": acc["m"], "focal_p2_mean_fp64": acc["d"],
            "focal_p2_rel_err": abs(acc["m"] - acc["d"]) / acc["d"],
            "R_over_R0_percentiles": {f"p{p}": [float(a), float(b)] for p, a, b in zip(pct, qm, qd)},
            "R_over_R0_percentile_rel_err_max": float(np.max(np.abs(qm - qd) / qd)),
            "expanded_2R0_fraction": [float(np.mean(rm > 2.0)), float(np.mean(rd > 2.0))],
            "p0_rel_l2_pointwise": cases.rel_l2(sim_m.field("p0"), sim_d.field("p0")),
            "R_t_rel_l2_pointwise": cases.rel_l2(sim_m.bubbles("R_t"), sim_d.bubbles("R_t"))}


def oracle_threads():
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return env or len(os.sched_getaffinity(0))


def oracle_seeded(cfg, bub, state, steps, threads):
    """The FP64 oracle moved to a GPU state and advanced `steps` steps."""
    import oracle_lib as ol
    ocfg = cases.abi.Config.from_buffer_copy(cfg)
    ocfg.precision = cases.abi.PREC_FP64
    o = ol.OracleSim(ocfg, threads=threads)
    o.set_bubbles(*bub)
    o.prepare()
    o.import_state(state)
    t0 = time.perf_counter()
    st = o.step(steps)
    return o, st, time.perf_counter() - t0


def oracle_leg(brp, sim_m, cfg, bub, state, steps, threads):
    """Error of the mixed GPU run against the FP64 oracle, both started from
    the GPU's state at the window start (slot 0) and advanced `steps` steps;
    the oracle's time is the CPU baseline."""
    n, nb = cfg.grid_x, len(bub[0])
    log(f"oracle: seeded at the window start, {steps} steps on {threads} threads")
    o, ost, wall = oracle_seeded(cfg, bub, state, steps, threads)
    log(f"  oracle: {wall:.1f} s")
    sim_m.restore_state(0)
    gst = sim_m.step(steps)
    subset = cases.rng(6).choice(nb, min(4096, nb), replace=False)
    err = {"steps": steps,
           "p0_rel_l2": cases.rel_l2(sim_m.field("p0"), o.field("p0")),
           "p0_rel_linf": cases.rel_linf(sim_m.field("p0"), o.field("p0")),
           "vx_rel_l2": cases.rel_l2(sim_m.field("vx"), o.field("vx")),
           "fg_rel_l2": cases.rel_l2(sim_m.field("fg"), o.field("fg")),
           "R_t_rel_l2_subset4096": cases.rel_l2(sim_m.bubbles("R_t")[subset],
                                                 o.bubbles("R_t")[subset]),
           "picard_gpu": [s["n_pic"] for s in gst], "picard_oracle": [s["n_pic"] for s in ost]}
    o.close()
    base = {"value": n * n * steps / wall, "unit": "cell-updates/s", "cores": threads,
            "kind": "port",
            "sample": f"the FP64 oracle (g++ -O2, OpenMP) seeded with the GPU state at the start "
                      f"of the timed window (cavitating), {steps} step(s), {wall:.1f} s",
            "bubble_steps_per_s": nb * steps / wall,
            "nproc": len(os.sched_getaffinity(0)), "os_cpu_count": os.cpu_count()}
    return err, base


def single_thread_baseline(brp):
    """The oracle on one core on the C1 workload (256^2, 1024 bubbles) over
    steps 301-340 (cavitating), beside the GPU (FP64) on the same steps."""
    import oracle_lib as ol
    cfg, bub, src = cases.c1_case()
    out = {}
    for tag, mk in (("cpu", lambda c: ol.OracleSim(c, threads=1)), ("gpu", brp.Simulation)):
        s = cases.make(mk, cfg, bub, src)
        s.step(300)
        t0 = time.perf_counter()
        s.step(40)
        out[tag] = time.perf_counter() - t0
        s.close()
    n = cfg.grid_x
    return {"value": n * n * 40 / out["cpu"], "unit": "cell-updates/s", "cores": 1,
            "sample": f"C1 workload (256^2, 1024 bubbles, 0.5 MPa), steps 301-340, "
                      f"{out['cpu']:.1f} s", "gpu_same_sample_value": n * n * 40 / out["gpu"]}


def focused_2048_leg(brp, warmup, steps):
    """The 2048^2 workload (2^20 bubbles, 1.5 MPa), mixed and FP64, same
    window after the pre-roll."""
    out = {}
    for prec, name in ((1, "mixed"), (0, "fp64")):
        cfg, bub = cases.c2_case(precision=prec)
        sim, st, wall, ph, _ = timed_window(brp, cfg, bub, warmup, steps, f"2048 {name}")
        sm = summarize(st, wall, ph, cfg.grid_x, len(bub[0]), name)
        sm["grid_roofline"] = grid_roofline(ph, cfg.grid_x, steps, prec == 1)
        out[name] = sm
        out[f"_{name}_sim"] = sim
    e = {"p0_rel_l2": cases.rel_l2(out["_mixed_sim"].field("p0"), out["_fp64_sim"].field("p0")),
         "R_t_rel_l2": cases.rel_l2(out["_mixed_sim"].bubbles("R_t"),
                                    out["_fp64_sim"].bubbles("R_t"))}
    for k in ("_mixed_sim", "_fp64_sim"):
        out.pop(k).close()
    return {"workload": "2048x2048 grid, 32-cell PML, 1 MHz 1.5 MPa focused CW, 1048576 "
                        "log-normal bubbles in an 8x8 mm box",
            "steps_window": f"{PREROLL + warmup + 1}-{PREROLL + warmup + steps}",
            "mixed": out["mixed"], "fp64": out["fp64"], "mixed_vs_fp64_at_window_end": e}


def thermal_leg(brp, warmup, steps):
    """The 2048^2 workload with the one-way temperature solve and focal search
    on, beside the same run without it: the difference of their "other"
    phase is the temperature solve.  Algorithmic bytes per cell: increment
    (T, f_g, f_g(n), heat 8 each; p.x, p.y, pn.x, pn.y 4 each; dT written 8),
    add (T, dT read, T written: 24), argmax (dT read: 8): 88 bytes."""
    res = {}
    for th in (0, 1):
        cfg, bub = cases.c2_case(precision=1)
        cfg.thermal = th
        sim, st, wall, ph, _ = timed_window(brp, cfg, bub, warmup, steps, f"thermal={th}")
        res[th] = (st, wall, ph)
        sim.close()
    n = 2048
    th_ms = max((res[1][2]["other"] - res[0][2]["other"]) / steps, 1e-9)
    achieved = 88.0 * n * n / (th_ms * 1e-3) / 1e9
    return {"ms_per_step": res[1][1] / steps * 1e3, "thermal_ms_per_step": th_ms,
            "focal_y_last": res[1][0][-1]["focal_y"],
            "roofline": {"kernels": "k_bubble_heat + k_temperature_increment + "
                                    "k_temperature_add + k_temperature_ghosts + k_focal_*",
                         "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "bytes_per_cell": 88.0}}


def adjoint_leg(brp, steps, window, every):
    """brp_adjoint on the 1024^2 adjoint workload beside a plain forward of
    the same steps: adjoint/forward time; the library refuses (raises) when
    its replay of the RP schedule differs from the forward's."""
    abi = brp.abi
    cfg, bub, src, box = cases.c4_case(precision=abi.PREC_MIXED)
    f = cases.make(brp.Simulation, cfg, bub, src)
    t0 = time.perf_counter()
    done, sub_total, pic_total, sub_max = 0, 0.0, 0, 0
    nb = len(bub[0])
    while done < steps:
        m = min(500, steps - done)
        for st in f.step(m):
            sub_total += st["rp_mean_substeps"] * nb * st["n_pic"]
            pic_total += st["n_pic"]
            sub_max = max(sub_max, st["rp_max_substeps"])
        done += m
        log(f"  adjoint workload forward step {done}: "
            f"{(time.perf_counter() - t0) / done * 1e3:.2f} ms/step")
    t_fwd = time.perf_counter() - t0
    f.close()
    a = cases.make(brp.Simulation, cfg, bub, src)
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
            "schedule_mismatch": r["schedule_mismatch"],
            "forward_rp_substeps": sub_total, "forward_rp_launches": pic_total,
            "forward_rp_max_substeps": sub_max,
            "grad_norm": float(np.linalg.norm(g)),
            "grad_fg0_abs_max": float(np.abs(r["grad_fg0"]).max())}


def c5_leg(brp, steps, window):
    """brp_adjoint on the adjoint workload at scale (4096^2, 2^20 bubbles,
    256 elements, checkpoints every ceil(sqrt(steps)) steps) beside a plain
    forward of the same steps.  Its bubble tape (192 B per bubble-step) does
    not fit the library's 96 GB tape budget, so the reverse sweep integrates
    RP again between checkpoints.  Device memory is the library's account of
    what the adjoint holds (checkpoints, step records, RP Jacobians) plus
    the handle."""
    abi = brp.abi
    cfg, bub, src, box = cases.c5_case(precision=abi.PREC_MIXED)
    out = {"workload": "4096x4096 grid, 32-cell PML, 1048576 log-normal bubbles in an 8x8 mm box, "
                       "256 elements amp~U(0.3,0.7) MPa phase~U(0,2pi), 1 MHz, focal length 51.2 mm, "
                       "mixed precision",
           "steps": steps, "window": window}
    f = cases.make(brp.Simulation, cfg, bub, src)
    out["handle_gb"] = f.info()["device_bytes"] / 1e9
    t0 = time.perf_counter()
    done, sub_total, sub_max = 0, 0.0, 0
    while done < steps:
        m = min(500, steps - done)
        for st in f.step(m):
            sub_total += st["rp_mean_substeps"] * len(bub[0]) * st["n_pic"]
            sub_max = max(sub_max, st["rp_max_substeps"])
        done += m
        log(f"  at-scale forward step {done}: {(time.perf_counter() - t0) / done * 1e3:.2f} ms/step")
    out["forward_s"] = time.perf_counter() - t0
    out["forward_rp_substeps"] = sub_total
    out["forward_rp_max_substeps"] = sub_max
    f.close()
    a = cases.make(brp.Simulation, cfg, bub, src)
    log(f"  at-scale adjoint: {steps} steps, window {window}")
    t0 = time.perf_counter()
    try:
        r = a.adjoint(box, steps, window, 0)
    except brp.BrpError as e:
        out["error"] = str(e)
        a.close()
        return out
    out["adjoint_s"] = time.perf_counter() - t0
    a.close()
    g = np.concatenate([r["grad_amp"], r["grad_phase"]])
    out.update({"adjoint_over_forward": out["adjoint_s"] / out["forward_s"],
                "adjoint_forward_sweep_s": r["forward_ms"] / 1e3,
                "adjoint_reverse_sweep_s": r["reverse_ms"] / 1e3,
                "J": r["J"], "adjoint_memory_gb": r["memory_bytes"] / 1e9,
                "device_memory_gb": r["memory_bytes"] / 1e9 + out["handle_gb"],
                "n_checkpoints": r["n_checkpoints"], "checkpoint_every": r["segment"],
                "schedule_mismatch": r["schedule_mismatch"],
                "grad_norm": float(np.linalg.norm(g)),
                "grad_fg0_abs_max": float(np.abs(r["grad_fg0"]).max())})
    log(f"  at-scale adjoint: {out['adjoint_s']:.1f} s vs forward {out['forward_s']:.1f} s, "
        f"{out['device_memory_gb']:.1f} GB")
    return out


def vf_accumulation(brp, sim):
    """The void fraction of the simulation's current radii scattered with
    each accumulation scheme of brp_scatter_variant, against FP64 atomics
    (rounding of order 1e-16 per add; order-dependent, so a second FP64
    scatter is compared too).  The library's own f_g is the int64 scheme."""
    abi = brp.abi
    ref = sim.scatter_variant(abi.VF_FP64_ATOMIC)
    out = {"reference": "FP64 atomic scatter of the same radii"}
    for name, mode in (("fp64_atomic_repeat", abi.VF_FP64_ATOMIC),
                       ("int64_fixed_point", abi.VF_INT64),
                       ("fp32_compensated", abi.VF_FP32_COMPENSATED),
                       ("fp32_atomic", abi.VF_FP32_ATOMIC)):
        v = sim.scatter_variant(mode)
        out[name] = {"rel_l2": cases.rel_l2(v, ref), "rel_linf": cases.rel_linf(v, ref)}
    out["int64_equals_fg"] = bool(np.array_equal(sim.scatter_variant(abi.VF_INT64),
                                                 sim.field("fg")))
    return out


def ascii_writer_leg(n=4096):
    """The driver's ASCII field writer (cuda-rp_amd/bin/rawdata_bench): one
    n x n field of save_ascii text, MB/s of the formatting alone and of the
    whole write, with the host's threads and with one (the tool also checks
    the bytes against the serial ostream text)."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "cuda-rp_amd", "bin", "rawdata_bench")
    threads = oracle_threads()
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe, d, str(n), str(threads)], capture_output=True, text=True,
                           timeout=600)
    out = {"grid": n, "ok": r.returncode == 0 and r.stdout.strip().endswith("OK")}
    for line in r.stdout.splitlines():
        if line.startswith("threads"):
            t = int(line.split()[1].rstrip(":"))
            out[f"write_MB_per_s_{t}_threads"] = float(line.split("=")[1].split()[0])
            out["MB_per_field"] = float(line.split(":")[1].split()[0])
        elif line.startswith("format threads"):
            t = int(line.split()[2].rstrip(":"))
            out[f"format_MB_per_s_{t}_threads"] = float(line.split("=")[1].split()[0])
    return out


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
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline window (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain copy kernel of 1 GiB (counter calibration)")
    ap.add_argument("--stat-steps", type=int, default=20)
    ap.add_argument("--oracle-steps", type=int, default=1)
    ap.add_argument("--leg-steps", type=int, default=50)
    ap.add_argument("--adjoint-steps", type=int, default=4000)
    ap.add_argument("--adjoint-window", type=int, default=1000)
    ap.add_argument("--adjoint-checkpoint", type=int, default=200)
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("--no-legs", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only the adjoint leg (for profiling)")
    ap.add_argument("--at-scale", action="store_true",
                    help="only the adjoint at scale (4096^2, 2^20 bubbles, 256 elements)")
    ap.add_argument("--at-scale-steps", type=int, default=8000)
    ap.add_argument("--at-scale-window", type=int, default=1000)
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported "
              "(the coupled step runs on one MI355X)", file=sys.stderr)
        sys.exit(2)

    import importlib
    brp = importlib.import_module("cuda-rp_amd")
    abi = brp.abi
    if args.at_scale:
        log("adjoint at scale")
        print(json.dumps({"adjoint_at_scale": c5_leg(brp, args.at_scale_steps,
                                                      args.at_scale_window)}), flush=True)
        return
    if args.adjoint_only:
        log("adjoint leg")
        print(json.dumps({"adjoint": adjoint_leg(brp, args.adjoint_steps, args.adjoint_window,
                                                 args.adjoint_checkpoint)}), flush=True)
        return
    K, W = args.steps, args.warmup

    cfg_m, bub = cases.c3_case(precision=abi.PREC_MIXED)
    n, nb = cfg_m.grid_x, len(bub[0])
    log(f"headline: {n}^2 cells, {nb} clustered bubbles, pre-roll {PREROLL} + warmup {W}, "
        f"timed {K}")
    keep = not (args.headline_only or args.no_oracle)
    sim_m, st_m, wall_m, ph_m, state = timed_window(brp, cfg_m, bub, W, K, "mixed", keep)
    sm = summarize(st_m, wall_m, ph_m, n, nb, "mixed")
    log(f"  mixed: {sm['ms_per_step']:.3f} ms/step, RP launch {sm['rp_launch_ms']:.3f} ms, "
        f"Picard {sm['picard_per_step']:.2f}/step")
    window = f"{PREROLL + W + 1}-{PREROLL + W + K}"
    out = {
        "metric": METRIC,
        "value": sm["cell_updates_per_s"],
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": sm["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields + f64 RP / int64 void fraction (mixed)",
        "data": "synthetic (seeded clustered bubble population, analytic source)",
        "config": {"workload": f"{n}x{n} grid, 32-cell PML, 2 MHz 3 MPa focused CW, {nb} "
                               "log-normal bubbles in 16 Gaussian clouds (sigma 0.25 mm) in an "
                               "8x8 mm box, CFL 0.5",
                   "grid": n, "bubbles": nb, "precision": "mixed",
                   "preroll_steps": PREROLL, "steps_window": window},
        "bubble_steps_per_s": sm["bubble_steps_per_s"],
        "rp_substeps_per_s": sm["rp_substeps_per_s"],
        "picard_per_step": sm["picard_per_step"],
        "rp_max_substeps": sm["rp_max_substeps"],
        "rp_lean_launches": sm["rp_lean_launches"],
        "rp_kept_fraction": sm["rp_kept_fraction"],
        "timed_picard_counts": [s["n_pic"] for s in st_m],
        "phase_ms_per_step": sm["phase_ms_per_step"],
        "roofline": rp_roofline(sm, nb, measured_traffic("k_bubble_persistent_4096_counters.md")),
        "step_roofline": step_roofline(sm, "mixed"),
        "grid_roofline": grid_roofline(ph_m, n, K, mixed=True),
    }
    if args.headline_only:
        print(json.dumps(out), flush=True)
        return

    if not args.no_fp64:
        cfg_d, _ = cases.c3_case(precision=abi.PREC_FP64)
        sim_d, st_d, wall_d, ph_d, _ = timed_window(brp, cfg_d, bub, W, K, "fp64")
        sd = summarize(st_d, wall_d, ph_d, n, nb, "fp64")
        out["fp64"] = {"value": sd["cell_updates_per_s"], "ms_per_step": sd["ms_per_step"],
                       "bubble_steps_per_s": sd["bubble_steps_per_s"],
                       "picard_per_step": sd["picard_per_step"],
                       "rp_launch_ms": sd["rp_launch_ms"],
                       "rp_kept_fraction": sd["rp_kept_fraction"],
                       "phase_ms_per_step": sd["phase_ms_per_step"],
                       "step_roofline": step_roofline(sd, "fp64"),
                       "grid_roofline": grid_roofline(ph_d, n, K, mixed=False)}
        mvd = {"speedup": sm["cell_updates_per_s"] / sd["cell_updates_per_s"],
               "picard_count_mismatch_steps":
                   int(sum(a["n_pic"] != b["n_pic"] for a, b in zip(st_m, st_d))),
               "at_window_end": {
                   "p0_rel_l2": cases.rel_l2(sim_m.field("p0"), sim_d.field("p0")),
                   "R_t_rel_l2": cases.rel_l2(sim_m.bubbles("R_t"), sim_d.bubbles("R_t"))}}
        log("window statistics, mixed vs FP64")
        mvd["window_after"] = window_statistics(sim_m, sim_d, cfg_m, args.stat_steps)
        out["mixed_vs_fp64"] = mvd
        log("void-fraction accumulation schemes (FP64 run, end of its window)")
        out["vf_accumulation"] = vf_accumulation(brp, sim_d)
        sim_d.close()

    if keep:
        threads = oracle_threads()
        err, base = oracle_leg(brp, sim_m, cfg_m, bub, state, args.oracle_steps, threads)
        out["mixed_vs_oracle"] = err
        out["cpu_baseline"] = base
        del state
        log("cpu baseline: one thread, C1 workload")
        out["cpu_baseline"]["single_thread"] = single_thread_baseline(brp)
    sim_m.close()

    if not args.no_legs:
        log("ASCII writer leg")
        out["ascii_writer"] = ascii_writer_leg()
        log("2048^2 leg")
        out["focused_2048"] = focused_2048_leg(brp, W, args.leg_steps)
        log("thermal leg")
        out["thermal"] = thermal_leg(brp, W, args.leg_steps)
    if not args.no_adjoint:
        log("adjoint leg")
        out["adjoint"] = adjoint_leg(brp, args.adjoint_steps, args.adjoint_window,
                                     args.adjoint_checkpoint)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
