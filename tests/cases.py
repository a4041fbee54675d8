"""Seeded parity cases shared by the GPU tests, smoke() and bench.py.

Each case is a configuration plus (optionally) an explicit bubble population,
small enough for the FP64 oracle to finish in seconds.
"""
import importlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
abi = importlib.import_module("cuda-rp_amd.abi")


def rng(seed):
    return np.random.default_rng(seed)


def acoustic_case(n=64, precision=abi.PREC_FP64, amp_pa=0.5e6):
    """Bubbles off: focused source on a PML-bounded square (E1, E2)."""
    focal = n * 1e-4 * 0.62 / (0.5 * np.sqrt(3.0))
    return abi.water_config(n, npml=8, focal_dist=focal, amp_pa=amp_pa, bubbles=False,
                            precision=precision)


def sample_like_case(n=48, precision=abi.PREC_FP64, cylindrical=0):
    """The boundary set of the reference's confs/Sample.txt (PML on x1 and y1
    only, mirrors on x0 and y0, focused pressure source on the y0 row,
    source written only because pml.y0 = 0)."""
    c = abi.default_config()
    c.update(grid_x=n, grid_y=2 * n, grid_lx=n * 1e-4, grid_ly=2 * n * 1e-4,
             pml_x0=0, pml_x1=1, pml_y0=0, pml_y1=1, pml_grid_number=6, pml_sigma=0.1,
             pml_order=2, sim_cfl=0.1, sim_order=2, sim_smooth_delta_band=2,
             tr_amp=1.5e-2 * 20, tr_freq=1.0e6, tr_focal_dist=2 * n * 1e-4 * 0.6,
             tr_box_size=1.0e-3, tr_cylindrical=cylindrical, tr_on_wave=2, tr_off_wave=0,
             tr_focused_p=1, bub_enable=1, bub_fg0=1e-7, bub_R0=6.5e-7, bub_P0=0.1e6,
             bub_T0=293.15, bub_gamma=1.4, bub_sigma=7.2e-2, bub_mu=891.0e-6,
             bub_rho=1.0e3, bub_K0=26.2e-3, precision=precision)
    return c


def bubbly_case(n=64, nb=160, seed=1, precision=abi.PREC_FP64, amp_pa=1.0e6,
                box=1.6e-3, R_lo=1e-6, R_hi=5e-6, fg0=1e-6):
    """Bubbles in a box around the focus: seeded uniform positions and R0
    (extension E3), n_B from fg_in = fg0 * cells_in_box / nb."""
    focal = n * 1e-4 * 0.62 / (0.5 * np.sqrt(3.0))
    c = abi.water_config(n, npml=8, focal_dist=focal, amp_pa=amp_pa, bubbles=True,
                         fg0=fg0, precision=precision)
    fp_x = c.ext_focus_x
    fp_y = focal * 0.5 * np.sqrt(3.0)
    r = rng(seed)
    x = fp_x + (r.random(nb) - 0.5) * box
    y = fp_y + (r.random(nb) - 0.5) * box
    R0 = R_lo + (R_hi - R_lo) * r.random(nb)
    cells = (box / 1e-4) ** 2
    fg_in = fg0 * cells / nb
    return c, (x, y, R0, None, fg_in)


def make(sim_cls, cfg, bubbles=None, source=None):
    s = sim_cls(cfg)
    if bubbles is not None:
        x, y, R0, nB, fg_in = bubbles
        s.set_bubbles(x, y, R0, nB, fg_in)
    if source is not None:
        s.set_source(*source)
    s.prepare()
    return s


def rel_l2(a, b):
    d = np.linalg.norm((a - b).ravel())
    n = np.linalg.norm(b.ravel())
    return d / n if n > 0 else d


def rel_linf(a, b):
    m = np.max(np.abs(b))
    return np.max(np.abs(a - b)) / m if m > 0 else np.max(np.abs(a - b))


def adjoint_case(n=48, nb=40, n_elem=8, seed=3, precision=abi.PREC_FP64, amp_pa=1.2e6):
    """A bubbly case driven by an n_elem-element aperture (extension E4) with
    seeded per-element amplitudes (0.6..1.4 x amp_pa) and phases (0..2 pi)."""
    c, b = bubbly_case(n, nb, seed, precision, amp_pa)
    r = rng(seed + 100)
    amp = amp_pa * (0.6 + 0.8 * r.random(n_elem))
    phase = 2.0 * np.pi * r.random(n_elem)
    return c, b, (amp, phase)


def focal_box(cfg, half=3):
    """Cells (i0, i1, j0, j1), inclusive, of a (2 half) x (2 half) box centred
    on the focus."""
    dx = cfg.grid_lx / cfg.grid_x
    fx = cfg.ext_focus_x
    fy = cfg.tr_focal_dist * 0.5 * np.sqrt(3.0)
    ic = int(round(fx / dx + 0.5))
    jc = int(round(fy / dx + 0.5))
    return ic - half, ic + half - 1, jc - half, jc + half - 1


def focal_objective(sim, steps, box, window):
    """Mean of p0^2 over the box cells and over the last `window` of `steps`
    further steps (one step at a time; the p0 after each step counts)."""
    i0, i1, j0, j1 = box
    acc = 0.0
    for s in range(steps):
        sim.step(1)
        if s >= steps - window:
            p = sim.field("p0")[j0:j1 + 1, i0:i1 + 1]
            acc += float(np.sum(p * p))
    return acc / (window * (i1 - i0 + 1) * (j1 - j0 + 1))
