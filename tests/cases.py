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
