"""Writes the committed golden fixtures from the FP64 oracle.

    python tests/golden/make_golden.py

c1mini.npz     32^2 grid, 16 seeded bubbles (R0 ~ U(1, 5) um), focused 1 MPa
               source, 80 steps: config fields, bubble inputs, final p0 / fg /
               vx / vy, bubble R_t / x / y, per-step Picard counts.
pml_sample.npz the PML profiles mx, my, nx, ny of confs/Sample.txt.
rp_single.npz  one 3 um bubble driven by PL = 1.2 MPa sin(2 pi 1 MHz t) for
               300 mixture steps: R_t and sub-step count per step.
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import cases  # noqa: E402
import oracle_lib as ol  # noqa: E402


def c1mini():
    cfg, b = cases.bubbly_case(32, nb=16, seed=21, amp_pa=1.0e6, box=0.8e-3)
    o = cases.make(ol.OracleSim, cfg, b)
    st = o.step(80)
    out = {k: np.array(v) for k, v in cfg.as_dict().items()}
    out.update(bx=b[0], by=b[1], bR0=b[2], fg_in=np.array(b[4]), steps=np.array(80),
               n_pic=np.array([s["n_pic"] for s in st]), p0=o.field("p0"), fg=o.field("fg"),
               vx=o.field("vx"), vy=o.field("vy"), R_t=o.bubbles("R_t"), x=o.bubbles("x"),
               y=o.bubbles("y"))
    np.savez_compressed(os.path.join(HERE, "c1mini.npz"), **out)


def pml_sample():
    import test_oracle_pins as pins
    c = pins.sample_config()
    c.bub_enable = 0
    o = ol.OracleSim(c)
    np.savez_compressed(os.path.join(HERE, "pml_sample.npz"), mx=o.sigma(0), my=o.sigma(1),
                        nx=o.sigma(2), ny=o.sigma(3), scalars=o.scalars())


def rp_single():
    dt = 1e-4 * 0.5 / 1500.0
    bp = ol.bubble_params(3e-6, dt_mix=dt)
    PL = 1.2e6 * np.sin(2 * math.pi * 1e6 * dt * np.arange(1, 301))
    st = [3e-6, 3e-6, 0.0, bp.PG0, bp.dt0, 0.0, 0.0, 0.0, 0.0]
    Rt, subs = [], []
    for k in range(300):
        st[6], st[7], st[8] = st[7], st[8], PL[k]
        n, out = ol.rp_integrate(bp, dt, st)
        Rt.append(out[0])
        subs.append(n)
        st[:6] = list(out[1:7])
    np.savez_compressed(os.path.join(HERE, "rp_single.npz"), PL=PL, R_t=np.array(Rt),
                        subs=np.array(subs), dt=np.array(dt))


if __name__ == "__main__":
    c1mini()
    pml_sample()
    rp_single()
    print("fixtures written to", HERE)
