"""The oracle reproduces the committed fixtures (regression of the checker
itself; the GPU side is checked against the same fixtures in
test_gpu_parity.py)."""
import math
import os

import numpy as np

import cases
import oracle_lib as ol
import test_oracle_pins as pins

HERE = os.path.join(os.path.dirname(__file__), "golden")


def test_c1mini_fixture():
    z = np.load(os.path.join(HERE, "c1mini.npz"))
    cfg = cases.abi.Config()
    for f, _ in cases.abi.Config._fields_:
        if f in z.files:
            setattr(cfg, f, z[f].item())
    o = cases.make(ol.OracleSim, cfg, (z["bx"], z["by"], z["bR0"], None, z["fg_in"].item()))
    st = o.step(int(z["steps"]))
    assert [s["n_pic"] for s in st] == list(z["n_pic"])
    for n in ("p0", "fg", "vx", "vy"):
        np.testing.assert_allclose(o.field(n), z[n], rtol=1e-12, atol=1e-12 * np.abs(z[n]).max())
    np.testing.assert_allclose(o.bubbles("R_t"), z["R_t"], rtol=1e-12)


def test_pml_sample_fixture():
    z = np.load(os.path.join(HERE, "pml_sample.npz"))
    c = pins.sample_config()
    c.bub_enable = 0
    o = ol.OracleSim(c)
    for k, n in enumerate(("mx", "my", "nx", "ny")):
        np.testing.assert_allclose(o.sigma(k), z[n], rtol=1e-15)
    np.testing.assert_allclose(o.scalars(), z["scalars"], rtol=1e-15)


def test_rp_single_fixture():
    z = np.load(os.path.join(HERE, "rp_single.npz"))
    dt = float(z["dt"])
    bp = ol.bubble_params(3e-6, dt_mix=dt)
    st = [3e-6, 3e-6, 0.0, bp.PG0, bp.dt0, 0.0, 0.0, 0.0, 0.0]
    Rt, subs = [], []
    for k in range(len(z["PL"])):
        st[6], st[7], st[8] = st[7], st[8], z["PL"][k]
        n, out = ol.rp_integrate(bp, dt, st)
        Rt.append(out[0])
        subs.append(n)
        st[:6] = list(out[1:7])
    np.testing.assert_array_equal(subs, z["subs"])
    np.testing.assert_allclose(Rt, z["R_t"], rtol=1e-12)
    assert max(subs) > 1000  # the violent-collapse branch of the step control is exercised
