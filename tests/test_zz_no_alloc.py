"""brp_get_bubbles_noalloc (the entry point that allocates nothing) against
brp_get_bubbles, called through the C ABI: every bubble array, bit for bit,
on two handles in a row, the second with fewer bubbles than the first (so
stale staging contents from the first call would show), plus its size limit
and its refusal before brp_prepare."""
import ctypes as C

import numpy as np
import pytest

import cases

pytestmark = pytest.mark.gpu


def _both(brp, sim, name):
    lib = brp.load_library()
    n = sim.info()["n_bubbles"]
    a = np.full(n + 8, -7.0)
    b = np.full(n + 8, -7.0)
    p = lambda v: v.ctypes.data_as(C.POINTER(C.c_double))
    assert lib.brp_get_bubbles(sim.h, brp.abi.BUB[name], p(a)) == 0, sim.last_error()
    assert lib.brp_get_bubbles_noalloc(sim.h, brp.abi.BUB[name], p(b)) == 0, sim.last_error()
    assert np.all(b[n:] == -7.0)  # nothing written past n
    return a[:n], b[:n]


@pytest.mark.parametrize("prec", [0, 1])
def test_noalloc_matches_get_bubbles_twice(brp, prec):
    for nb, steps in ((3000, 40), (700, 25)):  # the second call has a smaller input
        cfg, bub = cases.bubbly_case(64, nb=nb, seed=nb, precision=prec)
        sim = cases.make(brp.Simulation, cfg, bub)
        sim.step(steps)
        for name in brp.abi.BUB_NAMES:
            a, b = _both(brp, sim, name)
            np.testing.assert_array_equal(a, b, err_msg=f"{name} nb={nb}")
        print(f"noalloc prec={prec} nb={nb}: {len(brp.abi.BUB_NAMES)} arrays equal", flush=True)
        sim.close()


def test_noalloc_refuses_before_prepare_and_beyond_its_limit(brp):
    lib = brp.load_library()
    cfg, bub = cases.bubbly_case(64, nb=100)
    sim = brp.Simulation(cfg)
    sim.set_bubbles(*bub)
    out = np.zeros(100)
    rc = lib.brp_get_bubbles_noalloc(sim.h, brp.abi.BUB["R_t"],
                                     out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == brp.abi.ERR_STATE and "prepare" in sim.last_error()
    sim.close()
    n = (1 << 22) + 1
    cfg, _ = cases.bubbly_case(64, nb=4)
    sim = brp.Simulation(cfg)
    r = np.random.default_rng(0)
    sim.set_bubbles(3e-3 + 1e-3 * r.random(n), 3e-3 + 1e-3 * r.random(n), np.full(n, 3e-6),
                    None, 1e-9)
    sim.prepare()
    out = np.zeros(n)
    rc = lib.brp_get_bubbles_noalloc(sim.h, brp.abi.BUB["x"],
                                     out.ctypes.data_as(C.POINTER(C.c_double)))
    assert rc == brp.abi.ERR_INVALID and "4194304" in sim.last_error()
    sim.close()
