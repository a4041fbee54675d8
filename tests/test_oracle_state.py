"""State export / import on the oracle (CPU): a run moved to another run's
state continues bit for bit like it, so the seeded GPU-vs-oracle
comparisons at C2/C3 start the oracle from exactly the state it is given."""
import numpy as np

import cases
import oracle_lib as ol


def test_oracle_state_roundtrip_is_exact():
    cfg, bub = cases.bubbly_case(48, nb=200, amp_pa=1.5e6)
    a = cases.make(ol.OracleSim, cfg, bub)
    a.step(70)
    st = a.export_state()
    b = cases.make(ol.OracleSim, cfg, bub)
    b.import_state(st)
    sa, sb = a.step(30), b.step(30)
    assert max(s["rp_max_substeps"] for s in sa) > 30  # the bubbles are driven
    assert [s["n_pic"] for s in sa] == [s["n_pic"] for s in sb]
    for n in ("p0", "px", "py", "vx", "vy", "fg", "fgn", "fgm", "rho_m"):
        np.testing.assert_array_equal(a.field(n), b.field(n), err_msg=n)
    for n in ("R_t", "R_p", "R_n", "d1R", "PG", "x", "y", "Q_B", "ibn_x", "ibm_y"):
        np.testing.assert_array_equal(a.bubbles(n), b.bubbles(n), err_msg=n)
    assert a.clock() == b.clock()


def test_oracle_state_import_moves_the_clock_and_bins():
    cfg, bub = cases.bubbly_case(48, nb=20)
    a = cases.make(ol.OracleSim, cfg, bub)
    a.set_clock(17, 5.5e-7, 1.25e-23)
    assert a.clock() == (17, 5.5e-7, 1.25e-23)
    x = a.bubbles("x") + 3e-4
    a.set_bubble_array("x", x)
    np.testing.assert_array_equal(a.bubbles("ibn_x"), np.floor(x * (1.0 / 1e-4) + 0.5))
