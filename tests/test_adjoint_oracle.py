"""The oracle's frozen-schedule mode: the reference map of one run is
piecewise smooth (Picard counts and RP step sizes are discrete decisions of
the state); freezing them at a recorded run gives the smooth map whose
central differences the adjoint is checked against."""
import numpy as np

import cases
import oracle_lib as ol

STEPS, WINDOW = 90, 30


def _run(cfg, bub, src, mode, rec=None, fg_map=None):
    o = cases.make(ol.OracleSim, cfg, bub, src)
    if fg_map is not None:
        o.set_field("fg", fg_map)
    if rec is not None:
        o.schedule_from(rec)
    o.schedule_mode(mode)
    J = cases.focal_objective(o, STEPS, cases.focal_box(cfg), WINDOW)
    return o, J


def test_replay_reproduces_the_recorded_run_bit_exactly():
    cfg, bub, src = cases.adjoint_case()
    rec, J0 = _run(cfg, bub, src, 1)
    free, Jf = _run(cfg, bub, src, 0)
    rep, J1 = _run(cfg, bub, src, 2, rec)
    assert J0 == J1 == Jf
    for n in ("p0", "fg", "vx"):
        np.testing.assert_array_equal(rep.field(n), rec.field(n))
    np.testing.assert_array_equal(rep.bubbles("R_t"), rec.bubbles("R_t"))
    assert J0 > 0


def test_frozen_map_is_smooth_in_phase_and_amplitude():
    """Central differences at h and h/2 agree to O(h^2) on the frozen map,
    while the cavitating run really uses an adaptive schedule."""
    cfg, bub, (amp, phase) = cases.adjoint_case()
    rec, J0 = _run(cfg, bub, (amp, phase), 1)
    st = rec.step(1)[0]
    assert st["rp_max_substeps"] > 10  # adaptive sub-stepping is active
    for which, e, base in (("phase", 2, 1.0), ("amp", 5, amp[5])):
        d = []
        for h in (1e-4 * base, 0.5e-4 * base):
            Jpm = []
            for sgn in (1, -1):
                a, p = amp.copy(), phase.copy()
                (p if which == "phase" else a)[e] += sgn * h
                Jpm.append(_run(cfg, bub, (a, p), 2, rec)[1])
            d.append((Jpm[0] - Jpm[1]) / (2 * h))
        assert abs(d[0] - d[1]) <= 1e-6 * abs(d[1]) + 1e-30, (which, d)
