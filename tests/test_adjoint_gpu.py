"""The discrete adjoint (brp_adjoint) against central finite differences.

The reference map is piecewise smooth: Picard counts and RP sub-step sizes
are discrete decisions of the state.  The adjoint replays the forward's
schedule, so it is the exact derivative of the frozen-schedule map; the
oracle's frozen-schedule mode (record at the base point, replay at the
perturbed points) gives that map's central differences in FP64.  Bubble-free
runs have no schedule, and there the GPU's own central differences are the
check."""
import importlib

import numpy as np
import pytest

import cases
import oracle_lib as ol

pytestmark = pytest.mark.gpu
STEPS, WINDOW = 90, 30


def _oracle_J(cfg, bub, src, rec=None, fg_map=None, mode=2):
    o = cases.make(ol.OracleSim, cfg, bub, src)
    if fg_map is not None:
        o.set_field("fg", fg_map)
    if rec is not None:
        o.schedule_from(rec)
    o.schedule_mode(mode)
    J = cases.focal_objective(o, STEPS, cases.focal_box(cfg), WINDOW)
    return o, J


def _gpu_adjoint(brp, cfg, bub, src, every=20, fg_map=None):
    g = cases.make(brp.Simulation, cfg, bub, src)
    if fg_map is not None:
        g.set_field("fg", fg_map)
    r = g.adjoint(cases.focal_box(cfg), STEPS, WINDOW, every)
    return g, r


def test_adjoint_matches_frozen_schedule_differences(brp):
    """FP64, cavitating bubbles, 8 source elements: dJ/d amplitude and phase
    of every element and dJ/d f_g0 at cells in and out of the bubble cloud
    agree with the oracle's frozen-schedule central differences."""
    cfg, bub, (amp, phase) = cases.adjoint_case()
    rec, J0 = _oracle_J(cfg, bub, (amp, phase), mode=1)
    g, r = _gpu_adjoint(brp, cfg, bub, (amp, phase))
    assert r["J"] == pytest.approx(J0, rel=1e-9)
    assert r["n_checkpoints"] == 5
    assert r["schedule_mismatch"] == 0  # the replay took the forward's sub-steps
    worst = 0.0
    for e in range(len(amp)):
        for which, base in (("amp", amp[e]), ("phase", 1.0)):
            h = 1e-4 * base
            Jpm = []
            for sgn in (1, -1):
                a, p = amp.copy(), phase.copy()
                (a if which == "amp" else p)[e] += sgn * h
                Jpm.append(_oracle_J(cfg, bub, (a, p), rec)[1])
            fd = (Jpm[0] - Jpm[1]) / (2 * h)
            got = r["grad_" + which][e]
            scale = max(abs(fd), 1e-3 * abs(J0) / base)
            worst = max(worst, abs(got - fd) / scale)
            assert abs(got - fd) <= 1e-3 * scale, (which, e, got, fd)
    print(f"source gradients: worst relative error {worst:.2e}")
    o0 = cases.make(ol.OracleSim, cfg, bub, (amp, phase))
    base = o0.field("fg")
    box = cases.focal_box(cfg)
    ic, jc = (box[0] + box[1]) // 2, (box[2] + box[3]) // 2
    nz = np.argwhere(base > 0)
    cells = [tuple(nz[len(nz) // 2]), tuple(nz[0]), (jc + 12, ic - 15), (jc - 6, ic + 9)]
    for (j, i) in cells:
        h = 1e-10
        Jpm = []
        for sgn in (1, -1):
            m = base.copy()
            m[j, i] += sgn * h
            Jpm.append(_oracle_J(cfg, bub, (amp, phase), rec, fg_map=m)[1])
        fd = (Jpm[0] - Jpm[1]) / (2 * h)
        got = r["grad_fg0"][j, i]
        assert abs(got - fd) <= 1e-3 * abs(fd) + 1e-6 * abs(J0), ((j, i), got, fd)


def test_adjoint_checkpointing_and_final_state(brp):
    """The gradient does not depend on the checkpoint spacing nor on whether
    the reverse sweep replays a bubble tape or integrates RP again, and the
    handle ends at exactly the state a plain forward run reaches."""
    cfg, bub, src = cases.adjoint_case(nb=24)
    g1, r1 = _gpu_adjoint(brp, cfg, bub, src, every=0)   # sqrt spacing
    g2, r2 = _gpu_adjoint(brp, cfg, bub, src, every=STEPS)  # one segment
    g3, r3 = _gpu_adjoint(brp, cfg, bub, src, every=7)
    g4 = cases.make(brp.Simulation, cfg, bub, src)   # RP recomputed instead of taped
    r4 = g4.adjoint(cases.focal_box(cfg), STEPS, WINDOW, 20, tape=False)
    f = cases.make(brp.Simulation, cfg, bub, src)
    f.step(STEPS)
    for r in (r2, r3, r4):
        assert r["J"] == r1["J"]
        for k in ("grad_amp", "grad_phase", "grad_fg0"):
            np.testing.assert_allclose(r[k], r1[k], rtol=1e-11,
                                       atol=1e-11 * np.abs(r1[k]).max())
    for n in ("p0", "vx", "fg"):
        np.testing.assert_array_equal(g1.field(n), f.field(n))
    np.testing.assert_array_equal(g1.bubbles("R_t"), f.bubbles("R_t"))
    assert g1.info()["nstep"] == STEPS


def test_adjoint_without_bubbles_matches_gpu_differences(brp):
    """Linear acoustics (no schedule): the adjoint equals the GPU's own
    central differences, FP64 and mixed."""
    cfg, _, (amp, phase) = cases.adjoint_case()
    cfg.update(bub_enable=0)
    for prec, tol in ((0, 1e-6), (1, 2e-2)):
        cfg.update(precision=prec)
        _, r = _gpu_adjoint(brp, cfg, None, (amp, phase))
        for e in (1, 6):
            h = 1e-3
            Jpm = []
            for sgn in (1, -1):
                p = phase.copy()
                p[e] += sgn * h
                s = cases.make(brp.Simulation, cfg, None, (amp, p))
                Jpm.append(cases.focal_objective(s, STEPS, cases.focal_box(cfg), WINDOW))
            fd = (Jpm[0] - Jpm[1]) / (2 * h)
            # p0 is linear in the source: the O(h^2) truncation is about 1e-7 of dJ
            assert abs(r["grad_phase"][e] - fd) <= tol * abs(fd), (prec, e, r["grad_phase"][e], fd)


def test_mixed_adjoint_close_to_fp64(brp):
    cfg, bub, src = cases.adjoint_case()
    _, rd = _gpu_adjoint(brp, cfg, bub, src)
    cfg.update(precision=1)
    _, rm = _gpu_adjoint(brp, cfg, bub, src)
    for k in ("grad_amp", "grad_phase"):
        assert cases.rel_l2(rm[k], rd[k]) < 1e-2, k
    assert rm["J"] == pytest.approx(rd["J"], rel=1e-3)


def test_autograd_function(brp):
    torch = pytest.importorskip("torch")
    ag = importlib.import_module("cuda-rp_amd.autograd")
    cfg, bub, (amp, phase) = cases.adjoint_case(nb=24)
    box = cases.focal_box(cfg)

    def make_sim(a, p):
        return cases.make(brp.Simulation, cfg, bub, (a, p))

    s0 = make_sim(amp, phase)
    fg0 = torch.tensor(s0.field("fg"), requires_grad=True)
    s0.close()
    A = torch.tensor(amp, requires_grad=True)
    P = torch.tensor(phase, requires_grad=True)
    J = ag.focal_objective(A, P, fg0, make_sim, box, STEPS, WINDOW, 30)
    (2.0 * J).backward()
    g = make_sim(amp, phase)
    r = g.adjoint(box, STEPS, WINDOW, 30)
    assert float(J.detach()) == r["J"]
    np.testing.assert_allclose(A.grad.numpy(), 2.0 * r["grad_amp"], rtol=1e-12)
    np.testing.assert_allclose(P.grad.numpy(), 2.0 * r["grad_phase"], rtol=1e-12)
    np.testing.assert_allclose(fg0.grad.numpy(), 2.0 * r["grad_fg0"], rtol=1e-12,
                               atol=1e-12 * np.abs(r["grad_fg0"]).max())
