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


def test_adjoint_refuses_exact_rp_and_schedule_modes(brp):
    """brp_adjoint linearises the shortened-chain sub-step; with rp_exact = 1
    (another sub-step count on a few percent of bubbles) it refuses instead of
    returning the gradient of another trajectory; likewise while a frozen
    schedule is being recorded or replayed."""
    cfg, bub, src = cases.adjoint_case(nb=24)
    cfg.rp_exact = 1
    g = cases.make(brp.Simulation, cfg, bub, src)
    with pytest.raises(brp.BrpError) as e:
        g.adjoint(cases.focal_box(cfg), STEPS, WINDOW, 20)
    assert e.value.code == brp.abi.ERR_INVALID and "rp_exact" in str(e.value)
    cfg.rp_exact = 0
    g = cases.make(brp.Simulation, cfg, bub, src)
    g.schedule_mode(1)
    with pytest.raises(brp.BrpError) as e:
        g.adjoint(cases.focal_box(cfg), STEPS, WINDOW, 20)
    assert e.value.code == brp.abi.ERR_STATE


def _gpu_frozen_J(brp, cfg, bub, src, base, fg_map=None, steps=STEPS, window=WINDOW, box=None):
    s = cases.make(brp.Simulation, cfg, bub, src)
    if fg_map is not None:
        s.set_field("fg", fg_map)
    s.schedule_from(base)
    s.schedule_mode(2)
    J = cases.focal_objective(s, steps, box or cases.focal_box(cfg), window)
    s.close()
    return J


def test_gpu_frozen_schedule(brp):
    """The GPU's frozen-schedule mode: replaying its own record at the same
    parameters reproduces the recorded run bit for bit, and at perturbed
    source parameters its J agrees with the oracle's own frozen-schedule J
    (each replaying the schedule it recorded at the base point; FP64)."""
    cfg, bub, (amp, phase) = cases.adjoint_case()
    box = cases.focal_box(cfg)
    base = cases.make(brp.Simulation, cfg, bub, (amp, phase))
    base.schedule_mode(1)
    J0 = cases.focal_objective(base, STEPS, box, WINDOW)
    rec, J0o = _oracle_J(cfg, bub, (amp, phase), mode=1)
    assert J0 == pytest.approx(J0o, rel=1e-9)
    again = cases.make(brp.Simulation, cfg, bub, (amp, phase))
    again.schedule_from(base)
    again.schedule_mode(2)
    assert cases.focal_objective(again, STEPS, box, WINDOW) == J0
    for n in ("p0", "fg"):
        np.testing.assert_array_equal(again.field(n), base.field(n))
    np.testing.assert_array_equal(again.bubbles("R_t"), base.bubbles("R_t"))
    np.testing.assert_array_equal(again.bubbles("sub"), base.bubbles("sub"))
    for e in (0, 5):
        a = amp.copy()
        a[e] *= 1.0 + 1e-3
        Jg = _gpu_frozen_J(brp, cfg, bub, (a, phase), base)
        Jo = _oracle_J(cfg, bub, (a, phase), rec)[1]
        print(f"element {e}: frozen J GPU {Jg:.12e} oracle {Jo:.12e}", flush=True)
        assert Jg == pytest.approx(Jo, rel=1e-9)


def test_c4_adjoint_against_frozen_differences(brp):
    """The adjoint workload at its own size (1024^2, 65536 bubbles with
    R0 ~ U(1, 5) um, 64 elements with amp ~ U(0.3, 0.7) MPa, phase ~ U(0, 2 pi),
    the 2 x 2 mm objective box), FP64: brp_adjoint against the GPU's own
    frozen-schedule central differences (h = 1e-4 relative) on 8 seeded
    parameters (3 amplitudes, 3 phases, and 2 random +-1 directions over the
    cloud's cells of the initial void-fraction map, each cell moved by 1e-4
    of its value), agreement 1e-3.  The horizon is 400 steps with J over the last 100:
    the wavefront reaches the bubbles near step 300, and over the 4000-step
    window of the benchmark the cavitating cloud is chaotic (the gradient
    norm grows to ~1e19), so a perturbation of 1e-4 leaves the linear range
    and no finite difference can check the derivative there."""
    cfg, bub, src, box = cases.c4_case(precision=brp.abi.PREC_FP64)
    amp, phase = src
    steps, window = 400, 100
    base = cases.make(brp.Simulation, cfg, bub, src)
    base.schedule_mode(1)
    J0 = cases.focal_objective(base, steps, box, window)
    fg_base = cases.make(brp.Simulation, cfg, bub, src).field("fg")
    g = cases.make(brp.Simulation, cfg, bub, src)
    r = g.adjoint(box, steps, window, 50)
    assert r["J"] == pytest.approx(J0, rel=1e-12)
    assert r["schedule_mismatch"] == 0
    rng = cases.rng(44)
    picks = [("amp", int(k)) for k in rng.choice(64, 3, replace=False)] + \
            [("phase", int(k)) for k in rng.choice(64, 3, replace=False)]
    cloud = fg_base > 0.1 * fg_base.max()  # cells inside the cloud
    picks += [("fg0", int(k)) for k in (1, 2)]
    worst = 0.0
    # the map is stored as int64 multiples of 2^-60: differences are taken
    # over the perturbation actually applied
    q = lambda v: np.rint(v * 2.0 ** 60) * 2.0 ** -60

    def central(which, k, rel):
        Jpm = []
        for sgn in (1, -1):
            a, p, m = amp.copy(), phase.copy(), None
            if which == "amp":
                h = rel * amp[k]
                a[k] += sgn * h
            elif which == "phase":
                h = rel
                p[k] += sgn * h
            else:
                m = fg_base.copy()
                m[cloud] = q(fg_base[cloud] * (1 + sgn * rel * direction[cloud]))
                h = 1.0
            Jpm.append(_gpu_frozen_J(brp, cfg, bub, (a, p), base, fg_map=m, steps=steps,
                                     window=window, box=box))
        return (Jpm[0] - Jpm[1]) / (2 * h)

    for which, k in picks:
        if which == "fg0":
            # one cell of the map moves J by ~1e-11 of itself, the rounding
            # floor of 400 cavitating steps; a seeded random direction over
            # every cell of the cloud sums thousands of such responses.
            # Checked: dJ/2 against sum_i grad_i (m+_i - m-_i) / 2.
            direction = np.sign(cases.rng(100 + k).standard_normal(fg_base.shape))
            dm = (q(fg_base * (1 + 1e-4 * direction)) - q(fg_base * (1 - 1e-4 * direction))) / 2
            got = float(np.sum(r["grad_fg0"][cloud] * dm[cloud]))
        else:
            got = r["grad_" + which][k]
        fd = central(which, k, 1e-4)
        err = abs(got - fd) / max(abs(fd), 1e-30)
        worst = max(worst, err)
        print(f"{which} {k}: adjoint {got:.8e} FD {fd:.8e} rel err {err:.2e}", flush=True)
        assert err < 1e-3, (which, k, got, fd)
    print(f"C4 adjoint vs frozen-schedule differences: worst relative error {worst:.2e}")


def test_c5_replica_against_oracle(brp):
    """The adjoint workload at scale, reduced to 256^2 cells (256 elements,
    4096 bubbles, FP64, 550 steps with J over the last 100, checkpoints every
    ceil(sqrt(550)) steps): brp_adjoint against the oracle's frozen-schedule
    central differences along two seeded directions -- every one of the 512
    amplitudes and phases moved by +-1e-4 (relative, and radians), and every
    cloud cell of the initial void-fraction map by +-1e-4 of its value --
    agreement 1e-3.  A direction sums the responses of all its parameters,
    so one pair of oracle runs checks the whole gradient against the
    reference's equations."""
    cfg, bub, (amp, phase), box = cases.c5_replica_case()
    steps, window, h = 550, 100, 1e-4
    g = cases.make(brp.Simulation, cfg, bub, (amp, phase))
    fg_base = g.field("fg")
    r = g.adjoint(box, steps, window, 0)
    g.close()
    assert r["schedule_mismatch"] == 0
    rec = cases.make(ol.OracleSim, cfg, bub, (amp, phase))
    rec.schedule_mode(1)
    J0 = cases.focal_objective(rec, steps, box, window)
    assert r["J"] == pytest.approx(J0, rel=1e-9)
    q = lambda v: np.rint(v * 2.0 ** 60) * 2.0 ** -60
    cloud = fg_base > 0.1 * fg_base.max()

    def oracle_J(a, p, m=None):
        o = cases.make(ol.OracleSim, cfg, bub, (a, p))
        if m is not None:
            o.set_field("fg", m)
        o.schedule_from(rec)
        o.schedule_mode(2)
        J = cases.focal_objective(o, steps, box, window)
        o.close()
        return J

    r6 = cases.rng(6)
    sa = np.sign(r6.standard_normal(len(amp)))
    sp = np.sign(r6.standard_normal(len(phase)))
    fd = (oracle_J(amp * (1 + h * sa), phase + h * sp) -
          oracle_J(amp * (1 - h * sa), phase - h * sp)) / 2
    got = float(np.dot(r["grad_amp"], h * sa * amp) + np.dot(r["grad_phase"], h * sp))
    err_src = abs(got - fd) / abs(fd)
    print(f"source direction: adjoint {got:.8e} oracle FD {fd:.8e} rel err {err_src:.2e}",
          flush=True)
    d = np.sign(r6.standard_normal(fg_base.shape))
    mp, mm = fg_base.copy(), fg_base.copy()
    mp[cloud] = q(fg_base[cloud] * (1 + h * d[cloud]))
    mm[cloud] = q(fg_base[cloud] * (1 - h * d[cloud]))
    fd_f = (oracle_J(amp, phase, mp) - oracle_J(amp, phase, mm)) / 2
    got_f = float(np.sum(r["grad_fg0"][cloud] * (mp - mm)[cloud] / 2))
    err_fg = abs(got_f - fd_f) / abs(fd_f)
    print(f"void-fraction direction: adjoint {got_f:.8e} oracle FD {fd_f:.8e} "
          f"rel err {err_fg:.2e}", flush=True)
    rec.close()
    assert err_src < 1e-3 and err_fg < 1e-3
