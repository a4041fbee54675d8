"""Parity of libbubbles_rp (HIP, MI355X) with the FP64 oracle through the C ABI.

Tolerances (relative L2 unless stated):
  FP64 mode, bubbles off        1e-12   (only sin/cos last-ulp differences)
  FP64 mode, cavitating bubbles 1e-9    (+ int64 fixed-point void fraction)
  mixed mode (FP32 fields)      1e-4    pressure; 1e-4 bubble radius
  binning / bubble indices      bit-exact on identical FP64 positions
  Picard iteration counts       identical per step
"""
import os

import numpy as np
import pytest

import cases
import oracle_lib as ol
from cases import rel_l2, rel_linf

pytestmark = pytest.mark.gpu

FIELDS = ("p0", "px", "py", "vx", "vy")


def run_pair(brp, cfg, bubbles=None, steps=100, source=None):
    g = cases.make(brp.Simulation, cfg, bubbles, source)
    o = cases.make(ol.OracleSim, cfg, bubbles, source)
    gs = g.step(steps)
    os_ = o.step(steps)
    return g, o, gs, os_


# Split pressures and velocity components are normalised by the scale of the
# whole pressure / velocity field: in a plane wave p.x and v.x are ~1e-12 Pa
# of noise, and the int64 void fraction rounds the oracle's ~1e-23 FP64 jitter
# of f_g to exactly zero, so their own relative L2 would measure noise.
SCALE = {"px": "p0", "py": "p0", "vx": ("vx", "vy"), "vy": ("vx", "vy")}


def report(tag, g, o, names):
    out = {}
    for n in names:
        a, b = g.field(n), o.field(n)
        if n in SCALE:
            ref = SCALE[n]
            ref = ref if isinstance(ref, tuple) else (ref,)
            scale = max(np.linalg.norm(o.field(r)) for r in ref)
            out[n] = np.linalg.norm(a - b) / scale if scale > 0 else np.linalg.norm(a - b)
        else:
            out[n] = rel_l2(a, b)
    print(tag, {k: f"{v:.2e}" for k, v in out.items()}, flush=True)
    return out


@pytest.mark.parametrize("prec,tol", [(0, 1e-12), (1, 1e-4)])
def test_acoustic_focused_source(brp, prec, tol):
    cfg = cases.acoustic_case(64, precision=prec)
    g, o, gs, os_ = run_pair(brp, cfg, steps=150)
    err = report(f"acoustic prec={prec}", g, o, FIELDS)
    assert np.abs(o.field("p0")).max() > 1e5  # the wave is there
    for n in ("p0", "vx", "vy"):
        assert err[n] < tol, n
    for n in ("px", "py"):
        assert err[n] < 10 * tol, n
    assert [s["resimax0"] > 0 for s in gs] == [s["resimax0"] > 0 for s in os_]


@pytest.mark.parametrize("cyl", [0, 1])
def test_sample_boundaries_fp64(brp, cyl):
    """confs/Sample.txt's boundary set: mirrors on x0/y0, PML on x1/y1, the
    source row written because pml.y0 = 0; optionally cylindrical."""
    cfg = cases.sample_like_case(48, cylindrical=cyl)
    g, o, gs, os_ = run_pair(brp, cfg, steps=150)
    err = report(f"sample cyl={cyl}", g, o, FIELDS + ("fg",))
    for n in FIELDS:
        assert err[n] < 1e-12, n
    assert err["fg"] < 1e-10
    assert [s["n_pic"] for s in gs] == [s["n_pic"] for s in os_]
    for n in ("R_t", "x", "y"):
        assert rel_l2(g.bubbles(n), o.bubbles(n)) < 1e-12


def test_all_mirror_walls_and_plane_sources(brp):
    """No PML at all: every mirror branch of VelocityBoundaryKernel and
    MixtureBoundaryPressureKernel runs, including the reference's right vy ghost
    column (kernel.cuh:808) and y1 pressure mirror (kernel.cuh:929-932), with plane velocity
    and pressure sources on the y0 row."""
    cfg = cases.sample_like_case(40)
    cfg.update(grid_y=40, grid_ly=40e-4, pml_x1=0, pml_y1=0, tr_focused_p=0, tr_plane_p=1,
               tr_plane_v=1, tr_focal_dist=3e-3, tr_box_size=1.0e-3, tr_cylindrical=0)
    g, o, gs, os_ = run_pair(brp, cfg, steps=120)
    err = report("mirrors", g, o, FIELDS + ("fg",))
    for n in FIELDS:
        assert err[n] < 1e-12, n
    assert err["fg"] < 1e-10


@pytest.mark.parametrize("prec,exact", [(0, 1), (0, 0), (1, 0)])
def test_cavitating_bubbles(brp, prec, exact):
    """rp_exact = 1 integrates RP in the reference's operation order; 0 (the
    default) uses the shortened dependency chain, equal up to rounding."""
    cfg, b = cases.bubbly_case(64, nb=160, precision=prec)
    cfg.rp_exact = exact
    g, o, gs, os_ = run_pair(brp, cfg, b, steps=120)
    err = report(f"bubbly prec={prec}", g, o, FIELDS + ("fg",))
    rt = rel_l2(g.bubbles("R_t"), o.bubbles("R_t"))
    print("R_t relL2", rt, "max sub-steps", gs[-1]["rp_max_substeps"], flush=True)
    assert max(s["rp_max_substeps"] for s in os_) > 100  # inertial collapse is exercised
    assert [s["n_pic"] for s in gs] == [s["n_pic"] for s in os_]
    tol = 1e-9 if prec == 0 else 1e-4
    assert err["p0"] < tol and err["fg"] < tol and rt < tol
    for n in ("ibn_x", "ibn_y", "ibm_x", "ibm_y"):
        if prec == 0:
            np.testing.assert_array_equal(g.bubbles(n), o.bubbles(n))
    # the RP sub-step schedule of the last call, bubble by bubble
    sub_mismatch = np.mean(g.bubbles("sub") != o.bubbles("sub"))
    print("sub-step count mismatch fraction", sub_mismatch, flush=True)
    if prec == 0:
        assert sub_mismatch <= (0.0 if exact else 0.05)
    if prec == 0 and exact:
        for n in ("PG", "PL_p", "PL_n", "PL_m", "d1R", "R_p", "R_n", "dt", "re", "Q_B"):
            a, r = g.bubbles(n), o.bubbles(n)
            assert rel_l2(a, r) < 1e-8, n


def test_binning_bit_exact_at_bin_edges(brp):
    cfg = cases.abi.water_config(32, npml=4, focal_dist=2e-3)
    dx = cfg.grid_lx / 32
    xs = []
    for k in range(1, 32):
        for base in (k * dx, (k - 0.5) * dx, (k + 0.5) * dx):
            for d in range(-3, 4):
                xs.append(base + d * np.spacing(base))
    xs = np.array(xs)
    ys = np.roll(xs, 7)
    bub = (xs, ys, np.full(len(xs), 2e-6), None, 1e-6)
    g = cases.make(brp.Simulation, cfg, bub)
    o = cases.make(ol.OracleSim, cfg, bub)
    for n in ("ibn_x", "ibn_y", "ibm_x", "ibm_y"):
        np.testing.assert_array_equal(g.bubbles(n), o.bubbles(n))
    # the initial void fraction of coincident-edge bubbles
    assert rel_l2(g.field("fg"), o.field("fg")) < 1e-12


def test_source_elements(brp):
    """Extension E4: 16 aperture elements with their own amplitude and phase."""
    cfg = cases.acoustic_case(64)
    r = np.random.default_rng(7)
    amp = r.uniform(0.3e6, 0.7e6, 16)
    ph = r.uniform(0, 2 * np.pi, 16)
    g, o, gs, os_ = run_pair(brp, cfg, steps=120, source=(amp, ph))
    err = report("elements", g, o, FIELDS)
    assert err["p0"] < 1e-12


def test_deterministic_and_restore_bit_exact(brp):
    cfg, b = cases.bubbly_case(48, nb=400, seed=3, precision=1)
    runs = []
    for _ in range(2):
        g = cases.make(brp.Simulation, cfg, b)
        g.step(60)
        runs.append((g.field("p0"), g.field("fg"), g.bubbles("R_t")))
        g.close()
    for a, c in zip(*runs):
        np.testing.assert_array_equal(a, c)
    g = cases.make(brp.Simulation, cfg, b)
    g.step(30)
    g.save_state(3)
    g.step(30)
    first = (g.field("p0"), g.bubbles("R_t"), g.info()["time"])
    g.restore_state(3)
    g.step(30)
    second = (g.field("p0"), g.bubbles("R_t"), g.info()["time"])
    np.testing.assert_array_equal(first[0], second[0])
    np.testing.assert_array_equal(first[1], second[1])
    assert first[2] == second[2]
    np.testing.assert_array_equal(first[0], runs[0][0])


def test_clustered_and_lattice_scatter(brp):
    """The RP kernel's per-wavefront LDS aggregation of the void-fraction
    scatter: 6000 bubbles in a 2 x 2-cell cluster (thousands of contributions
    per cell through one table slot) and a 40 x 40 lattice two cells apart
    (every 64-bubble batch touches 256 distinct cells, so tables fill past
    their flush point and contributions fall back to direct atomics).  f_g
    against the oracle and bit-exact between two runs."""
    cfg, _ = cases.bubbly_case(128, nb=10, amp_pa=0.3e6)
    fp_x = cfg.ext_focus_x
    fp_y = cfg.tr_focal_dist * 0.5 * np.sqrt(3.0)
    r = np.random.default_rng(11)
    xc = fp_x + (r.random(6000) - 0.5) * 2e-4
    yc = fp_y + (r.random(6000) - 0.5) * 2e-4
    gx, gy = np.meshgrid(np.arange(40) * 2e-4, np.arange(40) * 2e-4)
    xl = fp_x - 4e-3 + gx.ravel() + 0.3e-4
    yl = fp_y - 4e-3 + gy.ravel() + 0.3e-4
    x, y = np.concatenate([xc, xl]), np.concatenate([yc, yl])
    R0 = 1e-6 + 4e-6 * r.random(len(x))
    bub = (x, y, R0, None, 1e-9)
    g, o, gs, os_ = run_pair(brp, cfg, bub, steps=40)
    err = report("clustered+lattice", g, o, ("p0", "fg"))
    assert [s["n_pic"] for s in gs] == [s["n_pic"] for s in os_]
    assert err["fg"] < 1e-10 and err["p0"] < 1e-10
    g2 = cases.make(brp.Simulation, cfg, bub)
    g2.step(40)
    np.testing.assert_array_equal(g.field("fg"), g2.field("fg"))
    np.testing.assert_array_equal(g.field("p0"), g2.field("p0"))


def test_cavitating_4096_bubbles_multi_workgroup(brp):
    """4096 cavitating bubbles (16 RP workgroups, up to 16 dedicated
    wavefronts, batched refills across many wavefronts) against the oracle,
    FP64, 120 steps."""
    cfg, b = cases.bubbly_case(64, nb=4096, seed=7)
    g, o, gs, os_ = run_pair(brp, cfg, b, steps=120)
    err = report("bubbly 4096", g, o, FIELDS + ("fg",))
    rt = rel_l2(g.bubbles("R_t"), o.bubbles("R_t"))
    print("R_t relL2", rt, "max sub-steps", max(s["rp_max_substeps"] for s in os_), flush=True)
    assert max(s["rp_max_substeps"] for s in os_) > 512
    assert [s["n_pic"] for s in gs] == [s["n_pic"] for s in os_]
    assert err["p0"] < 1e-9 and err["fg"] < 1e-9 and rt < 1e-9
    assert np.mean(g.bubbles("sub") != o.bubbles("sub")) <= 0.05


def test_many_wavefront_schedule_bit_exact(brp):
    """65 536 bubbles through an inertial collapse: the RP launch runs on a
    whole-device grid with dedicated wavefronts for the longest chains,
    batched refills and both grid shapes as the sub-step counts change, and
    the queue order changes every launch.  The results must not depend on
    that schedule: two runs agree bit for bit (fields, radii, sub-step
    counts, Picard counts)."""
    cfg, b = cases.bubbly_case(64, nb=65536, seed=7)
    runs = []
    for _ in range(2):
        g = cases.make(brp.Simulation, cfg, b)
        st = g.step(120)
        runs.append((st, g.field("p0"), g.field("fg"), g.bubbles("R_t"), g.bubbles("sub")))
        g.close()
    (s1, *a1), (s2, *a2) = runs
    lean = sum(s["rp_lean_launches"] for s in s1)
    launches = sum(s["n_pic"] for s in s1)
    print("max sub-steps", max(s["rp_max_substeps"] for s in s1), "lean launches", lean, "of",
          launches, flush=True)
    assert max(s["rp_max_substeps"] for s in s1) > 512  # long chains, dedicated wavefronts
    assert 0 < lean < launches  # both grid shapes were used
    assert [s["n_pic"] for s in s1] == [s["n_pic"] for s in s2]
    assert [s["rp_lean_launches"] for s in s1] == [s["rp_lean_launches"] for s in s2]
    for x, y in zip(a1, a2):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("nb", [400, 65536])
def test_kept_rp_results_equal_a_fresh_integration(brp, nb):
    """Mixed precision, cavitating: from a step's second Picard iteration on,
    a bubble whose gathered pressure is the same as in the iteration before,
    bit for bit (most of them once the FP32 pressure field stops changing),
    keeps that iteration's RP results instead of integrating again.  In
    frozen-schedule record mode a separate kernel integrates every bubble of
    every launch again and compares the sub-step count and all results with
    what the launch left, bit for bit; a difference fails the step.  The run
    must keep results in some iterations and show no difference, and its
    fields, radii and Picard counts must equal those of the same run outside
    record mode."""
    cfg, b = cases.bubbly_case(64, nb=nb, seed=7, precision=brp.abi.PREC_MIXED)
    rec = cases.make(brp.Simulation, cfg, b)
    rec.schedule_mode(1)
    s1 = rec.step(120)  # raises on a recording difference (BRP_ERR_STATE)
    plain = cases.make(brp.Simulation, cfg, b)
    s2 = plain.step(120)
    kept = sum(s["rp_reused"] for s in s1)
    calls = nb * sum(s["n_pic"] for s in s1)
    print(f"{nb} bubbles: {kept} of {calls} RP calls kept the previous iteration's results",
          flush=True)
    assert kept > 0
    assert [s["n_pic"] for s in s1] == [s["n_pic"] for s in s2]
    assert [s["rp_reused"] for s in s1] == [s["rp_reused"] for s in s2]
    for n in ("p0", "fg"):
        np.testing.assert_array_equal(rec.field(n), plain.field(n))
    for n in ("R_t", "R_p", "PG", "sub"):
        np.testing.assert_array_equal(rec.bubbles(n), plain.bubbles(n))


def test_wide_band_and_edge_bubbles(brp):
    """smooth_delta_band = 4, bubbles straddling the domain edges and outside it."""
    cfg, _ = cases.bubbly_case(48, nb=10)
    cfg.sim_smooth_delta_band = 4
    L = cfg.grid_lx
    x = np.array([0.3e-4, L - 0.2e-4, L / 2, -1e-3, L / 3, 2.1e-4, L + 1e-3])
    y = np.array([L / 2, L / 2, 0.4e-4, L / 2, L - 0.3e-4, 1.7e-4, L / 2])
    bub = (x, y, np.full(len(x), 3e-6), None, 1e-5)
    g, o, gs, os_ = run_pair(brp, cfg, bub, steps=80)
    err = report("band4/edges", g, o, FIELDS + ("fg",))
    assert err["p0"] < 1e-12 and err["fg"] < 1e-9


def test_error_paths(brp):
    cfg, b = cases.bubbly_case(48, nb=50)
    cfg.max_picard = 1
    g = cases.make(brp.Simulation, cfg, b)
    with pytest.raises(brp.BrpError) as e:
        g.step(100)
    assert e.value.code == brp.abi.ERR_PICARD_CAP
    cfg, b = cases.bubbly_case(48, nb=50)
    cfg.max_substeps = 5
    g = cases.make(brp.Simulation, cfg, b)
    with pytest.raises(brp.BrpError) as e:
        g.step(100)
    assert e.value.code == brp.abi.ERR_SUBSTEP_CAP
    bad = cases.acoustic_case(32)
    bad.sim_order = 4
    with pytest.raises(brp.BrpError):
        brp.Simulation(bad)


def test_nonfinite_pressure_is_reported(brp):
    """BRP_ERR_NONFINITE: a NaN planted in the pressure at the bubble box
    reaches the velocities, the pressures and, through the gather, the RP
    inputs; the step that meets it reports it and every kernel still ends
    (the sub-step sizes stay finite, so each RP integration ends)."""
    cfg, b = cases.bubbly_case(48, nb=50)
    g = cases.make(brp.Simulation, cfg, b)
    g.step(5)
    p = g.field("p0")
    ny, nx = p.shape
    p[ny // 2, nx // 2] = np.nan
    g.set_field("p0", p)
    with pytest.raises(brp.BrpError) as e:
        g.step(20)
    assert e.value.code == brp.abi.ERR_NONFINITE, str(e.value)
    g.close()


def test_zero_radius_hits_the_substep_cap_like_the_oracle(brp):
    """A bubble radius set to 0: the step-size rule 0.01 R / |dR/dt|
    (kernel.cuh:245-246) gives a sub-step of 0, so the integration can never
    reach the sync time; GPU and oracle both stop at max_substeps and report
    BRP_ERR_SUBSTEP_CAP (the GPU also flags the non-finite radius it wrote)."""
    cfg, b = cases.bubbly_case(48, nb=50, amp_pa=1.5e6)
    cfg.max_substeps = 3000
    codes = []
    for cls in (brp.Simulation, ol.OracleSim):
        s = cases.make(cls, cfg, b)
        s.step(100)  # the wavefront has reached the bubbles: dR/dt != 0
        R = s.bubbles("R_p")
        R[7] = 0.0
        s.set_bubble_array("R_p", R)
        with pytest.raises(brp.BrpError) as e:
            s.step(5)
        codes.append(e.value.code)
        s.close()
    assert codes == [brp.abi.ERR_SUBSTEP_CAP] * 2


def test_standing_wave_closed_form(brp):
    """The GPU FDTD reproduces the staggered scheme's exact discrete
    standing-wave solution (the oracle pin of the same case)."""
    import test_oracle_pins as pins
    n, m = 32, 256
    c = pins.standing_wave_case(n, m, 1)
    g = brp.Simulation(c)
    g.prepare()
    prof = pins.standing_wave_init(g, n, m, 1)
    g.step(300)
    amp = pins.standing_wave_exact(c, n, 1, 300)
    np.testing.assert_allclose(g.field("p0")[10, 1:n + 1], amp * prof[1:n + 1], atol=1e-11)


def test_golden_fixture(brp):
    """The committed 32^2 / 16-bubble / 50-step fixture (tests/golden)."""
    path = os.path.join(os.path.dirname(__file__), "golden", "c1mini.npz")
    z = np.load(path)
    cfg = cases.abi.Config()
    for k in cases.abi.Config._fields_:
        if k[0] in z.files:
            setattr(cfg, k[0], z[k[0]].item())
    for prec, tol in ((0, 1e-9), (1, 1e-4)):
        cfg.precision = prec
        g = cases.make(brp.Simulation, cfg, (z["bx"], z["by"], z["bR0"], None, z["fg_in"].item()))
        st = g.step(int(z["steps"]))
        assert [s["n_pic"] for s in st] == list(z["n_pic"])
        assert rel_l2(g.field("p0"), z["p0"]) < tol
        assert rel_l2(g.field("fg"), z["fg"]) < tol
        assert rel_l2(g.bubbles("R_t"), z["R_t"]) < tol


def _mirror_case():
    cfg = cases.sample_like_case(40)
    cfg.update(grid_y=40, grid_ly=40e-4, pml_x1=0, pml_y1=0, tr_focused_p=0, tr_plane_p=1,
               tr_plane_v=1, tr_focal_dist=3e-3, tr_box_size=1.0e-3, tr_cylindrical=0)
    return cfg, None


@pytest.mark.parametrize("case,prec,tol", [
    ("mirrors", 0, 1e-11),       # all four temperature boundary parts, absorption, bubble heat
    ("bubbly", 0, 1e-8),         # cavitating bubbles: Q_B from the fast RP chain
    ("bubbly", 1, 1e-3),         # mixed: the absorption term squares FP32 pressure rates
    ("cylindrical", 0, 1e-9),    # bubble heat / x
])
def test_temperature(brp, case, prec, tol):
    """calculate_temperature (bubble_CUDA_kernel.cuh:955-1102,1347-1380): the
    T field after 120 steps against the oracle."""
    if case == "mirrors":
        cfg, b = _mirror_case()
    elif case == "bubbly":
        cfg, b = cases.bubbly_case(64, nb=160, precision=prec)
    else:
        cfg, b = cases.sample_like_case(48, cylindrical=1), None
    cfg.update(thermal=1, precision=prec)
    g, o, gs, os_ = run_pair(brp, cfg, b, steps=120)
    T_g, T_o = g.field("T"), o.field("T")
    assert np.abs(T_o).max() > 0
    err = rel_l2(T_g, T_o)
    print(f"T {case} prec={prec}: rel L2 {err:.2e}, max {np.abs(T_o).max():.3e} K", flush=True)
    assert err < tol
    if prec == 0:  # the temperature is one-way: the acoustic fields are unchanged
        assert rel_l2(g.field("p0"), o.field("p0")) < (1e-12 if b is None else 1e-9)


def test_void_fraction_accumulation_schemes(brp):
    """The void-fraction scatter of one population (a 6000-bubble cluster in
    2 x 2 cells, thousands of contributions per cell, plus a lattice) with
    the four accumulation schemes of brp_scatter_variant, against the
    oracle's serial FP64 sum of the same contributions: the library's int64
    fixed point is exact to its 2^-60 resolution (order-free), FP64 atomics
    to FP64 rounding (order-dependent), compensated FP32 to FP32 rounding of
    the final value, plain FP32 atomics only to the FP32 rounding of every
    add."""
    cfg, _ = cases.bubbly_case(128, nb=10, amp_pa=0.3e6)
    fp_x = cfg.ext_focus_x
    fp_y = cfg.tr_focal_dist * 0.5 * np.sqrt(3.0)
    r = np.random.default_rng(12)
    xc = fp_x + (r.random(6000) - 0.5) * 2e-4
    yc = fp_y + (r.random(6000) - 0.5) * 2e-4
    gx, gy = np.meshgrid(np.arange(40) * 2e-4, np.arange(40) * 2e-4)
    x = np.concatenate([xc, fp_x - 4e-3 + gx.ravel() + 0.3e-4])
    y = np.concatenate([yc, fp_y - 4e-3 + gy.ravel() + 0.3e-4])
    R0 = 1e-6 + 4e-6 * r.random(len(x))
    bub = (x, y, R0, None, 1e-9)
    g = cases.make(brp.Simulation, cfg, bub)
    st = g.step(180)  # the wavefront has reached the bubbles
    assert st[-1]["n_pic"] > 0
    o = cases.make(ol.OracleSim, cfg, bub)
    for n in ("x", "y", "R_t"):
        o.set_bubble_array(n, g.bubbles(n))
    dx = cfg.grid_lx / cfg.grid_x
    q = 4.0 / 3.0 * np.pi * o.bubbles("n_B") * g.bubbles("R_t") ** 3 / (dx * dx)
    ref = o.scatter_values(q)
    errs = {}
    for mode, tol in ((brp.abi.VF_INT64, 1e-9), (brp.abi.VF_FP64_ATOMIC, 1e-14),
                      (brp.abi.VF_FP32_COMPENSATED, 1e-7), (brp.abi.VF_FP32_ATOMIC, 1e-3)):
        v = g.scatter_variant(mode)
        errs[mode] = (rel_l2(v, ref), rel_linf(v, ref))
        assert errs[mode][0] < tol, (mode, errs[mode])
    print("accumulation schemes vs serial FP64 (rel L2, rel Linf):",
          {k: f"{a:.2e}/{b:.2e}" for k, (a, b) in errs.items()}, flush=True)
    assert errs[brp.abi.VF_FP32_COMPENSATED][0] < 0.1 * errs[brp.abi.VF_FP32_ATOMIC][0]
    # the product's own f_g is the int64 scheme of the last scatter
    np.testing.assert_array_equal(g.scatter_variant(brp.abi.VF_INT64), g.field("fg"))
    g.close()
    o.close()
