"""Pins the FP64 oracle to facts it does not compute itself.

The reference ships no tests, golden vectors or fixtures for this path and
cannot be built here (no nvcc, no Boost, inc/thrust_tuples.cuh missing), so
each oracle function is pinned by a closed form, a physical limit, an
invariant or a brute-force evaluation, chosen so that a dropped term, a wrong
sign or index, or a transposed operand fails.  CPU only.
"""
import math

import numpy as np
import pytest

import cases
import oracle_lib as ol

abi = cases.abi


def sample_config():
    """confs/Sample.txt of the reference, key for key (transducer.plane_wave_*
    are not registered keys and are ignored, as in the reference)."""
    c = abi.default_config()
    c.update(grid_x=300, grid_y=600, grid_lx=30e-3, grid_ly=60e-3, pml_x0=0, pml_x1=1,
             pml_y0=0, pml_y1=1, pml_grid_number=10, pml_sigma=0.10, pml_order=2,
             sim_cfl=0.10, sim_dt0=0.10, sim_n_step_max=20000, sim_t_step_max=0,
             sim_save_interval=100, sim_order=2, sim_smooth_delta_band=2, tr_amp=1.5e-2,
             tr_freq=1.0e6, tr_focal_dist=40.0e-3, tr_pid=1, tr_pid_start_step=5000,
             tr_pid_iterate=2, tr_box_size=30.0e-3, tr_cylindrical=1, tr_on_wave=2,
             tr_off_wave=0, tr_focused_p=1, bub_enable=1, bub_fg0=1.0e-7, bub_R0=6.5e-7,
             bub_P0=0.1e6, bub_T0=293.15, bub_gamma=1.4, bub_sigma=7.2e-2,
             bub_mu=891.0e-6, bub_rho=1.0e3, bub_K0=26.2e-3)
    return c


# ---- a1: index ranges -------------------------------------------------------
def test_array_index_closed_form():
    c = sample_config()
    c.bub_enable = 0
    o = ol.OracleSim(c)
    X, Y = 300, 600
    # order 2: ms=0, me=1, ns=-1, ne=0; cell 1m/2m ranges [0, X+1], node 2n [-1, X+1]
    want = [0, 1, -1, 0, 1, X, 0, X, 0, X + 1, 0, X, 0, X + 1, -1, X + 1,
            1, Y, 0, Y, 0, Y + 1, 0, Y, 0, Y + 1, -1, Y + 1]
    assert o.array_index() == want


# ---- a2: PML profiles ---------------------------------------------------------
@pytest.mark.parametrize("npml,R", [(10, 0.1), (16, 0.1), (8, 0.01)])
def test_pml_profile_closed_form(npml, R):
    n = 64
    c = abi.water_config(n, npml=npml, bubbles=False)
    c.pml_sigma = R
    o = ol.OracleSim(c)
    dx = c.grid_lx / n
    smax = -math.log(R) * 0.5 * 3.0 / (dx * npml)
    mx, nx = o.sigma(0), o.sigma(2)
    assert len(mx) == n + 2 and len(nx) == n + 3
    # cells (index i - ista1m = i): X0 layer i in [1, npml], X1 in [n-npml+1, n]
    for i in range(n + 2):
        if 1 <= i <= npml:
            k = npml - i + 1
        elif n - npml + 1 <= i <= n:
            k = i - n + npml
        else:
            assert mx[i] == 0.0
            continue
        assert mx[i] == pytest.approx(smax * ((k - 0.5) / npml) ** 2, rel=1e-14)
    # nodes (index i - ista2n = i + 1): X0 nodes 0..npml-1, X1 nodes n-npml+1..n
    for i in range(-1, n + 2):
        if 0 <= i <= npml - 1:
            k = npml - i
        elif n - npml + 1 <= i <= n:
            k = i - n + npml
        else:
            assert nx[i + 1] == 0.0
            continue
        assert nx[i + 1] == pytest.approx(smax * (k / npml) ** 2, rel=1e-14)
    # design reflection: the one-way integral of the cell profile over one layer
    # is -ln(R)/2 * (1 - 1/(4 npml^2)) for a quadratic grading
    layer = mx[n - npml + 1:n + 1].sum() * dx
    assert layer == pytest.approx(-math.log(R) / 2 * (1 - 1 / (4 * npml ** 2)), rel=1e-12)


def test_pml_flags_gate_each_side():
    c = sample_config()
    c.bub_enable = 0
    o = ol.OracleSim(c)
    mx, my, nx, ny = (o.sigma(k) for k in range(4))
    assert np.all(mx[:291] == 0) and np.all(mx[291:301] > 0) and mx[301] == 0
    assert np.all(my[:591] == 0) and np.all(my[591:601] > 0)
    assert np.all(nx[:292] == 0) and np.all(nx[292:302] > 0)


# ---- a3: derived scalars ------------------------------------------------------
def test_derived_scalars_sample():
    c = sample_config()
    o = ol.OracleSim(c)
    inf = o.info()
    assert inf["dx"] == 30e-3 / 300 and inf["dy"] == 60e-3 / 600
    assert inf["dt"] == 0.10 * min(inf["dx"], inf["dy"]) / 1500.0
    assert inf["dt0_rp"] == 0.1 * inf["dt"]
    assert inf["fp_x"] == 0.0
    assert inf["fp_y"] == pytest.approx(40e-3 * math.sqrt(3) / 2, rel=1e-15)
    assert inf["omega"] == pytest.approx(2 * math.pi * 1e6, rel=1e-15)
    dcx, dcy, dcz, cdist, pi4r3 = o.scalars()[:5]
    assert dcx == 0.5
    assert dcy == pytest.approx(math.pi / inf["dx"], rel=1e-15)
    assert dcz == pytest.approx(math.pi / inf["dy"], rel=1e-15)
    assert cdist == pytest.approx(40e-3, rel=2e-16 * 4)
    assert pi4r3 == pytest.approx(4 * math.pi / 3, rel=1e-15)
    # default placement: one bubble per cell centre inside the 30 mm box around
    # (0, fp.y): x < 15 mm -> 150 columns, |y - fp.y| < 15 mm -> 300 rows
    assert inf["n_bubbles"] == 150 * 300


# ---- a5: binning brute force -----------------------------------------------------
def test_binning_matches_ieee_floor_at_bin_edges():
    c = abi.water_config(32, npml=4, focal_dist=2e-3)
    c.tr_box_size = 0
    o = ol.OracleSim(c)
    dx = c.grid_lx / 32
    rdx = 1.0 / dx
    xs = []
    for k in range(1, 32):
        for base in (k * dx, (k - 0.5) * dx, (k + 0.5) * dx):
            for d in (-2, -1, 0, 1, 2):
                xs.append(base + d * np.spacing(base))
    xs = np.array(xs)
    ys = xs[::-1].copy()
    o.set_bubbles(xs, ys, np.full(len(xs), 2e-6), None, 1e-6)
    o.prepare()
    for name, v, c0 in (("ibn_x", xs, 0.5), ("ibm_x", xs, 1.0), ("ibn_y", ys, 0.5),
                        ("ibm_y", ys, 1.0)):
        want = np.floor(v * rdx + c0)  # numpy: IEEE multiply then add, no FMA
        np.testing.assert_array_equal(o.bubbles(name), want)


# ---- a14, a13, a18: smooth delta, gather, scatter -------------------------------
@pytest.mark.parametrize("band", [2, 4])
def test_smooth_delta_partition_of_unity(band):
    c = abi.water_config(32, npml=4, focal_dist=2e-3)
    c.sim_smooth_delta_band = band
    o = ol.OracleSim(c)
    dx = c.grid_lx / 32
    mbs = -band // 2 + 1
    r = np.random.default_rng(3)
    for b in r.uniform(5 * dx, 25 * dx, 50):
        ibn = math.floor(b / dx + 0.5)
        w = [o.smooth_delta(0, i, b) for i in range(ibn + mbs, ibn + mbs + band)]
        # the cos argument ((i - 0.5) dx - b) carries ~1e-15 relative rounding
        assert sum(w) == pytest.approx(1.0, abs=2e-14)
        assert min(w) >= 0.0


def _bubble_oracle(seed=5, n=40, nb=60, band=2):
    c, bub = cases.bubbly_case(n, nb=nb, seed=seed, box=1.2e-3)
    c.sim_smooth_delta_band = band
    o = ol.OracleSim(c)
    x, y, R0, nB, fg_in = bub
    o.set_bubbles(x, y, R0, nB, fg_in)
    o.prepare()
    return c, o, bub


def test_gather_reproduces_constant_field():
    c, o, _ = _bubble_oracle()
    P = 123456.75
    o.set_field("p0", np.full(o.field("p0").shape, P))
    np.testing.assert_allclose(o.gather_p0(), P, rtol=2e-14)


@pytest.mark.parametrize("band", [2, 4])
def test_scatter_conserves_volume_and_is_gather_adjoint(band):
    c, o, _ = _bubble_oracle(band=band)
    r = np.random.default_rng(11)
    q = r.uniform(1e-7, 1e-6, o.info()["n_bubbles"])
    fg = o.scatter_values(q)
    assert fg.sum() == pytest.approx(q.sum(), rel=1e-14)
    p = r.normal(0, 1e5, fg.shape)
    o.set_field("p0", p)
    lhs = np.sum(fg * p)            # <S q, p>
    rhs = np.dot(q, o.gather_p0())  # <q, G p>
    assert lhs == pytest.approx(rhs, rel=1e-12)


def test_initial_void_fraction_is_scatter_of_R0():
    c, o, (x, y, R0, _, fg_in) = _bubble_oracle()
    nB = o.bubbles("n_B")
    dx = c.grid_lx / c.grid_x
    # n_B = fg_in dx dy / (4/3 pi R0^3)  =>  each bubble's volume fraction is fg_in
    np.testing.assert_allclose(nB * 4 / 3 * math.pi * R0 ** 3 / dx ** 2, fg_in, rtol=1e-14)
    assert o.field("fg").sum() == pytest.approx(fg_in * len(x), rel=1e-13)


# ---- a10: the void-fraction source term ------------------------------------------
def test_pressure_source_closed_form():
    n = 24
    c = abi.water_config(n, npml=4, bubbles=False)
    c.update(pml_x0=0, pml_x1=0, pml_y0=0, pml_y1=0, tr_focused_p=0)
    o = ol.OracleSim(c)
    o.prepare()
    f, d = 2.0e-5, 3.0e-8
    shp = o.field("fg").shape
    o.set_field("fg", np.full(shp, f))
    o.set_field("fgm", np.full(shp, f + d))  # prediction 3f - 3f + (f + d) = f + d
    o.step(1)
    p0 = o.field("p0")
    want = 1000.0 * 1500.0 ** 2 * d / (1.0 - (f + 0.5 * d))
    # rows 1..n-1: dp0 = rho c^2 dfg / (1 - mean fg)
    np.testing.assert_allclose(p0[1:n, 1:n + 1], want, rtol=1e-12)
    # row n: the reference's y1 mirror (pml.y1 = 0) copies ghost row n+1 (still 0)
    # into interior row n (bubble_CUDA_kernel.cuh:929-932)
    np.testing.assert_array_equal(p0[n, 1:n + 1], 0.0)
    np.testing.assert_allclose(o.field("fg"), f + d, rtol=1e-15)
    np.testing.assert_allclose(o.field("fgn"), f, rtol=0)


# ---- a8-a11: the staggered scheme's exact discrete standing wave -----------------
def standing_wave_case(n=32, m=256, mode=1):
    c = abi.water_config(n, npml=8, bubbles=False)
    c.update(grid_y=m, grid_ly=m * 1e-4, pml_x0=0, pml_x1=0, pml_y0=0, pml_y1=1,
             tr_focused_p=0, ext_source_in_pml=0)
    return c


def standing_wave_init(sim, n, m, mode):
    i = np.arange(0, n + 2)
    prof = np.cos(math.pi * mode * (np.clip(i, 1, n) - 0.5) / n)
    p = np.tile(prof, (m + 2, 1))
    p[m + 1, :] = 0.0
    sim.set_field("p0", p)
    sim.set_field("px", p)
    sim.set_field("py", np.zeros_like(p))
    return prof


def standing_wave_exact(c, n, mode, steps):
    dx = c.grid_lx / n
    dt = c.sim_cfl * dx / 1500.0
    lam = 4 * math.sin(math.pi * mode / (2 * n)) ** 2 / dx ** 2
    th = math.acos(1 - lam * 1500.0 ** 2 * dt ** 2 / 2)
    return math.cos(steps * th + th / 2) / math.cos(th / 2)


@pytest.mark.parametrize("mode", [1, 3])
def test_fdtd_standing_wave_exact_solution(mode):
    n, m = 32, 256
    c = standing_wave_case(n, m, mode)
    o = ol.OracleSim(c)
    o.prepare()
    prof = standing_wave_init(o, n, m, mode)
    done = 0
    for steps in (25, 100, 300):
        o.step(steps - done)
        done = steps
        amp = standing_wave_exact(c, n, mode, steps)
        row = o.field("p0")[10, 1:n + 1]
        np.testing.assert_allclose(row, amp * prof[1:n + 1], atol=1e-11, rtol=0)


# ---- a15-a17: reduced-order bubble model ------------------------------------------
def test_coth_identity():
    for z in [0.3 + 0.3j, 1.7 + 1.7j, 2.5 - 0.4j, 0.11 + 3.0j]:
        assert ol.coth(z) == pytest.approx(1 / np.tanh(z), rel=1e-14)


def test_upsilon_adiabatic_and_isothermal_limits():
    gam = 1.4
    assert ol.upsilon((1 + 1j) * 1e7, gam) == pytest.approx(3 * gam, rel=1e-6)
    assert ol.upsilon((1 + 1j) * 1e-4, gam) == pytest.approx(3.0, rel=1e-8)


def test_series_branches_continuous_at_switch():
    gam = 1.4
    for ph in (0.25 * math.pi, 0.1, 1.3):
        a_in = (0.1 - 1e-13) * complex(math.cos(ph), math.sin(ph))
        a_out = (0.1 + 1e-13) * complex(math.cos(ph), math.sin(ph))
        assert ol.upsilon(a_in, gam) == pytest.approx(ol.upsilon(a_out, gam), rel=1e-11)
        assert ol.solve_lp(a_in, 3e-6) == pytest.approx(ol.solve_lp(a_out, 3e-6), rel=1e-10)
    assert ol.solve_lp(1e-5 * (1 + 1j), 1.0) == pytest.approx(0.2, rel=1e-9)


@pytest.mark.parametrize("K0,kappa", [(1e-14, 1.4), (1e10, 1.0)])
def test_natural_frequency_minnaert_limits(K0, kappa):
    """Adiabatic (K0 -> 0) and isothermal (K0 -> inf) limits of solveOmegaN:
    w^2 = (3 kappa PG0 - 2 sigma / R0) / (rho R0^2) - 4 mu^2 / (rho^2 R0^4)."""
    R0, sig, mu, rho, PL0 = 3e-6, 0.0725, 1e-3, 1000.0, 101325.0
    bp = ol.bubble_params(R0, PL0=PL0, sig=sig, mu=mu, rho=rho, K0=K0)
    w, _ = ol.omega_n(bp.PG0, R0, bp)
    PG0 = PL0 + 2 * sig / R0
    want = math.sqrt((3 * kappa * PG0 - 2 * sig / R0) / (rho * R0 ** 2)
                     - 4 * mu ** 2 / (rho ** 2 * R0 ** 4))
    assert w == pytest.approx(want, rel=1e-6)


def _chain(bp, dt, R_start, PLs):
    st = [R_start, R_start, 0.0, bp.PG0, bp.dt0, 0.0, 0.0, 0.0, 0.0]
    Rt, subs = [], []
    for k in range(len(PLs)):
        st[6], st[7], st[8] = st[7], st[8], PLs[k]
        n, out = ol.rp_integrate(bp, dt, st)
        Rt.append(out[0])
        subs.append(n)
        st[:6] = [out[1], out[2], out[3], out[4], out[5], out[6]]
    return np.array(Rt), np.array(subs)


def test_rp_equilibrium_and_schedule():
    dt = 1e-4 * 0.5 / 1500.0
    for R0 in (1e-6, 3e-6, 5e-6):
        bp = ol.bubble_params(R0, dt_mix=dt)
        Rt, subs = _chain(bp, dt, R0, np.zeros(300))
        assert np.max(np.abs(Rt / R0 - 1)) < 1e-12
        assert subs.min() >= 10 and subs.max() <= 11  # dt / dt0 = 10 sub-steps
        n, out = ol.rp_integrate(bp, dt, [R0, R0, 0, bp.PG0, bp.dt0, 0, 0, 0, 0])
        assert out[6] == 0.0       # the clamped sub-steps end exactly on the sync time
        assert out[4] == pytest.approx(bp.PG0, rel=1e-13)


def _model_ode_frequency(R0, sig, mu, rho, PL0, gam, K0, T0, bp):
    """Free-oscillation frequency of the continuous model that solvePG and
    solveRayleighPlesset discretise, integrated independently (scipy DOP853):
      R R'' + 1.5 R'^2 = (PG - PL0 - 2 sig/R - 4 mu R'/R) / rho
      PG' = -3 gam PG R'/R - A [Re(1/Lp) (Z - PG0) - Im(1/Lp)/w Z'],
      Z = (R/R0)^3 PG,  A = 3 (gam-1) K0 T0 / (R PG0),
    with w = omega_N and Lp frozen at equilibrium."""
    from scipy.integrate import solve_ivp
    PG0 = bp.PG0
    wN, al = ol.omega_n(PG0, R0, bp)
    Lp = ol.solve_lp(al, R0)
    lnr, lni = Lp.real / abs(Lp) ** 2, Lp.imag / abs(Lp) ** 2

    def rhs(t, y):
        R, Rd, PG = y
        A = 3 * (gam - 1) * K0 * T0 / (R * PG0)
        Rx = R / R0
        num = -3 * gam * PG * Rd / R - A * lnr * (Rx ** 3 * PG - PG0) \
            + A * (lni / wN) * 3 * Rx ** 2 * (Rd / R0) * PG
        return [Rd, ((PG - PL0 - 2 * sig / R - 4 * mu * Rd / R) / rho - 1.5 * Rd * Rd) / R,
                num / (1 - A * (lni / wN) * Rx ** 3)]

    T = np.linspace(0, 15 * 2 * math.pi / wN, 3000)
    sol = solve_ivp(rhs, (0, T[-1]), [R0 * (1 + 1e-5), 0.0, PG0], t_eval=T, rtol=1e-11,
                    atol=[1e-20, 1e-14, 1e-9], method="DOP853")
    return _zero_crossing_freq(T, sol.y[0] - R0)


def _zero_crossing_freq(t, d):
    s = np.sign(d)
    idx = np.where(s[1:] != s[:-1])[0]
    tz = t[idx] - d[idx] * (t[idx + 1] - t[idx]) / (d[idx + 1] - d[idx])
    return 2 * math.pi / (2 * np.mean(np.diff(tz)))


def test_rp_free_oscillation_matches_model_ode():
    """The sub-stepped RP + gas model oscillates at the continuous model's
    frequency, in bubble time.  Bubble time runs n_sub * dt0 per mixture step:
    when dt - 9 dt0 rounds above dt0, an 11th sub-step is clamped for R_t but
    the state and `time` advance by the unclamped dt0
    (bubble_CUDA_kernel.cuh:248-256, 512)."""
    dt = 1e-4 * 0.5 / 1500.0
    R0, sig, mu, rho, PL0, gam, K0, T0 = 3e-6, 0.0725, 1e-3, 1000.0, 101325.0, 1.4, 26.2e-3, 293.15
    bp = ol.bubble_params(R0, dt_mix=dt)
    w_ode = _model_ode_frequency(R0, sig, mu, rho, PL0, gam, K0, T0, bp)
    w_n, _ = ol.omega_n(bp.PG0, R0, bp)
    assert w_ode == pytest.approx(w_n, rel=2e-3)   # the model's design: free oscillation at omega_N
    Rt, subs = _chain(bp, dt, R0 * (1 + 1e-5), np.zeros(400))
    assert np.all(subs == 11)
    w_mix = _zero_crossing_freq((np.arange(len(Rt)) + 1) * dt, Rt - R0)
    assert w_mix == pytest.approx(w_ode * 11 * bp.dt0 / dt, rel=2e-3)


def test_radius_update_exact_for_quadratic_motion():
    """solveRayleighPlesset's three-level update reproduces R(t) = R0 + v t + a t^2/2
    and R'(t) exactly when R R'' = R a (sigma = mu = 0)."""
    dtm = 3.0e-9
    dtp = 1.01 * dtm
    R0, v, a, rho, PL = 3e-6, 0.7, 2.0e5, 1000.0, 101325.0
    bp = ol.bubble_params(R0, sig=0.0, mu=0.0, rho=rho, dt_mix=dtp * 10 / 1.01 * 1.01)
    bp.dt0 = dtp
    R = lambda t: R0 + v * t + 0.5 * a * t * t
    Rn, Rm = R(0.0), R(-dtm)
    PG = PL + rho * (Rn * a + 1.5 * v * v)
    io = ol.rp_radius(bp, PG, PL, [0.0, Rn, Rm, v, dtm, 1.0])
    assert io[4] == dtp
    assert io[1] == pytest.approx(R(dtp), rel=1e-14)
    assert io[3] == pytest.approx(v + a * dtp, rel=1e-9)
    assert io[5] == pytest.approx(1.0 - dtp, rel=1e-15)


def test_clock_two_sum():
    c = abi.water_config(16, npml=4, bubbles=False)
    o = ol.OracleSim(c)
    o.step(1000)
    inf = o.info()
    assert abs(inf["time"] - 1000 * inf["dt"]) <= 2 * np.spacing(1000 * inf["dt"])


def test_rp_flop_model_is_per_substep_linear():
    dt = 1e-4 * 0.5 / 1500.0
    bp = ol.bubble_params(3e-6, dt_mix=dt)
    n, _ = ol.rp_integrate(bp, dt, [3e-6, 3e-6, 0, bp.PG0, bp.dt0, 0, 0, 0, 0])
    f = ol.rp_flops_last()
    assert 400 * n < f < 560 * n


def test_picard_skipped_on_first_step():
    """The first pressure evaluation of step 1 has zero residual, so RP
    does not advance on that step (bubble_CUDA.cu:243)."""
    c, b = cases.bubbly_case(48, nb=40)
    o = cases.make(ol.OracleSim, c, b)
    st = o.step(1)
    assert st[0]["n_pic"] == 0 and st[0]["resimax0"] == 0.0
    np.testing.assert_array_equal(o.bubbles("R_p"), b[2])
    np.testing.assert_array_equal(o.bubbles("R_t"), b[2])


def test_conduction_stencil_single_hot_cell():
    """One explicit conduction step of the temperature solve from a single
    hot cell, no sources: the centre loses 4 q and each face neighbour gains
    q = dt k_h (T0 / dx^2) / (rho c_p), with k_h the harmonic mean of the
    water conductivity at T_inf + T0 and T_inf (MixtureEnergyKernel /
    MixtureTemperatureKernel, bubble_CUDA_kernel.cuh:1009-1061); nothing
    else changes."""
    c = cases.acoustic_case(24, amp_pa=0.0)
    c.update(thermal=1)
    o = cases.make(ol.OracleSim, c)
    T = o.field("T")
    ic, jc, T0 = 12, 11, 1.0
    T[jc, ic] = T0
    o.set_field("T", T)
    o.step(1)
    d = o.field("T") - T
    T_inf, dx = 293.15, c.grid_lx / c.grid_x
    dt = o.info()["dt"]
    f = lambda v: float(np.float32(v))  # the reference's single-precision literals
    k = lambda t: 0.76760 + f(7.5390e-3) * t - f(9.8250e-6) * t * t
    cp = (917.5 - f(10.1016) * T_inf + f(0.0454134) * T_inf**2 - 9.07517e-5 * T_inf**3
          + f(6.80700e-8) * T_inf**4) / f(18.0e-3)
    kh = 2 * k(T_inf + T0) * k(T_inf) / (k(T_inf + T0) + k(T_inf))
    q = dt * kh * T0 / dx**2 / (1000.0 * cp)
    assert d[jc, ic] == pytest.approx(-4 * q, rel=1e-12)
    for (j, i) in ((jc + 1, ic), (jc - 1, ic), (jc, ic + 1), (jc, ic - 1)):
        assert d[j, i] == pytest.approx(q, rel=1e-12)
    d[jc, ic] = d[jc + 1, ic] = d[jc - 1, ic] = d[jc, ic + 1] = d[jc, ic - 1] = 0.0
    assert np.abs(d).max() == 0.0
