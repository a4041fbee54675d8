"""Pins of the oracle's field, source and coupling functions against facts it
does not compute itself (CPU only): the focused source's delay law and
gating (bubble_CUDA_kernel.cuh:903-925), the PML updates (635-677, 816-867)
one step at a time and the reflection they produce, the velocity gather's
half-cell shift (398-448, Q7), exact energy conservation and the 2-D
dispersion relation of the staggered scheme in a closed box, mirror symmetry
of a symmetric source, and the heat balance of the temperature solve
(972-1061)."""
import math

import numpy as np
import pytest

import cases
import oracle_lib as ol

abi = cases.abi
RHO, C = 1000.0, 1500.0


def _closed_box(n, m=None):
    """Rigid walls on all four sides, no PML, no source.  grid.x = grid.y + 1
    makes the reference's y1 pressure mirror (it loops over rows X and X+1,
    kernel.cuh:929-932) a true mirror: row Y+1 copies row Y and row Y+2 is
    outside the array."""
    m = n - 1 if m is None else m
    c = abi.water_config(n, npml=4, bubbles=False, amp_pa=0.0)
    c.update(grid_y=m, grid_ly=m * 1e-4, pml_x0=0, pml_x1=0, pml_y0=0, pml_y1=0,
             tr_focused_p=0, ext_source_in_pml=0)
    return c


# ---- a11: the focused source ------------------------------------------------
def _source_case(n=128, on=1000000, off=0, n_elem=0):
    c = cases.acoustic_case(n, amp_pa=0.5e6)
    c.update(tr_on_wave=on, tr_off_wave=off)
    o = ol.OracleSim(c)
    if n_elem:
        r = cases.rng(9)
        amp = 0.2e6 + 0.6e6 * r.random(n_elem)
        ph = 2 * math.pi * r.random(n_elem)
        o.set_source(amp, ph)
    o.prepare()
    dx = c.grid_lx / n
    fx, fy = c.ext_focus_x, c.tr_focal_dist * 0.5 * math.sqrt(3.0)
    xm = (np.arange(1, n + 1) - 0.5) * dx
    s1 = np.hypot(fx - xm, fy + 0.5 * dx)  # cell (i, 0) centre at y = -dy/2
    return c, o, s1


def _row0(o, n):
    return o.field("p0")[0, 1:n + 1].copy()


def test_source_delay_law_focuses_every_aperture_cell():
    """Each aperture cell's signal is a sinusoid delayed by tau_i such that
    tau_i + |fp - x_i| / c is the same for every cell (= f_dist / c: all
    wavelets arrive at the focus in phase); cells farther than f_dist from the
    focus stay at zero; before its delay a cell is silent."""
    n = 128
    c, o, s1 = _source_case(n)
    w = 2 * math.pi * c.tr_freq
    A = RHO * C * c.tr_amp
    f_dist = c.tr_focal_dist
    ap = s1 <= f_dist
    assert 60 < ap.sum() < n  # a proper sub-aperture
    o.step(4)
    t_early = o.clock()[1]
    early = _row0(o, n)
    tau = (f_dist - s1) / C
    assert np.all(early[~ap] == 0.0)
    assert np.all(early[ap & (tau > t_early + 1e-12)] == 0.0)
    assert np.all(early[ap & (tau < t_early - 1e-9)] != 0.0)
    o.step(60)
    t1, v1 = o.clock()[1], _row0(o, n)
    o.step(1)
    t2, v2 = o.clock()[1], _row0(o, n)
    assert np.all(v1[~ap] == 0.0) and np.all(v2[~ap] == 0.0)
    # v = A sin(w (t - tau)): recover the phase from two samples
    wd = w * (t2 - t1)
    s = v1[ap] / A
    cth = (v2[ap] / A - s * math.cos(wd)) / math.sin(wd)
    np.testing.assert_allclose(np.hypot(s, cth), 1.0, atol=1e-9)  # amplitude A
    tau_meas = t1 - np.arctan2(s, cth) / w
    arrive = tau_meas + s1[ap] / C - f_dist / C  # should be 0 modulo one period
    T = 1.0 / c.tr_freq
    arrive = (arrive + 0.5 * T) % T - 0.5 * T
    np.testing.assert_allclose(arrive, 0.0, atol=1e-15)


def test_source_gating_on_off_cycles():
    """With on_wave = 2, off_wave = 1 a cell emits for two periods of every
    three, counted from its own delayed start."""
    n = 128
    c, o, s1 = _source_case(n, on=2, off=1)
    ap = s1 <= c.tr_focal_dist
    tau = (c.tr_focal_dist - s1) / C
    T = 1.0 / c.tr_freq
    checked = 0
    for steps in (80, 70, 45, 50):
        o.step(steps)
        t = o.clock()[1]
        v = _row0(o, n)
        ph = np.mod(np.maximum(t - tau, 0.0), 3 * T)
        on = ap & (ph < 2 * T - 1e-12) & (t - tau > 1e-12)
        off = ap & (ph > 2 * T + 1e-12)
        assert np.all(v[off] == 0.0)
        assert np.all(v[on & (np.abs(np.sin(2 * math.pi * ph / T)) > 1e-6)] != 0.0)
        checked += off.sum()
    assert checked > 50  # the off phase was actually seen


def test_source_elements_partition_the_aperture():
    """Extension E4: n_elem contiguous, equal parts of the aperture in cell
    order, each with its own amplitude and phase on top of the delay law."""
    n, ne = 128, 4
    c, o, s1 = _source_case(n, n_elem=ne)
    r = cases.rng(9)
    amp = 0.2e6 + 0.6e6 * r.random(ne)
    ph = 2 * math.pi * r.random(ne)
    ap = np.nonzero(s1 <= c.tr_focal_dist)[0]
    w = 2 * math.pi * c.tr_freq
    o.step(64)
    t1, v1 = o.clock()[1], _row0(o, n)
    o.step(1)
    t2, v2 = o.clock()[1], _row0(o, n)
    wd = w * (t2 - t1)
    tau = (c.tr_focal_dist - s1) / C
    for k, i in enumerate(ap):
        e = k * ne // len(ap)
        a = v1[i]
        b = (v2[i] - v1[i] * math.cos(wd)) / math.sin(wd)
        assert math.hypot(a, b) == pytest.approx(amp[e], rel=1e-9)
        d = math.atan2(a, b) - (w * (t1 - tau[i]) + ph[e])
        assert abs((d + math.pi) % (2 * math.pi) - math.pi) < 1e-9


# ---- a8, a10: the PML updates one step at a time ----------------------------
def _pml_case(n=32, npml=8):
    c = cases.acoustic_case(n, amp_pa=0.0)
    c.update(pml_grid_number=npml, tr_focused_p=0)
    o = ol.OracleSim(c)
    o.prepare()
    return c, o


def test_pml_pressure_damping_is_undivided():
    """v = 0 and uniform split pressures P, Q: after one step p.x = P (1 - s2x)
    and p.y = Q (1 - s2y), with s2 = dt c sigma_m / 2 of the cell's own column
    (x) or row (y) -- the damping factor is not divided by (1 + s2) (Q2) and
    each split part sees only its own direction's profile."""
    n = 32
    c, o = _pml_case(n)
    P, Q = 3.0e5, -1.25e5
    shp = o.field("p0").shape
    o.set_field("px", np.full(shp, P))
    o.set_field("py", np.full(shp, Q))
    o.set_field("p0", np.full(shp, P + Q))
    dt = o.info()["dt"]
    sx, sy = o.sigma(0), o.sigma(1)  # cell profiles, index i (or j) = 0..n+1
    assert sx.max() > 0 and sx[n // 2] == 0.0
    o.step(1)
    px, py = o.field("px"), o.field("py")
    s2x = 0.5 * dt * C * sx
    s2y = 0.5 * dt * C * sy
    np.testing.assert_allclose(px[1:n + 1, 1:n + 1], np.tile(P * (1 - s2x[1:n + 1]), (n, 1)),
                               rtol=1e-14)
    np.testing.assert_allclose(py[1:n + 1, 1:n + 1],
                               np.tile((Q * (1 - s2y[1:n + 1]))[:, None], (1, n)), rtol=1e-14)


def test_pml_velocity_damping_is_divided():
    """Uniform p0 and vx = V, vy = 0: after one step vx = V (1 - s3)/(1 + s3)
    with s3 = dt c sigma_n / 2 of the node's own column; vy stays 0."""
    n = 32
    c, o = _pml_case(n)
    V = 0.75
    o.set_field("vx", np.full(o.field("vx").shape, V))
    dt = o.info()["dt"]
    sn = o.sigma(2)  # node profile, index in + 1 for in = -1..n+1
    o.step(1)
    vx = o.field("vx")  # columns in = -1..n+1, rows j = 0..n+1
    s3 = 0.5 * dt * C * sn[1:n + 2]  # nodes 0..n
    want = V * (1 - s3) / (1 + s3)
    np.testing.assert_allclose(vx[1:n + 1, 1:n + 2], np.tile(want, (n, 1)), rtol=1e-14)
    assert np.all(o.field("vy") == 0.0)


def _packet_reflection(axis, n_long=1500, npml=200, lam=24.0, env=40.0, x0=300.0, probe=1000,
                       steps=4000):
    """A right-going narrow-band wave packet (carrier wavelength lam cells,
    Gaussian envelope of env cells) on a 1-D strip (rigid walls across, a
    PML of npml cells at the far end, whose last ghost holds p = 0): the
    ratio of the reflected to the incident envelope peak at a probe cell, and
    the packet's discrete group velocity over c."""
    c = abi.water_config(n_long, npml=npml, bubbles=False, amp_pa=0.0)
    if axis == 0:
        c.update(grid_y=6, grid_ly=6e-4, pml_x0=0, pml_x1=1, pml_y0=0, pml_y1=0,
                 tr_focused_p=0, ext_source_in_pml=0)
    else:
        c.update(grid_x=6, grid_lx=6e-4, pml_x0=0, pml_x1=0, pml_y0=0, pml_y1=1,
                 tr_focused_p=0, ext_source_in_pml=0)
    o = ol.OracleSim(c)
    o.prepare()
    k = 2 * math.pi / lam
    wdt = 2 * math.asin(c.sim_cfl * math.sin(k / 2))  # sin(w dt/2) = cfl sin(k dx/2)
    pkt = lambda xc, shift: (np.exp(-0.5 * ((xc - x0) / env) ** 2)
                             * np.cos(k * (xc - x0) + shift))
    p = pkt(np.arange(0, n_long + 2) - 0.5, 0.0)  # cell centres
    p[0] = p[n_long + 1] = 0.0
    v = pkt(np.arange(-1, n_long + 2).astype(float), 0.5 * wdt) / (RHO * C)  # nodes, t - dt/2
    v[0] = 0.0
    if axis == 0:
        P = np.tile(p, (8, 1))
        o.set_field("p0", P)
        o.set_field("px", P)
        o.set_field("vx", np.tile(v, (8, 1)))
    else:
        P = np.tile(p[:, None], (1, 8))
        o.set_field("p0", P)
        o.set_field("py", P)
        o.set_field("vy", np.tile(v[:, None], (1, 8)))
    trace = np.zeros(steps)
    for s in range(steps):
        o.step(1)
        f = o.field("p0")
        trace[s] = f[3, probe] if axis == 0 else f[probe, 3]
    split = int((probe - x0) / c.sim_cfl + 1200)
    vg = math.cos(k / 2) / math.cos(wdt / 2)
    return np.abs(trace[:split]).max(), np.abs(trace[split:]).max(), vg, c


@pytest.mark.parametrize("axis", [0, 1])
def test_pml_reflection_against_its_design(axis):
    """init_sigma grades the layer so that the integral of sigma across it is
    -ln(pml.sigma)/2: a symmetric split-field PML, damping pressure and
    velocity both at the rate c sigma, would return a normally incident wave
    with amplitude pml.sigma after the round trip.  The reference's pressure
    update damps p by (1 - s2) ~ exp(-c sigma dt / 2) per step (only the flux
    term is divided by 1 + s2, Q2), half the velocity's rate
    (1 - s3)/(1 + s3) ~ exp(-c sigma dt), so a wave decays in time at 3/4 of
    c sigma, in space at 3/4 sigma c / v_g, and the round trip returns
    pml.sigma^(3/4 c / v_g) (v_g: the discrete group velocity).  A gently
    graded layer and a narrow-band packet keep the layer's own grading
    reflections and the damping per period small."""
    inc, refl, vg, c = _packet_reflection(axis)
    ratio = refl / inc
    want = c.pml_sigma ** (0.75 / vg)
    print(f"axis {axis}: incident {inc:.4f}, reflected {refl:.4f}, ratio {ratio:.4f}, "
          f"sigma^(3/4 c/v_g) {want:.4f}, sigma {c.pml_sigma:.4f}")
    assert inc == pytest.approx(1.0, abs=0.01)
    assert ratio == pytest.approx(want, rel=0.04)
    assert ratio > 1.5 * c.pml_sigma  # not the symmetric PML's round trip


# ---- a6: the velocity gather's half-cell shift (Q7) -------------------------
def test_velocity_gather_half_cell_shift():
    """The bubble velocity gathers vx (nodes at x = i dx) and vy (nodes at
    y = j dy) with the cell-centred smooth delta (Q7): a bubble at a cell
    centre ((k - 1/2) dx, (l - 1/2) dy) picks up vx(k, l) and vy(k, l) alone
    -- the node half a cell to its upper right, not the average of the two
    nodes beside it -- and a bubble on a node (k dx, (l - 1/2) dy) the mean of
    vx(k, l) and vx(k + 1, l)."""
    n = 40
    c = cases.acoustic_case(n, amp_pa=0.0)
    c.update(pml_grid_number=3, tr_focused_p=0, bub_enable=1, bub_fg0=1e-12, bub_R0=3e-6)
    dx = c.grid_lx / n
    ks = np.array([12, 17, 25, 30])
    ls = np.array([14, 21, 19, 28])
    x = np.concatenate([(ks - 0.5) * dx, ks * dx])
    y = np.concatenate([(ls - 0.5) * dx, (ls - 0.5) * dx])
    o = ol.OracleSim(c)
    o.set_bubbles(x, y, np.full(len(x), 3e-6), None, 1e-12)
    o.prepare()
    r = cases.rng(3)
    vx = r.standard_normal(o.field("vx").shape)  # [j, i + 1], i = -1..n+1
    vy = r.standard_normal(o.field("vy").shape)  # [j + 1, i], j = -1..n+1
    o.set_field("vx", vx)
    o.set_field("vy", vy)
    o.step(1)  # p0 = 0 and no PML where the bubbles are: vx, vy are unchanged
    np.testing.assert_array_equal(o.field("vx")[5:35, 5:35], vx[5:35, 5:35])
    vbx, vby = o.bubbles("vx"), o.bubbles("vy")
    m = len(ks)
    np.testing.assert_allclose(vbx[:m], vx[ls, ks + 1], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(vby[:m], vy[ls + 1, ks], rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(vbx[m:], 0.5 * (vx[ls, ks + 1] + vx[ls, ks + 2]), rtol=1e-13,
                               atol=1e-15)
    # the motion moves each bubble by its gathered velocity times dt
    dt = o.info()["dt"]
    np.testing.assert_array_equal(o.bubbles("x"), x + vbx * dt)


# ---- a8-a11: exact energy conservation and the 2-D dispersion relation ------
def _energy(o, p, vx_a, vx_b, vy_a, vy_b, n, m):
    """E = sum p^2 / (2 rho c^2) + rho/2 sum v^(n-1/2) v^(n+1/2) over the
    interior cells and the interior nodes (the walls' nodes are 0)."""
    e = np.sum(p[1:m + 1, 1:n + 1] ** 2) / (2 * RHO * C * C)
    e += 0.5 * RHO * np.sum(vx_a[1:m + 1, 2:n + 1] * vx_b[1:m + 1, 2:n + 1])
    e += 0.5 * RHO * np.sum(vy_a[2:m + 1, 1:n + 1] * vy_b[2:m + 1, 1:n + 1])
    return e


def test_closed_box_conserves_discrete_energy():
    """Rigid walls, no PML, no source, random smooth initial pressure: the
    staggered scheme's discrete energy (pressure at step n, the product of
    the velocities at n -/+ 1/2) is conserved to rounding over 400 steps."""
    n, m = 41, 40
    c = _closed_box(n, m)
    o = ol.OracleSim(c)
    o.prepare()
    r = cases.rng(12)
    shp = o.field("p0").shape
    p = np.zeros(shp)
    xx, yy = np.meshgrid(np.arange(shp[1]), np.arange(shp[0]))
    for _ in range(5):
        cx, cy, s = r.uniform(8, 33), r.uniform(8, 32), r.uniform(2, 5)
        p += r.uniform(-1, 1) * 1e5 * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    o.set_field("p0", p)
    o.set_field("px", p)
    o.step(1)
    prev = (o.field("p0"), o.field("vx"), o.field("vy"))
    es = []
    for _ in range(400):
        o.step(1)
        vx, vy = o.field("vx"), o.field("vy")
        es.append(_energy(o, prev[0], prev[1], vx, prev[2], vy, n, m))
        prev = (o.field("p0"), vx, vy)
    es = np.array(es)
    assert np.ptp(es) / es[0] < 1e-12, np.ptp(es) / es[0]


@pytest.mark.parametrize("mx,my", [(1, 1), (3, 2)])
def test_closed_box_mode_follows_2d_dispersion(mx, my):
    """A standing mode cos(pi mx (i - 1/2)/N) cos(pi my (j - 1/2)/M) of the
    closed box oscillates at the staggered scheme's 2-D dispersion relation,
    sin^2(w dt/2) = (c dt)^2 (sin^2(pi mx / 2N)/dx^2 + sin^2(pi my / 2M)/dy^2),
    with p^n / p^0 = cos(n w dt + w dt/2) / cos(w dt/2)."""
    n, m = 33, 32
    c = _closed_box(n, m)
    o = ol.OracleSim(c)
    o.prepare()
    dx = c.grid_lx / n
    dt = o.info()["dt"]
    i = np.clip(np.arange(0, n + 2), 1, n)
    j = np.clip(np.arange(0, m + 2), 1, m)
    prof = np.outer(np.cos(math.pi * my * (j - 0.5) / m), np.cos(math.pi * mx * (i - 0.5) / n))
    o.set_field("p0", prof)
    o.set_field("px", prof)
    lam = 4 * (math.sin(math.pi * mx / (2 * n)) ** 2 + math.sin(math.pi * my / (2 * m)) ** 2)
    th = 2 * math.asin(math.sqrt(lam) * C * dt / (2 * dx))
    done = 0
    for steps in (30, 170, 400):
        o.step(steps - done)
        done = steps
        amp = math.cos(steps * th + th / 2) / math.cos(th / 2)
        np.testing.assert_allclose(o.field("p0")[1:m + 1, 1:n + 1], amp * prof[1:m + 1, 1:n + 1],
                                   atol=1e-11)


# ---- mirror symmetry ----------------------------------------------------------
def test_symmetric_source_gives_a_mirror_symmetric_field():
    """The focused source centred in x on a PML-bounded square: p0 is
    mirror-symmetric about the focus column and vx antisymmetric, to
    rounding, after the wavefront has crossed the focus.  (Moving bubbles
    break this symmetry in the reference itself: the velocity gather reads
    the nodes half a cell to the bubble's right, Q7, pinned above.)"""
    n = 64
    cfg = cases.acoustic_case(n, amp_pa=0.5e6)
    o = cases.make(ol.OracleSim, cfg)
    o.step(200)
    p, vx, vy = o.field("p0"), o.field("vx"), o.field("vy")
    scale = np.abs(p).max()
    assert scale > 1e5
    np.testing.assert_allclose(p[:, 1:n + 1], p[:, n:0:-1], atol=1e-12 * scale)
    # vx nodes in = 0..n at columns in + 1; node in mirrors to node n - in
    np.testing.assert_allclose(vx[:, 1:n + 2], -vx[:, n + 1:0:-1], atol=1e-12 * np.abs(vx).max())
    np.testing.assert_allclose(vy[:, 1:n + 1], vy[:, n:0:-1], atol=1e-12 * np.abs(vy).max())


# ---- temperature: heat balance and the absorption term ----------------------
def _cp_water(T):
    f = lambda v: float(np.float32(v))  # the reference's single-precision literals
    return (917.5 - f(10.1016) * T + f(0.0454134) * T ** 2 - 9.07517e-5 * T ** 3
            + f(6.80700e-8) * T ** 4) / f(18.0e-3)


def test_heat_balance_of_the_first_thermal_step():
    """From T = 0, one temperature step deposits in each cell
    dt (bubble heat + absorption) / (rho_m c_p).  Cells without bubble heat
    get the absorbed acoustic power alone, 2 a (dp/dt)^2 / (rho c w^2) per
    unit volume with a = 0.22 ln(10)/20 (0.22 dB in nepers; the reference
    writes ln(10)/20 as the literal 0.115129255, nine digits); summed over the
    grid, rho_m c_p dT dx dy / dt minus the absorbed power equals the
    bubbles' total heat sum(n_B Q_B) (the smooth delta spreads each Q_B with
    weights summing to 1)."""
    n = 64
    cfg, bub = cases.bubbly_case(n, nb=120, amp_pa=1.2e6)
    o = cases.make(ol.OracleSim, cfg, bub)
    o.step(70)
    cfg2 = abi.Config.from_buffer_copy(cfg)
    cfg2.thermal = 1
    st = o.export_state()
    h = cases.make(ol.OracleSim, cfg2, bub)
    h.import_state(st)
    pn = h.field("px") + h.field("py")
    rho_m = h.field("rho_m")
    h.step(1)
    p = h.field("px") + h.field("py")
    dT = h.field("T")
    fg = h.field("fg")
    QB, nB = h.bubbles("Q_B"), h.bubbles("n_B")
    assert np.abs(QB).max() > 0
    dx = cfg.grid_lx / n
    dt = h.info()["dt"]
    w = 2 * math.pi * cfg.tr_freq
    cp = _cp_water(293.15)
    inner = (slice(1, n + 1), slice(1, n + 1))
    a_exact = 0.22 * math.log(10.0) / 20.0
    a_lit = 0.22 * round(math.log(10.0) / 20.0, 9)
    absorbed = 2 / (RHO * C * w * w) * ((p - pn) / dt) ** 2  # per unit a
    far = np.zeros_like(dT, bool)
    far[inner] = fg[inner] == 0.0  # no bubble stencil reaches these cells
    assert far.sum() > 3000
    np.testing.assert_allclose(dT[far], dt * a_exact * absorbed[far] / (rho_m[far] * cp),
                               rtol=4e-9)
    deposited = np.sum(rho_m[inner] * cp * dT[inner]) * dx * dx / dt
    bubbles = deposited - a_lit * np.sum(absorbed[inner]) * dx * dx
    print(f"deposited {deposited:.6e} W, by the bubbles {bubbles:.6e} W, "
          f"sum n_B Q_B {np.sum(nB * QB):.6e} W")
    assert bubbles == pytest.approx(np.sum(nB * QB), rel=1e-9)


def test_bubble_heat_is_zero_at_equilibrium():
    """A bubble at rest at its equilibrium pressure exchanges no heat and
    dissipates nothing: Q_B vanishes to rounding, against the Q_B of the same
    bubble driven by a pressure ramp."""
    dt = 1e-4 * 0.5 / 1500.0
    for R0 in (1e-6, 3e-6, 8e-6):
        bp = ol.bubble_params(R0, dt_mix=dt)
        _, rest = ol.rp_integrate(bp, dt, [R0, R0, 0.0, bp.PG0, bp.dt0, 0.0, 0.0, 0.0, 0.0])
        _, driven = ol.rp_integrate(bp, dt, [R0, R0, 0.0, bp.PG0, bp.dt0, 0.0, 0.0, 0.0,
                                             -2e5])
        assert abs(driven[7]) > 0
        assert abs(rest[7]) < 1e-9 * abs(driven[7]), (rest[7], driven[7])


# ---- the predictability horizon of a cavitating cloud ------------------------
def test_c1_cloud_is_chaotic():
    """The C1 workload (256^2, 1024 bubbles, 0.5 MPa) run twice by the oracle,
    the second time with one bubble's R0 moved by one ulp: the runs agree to
    rounding until the wavefront reaches the bubbles (~step 300), to ~1e-9 at
    step 500, and differ by more than 1e-4 at step 750 -- the bubbles'
    collapses amplify any rounding difference by about 1e6 per 250 steps, so
    two correct implementations that round differently agree pointwise only
    up to about step 500 of this run."""
    out = []
    for perturb in (False, True):
        cfg, (x, y, R0, _, fg_in), src = cases.c1_case()
        R0 = R0.copy()
        if perturb:
            R0[0] = np.nextafter(R0[0], 1.0)
        o = ol.OracleSim(cfg, threads=8)
        o.set_bubbles(x, y, R0, None, fg_in)
        o.set_source(*src)
        o.prepare()
        rec = {}
        for s in (250, 500, 750):
            o.step(s - (max(rec) if rec else 0))
            rec[s] = (o.field("p0"), o.bubbles("R_t"))
        out.append(rec)
        o.close()
    d = {s: (cases.rel_l2(out[1][s][0], out[0][s][0]), cases.rel_l2(out[1][s][1], out[0][s][1]))
         for s in out[0]}
    print({s: f"p0 {a:.1e} R_t {b:.1e}" for s, (a, b) in d.items()})
    assert max(d[250]) < 1e-12
    assert max(d[500]) < 1e-6
    assert min(d[750]) > 1e-4
