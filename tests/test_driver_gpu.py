"""The `bubbles` driver end to end on the GPU: its rawdata/ files (ASCII tuples
and SPH records) hold the fields the oracle computes for the same
configuration file."""
import os
import struct
import subprocess

import numpy as np
import pytest

import cases
import oracle_lib as ol

pytestmark = pytest.mark.gpu
BIN = os.path.join(cases.ROOT, "cuda-rp_amd", "bin", "bubbles")

CONF = """[grid]
x = 48
y = 48
lx = 4.8e-3
ly = 4.8e-3
[pml]
x0 = 0
x1 = 1
y0 = 0
y1 = 1
grid_number = 8
sigma = 0.1
order = 2
[simulation]
cfl = 0.5
dt0 = 0.1
n_step_max = 60
t_step_max = 0
save_interval = 30
order = 2
smooth_delta_band = 2
[transducer]
amp = 0.6
freq = 1.0e6
focal_dist = 3.4e-3
box_size = 0.6e-3
cylindrical = 0
on_wave = 1000
off_wave = 0
focused_wave_p = 1
[bubbles]
enable = 1
fg0 = 1.0e-6
R0 = 3.0e-6
P0 = 101325
T0 = 293.15
gamma = 1.4
liquid_surface_tension = 0.0725
liquid_viscosity = 1.0e-3
liquid_density = 1000
K0 = 26.2e-3
[debug]
fg = 1
T = 1
p = 1
v = 1
bubbles = 1
[extensions]
focus_x = 2.4e-3
"""


def conf_as_config():
    c = cases.abi.default_config()
    c.update(grid_x=48, grid_y=48, grid_lx=4.8e-3, grid_ly=4.8e-3, pml_x0=0, pml_x1=1, pml_y0=0,
             pml_y1=1, pml_grid_number=8, pml_sigma=0.1, pml_order=2, sim_cfl=0.5, sim_dt0=0.1,
             sim_n_step_max=60, sim_save_interval=30, sim_order=2, sim_smooth_delta_band=2,
             tr_amp=0.6, tr_freq=1.0e6, tr_focal_dist=3.4e-3, tr_box_size=0.6e-3, tr_on_wave=1000,
             tr_focused_p=1, bub_enable=1, bub_fg0=1e-6, bub_R0=3e-6, bub_P0=101325,
             bub_T0=293.15, bub_gamma=1.4, bub_sigma=0.0725, bub_mu=1e-3, bub_rho=1000,
             bub_K0=26.2e-3, ext_focus_x=2.4e-3, ext_focus_x_set=1, thermal=1)
    return c


def read_tuples(path, X, Y):
    a = np.loadtxt(path)
    assert a.shape == ((X + 1) * (Y + 1), 3)
    return a[:, 2].reshape(Y + 1, X + 1), a[:, 0].reshape(Y + 1, X + 1), a[:, 1].reshape(Y + 1, X + 1)


def read_sph(path):
    b = open(path, "rb").read()
    vals = struct.unpack_from("<4i", b, 0)
    assert vals == (8, 1, 1, 8)
    # records: [8][type][dtype][8] [12][dims][12] [12][origin][12] [12][pitch][12]
    #          [8][step][time][8] [4n][data][4n]  (SphDataType.cpp:153-182)
    dims = struct.unpack_from("<3i", b, 20)
    step, time = struct.unpack_from("<if", b, 80)
    n = dims[0] * dims[1]
    assert struct.unpack_from("<i", b, 92)[0] == 4 * n and len(b) == 96 + 4 * n + 4
    data = np.frombuffer(b, dtype="<f4", count=n, offset=96).reshape(dims[1], dims[0])
    return dims, step, time, data


def oracle_at(steps):
    o = ol.OracleSim(conf_as_config())
    o.prepare()
    out = {}
    done = 0
    for s in steps:
        o.step(s - done)
        done = s
        out[s] = {n: o.field(n) for n in ("p0", "fg", "vx", "vy", "T")}
        out[s]["R_t"] = o.bubbles("R_t")
        out[s]["x"] = o.bubbles("x")
        out[s]["time"] = o.info()["time"]
    return out


def test_driver_ascii_rawdata(tmp_path):
    conf = tmp_path / "conf.txt"
    conf.write_text(CONF)
    out = tmp_path / "raw"
    r = subprocess.run([BIN, "-a", "-d", str(out), str(conf)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = oracle_at([30, 60])
    X = Y = 48
    for s in (30, 60):
        p0, xs, ys = read_tuples(out / f"p0_step_{s:05d}.txt", X, Y)
        np.testing.assert_allclose(xs[0, :], np.arange(X + 1) * 1e-4, rtol=1e-5)
        np.testing.assert_allclose(ys[:, 0], np.arange(Y + 1) * 1e-4, rtol=1e-5)
        want = ref[s]["p0"][:Y + 1, :X + 1]
        # 6 significant digits (default ostream formatting)
        np.testing.assert_allclose(p0, want, rtol=5e-6, atol=1e-6 * np.abs(want).max())
        fg, _, _ = read_tuples(out / f"fg_step_{s:05d}.txt", X, Y)
        np.testing.assert_allclose(fg, ref[s]["fg"][:Y + 1, :X + 1], rtol=5e-6, atol=1e-12)
        vx, _, _ = read_tuples(out / f"vx_step_{s:05d}.txt", X, Y)
        # vx is stored from i = -1: column i sits at index i + 1
        want_vx = ref[s]["vx"][:Y + 1, 1:X + 2]
        np.testing.assert_allclose(vx, want_vx, rtol=5e-6, atol=1e-6 * np.abs(want_vx).max())
        T, _, _ = read_tuples(out / f"T_step_{s:05d}.txt", X, Y)
        want_T = ref[s]["T"][:Y + 1, :X + 1]
        assert np.abs(want_T).max() > 0
        np.testing.assert_allclose(T, want_T, rtol=5e-6, atol=1e-6 * np.abs(want_T).max())
        br = np.loadtxt(out / f"bubble_radii_{s:05d}.txt")
        np.testing.assert_allclose(br[:, 2], ref[s]["R_t"], rtol=5e-6)
        np.testing.assert_allclose(br[:, 0], ref[s]["x"], rtol=5e-6)
    mm = np.loadtxt(out / "p0_minmax.txt")
    assert mm[1] == pytest.approx(max(ref[s]["p0"][:Y + 1, :X + 1].max() for s in (30, 60)), rel=1e-5)
    assert "Finished in" in (out / "runtime.txt").read_text()


def test_driver_sph_rawdata(tmp_path):
    conf = tmp_path / "conf.txt"
    conf.write_text(CONF)
    out = tmp_path / "raw"
    r = subprocess.run([BIN, "-s", "-d", str(out), str(conf)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ref = oracle_at([60])
    dims, step, time, data = read_sph(out / "p0_step_00060.sph")
    assert dims == (49, 49, 1) and step == 60
    assert time == pytest.approx(ref[60]["time"], rel=1e-6)
    want = ref[60]["p0"][:49, :49].astype(np.float32)
    np.testing.assert_allclose(data, want, rtol=1e-6, atol=1e-6 * np.abs(want).max())
    assert not list(out.glob("*.txt")) or all("minmax" in p.name or "focal" in p.name or
                                              "runtime" in p.name for p in out.glob("*.txt"))


def _butterworth():
    xv, yv = [0.0] * 5, [0.0] * 5

    def f(u):
        xv[:4] = xv[1:]
        yv[:4] = yv[1:]
        xv[4] = u / 6.522112693e+08
        yv[4] = ((xv[0] + xv[4]) + 4 * (xv[1] + xv[3]) + 6 * xv[2] + (-0.9676955438 * yv[0])
                 + (3.9025587848 * yv[1]) + (-5.9020258615 * yv[2]) + (3.9671625959 * yv[3]))
        return yv[4]
    return f


def _pid_loop(c, steps, start, carry=None, filt=None):
    """The driver's per-pass loop on the oracle: the filtered focal y of
    every step and, after `start`, focal_PID's output re-aiming the array;
    carry = (control, err_total, d_err) from a previous pass, filt the
    Butterworth filter, whose state carries over between passes (the
    reference keeps it in file-scope statics)."""
    o = ol.OracleSim(c)
    o.prepare()
    info = o.info()
    fp = (info["fp_x"], info["fp_y"])
    filt = filt or _butterworth()
    focal = [0.0]
    if carry:
        control = [carry[0]]
        err_tot, d_err = list(carry[1]), list(carry[2])
        o.set_focus(fp[0], carry[0], carry[0] / 0.5 / np.sqrt(3.0))
    else:
        control = [fp[1]]
        err_tot, d_err = [0.0, 0.0], [0.0, 0.0]
    K, Pc = 0.2 * 0.60, 1000.0
    D, I = K * 0.125 * Pc, K / (0.5 * Pc)
    for n in range(1, steps + 1):
        st = o.step(1)[0]
        focal.append(filt(st["focal_y"]))
        if n > start:
            act, prev = (0.0, focal[-1]), (0.0, focal[-2])
            err = [fp[0] - act[0], fp[1] - act[1]]
            err_tot = [err_tot[0] + err[0], err_tot[1] + err[1]]
            d_err = [prev[0] - act[0], prev[1] - act[1]]
            uy = fp[1] + (K * err[1] + D * d_err[1] + I * err_tot[1])
            uy = min(max(uy, fp[1]), c.grid_ly)
            control.append(uy)
            o.set_focus(fp[0], uy, uy / 0.5 / np.sqrt(3.0))
    return focal, control, (control[-1], err_tot, d_err), o, filt


def test_driver_focal_tracking_and_pid(tmp_path):
    """transducer.pid = 1: the driver's focal-point series (filtered) and the
    controller's outputs follow the same loop run on the oracle
    (bubble_CUDA.cu:310-348, focal_PID and filter_loop of
    bubble_CUDA_kernel.cuh:118-135,1168-1184)."""
    conf = tmp_path / "conf.txt"
    conf.write_text(CONF.replace("focal_dist = 3.4e-3\n", "focal_dist = 3.4e-3\npid = 1\n"
                                 "pid_start_step = 10\n"))
    out = tmp_path / "raw"
    r = subprocess.run([BIN, "-a", "-d", str(out), str(conf)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    fa = np.loadtxt(out / "focalpoint_actual_T.txt")
    fc = np.loadtxt(out / "focalpoint_controller_T.txt")
    # the same loop on the oracle
    c = conf_as_config()
    c.update(tr_pid=1, tr_pid_start_step=10)
    o = ol.OracleSim(c)
    o.prepare()
    info = o.info()
    fp = (info["fp_x"], info["fp_y"])
    filt = _butterworth()
    focal, control = [0.0], [fp[1]]
    err_tot, d_err = [0.0, 0.0], [0.0, 0.0]
    K, Pc = 0.2 * 0.60, 1000.0
    D, I = K * 0.125 * Pc, K / (0.5 * Pc)
    for n in range(1, 61):
        st = o.step(1)[0]
        focal.append(filt(st["focal_y"]))
        if n > 10:
            act, prev = (0.0, focal[-1]), (0.0, focal[-2])
            err = [fp[0] - act[0], fp[1] - act[1]]
            err_tot = [err_tot[0] + err[0], err_tot[1] + err[1]]
            d_err = [prev[0] - act[0], prev[1] - act[1]]
            uy = fp[1] + (K * err[1] + D * d_err[1] + I * err_tot[1])
            uy = min(max(uy, fp[1]), c.grid_ly)
            control.append(uy)
            o.set_focus(fp[0], uy, uy / 0.5 / np.sqrt(3.0))
    assert fa.shape == (61, 2) and fc.shape == (51, 2)
    np.testing.assert_array_equal(fa[:, 0], np.arange(61))
    np.testing.assert_array_equal(fc[:, 0], np.arange(51) + 10)
    np.testing.assert_allclose(fa[:, 1], focal, rtol=5e-6, atol=1e-12)
    np.testing.assert_allclose(fc[:, 1], control, rtol=5e-6)
    assert len(set(np.round(fc[:, 1], 9))) > 1  # the controller moved the focus


def test_driver_pid_iterate_passes(tmp_path):
    """transducer.pid_iterate = 2: two passes into DIR_I/ and DIR_II/
    (bubbles.cpp:221-246); the second starts from the first pass's
    controller state (its last control point and error sums,
    bubble_CUDA.cu:470-475), re-aims the array from its first step and
    continues the first pass's Butterworth filter.  Both
    passes' focal-point and controller files against the same loop run on
    the oracle."""
    conf = tmp_path / "conf.txt"
    conf.write_text(CONF.replace("focal_dist = 3.4e-3\n", "focal_dist = 3.4e-3\npid = 1\n"
                                 "pid_start_step = 10\npid_iterate = 2\n"))
    out = tmp_path / "raw"
    r = subprocess.run([BIN, "-a", "-d", str(out), str(conf)], capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stdout + r.stderr
    d1, d2 = tmp_path / "raw_I", tmp_path / "raw_II"
    assert d1.is_dir() and d2.is_dir() and not out.exists()
    c = conf_as_config()
    c.update(tr_pid=1, tr_pid_start_step=10, tr_pid_iterate=2)
    f1, c1, carry, _, filt = _pid_loop(c, 60, 10)
    f2, c2, _, o2, _ = _pid_loop(c, 60, 10, carry, filt)
    for d, focal, control in ((d1, f1, c1), (d2, f2, c2)):
        fa = np.loadtxt(d / "focalpoint_actual_T.txt")
        fc = np.loadtxt(d / "focalpoint_controller_T.txt")
        np.testing.assert_allclose(fa[:, 1], focal, rtol=5e-6, atol=1e-12)
        np.testing.assert_allclose(fc[:, 1], control, rtol=5e-6)
        assert (d / "p0_step_00060.txt").exists() and (d / "runtime.txt").exists()
    assert np.loadtxt(d2 / "focalpoint_controller_T.txt")[0, 1] == pytest.approx(c1[-1], rel=5e-6)
    p0, _, _ = read_tuples(d2 / "p0_step_00060.txt", 48, 48)
    want = o2.field("p0")[:49, :49]
    np.testing.assert_allclose(p0, want, rtol=5e-6, atol=1e-6 * np.abs(want).max())


def test_driver_cylindrical_ascii_is_mirrored(tmp_path):
    """transducer.cylindrical = 1: the ASCII files hold the axisymmetric
    field mirrored about the axis (output_styles.cpp:196): columns
    i = -X..X at x = i dx, the value of column |i|, against the oracle's
    cylindrical run."""
    conf = tmp_path / "conf.txt"
    conf.write_text(CONF.replace("cylindrical = 0", "cylindrical = 1"))
    out = tmp_path / "raw"
    r = subprocess.run([BIN, "-a", "-d", str(out), str(conf)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    X = Y = 48
    a = np.loadtxt(out / "p0_step_00060.txt")
    assert a.shape == ((2 * X + 1) * (Y + 1), 3)
    v = a[:, 2].reshape(Y + 1, 2 * X + 1)
    xs = a[:, 0].reshape(Y + 1, 2 * X + 1)
    np.testing.assert_allclose(xs[0], np.arange(-X, X + 1) * 1e-4, rtol=1e-5, atol=1e-12)
    np.testing.assert_array_equal(v[:, :X], v[:, X + 1:][:, ::-1])  # value at |i|
    c = conf_as_config()
    c.update(tr_cylindrical=1)
    o = ol.OracleSim(c)
    o.prepare()
    o.step(60)
    want = o.field("p0")[:Y + 1, :X + 1]
    assert np.abs(want).max() > 1e4
    np.testing.assert_allclose(v[:, X:], want, rtol=5e-6, atol=1e-6 * np.abs(want).max())
