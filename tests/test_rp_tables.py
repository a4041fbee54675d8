"""The thermal-impedance tables of the shortened-chain RP sub-step
(cuda-rp_amd/csrc/rp_vfit.h, built on the host for the simulation's gam and
read by rp_fast.h) against the oracle's Upsilon and solveLp
(bubble_CUDA_kernel.cuh:296-323), against mpmath (coth at 40 digits) and
against closed forms.  CPU only: tests/rp_vfit_dump.cpp prints what
rp_vfit.h builds, and the tables are evaluated here the way the kernel
evaluates them (interval k = floor(8 s), Horner in t = 16 s - (2k + 1) with
fused multiply-adds).

  U(s) = (alpha coth alpha - 1) / alpha^2,  Upsilon = 3 gam / (1 + 3 (gam - 1) U)
  V(s) = conj(1 / (1 + 3 (gam - 1) U(s))) = conj(Upsilon) / (3 gam)
  Y(s) = 1 / conj(W(s)),  W = 1 / (alpha coth alpha - 1) - 3 / alpha^2,
  Lp = R W,  Lp / |Lp|^2 = Y / R (what solvePG uses),  alpha = s (1 + i).
"""
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "cuda-rp_amd", "csrc")
PER_UNIT, NINT, DEG = 8, 160, 9


@pytest.fixture(scope="module")
def dump(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("vfit") / "rp_vfit_dump")
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", CSRC, "-o", exe,
                    os.path.join(HERE, "rp_vfit_dump.cpp")], check=True, timeout=120)
    return exe


def build_table(exe, k3):
    out = subprocess.run([exe, "table", repr(k3)], capture_output=True, text=True, check=True,
                         timeout=60).stdout.split()
    return np.array([float.fromhex(x) for x in out]).reshape(2, NINT, DEG + 1, 2)


def _fma(a, b, c):
    """a b + c rounded once (exact with Python's rationals)."""
    from fractions import Fraction
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def table_eval(tab, f, s, exact_fma=False):
    k = max(0, min(int(s * PER_UNIT), NINT - 1))
    t = s * (2.0 * PER_UNIT) - (2.0 * k + 1.0)
    c = tab[f, k]
    pr, pi = c[DEG]
    for j in range(DEG - 1, -1, -1):
        if exact_fma:
            pr, pi = _fma(pr, t, c[j, 0]), _fma(pi, t, c[j, 1])
        else:
            pr, pi = pr * t + c[j, 0], pi * t + c[j, 1]
    return complex(pr, pi)


def u_oracle(s):
    a = s * (1 + 1j)
    return (a * ol.coth(a) - 1) / (a * a)


def w_oracle(s):
    return ol.solve_lp(s * (1 + 1j), 1.0)


def y_oracle(s):
    return 1 / np.conj(w_oracle(s))


GAMS = (1.4, 1.67, 1.1)


@pytest.fixture(scope="module")
def tables(dump):
    return {g: build_table(dump, 3.0 * (g - 1.0)) for g in GAMS}


def test_continued_fraction_matches_mpmath(dump):
    """U and W from the continued fraction, in extended precision, against
    coth at 40 digits: relative error below 1e-18 on 0 < s <= 20 (the
    table's nodes need them well below double precision)."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    ss = [0.0078125, 0.0625, 0.0703125, 0.25, 0.5, 1.0, 1.5, 2.375, 3.0, 5.5, 8.0, 11.25,
          14.0, 17.5, 19.875, 20.0]
    out = subprocess.run([dump, "cf"] + [repr(s) for s in ss], capture_output=True, text=True,
                         check=True, timeout=60).stdout.strip().split("\n")

    def hexl(x):  # glibc's %La: 0xh.hhhp+e, exact
        m = re.match(r"(-?)0x([0-9a-f]+)\.?([0-9a-f]*)p([+-]\d+)", x)
        sg, a, b, e = m.groups()
        v = mp.mpf(int(a + b, 16)) * mp.mpf(2) ** (int(e) - 4 * len(b))
        return -v if sg else v

    for s, line in zip(ss, out):
        p = line.split()
        u, w = mp.mpc(hexl(p[0]), hexl(p[1])), mp.mpc(hexl(p[2]), hexl(p[3]))
        a = mp.mpf(s) * mp.mpc(1, 1)
        ue = (a * mp.coth(a) - 1) / (a * a)
        we = 1 / (a * mp.coth(a) - 1) - 3 / (a * a)
        assert abs(u - ue) / abs(ue) < 1e-18, s
        assert abs(w - we) / abs(we) < 1e-18, s


def test_values_at_zero(dump):
    """U(0) = 1/3 and W(0) = 1/5, the constant terms of the reference's
    series: the continued fraction gives them exactly, and the tables give
    V(0) = 1 / (1 + k3 / 3) and Y(0) = 5."""
    p = subprocess.run([dump, "cf", "0"], capture_output=True, text=True, check=True,
                       timeout=60).stdout.split()
    assert float.fromhex(p[0]) == pytest.approx(1 / 3, rel=1e-16) and float.fromhex(p[1]) == 0.0
    assert float.fromhex(p[2]) == pytest.approx(1 / 5, rel=1e-16) and float.fromhex(p[3]) == 0.0
    tab = build_table(dump, 1.2)
    assert table_eval(tab, 0, 0.0) == pytest.approx(1 / (1 + 1.2 / 3), rel=1e-15)
    assert table_eval(tab, 1, 0.0) == pytest.approx(5.0, rel=1e-15)


def test_tables_against_mpmath_at_every_interval(dump, tables):
    """The double-precision Horner evaluation (fused multiply-adds) of the
    built tables against mpmath at 24 points per interval: error below
    2.5e-16 of |V| and |Y| for several gam."""
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    for g in GAMS:
        k3 = 3.0 * (g - 1.0)
        tab = tables[g]
        worst = [0.0, 0.0]
        for k in range(0, NINT, 3 if g != 1.4 else 1):
            for q in range(0, 24, 1 if g == 1.4 else 4):
                s = float(k / PER_UNIT + (1 / PER_UNIT) * q / 23)
                if s == 0.0:
                    continue
                a = mp.mpf(s) * mp.mpc(1, 1)
                ac = a * mp.coth(a)
                v = mp.conj(1 / (1 + k3 * (ac - 1) / (a * a)))
                y = 1 / mp.conj(1 / (ac - 1) - 3 / (a * a))
                for f, ex in ((0, v), (1, y)):
                    got = table_eval(tab, f, s, exact_fma=True)
                    worst[f] = max(worst[f], float(abs(mp.mpc(got.real, got.imag) - ex) / abs(ex)))
        assert max(worst) < 2.5e-16, (g, worst)


def test_tables_match_oracle_upsilon_and_lp(tables):
    """Where the reference's own expressions lose little to cancellation
    (|alpha| >= 0.7: relative error below about 45 eps / |alpha|^4)."""
    rng = np.random.default_rng(3)
    ss = np.concatenate([rng.uniform(0.5, 20.0, 3000), np.arange(0.5, 20.0, 0.25)])
    for g in GAMS:
        tab = tables[g]
        for s in ss[:: (1 if g == 1.4 else 10)]:
            tol = 1e-12 if s < 1.0 else 2e-14
            ups = 3 * g * np.conj(table_eval(tab, 0, s))
            assert abs(ups - ol.upsilon(s * (1 + 1j), g)) <= tol * abs(ups), (g, s)
            assert abs(table_eval(tab, 1, s) - y_oracle(s)) <= tol * abs(y_oracle(s)), s


def test_tables_agree_with_series_inside_the_taylor_branch(tables):
    """The kernel uses the reference's Taylor series for |alpha| <= 0.1 (the
    oracle's branch there too); the tables agree with the truncated series
    to its truncation error (the next terms, 2.14e-5 alpha^8 of U and
    6.0e-7 alpha^8 of W, 3.0e-6 alpha^8 of Y relative to |Y| = 5), so nothing
    jumps where the kernel switches."""
    g = 1.4
    k3 = 3.0 * (g - 1.0)
    tab = tables[g]
    for s in np.linspace(0.0, 0.1 / math.sqrt(2), 50):
        a = s * (1 + 1j)
        u_series = 1 / 3 + a * a * (-1 / 45 + a * a * (2 / 945 - a * a / 4725))
        v_series = np.conj(1 / (1 + k3 * u_series))
        a8 = abs(a) ** 8
        assert abs(table_eval(tab, 0, s) - v_series) < k3 * 2.2e-5 * a8 + 1e-15, s
        assert abs(table_eval(tab, 1, s) - y_oracle(s)) < 5 * (3.1e-6 * a8 + 1e-15), s


def test_closed_forms_beyond_twenty(tables):
    """For s > 20 the kernel uses U = (alpha - 1) / alpha^2 and
    Y = 1 / conj(W), W = 1 / (alpha - 1) - 3 / alpha^2 (coth alpha = 1 to below
    an ulp), Y as conj(N) / conj(D), N = alpha^2 (alpha - 1),
    D = alpha^2 - 3 alpha + 3, both over s^2: continuous with the tables at
    s = 20 and equal to the oracle beyond."""
    def u_far(s):
        return complex(0.5 / s, 0.5 / s * (1 / s - 1))

    def y_far(s):
        i_s = 1.0 / s
        a, b = -2.0 * s, -2.0 * (s - 1.0)
        c, d = 3.0 * i_s * (i_s - 1.0), 3.0 * i_s - 2.0
        den = c * c + d * d
        return complex((a * c + b * d) / den, (b * c - a * d) / den)

    for g in GAMS:
        k3 = 3.0 * (g - 1.0)
        v_far = np.conj(1 / (1 + k3 * u_far(20.0)))
        s = 20.0 - 1e-12
        assert abs(table_eval(tables[g], 0, s) - v_far) < 1e-13 * abs(v_far)
        assert abs(table_eval(tables[g], 1, s) - y_far(20.0)) < 1e-13 * abs(y_far(20.0))
    for s in (20.0, 20.5, 37.0, 300.0, 1e4):
        assert abs(u_far(s) - u_oracle(s)) < 1e-14 * abs(u_far(s))
        assert abs(y_far(s) - y_oracle(s)) < 1e-13 * abs(y_far(s))


def test_a_plausible_slip_fails():
    """The checks against the oracle catch a wrong k3 (gam - 1 instead of
    3 (gam - 1)) or a missing conjugate in V: both move Upsilon = 3 gam conj(V)
    far beyond the tolerances."""
    s, g = 2.0, 1.4
    ups = ol.upsilon(s * (1 + 1j), g)
    u = u_oracle(s)
    v_right = np.conj(1 / (1 + 3 * (g - 1) * u))
    v_wrong_k3 = np.conj(1 / (1 + (g - 1) * u))
    v_no_conj = 1 / (1 + 3 * (g - 1) * u)
    assert abs(3 * g * np.conj(v_right) - ups) < 1e-13 * abs(ups)
    assert abs(3 * g * np.conj(v_wrong_k3) - ups) > 1e-3 * abs(ups)
    assert abs(3 * g * np.conj(v_no_conj) - ups) > 1e-3 * abs(ups)
