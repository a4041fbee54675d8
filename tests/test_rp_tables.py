"""The thermal-impedance tables of the shortened-chain RP sub-step
(cuda-rp_amd/csrc/rp_vtable.h, read by rp_fast.h) against the oracle's
Upsilon and solveLp (bubble_CUDA_kernel.cuh:296-323) and against closed
forms.  CPU only: the header is parsed and evaluated in Python the way the
kernel evaluates it (interval k = floor(8 s), Horner in t = 16 s - (2k + 1)).

  U(s) = (alpha coth alpha - 1) / alpha^2,  Upsilon = 3 gam / (1 + 3 (gam - 1) U)
  Y(s) = 1 / conj(W(s)),  W = 1 / (alpha coth alpha - 1) - 3 / alpha^2,
  Lp = R W,  Lp / |Lp|^2 = Y / R (what solvePG uses),  alpha = s (1 + i).
"""
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle_lib as ol

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                    "cuda-rp_amd", "csrc")
HEADER = os.path.join(CSRC, "rp_vtable.h")


def _table():
    txt = open(HEADER).read()
    global PER_UNIT
    PER_UNIT = int(re.search(r"kVPerUnit = (\d+);", txt).group(1))
    nint = int(re.search(r"kVIntervals = (\d+);", txt).group(1))
    deg = int(re.search(r"kVDegree = (\d+);", txt).group(1))
    body = txt[txt.index("kVTable[2]"):]
    body = body[body.index("{"):]
    nums = [float(x) for x in re.findall(r"[-+]?\d+\.?\d*(?:e[-+]?\d+)?", body)]
    return np.array(nums).reshape(2, nint, deg + 1, 2)


TAB = _table()
NINT, DEG = TAB.shape[1], TAB.shape[2] - 1


def table_eval(f, s):
    k = max(0, min(int(s * PER_UNIT), NINT - 1))
    t = s * (2.0 * PER_UNIT) - (2.0 * k + 1.0)
    c = TAB[f, k]
    pr, pi = c[DEG]
    for j in range(DEG - 1, -1, -1):
        pr, pi = pr * t + c[j, 0], pi * t + c[j, 1]
    return complex(pr, pi)


def u_oracle(s):
    a = s * (1 + 1j)
    return (a * ol.coth(a) - 1) / (a * a)


def w_oracle(s):
    return ol.solve_lp(s * (1 + 1j), 1.0)


def y_oracle(s):
    return 1 / np.conj(w_oracle(s))


def test_table_shape_covers_zero_to_twenty():
    assert TAB.shape == (2, 160, 10, 2)
    assert NINT / PER_UNIT == 20.0


def test_values_at_zero():
    """U(0) = 1/3 and W(0) = 1/5 (Y(0) = 5), the constant terms of the
    reference's series."""
    assert table_eval(0, 0.0) == pytest.approx(1 / 3, rel=1e-15)
    assert table_eval(1, 0.0) == pytest.approx(5.0, rel=1e-15)


def test_tables_match_oracle_upsilon_and_lp():
    """Where the reference's own expressions lose little to cancellation
    (|alpha| >= 0.7: relative error below about 45 eps / |alpha|^4)."""
    rng = np.random.default_rng(3)
    ss = np.concatenate([rng.uniform(0.5, 20.0, 3000), np.arange(0.5, 20.0, 0.25)])
    for s in ss:
        tol = 1e-12 if s < 1.0 else 2e-14
        assert abs(table_eval(0, s) - u_oracle(s)) <= tol * abs(u_oracle(s)), s
        assert abs(table_eval(1, s) - y_oracle(s)) <= tol * abs(y_oracle(s)), s
    gam = 1.4
    for s in ss[:200]:
        ups = 3 * gam / (1 + 3 * (gam - 1) * table_eval(0, s))
        assert ups == pytest.approx(ol.upsilon(s * (1 + 1j), gam), rel=1e-13)


def test_tables_agree_with_series_inside_the_taylor_branch():
    """The kernel uses the reference's Taylor series for |alpha| <= 0.1 (the
    oracle's branch there too); the tables agree with the truncated series
    to its truncation error (the next terms, 2.14e-5 alpha^8 of U and
    6.0e-7 alpha^8 of W, 3.0e-6 alpha^8 of Y relative to |Y| = 5), so nothing
    jumps where the kernel switches."""
    for s in np.linspace(0.0, 0.1 / math.sqrt(2), 50):
        a = s * (1 + 1j)
        u_series = 1 / 3 + a * a * (-1 / 45 + a * a * (2 / 945 - a * a / 4725))
        a8 = abs(a) ** 8
        assert abs(table_eval(0, s) - u_series) < 2.2e-5 * a8 + 1e-15, s
        assert abs(table_eval(1, s) - y_oracle(s)) < 5 * (3.1e-6 * a8 + 1e-15), s


def test_closed_forms_beyond_twenty():
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

    s = 20.0 - 1e-12
    assert abs(table_eval(0, s) - u_far(20.0)) < 1e-15 * abs(u_far(20.0)) + 1e-13 * abs(u_far(20.0))
    assert abs(table_eval(1, s) - y_far(20.0)) < 1e-13 * abs(y_far(20.0))
    for s in (20.0, 20.5, 37.0, 300.0, 1e4):
        assert abs(u_far(s) - u_oracle(s)) < 1e-14 * abs(u_far(s))
        assert abs(y_far(s) - y_oracle(s)) < 1e-13 * abs(y_far(s))


def test_generator_reproduces_header():
    pytest.importorskip("mpmath")
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_rp_vtable.py")],
                         capture_output=True, text=True, check=True, timeout=120)
    assert out.stdout == open(HEADER).read()
    errs = [float(x) for x in re.findall(r"[UY] ([0-9.e+-]+)", out.stderr)]
    assert len(errs) == 2 and max(errs) < 2e-16
