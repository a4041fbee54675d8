"""ctypes wrapper of the FP64 host oracle (oracle/_build/liboracle.so).

The oracle is the checker: tests compare libbubbles_rp against it on the
same inputs.  OracleSim exposes the same Python surface as
``Simulation`` of the product package, so a parity test drives both with the
same calls.
"""
import ctypes as C
import importlib
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
brp = importlib.import_module("cuda-rp_amd")
abi = brp.abi

ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                       stdout=subprocess.DEVNULL)
    lib = C.CDLL(ORACLE_SO)
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "oracle_create": (vp, [C.POINTER(abi.Config), C.c_char_p, C.c_int]),
        "oracle_destroy": (None, [vp]),
        "oracle_last_error": (C.c_char_p, [vp]),
        "oracle_set_threads": (None, [vp, i32]),
        "oracle_get_info": (C.c_int, [vp, C.POINTER(abi.Info)]),
        "oracle_set_bubbles": (C.c_int, [vp, i32, dp, dp, dp, dp, C.c_double]),
        "oracle_set_source": (C.c_int, [vp, i32, dp, dp]),
        "oracle_prepare": (C.c_int, [vp]),
        "oracle_step": (C.c_int, [vp, i32, C.POINTER(abi.StepStats)]),
        "oracle_field_shape": (C.c_int, [vp, i32, C.POINTER(i32), C.POINTER(i32),
                                         C.POINTER(i32), C.POINTER(i32)]),
        "oracle_get_field": (C.c_int, [vp, i32, dp, i64]),
        "oracle_set_field": (C.c_int, [vp, i32, dp, i64]),
        "oracle_get_bubbles": (C.c_int, [vp, i32, dp]),
        "oracle_array_index": (C.c_int, [vp, C.POINTER(i32)]),
        "oracle_sigma": (C.c_int, [vp, i32, dp, C.POINTER(i32)]),
        "oracle_scalars": (C.c_int, [vp, dp]),
        "oracle_smooth_delta": (C.c_double, [vp, i32, i32, C.c_double]),
        "oracle_coth": (None, [C.c_double, C.c_double, dp]),
        "oracle_upsilon": (None, [C.c_double, C.c_double, C.c_double, dp]),
        "oracle_solve_lp": (None, [C.c_double, C.c_double, C.c_double, dp]),
        "oracle_solve_omega_n": (C.c_double, [C.c_double, C.c_double, vp, dp]),
        "oracle_bubble_params_init": (None, [vp] + [C.c_double] * 9),
        "oracle_rp_integrate": (C.c_int, [vp, C.c_double, dp, dp, i32]),
        "oracle_rp_flops_last": (C.c_int64, []),
        "oracle_gather_p0": (C.c_int, [vp, dp]),
        "oracle_scatter_values": (C.c_int, [vp, dp, dp]),
        "oracle_rp_radius": (None, [vp, C.c_double, C.c_double, dp]),
        "oracle_schedule_mode": (C.c_int, [vp, i32]),
        "oracle_schedule_copy": (C.c_int, [vp, vp]),
        "oracle_set_focus": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "oracle_set_bubble_array": (C.c_int, [vp, i32, dp]),
        "oracle_get_clock": (C.c_int, [vp, C.POINTER(i64), dp, dp]),
        "oracle_set_clock": (C.c_int, [vp, i64, C.c_double, C.c_double]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class BubbleParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("R0", "R03", "PL0", "PG0", "T0", "gam", "sig", "mu", "rho", "K0",
                 "coeff_alpha", "dt0")]


def bubble_params(R0, PL0=101325.0, T0=293.15, gam=1.4, sig=0.0725, mu=1e-3, rho=1000.0,
                  K0=26.2e-3, dt_mix=1e-4 * 0.5 / 1500.0):
    lib = load()
    bp = BubbleParams()
    lib.oracle_bubble_params_init(C.byref(bp), R0, PL0, T0, gam, sig, mu, rho, K0, dt_mix)
    return bp


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


class OracleSim(brp._Backend):
    prefix = "oracle_"

    def __init__(self, cfg, threads=1):
        self.lib = load()
        err = C.create_string_buffer(512)
        h = self.lib.oracle_create(C.byref(cfg), err, 512)
        if not h:
            raise brp.BrpError(abi.ERR_INVALID, err.value.decode())
        self.h = C.c_void_p(h)
        self.cfg = cfg
        self.lib.oracle_set_threads(self.h, threads)

    def last_error(self):
        return self.lib.oracle_last_error(self.h).decode()

    def schedule_mode(self, mode):
        """0 reference behaviour, 1 record, 2 replay the recorded schedule."""
        self._check(self.lib.oracle_schedule_mode(self.h, mode))

    def schedule_from(self, other):
        self._check(self.lib.oracle_schedule_copy(self.h, other.h))

    def array_index(self):
        out = (C.c_int32 * 28)()
        self._check(self.lib.oracle_array_index(self.h, out))
        return list(out)

    def sigma(self, which):
        n = C.c_int32()
        self._check(self.lib.oracle_sigma(self.h, which, None, C.byref(n)))
        out = np.zeros(n.value)
        self._check(self.lib.oracle_sigma(self.h, which, _d(out), C.byref(n)))
        return out

    def scalars(self):
        out = np.zeros(8)
        self._check(self.lib.oracle_scalars(self.h, _d(out)))
        return out

    def smooth_delta(self, axis, pos, b):
        return self.lib.oracle_smooth_delta(self.h, axis, pos, b)

    def gather_p0(self):
        out = np.zeros(self.info()["n_bubbles"])
        self._check(self.lib.oracle_gather_p0(self.h, _d(out)))
        return out

    def scatter_values(self, q):
        nx, ny, _, _ = self.shape("fg")
        out = np.zeros((ny, nx))
        q = np.ascontiguousarray(q, dtype=np.float64)
        self._check(self.lib.oracle_scatter_values(self.h, _d(q), _d(out)))
        return out

    def close(self):
        if self.h:
            self.lib.oracle_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def coth(z):
    out = np.zeros(2)
    load().oracle_coth(z.real, z.imag, _d(out))
    return complex(out[0], out[1])


def upsilon(z, gam):
    out = np.zeros(2)
    load().oracle_upsilon(z.real, z.imag, gam, _d(out))
    return complex(out[0], out[1])


def solve_lp(z, R):
    out = np.zeros(2)
    load().oracle_solve_lp(z.real, z.imag, R, _d(out))
    return complex(out[0], out[1])


def omega_n(PG, R, bp):
    al = np.zeros(2)
    w = load().oracle_solve_omega_n(PG, R, C.byref(bp), _d(al))
    return w, complex(al[0], al[1])


def rp_integrate(bp, dt_mix, state, max_substeps=0):
    """state = [R_pn, R_nn, d1_R_n, PG_n, dt_n, re_n, PL_m, PL_n, PL_p];
    returns (substeps, [R_t, R_p, R_n, d1_R_p, PG_p, dt, re, Q_B])."""
    st = np.ascontiguousarray(state, dtype=np.float64)
    out = np.zeros(8)
    n = load().oracle_rp_integrate(C.byref(bp), dt_mix, _d(st), _d(out), max_substeps)
    return n, out


def rp_flops_last():
    return load().oracle_rp_flops_last()


def rp_radius(bp, PG, PL, io):
    """One solveRayleighPlesset call; io = [Rt, Rp, Rn, d1R, dt, remain]."""
    a = np.ascontiguousarray(io, dtype=np.float64)
    load().oracle_rp_radius(C.byref(bp), PG, PL, _d(a))
    return a
