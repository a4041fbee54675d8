"""Python binding of libbubbles_rp (the MI355X coupled FDTD + Rayleigh-Plesset
step) over its C ABI, include/bubbles_rp.h.

Import with ``importlib.import_module("cuda-rp_amd")`` (the directory name is
not a Python identifier).  The binding loads ``lib/libbubbles_rp.so`` next to
this file and raises ``RuntimeError`` if it is missing: there is no CPU
fallback.  Build it with ``make -C cuda-rp_amd`` or ``__graft_entry__.build()``.

    sim = Simulation(config)          # brp_create
    sim.set_bubbles(x, y, R0)         # optional, extension E3
    stats = sim.step(100)             # brp_step -> list of per-step dicts
    p0 = sim.field("p0")              # brp_get_field, reference logical layout
"""
import ctypes as C
import os

import numpy as np

from . import abi
from .abi import *  # noqa: F401,F403  (re-export the ABI mirror)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbubbles_rp.so")
_lib = None


def library_path():
    return LIB_PATH


def load_library():
    """Loads libbubbles_rp.so (once) and declares its signatures."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libbubbles_rp.so not found at {LIB_PATH}; build it with "
            "`make -C cuda-rp_amd` (there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    vp, i32, i64, dp = C.c_void_p, C.c_int32, C.c_int64, C.POINTER(C.c_double)
    sig = {
        "brp_config_default": (None, [C.POINTER(abi.Config)]),
        "brp_create": (C.c_int, [C.POINTER(abi.Config), C.POINTER(vp)]),
        "brp_destroy": (None, [vp]),
        "brp_last_error": (C.c_char_p, [vp]),
        "brp_get_info": (C.c_int, [vp, C.POINTER(abi.Info)]),
        "brp_set_bubbles": (C.c_int, [vp, i32, dp, dp, dp, dp, C.c_double]),
        "brp_set_source": (C.c_int, [vp, i32, dp, dp]),
        "brp_prepare": (C.c_int, [vp]),
        "brp_step": (C.c_int, [vp, i32, C.POINTER(abi.StepStats)]),
        "brp_field_shape": (C.c_int, [vp, i32, C.POINTER(i32), C.POINTER(i32),
                                      C.POINTER(i32), C.POINTER(i32)]),
        "brp_get_field": (C.c_int, [vp, i32, dp, i64]),
        "brp_set_field": (C.c_int, [vp, i32, dp, i64]),
        "brp_get_bubbles": (C.c_int, [vp, i32, dp]),
        "brp_save_state": (C.c_int, [vp, i32]),
        "brp_restore_state": (C.c_int, [vp, i32]),
        "brp_get_phase_times": (C.c_int, [vp, dp]),
        "brp_set_focus": (C.c_int, [vp, C.c_double, C.c_double, C.c_double]),
        "brp_set_bubble_array": (C.c_int, [vp, i32, dp]),
        "brp_scatter_variant": (C.c_int, [vp, i32, dp, i64]),
        "brp_schedule_mode": (C.c_int, [vp, i32]),
        "brp_schedule_copy": (C.c_int, [vp, vp]),
        "brp_get_bubbles_noalloc": (C.c_int, [vp, i32, dp]),
        "brp_get_clock": (C.c_int, [vp, C.POINTER(i64), dp, dp]),
        "brp_set_clock": (C.c_int, [vp, i64, C.c_double, C.c_double]),
        "brp_adjoint": (C.c_int, [vp, C.POINTER(abi.Objective), dp, dp, dp, i64,
                                  C.POINTER(abi.AdjointStats)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def stats_to_dicts(arr):
    return [{k: getattr(s, k) for k, _ in abi.StepStats._fields_} for s in arr]


class BrpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"status {code}: {msg}")
        self.code = code


class _Backend:
    """Shared Python surface of a C-ABI simulation handle; subclasses bind a
    library and its symbol prefix."""
    prefix = "brp_"

    def _f(self, name):
        return getattr(self.lib, self.prefix + name)

    def _check(self, rc):
        if rc != 0:
            raise BrpError(rc, self.last_error())
        return rc

    def info(self):
        inf = abi.Info()
        self._check(self._f("get_info")(self.h, C.byref(inf)))
        return {k: getattr(inf, k) for k, _ in abi.Info._fields_}

    def set_bubbles(self, x, y, R0, n_B=None, fg_in=0.0):
        x, y, R0 = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, R0))
        nb = None if n_B is None else np.ascontiguousarray(n_B, dtype=np.float64)
        self._check(self._f("set_bubbles")(self.h, len(x), _dptr(x), _dptr(y), _dptr(R0),
                                           _dptr(nb), float(fg_in)))

    def set_source(self, amp, phase):
        amp = np.ascontiguousarray(amp, dtype=np.float64)
        phase = np.ascontiguousarray(phase, dtype=np.float64)
        self._check(self._f("set_source")(self.h, len(amp), _dptr(amp), _dptr(phase)))
        self._n_elem = len(amp)

    def prepare(self):
        self._check(self._f("prepare")(self.h))

    def step(self, k=1, check=True):
        arr = (abi.StepStats * max(k, 1))()
        rc = self._f("step")(self.h, k, arr)
        out = stats_to_dicts(arr[:k])
        if check:
            self._check(rc)
        return out

    def set_focus(self, fp_x, fp_y, f_dist):
        self._check(self._f("set_focus")(self.h, fp_x, fp_y, f_dist))

    def shape(self, name):
        fid = abi.FIELD_NAMES[name] if isinstance(name, str) else name
        nx, ny, i0, j0 = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
        self._check(self._f("field_shape")(self.h, fid, C.byref(nx), C.byref(ny),
                                           C.byref(i0), C.byref(j0)))
        return nx.value, ny.value, i0.value, j0.value

    def field(self, name):
        """Returns the field as an (ny, nx) array; element [j - j0, i - i0]."""
        fid = abi.FIELD_NAMES[name] if isinstance(name, str) else name
        nx, ny, _, _ = self.shape(fid)
        out = np.zeros((ny, nx), dtype=np.float64)
        self._check(self._f("get_field")(self.h, fid, _dptr(out), nx))
        return out

    def set_field(self, name, arr):
        fid = abi.FIELD_NAMES[name] if isinstance(name, str) else name
        a = np.ascontiguousarray(arr, dtype=np.float64)
        self._check(self._f("set_field")(self.h, fid, _dptr(a), a.shape[1]))

    def bubbles(self, name):
        n = self.info()["n_bubbles"]
        out = np.zeros(n, dtype=np.float64)
        if n:
            self._check(self._f("get_bubbles")(self.h, abi.BUB[name], _dptr(out)))
        return out

    def set_bubble_array(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self._f("set_bubble_array")(self.h, abi.BUB[name], _dptr(v)))

    def clock(self):
        n, t, c = C.c_int64(), C.c_double(), C.c_double()
        self._check(self._f("get_clock")(self.h, C.byref(n), C.byref(t), C.byref(c)))
        return n.value, t.value, c.value

    def set_clock(self, nstep, time, carry):
        self._check(self._f("set_clock")(self.h, nstep, time, carry))

    # The state one step starts from: the fields the step reads and the bubble
    # arrays of bubble_t it carries over (positions, RP state, pressure
    # history, outputs), plus the clock.
    STATE_FIELDS = ("p0", "px", "py", "vx", "vy", "fg", "fgn", "fgm", "rho_m")
    STATE_BUBBLES = ("x", "y", "R_t", "R_p", "R_n", "d1R", "PG", "PL_p", "PL_n", "PL_m",
                     "Q_B", "dt", "re", "vx", "vy")

    def export_state(self):
        return {"fields": {n: self.field(n) for n in self.STATE_FIELDS},
                "bubbles": {n: self.bubbles(n) for n in self.STATE_BUBBLES}
                if self.info()["n_bubbles"] else {},
                "clock": self.clock()}

    def import_state(self, st):
        """Moves this run to a state exported by another run of the same
        configuration and bubble population (f_g first: setting it sets f_gn
        and f_gm too)."""
        for n in self.STATE_FIELDS:
            if n == "rho_m" and self.prefix == "brp_":
                continue  # the library derives rho_m from f_g
            self.set_field(n, st["fields"][n])
        for n, v in st["bubbles"].items():
            self.set_bubble_array(n, v)
        self.set_clock(*st["clock"])

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


class Simulation(_Backend):
    """One libbubbles_rp handle on the current HIP device."""

    def __init__(self, cfg):
        self.lib = load_library()
        self.h = C.c_void_p()
        rc = self.lib.brp_create(C.byref(cfg), C.byref(self.h))
        if rc != 0:
            raise BrpError(rc, self.lib.brp_last_error(None).decode())
        self.cfg = cfg

    def last_error(self):
        return self.lib.brp_last_error(self.h).decode()

    def save_state(self, slot):
        self._check(self.lib.brp_save_state(self.h, slot))

    def restore_state(self, slot):
        self._check(self.lib.brp_restore_state(self.h, slot))

    def adjoint(self, box, n_steps, window, checkpoint_every=0, n_elem=None, tape=True):
        """brp_adjoint from the current state: J = mean p0^2 over the cell box
        (i0, i1, j0, j1) and the last `window` of `n_steps` steps, with dJ/d
        element amplitude, dJ/d element phase and dJ/d initial void-fraction
        map.  The handle ends at the state after n_steps."""
        if n_elem is None:
            n_elem = getattr(self, "_n_elem", 0)
        ob = abi.Objective(box[0], box[1], box[2], box[3], n_steps, window, checkpoint_every,
                           0 if tape else abi.ADJ_NO_TAPE)
        ga = np.zeros(max(n_elem, 1))
        gp = np.zeros(max(n_elem, 1))
        nx, ny = self.shape("fg")[:2]
        gf = np.zeros((ny, nx))
        st = abi.AdjointStats()
        self._check(self.lib.brp_adjoint(self.h, C.byref(ob), _dptr(ga), _dptr(gp), _dptr(gf), nx,
                                         C.byref(st)))
        return {"J": st.J, "grad_amp": ga[:n_elem], "grad_phase": gp[:n_elem], "grad_fg0": gf,
                "forward_ms": st.forward_ms, "reverse_ms": st.reverse_ms,
                "memory_bytes": st.memory_bytes, "n_checkpoints": st.n_checkpoints,
                "segment": st.segment, "schedule_mismatch": st.schedule_mismatch}

    def schedule_mode(self, mode):
        """0 reference behaviour, 1 record, 2 replay the recorded schedule."""
        self._check(self.lib.brp_schedule_mode(self.h, mode))

    def schedule_from(self, other):
        """Replay other's recorded schedule (shared, same device)."""
        self._check(self.lib.brp_schedule_copy(self.h, other.h))

    def scatter_variant(self, mode):
        """brp_scatter_variant: the current R_t scattered with accumulation
        scheme `mode` (abi.VF_*), as an f_g-shaped array."""
        nx, ny = self.shape("fg")[:2]
        out = np.zeros((ny, nx))
        self._check(self.lib.brp_scatter_variant(self.h, mode, _dptr(out), nx))
        return out

    def phase_times(self):
        a = (C.c_double * 5)()
        self._check(self.lib.brp_get_phase_times(self.h, a))
        return dict(zip(["velocity", "motion", "pressure", "bubble_rp", "other"], list(a)))

    def close(self):
        if self.h:
            self.lib.brp_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_config():
    return abi.default_config()
