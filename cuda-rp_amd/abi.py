"""ctypes mirror of include/bubbles_rp.h (structs, enums, defaults).

Pure data: importing this module loads no library, so it is usable on a
machine without a GPU.  The field order must match the C header exactly;
tests/test_abi.py checks sizes and offsets against the compiled library.
"""
import ctypes as C

ABI_VERSION = 2

OK = 0
ERR_INVALID = 1
ERR_DEVICE = 2
ERR_NONFINITE = 3
ERR_PICARD_CAP = 4
ERR_SUBSTEP_CAP = 5
ERR_NO_DEVICE = 6
ERR_STATE = 7

PREC_FP64 = 0
PREC_MIXED = 1

VF_INT64, VF_FP64_ATOMIC, VF_FP32_COMPENSATED, VF_FP32_ATOMIC = range(4)

FIELD_P0, FIELD_PX, FIELD_PY, FIELD_VX, FIELD_VY, FIELD_FG, FIELD_FGN, FIELD_FGM, \
    FIELD_RHO_M, FIELD_T = range(10)
FIELD_NAMES = {"p0": FIELD_P0, "px": FIELD_PX, "py": FIELD_PY, "vx": FIELD_VX,
               "vy": FIELD_VY, "fg": FIELD_FG, "fgn": FIELD_FGN, "fgm": FIELD_FGM,
               "rho_m": FIELD_RHO_M, "T": FIELD_T}

BUB_NAMES = ["x", "y", "R_t", "R_p", "R_n", "d1R", "PG", "PL_p", "PL_n", "PL_m", "Q_B",
             "n_B", "dt", "re", "vx", "vy", "ibm_x", "ibm_y", "ibn_x", "ibn_y", "R0", "sub"]
BUB = {n: k for k, n in enumerate(BUB_NAMES)}


class Config(C.Structure):
    """brp_config: the Boost program_options keys of the reference's
    src/bubbles.cpp:75-133 plus extensions E1-E4 and run controls."""
    _fields_ = [
        ("grid_x", C.c_int32), ("grid_y", C.c_int32),
        ("grid_lx", C.c_double), ("grid_ly", C.c_double),
        ("pml_x0", C.c_int32), ("pml_x1", C.c_int32), ("pml_y0", C.c_int32),
        ("pml_y1", C.c_int32), ("pml_grid_number", C.c_int32), ("pml_order", C.c_int32),
        ("pml_sigma", C.c_double),
        ("sim_cfl", C.c_double), ("sim_dt0", C.c_double), ("sim_t_step_max", C.c_double),
        ("sim_n_step_max", C.c_int32), ("sim_save_interval", C.c_int32),
        ("sim_order", C.c_int32), ("sim_smooth_delta_band", C.c_int32),
        ("tr_amp", C.c_double), ("tr_freq", C.c_double), ("tr_focal_dist", C.c_double),
        ("tr_init_control", C.c_double), ("tr_box_size", C.c_double),
        ("tr_pid", C.c_int32), ("tr_pid_start_step", C.c_int32),
        ("tr_pid_iterate", C.c_int32), ("tr_cylindrical", C.c_int32),
        ("tr_on_wave", C.c_int32), ("tr_off_wave", C.c_int32),
        ("tr_plane_v", C.c_int32), ("tr_plane_p", C.c_int32),
        ("tr_focused_v", C.c_int32), ("tr_focused_p", C.c_int32),
        ("bub_enable", C.c_int32), ("pad0_", C.c_int32),
        ("bub_fg0", C.c_double), ("bub_R0", C.c_double), ("bub_P0", C.c_double),
        ("bub_T0", C.c_double), ("bub_gamma", C.c_double), ("bub_sigma", C.c_double),
        ("bub_mu", C.c_double), ("bub_rho", C.c_double), ("bub_K0", C.c_double),
        ("ext_focus_x", C.c_double), ("ext_focus_x_set", C.c_int32),
        ("ext_source_in_pml", C.c_int32), ("precision", C.c_int32),
        ("max_picard", C.c_int32), ("max_substeps", C.c_int32), ("thermal", C.c_int32),
        ("rp_exact", C.c_int32), ("pad1_", C.c_int32),
    ]

    def update(self, **kw):
        for k, v in kw.items():
            if not hasattr(self, k):
                raise KeyError(k)
            setattr(self, k, v)
        return self

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_ if not f[0].startswith("pad")}


class StepStats(C.Structure):
    _fields_ = [
        ("nstep", C.c_int64), ("time", C.c_double), ("resimax0", C.c_double),
        ("resimax", C.c_double), ("n_pic", C.c_int32), ("rp_max_substeps", C.c_int32),
        ("rp_mean_substeps", C.c_double), ("status", C.c_int32), ("rp_lean_launches", C.c_int32),
        ("focal_x", C.c_double), ("focal_y", C.c_double), ("rp_reused", C.c_int64),
    ]


class Info(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("n_bubbles", C.c_int32),
        ("precision", C.c_int32), ("dx", C.c_double), ("dy", C.c_double),
        ("rdx", C.c_double), ("rdy", C.c_double), ("dt", C.c_double),
        ("dt0_rp", C.c_double), ("fp_x", C.c_double), ("fp_y", C.c_double),
        ("omega", C.c_double), ("time", C.c_double), ("nstep", C.c_int64),
        ("device_bytes", C.c_int64),
    ]


class Objective(C.Structure):
    """brp_objective: mean p0^2 over a cell box and the last `window` steps."""
    _fields_ = [(n, C.c_int32) for n in ("box_i0", "box_i1", "box_j0", "box_j1", "n_steps",
                                          "window", "checkpoint_every", "flags")]


ADJ_NO_TAPE = 1


class AdjointStats(C.Structure):
    _fields_ = [("J", C.c_double), ("forward_ms", C.c_double), ("reverse_ms", C.c_double),
                ("memory_bytes", C.c_int64), ("n_checkpoints", C.c_int32),
                ("segment", C.c_int32), ("schedule_mismatch", C.c_int64)]


def default_config():
    """Python twin of brp_config_default(): the reference's declared defaults
    (src/bubbles.cpp:97-132; calloc'd structs are zero) plus run controls."""
    c = Config()
    c.sim_order = 2
    c.sim_smooth_delta_band = 2
    c.pml_order = 2
    c.pml_sigma = 0.1
    c.sim_save_interval = 100
    c.tr_pid_start_step = 1
    c.tr_pid_iterate = 1
    c.precision = PREC_FP64
    return c


def water_config(n, *, lx=None, npml=16, cfl=0.5, amp_pa=0.5e6, freq=1.0e6,
                 focal_dist=15e-3, bubbles=True, fg0=1e-6, R0=3e-6,
                 precision=PREC_FP64, source_in_pml=True):
    """The medium and source shared by the benchmark configurations: water
    (c = 1500 m/s, rho = 1000 kg/m^3, mu = 1e-3 Pa s, sigma = 0.0725 N/m,
    P0 = 101325 Pa, gamma = 1.4, T0 = 293.15 K, K0 = 26.2e-3), dx = 0.1 mm,
    PML on all sides, a focused continuous-wave pressure source on the y0 edge
    with the focus centred in x (extensions E1, E2)."""
    c = default_config()
    lx = n * 1.0e-4 if lx is None else lx
    c.update(grid_x=n, grid_y=n, grid_lx=lx, grid_ly=lx,
             pml_x0=1, pml_x1=1, pml_y0=1, pml_y1=1, pml_grid_number=npml,
             pml_sigma=0.1, pml_order=2, sim_cfl=cfl, sim_order=2,
             sim_smooth_delta_band=2,
             tr_amp=amp_pa / (1000.0 * 1500.0), tr_freq=freq, tr_focal_dist=focal_dist,
             tr_on_wave=1000000, tr_off_wave=0, tr_focused_p=1,
             bub_enable=1 if bubbles else 0, bub_fg0=fg0, bub_R0=R0, bub_P0=101325.0,
             bub_T0=293.15, bub_gamma=1.4, bub_sigma=0.0725, bub_mu=1.0e-3,
             bub_rho=1000.0, bub_K0=26.2e-3,
             ext_focus_x=lx / 2.0, ext_focus_x_set=1,
             ext_source_in_pml=1 if source_in_pml else 0, precision=precision)
    return c
