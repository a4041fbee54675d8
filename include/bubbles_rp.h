/*
 * bubbles_rp.h -- C ABI of the MI355X coupled FDTD + Rayleigh-Plesset step.
 *
 * One handle owns one simulation on one GPU: the staggered-grid mixture
 * fields (velocity, split pressure, void fraction), the PML profiles and the
 * bubble population.  Every call returns an int status (BRP_OK == 0); no call
 * exits the process, and no device pointer crosses this boundary.  A handle is
 * not internally synchronized: use one handle per host thread.  Several
 * handles may live in one process (there is no global state).
 *
 * Reference seam this replaces (AdamPan/CUDA-RP):
 *   - solve_bubbles()                     src/bubble_CUDA.cu:56-476
 *   - initialize_variables() + init_*()   src/bubble_CUDA.cu:479-943, 946-1205
 *   - initialize_CUDA_variables()         src/bubble_CUDA.cu:1231-1520
 *   - destroy_CUDA_variables()            src/bubble_CUDA.cu:1523-1623
 *   - the L3 launch wrappers              inc/bubble_CUDA_kernel.cuh:1186-1411
 *   - the parameter PODs                  inc/bubble_CUDA.h:48-156
 */
#ifndef BUBBLES_RP_H
#define BUBBLES_RP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BRP_ABI_VERSION 2

/* ---- status codes ------------------------------------------------------ */
enum {
    BRP_OK = 0,
    BRP_ERR_INVALID = 1,      /* bad argument or unsupported configuration   */
    BRP_ERR_DEVICE = 2,       /* HIP runtime error (message in last_error)   */
    BRP_ERR_NONFINITE = 3,    /* NaN/Inf or void fraction out of range       */
    BRP_ERR_PICARD_CAP = 4,   /* Picard sub-loop hit max_picard              */
    BRP_ERR_SUBSTEP_CAP = 5,  /* an RP integration hit max_substeps          */
    BRP_ERR_NO_DEVICE = 6,    /* no HIP device visible                       */
    BRP_ERR_STATE = 7         /* call out of order (e.g. no saved slot)      */
};

/* ---- precision modes --------------------------------------------------- */
enum {
    BRP_PREC_FP64 = 0,  /* all fields FP64 (the reference's arithmetic)      */
    BRP_PREC_MIXED = 1  /* FP32 fields + coupling, FP64 RP / void fraction   */
};

/*
 * Simulation parameters.  Field-for-field the Boost program_options keys of
 * src/bubbles.cpp:75-133 (section.key in the comment), plus the declared
 * extensions E1-E4 and the run controls.  Booleans are int32 0/1.
 * brp_config_default() fills the reference's declared defaults.
 */
typedef struct brp_config {
    /* [grid]  bubbles.cpp:76-79 -> grid_t (bubble_CUDA.h:70-76) */
    int32_t grid_x;              /* grid.x   cells in x                      */
    int32_t grid_y;              /* grid.y   cells in y                      */
    double grid_lx;              /* grid.lx  domain length in x (m)          */
    double grid_ly;              /* grid.ly  domain length in y (m)          */
    /* [pml]  bubbles.cpp:80-86 -> PML_t (bubble_CUDA.h:91-97) */
    int32_t pml_x0, pml_x1, pml_y0, pml_y1; /* pml.x0 .. pml.y1             */
    int32_t pml_grid_number;     /* pml.grid_number  NPML                    */
    int32_t pml_order;           /* pml.order        grading exponent        */
    double pml_sigma;            /* pml.sigma        target reflection       */
    /* [simulation]  bubbles.cpp:87-93 -> sim_params_t (bubble_CUDA.h:118-127) */
    double sim_cfl;              /* simulation.cfl                           */
    double sim_dt0;              /* simulation.dt0   (read, unused)          */
    double sim_t_step_max;       /* simulation.t_step_max                    */
    int32_t sim_n_step_max;      /* simulation.n_step_max                    */
    int32_t sim_save_interval;   /* simulation.save_interval                 */
    int32_t sim_order;           /* simulation.order  FDTD order (2 only)    */
    int32_t sim_smooth_delta_band; /* simulation.smooth_delta_band  B        */
    /* [transducer]  bubbles.cpp:94-108 -> transducer_t (bubble_CUDA.h:99-116) */
    double tr_amp;               /* transducer.amp   velocity amplitude      */
    double tr_freq;              /* transducer.freq                          */
    double tr_focal_dist;        /* transducer.focal_dist                    */
    double tr_init_control;      /* transducer.init_control                  */
    double tr_box_size;          /* transducer.box_size                      */
    int32_t tr_pid;              /* transducer.pid                           */
    int32_t tr_pid_start_step;   /* transducer.pid_start_step                */
    int32_t tr_pid_iterate;      /* transducer.pid_iterate                   */
    int32_t tr_cylindrical;      /* transducer.cylindrical                   */
    int32_t tr_on_wave;          /* transducer.on_wave                       */
    int32_t tr_off_wave;         /* transducer.off_wave                      */
    int32_t tr_plane_v;          /* transducer.transducer_v   (Plane_V)      */
    int32_t tr_plane_p;          /* transducer.transducer_p   (Plane_P)      */
    int32_t tr_focused_v;        /* transducer.focused_wave_v (Focused_V)    */
    int32_t tr_focused_p;        /* transducer.focused_wave_p (Focused_P)    */
    /* [bubbles]  bubbles.cpp:109-118 -> bub_params_t (bubble_CUDA.h:129-147) */
    int32_t bub_enable;          /* bubbles.enable                           */
    int32_t pad0_;
    double bub_fg0;              /* bubbles.fg0                              */
    double bub_R0;               /* bubbles.R0                               */
    double bub_P0;               /* bubbles.P0   (PL0)                       */
    double bub_T0;               /* bubbles.T0                               */
    double bub_gamma;            /* bubbles.gamma                            */
    double bub_sigma;            /* bubbles.liquid_surface_tension           */
    double bub_mu;               /* bubbles.liquid_viscosity                 */
    double bub_rho;              /* bubbles.liquid_density                   */
    double bub_K0;               /* bubbles.K0                               */
    /* ---- extensions (each 0 reproduces the reference exactly) ---------- */
    double ext_focus_x;          /* E2: focal point x when ext_focus_x_set   */
    int32_t ext_focus_x_set;     /* E2: 0 -> fp.x = 0 (bubble_CUDA.cu:548)   */
    int32_t ext_source_in_pml;   /* E1: apply the y0 source row even if pml.y0 */
    int32_t precision;           /* BRP_PREC_FP64 / BRP_PREC_MIXED           */
    int32_t max_picard;          /* Picard cap per step (0 -> 1000)          */
    int32_t max_substeps;        /* RP sub-step cap per call (0 -> 1000000)  */
    int32_t thermal;             /* run the one-way temperature solve
                                    (calculate_temperature, which the reference
                                    runs every step; 0 skips it)              */
    int32_t rp_exact;            /* 1: RP in the reference's operation order;
                                    0: same model, shorter dependency chain   */
    int32_t pad1_;
} brp_config;

/* Per-step statistics written by brp_step (one record per step). */
typedef struct brp_step_stats {
    int64_t nstep;               /* step counter after the step              */
    double time;                 /* simulation time after the step (s)       */
    double resimax0;             /* residual of the first pressure evaluation */
    double resimax;              /* residual of the last pressure evaluation */
    int32_t n_pic;               /* Picard iterations taken (0 -> RP skipped) */
    int32_t rp_max_substeps;     /* max RP sub-steps of any bubble this step */
    double rp_mean_substeps;     /* mean RP sub-steps integrated per bubble
                                    per call (kept results count 0)          */
    int32_t status;              /* BRP_OK or the error of this step         */
    int32_t rp_lean_launches;    /* RP launches of this step on the lean grid
                                    (one workgroup per compute unit)         */
    /* With thermal = 1: the cell (x = i dx, y = j dy) of the largest
     * temperature increment of the step over the 2m grid, the raw input of
     * the reference's focal-point tracking (determine_focal_point on Work
     * after calculate_temperature, bubble_CUDA.cu:310-324); 0 otherwise. */
    double focal_x, focal_y;
    /* RP calls of this step that kept the previous Picard iteration's
     * results: the bubble's gathered pressure was the same, bit for bit, so
     * the integration had the same inputs (extension: the reference
     * integrates every bubble in every iteration, with the same results). */
    int64_t rp_reused;
} brp_step_stats;

/* Derived scalars (init_grid_size, mix_set_time_increment, init_bub_params). */
typedef struct brp_info {
    int32_t nx, ny;              /* interior cells                           */
    int32_t n_bubbles;
    int32_t precision;
    double dx, dy, rdx, rdy;
    double dt;                   /* mixture time step                        */
    double dt0_rp;               /* bubble default sub-step (0.1 dt)         */
    double fp_x, fp_y;           /* focal point                              */
    double omega;
    double time;
    int64_t nstep;
    int64_t device_bytes;        /* device memory owned by the handle        */
} brp_info;

/* Field identifiers for brp_get_field / brp_set_field. */
enum {
    BRP_FIELD_P0 = 0,   /* total pressure, 1m layout [ista1m..iend1m]^2       */
    BRP_FIELD_PX = 1,   /* split pressure p.x                                 */
    BRP_FIELD_PY = 2,   /* split pressure p.y                                 */
    BRP_FIELD_VX = 3,   /* x velocity, [ista2n..iend2n] x [jsta2m..jend2m]    */
    BRP_FIELD_VY = 4,   /* y velocity, [ista2m..iend2m] x [jsta2n..jend2n]    */
    BRP_FIELD_FG = 5,   /* void fraction f_g, 2m layout                       */
    BRP_FIELD_FGN = 6,  /* f_g at step n                                      */
    BRP_FIELD_FGM = 7,  /* f_g at step n-1                                    */
    BRP_FIELD_RHO_M = 8,/* mixture density                                    */
    BRP_FIELD_T = 9,    /* temperature rise                                   */
    BRP_FIELD_COUNT = 10
};

/* Bubble array identifiers for brp_get_bubbles (bubble_t, bubble_CUDA.h:172-189). */
enum {
    BRP_BUB_X = 0, BRP_BUB_Y = 1, BRP_BUB_R_T = 2, BRP_BUB_R_P = 3,
    BRP_BUB_R_N = 4, BRP_BUB_D1R = 5, BRP_BUB_PG = 6, BRP_BUB_PL_P = 7,
    BRP_BUB_PL_N = 8, BRP_BUB_PL_M = 9, BRP_BUB_Q_B = 10, BRP_BUB_N_B = 11,
    BRP_BUB_DT = 12, BRP_BUB_RE = 13, BRP_BUB_VX = 14, BRP_BUB_VY = 15,
    BRP_BUB_IBM_X = 16, BRP_BUB_IBM_Y = 17, BRP_BUB_IBN_X = 18,
    BRP_BUB_IBN_Y = 19, BRP_BUB_R0 = 20,
    BRP_BUB_SUB = 21,   /* RP sub-steps of the bubble's last RP call (diagnostic) */
    BRP_BUB_COUNT = 22
};

typedef struct brp_handle brp_handle;

/* Fills *cfg with the reference defaults (bubbles.cpp:97-132; calloc'd
 * structs are zero) and this library's run-control defaults. */
void brp_config_default(brp_config *cfg);

/* Computes every derived parameter (init_grid_size, init_transducer,
 * init_array, init_sigma, init_grid_vector, init_mix,
 * mix_set_time_increment, init_bub_params, delta_coef:
 * bubble_CUDA.cu:479-521, 1467-1471), allocates device memory and places the
 * reference's default bubbles (init_bub_array, bubble_CUDA.cu:797-943).
 * Replaces solve_bubbles' set-up phase (bubble_CUDA.cu:92-188). */
int brp_create(const brp_config *cfg, brp_handle **out);

/* Frees everything (destroy_CUDA_variables, bubble_CUDA.cu:1523-1623). */
void brp_destroy(brp_handle *h);

/* Last error message of this handle ("" when none).  With h == NULL, the
 * message of the last failed brp_create on this thread. */
const char *brp_last_error(const brp_handle *h);

int brp_get_info(const brp_handle *h, brp_info *info);

/* Replaces the bubble population (extension E3: per-bubble R0 and arbitrary
 * positions).  x, y, R0 are n host doubles.  n_B may be NULL, in which case
 * each weight follows bubble_input (bubble_CUDA.cu:922-931) with fg_in.
 * The state of every bubble is reset as bubble_input does.  Replaces
 * init_bub_array / bubble_input (bubble_CUDA.cu:797-943). */
int brp_set_bubbles(brp_handle *h, int32_t n, const double *x, const double *y,
                    const double *R0, const double *n_B, double fg_in);

/* Extension E4: partitions the focused aperture into n_elem contiguous
 * elements with amplitude amp[e] (Pa) and phase[e] (rad).  n_elem == 0
 * restores the reference's single source pamp = rho*c*amp, phase 0
 * (MixtureBoundaryPressureKernel, bubble_CUDA_kernel.cuh:903-925). */
int brp_set_source(brp_handle *h, int32_t n_elem, const double *amp,
                   const double *phase);

/* Runs the initialisation kernels: bubble indices, the initial void-fraction
 * scatter and f_gn = f_gm = f_g (bubble_CUDA.cu:172-188).  brp_step calls it
 * on first use; call it explicitly to read the initial state. */
int brp_prepare(brp_handle *h);

/* Advances k coupled steps exactly as solve_bubbles' main loop
 * (bubble_CUDA.cu:203-286): clock, store/predict, velocity, bubble motion,
 * pressure, Picard sub-loop of gather -> RP -> scatter -> pressure, then
 * (optionally) temperature and properties.  stats may be NULL, else it
 * receives k records.  Replaces store_variables, calculate_velocity_field,
 * bubble_motion, calculate_pressure_field, interpolate_bubble_pressure,
 * solve_bubble_radii, calculate_void_fraction, calculate_temperature and
 * calculate_properties (bubble_CUDA_kernel.cuh:1198-1411). */
int brp_step(brp_handle *h, int32_t k, brp_step_stats *stats);

/* Logical extents of a field: columns i0..i0+nx-1, rows j0..j0+ny-1. */
int brp_field_shape(const brp_handle *h, int32_t field, int32_t *nx,
                    int32_t *ny, int32_t *i0, int32_t *j0);

/* Synchronous copy of a field into a caller-owned row-major double buffer
 * with leading dimension ld (>= nx).  Replaces the save-interval D2H copies
 * (bubble_CUDA.cu:364-392). */
int brp_get_field(brp_handle *h, int32_t field, double *dst, int64_t ld);

/* Overwrites a field from a caller buffer.  BRP_FIELD_FG sets f_g, f_gn and
 * f_gm together (the initial void-fraction map); FGN, FGM, P0, PX, PY, VX
 * and VY set that array alone.  Runs brp_prepare first if needed. */
int brp_set_field(brp_handle *h, int32_t field, const double *src, int64_t ld);

/* Copies one bubble array (n_bubbles values, in the order given to
 * brp_set_bubbles) into dst. */
int brp_get_bubbles(brp_handle *h, int32_t which, double *dst);

/* brp_get_bubbles without any allocation: no hipMalloc, hipFree or other
 * allocating call, so code that must not allocate (under HIP graph capture,
 * in a tight serving loop) can read the bubbles.  No counterpart in the
 * reference; it exists for those callers.  The handle must be prepared
 * (brp_prepare or a step), else BRP_ERR_STATE.  The values are staged in one
 * statically allocated device array per GPU of 4,194,304 doubles (32 MiB,
 * held until the process ends): more bubbles than that give
 * BRP_ERR_INVALID.  All calls on one GPU share that array, so calls that
 * overlap on one GPU (from several host threads) race. */
int brp_get_bubbles_noalloc(brp_handle *h, int32_t which, double *dst);

/* Frozen schedule, for finite differences of the piecewise-smooth map that
 * brp_adjoint differentiates: mode 1 records every step's Picard count and,
 * for every RP launch, every bubble's accepted RP step sizes; mode 2 replays
 * them in order instead of the residual test (bubble_CUDA.cu:243) and the
 * step-size control (kernel.cuh:241-247) -- the rest of the arithmetic is
 * unchanged; mode 0 is the reference behaviour.  Setting a mode rewinds the
 * replay; mode 1 starts a new record.  Shortened-chain RP (rp_exact = 0)
 * only; brp_adjoint needs mode 0.  brp_schedule_copy makes dst (same device,
 * same bubble count) replay src's record, which the two then share.  No
 * reference counterpart (the reference has no adjoint). */
int brp_schedule_mode(brp_handle *h, int32_t mode);
int brp_schedule_copy(brp_handle *dst, const brp_handle *src);

/* Accumulation schemes of the void-fraction scatter (brp_scatter_variant). */
enum {
    BRP_VF_INT64 = 0,           /* the library's: int64 fixed point (value * 2^60),
                                   exact integer sums, independent of order  */
    BRP_VF_FP64_ATOMIC = 1,     /* FP64 atomic adds, the reference's arithmetic
                                   (atomicAdd by CAS, kernel.cuh:95-114); the
                                   rounding depends on the order of the adds */
    BRP_VF_FP32_COMPENSATED = 2,/* FP32 atomic adds plus each add's exact
                                   rounding error summed in a second FP32 array */
    BRP_VF_FP32_ATOMIC = 3      /* plain FP32 atomic adds                    */
};

/* Diagnostic: the void-fraction scatter (VoidFractionReverseLookupKernel,
 * kernel.cuh:576-619) of the bubbles' current clamped radii R_t at their
 * current positions, accumulated with scheme `mode`, copied as doubles into
 * dst (the BRP_FIELD_FG layout, leading dimension ld).  The simulation state
 * is not changed.  Used to compare accumulation schemes on one population. */
int brp_scatter_variant(brp_handle *h, int32_t mode, double *dst, int64_t ld);

/* Overwrites one bubble array (n_bubbles host doubles in the order given to
 * brp_set_bubbles): X, Y, R_T, R_P, R_N, D1R, PG, PL_P, PL_N, PL_M, Q_B, DT,
 * RE, VX or VY -- the state a step starts from (bubble_t,
 * bubble_CUDA.h:172-189), where R_P/R_N/D1R/PG/DT/RE are the state after the
 * last step, before store_variables shifts it (kernel.cuh:1229-1262).  With
 * brp_set_field and brp_set_clock it moves a run to a state taken from
 * another run (a restart, or the oracle seeded with a state the GPU reached).
 * No reference counterpart (the reference restarts from nothing). */
int brp_set_bubble_array(brp_handle *h, int32_t which, const double *src);

/* The compensated clock (bubble_CUDA.cu:68-69, 208-212): step counter, time
 * and its two-sum carry.  brp_set_clock with nstep > 0 also marks the first
 * step as done (the mixture density is then (1 - f_g) rho, not rho).  No
 * reference counterpart (see brp_set_bubble_array). */
int brp_get_clock(const brp_handle *h, int64_t *nstep, double *time, double *carry);
int brp_set_clock(brp_handle *h, int64_t nstep, double time, double carry);

/* Full-state snapshot into slot (0..15) kept in device memory, and its
 * restore.  The snapshot holds every field, every bubble array, the clock
 * and the step counter, so a restored run reproduces the original bit for
 * bit.  No reference counterpart (the reference has no checkpointing). */
int brp_save_state(brp_handle *h, int32_t slot);
int brp_restore_state(brp_handle *h, int32_t slot);

/* Moves the focused source's control point and focal distance (the per-step
 * focal_point_c / focal_dist_c upload of bubble_CUDA.cu:213-215, driven by
 * the PID controller of bubble_CUDA.cu:326-348) for the following steps; the
 * aperture (the row-0 cells within f_dist of the point) and its element
 * partition are recomputed. */
int brp_set_focus(brp_handle *h, double fp_x, double fp_y, double f_dist);

/* Device time of the last brp_step call split by phase (ms): [0] velocity,
 * [1] bubble motion, [2] pressure, [3] gather+RP+scatter, [4] other (ordering of the bubble work queue). */
int brp_get_phase_times(const brp_handle *h, double ms[5]);

/* Objective of brp_adjoint: J = mean of p0^2 over the cells
 * [box_i0, box_i1] x [box_j0, box_j1] (inclusive, within 1..N x 1..M) and
 * over the final p0 of each of the last `window` of `n_steps` steps.  The
 * forward run keeps a checkpoint every `checkpoint_every` steps
 * (0: ceil(sqrt(n_steps))) and recomputes each segment during the reverse
 * sweep; the bubbles' RP results of every step are kept on a device tape
 * (when it fits in 96 GB and flags lacks BRP_ADJ_NO_TAPE), so that the
 * recomputation replays the grid only. */
typedef struct brp_objective {
    int32_t box_i0, box_i1, box_j0, box_j1;
    int32_t n_steps, window, checkpoint_every;
    int32_t flags;  /* BRP_ADJ_NO_TAPE: recompute RP in the reverse sweep
                       instead of keeping a tape of every step's bubble
                       intermediates (less memory, about one forward more) */
} brp_objective;
#define BRP_ADJ_NO_TAPE 1

typedef struct brp_adjoint_stats {
    double J;
    double forward_ms;     /* the checkpointed forward sweep, host wall time */
    double reverse_ms;     /* recomputation + adjoint kernels, host wall time */
    int64_t memory_bytes;  /* device bytes of checkpoints, tape and adjoint state */
    int32_t n_checkpoints, segment;
    int64_t schedule_mismatch; /* replayed RP calls whose sub-step count differed
                                  from the forward's (0 when the replay is exact) */
} brp_adjoint_stats;

/* Discrete adjoint of n_steps coupled steps from the current state (the
 * state must be the start of the run whose parameters are differentiated).
 * Returns J and its gradient with respect to the per-element source
 * amplitudes and phases set by brp_set_source (grad_amp, grad_phase: n_elem
 * doubles each, may be NULL) and to the initial void-fraction map
 * (grad_fg0: (N+2) x (M+2) logical cells i, j in [0, N+1], row-major with
 * leading dimension ld, may be NULL), where the map is f_g = f_gn = f_gm as
 * set by brp_set_field(BRP_FIELD_FG).  The Picard counts and the RP sub-step
 * schedule of the forward run are frozen (replayed, not differentiated).
 * On return the handle holds the state after n_steps.  Cartesian, order 2,
 * thermal off.  No reference counterpart (the reference has no adjoint). */
int brp_adjoint(brp_handle *h, const brp_objective *obj, double *grad_amp,
                double *grad_phase, double *grad_fg0, int64_t ld,
                brp_adjoint_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* BUBBLES_RP_H */
