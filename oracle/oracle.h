/*
 * oracle.h -- host FP64 restatement of the reference's coupled step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load liboracle.so, and only as the checker
 * or the CPU baseline; the product library never links or calls it.
 *
 * Every function follows AdamPan/CUDA-RP's arithmetic 1:1 (same operation
 * order, same FP64 literals, -ffp-contract=off) with consistent logical (i,j)
 * indexing; each definition in oracle.cpp cites the reference file:line it
 * restates.  The API mirrors include/bubbles_rp.h with an oracle_ prefix, so
 * a parity test drives both with the same calls.
 */
#ifndef BRP_ORACLE_H
#define BRP_ORACLE_H

#include "../include/bubbles_rp.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle oracle;

oracle *oracle_create(const brp_config *cfg, char *err, int errlen);
void oracle_destroy(oracle *o);
int oracle_get_info(const oracle *o, brp_info *info);
int oracle_set_bubbles(oracle *o, int32_t n, const double *x, const double *y,
                       const double *R0, const double *n_B, double fg_in);
int oracle_set_source(oracle *o, int32_t n_elem, const double *amp,
                      const double *phase);
int oracle_prepare(oracle *o);
int oracle_step(oracle *o, int32_t k, brp_step_stats *stats);
int oracle_field_shape(const oracle *o, int32_t field, int32_t *nx,
                       int32_t *ny, int32_t *i0, int32_t *j0);
int oracle_get_field(const oracle *o, int32_t field, double *dst, int64_t ld);
int oracle_set_field(oracle *o, int32_t field, const double *src, int64_t ld);
int oracle_get_bubbles(const oracle *o, int32_t which, double *dst);
/* Probes for the pinning tests (no state change): the p0 gather at every
 * bubble, the scatter of given per-bubble volumes, one radius update. */
int oracle_gather_p0(oracle *o, double *PL_out);
int oracle_scatter_values(oracle *o, const double *q, double *fg_out);
const char *oracle_last_error(const oracle *o);
/* Frozen schedule (for finite differences of a piecewise-smooth map):
 * mode 1 records every step's Picard count and every accepted RP step size,
 * mode 2 replays them in order instead of the residual test and the step
 * control, mode 0 is the reference behaviour.  Setting a mode rewinds the
 * replay position; mode 1 clears the record.  schedule_copy hands a record to
 * another instance with the same bubble count. */
int oracle_schedule_mode(oracle *o, int32_t mode);
/* brp_set_bubble_array, brp_get_clock, brp_set_clock of the oracle. */
int oracle_set_bubble_array(oracle *o, int32_t which, const double *src);
int oracle_get_clock(const oracle *o, int64_t *nstep, double *time, double *carry);
int oracle_set_clock(oracle *o, int64_t nstep, double time, double carry);
/* brp_set_focus of the oracle. */
int oracle_set_focus(oracle *o, double fp_x, double fp_y, double f_dist);
int oracle_schedule_copy(oracle *dst, const oracle *src);
/* OpenMP threads used by oracle_step (1 = serial; result is identical). */
void oracle_set_threads(oracle *o, int32_t n);

/* ---- introspection used by the pinning tests ------------------------- */
/* 28 ints of array_index_t in declaration order (bubble_CUDA.h:59-69). */
int oracle_array_index(const oracle *o, int32_t out[28]);
/* PML profile: which = 0 mx, 1 my, 2 nx, 3 ny (init_sigma). */
int oracle_sigma(const oracle *o, int32_t which, double *dst, int32_t *n);
/* delta_coef (bubble_CUDA.cu:1467-1471), control distance, Pi4r3. */
int oracle_scalars(const oracle *o, double out[8]);
/* smooth_delta_x / smooth_delta_y at grid index pos for coordinate b. */
double oracle_smooth_delta(const oracle *o, int32_t axis, int32_t pos, double b);

/* ---- single-bubble physics (bubble_CUDA_kernel.cuh:227-340) ----------- */
typedef struct oracle_bubble_params {
    double R0, R03, PL0, PG0, T0, gam, sig, mu, rho, K0, coeff_alpha, dt0;
} oracle_bubble_params;

/* Fills the per-bubble constants exactly as init_bub_params. */
void oracle_bubble_params_init(oracle_bubble_params *bp, double R0, double PL0,
                               double T0, double gam, double sig, double mu,
                               double rho, double K0, double dt_mix);
void oracle_coth(double re, double im, double out[2]);
void oracle_upsilon(double re, double im, double gam, double out[2]);
void oracle_solve_lp(double re, double im, double R, double out[2]);
double oracle_solve_omega_n(double PG, double R, const oracle_bubble_params *bp,
                            double alpha_out[2]);
/* One BubbleRadiusKernel body (bubble_CUDA_kernel.cuh:485-551) for a single
 * bubble.  st[] = {R_pn, R_nn, d1_R_n, PG_n, dt_n, re_n, PL_m, PL_n, PL_p}
 * in; out[] = {R_t, R_p, R_n, d1_R_p, PG_p, dt, re, Q_B} out.  Returns the
 * sub-step count. */
int oracle_rp_integrate(const oracle_bubble_params *bp, double dt_mix,
                        const double st[9], double out[8], int32_t max_substeps);
/* Total FP64 operation count of the last oracle_rp_integrate call, by the
 * counting rule of oracle.cpp (add/sub/mul/div/sqrt = 1, fma n/a). */
int64_t oracle_rp_flops_last(void);
/* One solveRayleighPlesset call; io = {Rt, Rp, Rn, d1R, dt, remain}. */
void oracle_rp_radius(const oracle_bubble_params *bp, double PG, double PL, double io[6]);

#ifdef __cplusplus
}
#endif
#endif
