// brp_internal.h -- device data layout and kernel entry points of
// libbubbles_rp (the MI355X coupled FDTD + Rayleigh-Plesset step).
//
// HBM layout
//   Every grid array lives in one common frame: logical columns
//   i in [-1, N+1] and rows j in [-1, M+1] (the union of the reference's
//   1m, 2m, 2n extents for simulation.order = 2), row-major, element
//   (i, j) at  (j + 1) * pitch + (i + kCol0).  kCol0 = 64 puts column i = 0
//   on a 256-byte boundary for both float and double, so a wavefront's 64
//   consecutive cells are one aligned 256/512-byte segment.  pitch is a
//   multiple of 64 elements.
//   Fields: vx, vy, p.x, p.y (two buffers, rotated as p / pn), p0 in the
//   precision's field type F (double or float); void fraction in four int64
//   fixed-point buffers (value * 2^60) rotated as f_g / f_gn / f_gm / spare.
//   Bubbles: structure of arrays in FP64, sorted by cell at set-up.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/bubbles_rp.h"
#include "rp_vfit.h"

namespace brp {

constexpr int kCol0 = 64;
constexpr double kFgScale = 1152921504606846976.0;       // 2^60
constexpr double kFgInvScale = 8.673617379884035e-19;    // 2^-60
constexpr int kMaxEval = 1024;                           // residual slots per step

// Constant parameters of the grid kernels (passed by value).
struct GridParams {
    int N, M;            // interior cells in x and y
    int pitch;           // elements per frame row
    double dx, dy, rdx, rdy;
    double dt;
    double rho_inf, cs_inf;
    double P_inf;
    int pml_x0, pml_x1, pml_y0, pml_y1;
    int plane_v, plane_p, focused_v, focused_p;
    int src_row;         // source row written (pml.y0 == 0 or extension E1)
    int cylindrical;
    double amp, freq, omega;
    int on_wave, off_wave;
    double fp_x, fp_y, fdist;  // control point and focal distance
    int ap_lo, ap_hi, n_elem;  // aperture cells and element count (E4)
    double s3;           // dt * rho_l * c * c   (MixturePressureKernel)
};

__host__ __device__ inline int64_t frame_off(int i, int j, int pitch) {
    return (int64_t)(j + 1) * pitch + (i + kCol0);
}

// Bit-exact binning: floor(x * rdx + c) with no FMA contraction
// (BubbleUpdateIndexKernel, bubble_CUDA_kernel.cuh:355-356).
__device__ inline int bin_index(double x, double rdx, double c) {
#pragma clang fp contract(off)
    double v = floor(x * rdx + c);
    return (v >= -1.0e9 && v <= 1.0e9) ? (int)v : -1000000000;
}

// Window of cells where any void-fraction buffer may be nonzero.
struct Window {
    int i0, i1, j0, j1;  // inclusive, within [0, N+1]
};

// Device-side per-step scalars written by the kernels and read by the host.
struct StepScalars {
    int bbox[4];                         // min i, max i, min j, max j of bubble stencils
    int flags;                           // bit 0 nonfinite p / f_g >= 1, bit 1 scatter
                                         // range / nonfinite R, bit 2 sub-step cap
    int sub_max;                         // max RP sub-steps this step
    unsigned long long sub_sum;          // sum of RP sub-steps this step
    unsigned int done;                   // bubbles written back by the RP kernel this step
    unsigned int reused;                 // of them, kept from the previous iteration
    double focal[2];                     // (x, y) of the largest temperature increment
    unsigned long long resid[kMaxEval];  // max |dp0| bit patterns (>= 0 doubles)
    unsigned int work[kMaxEval];         // bubble work-queue heads, one per Picard iteration
    double focal_v[256];                 // per-block maxima of the focal-point search
    int focal_i[256];
};

// Bubble arrays (device pointers; n padded to a multiple of 256).
struct BubbleArrays {
    int n;
    double *x, *y;
    double *R[2], *Rn[2], *d1R[2], *PG[2], *bdt[2], *re[2];  // RP state, double-buffered
    double *R_t, *Q_B;
    double *PL[3];        // rotating liquid-pressure history
    double *n_B, *R0, *PG0, *calpha;
    double *vx, *vy;      // v_B (= v_L)
    int *sub;             // sub-steps of the last RP call, per bubble
    int *order;           // work-queue order: bubbles with many sub-steps first
    int *key;             // work-queue keys of a step's first Picard iteration
    int *order_tmp;       // partition scratch: the extreme class before its sort
    int *sort_hist;       // partition scratch: 2048 bin counts / offsets of that sort
    int *part;            // partition scratch: per-block extreme and heavy offsets + totals
    int *bbox_blk;        // per-block stencil bounding boxes of k_motion (4 ints each)
};

// Bubble physics constants shared by every bubble.
struct BubbleParams {
    double PL0, T0, gam, sig, mu, rho, K0, dt0;
    int mbs, mbe, nbs, nbe;
    double dcx, dcy, dcz;   // delta_coef
    double pi4r3, pi;
    int max_substeps;
    int cylindrical;
    int rp_exact;
    int heavy_threshold;    // sub-steps that put a bubble at the queue front
    int extreme_threshold;  // sub-steps that give a bubble a wavefront of its own
    const double *rp_shared;  // device table of rp_table_values (kRpTable doubles)
};

// The persistent RP kernel's fixed parameters, in device memory: the kernel
// reads them where its refill and write-back code uses them instead of
// holding them in scalar registers across its sub-step loop.
struct RPLaunch {
    GridParams g;
    BubbleParams bp;
    BubbleArrays b;
};

// The shortened-chain sub-step's constant table: the shared constants
// (rp_table_values, 15 doubles, padded to 64), then the piecewise polynomials
// of the thermal-impedance functions V and Y of alpha = s (1 + i)
// (rp_vfit.h, built for the simulation's gam), which the kernels stage in LDS.
constexpr int kRpVOffset = 64;
constexpr int kRpVFunc = kVIntervals * (kVDegree + 1) * 2;  // doubles per function
constexpr int kRpVSize = 2 * kRpVFunc;
constexpr int kRpTable = kRpVOffset + kRpVSize;

// The constants shared by every bubble in the shortened-chain RP sub-step, in
// the field order of RPFastShared (rp_device.h): PL0, gam, mu, rho, 1/rho,
// dt0, 3 gam, 3 (gam - 1), 3 (gam - 1) K0 T0, 2 sig, 4 mu, mu^2,
// -2 sig rho, 4 pi K0, 16 pi mu; then (from kRpVOffset) the V and Y tables.
inline void rp_table_values(const BubbleParams &bp, double v[kRpTable]) {
    v[0] = bp.PL0;
    v[1] = bp.gam;
    v[2] = bp.mu;
    v[3] = bp.rho;
    v[4] = 1.0 / bp.rho;
    v[5] = bp.dt0;
    v[6] = 3.0 * bp.gam;
    v[7] = 3.0 * (bp.gam - 1.0);
    v[8] = 3.0 * (bp.gam - 1.0) * bp.K0 * bp.T0;
    v[9] = 2.0 * bp.sig;
    v[10] = 4.0 * bp.mu;
    v[11] = bp.mu * bp.mu;
    v[12] = -2.0 * bp.sig * bp.rho;
    v[13] = 4.0 * bp.pi * bp.K0;
    v[14] = 16.0 * bp.pi * bp.mu;
    for (int k = 15; k < kRpVOffset; k++) v[k] = 0.0;
    rp_vtable_build(v[7], v + kRpVOffset);
}

}  // namespace brp

// Kernel launchers (kernels_*.hip).  All run on stream s.
namespace brp {
template <typename F>
void launch_velocity(hipStream_t s, const GridParams &g, F *vx, F *vy, const F *p0,
                     const int64_t *fgn, Window w, int first_step, const F *cvx_a,
                     const F *cvx_b, const F *cvy_a, const F *cvy_b);
template <typename F>
void launch_velocity_ghosts(hipStream_t s, const GridParams &g, F *vx, F *vy, double t);
template <typename F>
void launch_pressure_first(hipStream_t s, const GridParams &g, const F *vx, const F *vy,
                           const F *pnx, const F *pny, F *px, F *py, F *p0,
                           const int64_t *fg_cur, const int64_t *fg_n, int64_t *fg_mp,
                           Window w, const F *ax, const F *bx, const F *ay, const F *by,
                           const double *xu, const double *rxp, StepScalars *sc);
template <typename F>
void launch_pressure_picard(hipStream_t s, const GridParams &g, const F *vx, const F *vy,
                            const F *pnx, const F *pny, F *px, F *py, F *p0,
                            const int64_t *fg_k, const int64_t *fg_n, int64_t *fg_prev,
                            Window w, const F *ax, const F *bx, const F *ay, const F *by,
                            const double *xu, const double *rxp, StepScalars *sc, int slot);
template <typename F>
void launch_pressure_ghosts(hipStream_t s, const GridParams &g, F *p0, const double *el_amp,
                            const double *el_phase, double t);
template <typename F>
void launch_motion(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                   const BubbleArrays &b, const F *vx, const F *vy, StepScalars *sc);
// The RP launch of Picard iteration `slot` over the queue order, with the
// gathered pressures of launch_queue_keys in PL[plp].
template <typename F>
void launch_bubble_step(hipStream_t s, const RPLaunch &host, const RPLaunch *dev, int cur,
                        int plm, int pln, int plp, int64_t *fg_target, StepScalars *sc, int slot,
                        bool lean, int cus, int per_cu);
// The device's compute units and the persistent RP kernel's workgroups per
// compute unit (its occupancy limit), queried once per handle.
void rp_grid_query(int device, int *cus, int *per_cu);
int bubble_step_blocks(int n_bubbles, bool lean, int cus, int per_cu);
// Whether the next RP launch is bound by its longest chains (lean grid): the
// longest chain of the last launch is long and the sub-steps per lane of a
// one-workgroup-per-CU grid are few against it.
bool bubble_step_lean(int sub_max, double sub_per_launch, int extreme_threshold, int cus);
// Pressure gather of a Picard iteration's bubbles into PL[plp] and its queue
// keys (predicted or previous sub-step counts; -1: the previous iteration's
// results stand).
template <typename F>
void launch_queue_keys(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                       const BubbleArrays &b, const F *p0, int cur, int plp, int slot, int ext);
void launch_order_partition(hipStream_t s, const BubbleArrays &b, const int *key, int threshold,
                            int ext);
void launch_scatter_initial(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                            const BubbleArrays &b, int64_t *fg_target, StepScalars *sc);
void launch_cyl_fold(hipStream_t s, const GridParams &g, int64_t *fg);
// Frozen schedule: record the accepted RP step sizes of the last RP launch
// (same inputs, same sub-step function) and replay them in a later run.
void launch_rp_record(hipStream_t s, const RPLaunch *dev, int n, int cur, int plm, int pln,
                      int plp, const int64_t *off, double *dts, unsigned int *mismatch);
template <typename F>
void launch_rp_replay(hipStream_t s, const RPLaunch *dev, int n, int cur, int plm, int pln,
                      int plp, const F *p0, int64_t *fg_target, StepScalars *sc,
                      const int64_t *off, const double *dts);
// Scatter of the current R_t with accumulation mode 1 (FP64 atomics),
// 2 (compensated FP32 atomics) or 3 (FP32 atomics) into out (zeroed).
void launch_scatter_variant(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                            const BubbleArrays &b, int mode, void *out, int64_t frame);
void launch_scatter_radii(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                          const BubbleArrays &b, const double *Rt, int64_t *fg, StepScalars *sc);
void launch_copy_d(hipStream_t s, double *dst, const double *src, int n);
void launch_bubble_index(hipStream_t s, const GridParams &g, const BubbleArrays &b,
                         double *out, int which);
void launch_zero_scalars(hipStream_t s, StepScalars *sc, int nslots);

// ---- temperature (kernels_thermal.hip) ----
template <typename F>
void launch_temperature(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                        const BubbleArrays &b, bool bubbles, Window w, double T_inf, double *T,
                        double *dT, double *heat, const int64_t *fg, const int64_t *fgn,
                        int first, const F *px, const F *py, const F *pnx, const F *pny);

void launch_focal_point(hipStream_t s, const GridParams &g, const double *dT, const double *heat,
                        StepScalars *sc);

// ---- adjoint (kernels_adjoint.hip) ----
// Adjoint fields, FP64, same frame as the forward fields.
struct AdjField {
    double *p0, *px, *py, *pnx, *pny, *vx, *vy;
};
// A recorded rectangle of a frame array: logical cells [i0, i0 + w) x
// [j0, j0 + h), stored row by row without padding; cells outside read 0.
template <typename T>
struct Rect {
    const T *p;
    int i0, j0, w, h;
    __host__ __device__ bool in(int i, int j) const {
        return (unsigned)(i - i0) < (unsigned)w && (unsigned)(j - j0) < (unsigned)h;
    }
    __device__ T at(int i, int j) const {
        return in(i, j) ? p[(int64_t)(j - j0) * w + (i - i0)] : T(0);
    }
};
template <typename F>
void launch_box_sumsq(hipStream_t s, const GridParams &g, const F *p0, const int box[4],
                      double *out);
template <typename F>
void launch_adj_seed(hipStream_t s, const GridParams &g, Rect<F> p0, double *Lp0,
                     const int box[4], double c);
void launch_adj_pressure_ghosts(hipStream_t s, const GridParams &g, double *L,
                                const double *el_amp, const double *el_phase, double t,
                                double *gA, double *gP);
template <typename F>
// win (may be null = the whole frame): the cells to evaluate; outside them
// the evaluation's adjoint inputs must be zero.
void launch_adj_pressure(hipStream_t s, const GridParams &g, bool first, const AdjField &a,
                         Rect<int64_t> fk, Rect<int64_t> gC, const double *extra, double *Lf,
                         double *LgC, double *LfN, double *LfM, const F *ax, const F *bx,
                         const F *ay, const F *by, const Window *win);
void launch_adj_velocity_ghosts(hipStream_t s, const GridParams &g, double *Lvx, double *Lvy);
template <typename F>
void launch_adj_velocity(hipStream_t s, const GridParams &g, double *Lvx, double *Lvy,
                         const F *p0, Rect<int64_t> fgn, int first, const F *cxa, const F *cxb,
                         const F *cya, const F *cyb, double *Lp0, double *LfC);
void launch_adj_scatter(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                        const BubbleArrays &b, const double *Rt, const double *Lf, double *lRt,
                        double *Lx, double *Ly);
template <typename F>
void launch_adj_gather(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                       const BubbleArrays &b, Rect<F> p0, const double *lPL,
                       const double *lPL_extra, double *Lp0, double *Lx, double *Ly);
template <typename F>
void launch_adj_motion(hipStream_t s, const GridParams &g, const BubbleParams &bp, int n,
                       const double *xs, const double *ys, Rect<F> vx, Rect<F> vy, double *Lvx,
                       double *Lvy, double *Lx, double *Ly);
// One Picard iteration of one step for the RP Jacobians of the reverse
// sweep: the step's RP start state and pressure history, the iteration's
// gathered pressure and the forward's sub-step counts (taped), and where its
// Jacobian columns go (35 x nb doubles when last, else 7 x nb).
struct JacSlot {
    const double *rp[6];  // R, Rn, d1R, PG, dt, re at the step's start
    const double *PLm, *PLn, *PLp;
    const int *sub;
    double *jac;
    int last;  // the step's last Picard iteration
};
constexpr int kJacCols = 7, kJacColsLast = 35;
// The Jacobians of n_slots iterations (slots in device memory), in the
// kernel's order (ord: n_slots x nb indices, bins: kAdjJacBins ints, head: two
// counters, buf: adj_rp_jac_buf_doubles(max_blocks) doubles of reverse-mode
// states); replays whose sub-step count differs from the taped one are
// counted in *mismatch.
constexpr int kAdjJacBins = 8193;
void launch_adj_rp_jacobians(hipStream_t s, const BubbleParams &bp, const BubbleArrays &b,
                             double dt_mix, const JacSlot *slots, int n_slots, unsigned *ord,
                             int *bins, unsigned *head, unsigned *mismatch, double *buf,
                             int max_blocks);
// Workgroups of the Jacobian kernel one compute unit holds (its occupancy),
// and the reverse-mode state buffer a grid of `blocks` workgroups needs.
int adj_rp_jac_per_cu();
size_t adj_rp_jac_buf_doubles(int blocks);
// The transpose of one iteration's RP map: LRe (the adjoint of the step's
// final RP state) only for the last iteration, else null.
void launch_adj_rp_apply(hipStream_t s, int nb, const double *jac, const double *lRt,
                         double *const LRe[4], double *const LRs[4], double *lPLp, double *lPLn,
                         double *lPLm);
void launch_adj_pl_history(hipStream_t s, int n, int picard, double *LA, double *LB,
                           double *lPLn, double *lPLm);
void launch_fill_d(hipStream_t s, double *p, size_t n, double v);
}  // namespace brp
