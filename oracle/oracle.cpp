// oracle.cpp -- host FP64 restatement of AdamPan/CUDA-RP's coupled step.
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  Compiled with -O2
// -ffp-contract=off so no FMA contraction changes the reference's rounding.
//
// Conventions
//   * Logical indices (i, j) as in the reference's array_index_t
//     (bubble_CUDA.cu:561-633).  Every array is stored with its own extents
//     and a consistent logical index; the reference's p0-with-p-pitch
//     aliasing in MixturePressureKernel (bubble_CUDA_kernel.cuh:826,862-864)
//     is pitch dependent and is NOT reproduced.
//   * Boundary kernels whose threads race in the reference
//     (VelocityBoundaryKernel, MixtureBoundaryPressureKernel) are defined as
//     their parts executed in program order, each part over all its threads
//     before the next part starts -- the order warp-mates observe.
//   * Expressions keep the reference's operation order and literals.
//
// Extensions (all off by default = reference behaviour):
//   E1 ext_source_in_pml : the y0 source row is written even when pml.y0 = 1
//   E2 ext_focus_x       : focal point x coordinate (reference: 0)
//   E3 set_bubbles       : per-bubble R0 and positions
//   E4 set_source        : aperture split into elements with amp/phase

#include "oracle.h"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---- operation counter for the RP FLOP model ----------------------------
// Counting rule: each IEEE add, sub, mul, div and sqrt is one operation;
// tanh and tan are one each; comparisons, min/max, abs and negation are 0.
thread_local int64_t g_ops = 0;
#define OPS(n) (g_ops += (n))

const double kPi = std::acos(-1.0);                 // bubble_CUDA.cu:13
const double kPi4r3 = kPi * 4.0 / 3.0;              // bubble_CUDA.cu:14

struct cplx {
    double re, im;
};
// complex.cuh:13-85
inline cplx mk(double a, double b) { return cplx{a, b}; }
inline cplx operator+(cplx a, cplx b) { OPS(2); return mk(a.re + b.re, a.im + b.im); }
inline cplx operator+(double a, cplx b) { OPS(1); return mk(a + b.re, b.im); }
inline cplx operator+(cplx a, double b) { OPS(1); return mk(a.re + b, a.im); }
inline cplx operator-(cplx a, cplx b) { OPS(2); return mk(a.re - b.re, a.im - b.im); }
inline cplx operator-(cplx a, double b) { OPS(1); return mk(a.re - b, a.im); }
inline cplx operator-(double a, cplx b) { OPS(1); return mk(a - b.re, -b.im); }
inline cplx operator*(cplx a, cplx b) {
    OPS(6);
    return mk(a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im);
}
inline cplx operator*(cplx a, double b) { OPS(2); return mk(a.re * b, a.im * b); }
inline cplx operator*(double a, cplx b) { OPS(2); return mk(a * b.re, a * b.im); }
inline cplx operator/(cplx a, cplx b) {
    OPS(11);
    double c = (b.re * b.re + b.im * b.im);
    return mk((a.re * b.re + a.im * b.im) / c, (a.im * b.re - a.re * b.im) / c);
}
inline cplx operator/(cplx a, double b) { OPS(2); return mk(a.re / b, a.im / b); }
inline cplx operator/(double a, cplx b) {
    OPS(7);
    double c = (b.re * b.re + b.im * b.im);
    return mk(a * b.re / c, -a * b.im / c);
}
inline double cabs(cplx a) { OPS(4); return std::sqrt(a.re * a.re + a.im * a.im); }
// complex.cuh:81-85
inline cplx coth(cplx a) {
    OPS(3);
    double x = std::tanh(a.re), y = std::tan(a.im);
    return mk(1.0, x * y) / mk(x, y);
}

// bubble_CUDA_kernel.cuh:145-149 -- an absolute DBL_EPSILON
inline double epsilon_(double) { return 2.22044604925031308e-016; }

struct BubParams {  // per-bubble slice of bub_params_t (bubble_CUDA.h:129-147)
    double R0, R03, PL0, PG0, T0, gam, sig, mu, rho, K0, coeff_alpha, dt0;
};

// bubble_CUDA_kernel.cuh:227-260
// forced_dtp: when not null, the step size is taken from *forced_dtp instead
// of the step control (frozen-schedule replay; not part of the reference).
void solveRayleighPlesset(double *Rt, double *Rp, double *Rn, double *d1R, double PG,
                          double PL, double *dt, double *remain, const BubParams &bp,
                          const double *forced_dtp = nullptr) {
    double Rm, Rnd2R, dtm, dtp;
    Rm = *Rn;
    *Rn = *Rp;
    Rnd2R = (PG - PL - 2.0 * bp.sig / (*Rn) - 4.0 * bp.mu / (*Rn) * (*d1R)) / bp.rho -
            1.5 * (*d1R) * (*d1R);
    OPS(12);
    dtm = *dt;
    dtp = std::min(1.01 * (*dt), bp.dt0);
    OPS(1);
    if (std::fabs(*d1R) > epsilon_(std::fabs(*d1R))) {
        dtp = std::min(dtp, 0.01 * (*Rn) / std::fabs(*d1R));
        OPS(2);
    }
    if (std::fabs(Rnd2R) > epsilon_(std::fabs(Rnd2R))) {
        dtp = std::min(dtp, 0.1 * (*Rn) / std::sqrt(std::fabs(Rnd2R)));
        OPS(3);
    }
    if (forced_dtp) dtp = *forced_dtp;
    *Rt = *Rp = (1.0 + dtp / dtm) * (*Rn) - dtp / dtm * Rm +
                (dtp / (*Rn)) * (0.5 * (dtp + dtm) * Rnd2R);
    OPS(12);
    *d1R = (dtm * dtm * (*Rp) + (-dtm * dtm + dtp * dtp) * (*Rn) - dtp * dtp * Rm) /
               (dtm * dtp * (dtm + dtp)) +
           (dtp / (*Rn)) * Rnd2R;
    OPS(17);
    *dt = dtp;
    if (dtp >= *remain) {
        dtp = *remain;
        *Rt = (1.0 + dtp / dtm) * (*Rn) - dtp / dtm * Rm +
              (dtp / (*Rn)) * (0.5 * (dtp + dtm) * Rnd2R);
        OPS(12);
    }
    *remain -= dtp;
    OPS(1);
}

// bubble_CUDA_kernel.cuh:291-294
cplx Alpha(double R, double omega_N, double coeff_alpha) {
    OPS(3);
    return mk(1.0, 1.0) * std::sqrt(coeff_alpha / R * omega_N);
}

// bubble_CUDA_kernel.cuh:297-307
cplx Upsilon(cplx a, double gam) {
    if (cabs(a) > 1.0e-1) {
        OPS(3);
        return (3.0 * gam) / (1.0 + (3.0 * ((gam - 1.0) * ((a * coth(a) - 1.0) / (a * a)))));
    } else {
        OPS(7);
        return (3.0 * gam) /
               (1.0 + (3.0 * ((gam - 1.0) *
                              ((1.0 / 3.0 + a * a * (-1.0 / 45.0 + a * a * (2.0 / 945.0 - a * a / 4725.0)))))));
    }
}

// bubble_CUDA_kernel.cuh:310-323
cplx solveLp(cplx a, double R) {
    cplx ctmp;
    if (cabs(a) > 1.0e-1) {
        ctmp = a * coth(a) - 1.0;
        ctmp = (a * a - 3.0 * ctmp) / (a * a * ctmp);
    } else {
        OPS(4);
        ctmp = 1.0 / 5.0 + a * a * (-1.0 / 175.0 + a * a * (2.0 / 7875.0 - a * a * 37.0 / 3031875.0));
    }
    return R * ctmp;
}

// bubble_CUDA_kernel.cuh:263-288
double solveOmegaN(cplx *alpha_N, double PG, double R, const BubParams &bp) {
    double Rx = R / bp.R0;
    double coef1 = bp.PG0 / (PG * Rx * Rx * Rx);
    double coef2 = PG * bp.rho;
    double coef3 = 4.0 / (R * R);
    double value1 = -2.0 * bp.sig * bp.rho / R;
    OPS(11);
    cplx Upsilon_N = mk(3.0 * bp.gam, 0.0) * coef1;
    OPS(1);
    double mu_eff = bp.mu;
    double eta = Upsilon_N.re * coef2 + value1 - mu_eff * mu_eff * coef3;
    double omega_N = std::sqrt(std::max(eta, 1.0e-6 * epsilon_(eta))) / (bp.rho * R);
    OPS(9);
    for (int i = 0; i < 3; i++) {
        *alpha_N = Alpha(R, omega_N, bp.coeff_alpha);
        Upsilon_N = Upsilon(*alpha_N, bp.gam) * coef1;
        mu_eff = bp.mu + Upsilon_N.im * PG / (4.0 * (omega_N));
        eta = Upsilon_N.re * coef2 + value1 - mu_eff * mu_eff * coef3;
        omega_N = std::sqrt(std::max(eta, 1.0e-6 * epsilon_(eta))) / (bp.rho * R);
        OPS(13);
    }
    *alpha_N = Alpha(R, omega_N, bp.coeff_alpha);
    return omega_N;
}

// bubble_CUDA_kernel.cuh:326-340
double solvePG(double PGn, double Rp, double Rn, double omega_N, double dt, cplx Lp_N,
               const BubParams &bp) {
    double R = 0.5 * (Rp + Rn);
    double Rx = R / bp.R0;
    double Rpx = Rp / bp.R0;
    double Rnx = Rn / bp.R0;
    double Lp_NR = Lp_N.re / (Lp_N.re * Lp_N.re + Lp_N.im * Lp_N.im);
    double Lp_NI = Lp_N.im / (Lp_N.re * Lp_N.re + Lp_N.im * Lp_N.im);
    double c0 = dt * 3.0 * (bp.gam - 1.0) * bp.K0 * bp.T0 / (R * bp.PG0);
    double c3 = 1.5 * bp.gam * (Rp - Rn) / R + c0 * Lp_NR * 0.5 * Rx * Rx * Rx;
    double c1 = 1.0 + c3 - c0 * Lp_NI / (omega_N * dt) * Rpx * Rpx * Rpx;
    double c2 = 1.0 - c3 - c0 * Lp_NI / (omega_N * dt) * Rnx * Rnx * Rnx;
    OPS(51);
    return (c2 * PGn + c0 * Lp_NR * bp.PG0) / c1;
}

// The body of BubbleRadiusKernel for one bubble (bubble_CUDA_kernel.cuh:485-551).
struct RPIn {
    double R_pn, R_nn, d1_R_n, PG_n, dt_n, re_n, PL_m, PL_n, PL_p;
};
struct RPOut {
    double R_t, R_p, R_n, d1_R_p, PG_p, dt, re, Q_B;
};
// sched: when not null, the accepted step sizes are appended to it
// (replay == false) or taken from it in order (replay == true).
int rp_integrate(const BubParams &bp, double dt_mix, const RPIn &in, RPOut *out,
                 int max_substeps, bool *capped, std::vector<double> *sched = nullptr,
                 bool replay = false) {
    double PGn, PL, Rt = 0.0, omega_N;
    double dTdr_R = 0, SumHeat, SumVis;
    cplx alpha_N, Lp_N;
    double PGp, Rp, Rn, d1Rp, dt_L, remain, time;
    double PC0, PC1, PC2;
    double s0, s1;
    int counter;

    Rp = in.R_pn;
    Rn = in.R_nn;
    d1Rp = in.d1_R_n;
    PGp = in.PG_n;
    dt_L = in.dt_n;
    remain = in.re_n + dt_mix;
    time = -in.re_n;
    PC0 = in.PL_n + bp.PL0;
    PC1 = 0.5 * (in.PL_p - in.PL_m) / dt_mix;
    PC2 = 0.5 * (in.PL_p + in.PL_m - 2.0 * in.PL_n) / (dt_mix * dt_mix);
    OPS(12);
    SumHeat = 0.0;
    SumVis = 0.0;
    counter = 0;
    *capped = false;
    while (remain > 0.0) {
        if (counter >= max_substeps) {
            *capped = true;
            break;
        }
        PGn = PGp;
        PL = PC2 * time * time + PC1 * time + PC0;
        OPS(5);
        if (replay && counter >= (int)sched->size()) {  // schedule exhausted
            *capped = true;
            break;
        }
        solveRayleighPlesset(&Rt, &Rp, &Rn, &d1Rp, PGn, PL, &dt_L, &remain, bp,
                             replay ? &(*sched)[counter] : nullptr);
        if (sched && !replay) sched->push_back(dt_L);
        time = time + dt_L;
        OPS(1);
        omega_N = solveOmegaN(&alpha_N, PGn, Rn, bp);
        Lp_N = solveLp(alpha_N, Rn);
        PGp = solvePG(PGn, Rp, Rn, omega_N, dt_L, Lp_N, bp);
        s0 = 0.5 * (Rp + Rn);
        s1 = 1 / (Lp_N.re * Lp_N.re + Lp_N.im * Lp_N.im) * bp.T0 / (bp.PG0 * bp.R03);
        dTdr_R = Lp_N.re * s1 * (bp.PG0 * bp.R03 - 0.5 * (PGp + PGn) * s0 * s0 * s0) +
                 Lp_N.im * s1 * (PGp * Rp * Rp * Rp - PGn * Rn * Rn * Rn) / (omega_N * dt_L);
        SumHeat -= 4.0 * kPi * Rp * Rp * bp.K0 * dTdr_R * dt_L;
        SumVis += 4.0 * kPi * Rp * Rp * 4.0 * bp.mu * d1Rp / Rp * d1Rp * dt_L;
        OPS(47);
        counter++;
    }
    out->R_t = Rt;
    out->R_p = Rp;
    out->R_n = Rn;
    out->d1_R_p = d1Rp;
    out->PG_p = PGp;
    out->dt = dt_L;
    out->re = remain;
    out->Q_B = (SumHeat + SumVis) / (dt_mix - remain + in.re_n);
    OPS(4);
    return counter;
}

// ---- water properties (bubble_CUDA_kernel.cuh:153-211) --------------------
inline double density_water(double, double) { return 1000.0; }
inline double adiabatic_sound_speed_water(double, double) { return 1500.0; }
inline double thermal_conductivity_water(double T) {
    return 0.76760 + (7.5390e-3f * T) - 9.8250e-6f * (T * T);
}
inline double specific_heat_water(double T) {
    double H2O = 18.0e-3f;
    return (917.5 - 10.1016f * T + 0.0454134f * (T * T) - 9.07517e-5 * (T * T * T) +
            6.80700e-8f * (T * T * T * T)) /
           H2O;
}

struct ArrayIndex {  // array_index_t (bubble_CUDA.h:59-69)
    int ms, me, ns, ne;
    int istam, iendm, istan, iendn;
    int ista1m, iend1m, ista1n, iend1n;
    int ista2m, iend2m, ista2n, iend2n;
    int jstam, jendm, jstan, jendn;
    int jsta1m, jend1m, jsta1n, jend1n;
    int jsta2m, jend2m, jsta2n, jend2n;
};

// A 2-D array with logical index ranges [i0, i0+nx) x [j0, j0+ny).
struct Grid {
    int i0 = 0, j0 = 0, nx = 0, ny = 0;
    std::vector<double> v;
    void init(int i0_, int i1_, int j0_, int j1_, double fill) {
        i0 = i0_;
        j0 = j0_;
        nx = i1_ - i0_ + 1;
        ny = j1_ - j0_ + 1;
        v.assign((size_t)nx * ny, fill);
    }
    bool in(int i, int j) const { return i >= i0 && i < i0 + nx && j >= j0 && j < j0 + ny; }
    double &operator()(int i, int j) { return v[(size_t)(j - j0) * nx + (i - i0)]; }
    double operator()(int i, int j) const { return v[(size_t)(j - j0) * nx + (i - i0)]; }
};

}  // namespace

struct oracle {
    brp_config cfg{};
    std::string err;
    int nthreads = 1;

    // derived parameters
    ArrayIndex a{};
    double dx = 0, dy = 0, rdx = 0, rdy = 0;
    double fp_x = 0, fp_y = 0, omega = 0;
    double control_x = 0, control_y = 0, control_dist = 0;
    double T_inf = 0, P_inf = 0, rho_inf = 0, cs_inf = 0, dt = 0;
    double dt0_rp = 0;
    int mbs = 0, mbe = 0, nbs = 0, nbe = 0;
    double dcoef_x = 0, dcoef_y = 0, dcoef_z = 0;
    std::vector<double> sig_mx, sig_my, sig_nx, sig_ny;  // index i - ista1m / i - ista2n
    std::vector<double> xu, rxp;                            // init_grid_vector
    // source elements (E4)
    int n_elem = 0;
    std::vector<double> el_amp, el_phase;
    int ap_lo = 1, ap_hi = 0;

    // fields
    Grid p0, px, py, pnx, pny, vx, vy, c_sl, rho_m, rho_l, fg, fgn, fgm, Work, T, k_m, C_pm, Ex, Ey;

    // bubbles (bubble_t, bubble_CUDA.h:172-189) + per-bubble constants (E3)
    int nb = 0;
    std::vector<double> bx, by, R_t, R_p, R_pn, R_n, R_nn, d1_R_p, d1_R_n, PG_p, PG_n, PL_p,
        PL_n, PL_m, Q_B, n_B, bdt, bdt_n, re, re_n, vBx, vBy, vLx, vLy;
    std::vector<int> ibmx, ibmy, ibnx, ibny;
    std::vector<int> sub;  // RP sub-steps of the last RP call (diagnostic)
    std::vector<BubParams> bp;

    // frozen-schedule record / replay (mode 1 / 2): Picard counts per step and
    // the accepted RP step sizes of every (RP call, bubble)
    int sched_mode = 0;
    std::vector<int> sched_npic;
    std::vector<std::vector<double>> sched_dt;
    size_t sched_step = 0, sched_call = 0;

    // clock (bubble_CUDA.cu:68-69,208-212)
    int64_t nstep = 0;
    double tstep = 0.0, tstepx = 0.0;
    bool prepared = false;
};

namespace {

int fail(oracle *o, int code, const char *fmt, const char *detail = "") {
    char buf[512];
    std::snprintf(buf, sizeof buf, fmt, detail);
    o->err = buf;
    return code;
}

// init_array (bubble_CUDA.cu:561-633)
ArrayIndex init_array(int X, int Y, int order) {
    ArrayIndex a;
    a.ms = -order / 2 + 1;
    a.me = order + a.ms - 1;
    a.ns = -order / 2;
    a.ne = order + a.ns - 1;
    a.istam = 1;
    a.iendm = X;
    a.istan = a.istam - 1;
    a.iendn = a.iendm;
    a.ista1m = a.istan + a.ms;
    a.iend1m = a.iendn + a.me;
    a.ista1n = a.istam + a.ns;
    a.iend1n = a.iendm + a.ne;
    a.ista2m = a.ista1n + a.ms;
    a.iend2m = a.iend1n + a.me;
    a.ista2n = a.ista1m + a.ns;
    a.iend2n = a.iend1m + a.ne;
    a.jstam = 1;
    a.jendm = Y;
    a.jstan = a.jstam - 1;
    a.jendn = a.jendm;
    a.jsta1m = a.jstan + a.ms;
    a.jend1m = a.jendn + a.me;
    a.jsta1n = a.jstam + a.ns;
    a.jend1n = a.jendm + a.ne;
    a.jsta2m = a.jsta1n + a.ms;
    a.jend2m = a.jend1n + a.me;
    a.jsta2n = a.jsta1m + a.ns;
    a.jend2n = a.jend1m + a.ne;
    return a;
}

// init_sigma (bubble_CUDA.cu:946-1205)
void init_sigma(oracle *o) {
    const ArrayIndex &a = o->a;
    const brp_config &c = o->cfg;
    o->sig_mx.assign(a.iend1m - a.ista1m + 1, 0.0);
    o->sig_my.assign(a.jend1m - a.jsta1m + 1, 0.0);
    o->sig_nx.assign(a.iend2n - a.ista2n + 1, 0.0);
    o->sig_ny.assign(a.jend2n - a.jsta2n + 1, 0.0);
    int n;
    int npml = c.pml_grid_number;
    double sig = c.pml_sigma;
    double order = c.pml_order;
    double dx = o->dx, dy = o->dy;
    int nx = c.grid_x, ny = c.grid_y;
    double sxm, sym;
    if (c.pml_x0) {
        sxm = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dx * (double)npml);
        if (a.istam <= npml) {
            int s = std::max(a.istam, a.ms), e = std::min(a.iendm, npml);
            for (int i = s; i <= e; i++) {
                n = npml - i + 1;
                o->sig_mx[i - a.ista1m] = sxm * std::pow(((double)n - 0.5) / ((double)npml), order);
            }
        }
    }
    if (c.pml_x1) {
        sxm = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dx * (double)npml);
        if (nx - npml + 1 <= a.iendm) {
            int s = std::max(a.istam, nx - npml + 1), e = std::min(a.iendm, nx + a.me);
            for (int i = s; i <= e; i++) {
                n = i - nx + npml;
                o->sig_mx[i - a.ista1m] = sxm * std::pow(((double)n - 0.5) / ((double)npml), order);
            }
        }
    }
    if (c.pml_y0) {
        sym = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dy * (double)npml);
        if (a.jstam <= npml) {
            int s = std::max(a.jstam, a.ms), e = std::min(a.jendm, npml);
            for (int j = s; j <= e; j++) {
                n = npml - j + 1;
                o->sig_my[j - a.jsta1m] = sym * std::pow(((double)n - 0.5) / ((double)npml), order);
            }
        }
    }
    if (c.pml_y1) {
        sym = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dy * (double)npml);
        if (ny - npml + 1 <= a.jendm) {
            int s = std::max(a.jstam, ny - npml + 1), e = std::min(a.jendm, ny + a.me);
            for (int j = s; j <= e; j++) {
                n = j - ny + npml;
                o->sig_my[j - a.jsta1m] = sym * std::pow(((double)n - 0.5) / ((double)npml), order);
            }
        }
    }
    if (c.pml_x0) {
        sxm = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dx * (double)npml);
        if (a.istan <= npml - 1) {
            int s = std::max(a.istan, a.ms + a.ns), e = std::min(a.iendn, npml - 1);
            for (int i = s; i <= e; i++) {
                n = npml - i;
                o->sig_nx[i - a.ista2n] = sxm * std::pow(((double)n) / ((double)npml), order);
            }
        }
    }
    if (c.pml_x1) {
        sxm = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dx * (double)npml);
        if (nx - npml + 1 <= a.iendn) {
            int s = std::max(a.istan, nx - npml + 1), e = std::min(a.iendn, nx + a.me + a.ne + 1);
            for (int i = s; i <= e; i++) {
                n = i - nx + npml;
                o->sig_nx[i - a.ista2n] = sxm * std::pow(((double)n) / ((double)npml), order);
            }
        }
    }
    if (c.pml_y0) {
        sym = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dy * (double)npml);
        if (a.jstan <= npml - 1) {
            int s = std::max(a.jstan, a.ms + a.ns), e = std::min(a.jendn, npml - 1);
            for (int j = s; j <= e; j++) {
                n = npml - j;
                o->sig_ny[j - a.jsta2n] = sym * std::pow(((double)n) / ((double)npml), order);
            }
        }
    }
    if (c.pml_y1) {
        sym = -std::log(sig) * 0.5 * ((double)order + 1.0) / (dy * (double)npml);
        if (ny - npml + 1 <= a.jendn) {
            int s = std::max(a.jstan, ny - npml + 1), e = std::min(a.jendn, ny + a.me + a.ne + 1);
            for (int j = s; j <= e; j++) {
                n = j - ny + npml;
                o->sig_ny[j - a.jsta2n] = sym * std::pow(((double)n) / ((double)npml), order);
            }
        }
    }
}

// init_bub_params' per-bubble part (bubble_CUDA.cu:661-675) with E3's R0.
BubParams make_bub_params(const brp_config &c, double R0, double dt_mix) {
    BubParams b;
    b.R0 = R0;
    b.R03 = R0 * R0 * R0;
    b.PL0 = c.bub_P0;
    b.PG0 = b.PL0 + 2.0 * c.bub_sigma / b.R0;
    b.T0 = c.bub_T0;
    b.gam = c.bub_gamma;
    b.sig = c.bub_sigma;
    b.mu = c.bub_mu;
    b.rho = c.bub_rho;
    b.K0 = c.bub_K0;
    b.coeff_alpha = b.gam * b.PG0 * b.R03 / (2.0 * (b.gam - 1.0) * b.T0 * b.K0);
    b.dt0 = 0.1 * dt_mix;
    return b;
}

void resize_bubbles(oracle *o, int n) {
    o->nb = n;
    for (auto *v : {&o->bx, &o->by, &o->R_t, &o->R_p, &o->R_pn, &o->R_n, &o->R_nn, &o->d1_R_p,
                    &o->d1_R_n, &o->PG_p, &o->PG_n, &o->PL_p, &o->PL_n, &o->PL_m, &o->Q_B,
                    &o->n_B, &o->bdt, &o->bdt_n, &o->re, &o->re_n, &o->vBx, &o->vBy, &o->vLx,
                    &o->vLy})
        v->assign(n, 0.0);
    for (auto *v : {&o->ibmx, &o->ibmy, &o->ibnx, &o->ibny, &o->sub}) v->assign(n, 0);
    o->bp.assign(n, BubParams{});
}

// bubble_input (bubble_CUDA.cu:900-943) for bubble b with radius R0.
void bubble_input(oracle *o, int b, double x, double y, double R0, double fg_in,
                  const double *nB_given) {
    o->bx[b] = x;
    o->by[b] = y;
    o->bp[b] = make_bub_params(o->cfg, R0, o->dt);
    const BubParams &p = o->bp[b];
    double Pi = std::acos(-1.0);
    o->R_t[b] = p.R0;
    o->R_p[b] = o->R_pn[b] = p.R0;
    o->R_n[b] = o->R_nn[b] = p.R0;
    o->d1_R_p[b] = o->d1_R_n[b] = 0.0;
    o->PG_p[b] = o->PG_n[b] = p.PG0;
    o->PL_p[b] = o->PL_n[b] = o->PL_m[b] = 0.0;
    if (nB_given) {
        o->n_B[b] = *nB_given;
    } else if (o->cfg.tr_cylindrical) {
        o->n_B[b] = fg_in * (x * o->dx * o->dy) / (4.0 / 3.0 * Pi * std::pow(o->R_t[b], 3));
    } else {
        o->n_B[b] = fg_in * (o->dx * o->dy) / (4.0 / 3.0 * Pi * std::pow(o->R_t[b], 3));
    }
    o->Q_B[b] = 0.0;
    o->bdt[b] = o->bdt_n[b] = p.dt0;
    o->re[b] = o->re_n[b] = 0.0;
    o->vBx[b] = o->vBy[b] = o->vLx[b] = o->vLy[b] = 0.0;
    o->ibmx[b] = o->ibmy[b] = o->ibnx[b] = o->ibny[b] = 0;
}

// Aperture extent for E4: the contiguous run of row-0 cells with s1 <= f_dist.
void find_aperture(oracle *o) {
    o->ap_lo = 1;
    o->ap_hi = 0;
    double ym = -0.5 * o->dy;
    for (int i = o->a.istam; i <= o->a.iendm; i++) {
        double xm = (i - 0.5) * o->dx;
        double s1 = std::sqrt((o->control_x - xm) * (o->control_x - xm) +
                              (o->control_y - ym) * (o->control_y - ym));
        if (s1 <= o->control_dist) {
            if (o->ap_hi < o->ap_lo) o->ap_lo = i;
            o->ap_hi = i;
        }
    }
}

// Source value at time t for row-0 cell i (the Focused_P branch of
// MixtureBoundaryPressureKernel, bubble_CUDA_kernel.cuh:903-925, with E4).
double focused_source(const oracle *o, int i, double t, double amp_scalar, double ym,
                      bool elements) {
    double xm = (i - 0.5) * o->dx;
    double s1 = std::sqrt((o->control_x - xm) * (o->control_x - xm) +
                          (o->control_y - ym) * (o->control_y - ym));
    const brp_config &c = o->cfg;
    if (s1 <= o->control_dist) {
        s1 = std::max(t - (o->control_dist - s1) / o->cs_inf, 0.0);
        double s2 = std::fmod(s1, ((double)c.tr_on_wave + (double)c.tr_off_wave) / c.tr_freq);
        if (s2 < (double)c.tr_on_wave / c.tr_freq) {
            if (elements && o->n_elem > 0 && o->ap_hi >= o->ap_lo) {
                long n_ap = o->ap_hi - o->ap_lo + 1;
                int e = (int)((long)(i - o->ap_lo) * o->n_elem / n_ap);
                return o->el_amp[e] * std::sin(o->omega * s2 + o->el_phase[e]);
            }
            return amp_scalar * std::sin(o->omega * s2);
        }
        return 0.0;
    }
    return 0.0;
}

// ---- kernels -----------------------------------------------------------

// smooth_delta_x / smooth_delta_y (bubble_CUDA_kernel.cuh:214-222)
inline double delta_x(const oracle *o, int pos, double b) {
    return o->dcoef_x * (1.0 + std::cos(o->dcoef_y * (((double)pos - 0.5) * o->dx - b)));
}
inline double delta_y(const oracle *o, int pos, double b) {
    return o->dcoef_x * (1.0 + std::cos(o->dcoef_z * (((double)pos - 0.5) * o->dy - b)));
}

// BubbleUpdateIndexKernel / the index part of BubbleMotionKernel
// (bubble_CUDA_kernel.cuh:355-356, 465-466).
inline void bin_bubble(oracle *o, int b) {
    o->ibmx[b] = (int)std::floor(o->bx[b] * o->rdx + 1.0);
    o->ibmy[b] = (int)std::floor(o->by[b] * o->rdy + 1.0);
    o->ibnx[b] = (int)std::floor(o->bx[b] * o->rdx + 0.5);
    o->ibny[b] = (int)std::floor(o->by[b] * o->rdy + 0.5);
}

// VoidFractionReverseLookupKernel (bubble_CUDA_kernel.cuh:576-619) after the
// f_g clear (kernel.cuh:1206), serial in bubble order; then the cylinder fold
// (VoidFractionCylinderKernel, kernel.cuh:559-573).
void calculate_void_fraction(oracle *o) {
    const ArrayIndex &a = o->a;
    std::fill(o->fg.v.begin(), o->fg.v.end(), 0.0);
    for (int b = 0; b < o->nb; b++) {
        int ibnx = o->ibnx[b], ibny = o->ibny[b];
        double px = o->bx[b], py = o->by[b];
        double R = o->R_t[b];
        double fg_temp;
        if (o->cfg.tr_cylindrical)
            fg_temp = kPi4r3 * o->n_B[b] * (R * R * R) * o->rdx * o->rdy / px;
        else
            fg_temp = kPi4r3 * o->n_B[b] * (R * R * R) * o->rdx * o->rdy;
        for (int i = ibnx + o->mbs; i <= ibnx + o->mbe; i++) {
            double Dx = delta_x(o, i, px);
            for (int j = ibny + o->mbs; j <= ibny + o->mbe; j++) {
                double Dy = delta_y(o, j, py);
                if (i >= a.ista2m && i <= a.iend2m && j >= a.jsta2m && j <= a.jend2m)
                    o->fg(i, j) += fg_temp * Dx * Dy;
            }
        }
    }
    if (o->cfg.tr_cylindrical) {
        for (int j = a.jsta2m; j <= a.jend2m; j++) {
            for (int count = 1 + a.ns + a.ms; count <= 0; count++) {
                o->fg(count, j) += o->fg(1 - count, j);
                o->fg(1 - count, j) = o->fg(count, j);
            }
        }
    }
}

// VelocityKernel (bubble_CUDA_kernel.cuh:635-677)
void velocity_kernel(oracle *o) {
    const ArrayIndex &a = o->a;
    const double dtm = o->dt;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int jn = a.jsta2n; jn <= a.jend2n; jn++) {
        for (int in = a.ista2n; in <= a.iend2n; in++) {
            bool ok1 = (in >= a.istan) && (in <= a.iendn) && (jn >= a.jstam) && (jn <= a.jendm);
            bool ok2 = (in >= a.istam) && (in <= a.iendm) && (jn >= a.jstan) && (jn <= a.jendn);
            double s1, s2, s3;
            if (ok1) {
                s1 = (o->rho_m(in, jn) + o->rho_m(in + 1, jn)) * 0.5;
                s2 = (-o->p0(in, jn) + o->p0(in + 1, jn)) * o->rdx;
                s3 = (o->c_sl(in, jn) + o->c_sl(in + 1, jn)) * 0.5;
                s3 = s3 * 0.5 * dtm * o->sig_nx[in - a.ista2n];
                o->vx(in, jn) = (o->vx(in, jn) * (1.0 - s3) - dtm / s1 * s2) / (1.0 + s3);
            }
            if (ok2) {
                s1 = (o->rho_m(in, jn) + o->rho_m(in, jn + 1)) * 0.5;
                s2 = (-o->p0(in, jn) + o->p0(in, jn + 1)) * o->rdx;  // rdx, as the reference (kernel.cuh:671)
                s3 = (o->c_sl(in, jn) + o->c_sl(in, jn + 1)) * 0.5;
                s3 = s3 * 0.5 * dtm * o->sig_ny[jn - a.jsta2n];
                o->vy(in, jn) = (o->vy(in, jn) * (1.0 - s3) - dtm / s1 * s2) / (1.0 + s3);
            }
        }
    }
}

// VelocityBoundaryKernel (bubble_CUDA_kernel.cuh:680-813), parts in program order.
void velocity_boundary(oracle *o) {
    const ArrayIndex &a = o->a;
    const brp_config &c = o->cfg;
    const int X = c.grid_x, Y = c.grid_y;
    const double t = o->tstep;
    // VX part a: rows j in [jstam, jendm]
    for (int j = a.jstam; j <= a.jendm; j++) {
        if (!c.pml_x0) {
            for (int i = a.ms + a.ns; i <= -1; i++) o->vx(i, j) = o->vx(-i, j);
            o->vx(0, j) = 0.0;
        }
        if (!c.pml_x1) {
            for (int i = X + 1; i <= X + a.me + a.ne; i++) o->vx(i, j) = -o->vx(X - (i - X), j);
            o->vx(X, j) = 0.0;
        }
    }
    // VX part b: columns i in [ista2n, iend2n]
    for (int i = a.ista2n; i <= a.iend2n; i++) {
        for (int j = 1 + a.ms + a.ns; j <= 0; j++) o->vx(i, j) = o->vx(i, 1 - j);
        for (int j = Y + 1; j <= Y + a.me + a.ne; j++) o->vx(i, j) = o->vx(i, Y - (j - Y - 1));
    }
    // VY part a: columns i in [istam, iendm]
    const bool src_row = !c.pml_y0 || c.ext_source_in_pml;
    for (int i = a.istam; i <= a.iendm; i++) {
        if (!c.pml_y0) {
            if (!c.tr_plane_p && !c.tr_focused_p) o->vy(i, 0) = 0.0;
            for (int j = a.ns + a.ms; j <= -1; j++) o->vy(i, j) = o->vy(i, -(j + 1));
        }
        if (src_row) {
            if (c.tr_plane_v) {
                if (std::fmod(t, ((double)c.tr_on_wave + (double)c.tr_off_wave) / c.tr_freq) <=
                    ((double)c.tr_on_wave) / c.tr_freq)
                    o->vy(i, 0) = c.tr_amp * std::sin(o->omega * t);
                else
                    o->vy(i, 0) = 0.0;
            }
            if (c.tr_focused_v) o->vy(i, 0) = focused_source(o, i, t, c.tr_amp, 0.0, false);
        }
        if (!c.pml_y1) {
            o->vy(i, Y) = 0.0;
            for (int j = Y + 1; j <= Y + a.ne + a.me; j++) o->vy(i, j) = o->vy(i, Y - (j - Y));
        }
    }
    // VY part b: rows j in [jsta2n, jend2n]
    for (int j = a.jsta2n; j <= a.jend2n; j++) {
        for (int i = 1 + a.ms + a.ns; i <= 0; i++) o->vy(i, j) = o->vy(1 - i, j);
        // The reference's source row index is the loop variable i, so every
        // row copies vy(X - (i - X - 1), jsta2n + i) = vy(X, Y) for order 2.
        for (int i = X + 1; i <= X + a.me + a.ne; i++) {
            int src_row_j = a.jsta2n + i;
            int src_col_i = X - (i - X - 1);
            o->vy(i, j) = o->vy.in(src_col_i, src_row_j) ? o->vy(src_col_i, src_row_j) : 0.0;
        }
    }
}

// VelocityBoundaryKernel's right-ghost read (kernel.cuh:808) reaches row jsta2n + X + 1, which is inside
// the vy array only while X + 1 <= jend2n - jsta2n, i.e. X <= Y + 1; beyond
// that the reference reads out of bounds and this restatement reads 0.

// BubbleInterpolationVelocityKernel + v_B <- v_L + BubbleMotionKernel
// (bubble_CUDA_kernel.cuh:398-468, 1282-1297).
void bubble_motion(oracle *o) {
    const ArrayIndex &a = o->a;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int b = 0; b < o->nb; b++) {
        double vLx = 0.0, vLy = 0.0;
        double px = o->bx[b], py = o->by[b];
        for (int i = o->ibmx[b] + o->nbs; i <= o->ibmx[b] + o->nbe; i++) {
            double Dx = delta_x(o, i, px);
            for (int j = o->ibny[b] + o->mbs; j <= o->ibny[b] + o->mbe; j++) {
                double Dy = delta_y(o, j, py);
                if (i >= a.ista2n && i <= a.iend2n && j >= a.jsta2m && j <= a.jend2m)
                    vLx += o->vx(i, j) * Dx * Dy;
            }
        }
        for (int i = o->ibnx[b] + o->mbs; i <= o->ibnx[b] + o->mbe; i++) {
            double Dx = delta_x(o, i, px);
            for (int j = o->ibmy[b] + o->nbs; j <= o->ibmy[b] + o->nbe; j++) {
                double Dy = delta_y(o, j, py);
                if (i >= a.ista2m && i <= a.iend2m && j >= a.jsta2n && j <= a.jend2n)
                    vLy += o->vy(i, j) * Dx * Dy;
            }
        }
        o->vLx[b] = vLx;
        o->vLy[b] = vLy;
        o->vBx[b] = vLx;
        o->vBy[b] = vLy;
        o->bx[b] = o->bx[b] + o->vBx[b] * o->dt;
        o->by[b] = o->by[b] + o->vBy[b] * o->dt;
        bin_bubble(o, b);
    }
}

// MixturePressureKernel (bubble_CUDA_kernel.cuh:816-867) after the Work
// clear (kernel.cuh:1310).
void pressure_kernel(oracle *o) {
    const ArrayIndex &a = o->a;
    std::fill(o->Work.v.begin(), o->Work.v.end(), 0.0);
    const double dtm = o->dt;
    const bool cyl = o->cfg.tr_cylindrical;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int j = a.jstam; j <= a.jendm; j++) {
        for (int i = a.istam; i <= a.iendm; i++) {
            double s1x, s1y, s2x, s2y, s3, s4, s5;
            if (cyl) {
                s1x = (-o->xu[i - 1 - a.ista2n] * o->vx(i - 1, j) + o->xu[i - a.ista2n] * o->vx(i, j)) *
                      o->rdx * o->rxp[i - a.ista2m];
            } else {
                s1x = (-o->vx(i - 1, j) + o->vx(i, j)) * o->rdx;
            }
            s1y = (-o->vy(i, j - 1) + o->vy(i, j)) * o->rdy;
            s3 = dtm * o->rho_l(i, j) * o->c_sl(i, j) * o->c_sl(i, j);
            s4 = 0.5 * dtm * o->c_sl(i, j);
            s5 = -(o->fg(i, j) - o->fgn(i, j)) / dtm / (1.0 - 0.5 * (o->fg(i, j) + o->fgn(i, j)));
            s2x = s4 * o->sig_mx[i - a.ista1m];
            s2y = s4 * o->sig_my[j - a.jsta1m];
            double ppx = o->pnx(i, j) * (1.0 - s2x) - s3 * (s1x + 0.5 * s5) / (1.0 + s2x);
            double ppy = o->pny(i, j) * (1.0 - s2y) - s3 * (s1y + 0.5 * s5) / (1.0 + s2y);
            o->px(i, j) = ppx;
            o->py(i, j) = ppy;
            o->Work(i, j) = std::fabs(o->p0(i, j) - (ppx + ppy));
            o->p0(i, j) = ppx + ppy;
        }
    }
}

// MixtureBoundaryPressureKernel (bubble_CUDA_kernel.cuh:870-952), parts in
// program order; E1 lets the source row be written with pml.y0 = 1.
void pressure_boundary(oracle *o) {
    const ArrayIndex &a = o->a;
    const brp_config &c = o->cfg;
    const int X = c.grid_x, Y = c.grid_y;
    const double t = o->tstep;
    const double pamp = o->rho_inf * o->cs_inf * c.tr_amp;
    const bool src_row = !c.pml_y0 || c.ext_source_in_pml;
    for (int i = a.istam; i <= a.iendm; i++) {
        if (!c.pml_y0) {
            for (int j = a.ms; j <= 0; j++) o->p0(i, j) = o->p0(i, 1 - j);
        }
        if (src_row) {
            if (c.tr_plane_p) {
                if (std::fmod(t, ((double)c.tr_on_wave + (double)c.tr_off_wave) / c.tr_freq) <=
                    ((double)c.tr_on_wave) / c.tr_freq)
                    o->p0(i, 0) = pamp * std::sin(o->omega * t);
                else
                    o->p0(i, 0) = 0.0;
            }
            if (c.tr_focused_p) o->p0(i, 0) = focused_source(o, i, t, pamp, -0.5 * o->dy, true);
        }
        if (!c.pml_y1) {
            // As the reference (kernel.cuh:929-932): rows X..X+me copy row 2Y+1-j (rows outside the
            // array are skipped).
            for (int j = X; j <= X + a.me; j++) {
                int src = Y - (j - Y - 1 - a.jsta1m);
                if (o->p0.in(i, j) && o->p0.in(i, src)) o->p0(i, j) = o->p0(i, src);
            }
        }
    }
    for (int j = a.jsta1m; j <= a.jend1m; j++) {
        if (!c.pml_x0) {
            for (int i = a.ms; i <= 0; i++) o->p0(i, j) = o->p0(1 - i, j);
        }
        if (!c.pml_x1) {
            for (int i = X + 1; i <= X + a.me; i++) o->p0(i, j) = o->p0(X - (i - X - 1), j);
        }
    }
}

// calculate_pressure_field (bubble_CUDA_kernel.cuh:1299-1333): returns resimax.
double calculate_pressure_field(oracle *o) {
    pressure_kernel(o);
    pressure_boundary(o);
    double resimax = 0;
    for (double w : o->Work.v)
        if (resimax < w) resimax = w;
    return resimax / o->P_inf;
}

// BubbleInterpolationScalarKernel (bubble_CUDA_kernel.cuh:361-395)
void interpolate_bubble_pressure(oracle *o) {
    const ArrayIndex &a = o->a;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int b = 0; b < o->nb; b++) {
        double px = o->bx[b], py = o->by[b];
        double PL = 0.0;
        for (int i = o->ibnx[b] + o->mbs; i <= o->ibnx[b] + o->mbe; i++) {
            double Dx = delta_x(o, i, px);
            for (int j = o->ibny[b] + o->mbs; j <= o->ibny[b] + o->mbe; j++) {
                double Dy = delta_y(o, j, py);
                if (i >= a.ista1m && i <= a.iend1m && j >= a.jsta1m && j <= a.jend1m)
                    PL += o->p0(i, j) * Dx * Dy;
            }
        }
        o->PL_p[b] = PL;
    }
}

// solve_bubble_radii (bubble_CUDA_kernel.cuh:1396-1411) -> max sub-steps.
int solve_bubble_radii(oracle *o, double *mean_sub, bool *capped) {
    int maxc = 0;
    int64_t sum = 0;
    bool cap_any = false;
    const int cap = o->cfg.max_substeps > 0 ? o->cfg.max_substeps : 1000000;
    std::vector<double> *sched = nullptr;
    if (o->sched_mode) {
        const size_t base = o->sched_call * (size_t)o->nb;
        if (o->sched_mode == 1) o->sched_dt.resize(base + o->nb);
        if (base + o->nb > o->sched_dt.size()) {
            *mean_sub = 0.0;
            *capped = true;
            return 0;
        }
        sched = o->sched_dt.data() + base;
        o->sched_call++;
    }
#pragma omp parallel for num_threads(o->nthreads) schedule(dynamic, 256) reduction(max : maxc) reduction(+ : sum) reduction(|| : cap_any)
    for (int b = 0; b < o->nb; b++) {
        RPIn in{o->R_pn[b], o->R_nn[b], o->d1_R_n[b], o->PG_n[b], o->bdt_n[b],
                o->re_n[b], o->PL_m[b], o->PL_n[b], o->PL_p[b]};
        RPOut out;
        bool cp = false;
        int cnt = rp_integrate(o->bp[b], o->dt, in, &out, cap, &cp, sched ? sched + b : nullptr,
                               o->sched_mode == 2);
        o->R_t[b] = out.R_t;
        o->R_p[b] = out.R_p;
        o->R_n[b] = out.R_n;
        o->d1_R_p[b] = out.d1_R_p;
        o->PG_p[b] = out.PG_p;
        o->bdt[b] = out.dt;
        o->re[b] = out.re;
        o->Q_B[b] = out.Q_B;
        o->sub[b] = cnt;
        maxc = std::max(maxc, cnt);
        sum += cnt;
        cap_any = cap_any || cp;
    }
    *mean_sub = o->nb ? (double)sum / o->nb : 0.0;
    *capped = cap_any;
    return maxc;
}

// store_variables (bubble_CUDA_kernel.cuh:1229-1262)
void store_variables(oracle *o) {
    o->R_pn = o->R_p;
    o->R_nn = o->R_n;
    o->d1_R_n = o->d1_R_p;
    o->PG_n = o->PG_p;
    o->PL_m = o->PL_n;
    o->PL_n = o->PL_p;
    o->bdt_n = o->bdt;
    o->re_n = o->re;
    // VFPredictionKernel (kernel.cuh:622-633) into Work, then the copies.
    for (size_t k = 0; k < o->Work.v.size(); k++)
        o->Work.v[k] = 3.0 * o->fg.v[k] - 3.0 * o->fgn.v[k] + o->fgm.v[k];
    o->pnx.v = o->px.v;
    o->pny.v = o->py.v;
    o->fgm.v = o->fgn.v;
    o->fgn.v = o->fg.v;
    o->fg.v = o->Work.v;
}

// ---- temperature (calculate_temperature, bubble_CUDA_kernel.cuh:1347-1380) --
// MixtureKMKernel (bubble_CUDA_kernel.cuh:955-969): k_m on the 2m range.
void mixture_km(oracle *o) {
    const ArrayIndex &a = o->a;
    for (int j = a.jsta2m; j <= a.jend2m; j++)
        for (int i = a.ista2m; i <= a.iend2m; i++)
            o->k_m(i, j) = thermal_conductivity_water(o->T(i, j) + o->T_inf) *
                           (1.0 - 3.0 * o->fg(i, j) / (2.0 + o->fg(i, j)));
}

// BubbleHeatKernel (bubble_CUDA_kernel.cuh:972-1006): Q_B of every bubble
// spread over its B x B stencil into Work, in bubble order.
void bubble_heat(oracle *o) {
    const ArrayIndex &a = o->a;
    for (int b = 0; b < o->nb; b++) {
        double Q_temp;
        if (o->cfg.tr_cylindrical)
            Q_temp = o->n_B[b] * o->Q_B[b] * o->rdx * o->rdy / o->bx[b];
        else
            Q_temp = o->n_B[b] * o->Q_B[b] * o->rdx * o->rdy;
        for (int i = o->ibnx[b] + o->mbs; i <= o->ibnx[b] + o->mbe; i++) {
            const double Dx = delta_x(o, i, o->bx[b]);
            for (int j = o->ibny[b] + o->mbs; j <= o->ibny[b] + o->mbe; j++) {
                const double Dy = delta_y(o, j, o->by[b]);
                if (i >= a.ista2m && i <= a.iend2m && j >= a.jsta2m && j <= a.jend2m)
                    o->Work(i, j) += Q_temp * Dx * Dy;
            }
        }
    }
}

// MixtureEnergyKernel (bubble_CUDA_kernel.cuh:1009-1033): conductive fluxes
// with the harmonic mean of k_m on the faces.
void mixture_energy(oracle *o) {
    const ArrayIndex &a = o->a;
    for (int j = a.jsta1m; j <= a.jend1m; j++)
        for (int i = a.ista1n; i <= a.iend1n; i++) {
            double s0 = 2.0 * o->k_m(i, j) * o->k_m(i + 1, j) / (o->k_m(i, j) + o->k_m(i + 1, j));
            double s1 = (-o->T(i, j) + o->T(i + 1, j)) * o->rdx;
            o->Ex(i, j) = s0 * s1;
        }
    for (int j = a.jsta1n; j <= a.jend1n; j++)
        for (int i = a.ista1m; i <= a.iend1m; i++) {
            double s0 = 2.0 * o->k_m(i, j) * o->k_m(i, j + 1) / (o->k_m(i, j) + o->k_m(i, j + 1));
            double s1 = (-o->T(i, j) + o->T(i, j + 1)) * o->rdy;
            o->Ey(i, j) = s0 * s1;
        }
}

// MixtureTemperatureKernel (bubble_CUDA_kernel.cuh:1036-1061): absorption of
// the acoustic field, bubble heat and conduction, explicit in time; rho_m
// and C_pm are the previous step's properties.
void mixture_temperature(oracle *o) {
    const ArrayIndex &a = o->a;
    for (int j = a.jstam; j <= a.jendm; j++)
        for (int i = a.istam; i <= a.iendm; i++) {
            double s0 = (o->px(i, j) + o->py(i, j) - o->pnx(i, j) - o->pny(i, j)) / o->dt;
            s0 = 2.0 * (0.22) * 0.115129255 / (o->rho_inf * o->cs_inf * o->omega * o->omega) * s0 * s0;
            double Q = o->Work(i, j) + s0;
            double s1 = (o->Ex(i, j) - o->Ex(i - 1, j)) * o->rdx;
            double s2 = (o->Ey(i, j) - o->Ey(i, j - 1)) * o->rdy;
            o->Work(i, j) = o->dt * (s1 + s2 + Q) / (o->rho_m(i, j) * o->C_pm(i, j));
            o->T(i, j) += o->Work(i, j);
        }
}

// MixtureBoundaryTemperatureKernel (bubble_CUDA_kernel.cuh:1064-1102), its
// four parts in program order (the reference runs them in one grid whose
// blocks race on the corner cells).  The y1 part starts at row Y, so row Y
// is overwritten with the ghost row Y + 1, as in the reference.
void mixture_boundary_temperature(oracle *o) {
    const ArrayIndex &a = o->a;
    const int X = o->cfg.grid_x, Y = o->cfg.grid_y;
    if (!o->cfg.pml_x0)
        for (int j = a.jstam; j <= a.jendm; j++)
            for (int i = 1 + a.ns + a.ms; i <= 0; i++) o->T(i, j) = o->T(1 - i, j);
    if (!o->cfg.pml_x1)
        for (int j = a.jsta1m; j <= a.jend1m; j++)
            for (int i = X + 1; i <= X + a.ne + a.me; i++) o->T(i, j) = o->T(X - (i - X - 1), j);
    if (!o->cfg.pml_y0)
        for (int i = a.ista2m; i <= a.iend2m; i++)
            for (int j = 1 + a.ns + a.ms; j <= 0; j++) o->T(i, j) = o->T(i, 1 - j);
    if (!o->cfg.pml_y1)
        for (int i = a.ista2m; i <= a.iend2m; i++)
            for (int j = Y; j <= Y + a.ne + a.me; j++) o->T(i, j) = o->T(i, Y - (j - Y - 1));
}

void calculate_temperature(oracle *o) {
    std::fill(o->Work.v.begin(), o->Work.v.end(), 0.0);
    if (o->cfg.bub_enable) bubble_heat(o);
    mixture_km(o);
    mixture_energy(o);
    mixture_temperature(o);
    mixture_boundary_temperature(o);
}

// MixturePropertiesKernel (bubble_CUDA_kernel.cuh:1110-1125)
void calculate_properties(oracle *o) {
    const ArrayIndex &a = o->a;
    for (int j = a.jsta1m; j <= a.jend1m; j++)
        for (int i = a.ista1m; i <= a.iend1m; i++) {
            o->rho_l(i, j) = o->rho_inf;
            o->c_sl(i, j) = o->cs_inf;
            o->rho_m(i, j) = (1.0 - o->fg(i, j)) * o->rho_inf;
            o->C_pm(i, j) = specific_heat_water(o->T(i, j) + o->T_inf);
        }
}

bool finite_fields(const oracle *o) {
    for (double v : o->p0.v)
        if (!std::isfinite(v)) return false;
    for (int b = 0; b < o->nb; b++)
        if (!std::isfinite(o->R_t[b])) return false;
    return true;
}

Grid *field_ptr(oracle *o, int id) {
    switch (id) {
        case BRP_FIELD_P0: return &o->p0;
        case BRP_FIELD_PX: return &o->px;
        case BRP_FIELD_PY: return &o->py;
        case BRP_FIELD_VX: return &o->vx;
        case BRP_FIELD_VY: return &o->vy;
        case BRP_FIELD_FG: return &o->fg;
        case BRP_FIELD_FGN: return &o->fgn;
        case BRP_FIELD_FGM: return &o->fgm;
        case BRP_FIELD_RHO_M: return &o->rho_m;
        case BRP_FIELD_T: return &o->T;
        default: return nullptr;
    }
}

}  // namespace

extern "C" {

oracle *oracle_create(const brp_config *cfg, char *err, int errlen) {
    auto set_err = [&](const char *m) {
        if (err && errlen > 0) std::snprintf(err, errlen, "%s", m);
    };
    if (!cfg) {
        set_err("null config");
        return nullptr;
    }
    const brp_config &c = *cfg;
    if (c.grid_x < 4 || c.grid_y < 4 || !(c.grid_lx > 0) || !(c.grid_ly > 0)) {
        set_err("grid.x, grid.y >= 4 and grid.lx, grid.ly > 0 required");
        return nullptr;
    }
    if (c.sim_order != 2) {
        set_err("simulation.order must be 2");
        return nullptr;
    }
    if (c.sim_smooth_delta_band < 1 || c.sim_smooth_delta_band > 8) {
        set_err("simulation.smooth_delta_band must be in [1, 8]");
        return nullptr;
    }
    if ((c.pml_x0 || c.pml_x1 || c.pml_y0 || c.pml_y1) &&
        (c.pml_grid_number < 1 || !(c.pml_sigma > 0.0))) {
        set_err("pml.grid_number >= 1 and pml.sigma > 0 required with PML");
        return nullptr;
    }
    if (!(c.sim_cfl > 0)) {
        set_err("simulation.cfl must be > 0");
        return nullptr;
    }
    oracle *o = new oracle();
    o->cfg = c;
    // init_grid_size (bubble_CUDA.cu:524-540)
    o->dx = (double)c.grid_lx / (double)c.grid_x;
    o->dy = (double)c.grid_ly / (double)c.grid_y;
    o->rdx = (double)1.0 / (double)o->dx;
    o->rdy = (double)1.0 / (double)o->dy;
    // init_transducer (bubble_CUDA.cu:543-558) with E2
    if (c.tr_focal_dist) {
        o->fp_x = 0.0;
        o->fp_y = c.tr_focal_dist * 0.5 * std::sqrt(3.0);
    } else {
        o->fp_x = 0.0;
        o->fp_y = c.grid_ly * 0.5;
    }
    if (c.ext_focus_x_set) o->fp_x = c.ext_focus_x;
    o->omega = 2.0 * std::acos(-1.0) * c.tr_freq;
    // control point and distance (bubble_CUDA.cu:132-153; PID passes not modelled)
    if (c.tr_pid && c.tr_init_control) {
        o->control_x = c.ext_focus_x_set ? o->fp_x : 0.0;
        o->control_y = c.tr_init_control;
    } else {
        o->control_x = o->fp_x;
        o->control_y = o->fp_y;
    }
    o->control_dist = o->control_y / 0.5 / std::sqrt(3.0);
    o->a = init_array(c.grid_x, c.grid_y, c.sim_order);
    init_sigma(o);
    // init_grid_vector (bubble_CUDA.cu:690-713)
    const ArrayIndex &a = o->a;
    o->rxp.assign(a.iend2m - a.ista2m + 1, 0.0);
    o->xu.assign(a.iend2n - a.ista2n + 1, 0.0);
    for (int i = a.ista2m; i <= a.iend2m; i++) o->rxp[i - a.ista2m] = 1.0 / (o->dx * ((double)i - 0.5));
    for (int i = a.ista2n; i <= a.iend2n; i++) o->xu[i - a.ista2n] = ((double)i) * o->dx;
    // init_mix, mix_set_time_increment (bubble_CUDA.cu:636-658)
    o->T_inf = 293.15;
    o->P_inf = 0.1e6;
    o->rho_inf = density_water(o->P_inf, o->T_inf);
    o->cs_inf = adiabatic_sound_speed_water(o->P_inf, o->T_inf);
    o->dt = c.sim_cfl * std::min(o->dx, o->dy) / o->cs_inf;
    // init_bub_params' shared part (bubble_CUDA.cu:661-675)
    o->dt0_rp = 0.1 * o->dt;
    o->mbs = -c.sim_smooth_delta_band / 2 + 1;
    o->mbe = c.sim_smooth_delta_band + o->mbs - 1;
    o->nbs = -c.sim_smooth_delta_band / 2;
    o->nbe = c.sim_smooth_delta_band + o->nbs - 1;
    // delta_coef (bubble_CUDA.cu:1467-1471)
    o->dcoef_x = 1.0 / ((double)c.sim_smooth_delta_band);
    o->dcoef_y = 2.0 * std::acos(-1.0) / ((double)c.sim_smooth_delta_band) * o->rdx;
    o->dcoef_z = 2.0 * std::acos(-1.0) / ((double)c.sim_smooth_delta_band) * o->rdy;
    // init_mix_array (bubble_CUDA.cu:716-794)
    o->p0.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, 0.0);
    o->px.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, 0.0);
    o->py.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, 0.0);
    o->pnx.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, 0.0);
    o->pny.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, 0.0);
    o->rho_m.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, o->rho_inf);
    o->rho_l.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, o->rho_inf);
    o->c_sl.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, o->cs_inf);
    o->C_pm.init(a.ista1m, a.iend1m, a.jsta1m, a.jend1m, specific_heat_water(o->T_inf));
    o->k_m.init(a.ista2m, a.iend2m, a.jsta2m, a.jend2m, thermal_conductivity_water(o->T_inf));
    o->vx.init(a.ista2n, a.iend2n, a.jsta2m, a.jend2m, 0.0);
    o->vy.init(a.ista2m, a.iend2m, a.jsta2n, a.jend2n, 0.0);
    o->Ex.init(a.ista1n, a.iend1n, a.jsta1m, a.jend1m, 0.0);
    o->Ey.init(a.ista1m, a.iend1m, a.jsta1n, a.jend1n, 0.0);
    o->T.init(a.ista2m, a.iend2m, a.jsta2m, a.jend2m, 0.0);
    o->fg.init(a.ista2m, a.iend2m, a.jsta2m, a.jend2m, 0.0);
    o->fgn.init(a.ista2m, a.iend2m, a.jsta2m, a.jend2m, 0.0);
    o->fgm.init(a.ista2m, a.iend2m, a.jsta2m, a.jend2m, 0.0);
    o->Work.init(a.ista2m, a.iend2m, a.jsta2m, a.jend2m, 0.0);
    // init_bub_array (bubble_CUDA.cu:797-897): one bubble per cell centre
    // inside the box (or everywhere when box_size == 0), i outer, j inner.
    if (c.bub_enable) {
        std::vector<double> xs, ys;
        for (int i = a.istam; i <= a.iendm; i++) {
            double px = ((double)i - 0.5) * o->dx;
            for (int j = a.jstam; j <= a.jendm; j++) {
                double py = ((double)j - 0.5) * o->dy;
                if (c.tr_box_size ? (std::fabs(px - o->fp_x) < 0.5 * c.tr_box_size &&
                                     std::fabs(py - o->fp_y) < 0.5 * c.tr_box_size)
                                  : true) {
                    xs.push_back(px);
                    ys.push_back(py);
                }
            }
        }
        resize_bubbles(o, (int)xs.size());
        for (int b = 0; b < o->nb; b++) bubble_input(o, b, xs[b], ys[b], c.bub_R0, c.bub_fg0, nullptr);
    }
    find_aperture(o);
    return o;
}

void oracle_destroy(oracle *o) { delete o; }

const char *oracle_last_error(const oracle *o) { return o ? o->err.c_str() : ""; }

void oracle_set_threads(oracle *o, int32_t n) {
    if (o) o->nthreads = n < 1 ? 1 : n;
}

int oracle_get_info(const oracle *o, brp_info *info) {
    if (!o || !info) return BRP_ERR_INVALID;
    std::memset(info, 0, sizeof *info);
    info->nx = o->cfg.grid_x;
    info->ny = o->cfg.grid_y;
    info->n_bubbles = o->nb;
    info->precision = BRP_PREC_FP64;
    info->dx = o->dx;
    info->dy = o->dy;
    info->rdx = o->rdx;
    info->rdy = o->rdy;
    info->dt = o->dt;
    info->dt0_rp = o->dt0_rp;
    info->fp_x = o->fp_x;
    info->fp_y = o->fp_y;
    info->omega = o->omega;
    info->time = o->tstep;
    info->nstep = o->nstep;
    return BRP_OK;
}

int oracle_set_bubbles(oracle *o, int32_t n, const double *x, const double *y, const double *R0,
                       const double *n_B, double fg_in) {
    if (!o || n < 0 || (n > 0 && (!x || !y || !R0))) return BRP_ERR_INVALID;
    if (o->prepared) return fail(o, BRP_ERR_STATE, "set_bubbles after prepare");
    if (!o->cfg.bub_enable) return fail(o, BRP_ERR_INVALID, "bubbles.enable is 0");
    for (int b = 0; b < n; b++)
        if (!(R0[b] > 0) || !std::isfinite(x[b]) || !std::isfinite(y[b]))
            return fail(o, BRP_ERR_INVALID, "bubble %s: R0 must be > 0, x and y finite", "input");
    resize_bubbles(o, n);
    for (int b = 0; b < n; b++) bubble_input(o, b, x[b], y[b], R0[b], fg_in, n_B ? &n_B[b] : nullptr);
    return BRP_OK;
}

int oracle_set_source(oracle *o, int32_t n_elem, const double *amp, const double *phase) {
    if (!o || n_elem < 0 || (n_elem > 0 && (!amp || !phase))) return BRP_ERR_INVALID;
    o->n_elem = n_elem;
    o->el_amp.assign(amp, amp + n_elem);
    o->el_phase.assign(phase, phase + n_elem);
    return BRP_OK;
}

// solve_bubbles' initialisation kernels (bubble_CUDA.cu:172-188)
int oracle_prepare(oracle *o) {
    if (!o) return BRP_ERR_INVALID;
    if (o->prepared) return BRP_OK;
    if (o->cfg.bub_enable) {
        for (int b = 0; b < o->nb; b++) bin_bubble(o, b);
        calculate_void_fraction(o);
    }
    o->fgn.v = o->fg.v;  // synchronize_void_fraction (kernel.cuh:1220-1227)
    o->fgm.v = o->fg.v;
    o->prepared = true;
    return BRP_OK;
}

// One iteration of solve_bubbles' main loop (bubble_CUDA.cu:205-286).
int oracle_step(oracle *o, int32_t k, brp_step_stats *stats) {
    if (!o || k < 0) return BRP_ERR_INVALID;
    if (!o->prepared) oracle_prepare(o);
    const double thr = (double)1.0e-7f;  // bubble_CUDA.cu:243
    const int max_pic = o->cfg.max_picard > 0 ? o->cfg.max_picard : 1000;
    for (int s = 0; s < k; s++) {
        brp_step_stats st{};
        o->nstep++;
        double s1 = o->tstep;
        o->tstep = o->tstep + o->dt + o->tstepx;
        double s2 = o->tstep - s1;
        o->tstepx = o->dt + o->tstepx - s2;
        store_variables(o);
        velocity_kernel(o);
        velocity_boundary(o);
        if (o->cfg.bub_enable) bubble_motion(o);
        double resimax = calculate_pressure_field(o);
        st.resimax0 = resimax;
        int status = BRP_OK;
        int maxsub = 0;
        double meansum = 0;
        if (o->cfg.bub_enable) {
            int loop = 0;
            int forced = -1;
            if (o->sched_mode == 2) {
                if (o->sched_step >= o->sched_npic.size()) {
                    o->err = "frozen schedule exhausted";
                    return BRP_ERR_STATE;
                }
                forced = o->sched_npic[o->sched_step++];
            }
            while (forced >= 0 ? loop < forced : resimax > thr) {
                if (loop >= max_pic) {
                    status = BRP_ERR_PICARD_CAP;
                    break;
                }
                loop++;
                interpolate_bubble_pressure(o);
                double mean_sub;
                bool capped;
                maxsub = std::max(maxsub, solve_bubble_radii(o, &mean_sub, &capped));
                meansum += mean_sub;
                if (capped) status = BRP_ERR_SUBSTEP_CAP;
                calculate_void_fraction(o);
                resimax = calculate_pressure_field(o);
                if (capped) break;
            }
            st.n_pic = loop;
            if (o->sched_mode == 1) o->sched_npic.push_back(loop);
        }
        if (o->cfg.thermal) {
            calculate_temperature(o);
            // determine_focal_point (bubble_CUDA_kernel.cuh:1134-1166) on Work:
            // its reduction functors are not in the reference tree; the
            // largest value, first in row-major order, at (i dx, j dy)
            const ArrayIndex &a = o->a;
            double best = -INFINITY;
            int bi = a.ista2m, bj = a.jsta2m;
            for (int j = a.jsta2m; j <= a.jend2m; j++)
                for (int i = a.ista2m; i <= a.iend2m; i++)
                    if (o->Work(i, j) > best) {
                        best = o->Work(i, j);
                        bi = i;
                        bj = j;
                    }
            st.focal_x = bi * o->dx;
            st.focal_y = bj * o->dy;
        }
        calculate_properties(o);
        if (status == BRP_OK && !finite_fields(o)) status = BRP_ERR_NONFINITE;
        st.nstep = o->nstep;
        st.time = o->tstep;
        st.resimax = resimax;
        st.rp_max_substeps = maxsub;
        st.rp_mean_substeps = st.n_pic ? meansum / st.n_pic : 0.0;
        st.status = status;
        if (stats) stats[s] = st;
        if (status != BRP_OK) {
            char buf[128];
            std::snprintf(buf, sizeof buf, "step %lld: status %d", (long long)o->nstep, status);
            o->err = buf;
            return status;
        }
    }
    return BRP_OK;
}

int oracle_set_focus(oracle *o, double fp_x, double fp_y, double f_dist) {
    if (!o) return BRP_ERR_INVALID;
    o->control_x = fp_x;
    o->control_y = fp_y;
    o->control_dist = f_dist;
    find_aperture(o);
    return BRP_OK;
}

int oracle_schedule_mode(oracle *o, int32_t mode) {
    if (!o || mode < 0 || mode > 2) return BRP_ERR_INVALID;
    o->sched_mode = mode;
    o->sched_step = 0;
    o->sched_call = 0;
    if (mode == 1) {
        o->sched_npic.clear();
        o->sched_dt.clear();
    }
    return BRP_OK;
}

int oracle_schedule_copy(oracle *dst, const oracle *src) {
    if (!dst || !src || dst->nb != src->nb) return BRP_ERR_INVALID;
    dst->sched_npic = src->sched_npic;
    dst->sched_dt = src->sched_dt;
    dst->sched_step = 0;
    dst->sched_call = 0;
    return BRP_OK;
}

int oracle_field_shape(const oracle *o, int32_t field, int32_t *nx, int32_t *ny, int32_t *i0,
                       int32_t *j0) {
    if (!o) return BRP_ERR_INVALID;
    Grid *g = field_ptr(const_cast<oracle *>(o), field);
    if (!g) return BRP_ERR_INVALID;
    if (nx) *nx = g->nx;
    if (ny) *ny = g->ny;
    if (i0) *i0 = g->i0;
    if (j0) *j0 = g->j0;
    return BRP_OK;
}

int oracle_get_field(const oracle *o, int32_t field, double *dst, int64_t ld) {
    if (!o || !dst) return BRP_ERR_INVALID;
    const Grid *g = field_ptr(const_cast<oracle *>(o), field);
    if (!g || ld < g->nx) return BRP_ERR_INVALID;
    for (int j = 0; j < g->ny; j++)
        std::memcpy(dst + (size_t)j * ld, &g->v[(size_t)j * g->nx], sizeof(double) * g->nx);
    return BRP_OK;
}

int oracle_set_field(oracle *o, int32_t field, const double *src, int64_t ld) {
    if (!o || !src) return BRP_ERR_INVALID;
    if (!o->prepared) oracle_prepare(o);
    Grid *g = field_ptr(o, field);
    if (!g || ld < g->nx) return BRP_ERR_INVALID;
    for (int j = 0; j < g->ny; j++)
        std::memcpy(&g->v[(size_t)j * g->nx], src + (size_t)j * ld, sizeof(double) * g->nx);
    if (field == BRP_FIELD_FG) {
        o->fgn.v = o->fg.v;
        o->fgm.v = o->fg.v;
    }
    return BRP_OK;
}

// Gather of the current p0 at every bubble (BubbleInterpolationScalarKernel)
// without advancing anything; PL_out receives n_bubbles values.
int oracle_gather_p0(oracle *o, double *PL_out) {
    if (!o || !PL_out) return BRP_ERR_INVALID;
    if (!o->prepared) oracle_prepare(o);
    std::vector<double> keep = o->PL_p;
    interpolate_bubble_pressure(o);
    std::memcpy(PL_out, o->PL_p.data(), sizeof(double) * o->nb);
    o->PL_p = keep;
    return BRP_OK;
}

// Scatter of per-bubble volumes q (as fg_temp) with the stencil of
// VoidFractionReverseLookupKernel into fg_out (2m layout), state untouched.
int oracle_scatter_values(oracle *o, const double *q, double *fg_out) {
    if (!o || !q || !fg_out) return BRP_ERR_INVALID;
    if (!o->prepared) oracle_prepare(o);
    const ArrayIndex &a = o->a;
    Grid g;
    g.init(a.ista2m, a.iend2m, a.jsta2m, a.jend2m, 0.0);
    for (int b = 0; b < o->nb; b++) {
        for (int i = o->ibnx[b] + o->mbs; i <= o->ibnx[b] + o->mbe; i++) {
            double Dx = delta_x(o, i, o->bx[b]);
            for (int j = o->ibny[b] + o->mbs; j <= o->ibny[b] + o->mbe; j++) {
                double Dy = delta_y(o, j, o->by[b]);
                if (i >= a.ista2m && i <= a.iend2m && j >= a.jsta2m && j <= a.jend2m)
                    g(i, j) += q[b] * Dx * Dy;
            }
        }
    }
    std::memcpy(fg_out, g.v.data(), sizeof(double) * g.v.size());
    return BRP_OK;
}

// One solveRayleighPlesset call.  io = {Rt, Rp, Rn, d1R, dt, remain} in/out.
void oracle_rp_radius(const oracle_bubble_params *p, double PG, double PL, double io[6]) {
    BubParams b{p->R0, p->R03, p->PL0, p->PG0, p->T0, p->gam, p->sig,
                p->mu, p->rho, p->K0,  p->coeff_alpha, p->dt0};
    solveRayleighPlesset(&io[0], &io[1], &io[2], &io[3], PG, PL, &io[4], &io[5], b);
}

int oracle_get_bubbles(const oracle *o, int32_t which, double *dst) {
    if (!o || !dst) return BRP_ERR_INVALID;
    const std::vector<double> *v = nullptr;
    const std::vector<int> *vi = nullptr;
    switch (which) {
        case BRP_BUB_X: v = &o->bx; break;
        case BRP_BUB_Y: v = &o->by; break;
        case BRP_BUB_R_T: v = &o->R_t; break;
        case BRP_BUB_R_P: v = &o->R_p; break;
        case BRP_BUB_R_N: v = &o->R_n; break;
        case BRP_BUB_D1R: v = &o->d1_R_p; break;
        case BRP_BUB_PG: v = &o->PG_p; break;
        case BRP_BUB_PL_P: v = &o->PL_p; break;
        case BRP_BUB_PL_N: v = &o->PL_n; break;
        case BRP_BUB_PL_M: v = &o->PL_m; break;
        case BRP_BUB_Q_B: v = &o->Q_B; break;
        case BRP_BUB_N_B: v = &o->n_B; break;
        case BRP_BUB_DT: v = &o->bdt; break;
        case BRP_BUB_RE: v = &o->re; break;
        case BRP_BUB_VX: v = &o->vBx; break;
        case BRP_BUB_VY: v = &o->vBy; break;
        case BRP_BUB_IBM_X: vi = &o->ibmx; break;
        case BRP_BUB_IBM_Y: vi = &o->ibmy; break;
        case BRP_BUB_IBN_X: vi = &o->ibnx; break;
        case BRP_BUB_IBN_Y: vi = &o->ibny; break;
        case BRP_BUB_SUB: vi = &o->sub; break;
        case BRP_BUB_R0:
            for (int b = 0; b < o->nb; b++) dst[b] = o->bp[b].R0;
            return BRP_OK;
        default: return BRP_ERR_INVALID;
    }
    if (v) std::memcpy(dst, v->data(), sizeof(double) * o->nb);
    if (vi)
        for (int b = 0; b < o->nb; b++) dst[b] = (*vi)[b];
    return BRP_OK;
}

// State injection (the counterpart of brp_set_bubble_array): one bubble
// array of the state a step starts from; positions re-bin the bubble.
int oracle_set_bubble_array(oracle *o, int32_t which, const double *src) {
    if (!o || !src) return BRP_ERR_INVALID;
    if (!o->prepared) oracle_prepare(o);
    std::vector<double> *v = nullptr;
    switch (which) {
        case BRP_BUB_X: v = &o->bx; break;
        case BRP_BUB_Y: v = &o->by; break;
        case BRP_BUB_R_T: v = &o->R_t; break;
        case BRP_BUB_R_P: v = &o->R_p; break;
        case BRP_BUB_R_N: v = &o->R_n; break;
        case BRP_BUB_D1R: v = &o->d1_R_p; break;
        case BRP_BUB_PG: v = &o->PG_p; break;
        case BRP_BUB_PL_P: v = &o->PL_p; break;
        case BRP_BUB_PL_N: v = &o->PL_n; break;
        case BRP_BUB_PL_M: v = &o->PL_m; break;
        case BRP_BUB_Q_B: v = &o->Q_B; break;
        case BRP_BUB_DT: v = &o->bdt; break;
        case BRP_BUB_RE: v = &o->re; break;
        case BRP_BUB_VX: v = &o->vBx; break;
        case BRP_BUB_VY: v = &o->vBy; break;
        default: return BRP_ERR_INVALID;
    }
    std::memcpy(v->data(), src, sizeof(double) * o->nb);
    if (which == BRP_BUB_X || which == BRP_BUB_Y)
        for (int b = 0; b < o->nb; b++) bin_bubble(o, b);
    return BRP_OK;
}

int oracle_get_clock(const oracle *o, int64_t *nstep, double *time, double *carry) {
    if (!o) return BRP_ERR_INVALID;
    if (nstep) *nstep = o->nstep;
    if (time) *time = o->tstep;
    if (carry) *carry = o->tstepx;
    return BRP_OK;
}

int oracle_set_clock(oracle *o, int64_t nstep, double time, double carry) {
    if (!o || nstep < 0) return BRP_ERR_INVALID;
    if (!o->prepared) oracle_prepare(o);
    o->nstep = nstep;
    o->tstep = time;
    o->tstepx = carry;
    return BRP_OK;
}

int oracle_array_index(const oracle *o, int32_t out[28]) {
    if (!o) return BRP_ERR_INVALID;
    const ArrayIndex &a = o->a;
    int32_t v[28] = {a.ms,     a.me,     a.ns,     a.ne,     a.istam,  a.iendm,  a.istan,
                     a.iendn,  a.ista1m, a.iend1m, a.ista1n, a.iend1n, a.ista2m, a.iend2m,
                     a.ista2n, a.iend2n, a.jstam,  a.jendm,  a.jstan,  a.jendn,  a.jsta1m,
                     a.jend1m, a.jsta1n, a.jend1n, a.jsta2m, a.jend2m, a.jsta2n, a.jend2n};
    std::memcpy(out, v, sizeof v);
    return BRP_OK;
}

int oracle_sigma(const oracle *o, int32_t which, double *dst, int32_t *n) {
    if (!o) return BRP_ERR_INVALID;
    const std::vector<double> *v = which == 0 ? &o->sig_mx : which == 1 ? &o->sig_my
                                 : which == 2 ? &o->sig_nx : which == 3 ? &o->sig_ny : nullptr;
    if (!v) return BRP_ERR_INVALID;
    if (n) *n = (int32_t)v->size();
    if (dst) std::memcpy(dst, v->data(), sizeof(double) * v->size());
    return BRP_OK;
}

int oracle_scalars(const oracle *o, double out[8]) {
    if (!o) return BRP_ERR_INVALID;
    out[0] = o->dcoef_x;
    out[1] = o->dcoef_y;
    out[2] = o->dcoef_z;
    out[3] = o->control_dist;
    out[4] = kPi4r3;
    out[5] = o->control_x;
    out[6] = o->control_y;
    out[7] = (double)(o->ap_hi - o->ap_lo + 1);
    return BRP_OK;
}

double oracle_smooth_delta(const oracle *o, int32_t axis, int32_t pos, double b) {
    return axis == 0 ? delta_x(o, pos, b) : delta_y(o, pos, b);
}

void oracle_bubble_params_init(oracle_bubble_params *out, double R0, double PL0, double T0,
                               double gam, double sig, double mu, double rho, double K0,
                               double dt_mix) {
    brp_config c{};
    c.bub_P0 = PL0;
    c.bub_T0 = T0;
    c.bub_gamma = gam;
    c.bub_sigma = sig;
    c.bub_mu = mu;
    c.bub_rho = rho;
    c.bub_K0 = K0;
    BubParams b = make_bub_params(c, R0, dt_mix);
    *out = oracle_bubble_params{b.R0, b.R03, b.PL0, b.PG0, b.T0, b.gam, b.sig,
                                b.mu, b.rho, b.K0,  b.coeff_alpha, b.dt0};
}

static BubParams to_bp(const oracle_bubble_params *p) {
    return BubParams{p->R0, p->R03, p->PL0, p->PG0, p->T0, p->gam, p->sig,
                     p->mu, p->rho, p->K0,  p->coeff_alpha, p->dt0};
}

void oracle_coth(double re, double im, double out[2]) {
    cplx r = coth(mk(re, im));
    out[0] = r.re;
    out[1] = r.im;
}

void oracle_upsilon(double re, double im, double gam, double out[2]) {
    cplx r = Upsilon(mk(re, im), gam);
    out[0] = r.re;
    out[1] = r.im;
}

void oracle_solve_lp(double re, double im, double R, double out[2]) {
    cplx r = solveLp(mk(re, im), R);
    out[0] = r.re;
    out[1] = r.im;
}

double oracle_solve_omega_n(double PG, double R, const oracle_bubble_params *p, double alpha_out[2]) {
    cplx al;
    double w = solveOmegaN(&al, PG, R, to_bp(p));
    if (alpha_out) {
        alpha_out[0] = al.re;
        alpha_out[1] = al.im;
    }
    return w;
}

static int64_t g_last_rp_ops = 0;

int oracle_rp_integrate(const oracle_bubble_params *p, double dt_mix, const double st[9],
                        double out[8], int32_t max_substeps) {
    RPIn in{st[0], st[1], st[2], st[3], st[4], st[5], st[6], st[7], st[8]};
    RPOut r;
    bool capped;
    g_ops = 0;
    int n = rp_integrate(to_bp(p), dt_mix, in, &r, max_substeps > 0 ? max_substeps : 1000000, &capped);
    g_last_rp_ops = g_ops;
    out[0] = r.R_t;
    out[1] = r.R_p;
    out[2] = r.R_n;
    out[3] = r.d1_R_p;
    out[4] = r.PG_p;
    out[5] = r.dt;
    out[6] = r.re;
    out[7] = r.Q_B;
    return n;
}

int64_t oracle_rp_flops_last(void) { return g_last_rp_ops; }

}  // extern "C"
