// rp_device.h -- FP64 reduced-order Rayleigh-Plesset model on one lane.
//
// Semantics follow the reference's device physics
// (inc/bubble_CUDA_kernel.cuh:145-340, inc/complex.cuh:1-90): the same
// expressions in the same operation order, so an FP64 lane reproduces the
// reference's rounding up to the last-ulp differences of tanh/tan/cos
// between math libraries.  Per-bubble constants (R0, R0^3, PG0, coeff_alpha)
// are arrays (extension E3); everything else is a BubbleParams scalar.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

namespace brp {

struct cd {
    double re, im;
};

__device__ inline cd c_mk(double a, double b) { return cd{a, b}; }
__device__ inline cd c_add(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ inline cd c_radd(double a, cd b) { return cd{a + b.re, b.im}; }
__device__ inline cd c_subr(cd a, double b) { return cd{a.re - b, a.im}; }
__device__ inline cd c_rsub(double a, cd b) { return cd{a - b.re, -b.im}; }
__device__ inline cd c_sub(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ inline cd c_mul(cd a, cd b) {
    return cd{a.re * b.re - a.im * b.im, a.im * b.re + a.re * b.im};
}
__device__ inline cd c_mulr(cd a, double b) { return cd{a.re * b, a.im * b}; }
__device__ inline cd c_rmul(double a, cd b) { return cd{a * b.re, a * b.im}; }
__device__ inline cd c_div(cd a, cd b) {
    double c = (b.re * b.re + b.im * b.im);
    return cd{(a.re * b.re + a.im * b.im) / c, (a.im * b.re - a.re * b.im) / c};
}
__device__ inline cd c_divr(cd a, double b) { return cd{a.re / b, a.im / b}; }
__device__ inline cd c_rdiv(double a, cd b) {
    double c = (b.re * b.re + b.im * b.im);
    return cd{a * b.re / c, -a * b.im / c};
}
__device__ inline double c_abs(cd a) { return sqrt(a.re * a.re + a.im * a.im); }
// coth(x + iy) = (1 + i tanh(x) tan(y)) / (tanh(x) + i tan(y))
__device__ inline cd c_coth(cd a) {
    double x = tanh(a.re), y = tan(a.im);
    return c_div(c_mk(1.0, x * y), c_mk(x, y));
}

struct RPConst {       // per-bubble constants
    double R0, R03, PG0, calpha;
};
struct RPShared {      // constants shared by all bubbles
    double PL0, T0, gam, sig, mu, rho, K0, dt0, pi;
};

constexpr double kDblEps = 2.22044604925031308e-016;  // absolute, as the reference

// One variable-step three-level radius update (solveRayleighPlesset).
__device__ inline void rp_radius(double *Rt, double *Rp, double *Rn, double *d1R, double PG,
                                 double PL, double *dt, double *remain, const RPShared &s) {
    double Rm = *Rn;
    *Rn = *Rp;
    double Rnd2R = (PG - PL - 2.0 * s.sig / (*Rn) - 4.0 * s.mu / (*Rn) * (*d1R)) / s.rho -
                   1.5 * (*d1R) * (*d1R);
    double dtm = *dt;
    double dtp = fmin(1.01 * (*dt), s.dt0);
    if (fabs(*d1R) > kDblEps) dtp = fmin(dtp, 0.01 * (*Rn) / fabs(*d1R));
    if (fabs(Rnd2R) > kDblEps) dtp = fmin(dtp, 0.1 * (*Rn) / sqrt(fabs(Rnd2R)));
    *Rt = *Rp = (1.0 + dtp / dtm) * (*Rn) - dtp / dtm * Rm +
                (dtp / (*Rn)) * (0.5 * (dtp + dtm) * Rnd2R);
    *d1R = (dtm * dtm * (*Rp) + (-dtm * dtm + dtp * dtp) * (*Rn) - dtp * dtp * Rm) /
               (dtm * dtp * (dtm + dtp)) +
           (dtp / (*Rn)) * Rnd2R;
    *dt = dtp;
    if (dtp >= *remain) {
        dtp = *remain;
        *Rt = (1.0 + dtp / dtm) * (*Rn) - dtp / dtm * Rm +
              (dtp / (*Rn)) * (0.5 * (dtp + dtm) * Rnd2R);
    }
    *remain -= dtp;
}

__device__ inline cd rp_alpha(double R, double omega, double calpha) {
    return c_mulr(c_mk(1.0, 1.0), sqrt(calpha / R * omega));
}

__device__ inline cd rp_upsilon(cd a, double gam) {
    if (c_abs(a) > 1.0e-1) {
        cd t = c_div(c_subr(c_mul(a, c_coth(a)), 1.0), c_mul(a, a));
        return c_rdiv(3.0 * gam, c_radd(1.0, c_rmul(3.0, c_rmul(gam - 1.0, t))));
    }
    cd aa = c_mul(a, a);
    cd u = c_rsub(2.0 / 945.0, c_divr(aa, 4725.0));
    u = c_radd(-1.0 / 45.0, c_mul(aa, u));
    u = c_radd(1.0 / 3.0, c_mul(aa, u));
    return c_rdiv(3.0 * gam, c_radd(1.0, c_rmul(3.0, c_rmul(gam - 1.0, u))));
}

__device__ inline cd rp_lp(cd a, double R) {
    cd ctmp;
    if (c_abs(a) > 1.0e-1) {
        ctmp = c_subr(c_mul(a, c_coth(a)), 1.0);
        ctmp = c_div(c_sub(c_mul(a, a), c_rmul(3.0, ctmp)), c_mul(c_mul(a, a), ctmp));
    } else {
        cd aa = c_mul(a, a);
        cd u = c_rsub(2.0 / 7875.0, c_divr(c_mulr(aa, 37.0), 3031875.0));
        u = c_radd(-1.0 / 175.0, c_mul(aa, u));
        ctmp = c_radd(1.0 / 5.0, c_mul(aa, u));
    }
    return c_rmul(R, ctmp);
}

__device__ inline double rp_omega(cd *alpha, double PG, double R, const RPConst &k,
                                  const RPShared &s) {
    double Rx = R / k.R0;
    double coef1 = k.PG0 / (PG * Rx * Rx * Rx);
    double coef2 = PG * s.rho;
    double coef3 = 4.0 / (R * R);
    double value1 = -2.0 * s.sig * s.rho / R;
    cd ups = c_mulr(c_mk(3.0 * s.gam, 0.0), coef1);
    double mu_eff = s.mu;
    double eta = ups.re * coef2 + value1 - mu_eff * mu_eff * coef3;
    double om = sqrt(fmax(eta, 1.0e-6 * kDblEps)) / (s.rho * R);
    for (int it = 0; it < 3; it++) {
        *alpha = rp_alpha(R, om, k.calpha);
        ups = c_mulr(rp_upsilon(*alpha, s.gam), coef1);
        mu_eff = s.mu + ups.im * PG / (4.0 * (om));
        eta = ups.re * coef2 + value1 - mu_eff * mu_eff * coef3;
        om = sqrt(fmax(eta, 1.0e-6 * kDblEps)) / (s.rho * R);
    }
    *alpha = rp_alpha(R, om, k.calpha);
    return om;
}

__device__ inline double rp_pg(double PGn, double Rp, double Rn, double om, double dt, cd Lp,
                               const RPConst &k, const RPShared &s) {
    double R = 0.5 * (Rp + Rn);
    double Rx = R / k.R0;
    double Rpx = Rp / k.R0;
    double Rnx = Rn / k.R0;
    double Lp_NR = Lp.re / (Lp.re * Lp.re + Lp.im * Lp.im);
    double Lp_NI = Lp.im / (Lp.re * Lp.re + Lp.im * Lp.im);
    double c0 = dt * 3.0 * (s.gam - 1.0) * s.K0 * s.T0 / (R * k.PG0);
    double c3 = 1.5 * s.gam * (Rp - Rn) / R + c0 * Lp_NR * 0.5 * Rx * Rx * Rx;
    double c1 = 1.0 + c3 - c0 * Lp_NI / (om * dt) * Rpx * Rpx * Rpx;
    double c2 = 1.0 - c3 - c0 * Lp_NI / (om * dt) * Rnx * Rnx * Rnx;
    return (c2 * PGn + c0 * Lp_NR * k.PG0) / c1;
}

struct RPState {
    double Rp, Rn, d1R, PG, dt, re;
};

// The integration state of one lane between sub-steps.
struct RPLane {
    double Rp, Rn, d1R, PG, dtL, remain, time, PC0, PC1, PC2, SumHeat, SumVis, Rt, re_n;
    int counter;
};

// Entry of BubbleRadiusKernel (bubble_CUDA_kernel.cuh:487-504): start state,
// remaining time and the quadratic liquid-pressure history through
// (PL_m, PL_n, PL_p).
__device__ inline void rp_begin(RPLane &L, const RPState &in, double PLm, double PLn, double PLp,
                                double dt_mix, const RPShared &s) {
    L.Rp = in.Rp;
    L.Rn = in.Rn;
    L.d1R = in.d1R;
    L.PG = in.PG;
    L.dtL = in.dt;
    L.re_n = in.re;
    L.remain = in.re + dt_mix;
    L.time = -in.re;
    L.PC0 = PLn + s.PL0;
    L.PC1 = 0.5 * (PLp - PLm) / dt_mix;
    L.PC2 = 0.5 * (PLp + PLm - 2.0 * PLn) / (dt_mix * dt_mix);
    L.SumHeat = 0.0;
    L.SumVis = 0.0;
    L.Rt = 0.0;
    L.counter = 0;
}

// One pass of the sub-step loop body (bubble_CUDA_kernel.cuh:508-538).
__device__ inline void rp_substep(RPLane &L, const RPConst &k, const RPShared &s) {
    double PGn = L.PG;
    double PL = L.PC2 * L.time * L.time + L.PC1 * L.time + L.PC0;
    rp_radius(&L.Rt, &L.Rp, &L.Rn, &L.d1R, PGn, PL, &L.dtL, &L.remain, s);
    L.time = L.time + L.dtL;
    cd alpha;
    double om = rp_omega(&alpha, PGn, L.Rn, k, s);
    cd Lp = rp_lp(alpha, L.Rn);
    double PGp = rp_pg(PGn, L.Rp, L.Rn, om, L.dtL, Lp, k, s);
    L.PG = PGp;
    double Rp = L.Rp, Rn = L.Rn;
    double s0 = 0.5 * (Rp + Rn);
    double s1 = 1 / (Lp.re * Lp.re + Lp.im * Lp.im) * s.T0 / (k.PG0 * k.R03);
    double dTdr = Lp.re * s1 * (k.PG0 * k.R03 - 0.5 * (PGp + PGn) * s0 * s0 * s0) +
                  Lp.im * s1 * (PGp * Rp * Rp * Rp - PGn * Rn * Rn * Rn) / (om * L.dtL);
    L.SumHeat -= 4.0 * s.pi * Rp * Rp * s.K0 * dTdr * L.dtL;
    L.SumVis += 4.0 * s.pi * Rp * Rp * 4.0 * s.mu * L.d1R / Rp * L.d1R * L.dtL;
    L.counter++;
}

// Exit of BubbleRadiusKernel (bubble_CUDA_kernel.cuh:542-550).
__device__ inline void rp_end(const RPLane &L, double dt_mix, RPState *out, double *R_t,
                              double *Q_B) {
    out->Rp = L.Rp;
    out->Rn = L.Rn;
    out->d1R = L.d1R;
    out->PG = L.PG;
    out->dt = L.dtL;
    out->re = L.remain;
    *R_t = L.Rt;
    *Q_B = (L.SumHeat + L.SumVis) / (dt_mix - L.remain + L.re_n);
}

// ---- constants of the shortened-chain sub-step (rp_fast.h) ---------------
struct RPFastConst {   // per bubble (rp_fast_const)
    double iR0, PG0, iPG0, calpha_irho, PG0R03, T0_PG0R03, A0, B0;
};
struct RPFastShared {  // shared by all bubbles
    double PL0, gam, mu, rho, irho, dt0, g3, k3, C0, two_sig, four_mu, mu2, m2sig_rho,
        four_pi_K0, sixteen_pi_mu;
    const double *vtab;  // the V and Y tables (rp_vfit.h) staged in LDS by the kernel
};

}  // namespace brp
