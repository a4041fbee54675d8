// rp_tangent.h -- tangent (forward-mode) linearisation of one RP sub-step.
//
// The adjoint of a bubble's Rayleigh-Plesset integration over one mixture step
// needs v^T J for the map
//   (Rp, Rn, d1R, PG | PL_p, PL_n, PL_m)  ->  (Rp', Rn', d1R', PG', R_t)
// at a frozen sub-step schedule (the accepted step sizes dtp, the clamp of the
// last sub-step to the sync time and the series/closed-form branch of the gas
// model are those of the primal run; they are not differentiated).  Seven
// inputs, five outputs: a lane carries (value, one tangent) pairs seeded with
// one input direction, so seven lanes of one bubble give the seven columns of
// J and each lane ends with one entry of v^T J.  The primal part repeats the
// shortened-chain sub-step of rp_device.h (rp_substep_fast) expression by
// expression, so the replayed schedule is the forward's.
#pragma once

#include "rp_device.h"

namespace brp {

struct dd {
    double v, d;  // value, tangent
};

__device__ inline dd dk(double v) { return dd{v, 0.0}; }
__device__ inline dd operator+(dd a, dd b) { return dd{a.v + b.v, a.d + b.d}; }
__device__ inline dd operator+(dd a, double b) { return dd{a.v + b, a.d}; }
__device__ inline dd operator+(double a, dd b) { return dd{a + b.v, b.d}; }
__device__ inline dd operator-(dd a, dd b) { return dd{a.v - b.v, a.d - b.d}; }
__device__ inline dd operator-(dd a, double b) { return dd{a.v - b, a.d}; }
__device__ inline dd operator-(double a, dd b) { return dd{a - b.v, -b.d}; }
__device__ inline dd operator-(dd a) { return dd{-a.v, -a.d}; }
__device__ inline dd operator*(dd a, dd b) { return dd{a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline dd operator*(dd a, double b) { return dd{a.v * b, a.d * b}; }
__device__ inline dd operator*(double a, dd b) { return dd{a * b.v, a * b.d}; }
__device__ inline dd operator/(dd a, dd b) {
    const double q = a.v / b.v;
    return dd{q, (a.d - q * b.d) / b.v};
}
__device__ inline dd operator/(dd a, double b) { return dd{a.v / b, a.d / b}; }
__device__ inline dd operator/(double a, dd b) {
    const double q = a / b.v;
    return dd{q, -q * b.d / b.v};
}
__device__ inline dd d_sqrt(dd a) {
    const double r = sqrt(a.v);
    return dd{r, r > 0.0 ? 0.5 * a.d / r : 0.0};
}
__device__ inline dd d_expm1(dd a) {
    const double m = expm1(a.v);
    return dd{m, (m + 1.0) * a.d};
}
__device__ inline void d_sincos(dd a, dd *s, dd *c) {
    double sn, cn;
    sincos(a.v, &sn, &cn);
    *s = dd{sn, cn * a.d};
    *c = dd{cn, -sn * a.d};
}
// max(a, floor) with the branch of the primal
__device__ inline dd d_fmax(dd a, double floor) { return a.v >= floor ? a : dk(floor); }

struct cdd {
    dd re, im;
};

__device__ inline cdd alpha_coth_m1_t(dd s) {
    const dd m = d_expm1(-2.0 * s);
    dd sn, cn;
    d_sincos(s, &sn, &cn);
    const dd e2 = 1.0 + m;
    const dd s2 = 2.0 * sn * cn;
    const dd c2 = 1.0 - 2.0 * sn * sn;
    const dd ei = -(e2 * s2);
    const dd one_m_er = 2.0 * sn * sn - m * c2;
    const dd iD = 1.0 / (one_m_er * one_m_er + ei * ei);
    const dd cr = -(m * (m + 2.0)) * iD;
    const dd ci = 2.0 * ei * iD;
    return cdd{s * (cr - ci) - 1.0, s * (cr + ci)};
}

__device__ inline cdd upsilon_t(dd sa, double gam, double k3) {
    dd tr, ti;
    if (sa.v * 1.4142135623730951 > 1.0e-1) {
        const cdd u = alpha_coth_m1_t(sa);
        const dd h = 0.5 / (sa * sa);
        tr = u.im * h;
        ti = -(u.re * h);
    } else {  // series in a^2 = 2 i s^2:  1/3 + a^2 (-1/45 + a^2 (2/945 - a^2/4725))
        const dd q = 2.0 * sa * sa;  // a^2 = i q
        // u = 2/945 - i q / 4725
        dd ur = dk(2.0 / 945.0), ui = -(q / 4725.0);
        // u = -1/45 + (i q) u = -1/45 - q ui + i q ur
        dd nr = -1.0 / 45.0 - q * ui, ni = q * ur;
        ur = nr;
        ui = ni;
        nr = 1.0 / 3.0 - q * ui;
        ni = q * ur;
        tr = nr;
        ti = ni;
    }
    const dd dr = 1.0 + k3 * tr, di = k3 * ti;
    const dd f = 3.0 * gam / (dr * dr + di * di);
    return cdd{dr * f, -(di * f)};
}

__device__ inline cdd lp_t(dd sa, dd R) {
    if (sa.v * 1.4142135623730951 > 1.0e-1) {
        const cdd u = alpha_coth_m1_t(sa);
        const dd iu = 1.0 / (u.re * u.re + u.im * u.im);
        return cdd{R * u.re * iu, R * (1.5 / (sa * sa) - u.im * iu)};
    }
    // 1/5 + a^2 (-1/175 + a^2 (2/7875 - 37 a^2 / 3031875)),  a^2 = i q
    const dd q = 2.0 * sa * sa;
    dd ur = dk(2.0 / 7875.0), ui = -(q * (37.0 / 3031875.0));
    dd nr = -1.0 / 175.0 - q * ui, ni = q * ur;
    ur = nr;
    ui = ni;
    nr = 1.0 / 5.0 - q * ui;
    ni = q * ur;
    return cdd{R * nr, R * ni};
}

// Lane state of the tangent integration: differentiable radius state and
// pressure polynomial, frozen schedule quantities.
struct RPLaneT {
    dd Rp, Rn, d1R, PG, Rt, PC0, PC1, PC2;
    double dtL, remain, time;
    int counter;
};

__device__ inline void rp_begin_t(RPLaneT &L, dd Rp, dd Rn, dd d1R, dd PG, double dt_n,
                                  double re_n, dd PLm, dd PLn, dd PLp, double dt_mix,
                                  double PL0) {
    L.Rp = Rp;
    L.Rn = Rn;
    L.d1R = d1R;
    L.PG = PG;
    L.dtL = dt_n;
    L.remain = re_n + dt_mix;
    L.time = -re_n;
    L.PC0 = PLn + PL0;
    L.PC1 = 0.5 * (PLp - PLm) / dt_mix;
    L.PC2 = 0.5 * (PLp + PLm - 2.0 * PLn) / (dt_mix * dt_mix);
    L.Rt = dk(0.0);
    L.counter = 0;
}

// rp_substep_fast in (value, tangent) arithmetic; the step size is chosen
// from the values exactly as there and is a constant of the tangent.
__device__ inline void rp_substep_t(RPLaneT &L, const RPFastConst &k, const RPFastShared &s) {
    const dd PGn = L.PG;
    const dd PL = L.PC2 * (L.time * L.time) + L.PC1 * L.time + L.PC0;
    const dd Rm = L.Rn;
    const dd Rn = L.Rp;
    const dd iRn = 1.0 / Rn;
    dd d1R = L.d1R;
    const dd Rnd2R = (PGn - PL - s.two_sig * iRn - s.four_mu * iRn * d1R) * s.irho -
                     1.5 * d1R * d1R;
    const double dtm = L.dtL;
    double dtp = fmin(1.01 * dtm, s.dt0);
    if (fabs(d1R.v) > kDblEps) dtp = fmin(dtp, 0.01 * Rn.v / fabs(d1R.v));
    if (fabs(Rnd2R.v) > kDblEps) dtp = fmin(dtp, 0.1 * Rn.v / sqrt(fabs(Rnd2R.v)));
    const double r = dtp / dtm;
    const dd acc = dtp * iRn * Rnd2R;
    const dd Rp = (1.0 + r) * Rn - r * Rm + acc * (0.5 * (dtp + dtm));
    d1R = (dtm * dtm * Rp + (dtp * dtp - dtm * dtm) * Rn - dtp * dtp * Rm) /
              (dtm * dtp * (dtm + dtp)) + acc;
    const double dtL = dtp;
    if (dtp >= L.remain) {
        const double d = L.remain;
        const double r2 = d / dtm;
        L.Rt = (1.0 + r2) * Rn - r2 * Rm + d * iRn * Rnd2R * (0.5 * (d + dtm));
        dtp = d;
    } else {
        L.Rt = Rp;
    }
    L.remain -= dtp;
    L.time = L.time + dtL;
    L.dtL = dtL;
    L.Rp = Rp;
    L.Rn = Rn;
    L.d1R = d1R;
    // solveOmegaN at (PGn, Rn)
    const dd Rx = Rn * k.iR0;
    const dd coef1 = k.PG0 / (PGn * Rx * Rx * Rx);
    const dd coef2 = PGn * s.rho;
    const dd coef3 = 4.0 * iRn * iRn;
    const dd value1 = s.m2sig_rho * iRn;
    const dd scale = iRn * s.irho;
    dd eta = s.g3 * coef1 * coef2 + value1 - s.mu2 * coef3;
    dd om = d_sqrt(d_fmax(eta, 1.0e-6 * kDblEps)) * scale;
    const dd ca = k.calpha * iRn;
    for (int it = 0; it < 3; it++) {
        const cdd ups = upsilon_t(d_sqrt(ca * om), s.gam, s.k3);
        const dd ur = ups.re * coef1, ui = ups.im * coef1;
        const dd mu_eff = s.mu + ui * PGn / (4.0 * om);
        eta = ur * coef2 + value1 - mu_eff * mu_eff * coef3;
        om = d_sqrt(d_fmax(eta, 1.0e-6 * kDblEps)) * scale;
    }
    const cdd Lp = lp_t(d_sqrt(ca * om), Rn);
    // solvePG
    const dd Rbar = 0.5 * (Rp + Rn);
    const dd iRbar = 1.0 / Rbar;
    const dd Rx_ = Rbar * k.iR0, Rpx = Rp * k.iR0, Rnx = Rn * k.iR0;
    const dd iL = 1.0 / (Lp.re * Lp.re + Lp.im * Lp.im);
    const dd Lnr = Lp.re * iL, Lni = Lp.im * iL;
    const dd iom = 1.0 / om;
    const dd cdt = s.C0 * iRbar * k.iPG0;
    const dd c0 = dtL * cdt;
    const dd c3 = 1.5 * s.gam * (Rp - Rn) * iRbar + c0 * Lnr * 0.5 * Rx_ * Rx_ * Rx_;
    const dd q = cdt * Lni * iom;
    const dd c1 = 1.0 + c3 - q * Rpx * Rpx * Rpx;
    const dd c2 = 1.0 - c3 - q * Rnx * Rnx * Rnx;
    L.PG = (c2 * PGn + c0 * Lnr * k.PG0) / c1;
    L.counter++;
}

}  // namespace brp
