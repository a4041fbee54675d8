// rp_fast.h -- the RP sub-step with a short dependency chain, generic over the
// scalar type.
//
// Same model as rp_device.h (bubble_CUDA_kernel.cuh:227-340), rearranged for
// latency, since one collapsing bubble's thousands of sequential sub-steps
// bound a whole launch:
//   * alpha = s (1 + i) always (rp_alpha), so
//       coth(alpha) = (1 - |E|^2 + 2 i Im E) / |1 - E|^2,  E = exp(-2 alpha),
//     from expm1(-2s) and sincos(s); for s > 20, |E| < 5e-18 and
//     alpha coth(alpha) - 1 = (s - 1, s) to within a relative 1e-17, so the
//     transcendentals are skipped (this is where collapsing bubbles spend
//     about half their sub-steps);
//   * 1 / alpha^2 = -i / (2 s^2); Lp = R (1/u + 1.5 i / s^2),
//     u = alpha coth(alpha) - 1;
//   * omega_N and 1/omega_N from one rsqrt of eta; sqrt(c_alpha omega / R)
//     and its reciprocal from one rsqrt;
//   * reciprocals by v_rcp_f64 and two Newton steps (drcp), reciprocal
//     square roots by v_rsq_f64 and one second-order step (drsqrt); every
//     fused multiply-add is written out (the file is compiled with
//     contraction off), so the instruction sequence is fixed by the source.
// A wavefront issues one instruction at a time whether one lane or 64 are
// active, so the sub-step of a lone collapsing bubble costs its instruction
// count: the polynomial coefficients are read from a constant table into
// scalar registers right where they are used (not kept live across the
// sub-step loop, where they would be spilled) and each Horner step is one
// three-operand v_fma_f64.
// T = double is the forward.  T = dd (value, tangent) is the adjoint's
// linearisation: every dd operation computes its value with the same single
// operation the double code uses, so the value part -- and with it the
// sub-step schedule -- is bit-identical to the forward's, and the tangent
// follows the same branches.
#pragma once

#include <type_traits>

#include "rp_device.h"

namespace brp {

// Reciprocal of a normal double: hardware estimate + two Newton steps.
__device__ inline double drcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Reciprocal square root of a positive normal double: hardware estimate y,
// e = 1 - x y^2, then y (1 + e/2 + 3 e^2/8) (second-order Newton-Raphson
// step for x^(-1/2); the estimate's error squared and cubed are far below
// one ulp).
__device__ inline double drsqrt(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(y * -x, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

struct dd {
    double v, d;  // value, tangent
    __device__ dd() = default;
    __device__ dd(double x) : v(x), d(0.0) {}
    __device__ dd(double x, double t) : v(x), d(t) {}
};

__device__ inline dd operator+(dd a, dd b) { return dd(a.v + b.v, a.d + b.d); }
__device__ inline dd operator-(dd a, dd b) { return dd(a.v - b.v, a.d - b.d); }
__device__ inline dd operator-(dd a) { return dd(-a.v, -a.d); }
__device__ inline dd operator*(dd a, dd b) { return dd(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ inline dd operator*(double a, dd b) { return dd(a * b.v, a * b.d); }
__device__ inline dd operator*(dd a, double b) { return dd(a.v * b, a.d * b); }
__device__ inline dd operator/(dd a, double b) { return dd(a.v / b, a.d / b); }

__device__ inline double val(double x) { return x; }
__device__ inline double val(dd x) { return x.v; }
__device__ inline double fma_(double a, double b, double c) { return fma(a, b, c); }
__device__ inline dd fma_(dd a, dd b, dd c) {
    return dd(fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d);
}
__device__ inline double rcp_(double x) { return drcp(x); }
__device__ inline dd rcp_(dd x) {
    const double r = drcp(x.v);
    return dd(r, -r * r * x.d);
}
__device__ inline double rsqrt_(double x) { return drsqrt(x); }
__device__ inline dd rsqrt_(dd x) {
    const double r = drsqrt(x.v);
    return dd(r, -0.5 * r * r * r * x.d);
}
// Constant (scalar-loadable) memory.
typedef __attribute__((address_space(4))) const double cdouble;

// The shared constants from their device table (rp_fast_shared_values),
// read where the sub-step uses them rather than held in scalar registers
// across the sub-step loop.
__device__ inline RPFastShared rp_shared_load(const double *tab) {
    cdouble *t = (cdouble *)tab;
    asm volatile("" : "+s"(t));
    return RPFastShared{t[0], t[1], t[2],  t[3],  t[4],  t[5],  t[6], t[7],
                        t[8], t[9], t[10], t[11], t[12], t[13], t[14], tab + 16};
}

// The series coefficients (kRpPoly), read from the table where each use
// needs them: the pointer is made opaque per use so that the loads are
// neither hoisted nor shared across uses (which would keep 56 scalar
// registers live across the sub-step).
__device__ inline cdouble *rp_poly(const double *poly) {
    cdouble *c = (cdouble *)poly;
    asm volatile("" : "+s"(c));
    return c;
}

#ifndef BRP_OWN_TRANSCENDENTALS
#define BRP_OWN_TRANSCENDENTALS 1
#endif

#if BRP_OWN_TRANSCENDENTALS
// a b + c with c in scalar registers, as one v_fma_f64 (the compiler would
// otherwise copy c into the destination for a two-operand v_fmac_f64).
__device__ inline double fma_sc(double a, double b, double c) {
    double d;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
    return d;
}

// exp(x) - 1 for -40 <= x <= -0.14 (the range alpha_coth_m1_g uses):
// x = n ln2 + r, |r| <= ln2 / 2 (Cody-Waite reduction, ln2 in two parts),
// e^r - 1 by its Taylor series to r^14 (truncation < 2e-19), then
// 2^n (1 + p) - 1, which for these x is in [-1, -0.13] (no cancellation
// beyond one ulp of 1).
__device__ inline double expm1_neg(double x, const double *poly) {
    cdouble *c = rp_poly(poly);
    const double n = rint(x * 1.4426950408889634);
    double r = fma(-n, 6.93147180369123816490e-01, x);
    r = fma(-n, 1.90821492927058770002e-10, r);
    double p = c[0];
#pragma unroll
    for (int k = 1; k < 12; k++) p = fma_sc(p, r, c[k]);
    p = fma(p, r, 0.5);
    p = fma(p * r, r, r);  // e^r - 1
    return n == 0.0 ? p : fma(ldexp(1.0, (int)n), p, ldexp(1.0, (int)n) - 1.0);
}

// For 0 <= x <= 64: sa = +-sin x and sc = sin x cos x, the two quantities
// sin 2x = 2 sc and cos 2x = 1 - 2 sa^2 are made of.  x = k pi/2 + r,
// |r| <= pi/4 (Cody-Waite reduction, pi/2 in two parts, each step one
// rounding through fma), Taylor series of sin r to r^17 and cos r to r^18
// (truncation < 1e-19); sin x = +-(sin r or cos r) by the quadrant k, and
// sin x cos x = (-1)^k sin r cos r.
__device__ inline void trig_parts(double x, const double *poly, double *sa, double *sc) {
    cdouble *t = rp_poly(poly);
    const double k = rint(x * 0.63661977236758134308);
    double r = fma(-k, 1.57079632679489655800e+00, x);
    r = fma(-k, 6.12323399573676603587e-17, r);
    const double r2 = r * r;
    double ps = t[12];
#pragma unroll
    for (int q = 13; q < 20; q++) ps = fma_sc(ps, r2, t[q]);
    ps = fma(-(ps * r2), r, r);  // r - r^3/6 + ...
    double pc = t[20];
#pragma unroll
    for (int q = 21; q < 28; q++) pc = fma_sc(pc, r2, t[q]);
    pc = fma(pc, r2, 0.5);
    pc = fma(-pc, r2, 1.0);  // 1 - r^2/2 + ...
    const bool odd = (((int)k) & 1) != 0;
    const double p = ps * pc;
    *sa = odd ? pc : ps;
    *sc = odd ? -p : p;
}
__device__ inline double expm1_(double x, const double *poly) { return expm1_neg(x, poly); }
__device__ inline void trig_parts_(double x, const double *poly, double *sa, double *sc) {
    trig_parts(x, poly, sa, sc);
}
#else
__device__ inline double expm1_(double x, const double *) { return expm1(x); }
__device__ inline void trig_parts_(double x, const double *, double *sa, double *sc) {
    double sn, cn;
    sincos(x, &sn, &cn);
    *sa = sn;
    *sc = sn * cn;
}
#endif
__device__ inline dd expm1_(dd x, const double *poly) {
    const double m = expm1_(x.v, poly);
    return dd(m, (m + 1.0) * x.d);
}
// The tangent form: sa = sin x with its tangent cos x dx (the value may
// differ from the double form's in sign only, which no caller sees: they use
// sa^2 and sc), sc = sin x cos x with value equal to the double form's.
__device__ inline void trig_parts_(dd x, const double *poly, dd *sa, dd *sc) {
    double a, p;
    trig_parts_(x.v, poly, &a, &p);
    // sin x and cos x up to a common sign pattern: a = +-sin x, and
    // p / a = cos x when a = sin x; recover cos x from a and p without a
    // division by choosing the tangent of sa consistently with its sign
    const double c = a != 0.0 ? p / a : 1.0;  // cos x * (sign of sa / sign of sin x)
    *sa = dd(a, c * x.d);
    *sc = dd(p, (c * c - a * a) * x.d);
}

template <typename T>
struct cx {
    T re, im;
};

// alpha coth(alpha) - 1 = (re + i im) / d for alpha = s (1 + i), s > 0,
// with the division left to the caller, which folds it into its own
// reciprocal (one reciprocal per caller instead of two).  The callers use
// it for s > 0.1 / sqrt(2) only, so expm1_ sees -40 <= -2s < -0.14.
template <typename T>
struct cxd {
    T re, im, d;
};
template <typename T>
__device__ inline cxd<T> alpha_coth_m1_g(T s, const double *poly) {
    if (val(s) > 20.0) return cxd<T>{s - T(1.0), s, T(1.0)};
    const T m = expm1_(-2.0 * s, poly);  // exp(-2s) - 1
    T sa, sc;                            // +-sin s, sin s cos s
    trig_parts_(s, poly, &sa, &sc);
    const T e2 = T(1.0) + m;
    const T s2 = 2.0 * sc;                     // sin 2s
    const T c2 = fma_(-2.0 * sa, sa, T(1.0));  // cos 2s
    const T ei = -(e2 * s2);                   // Im E
    const T one_m_er = fma_(2.0 * sa, sa, -(m * c2));  // 1 - Re E
    const T D = fma_(one_m_er, one_m_er, ei * ei);     // |1 - E|^2
    const T nr = -(m * (m + T(2.0)));                  // 1 - |E|^2
    const T ni = 2.0 * ei;
    // coth(alpha) = (nr + i ni) / D
    return cxd<T>{fma_(s, nr - ni, -D), s * (nr + ni), D};
}

// Upsilon(alpha) (bubble_CUDA_kernel.cuh:296-311) from s and 1/s:
// 3 gam / (1 + k3 (alpha coth alpha - 1) / alpha^2) with 1/alpha^2 =
// -i / (2 s^2), written as 3 gam D (A - i B) / (A^2 + B^2).
template <typename T>
__device__ inline cx<T> upsilon_g(T s, T is, double gam, double k3, const double *poly) {
    T A, B, D;
    if (val(s) * 1.4142135623730951 > 1.0e-1) {
        const cxd<T> u = alpha_coth_m1_g(s, poly);
        const T q = k3 * (0.5 * (is * is));  // k3 / (2 s^2)
        A = fma_(q, u.im, u.d);
        B = -(q * u.re);
        D = u.d;
    } else {  // 1/3 + a^2 (-1/45 + a^2 (2/945 - a^2/4725)),  a^2 = i q
        const T q = 2.0 * (s * s);
        T ur = T(2.0 / 945.0), ui = -(q * (1.0 / 4725.0));
        T nr = fma_(-q, ui, T(-1.0 / 45.0)), ni = q * ur;
        const T tr = fma_(-q, ni, T(1.0 / 3.0));
        const T ti = q * nr;
        A = fma_(T(k3), tr, T(1.0));
        B = k3 * ti;
        D = T(1.0);
    }
    const T f = ((3.0 * gam) * D) * rcp_(fma_(A, A, B * B));
    return cx<T>{A * f, -(B * f)};
}

// solveLp (bubble_CUDA_kernel.cuh:314-320) from s and 1/s:
// R (1 / (alpha coth alpha - 1) - 3 / alpha^2).
template <typename T>
__device__ inline cx<T> lp_g(T s, T is, T R, const double *poly) {
    if (val(s) * 1.4142135623730951 > 1.0e-1) {
        const cxd<T> u = alpha_coth_m1_g(s, poly);
        const T g = u.d * rcp_(fma_(u.re, u.re, u.im * u.im));
        return cx<T>{R * (u.re * g), R * fma_(1.5 * is, is, -(u.im * g))};
    }
    // 1/5 + a^2 (-1/175 + a^2 (2/7875 - 37 a^2 / 3031875)),  a^2 = i q
    const T q = 2.0 * (s * s);
    T ur = T(2.0 / 7875.0), ui = -(q * (37.0 / 3031875.0));
    T nr = fma_(-q, ui, T(-1.0 / 175.0)), ni = q * ur;
    return cx<T>{R * fma_(-q, ni, T(1.0 / 5.0)), R * (q * nr)};
}

// omega = sqrt(max(eta, floor)) / (rho R) and 1 / omega from one rsqrt.
template <typename T>
__device__ inline void omega_g(T eta, T scale, T iscale, T *om, T *iom) {
    const T e = val(eta) >= 1.0e-6 * kDblEps ? eta : T(1.0e-6 * kDblEps);
    const T r = rsqrt_(e);
    *om = (e * r) * scale;
    *iom = r * iscale;
}

template <typename T>
struct RPLaneF {
    T Rp, Rn, d1R, PG, Rt, PC0, PC1, PC2;
    double dtL, remain, time, re_n, SumHeat, SumVis;
    int counter;
};

// Entry of BubbleRadiusKernel (bubble_CUDA_kernel.cuh:487-504).
template <typename T>
__device__ inline void rp_begin_g(RPLaneF<T> &L, T Rp, T Rn, T d1R, T PG, double dt_n, double re_n,
                                  T PLm, T PLn, T PLp, double dt_mix, double PL0) {
    L.Rp = Rp;
    L.Rn = Rn;
    L.d1R = d1R;
    L.PG = PG;
    L.dtL = dt_n;
    L.re_n = re_n;
    L.remain = re_n + dt_mix;
    L.time = -re_n;
    L.PC0 = PLn + T(PL0);
    L.PC1 = (0.5 * (PLp - PLm)) / dt_mix;
    L.PC2 = (0.5 * (PLp + PLm - 2.0 * PLn)) / (dt_mix * dt_mix);
    L.SumHeat = 0.0;
    L.SumVis = 0.0;
    L.Rt = T(0.0);
    L.counter = 0;
}

// One pass of the sub-step loop body (bubble_CUDA_kernel.cuh:508-538).
template <typename T>
__device__ inline void rp_substep_g(RPLaneF<T> &L, const RPFastConst &k, const RPFastShared &s) {
    // solveRayleighPlesset (bubble_CUDA_kernel.cuh:227-260)
    const T PGn = L.PG;
    const double t = L.time;
    const T PL = fma_(L.PC2, T(t * t), fma_(L.PC1, T(t), L.PC0));
    const T Rm = L.Rn;
    const T Rn = L.Rp;
    const T iRn = rcp_(Rn);
    T d1R = L.d1R;
    const T sv = iRn * fma_(T(s.four_mu), d1R, T(s.two_sig));  // (2 sig + 4 mu dR) / R
    const T Rnd2R = fma_(-1.5 * d1R, d1R, (PGn - PL - sv) * s.irho);
    const double dtm = L.dtL;
    double dtp = fmin(1.01 * dtm, s.dt0);
    const double ad = fabs(val(d1R)), aa = fabs(val(Rnd2R));
    if (ad > kDblEps) dtp = fmin(dtp, (0.01 * val(Rn)) * drcp(ad));
    if (aa > kDblEps) dtp = fmin(dtp, (0.1 * val(Rn)) * drsqrt(aa));
    const double idtm = drcp(dtm);
    const double r = dtp * idtm;
    const T acc = (dtp * Rnd2R) * iRn;  // (dtp / Rn) Rnd2R
    const T Rp = fma_(acc, T(0.5 * (dtp + dtm)), fma_(T(1.0 + r), Rn, -r * Rm));
    const double dm2 = dtm * dtm, dp2 = dtp * dtp;
    const double iden = drcp(dtm * dtp * (dtm + dtp));
    d1R = fma_(fma_(T(dm2), Rp, fma_(T(dp2 - dm2), Rn, -dp2 * Rm)), T(iden), acc);
    const double dtL = dtp;
    if (dtp >= L.remain) {  // last sub-step: R_t at the sync time
        const double d = L.remain;
        const double r2 = d * idtm;
        L.Rt = fma_((d * Rnd2R) * iRn, T(0.5 * (d + dtm)), fma_(T(1.0 + r2), Rn, -r2 * Rm));
        dtp = d;
    } else {
        L.Rt = Rp;
    }
    L.remain -= dtp;
    L.time = t + dtL;
    L.dtL = dtL;
    L.Rp = Rp;
    L.Rn = Rn;
    L.d1R = d1R;
    // solveOmegaN at (PGn, Rn) (bubble_CUDA_kernel.cuh:263-289)
    const T Rx = Rn * k.iR0;
    const T coef1 = k.PG0 * rcp_(PGn * (Rx * Rx * Rx));
    const T coef2 = PGn * s.rho;
    const T coef3 = 4.0 * (iRn * iRn);
    const T value1 = s.m2sig_rho * iRn;
    const T scale = iRn * s.irho;
    const T iscale = Rn * s.rho;
    T eta = fma_(s.g3 * coef1, coef2, fma_(T(-s.mu2), coef3, value1));
    T om, iom;
    omega_g(eta, scale, iscale, &om, &iom);
    const T ca = k.calpha * iRn;
    for (int it = 0; it < 3; it++) {
        const T x = ca * om;
        const T isa = rsqrt_(x);
        const cx<T> ups = upsilon_g(x * isa, isa, s.gam, s.k3, s.poly);
        const T ur = ups.re * coef1, ui = ups.im * coef1;
        const T mu_eff = fma_(ui * PGn, 0.25 * iom, T(s.mu));  // mu + ui PG / (4 omega)
        eta = fma_(ur, coef2, fma_(-(mu_eff * mu_eff), coef3, value1));
        omega_g(eta, scale, iscale, &om, &iom);
    }
    const T x = ca * om;
    const T isa = rsqrt_(x);
    const cx<T> Lp = lp_g(x * isa, isa, Rn, s.poly);
    // solvePG (bubble_CUDA_kernel.cuh:322-340)
    const T Rbar = 0.5 * (Rp + Rn);
    const T iRbar = rcp_(Rbar);
    const T Rx_ = Rbar * k.iR0, Rpx = Rp * k.iR0, Rnx = Rn * k.iR0;
    const T iL = rcp_(fma_(Lp.re, Lp.re, Lp.im * Lp.im));
    const T Lnr = Lp.re * iL, Lni = Lp.im * iL;
    const T cdt = (s.C0 * iRbar) * k.iPG0;  // c0 / dt
    const T c0 = dtL * cdt;
    const T c3 = fma_(c0 * Lnr, 0.5 * (Rx_ * Rx_ * Rx_), ((1.5 * s.gam) * (Rp - Rn)) * iRbar);
    const T q = (cdt * Lni) * iom;  // c0 Lp_NI / (omega dt)
    const T c1 = fma_(-q, Rpx * Rpx * Rpx, T(1.0) + c3);
    const T c2 = fma_(-q, Rnx * Rnx * Rnx, T(1.0) - c3);
    const T PGp = fma_(c2, PGn, (c0 * Lnr) * k.PG0) * rcp_(c1);
    L.PG = PGp;
    if constexpr (std::is_same<T, double>::value) {
        // heat and viscous dissipation (Q_B): dTdr dt with the 1/dt of its
        // second term cancelled against the dt factor
        const double s1 = iL * k.T0_PG0R03;
        const double dTdr_dt =
            fma(Lp.re * s1 * dtL, fma(-0.5 * (PGp + PGn), Rbar * Rbar * Rbar, k.PG0R03),
                (Lp.im * s1 * iom) * fma(PGp, Rp * Rp * Rp, -(PGn * (Rn * Rn * Rn))));
        L.SumHeat = fma(-s.four_pi_K0 * (Rp * Rp), dTdr_dt, L.SumHeat);
        L.SumVis = fma(s.sixteen_pi_mu * Rp * dtL, d1R * d1R, L.SumVis);
    }
    L.counter++;
}

// Exit of BubbleRadiusKernel (bubble_CUDA_kernel.cuh:542-550).
__device__ inline void rp_end_g(const RPLaneF<double> &L, double dt_mix, RPState *out, double *R_t,
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

}  // namespace brp
