// rp_fast.h -- the RP sub-step with a short dependency chain, generic over the
// scalar type.
//
// Same model as rp_device.h (bubble_CUDA_kernel.cuh:227-340), rearranged for
// latency, since one collapsing bubble's thousands of sequential sub-steps
// bound a whole launch:
//   * alpha = s (1 + i) always (rp_alpha), so the thermal-impedance terms
//     are functions of s alone: Upsilon's U = (alpha coth alpha - 1)/alpha^2
//     and Lp's W = 1/(alpha coth alpha - 1) - 3/alpha^2, of which solvePG
//     needs only Lp/|Lp|^2 = Y/R, Y = 1/conj(W) (U and Y are finite at
//     s = 0, where the two terms of W cancel).  For |alpha| <= 0.1 the
//     reference's own Taylor series; up to s = 20 one degree-9 polynomial
//     per interval of width 1/8 (rp_vtable.h, staged in LDS; error below
//     2e-16 of |U| and |Y|); for s > 20 coth alpha = 1 to below an ulp, so
//     both follow from alpha - 1 in closed form (collapsing bubbles spend
//     about half their sub-steps there);
//   * omega_N and 1/omega_N from one rsqrt of eta; sqrt(c_alpha omega / R)
//     and its reciprocal from one rsqrt;
//   * reciprocals by v_rcp_f64 and two Newton steps (drcp), reciprocal
//     square roots by v_rsq_f64 and one second-order step (drsqrt); every
//     fused multiply-add is written out (the file is compiled with
//     contraction off), so the instruction sequence is fixed by the source.
// A wavefront issues one instruction at a time whether one lane or 64 are
// active, so the sub-step of a lone collapsing bubble costs its instruction
// count: the polynomial replaces an exp, a sin, a cos and the complex
// arithmetic joining them (about 110 instructions per evaluation, four per
// sub-step) by 18 fused multiply-adds on LDS operands.
// T = double is the forward.  T = dd (value, tangent) is the adjoint's
// linearisation: every dd operation computes its value with the same single
// operation the double code uses, so the value part -- and with it the
// sub-step schedule -- is bit-identical to the forward's, and the tangent
// follows the same branches.
#pragma once

#include <type_traits>

#include "brp_internal.h"
#include "rp_device.h"

namespace brp {

// Reciprocal of a normal double: hardware estimate r, e = 1 - x r, then
// r (1 + e + e^2) (third-order step: the estimate's error cubed is far below
// one ulp; one Newton step, r (1 + e), leaves up to 11 ulp).
__device__ inline double drcp(double x) {
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    return fma(r, fma(e, e, e), r);
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
// Per-bubble constants of the sub-step.  A0 and B0 are the Omega iteration's
// (rp_substep_g): 3 gam rho PG0 R0^3 and -3 gam PG0 R0^3 / 4.
__device__ inline RPFastConst rp_fast_const(double R0, double PG0, double calpha, double T0,
                                            const RPFastShared &s) {
    const double PG0R03 = PG0 * (R0 * R0 * R0);
    return RPFastConst{1.0 / R0,      PG0,    1.0 / PG0, calpha, PG0R03, T0 / PG0R03,
                       (s.g3 * s.rho) * PG0R03, (-0.25 * s.g3) * PG0R03};
}

// Constant (scalar-loadable) memory.
typedef __attribute__((address_space(4))) const double cdouble;

// The shared constants from their device table (rp_table_values), read
// where the sub-step uses them rather than held in scalar registers across
// the sub-step loop; vtab is the kernel's LDS copy of the U and Y tables.
__device__ inline RPFastShared rp_shared_load(const double *tab, const double *vtab) {
    cdouble *t = (cdouble *)tab;
    asm volatile("" : "+s"(t));
    return RPFastShared{t[0], t[1], t[2],  t[3],  t[4],  t[5],  t[6], t[7],
                        t[8], t[9], t[10], t[11], t[12], t[13], t[14], vtab};
}

#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) const double ldouble;  // LDS
#else
typedef const double ldouble;  // the host pass only parses device code
#endif

// Copies the U and Y tables (kRpVSize doubles from the device table) into the
// workgroup's LDS array; the caller synchronises the workgroup afterwards.
__device__ inline void rp_vtab_stage(const double *tab, double *lds) {
    for (int k = threadIdx.x; k < kRpVSize; k += blockDim.x) lds[k] = tab[kRpVOffset + k];
}

struct cx_d {
    double re, im;
};

// U or Y (ftab = that function's table in LDS) for 0 <= s <= 20: interval
// k = floor(8 s), t = 16 s - (2k + 1), Horner in t for the real and
// imaginary parts (two independent chains); with dpr/dpi also d/ds.
__device__ inline cx_d v_poly(double s, const double *ftab, double *dpr = nullptr,
                              double *dpi = nullptr) {
    static_assert(kVPerUnit == 8, "t below scales s by 2^4");
    const int k = max(0, min((int)ldexp(s, 3), kVIntervals - 1));
    const double t = ldexp(s, 4) - (2.0 * k + 1.0);  // exact product, one rounding
    ldouble *c = (ldouble *)ftab + k * (2 * (kVDegree + 1));
    double cr[kVDegree + 1], ci[kVDegree + 1];
#pragma unroll
    for (int j = kVDegree; j >= 0; j--) {
        cr[j] = c[2 * j];
        ci[j] = c[2 * j + 1];
    }
    if (!dpr) {
        // all ten LDS reads first, then the two Horner chains: the chains
        // wait only for the first read's latency (the scheduler otherwise
        // interleaves reads and waits along the chain)
        __builtin_amdgcn_sched_group_barrier(0x0100, kVDegree + 1, 0);  // DS reads
        __builtin_amdgcn_sched_group_barrier(0x0002, 2 * kVDegree, 0);  // VALU
    }
    double pr = cr[kVDegree], pi = ci[kVDegree];
    double qr = 0.0, qi = 0.0;  // d/dt
#pragma unroll
    for (int j = kVDegree - 1; j >= 0; j--) {
        if (dpr) {
            qr = fma(qr, t, pr);
            qi = fma(qi, t, pi);
        }
        pr = fma(pr, t, cr[j]);
        pi = fma(pi, t, ci[j]);
    }
    if (dpr) {
        *dpr = (2.0 * kVPerUnit) * qr;  // dt/ds
        *dpi = (2.0 * kVPerUnit) * qi;
    }
    return cx_d{pr, pi};
}

template <typename T>
struct cx {
    T re, im;
};

__device__ inline cx<double> v_table(double s, const double *ftab) {
    const cx_d v = v_poly(s, ftab);
    return cx<double>{v.re, v.im};
}
// The tangent form: the value is the double form's, the tangent d/ds ds.
__device__ inline cx<dd> v_table(dd s, const double *ftab) {
    double dr, di;
    const cx_d v = v_poly(s.v, ftab, &dr, &di);
    return cx<dd>{dd(v.re, dr * s.d), dd(v.im, di * s.d)};
}

// |alpha| <= 0.1: the reference's Taylor branches (bubble_CUDA_kernel.cuh:
// 302-306, 317-319).  They are kept rather than the tables because there
// omega, and with it s^2, can be as small as the eta floor allows, and the
// sub-step divides the imaginary parts of Upsilon and Lp / |Lp|^2 (of order
// s^2) by omega: the series give them to relative accuracy, the tables only
// to accuracy relative to |U| and |Y|.
// |alpha| = sqrt(2) s, so |alpha| > 0.1 is s > 0.1 / sqrt(2) (one literal;
// the two forms can differ only within an ulp of the threshold, where the
// series and the table agree to the series' truncation error).
template <typename T>
__device__ inline bool small_alpha(T s) {
    return !(val(s) > 0.070710678118654752);
}

// Upsilon(alpha) = 3 gam / (1 + k3 U) (bubble_CUDA_kernel.cuh:296-311) from
// s and is = 1/s, U = (alpha coth alpha - 1) / alpha^2: the table up to
// s = 20; beyond it coth alpha = 1 to below an ulp, so
// U = (alpha - 1) / alpha^2 = (is / 2) (1 + i (is - 1)).  Returned as
// conj(Upsilon) / (3 gam) = conj(1 / (1 + k3 U)): the caller's constants carry
// the factor 3 gam and the sign of the imaginary part.
template <typename T>
__device__ inline cx<T> upsilon_g(T s, T is, double k3, const double *vtab) {
    cx<T> u;
    if (val(s) > 20.0) {
        const T h = 0.5 * is;
        u = cx<T>{h, h * (is - T(1.0))};
    } else if (small_alpha(s)) {  // 1/3 + a^2 (-1/45 + a^2 (2/945 - a^2/4725)), a^2 = i q
        const T q = 2.0 * (s * s);
        const T ur = T(2.0 / 945.0), ui = -(q * (1.0 / 4725.0));
        const T nr = fma_(-q, ui, T(-1.0 / 45.0)), ni = q * ur;
        u = cx<T>{fma_(-q, ni, T(1.0 / 3.0)), q * nr};
    } else {
        u = v_table(s, vtab);
    }
    const T dr = fma_(T(k3), u.re, T(1.0)), di = k3 * u.im;
    const T f = rcp_(fma_(dr, dr, di * di));
    return cx<T>{dr * f, di * f};
}

// solveLp (bubble_CUDA_kernel.cuh:314-320) as solvePG uses it
// (bubble_CUDA_kernel.cuh:328-329): Lp / |Lp|^2 = Y / R with Lp = R W,
// W = 1 / (alpha coth alpha - 1) - 3 / alpha^2 and Y = 1 / conj(W); iR = 1/R.
// The table up to s = 20; beyond it W = 1/(alpha - 1) - 3/alpha^2, so
// Y = conj(N) / conj(D) with N = alpha^2 (alpha - 1), D = alpha^2 - 3 alpha + 3,
// both divided by s^2: conj(N)/s^2 = -2s - 2i(s - 1),
// conj(D)/s^2 = 3 is (is - 1) - i (2 - 3 is).
template <typename T>
__device__ inline cx<T> lp_inv_g(T s, T is, T iR, const double *vtab) {
    cx<T> y;
    if (val(s) > 20.0) {
        const T a = -2.0 * s, b = -2.0 * (s - T(1.0));
        const T c = (3.0 * is) * (is - T(1.0)), d = fma_(T(3.0), is, T(-2.0));
        const T id = rcp_(fma_(c, c, d * d));
        y = cx<T>{fma_(a, c, b * d) * id, fma_(b, c, -(a * d)) * id};
    } else if (small_alpha(s)) {  // W = 1/5 + a^2 (-1/175 + a^2 (2/7875 - 37 a^2 / 3031875))
        const T q = 2.0 * (s * s);
        const T ur = T(2.0 / 7875.0), ui = -(q * (37.0 / 3031875.0));
        const T nr = fma_(-q, ui, T(-1.0 / 175.0)), ni = q * ur;
        const T wr = fma_(-q, ni, T(1.0 / 5.0)), wi = q * nr;
        const T iw = rcp_(fma_(wr, wr, wi * wi));
        y = cx<T>{wr * iw, wi * iw};
    } else {
        y = v_table(s, vtab + kRpVFunc);
    }
    return cx<T>{y.re * iR, y.im * iR};
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
// FORCED: the step size is forced_dtp instead of the step-size control's
// (the replay of a recorded schedule; the rest of the sub-step is unchanged).
template <typename T, bool FORCED = false>
__device__ inline void rp_substep_g(RPLaneF<T> &L, const RPFastConst &k, const RPFastShared &s,
                                    double forced_dtp = 0.0) {
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
    if constexpr (FORCED) dtp = forced_dtp;
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
    // solveOmegaN at (PGn, Rn) (bubble_CUDA_kernel.cuh:263-289).  PGn cancels
    // from the iteration: with coef1 = PG0 / (PGn Rx^3) and coef2 = PGn rho,
    // Upsilon.re coef1 coef2 = (Upsilon.re / 3 gam) A and
    // Upsilon.im coef1 PGn / 4 = (-Upsilon.im / 3 gam) B with A = A0 / Rn^3,
    // B = B0 / Rn^3 (rp_fast_const), so the iteration depends on Rn alone.
    const T iRn2 = iRn * iRn;
    const T iRn3 = iRn2 * iRn;
    const T A = k.A0 * iRn3, B = k.B0 * iRn3;
    const T coef3 = 4.0 * iRn2;
    const T value1 = s.m2sig_rho * iRn;
    const T scale = iRn * s.irho;
    const T iscale = Rn * s.rho;
    T eta = A + fma_(T(-s.mu2), coef3, value1);
    T om, iom;
    omega_g(eta, scale, iscale, &om, &iom);
    const T ca = k.calpha * iRn;
    for (int it = 0; it < 3; it++) {
        const T x = ca * om;
        const T isa = rsqrt_(x);
        const cx<T> ups = upsilon_g(x * isa, isa, s.k3, s.vtab);
        const T mu_eff = fma_(ups.im * B, iom, T(s.mu));  // mu + ui PG / (4 omega)
        eta = fma_(ups.re, A, fma_(-(mu_eff * mu_eff), coef3, value1));
        omega_g(eta, scale, iscale, &om, &iom);
    }
    const T x = ca * om;
    const T isa = rsqrt_(x);
    const cx<T> Ln = lp_inv_g(x * isa, isa, iRn, s.vtab);  // Lp / |Lp|^2
    // solvePG (bubble_CUDA_kernel.cuh:322-340)
    const T Rbar = 0.5 * (Rp + Rn);
    const T iRbar = rcp_(Rbar);
    const T Rx_ = Rbar * k.iR0, Rpx = Rp * k.iR0, Rnx = Rn * k.iR0;
    const T Lnr = Ln.re, Lni = Ln.im;
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
        // Lp.re / |Lp|^2 = Lnr, Lp.im / |Lp|^2 = Lni
        const double dTdr_dt =
            fma(Lnr * k.T0_PG0R03 * dtL, fma(-0.5 * (PGp + PGn), Rbar * Rbar * Rbar, k.PG0R03),
                (Lni * k.T0_PG0R03 * iom) * fma(PGp, Rp * Rp * Rp, -(PGn * (Rn * Rn * Rn))));
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
