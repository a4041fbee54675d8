// rp_fast.h -- the RP sub-step with a short dependency chain, generic over the
// scalar type.
//
// Same model as rp_device.h (bubble_CUDA_kernel.cuh:227-340), rearranged for
// latency, since one collapsing bubble's thousands of sequential sub-steps
// bound a whole launch:
//   * alpha = s (1 + i) always (rp_alpha), so the thermal-impedance terms
//     are functions of s alone: Upsilon = 3 gam / (1 + k3 U) with
//     U = (alpha coth alpha - 1)/alpha^2, and Lp's W = 1/(alpha coth alpha -
//     1) - 3/alpha^2, of which solvePG needs only Lp/|Lp|^2 = Y/R,
//     Y = 1/conj(W) (U and Y are finite at s = 0, where the two terms of W
//     cancel).  For |alpha| <= 0.1 the reference's own Taylor series; up to
//     s = 20 one degree-9 polynomial per interval of width 1/8 of
//     V = conj(1 / (1 + k3 U)) and of Y (rp_vfit.h, built for the
//     simulation's gam and staged in LDS; error below 2e-16 of |V| and |Y|),
//     so that no division follows the table; for s > 20 coth alpha = 1 to
//     below an ulp, so both follow from alpha - 1 in closed form (collapsing
//     bubbles spend about half their sub-steps there);
//   * omega_N and 1/omega_N from one rsqrt of eta; c_alpha omega / R, its
//     square root and reciprocal square root from the same rsqrt and one
//     more;
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
// (rp_substep_g): 3 gam rho PG0 R0^3 and -3 gam PG0 R0^3 / 4; calpha_irho is
// c_alpha / rho.
__device__ inline RPFastConst rp_fast_const(double R0, double PG0, double calpha, double T0,
                                            const RPFastShared &s) {
    const double PG0R03 = PG0 * (R0 * R0 * R0);
    return RPFastConst{1.0 / R0,      PG0,    1.0 / PG0, calpha * s.irho, PG0R03, T0 / PG0R03,
                       (s.g3 * s.rho) * PG0R03, (-0.25 * s.g3) * PG0R03};
}

// Constant (scalar-loadable) memory.
typedef __attribute__((address_space(4))) const double cdouble;

// The shared constants from their device table (rp_table_values), read
// where the sub-step uses them rather than held in scalar registers across
// the sub-step loop; vtab is the kernel's LDS copy of the V and Y tables.
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

// Copies the V and Y tables (kRpVSize doubles from the device table) into the
// workgroup's LDS array; the caller synchronises the workgroup afterwards.
__device__ inline void rp_vtab_stage(const double *tab, double *lds) {
    for (int k = threadIdx.x; k < kRpVSize; k += blockDim.x) lds[k] = tab[kRpVOffset + k];
}

struct cx_d {
    double re, im;
};

// V or Y (ftab = that function's table in LDS) for 0 <= s <= 20: interval
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

// A branch condition of the sub-step.  UNI: the caller runs the sub-step on
// one lane alone (a dedicated wavefront), so the condition of the active
// lanes' ballot is that lane's and the branch is taken on the scalar unit,
// without saving and restoring the execution mask around each arm.
template <bool UNI>
__device__ inline bool cond(bool c) {
    if constexpr (UNI)
        return __builtin_amdgcn_ballot_w64(c) != 0ull;
    else
        return c;
}

// Upsilon(alpha) = 3 gam / (1 + k3 U) (bubble_CUDA_kernel.cuh:296-311) from
// s and is = 1/s, U = (alpha coth alpha - 1) / alpha^2, returned as
// V = conj(Upsilon) / (3 gam) = conj(1 / (1 + k3 U)): the caller's constants
// carry the factor 3 gam and the sign of the imaginary part.  V's table up to
// s = 20 (k3 is the simulation's; the table was built for it); beyond it
// coth alpha = 1 to below an ulp, so U = (alpha - 1) / alpha^2 =
// (is / 2) (1 + i (is - 1)).
template <typename T, bool UNI = false>
__device__ inline cx<T> upsilon_g(T s, T is, double k3, const double *vtab) {
    cx<T> v;
    const bool far = cond<UNI>(val(s) > 20.0), small = cond<UNI>(small_alpha(s));
    if (!far && !small) {  // the table
        v = v_table(s, vtab);
    } else {
        cx<T> u;
        if (far) {
            const T h = 0.5 * is;
            u = cx<T>{h, h * (is - T(1.0))};
        } else {  // 1/3 + a^2 (-1/45 + a^2 (2/945 - a^2/4725)), a^2 = i q
            const T q = 2.0 * (s * s);
            const T ur = T(2.0 / 945.0), ui = -(q * (1.0 / 4725.0));
            const T nr = fma_(-q, ui, T(-1.0 / 45.0)), ni = q * ur;
            u = cx<T>{fma_(-q, ni, T(1.0 / 3.0)), q * nr};
        }
        const T dr = fma_(T(k3), u.re, T(1.0)), di = k3 * u.im;
        const T f = rcp_(fma_(dr, dr, di * di));
        v = cx<T>{dr * f, di * f};
    }
    return v;
}

// solveLp (bubble_CUDA_kernel.cuh:314-320) as solvePG uses it
// (bubble_CUDA_kernel.cuh:328-329): Lp / |Lp|^2 = Y / R with Lp = R W,
// W = 1 / (alpha coth alpha - 1) - 3 / alpha^2 and Y = 1 / conj(W); iR = 1/R.
// The table up to s = 20; beyond it W = 1/(alpha - 1) - 3/alpha^2, so
// Y = conj(N) / conj(D) with N = alpha^2 (alpha - 1), D = alpha^2 - 3 alpha + 3,
// both divided by s^2: conj(N)/s^2 = -2s - 2i(s - 1),
// conj(D)/s^2 = 3 is (is - 1) - i (2 - 3 is).
template <typename T, bool UNI = false>
__device__ inline cx<T> lp_inv_g(T s, T is, T iR, const double *vtab) {
    cx<T> y;
    const bool far = cond<UNI>(val(s) > 20.0), small = cond<UNI>(small_alpha(s));
    if (far) {
        const T a = -2.0 * s, b = -2.0 * (s - T(1.0));
        const T c = (3.0 * is) * (is - T(1.0)), d = fma_(T(3.0), is, T(-2.0));
        const T id = rcp_(fma_(c, c, d * d));
        y = cx<T>{fma_(a, c, b * d) * id, fma_(b, c, -(a * d)) * id};
    } else if (small) {  // W = 1/5 + a^2 (-1/175 + a^2 (2/7875 - 37 a^2 / 3031875))
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

// omega = sqrt(max(eta, floor)) / (rho R) enters only through
// x = c_alpha omega / R, and 1 / omega; both from one rsqrt, with
// xscale = c_alpha / (rho R^2) and iscale = rho R.
template <typename T>
__device__ inline void omega_g(T eta, T xscale, T iscale, T *x, T *iom) {
    const T e = val(eta) >= 1.0e-6 * kDblEps ? eta : T(1.0e-6 * kDblEps);
    const T r = rsqrt_(e);
    *x = (e * r) * xscale;
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

// The step-size control of solveRayleighPlesset (bubble_CUDA_kernel.cuh:
// 235-243): the previous step times 1.01, capped by dt0, by 1 % of R over
// |dR/dt| and by 10 % of R over sqrt|R d2R/dt2| (each limit only where its
// denominator exceeds DBL_EPSILON).  The two limits share one reciprocal
// square root: min(0.01 R / |dR/dt|, 0.1 R / sqrt|R d2R/dt2|) =
// R / sqrt(max(1e4 (dR/dt)^2, 100 |R d2R/dt2|)).
template <bool UNI = false>
__device__ inline double rp_dtp(double dtm, double Rn, double d1R, double Rnd2R,
                                const RPFastShared &s) {
    double dtp = fmin(1.01 * dtm, s.dt0);
    const double ad = fabs(d1R), aa = fabs(Rnd2R);
    const double m = fmax(ad > kDblEps ? 1.0e4 * (ad * ad) : 0.0, aa > kDblEps ? 100.0 * aa : 0.0);
    if (__builtin_expect(cond<UNI>(m > 0.0), 1)) dtp = fmin(dtp, Rn * drsqrt(m));
    return dtp;
}

// solveOmegaN at (PGn, Rn) (bubble_CUDA_kernel.cuh:263-289) and solveLp as
// solvePG uses it (Lp / |Lp|^2), with iRn = 1/Rn: *Ln = Lp / |Lp|^2 and
// *iom = 1/omega_N.  PGn cancels from the iteration: with
// coef1 = PG0 / (PGn Rx^3) and coef2 = PGn rho,
// Upsilon.re coef1 coef2 = (Upsilon.re / 3 gam) A and
// Upsilon.im coef1 PGn / 4 = (-Upsilon.im / 3 gam) B with A = A0 / Rn^3,
// B = B0 / Rn^3 (rp_fast_const), so both depend on Rn alone.
template <typename T, bool UNI = false>
__device__ inline void omega_lp_g(T Rn, T iRn, const RPFastConst &k, const RPFastShared &s,
                                  cx<T> *Ln, T *iom_out) {
    const T iRn2 = iRn * iRn;
    const T iRn3 = iRn2 * iRn;
    const T A = k.A0 * iRn3, B = k.B0 * iRn3;
    const T coef3 = 4.0 * iRn2;
    const T value1 = s.m2sig_rho * iRn;
    const T xscale = k.calpha_irho * iRn2;  // c_alpha / (rho Rn^2)
    const T iscale = Rn * s.rho;
    T eta = A + fma_(T(-s.mu2), coef3, value1);
    T x, iom;  // x = c_alpha omega / Rn = s^2
    omega_g(eta, xscale, iscale, &x, &iom);
#pragma unroll
    for (int it = 0; it < 3; it++) {
        const T isa = rsqrt_(x);
        const cx<T> ups = upsilon_g<T, UNI>(x * isa, isa, s.k3, s.vtab);
        const T mu_eff = fma_(ups.im * B, iom, T(s.mu));  // mu + ui PG / (4 omega)
        eta = fma_(ups.re, A, fma_(-(mu_eff * mu_eff), coef3, value1));
        omega_g(eta, xscale, iscale, &x, &iom);
    }
    const T isa = rsqrt_(x);
    *Ln = lp_inv_g<T, UNI>(x * isa, isa, iRn, s.vtab);
    *iom_out = iom;
}

// One pass of the sub-step loop body (bubble_CUDA_kernel.cuh:508-538).
// FORCED: the step size is forced_dtp instead of the step-size control's
// (the replay of a recorded schedule; the rest of the sub-step is unchanged).
// HEAT: accumulate the heat and viscous dissipation of Q_B (T = double).
// UNI: one lane runs the sub-step alone (branches on the scalar unit, cond).
template <typename T, bool FORCED = false, bool HEAT = true, bool UNI = false>
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
    double dtp = rp_dtp<UNI>(dtm, val(Rn), val(d1R), val(Rnd2R), s);
    if constexpr (FORCED) dtp = forced_dtp;
    const double idtm = drcp(dtm);
    const double r = dtp * idtm;
    const T acc = (dtp * Rnd2R) * iRn;  // (dtp / Rn) Rnd2R
    const T Rp = fma_(acc, T(0.5 * (dtp + dtm)), fma_(T(1.0 + r), Rn, -r * Rm));
    const double dm2 = dtm * dtm, dp2 = dtp * dtp;
    const double iden = drcp(dtm * dtp * (dtm + dtp));
    d1R = fma_(fma_(T(dm2), Rp, fma_(T(dp2 - dm2), Rn, -dp2 * Rm)), T(iden), acc);
    const double dtL = dtp;
    if (__builtin_expect(cond<UNI>(dtp >= L.remain), 0)) {  // last sub-step: R_t at the sync time
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
    cx<T> Ln;  // Lp / |Lp|^2
    T iom;
    omega_lp_g<T, UNI>(Rn, iRn, k, s, &Ln, &iom);
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
    if constexpr (HEAT && std::is_same<T, double>::value) {
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

// ---- reverse mode: adjoints of the RP map of one Picard iteration ----------
// The lane state before a sub-step, as the reverse pass restores it.
struct RPPre {
    double Rp, Rn, d1R, PG, dtL, remain, time;
};

// Reverse of rp_substep_g<double> at the frozen schedule (the step size, the
// previous one and the time held at the forward's values), for NOUT adjoint
// vectors at once.  On entry l[.][o] holds the adjoints of the sub-step's
// outputs (Rp, Rn, d1R, PG) and lRt[o] that of R_t, used when the sub-step
// is the integration's final one (`final`); on exit l[.][o] holds the
// adjoints of its inputs, and lPC[.][o] has the adjoints of the pressure
// coefficients PC0, PC1, PC2 added.  The Omega iteration and Lp depend on Rn
// alone (omega_lp_g), so their derivative is one tangent along Rn;
// solveRayleighPlesset and solvePG are reversed statement by statement.
template <int NOUT>
__device__ inline void rp_substep_rev(const RPPre &P, bool final, const double (&lRt)[NOUT],
                                      double PC0, double PC1, double PC2, const RPFastConst &k,
                                      const RPFastShared &s, double (&l)[4][NOUT],
                                      double (&lPC)[3][NOUT]) {
    // the sub-step's values, as rp_substep_g<double> computes them
    const double PGn = P.PG, t = P.time, dtm = P.dtL;
    const double PL = fma(PC2, t * t, fma(PC1, t, PC0));
    const double Rm = P.Rn, Rn = P.Rp, d1R = P.d1R;
    const double iRn = drcp(Rn);
    const double sva = fma(s.four_mu, d1R, s.two_sig);
    const double Rnd2R = fma(-1.5 * d1R, d1R, (PGn - PL - iRn * sva) * s.irho);
    const double dtp = rp_dtp(dtm, Rn, d1R, Rnd2R, s);
    const double idtm = drcp(dtm);
    const double r = dtp * idtm, h1 = 0.5 * (dtp + dtm);
    const double acc = (dtp * Rnd2R) * iRn;
    const double Rp = fma(acc, h1, fma(1.0 + r, Rn, -r * Rm));
    const double dm2 = dtm * dtm, dp2 = dtp * dtp;
    const double iden = drcp(dtm * dtp * (dtm + dtp));
    cx<dd> Ld;
    dd iomd;
    const dd Rnd(Rn, 1.0);
    omega_lp_g<dd>(Rnd, rcp_(Rnd), k, s, &Ld, &iomd);
    const double Lnr = Ld.re.v, Lni = Ld.im.v, iom = iomd.v;
    const double Rbar = 0.5 * (Rp + Rn), iRbar = drcp(Rbar);
    const double Rx = Rbar * k.iR0, Rpx = Rp * k.iR0, Rnx = Rn * k.iR0;
    const double Rx2 = Rx * Rx, Rpx2 = Rpx * Rpx, Rnx2 = Rnx * Rnx;
    const double cdt = (s.C0 * iRbar) * k.iPG0, c0 = dtp * cdt;
    const double hg = 1.5 * s.gam, hx3 = 0.5 * (Rx2 * Rx);
    const double c3 = fma(c0 * Lnr, hx3, (hg * (Rp - Rn)) * iRbar);
    const double q = (cdt * Lni) * iom;
    const double c1 = fma(-q, Rpx2 * Rpx, 1.0 + c3), c2 = fma(-q, Rnx2 * Rnx, 1.0 - c3);
    const double ic1 = drcp(c1);
    const double PGp = fma(c2, PGn, (c0 * Lnr) * k.PG0) * ic1;
    // R_t of the final sub-step: at the sync time, or Rp when the sub-step
    // cap ended the integration
    const bool sync = final && dtp >= P.remain;
    const double d = P.remain, r2 = d * idtm, h2 = 0.5 * (d + dtm);
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
        // solvePG: PGp = (c2 PGn + c0 Lnr PG0) / c1
        const double lN = l[3][o] * ic1, lc1 = -l[3][o] * PGp * ic1, lc2 = lN * PGn;
        double lPGn = lN * c2;
        double lc0 = lN * Lnr * k.PG0, lLnr = lN * c0 * k.PG0;
        const double lc3 = lc1 - lc2;
        const double lq = -(lc1 * (Rpx2 * Rpx) + lc2 * (Rnx2 * Rnx));
        const double lRpx = -3.0 * lc1 * q * Rpx2, lRnx = -3.0 * lc2 * q * Rnx2;
        double lcdt = lq * Lni * iom;
        const double lLni = lq * cdt * iom, liom = lq * cdt * Lni;
        lc0 += lc3 * Lnr * hx3;
        lLnr += lc3 * c0 * hx3;
        const double lRx = 1.5 * lc3 * c0 * Lnr * Rx2;
        double liRbar = lc3 * hg * (Rp - Rn);
        double lRp = lc3 * hg * iRbar, lRn = -lc3 * hg * iRbar;
        lcdt += lc0 * dtp;
        liRbar += lcdt * s.C0 * k.iPG0;
        const double lRbar = lRx * k.iR0 - liRbar * iRbar * iRbar;
        lRp += lRpx * k.iR0 + 0.5 * lRbar;
        lRn += lRnx * k.iR0 + 0.5 * lRbar;
        // the Omega iteration and Lp
        lRn += lLnr * Ld.re.d + lLni * Ld.im.d + liom * iomd.d;
        // solveRayleighPlesset: d1R' = (dm2 Rp + (dp2 - dm2) Rn - dp2 Rm) iden
        // + acc, Rp = acc h1 + (1 + r) Rn - r Rm, acc = dtp Rnd2R / Rn
        const double lS = l[2][o] * iden;
        double lacc = l[2][o];
        lRp += l[0][o] + lS * dm2;
        lRn += l[1][o] + lS * (dp2 - dm2);
        double lRm = -lS * dp2;
        double lRnd2R = 0.0, liRn = 0.0;
        if (sync) {
            lRnd2R += lRt[o] * d * iRn * h2;
            liRn += lRt[o] * d * Rnd2R * h2;
            lRn += lRt[o] * (1.0 + r2);
            lRm -= lRt[o] * r2;
        } else if (final) {
            lRp += lRt[o];
        }
        lacc += lRp * h1;
        lRn += lRp * (1.0 + r);
        lRm -= lRp * r;
        lRnd2R += lacc * dtp * iRn;
        liRn += lacc * dtp * Rnd2R;
        // Rnd2R = -1.5 d1R^2 + (PGn - PL - sva / Rn) / rho
        const double lQ = lRnd2R * s.irho;
        const double ld1R = -3.0 * lRnd2R * d1R - lQ * iRn * s.four_mu;
        lPGn += lQ;
        liRn -= lQ * sva;
        lRn -= liRn * iRn * iRn;
        lPC[0][o] -= lQ;
        lPC[1][o] -= lQ * t;
        lPC[2][o] -= lQ * (t * t);
        l[0][o] = lRn;  // the input Rp is this sub-step's Rn
        l[1][o] = lRm;  // the input Rn is its Rm
        l[2][o] = ld1R;
        l[3][o] = lPGn;
    }
}

// Per-lane storage of the reverse pass: `ck` records every m-th lane state
// (c of them), `blk` the m states of the block being reversed; each record is
// seven doubles, record r field f at base[(r * 7 + f) * stride].
struct RPRevBuf {
    double *ck, *blk;
    size_t stride;
    int m, c;
};
__device__ inline void rev_put(double *p, size_t st, const RPLaneF<double> &L) {
    p[0] = L.Rp;
    p[st] = L.Rn;
    p[2 * st] = L.d1R;
    p[3 * st] = L.PG;
    p[4 * st] = L.dtL;
    p[5 * st] = L.remain;
    p[6 * st] = L.time;
}
__device__ inline RPPre rev_get(const double *p, size_t st) {
    return RPPre{p[0], p[st], p[2 * st], p[3 * st], p[4 * st], p[5 * st], p[6 * st]};
}

// Jacobian rows of the RP map of one Picard iteration on one lane:
// J[(7 o + d) * nb] = dOut_o / dIn_d for the outputs o < nout <= NOUT (R_t,
// then R', Rn', d1R', PG') and the inputs d (R, Rn, d1R, PG, PL_p, PL_n,
// PL_m).  The forward pass repeats the forward's sub-steps (their count must
// be K, the forward's; returns false otherwise) and keeps every lane state,
// or every m-th when K > m, whose blocks are integrated again from there
// during the reverse pass, which carries all NOUT adjoint vectors at once.
template <int NOUT>
__device__ inline bool rp_jac_rev(const RPLaneF<double> &L0, int K, int nout, int max_substeps,
                                  double dt_mix, const RPFastConst &k, const RPFastShared &s,
                                  const RPRevBuf &B, double *J, size_t nb) {
    const bool two = K > B.m;
    const int cap = two ? B.c * B.m : B.m;
    RPLaneF<double> L = L0;
    while (L.remain > 0.0 && L.counter < max_substeps && L.counter < cap) {
        const int kk = L.counter;
        if (!two)
            rev_put(B.blk + (size_t)kk * 7 * B.stride, B.stride, L);
        else if (kk % B.m == 0)
            rev_put(B.ck + (size_t)(kk / B.m) * 7 * B.stride, B.stride, L);
        rp_substep_g<double, false, false>(L, k, s);
    }
    if (L.counter != K || (L.remain > 0.0 && L.counter < max_substeps)) return false;
    double l[4][NOUT], lPC[3][NOUT], lRt[NOUT];
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
#pragma unroll
        for (int j = 0; j < 4; j++) l[j][o] = o == j + 1 ? 1.0 : 0.0;
#pragma unroll
        for (int j = 0; j < 3; j++) lPC[j][o] = 0.0;
        lRt[o] = o == 0 ? 1.0 : 0.0;
    }
    for (int b = two ? (K - 1) / B.m : 0; b >= 0; b--) {
        const int k0 = two ? b * B.m : 0, k1 = two ? min(K, k0 + B.m) : K;
        if (two) {  // the block's states again, from its checkpoint
            RPLaneF<double> R = L0;
            const RPPre c = rev_get(B.ck + (size_t)b * 7 * B.stride, B.stride);
            R.Rp = c.Rp;
            R.Rn = c.Rn;
            R.d1R = c.d1R;
            R.PG = c.PG;
            R.dtL = c.dtL;
            R.remain = c.remain;
            R.time = c.time;
            R.counter = k0;
            for (int kk = k0; kk < k1; kk++) {
                rev_put(B.blk + (size_t)(kk - k0) * 7 * B.stride, B.stride, R);
                rp_substep_g<double, false, false>(R, k, s);
            }
        }
        for (int kk = k1 - 1; kk >= k0; kk--)
            rp_substep_rev<NOUT>(rev_get(B.blk + (size_t)(kk - k0) * 7 * B.stride, B.stride),
                                 kk == K - 1, lRt, L0.PC0, L0.PC1, L0.PC2, k, s, l, lPC);
    }
    // PC0 = PL_n + PL0, PC1 = (PL_p - PL_m) / (2 dt), PC2 = (PL_p + PL_m - 2 PL_n) / (2 dt^2)
    const double idt = 1.0 / dt_mix;
#pragma unroll
    for (int o = 0; o < NOUT; o++) {
        if (o >= nout) break;
        double *Jo = J + (size_t)(7 * o) * nb;
        Jo[0] = l[0][o];
        Jo[nb] = l[1][o];
        Jo[2 * nb] = l[2][o];
        Jo[3 * nb] = l[3][o];
        Jo[4 * nb] = 0.5 * idt * (lPC[1][o] + lPC[2][o] * idt);
        Jo[5 * nb] = lPC[0][o] - lPC[2][o] * idt * idt;
        Jo[6 * nb] = 0.5 * idt * (lPC[2][o] * idt - lPC[1][o]);
    }
    return true;
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
