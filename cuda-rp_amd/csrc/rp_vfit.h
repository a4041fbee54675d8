// rp_vfit.h -- the thermal-impedance tables of the shortened-chain RP
// sub-step (rp_fast.h), built on the host when a simulation is created.
//
// With alpha = s (1 + i) (bubble_CUDA_kernel.cuh:263-320 always forms alpha
// this way), the sub-step needs, on 0 <= s <= 20,
//   V(s) = conj(1 / (1 + k3 U(s))),  U = (alpha coth alpha - 1) / alpha^2,
//          k3 = 3 (gam - 1)      (Upsilon = 3 gam conj(V), Upsilon at :296-311),
//   Y(s) = 1 / conj(W(s)),       W = 1 / (alpha coth alpha - 1) - 3 / alpha^2
//          (Lp = R W, :314-320; solvePG uses only Lp / |Lp|^2 = Y / R).
// V depends on the gas's gam, so the table is built per simulation rather
// than compiled in.  Both functions come from Lambert's continued fraction
// tanh a = a / (1 + a^2 / (3 + a^2 / (5 + ...))):
//   U = 1 / (3 + a^2 / (5 + a^2 / (7 + ...))),
//   W = 1 / (5 + a^2 / (7 + a^2 / (9 + ...))),
// evaluated bottom-up in x87 extended precision (64-bit significand).  Unlike
// the closed forms, these have no cancellation at small s (U(0) = 1/3,
// W(0) = 1/5).  64 levels converge to below 1e-40 for s <= 20 (the last
// denominator, 131 + a^2 / ..., is far beyond |alpha| <= 28.3); the
// backward evaluation is the stable direction (the tails are ratios of
// modified spherical Bessel functions, whose zeros lie on the imaginary
// axis, away from alpha's diagonal).
//
// On each interval k = [k/8, (k+1)/8] the degree-9 interpolant at the 10
// Chebyshev nodes (the singularities of U and W are at least pi/2 from the
// real s axis, 25 half-widths) is stored as monomial coefficients in
// t = 16 s - (2k + 1) in [-1, 1]: table[f][k][j][0] the real and [1] the
// imaginary part of coefficient j of function f (0: V, 1: Y).
#pragma once

#include <cmath>

namespace brp {

constexpr int kVPerUnit = 8;      // intervals per unit of s
constexpr int kVIntervals = 160;  // 0 <= s <= 20
constexpr int kVDegree = 9;
constexpr int kVCfLevels = 64;    // continued-fraction depth

struct RpCxl {
    long double re, im;
};

inline RpCxl rp_cxl_div(RpCxl a, RpCxl b) {
    const long double d = b.re * b.re + b.im * b.im;
    return RpCxl{(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
}

// 1 / (m + z / (m + 2 + z / (m + 4 + ...))) with z = alpha^2 = 2 i s^2:
// U(s) for m = 3, W(s) for m = 5.
inline RpCxl rp_impedance_cf(long double s, int m) {
    const RpCxl z{0.0L, 2.0L * s * s};
    RpCxl d{0.0L, 0.0L};
    for (int n = m + 2 * kVCfLevels; n > m; n -= 2) d = rp_cxl_div(z, RpCxl{n + d.re, d.im});
    return rp_cxl_div(RpCxl{1.0L, 0.0L}, RpCxl{m + d.re, d.im});
}

// V(s) and Y(s) above, for k3 = 3 (gam - 1).
inline RpCxl rp_v_value(long double s, double k3) {
    const RpCxl u = rp_impedance_cf(s, 3);
    const RpCxl v = rp_cxl_div(RpCxl{1.0L, 0.0L}, RpCxl{1.0L + k3 * u.re, k3 * u.im});
    return RpCxl{v.re, -v.im};
}
inline RpCxl rp_y_value(long double s) {
    const RpCxl w = rp_impedance_cf(s, 5);
    return rp_cxl_div(RpCxl{1.0L, 0.0L}, RpCxl{w.re, -w.im});
}

// Fills out[2][kVIntervals][kVDegree + 1][2] (layout above) for k3.
inline void rp_vtable_build(double k3, double *out) {
    constexpr int n = kVDegree + 1;
    const long double pi = 3.14159265358979323846264338327950288L;
    long double node[n], cheb[n][n], mono[n][n] = {};
    for (int j = 0; j < n; j++) node[j] = std::cos(pi * (j + 0.5L) / n);
    for (int m = 0; m < n; m++)
        for (int j = 0; j < n; j++) cheb[m][j] = std::cos(pi * m * (j + 0.5L) / n);  // T_m(node j)
    // monomial coefficients of T_m: T_0 = 1, T_1 = t, T_m = 2 t T_(m-1) - T_(m-2)
    mono[0][0] = 1.0L;
    mono[1][1] = 1.0L;
    for (int m = 2; m < n; m++)
        for (int i = 0; i < n; i++)
            mono[m][i] = (i ? 2.0L * mono[m - 1][i - 1] : 0.0L) - mono[m - 2][i];
    for (int k = 0; k < kVIntervals; k++) {
        RpCxl val[2][n];
        for (int j = 0; j < n; j++) {
            const long double s = (2.0L * k + 1.0L + node[j]) / (2.0L * kVPerUnit);
            val[0][j] = rp_v_value(s, k3);
            val[1][j] = rp_y_value(s);
        }
        for (int f = 0; f < 2; f++) {
            RpCxl a[n];  // Chebyshev coefficients of the interpolant
            for (int m = 0; m < n; m++) {
                long double re = 0.0L, im = 0.0L;
                for (int j = 0; j < n; j++) {
                    re += val[f][j].re * cheb[m][j];
                    im += val[f][j].im * cheb[m][j];
                }
                const long double w = (m ? 2.0L : 1.0L) / n;
                a[m] = RpCxl{w * re, w * im};
            }
            double *c = out + ((size_t)(f * kVIntervals + k) * n) * 2;
            for (int i = 0; i < n; i++) {
                long double re = 0.0L, im = 0.0L;
                for (int m = i; m < n; m++) {
                    re += a[m].re * mono[m][i];
                    im += a[m].im * mono[m][i];
                }
                c[2 * i] = (double)re;
                c[2 * i + 1] = (double)im;
            }
        }
    }
}

}  // namespace brp
