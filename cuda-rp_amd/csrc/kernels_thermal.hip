// kernels_thermal.hip -- the one-way temperature solve (calculate_temperature,
// bubble_CUDA_kernel.cuh:1347-1380) for gfx950.
//
// The temperature rise T is FP64 in both precision modes (it accumulates
// increments many orders below its value).  The reference's k_m, Ex, Ey and
// C_pm arrays are not stored: the conductivities, face fluxes and heat
// capacity are recomputed from T and f_g where a cell needs them, with the
// reference's expressions, so one pass computes every cell's increment and a
// second adds it (the increment of a cell reads its neighbours' old T).
// Bubble heat is scattered with FP64 atomics, so T is reproducible to the
// rounding of the order of those additions.
#include "brp_internal.h"

#include <math.h>

#include <algorithm>

namespace brp {
namespace {

constexpr int kTX = 64, kTY = 4;

// Water properties with the reference's single-precision literals
// (bubble_CUDA_kernel.cuh:174-193).
__device__ inline double k_water(double T) {
    return 0.76760 + ((double)7.5390e-3f * T) - (double)9.8250e-6f * (T * T);
}
__device__ inline double cp_water(double T) {
    const double H2O = (double)18.0e-3f;
    return (917.5 - (double)10.1016f * T + (double)0.0454134f * (T * T) -
            9.07517e-5 * (T * T * T) + (double)6.80700e-8f * (T * T * T * T)) /
           H2O;
}

__device__ inline double fgv(const int64_t *f, int64_t o) { return (double)f[o] * kFgInvScale; }

// smooth_delta_x / smooth_delta_y (bubble_CUDA_kernel.cuh:214-222)
__device__ inline double delta_t(double coef_x, double coef_k, int pos, double h, double b) {
#pragma clang fp contract(off)
    return coef_x * (1.0 + cos(coef_k * (((double)pos - 0.5) * h - b)));
}

__global__ void k_zero_window(double *q, Window w, int pitch) {
    const int i = w.i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = w.j0 + (int)blockIdx.y;
    if (i > w.i1 || j > w.j1) return;
    q[frame_off(i, j, pitch)] = 0.0;
}

// BubbleHeatKernel (bubble_CUDA_kernel.cuh:972-1006).  Bubbles are
// cell-sorted, so a workgroup's 256 stencils cover a small patch of cells:
// the contributions are summed in an LDS tile over the workgroup's bounding
// box and flushed with one global atomic per touched cell (a patch larger
// than the tile falls back to global atomics per contribution).  The sums
// are FP64 in either path; only their order differs.
constexpr int kHeatTile = 4096;  // cells (32 KB of LDS)
constexpr int kHeatBand = 8;     // the largest smooth_delta_band (checked at set-up)
__global__ __launch_bounds__(256) void k_bubble_heat(GridParams g, BubbleParams bp, BubbleArrays b,
                                                      double *heat) {
    __shared__ double tile[kHeatTile];
    __shared__ int box[4];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = idx < b.n;
    double x = 0.0, y = 0.0, Q = 0.0;
    int ibnx = 0, ibny = 0;
    if (valid) {
        x = b.x[idx];
        y = b.y[idx];
        {
#pragma clang fp contract(off)
            Q = bp.cylindrical ? b.n_B[idx] * b.Q_B[idx] * g.rdx * g.rdy / x
                               : b.n_B[idx] * b.Q_B[idx] * g.rdx * g.rdy;
        }
        double v;
        {
#pragma clang fp contract(off)
            v = floor(x * g.rdx + 0.5);
        }
        ibnx = (v >= -1.0e9 && v <= 1.0e9) ? (int)v : -1000000000;
        {
#pragma clang fp contract(off)
            v = floor(y * g.rdy + 0.5);
        }
        ibny = (v >= -1.0e9 && v <= 1.0e9) ? (int)v : -1000000000;
    }
    // this bubble's stencil clipped to the frame's cells [0, N+1] x [0, M+1]
    const int i0 = max(ibnx + bp.mbs, 0), i1 = min(ibnx + bp.mbe, g.N + 1);
    const int j0 = max(ibny + bp.mbs, 0), j1 = min(ibny + bp.mbe, g.M + 1);
    const bool any = valid && i0 <= i1 && j0 <= j1;
    if (threadIdx.x == 0) {
        box[0] = 0x7fffffff;
        box[1] = -1;
        box[2] = 0x7fffffff;
        box[3] = -1;
    }
    __syncthreads();
    if (any) {
        atomicMin(&box[0], i0);
        atomicMax(&box[1], i1);
        atomicMin(&box[2], j0);
        atomicMax(&box[3], j1);
    }
    __syncthreads();
    const int bx0 = box[0], by0 = box[2];
    const int bw = box[1] - bx0 + 1, bh = box[3] - by0 + 1;
    if (bw <= 0 || bh <= 0) return;  // workgroup-uniform: no stencil in the frame
    const bool binned = (int64_t)bw * bh <= kHeatTile;  // workgroup-uniform
    if (binned) {
        for (int k = threadIdx.x; k < bw * bh; k += blockDim.x) tile[k] = 0.0;
        __syncthreads();
    }
    if (any) {
        double Dy[kHeatBand];
        for (int j = j0; j <= j1; j++) Dy[j - j0] = delta_t(bp.dcx, bp.dcz, j, g.dy, y);
        for (int i = i0; i <= i1; i++) {
            const double Dx = delta_t(bp.dcx, bp.dcy, i, g.dx, x);
            for (int j = j0; j <= j1; j++) {
                double c;
                {
#pragma clang fp contract(off)
                    c = Q * Dx * Dy[j - j0];
                }
                if (binned)
                    atomicAdd(&tile[(j - by0) * bw + (i - bx0)], c);
                else
                    atomicAdd(&heat[frame_off(i, j, g.pitch)], c);
            }
        }
    }
    if (binned) {
        __syncthreads();
        for (int k = threadIdx.x; k < bw * bh; k += blockDim.x) {
            const double v = tile[k];
            if (v != 0.0) atomicAdd(&heat[frame_off(bx0 + k % bw, by0 + k / bw, g.pitch)], v);
        }
    }
}

// MixtureKMKernel + MixtureEnergyKernel + MixtureTemperatureKernel
// (bubble_CUDA_kernel.cuh:955-969, 1009-1061) for interior cell (i, j): the
// increment dT, with rho_m = (1 - f_gn) rho_inf and C_pm = c_p(T + T_inf) of
// the previous step (rho_inf and c_p(T_inf) before the first).
template <typename F>
__global__ __launch_bounds__(256) void k_temperature_increment(
    GridParams g, double T_inf, const double *__restrict__ T, const int64_t *__restrict__ fg,
    const int64_t *__restrict__ fgn, int first, const F *__restrict__ px,
    const F *__restrict__ py, const F *__restrict__ pnx, const F *__restrict__ pny,
    const double *__restrict__ heat, double *__restrict__ dT) {
#pragma clang fp contract(off)
    const int i = 1 + (int)(blockIdx.x * kTX + threadIdx.x);
    const int j = 1 + (int)(blockIdx.y * kTY + threadIdx.y);
    if (i > g.N || j > g.M) return;
    const int P = g.pitch;
    auto km = [&](int a, int c) {
        const int64_t q = frame_off(a, c, P);
        const double f = fgv(fg, q);
        return k_water(T[q] + T_inf) * (1.0 - 3.0 * f / (2.0 + f));
    };
    auto tv = [&](int a, int c) { return T[frame_off(a, c, P)]; };
    const double kc = km(i, j), kw = km(i - 1, j), ke = km(i + 1, j), ks = km(i, j - 1),
                 kn = km(i, j + 1);
    const double Tc = tv(i, j);
    const double Ex_e = 2.0 * kc * ke / (kc + ke) * ((-Tc + tv(i + 1, j)) * g.rdx);
    const double Ex_w = 2.0 * kw * kc / (kw + kc) * ((-tv(i - 1, j) + Tc) * g.rdx);
    const double Ey_n = 2.0 * kc * kn / (kc + kn) * ((-Tc + tv(i, j + 1)) * g.rdy);
    const double Ey_s = 2.0 * ks * kc / (ks + kc) * ((-tv(i, j - 1) + Tc) * g.rdy);
    const int64_t o = frame_off(i, j, P);
    double s0 = ((double)px[o] + (double)py[o] - (double)pnx[o] - (double)pny[o]) / g.dt;
    s0 = 2.0 * (0.22) * 0.115129255 / (g.rho_inf * g.cs_inf * g.omega * g.omega) * s0 * s0;
    const double Q = heat[o] + s0;
    const double s1 = (Ex_e - Ex_w) * g.rdx;
    const double s2 = (Ey_n - Ey_s) * g.rdy;
    const double rho_m = first ? g.rho_inf : (1.0 - fgv(fgn, o)) * g.rho_inf;
    const double C_pm = cp_water(Tc + T_inf);
    dT[o] = g.dt * (s1 + s2 + Q) / (rho_m * C_pm);
}

__global__ void k_temperature_add(GridParams g, double *T, const double *dT) {
    const int i = 1 + (int)(blockIdx.x * kTX + threadIdx.x);
    const int j = 1 + (int)(blockIdx.y * kTY + threadIdx.y);
    if (i > g.N || j > g.M) return;
    const int64_t o = frame_off(i, j, g.pitch);
    T[o] += dT[o];
}

// MixtureBoundaryTemperatureKernel (bubble_CUDA_kernel.cuh:1064-1102), its
// four parts in program order in one workgroup.  The y1 part starts at row
// Y, so row Y takes the ghost row Y + 1, as in the reference.
__global__ __launch_bounds__(1024) void k_temperature_ghosts(GridParams g, double *T) {
    const int N = g.N, M = g.M, P = g.pitch;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (!g.pml_x0)
        for (int j = 1 + tid; j <= M; j += nt) T[frame_off(0, j, P)] = T[frame_off(1, j, P)];
    __syncthreads();
    if (!g.pml_x1)
        for (int j = tid; j <= M + 1; j += nt) T[frame_off(N + 1, j, P)] = T[frame_off(N, j, P)];
    __syncthreads();
    if (!g.pml_y0)
        for (int i = tid; i <= N + 1; i += nt) T[frame_off(i, 0, P)] = T[frame_off(i, 1, P)];
    __syncthreads();
    if (!g.pml_y1)
        for (int i = tid; i <= N + 1; i += nt) {
            T[frame_off(i, M, P)] = T[frame_off(i, M + 1, P)];
            T[frame_off(i, M + 1, P)] = T[frame_off(i, M, P)];
        }
}

// determine_focal_point (bubble_CUDA_kernel.cuh:1134-1166) on the reference's
// Work array after calculate_temperature: the temperature increment in the
// interior, the bubble heat deposited in the ghost cells.  The reference's
// reduction functors are not in its tree; here the focal point is the cell
// of the largest value (the first in row-major order on ties) at
// (x, y) = (i dx, j dy), the coordinates of the reference's output files.
__device__ inline void better(double &v, int &k, double v2, int k2) {
    if (v2 > v || (v2 == v && k2 < k)) {
        v = v2;
        k = k2;
    }
}

__global__ __launch_bounds__(256) void k_focal_blocks(GridParams g, const double *dT,
                                                      const double *heat, StepScalars *sc) {
    const int W = g.N + 2, n = W * (g.M + 2);
    double v = -INFINITY;
    int k = 0x7fffffff;
    for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < n; q += gridDim.x * blockDim.x) {
        const int i = q % W, j = q / W;
        const int64_t o = frame_off(i, j, g.pitch);
        const bool in = i >= 1 && i <= g.N && j >= 1 && j <= g.M;
        better(v, k, in ? dT[o] : (heat ? heat[o] : 0.0), q);
    }
    __shared__ double sv[256];
    __shared__ int sk[256];
    sv[threadIdx.x] = v;
    sk[threadIdx.x] = k;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) better(sv[threadIdx.x], sk[threadIdx.x], sv[threadIdx.x + s], sk[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        sc->focal_v[blockIdx.x] = sv[0];
        sc->focal_i[blockIdx.x] = sk[0];
    }
}

__global__ __launch_bounds__(256) void k_focal_final(GridParams g, int nblk, StepScalars *sc) {
    __shared__ double sv[256];
    __shared__ int sk[256];
    const int t = threadIdx.x;
    sv[t] = t < nblk ? sc->focal_v[t] : -INFINITY;
    sk[t] = t < nblk ? sc->focal_i[t] : 0x7fffffff;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) better(sv[t], sk[t], sv[t + s], sk[t + s]);
        __syncthreads();
    }
    if (t == 0) {
        const int W = g.N + 2;
        const int k = sk[0] == 0x7fffffff ? 0 : sk[0];
        sc->focal[0] = (double)(k % W) * g.dx;
        sc->focal[1] = (double)(k / W) * g.dy;
    }
}

}  // namespace

void launch_focal_point(hipStream_t s, const GridParams &g, const double *dT, const double *heat,
                        StepScalars *sc) {
    const int n = (g.N + 2) * (g.M + 2);
    const int nblk = std::min(256, (n + 255) / 256);
    hipLaunchKernelGGL(k_focal_blocks, dim3(nblk), dim3(256), 0, s, g, dT, heat, sc);
    hipLaunchKernelGGL(k_focal_final, dim3(1), dim3(256), 0, s, g, nblk, sc);
}

template <typename F>
void launch_temperature(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                        const BubbleArrays &b, bool bubbles, Window w, double T_inf, double *T,
                        double *dT, double *heat, const int64_t *fg, const int64_t *fgn,
                        int first, const F *px, const F *py, const F *pnx, const F *pny) {
    if (bubbles && w.i0 <= w.i1 && w.j0 <= w.j1) {
        hipLaunchKernelGGL(k_zero_window, dim3((w.i1 - w.i0 + 256) / 256, w.j1 - w.j0 + 1),
                           dim3(256), 0, s, heat, w, g.pitch);
        if (b.n > 0)
            hipLaunchKernelGGL(k_bubble_heat, dim3((b.n + 255) / 256), dim3(256), 0, s, g, bp, b,
                               heat);
    }
    const dim3 blk(kTX, kTY), grid((g.N + kTX - 1) / kTX, (g.M + kTY - 1) / kTY);
    hipLaunchKernelGGL(k_temperature_increment<F>, grid, blk, 0, s, g, T_inf, T, fg, fgn, first,
                       px, py, pnx, pny, heat, dT);
    hipLaunchKernelGGL(k_temperature_add, grid, blk, 0, s, g, T, dT);
    hipLaunchKernelGGL(k_temperature_ghosts, dim3(1), dim3(1024), 0, s, g, T);
}

template void launch_temperature<float>(hipStream_t, const GridParams &, const BubbleParams &,
                                        const BubbleArrays &, bool, Window, double, double *,
                                        double *, double *, const int64_t *, const int64_t *, int,
                                        const float *, const float *, const float *,
                                        const float *);
template void launch_temperature<double>(hipStream_t, const GridParams &, const BubbleParams &,
                                         const BubbleArrays &, bool, Window, double, double *,
                                         double *, double *, const int64_t *, const int64_t *,
                                         int, const double *, const double *, const double *,
                                         const double *);

}  // namespace brp
