// kernels_grid.hip -- staggered-grid FDTD kernels for gfx950.
//
// All grid kernels are HBM-bandwidth bound stencils (no MFMA).  A block is
// 64 x 4 threads: one wavefront per frame row segment of 64 consecutive cells
// (one aligned 256/512-byte line), four rows per block, RB rows per block
// walked in steps of four.  The precision is a template parameter F (double
// for the FP64 mode, float for the mixed mode); the void fraction is always
// int64 fixed point (value * 2^60) so its sums are exact and order-free.
#include "brp_internal.h"

#include <math.h>

namespace brp {
namespace {

constexpr int kBX = 64;
constexpr int kBY = 4;
constexpr int kRB = 16;  // rows per block

__device__ inline double fg_to_d(int64_t v) { return (double)v * kFgInvScale; }

// Max-reduce a non-negative double over the block and fold it into *slot with
// one 64-bit atomicMax of its bit pattern (non-negative IEEE doubles order as
// unsigned integers).  Skips the atomic when the slot already holds a larger
// value, so most blocks issue none.
__device__ inline void block_max_to(double v, unsigned long long *slot) {
    __shared__ double red[kBX * kBY / 64];
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int tid = threadIdx.y * kBX + threadIdx.x;
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double m = red[0];
        for (int k = 1; k < kBX * kBY / 64; k++) m = fmax(m, red[k]);
        unsigned long long bits = (unsigned long long)__double_as_longlong(m);
        if (bits > __hip_atomic_load(slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(slot, bits);
    }
}

template <typename F>
__device__ inline F rho_m_at(const GridParams &g, const int64_t *fgn, Window w, int first,
                             int i, int j) {
    // MixturePropertiesKernel (bubble_CUDA_kernel.cuh:1122): rho_m = (1 - f_g) rho_inf,
    // evaluated from the previous step's final f_g (now f_gn); the initial
    // rho_m is rho_inf (bubble_CUDA.cu:761).
    if (first || i < w.i0 || i > w.i1 || j < w.j0 || j > w.j1) return (F)g.rho_inf;
    double f = fg_to_d(fgn[frame_off(i, j, g.pitch)]);
    return (F)((1.0 - f) * g.rho_inf);
}

// Four consecutive cells of one frame row as one vector access (16 bytes in
// FP32, 32 in FP64).  Frame rows start 256-byte aligned and a thread's first
// column is a multiple of 4, so the accesses are aligned; the pitch leaves
// room for the last group of a row (brp_create sizes it).
template <typename F>
struct Vec4 {
    typedef F type __attribute__((ext_vector_type(4)));
};
template <typename F>
__device__ inline typename Vec4<F>::type ld4(const F *p) {
    return *(const typename Vec4<F>::type *)p;
}
template <typename F>
__device__ inline void st4(F *p, typename Vec4<F>::type v) {
    *(typename Vec4<F>::type *)p = v;
}

constexpr int kGW = 4;     // cells per thread along a row
constexpr int kGRows = 4;  // consecutive rows per thread
constexpr int kGBlockRows = kBY * kGRows;  // rows per block (wavefronts stacked in y)

// VelocityKernel (bubble_CUDA_kernel.cuh:635-677): vx at nodes i in [0, N],
// rows j in [1, M]; vy at cells i in [1, N], node rows j in [0, M].  A
// wavefront covers 256 consecutive cells of a row (4 per lane); each thread
// walks kGRows consecutive rows, so the p0 row j + 1 it loads for the vy
// gradient is its next row's p0; the p0 column i + 4 of the vx gradient comes
// from the next lane.  Per cell the arithmetic is the scalar update's.
template <typename F>
__global__ __launch_bounds__(256) void k_velocity(GridParams g, F *__restrict__ vx,
                                                   F *__restrict__ vy,
                                                   const F *__restrict__ p0,
                                                   const int64_t *__restrict__ fgn, Window w,
                                                   int first, const F *__restrict__ cxa,
                                                   const F *__restrict__ cxb,
                                                   const F *__restrict__ cya,
                                                   const F *__restrict__ cyb) {
    typedef typename Vec4<F>::type V;
    const int lane = threadIdx.x;
    const int i0 = (blockIdx.x * kBX + lane) * kGW;
    const int jb = blockIdx.y * kGBlockRows + threadIdx.y * kGRows;
    if (jb > g.M) return;  // wave-uniform
    const bool live = i0 <= g.N;
    const F dt = (F)g.dt;
    const F rdx = (F)g.rdx;
    const F half = (F)0.5;
    const F dt_rho = dt / (((F)g.rho_inf + (F)g.rho_inf) * half);
    F ca[kGW], cb[kGW];
#pragma unroll
    for (int e = 0; e < kGW; e++) {
        const int i = min(i0 + e, g.N);
        ca[e] = cxa[i + 1];
        cb[e] = cxb[i + 1];
    }
    const int jend = min(jb + kGRows - 1, g.M);
    V pc = live ? ld4(p0 + frame_off(i0, jb, g.pitch)) : V{};
    for (int j = jb; j <= jend; j++) {
        const int64_t o = frame_off(i0, j, g.pitch);
        const V pn = live ? ld4(p0 + o + g.pitch) : V{};  // row j + 1
        F pr = __shfl_down(pc.x, 1, 64);                   // p0(i0 + 4, j)
        if (lane == 63 && live) pr = p0[o + kGW];
        if (live) {
            const F pcs[kGW + 1] = {pc.x, pc.y, pc.z, pc.w, pr};
            const F pns[kGW] = {pn.x, pn.y, pn.z, pn.w};
            const bool row_w = !first && j + 1 >= w.j0 && j <= w.j1;
            const F ya = cya[j + 1], yb = cyb[j + 1];
            V ox = ld4(vx + o), oy = ld4(vy + o);
            F nx[kGW], ny[kGW];
            const F oxs[kGW] = {ox.x, ox.y, ox.z, ox.w};
            const F oys[kGW] = {oy.x, oy.y, oy.z, oy.w};
#pragma unroll
            for (int e = 0; e < kGW; e++) {
                const int i = i0 + e;
                const bool in_w = row_w && i + 1 >= w.i0 && i <= w.i1;  // a density not rho_inf
                const F rc = in_w ? rho_m_at<F>(g, fgn, w, first, i, j) : (F)g.rho_inf;
                {
                    const F s2 = (-pcs[e] + pcs[e + 1]) * rdx;
                    const F q = in_w ? dt / ((rc + rho_m_at<F>(g, fgn, w, first, i + 1, j)) * half)
                                     : dt_rho;
                    nx[e] = (oxs[e] * ca[e] - q * s2) / cb[e];
                }
                {
                    const F s2 = (-pcs[e] + pns[e]) * rdx;  // rdx, as the reference (kernel.cuh:671)
                    const F q = in_w ? dt / ((rc + rho_m_at<F>(g, fgn, w, first, i, j + 1)) * half)
                                     : dt_rho;
                    ny[e] = (oys[e] * ya - q * s2) / yb;
                }
            }
            // vx at nodes 0..N of rows 1..M; vy at cells 1..N of rows 0..M
            if (j >= 1) {
                if (i0 + kGW - 1 <= g.N)
                    st4(vx + o, V{nx[0], nx[1], nx[2], nx[3]});
                else
                    for (int e = 0; e < kGW; e++)
                        if (i0 + e <= g.N) vx[o + e] = nx[e];
            }
            if (i0 >= 1 && i0 + kGW - 1 <= g.N)
                st4(vy + o, V{ny[0], ny[1], ny[2], ny[3]});
            else
                for (int e = 0; e < kGW; e++)
                    if (i0 + e >= 1 && i0 + e <= g.N) vy[o + e] = ny[e];
        }
        pc = pn;
    }
}

// Focused source value of row-0 cell i at time t (MixtureBoundaryPressureKernel
// Focused_P branch, bubble_CUDA_kernel.cuh:905-925, and VelocityBoundaryKernel
// Focused_V, 763-787).  Extension E4 splits the aperture into n_elem
// contiguous elements with their own amplitude and phase.
__device__ inline double focused_value(const GridParams &g, int i, double t, double amp,
                                       double ym, const double *el_amp,
                                       const double *el_phase) {
    double xm = (i - 0.5) * g.dx;
    double s1 = sqrt((g.fp_x - xm) * (g.fp_x - xm) + (g.fp_y - ym) * (g.fp_y - ym));
    if (s1 <= g.fdist) {
        s1 = fmax(t - (g.fdist - s1) / g.cs_inf, 0.0);
        double s2 = fmod(s1, ((double)g.on_wave + (double)g.off_wave) / g.freq);
        if (s2 < (double)g.on_wave / g.freq) {
            if (el_amp && g.n_elem > 0 && g.ap_hi >= g.ap_lo) {
                long n_ap = g.ap_hi - g.ap_lo + 1;
                int e = (int)((long)(i - g.ap_lo) * g.n_elem / n_ap);
                return el_amp[e] * sin(g.omega * s2 + el_phase[e]);
            }
            return amp * sin(g.omega * s2);
        }
        return 0.0;
    }
    return 0.0;
}

// VelocityBoundaryKernel (bubble_CUDA_kernel.cuh:680-813) for order 2.  The
// reference runs its four parts in one racing grid; here one workgroup runs
// them in program order with a barrier between parts, which is the order
// warp-mates of the reference observe.
template <typename F>
__global__ __launch_bounds__(1024) void k_velocity_ghosts(GridParams g, F *vx, F *vy,
                                                          double t) {
    const int N = g.N, M = g.M, P = g.pitch;
    const int tid = threadIdx.x, nt = blockDim.x;
    // VX part a: rows j in [1, M]
    for (int j = 1 + tid; j <= M; j += nt) {
        if (!g.pml_x0) {
            vx[frame_off(-1, j, P)] = vx[frame_off(1, j, P)];
            vx[frame_off(0, j, P)] = (F)0;
        }
        if (!g.pml_x1) {
            vx[frame_off(N + 1, j, P)] = -vx[frame_off(N - 1, j, P)];
            vx[frame_off(N, j, P)] = (F)0;
        }
    }
    __syncthreads();
    // VX part b: columns i in [-1, N+1]
    for (int i = -1 + tid; i <= N + 1; i += nt) {
        vx[frame_off(i, 0, P)] = vx[frame_off(i, 1, P)];
        vx[frame_off(i, M + 1, P)] = vx[frame_off(i, M, P)];
    }
    __syncthreads();
    // VY part a: columns i in [1, N]
    for (int i = 1 + tid; i <= N; i += nt) {
        if (!g.pml_y0) {
            if (!g.plane_p && !g.focused_p) vy[frame_off(i, 0, P)] = (F)0;
            vy[frame_off(i, -1, P)] = vy[frame_off(i, 0, P)];
        }
        if (g.src_row) {
            if (g.plane_v) {
                double v = 0.0;
                if (fmod(t, ((double)g.on_wave + (double)g.off_wave) / g.freq) <=
                    ((double)g.on_wave) / g.freq)
                    v = g.amp * sin(g.omega * t);
                vy[frame_off(i, 0, P)] = (F)v;
            }
            if (g.focused_v)
                vy[frame_off(i, 0, P)] = (F)focused_value(g, i, t, g.amp, 0.0, nullptr, nullptr);
        }
        if (!g.pml_y1) {
            vy[frame_off(i, M, P)] = (F)0;
            vy[frame_off(i, M + 1, P)] = vy[frame_off(i, M - 1, P)];
        }
    }
    __syncthreads();
    // VY part b: rows j in [-1, M+1].  The right ghost column copies
    // vy(N, jsta2n + N + 1) = vy(N, N) in every row: the reference reads row
    // vy_width*i with the loop variable i (bubble_CUDA_kernel.cuh:808).
    const int qj = N;  // logical row jsta2n + (N + 1)
    const bool q_in = qj <= M + 1;
    F q = q_in ? vy[frame_off(N, qj, P)] : (F)0;
    __syncthreads();
    for (int j = -1 + tid; j <= M + 1; j += nt) {
        vy[frame_off(0, j, P)] = vy[frame_off(1, j, P)];
        vy[frame_off(N + 1, j, P)] = q;
    }
}

// MixturePressureKernel (bubble_CUDA_kernel.cuh:816-867) for the first
// evaluation of a step, fused with VFPredictionKernel (622-633) and the
// f_g / f_gn / f_gm rotation of store_variables (1250-1259):
//   pred = 3 f_g - 3 f_gn + f_gm  (exact in int64)  -> written over f_gm,
//   which becomes the step's f_g; the old f_g becomes f_gn.
// The residual max |p0_old - p0_new| goes to sc->resid[0].  Layout as
// k_velocity: 4 cells per lane, kGRows consecutive rows per thread; the vy row
// j - 1 of the divergence is the previous row's vy, the vx column i0 - 1 comes
// from the previous lane.  Per cell the arithmetic is the scalar update's.
template <typename F>
__global__ __launch_bounds__(256) void k_pressure_first(
    GridParams g, const F *__restrict__ vx, const F *__restrict__ vy, const F *__restrict__ pnx,
    const F *__restrict__ pny, F *__restrict__ px, F *__restrict__ py, F *__restrict__ p0,
    const int64_t *__restrict__ fg_cur, const int64_t *__restrict__ fg_n,
    int64_t *__restrict__ fg_mp, Window w, const F *__restrict__ ax, const F *__restrict__ bx,
    const F *__restrict__ ay, const F *__restrict__ by, const double *__restrict__ xu,
    const double *__restrict__ rxp, StepScalars *sc) {
    typedef typename Vec4<F>::type V;
    const int lane = threadIdx.x;
    const int i0 = (blockIdx.x * kBX + lane) * kGW;  // cells i0 .. i0 + 3 of [0, N+1]
    const int jb = blockIdx.y * kGBlockRows + threadIdx.y * kGRows;
    double rmax = 0.0;
    bool bad = false;
    if (jb <= g.M + 1) {  // wave-uniform
        const bool live = i0 <= g.N + 1;
        const F s3 = (F)g.s3;
        const F rdx = (F)g.rdx, rdy = (F)g.rdy;
        F xa[kGW], xb[kGW];
        double xum[kGW], xup[kGW], rx[kGW];
#pragma unroll
        for (int e = 0; e < kGW; e++) {
            const int i = min(max(i0 + e, 1), g.N);
            xa[e] = ax[i + 1];
            xb[e] = bx[i + 1];
            if (g.cylindrical) {
                xum[e] = xu[i];
                xup[e] = xu[i + 1];
                rx[e] = rxp[i + 1];
            }
        }
        const bool col_w = i0 + kGW - 1 >= w.i0 && i0 <= w.i1;
        const int jend = min(jb + kGRows - 1, g.M + 1);
        V vyl = live ? ld4(vy + frame_off(i0, jb - 1, g.pitch)) : V{};  // vy row jb - 1
        for (int j = jb; j <= jend; j++) {
            const int64_t o = frame_off(i0, j, g.pitch);
            const V vxc = live ? ld4(vx + o) : V{};
            const V vyc = live ? ld4(vy + o) : V{};
            F vxl = __shfl_up(vxc.w, 1, 64);  // vx(i0 - 1, j)
            if (lane == 0 && live) vxl = vx[o - 1];
            if (live) {
                int64_t fpred[kGW] = {0, 0, 0, 0}, fold[kGW] = {0, 0, 0, 0};
                if (col_w && j >= w.j0 && j <= w.j1) {
#pragma unroll
                    for (int e = 0; e < kGW; e++) {
                        const int i = i0 + e;
                        if (i >= w.i0 && i <= w.i1) {
                            const int64_t c = fg_cur[o + e];
                            fpred[e] = 3 * c - 3 * fg_n[o + e] + fg_mp[o + e];
                            fold[e] = c;
                            fg_mp[o + e] = fpred[e];
                        }
                    }
                }
                if (j >= 1 && j <= g.M) {
                    const V qx = ld4(pnx + o), qy = ld4(pny + o), qo = ld4(p0 + o);
                    const F pxs[kGW] = {qx.x, qx.y, qx.z, qx.w};
                    const F pys[kGW] = {qy.x, qy.y, qy.z, qy.w};
                    const F p0s[kGW] = {qo.x, qo.y, qo.z, qo.w};
                    const F vxs[kGW + 1] = {vxl, vxc.x, vxc.y, vxc.z, vxc.w};
                    const F vys[kGW] = {vyc.x, vyc.y, vyc.z, vyc.w};
                    const F vyls[kGW] = {vyl.x, vyl.y, vyl.z, vyl.w};
                    const F ya = ay[j + 1], yb = by[j + 1];
                    F npx[kGW], npy[kGW], np0[kGW];
#pragma unroll
                    for (int e = 0; e < kGW; e++) {
                        // outside the window fpred = fold = 0 and the source term is -0
                        F s5h = (F)(-0.0);
                        if (fpred[e] != 0 || fold[e] != 0) {
                            const double s5d = -fg_to_d(fpred[e] - fold[e]) / g.dt /
                                               (1.0 - 0.5 * fg_to_d(fpred[e] + fold[e]));
                            s5h = (F)(0.5 * s5d);
                        }
                        F s1x;
                        if (g.cylindrical)
                            s1x = (F)((-xum[e] * (double)vxs[e] + xup[e] * (double)vxs[e + 1]) *
                                      g.rdx * rx[e]);
                        else
                            s1x = (-vxs[e] + vxs[e + 1]) * rdx;
                        const F s1y = (-vyls[e] + vys[e]) * rdy;
                        npx[e] = pxs[e] * xa[e] - s3 * (s1x + s5h) / xb[e];
                        npy[e] = pys[e] * ya - s3 * (s1y + s5h) / yb;
                        np0[e] = npx[e] + npy[e];
                        const int i = i0 + e;
                        if (i >= 1 && i <= g.N) {
                            rmax = fmax(rmax, fabs((double)p0s[e] - (double)np0[e]));
                            bad |= !isfinite((double)np0[e]) || fabs(fg_to_d(fpred[e])) >= 1.0;
                        }
                    }
                    if (i0 >= 1 && i0 + kGW - 1 <= g.N) {
                        st4(px + o, V{npx[0], npx[1], npx[2], npx[3]});
                        st4(py + o, V{npy[0], npy[1], npy[2], npy[3]});
                        st4(p0 + o, V{np0[0], np0[1], np0[2], np0[3]});
                    } else {
                        for (int e = 0; e < kGW; e++)
                            if (i0 + e >= 1 && i0 + e <= g.N) {
                                px[o + e] = npx[e];
                                py[o + e] = npy[e];
                                p0[o + e] = np0[e];
                            }
                    }
                }
            }
            vyl = vyc;
        }
    }
    if (bad) atomicOr(&sc->flags, 1);
    block_max_to(rmax, &sc->resid[0]);
}

// MixturePressureKernel for Picard iteration k >= 1 over the void-fraction
// window only: outside it f_g = f_gn = 0 in every iterate, so p and p0 are
// unchanged there and the residual is exactly 0.  fg_k is the scatter result
// of this iteration; fg_prev (the previous iterate) is cleared for the next
// scatter.  FP64 residual: |p0_old - p0_new| as the reference.  Mixed: the
// same quantity evaluated exactly in FP64 from the change of the source
// term, s3 * 0.5 * (s5_k - s5_{k-1}) * (1/(1+s2x) + 1/(1+s2y)), since the
// FP32 pressures cannot resolve the reference's 0.01 Pa threshold.
template <typename F>
__global__ __launch_bounds__(256) void k_pressure_picard(
    GridParams g, const F *__restrict__ vx, const F *__restrict__ vy, const F *__restrict__ pnx,
    const F *__restrict__ pny, F *__restrict__ px, F *__restrict__ py, F *__restrict__ p0,
    const int64_t *__restrict__ fg_k, const int64_t *__restrict__ fg_n,
    int64_t *__restrict__ fg_prev, Window w, const F *__restrict__ ax, const F *__restrict__ bx,
    const F *__restrict__ ay, const F *__restrict__ by, const double *__restrict__ xu,
    const double *__restrict__ rxp, StepScalars *sc, int slot) {
    const int i = w.i0 + (int)(blockIdx.x * kBX + threadIdx.x);
    double rmax = 0.0;
    bool bad = false;
    if (i <= w.i1) {
        const int j0 = w.j0 + (int)blockIdx.y * kRB;
        const int jend = min(j0 + kRB - 1, w.j1);
        const F s3 = (F)g.s3;
        const F rdx = (F)g.rdx, rdy = (F)g.rdy;
        for (int j = j0 + threadIdx.y; j <= jend; j += kBY) {
            const int64_t o = frame_off(i, j, g.pitch);
            const int64_t fk = fg_k[o];
            const int64_t fn = fg_n[o];
            const int64_t fprev = fg_prev[o];
            fg_prev[o] = 0;
            if (i < 1 || i > g.N || j < 1 || j > g.M) continue;
            const double s5d = -fg_to_d(fk - fn) / g.dt / (1.0 - 0.5 * fg_to_d(fk + fn));
            const F s5h = (F)(0.5 * s5d);
            F s1x;
            if (g.cylindrical)
                s1x = (F)((-xu[i] * (double)vx[o - 1] + xu[i + 1] * (double)vx[o]) * g.rdx *
                          rxp[i + 1]);
            else
                s1x = (-vx[o - 1] + vx[o]) * rdx;
            const F s1y = (-vy[o - g.pitch] + vy[o]) * rdy;
            const F ppx = pnx[o] * ax[i + 1] - s3 * (s1x + s5h) / bx[i + 1];
            const F ppy = pny[o] * ay[j + 1] - s3 * (s1y + s5h) / by[j + 1];
            px[o] = ppx;
            py[o] = ppy;
            const F pnew = ppx + ppy;
            double r;
            if (sizeof(F) == sizeof(double)) {
                r = fabs((double)p0[o] - (double)pnew);
            } else {
                const double s5p =
                    -fg_to_d(fprev - fn) / g.dt / (1.0 - 0.5 * fg_to_d(fprev + fn));
                const double d = g.s3 * 0.5 * (s5d - s5p);
                r = fabs(d / (double)bx[i + 1] + d / (double)by[j + 1]);
            }
            rmax = fmax(rmax, r);
            bad |= !isfinite((double)pnew) || fabs(fg_to_d(fk)) >= 1.0;
            p0[o] = pnew;
        }
    }
    if (bad) atomicOr(&sc->flags, 1);
    block_max_to(rmax, &sc->resid[slot]);
}

// MixtureBoundaryPressureKernel (bubble_CUDA_kernel.cuh:870-952) for order 2,
// parts in program order in one workgroup.  Extension E1 (g.src_row) writes
// the source row even when pml.y0 = 1.
template <typename F>
__global__ __launch_bounds__(1024) void k_pressure_ghosts(GridParams g, F *p0,
                                                          const double *el_amp,
                                                          const double *el_phase, double t) {
    const int N = g.N, M = g.M, P = g.pitch;
    const int tid = threadIdx.x, nt = blockDim.x;
    const double pamp = g.rho_inf * g.cs_inf * g.amp;
    for (int i = 1 + tid; i <= N; i += nt) {
        if (!g.pml_y0) p0[frame_off(i, 0, P)] = p0[frame_off(i, 1, P)];
        if (g.src_row) {
            if (g.plane_p) {
                double v = 0.0;
                if (fmod(t, ((double)g.on_wave + (double)g.off_wave) / g.freq) <=
                    ((double)g.on_wave) / g.freq)
                    v = pamp * sin(g.omega * t);
                p0[frame_off(i, 0, P)] = (F)v;
            }
            if (g.focused_p)
                p0[frame_off(i, 0, P)] =
                    (F)focused_value(g, i, t, pamp, -0.5 * g.dy, el_amp, el_phase);
        }
        if (!g.pml_y1) {
            // As the reference (kernel.cuh:929-932): rows X..X+1 copy row 2Y+1-j (rows outside [0, M+1] skipped).
            for (int j = N; j <= N + 1; j++) {
                int src = 2 * M + 1 - j;
                if (j >= 0 && j <= M + 1 && src >= 0 && src <= M + 1)
                    p0[frame_off(i, j, P)] = p0[frame_off(i, src, P)];
            }
        }
    }
    __syncthreads();
    for (int j = tid; j <= M + 1; j += nt) {
        if (!g.pml_x0) p0[frame_off(0, j, P)] = p0[frame_off(1, j, P)];
        if (!g.pml_x1) p0[frame_off(N + 1, j, P)] = p0[frame_off(N, j, P)];
    }
}

__global__ void k_zero_scalars(StepScalars *sc, int nslots) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < nslots) {
        sc->resid[t] = 0ull;
        sc->work[t] = 0u;
    }
    if (t == 0) {
        sc->bbox[0] = 0x7fffffff;
        sc->bbox[1] = -0x7fffffff;
        sc->bbox[2] = 0x7fffffff;
        sc->bbox[3] = -0x7fffffff;
        sc->flags = 0;
        sc->sub_max = 0;
        sc->sub_sum = 0ull;
        sc->done = 0u;
        sc->reused = 0u;
    }
}

}  // namespace

template <typename F>
void launch_velocity(hipStream_t s, const GridParams &g, F *vx, F *vy, const F *p0,
                     const int64_t *fgn, Window w, int first_step, const F *cxa, const F *cxb,
                     const F *cya, const F *cyb) {
    dim3 blk(kBX, kBY);  // columns 0..N, rows 0..M
    dim3 grd((g.N + 1 + kBX * kGW - 1) / (kBX * kGW), (g.M + 1 + kGBlockRows - 1) / kGBlockRows);
    hipLaunchKernelGGL(k_velocity<F>, grd, blk, 0, s, g, vx, vy, p0, fgn, w, first_step, cxa,
                       cxb, cya, cyb);
}

template <typename F>
void launch_velocity_ghosts(hipStream_t s, const GridParams &g, F *vx, F *vy, double t) {
    hipLaunchKernelGGL(k_velocity_ghosts<F>, dim3(1), dim3(1024), 0, s, g, vx, vy, t);
}

template <typename F>
void launch_pressure_first(hipStream_t s, const GridParams &g, const F *vx, const F *vy,
                           const F *pnx, const F *pny, F *px, F *py, F *p0,
                           const int64_t *fg_cur, const int64_t *fg_n, int64_t *fg_mp, Window w,
                           const F *ax, const F *bx, const F *ay, const F *by, const double *xu,
                           const double *rxp, StepScalars *sc) {
    dim3 blk(kBX, kBY);  // columns 0..N+1, rows 0..M+1
    dim3 grd((g.N + 2 + kBX * kGW - 1) / (kBX * kGW), (g.M + 2 + kGBlockRows - 1) / kGBlockRows);
    hipLaunchKernelGGL(k_pressure_first<F>, grd, blk, 0, s, g, vx, vy, pnx, pny, px, py, p0,
                       fg_cur, fg_n, fg_mp, w, ax, bx, ay, by, xu, rxp, sc);
}

template <typename F>
void launch_pressure_picard(hipStream_t s, const GridParams &g, const F *vx, const F *vy,
                            const F *pnx, const F *pny, F *px, F *py, F *p0,
                            const int64_t *fg_k, const int64_t *fg_n, int64_t *fg_prev,
                            Window w, const F *ax, const F *bx, const F *ay, const F *by,
                            const double *xu, const double *rxp, StepScalars *sc, int slot) {
    if (w.i1 < w.i0 || w.j1 < w.j0) return;
    dim3 blk(kBX, kBY);
    dim3 grd((w.i1 - w.i0 + 1 + kBX - 1) / kBX, (w.j1 - w.j0 + 1 + kRB - 1) / kRB);
    hipLaunchKernelGGL(k_pressure_picard<F>, grd, blk, 0, s, g, vx, vy, pnx, pny, px, py, p0,
                       fg_k, fg_n, fg_prev, w, ax, bx, ay, by, xu, rxp, sc, slot);
}

template <typename F>
void launch_pressure_ghosts(hipStream_t s, const GridParams &g, F *p0, const double *el_amp,
                            const double *el_phase, double t) {
    hipLaunchKernelGGL(k_pressure_ghosts<F>, dim3(1), dim3(1024), 0, s, g, p0, el_amp,
                       el_phase, t);
}

void launch_zero_scalars(hipStream_t s, StepScalars *sc, int nslots) {
    hipLaunchKernelGGL(k_zero_scalars, dim3((nslots + 255) / 256), dim3(256), 0, s, sc, nslots);
}

#define BRP_INST(F)                                                                          \
    template void launch_velocity<F>(hipStream_t, const GridParams &, F *, F *, const F *,   \
                                     const int64_t *, Window, int, const F *, const F *,     \
                                     const F *, const F *);                                  \
    template void launch_velocity_ghosts<F>(hipStream_t, const GridParams &, F *, F *,      \
                                            double);                                         \
    template void launch_pressure_first<F>(                                                  \
        hipStream_t, const GridParams &, const F *, const F *, const F *, const F *, F *,   \
        F *, F *, const int64_t *, const int64_t *, int64_t *, Window, const F *, const F *, \
        const F *, const F *, const double *, const double *, StepScalars *);                \
    template void launch_pressure_picard<F>(                                                 \
        hipStream_t, const GridParams &, const F *, const F *, const F *, const F *, F *,   \
        F *, F *, const int64_t *, const int64_t *, int64_t *, Window, const F *, const F *, \
        const F *, const F *, const double *, const double *, StepScalars *, int);           \
    template void launch_pressure_ghosts<F>(hipStream_t, const GridParams &, F *,           \
                                            const double *, const double *, double);

BRP_INST(double)
BRP_INST(float)

}  // namespace brp
