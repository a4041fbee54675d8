// kernels_adjoint.hip -- discrete adjoint of the coupled step for gfx950.
//
// Reverse-mode kernels for the objective J = mean over a focal box and the
// last K steps of p0^2, differentiated with respect to the per-element source
// amplitudes and phases and the initial void-fraction map.  Each kernel is
// the transpose of one forward kernel at the forward's recorded state:
//   k_adj_pressure   MixturePressureKernel (+ VFPredictionKernel on eval 0)
//   k_adj_pressure_ghosts / k_adj_velocity_ghosts   the boundary kernels,
//                    their writes undone in reverse program order
//   k_adj_velocity_a / _b   VelocityKernel
//   k_adj_scatter / k_adj_gather / k_adj_motion   the bubble-grid coupling,
//                    including the derivative of the smooth-delta weights with
//                    respect to the bubble position (bins are piecewise
//                    constant and not differentiated)
//   k_adj_rp_jac / k_adj_rp_apply   Rayleigh-Plesset over one mixture step
//                    at the frozen sub-step schedule: Jacobian columns from
//                    seven tangent lanes per bubble, then their transpose
// Grid adjoints are FP64 in both precision modes; scatter-type transposes
// (gather adjoint, motion adjoint) use FP64 atomics, every other kernel is a
// gather and writes each cell from one thread.
#include "brp_internal.h"
#include "rp_fast.h"

#include <math.h>

#include <algorithm>

namespace brp {
namespace {

constexpr int kAX = 64, kAY = 4;

__device__ inline double fgd(int64_t v) { return (double)v * kFgInvScale; }

__device__ inline double wdelta(double c, double k, int pos, double h, double b) {
    return c * (1.0 + cos(k * (((double)pos - 0.5) * h - b)));
}
// d/db of wdelta
__device__ inline double wdelta_db(double c, double k, int pos, double h, double b) {
    return c * k * sin(k * (((double)pos - 0.5) * h - b));
}

__device__ inline int bin_idx(double x, double rdx, double c) {
#pragma clang fp contract(off)
    double v = floor(x * rdx + c);
    return (v >= -1.0e9 && v <= 1.0e9) ? (int)v : -1000000000;
}

// s5 = -(f - g) / dt / (1 - (f + g) / 2) and its partial derivatives.
__device__ inline void s5_partials(double f, double g, double dt, double *dfd, double *dgd) {
    const double D = 1.0 - 0.5 * (f + g);
    const double q = (f - g) / (dt * D * D);
    *dfd = -1.0 / (dt * D) - 0.5 * q;
    *dgd = 1.0 / (dt * D) - 0.5 * q;
}

// J accumulation: sum of p0^2 over the box into out[0] (one workgroup,
// fixed reduction order).
template <typename F>
__global__ __launch_bounds__(256) void k_box_sumsq(GridParams g, const F *p0, int i0, int i1,
                                                    int j0, int j1, double *out) {
    const int ni = i1 - i0 + 1, n = ni * (j1 - j0 + 1);
    double acc = 0.0;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double v = (double)p0[frame_off(i0 + k % ni, j0 + k / ni, g.pitch)];
        acc += v * v;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// dJ/dp0 of the step's final pressure: 2 c p0 over the box.
template <typename F>
__global__ void k_adj_seed(GridParams g, Rect<F> p0, double *Lp0, int i0, int i1, int j0,
                          int j1, double c) {
    const int i = i0 + (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int j = j0 + (int)blockIdx.y;
    if (i > i1 || j > j1) return;
    Lp0[frame_off(i, j, g.pitch)] += 2.0 * c * (double)p0.at(i, j);
}

// Adjoint of k_pressure_ghosts: part 2 undone first, then part 1 with each
// column's writes in reverse order.  The focused source of element e,
// A_e sin(omega s2 + phi_e), adds L * sin(.) to gA[i] and L * A_e cos(.) to
// gP[i] of its column.
__global__ __launch_bounds__(1024) void k_adj_pressure_ghosts(GridParams g, double *L,
                                                               const double *el_amp,
                                                               const double *el_phase, double t,
                                                               double *gA, double *gP) {
    const int N = g.N, M = g.M, P = g.pitch;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int j = tid; j <= M + 1; j += nt) {
        if (!g.pml_x1) {
            L[frame_off(N, j, P)] += L[frame_off(N + 1, j, P)];
            L[frame_off(N + 1, j, P)] = 0.0;
        }
        if (!g.pml_x0) {
            L[frame_off(1, j, P)] += L[frame_off(0, j, P)];
            L[frame_off(0, j, P)] = 0.0;
        }
    }
    __syncthreads();
    for (int i = 1 + tid; i <= N; i += nt) {
        if (!g.pml_y1) {
            for (int j = N + 1; j >= N; j--) {
                const int src = 2 * M + 1 - j;
                if (j >= 0 && j <= M + 1 && src >= 0 && src <= M + 1) {
                    L[frame_off(i, src, P)] += L[frame_off(i, j, P)];
                    L[frame_off(i, j, P)] = 0.0;
                }
            }
        }
        const int64_t o0 = frame_off(i, 0, P);
        if (g.src_row && g.focused_p) {
            const double lam = L[o0];
            const double xm = (i - 0.5) * g.dx, ym = -0.5 * g.dy;
            double s1 = sqrt((g.fp_x - xm) * (g.fp_x - xm) + (g.fp_y - ym) * (g.fp_y - ym));
            if (s1 <= g.fdist && el_amp && g.n_elem > 0 && g.ap_hi >= g.ap_lo) {
                s1 = fmax(t - (g.fdist - s1) / g.cs_inf, 0.0);
                const double s2 = fmod(s1, ((double)g.on_wave + (double)g.off_wave) / g.freq);
                if (s2 < (double)g.on_wave / g.freq) {
                    const long n_ap = g.ap_hi - g.ap_lo + 1;
                    const int e = (int)((long)(i - g.ap_lo) * g.n_elem / n_ap);
                    double sn, cn;
                    sincos(g.omega * s2 + el_phase[e], &sn, &cn);
                    gA[i] += lam * sn;
                    gP[i] += lam * el_amp[e] * cn;
                }
            }
            L[o0] = 0.0;
        } else if (g.src_row && g.plane_p) {
            L[o0] = 0.0;
        }
        if (!g.pml_y0) {
            L[frame_off(i, 1, P)] += L[o0];
            L[o0] = 0.0;
        }
    }
}

// Adjoint of one pressure evaluation p = P(pn, v, f, g) (k_pressure_first /
// k_pressure_picard), as a gather over cells (i, j) in [-1, N+1] x [-1, M+1]:
//   interior cells: gx = Lpx + Lp0, gy = Lpy + Lp0 ->  Lpn += g a,
//     a = -s3 gx / bx, b = -s3 gy / by,  ds5 = (a + b) / 2
//   vx node (i, j) gets  (a(i,j) - a(i+1,j)) rdx,  vy node (b(i,j) - b(i,j+1)) rdy
//   f cells [0, N+1]^2:  Lf = ds5 ds5/df (+ extra),  LgC += ds5 ds5/dg.
// FIRST: f is the prediction 3 fC - 3 fN + fM, so Lf goes to (LgC, LfN, LfM)
// with weights (3, -3, 1) instead of to Lf (LfM is written, not added).
template <typename F, bool FIRST>
__global__ __launch_bounds__(256) void k_adj_pressure(
    GridParams g, const double *__restrict__ Lp0, const double *__restrict__ Lpx,
    const double *__restrict__ Lpy, double *__restrict__ Lpnx, double *__restrict__ Lpny,
    double *__restrict__ Lvx, double *__restrict__ Lvy, Rect<int64_t> fk, Rect<int64_t> gC,
    const double *extra, double *__restrict__ Lf,
    double *__restrict__ LgC, double *__restrict__ LfN, double *LfM,
    const F *__restrict__ ax, const F *__restrict__ bx, const F *__restrict__ ay,
    const F *__restrict__ by, Window w) {
    const int i = w.i0 + (int)(blockIdx.x * kAX + threadIdx.x);
    const int j = w.j0 + (int)(blockIdx.y * kAY + threadIdx.y);
    const int N = g.N, M = g.M, P = g.pitch;
    if (i > w.i1 || j > w.j1) return;
    const double s3 = (double)(F)g.s3;
    const int64_t o = frame_off(i, j, P);
    auto interior = [&](int a, int b) { return a >= 1 && a <= N && b >= 1 && b <= M; };
    auto ax_of = [&](int a, int b) {
        const int64_t q = frame_off(a, b, P);
        return -s3 * (Lpx[q] + Lp0[q]) / (double)bx[a + 1];
    };
    auto by_of = [&](int a, int b) {
        const int64_t q = frame_off(a, b, P);
        return -s3 * (Lpy[q] + Lp0[q]) / (double)by[b + 1];
    };
    double a_c = 0.0, b_c = 0.0;
    const bool in_c = interior(i, j);
    if (in_c) {
        const double gx = Lpx[o] + Lp0[o], gy = Lpy[o] + Lp0[o];
        Lpnx[o] += gx * (double)ax[i + 1];
        Lpny[o] += gy * (double)ay[j + 1];
        a_c = -s3 * gx / (double)bx[i + 1];
        b_c = -s3 * gy / (double)by[j + 1];
    }
    if (i >= 0 && i <= N && j >= 1 && j <= M) {
        double v = a_c;
        if (interior(i + 1, j)) v -= ax_of(i + 1, j);
        Lvx[o] += v * (double)(F)g.rdx;
    }
    if (i >= 1 && i <= N && j >= 0 && j <= M) {
        double v = b_c;
        if (interior(i, j + 1)) v -= by_of(i, j + 1);
        Lvy[o] += v * (double)(F)g.rdy;
    }
    if (i >= 0 && j >= 0) {
        double lf = extra ? extra[o] : 0.0;
        if (in_c) {
            double df, dg;
            s5_partials(fgd(fk.at(i, j)), fgd(gC.at(i, j)), g.dt, &df, &dg);
            const double ds5 = 0.5 * (a_c + b_c);
            lf += ds5 * df;
            LgC[o] += ds5 * dg;
        }
        if (FIRST) {
            LgC[o] += 3.0 * lf;
            LfN[o] -= 3.0 * lf;
            LfM[o] = lf;
        } else {
            Lf[o] = lf;
        }
    }
}

// The evaluation overwrote p, p0 in the interior: their adjoints are consumed
// (over the interior cells of w).
__global__ void k_adj_zero_interior(GridParams g, double *Lp0, double *Lpx, double *Lpy,
                                    Window w) {
    const int i = w.i0 + (int)(blockIdx.x * kAX + threadIdx.x);
    const int j = w.j0 + (int)(blockIdx.y * kAY + threadIdx.y);
    if (i > w.i1 || j > w.j1) return;
    const int64_t o = frame_off(i, j, g.pitch);
    Lp0[o] = 0.0;
    Lpx[o] = 0.0;
    Lpy[o] = 0.0;
}

// Adjoint of k_velocity_ghosts: parts VY b, VY a, VX b, VX a undone in turn.
__global__ __launch_bounds__(1024) void k_adj_velocity_ghosts(GridParams g, double *Lvx,
                                                               double *Lvy) {
    const int N = g.N, M = g.M, P = g.pitch;
    const int tid = threadIdx.x, nt = blockDim.x;
    __shared__ double red[1024];
    // VY part b: (0, j) <- (1, j); (N+1, j) <- q = vy(N, N) read before the part
    double lq = 0.0;
    for (int j = -1 + tid; j <= M + 1; j += nt) {
        const int64_t oq = frame_off(N + 1, j, P);
        lq += Lvy[oq];
        Lvy[oq] = 0.0;
        Lvy[frame_off(1, j, P)] += Lvy[frame_off(0, j, P)];
        Lvy[frame_off(0, j, P)] = 0.0;
    }
    red[tid] = lq;
    __syncthreads();
    for (int s = nt / 2; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0 && N <= M + 1) Lvy[frame_off(N, N, P)] += red[0];
    __syncthreads();
    // VY part a
    for (int i = 1 + tid; i <= N; i += nt) {
        if (!g.pml_y1) {
            Lvy[frame_off(i, M - 1, P)] += Lvy[frame_off(i, M + 1, P)];
            Lvy[frame_off(i, M + 1, P)] = 0.0;
            Lvy[frame_off(i, M, P)] = 0.0;
        }
        if (g.src_row && (g.plane_v || g.focused_v)) Lvy[frame_off(i, 0, P)] = 0.0;
        if (!g.pml_y0) {
            Lvy[frame_off(i, 0, P)] += Lvy[frame_off(i, -1, P)];
            Lvy[frame_off(i, -1, P)] = 0.0;
            if (!g.plane_p && !g.focused_p) Lvy[frame_off(i, 0, P)] = 0.0;
        }
    }
    __syncthreads();
    // VX part b
    for (int i = -1 + tid; i <= N + 1; i += nt) {
        Lvx[frame_off(i, M, P)] += Lvx[frame_off(i, M + 1, P)];
        Lvx[frame_off(i, M + 1, P)] = 0.0;
        Lvx[frame_off(i, 1, P)] += Lvx[frame_off(i, 0, P)];
        Lvx[frame_off(i, 0, P)] = 0.0;
    }
    __syncthreads();
    // VX part a
    for (int j = 1 + tid; j <= M; j += nt) {
        if (!g.pml_x1) {
            Lvx[frame_off(N, j, P)] = 0.0;
            Lvx[frame_off(N - 1, j, P)] -= Lvx[frame_off(N + 1, j, P)];
            Lvx[frame_off(N + 1, j, P)] = 0.0;
        }
        if (!g.pml_x0) {
            Lvx[frame_off(0, j, P)] = 0.0;
            Lvx[frame_off(1, j, P)] += Lvx[frame_off(-1, j, P)];
            Lvx[frame_off(-1, j, P)] = 0.0;
        }
    }
}

// Adjoint of k_velocity, part a: p0 and f_gn adjoints as a gather over the
// cells (i, j) in [0, N+1] x [0, M+1] from the vx nodes (i-1, j), (i, j) and
// the vy nodes (i, j-1), (i, j).
template <typename F>
__global__ __launch_bounds__(256) void k_adj_velocity_a(
    GridParams g, const double *__restrict__ Lvx, const double *__restrict__ Lvy,
    const F *__restrict__ p0, Rect<int64_t> fgn, int first,
    const F *__restrict__ cxb, const F *__restrict__ cyb, double *__restrict__ Lp0,
    double *__restrict__ LfC) {
    const int i = (int)(blockIdx.x * kAX + threadIdx.x);
    const int j = (int)(blockIdx.y * kAY + threadIdx.y);
    const int N = g.N, M = g.M, P = g.pitch;
    if (i > N + 1 || j > M + 1) return;
    const double dt = (double)(F)g.dt, rdx = (double)(F)g.rdx, rho = g.rho_inf;
    auto rho_at = [&](int a, int b) {
        return first ? rho : (1.0 - fgd(fgn.at(a, b))) * rho;
    };
    auto pv = [&](int a, int b) { return (double)p0[frame_off(a, b, P)]; };
    double lp = 0.0, lr = 0.0;  // adjoints of p0(i, j) and rho(i, j)
    // vx node i' = i - 1: reads p0(i) with weight -, rho(i) in s1
    if (i - 1 >= 0 && i - 1 <= N && j >= 1 && j <= M) {
        const double c = Lvx[frame_off(i - 1, j, P)] / (double)cxb[i];
        const double s1 = 0.5 * (rho_at(i - 1, j) + rho_at(i, j));
        const double s2 = (pv(i, j) - pv(i - 1, j)) * rdx;
        lp -= c * dt * rdx / s1;
        lr += 0.5 * c * dt * s2 / (s1 * s1);
    }
    if (i <= N && j >= 1 && j <= M) {
        const double c = Lvx[frame_off(i, j, P)] / (double)cxb[i + 1];
        const double s1 = 0.5 * (rho_at(i, j) + rho_at(i + 1, j));
        const double s2 = (pv(i + 1, j) - pv(i, j)) * rdx;
        lp += c * dt * rdx / s1;
        lr += 0.5 * c * dt * s2 / (s1 * s1);
    }
    if (i >= 1 && i <= N && j - 1 >= 0 && j - 1 <= M) {
        const double c = Lvy[frame_off(i, j - 1, P)] / (double)cyb[j];
        const double s1 = 0.5 * (rho_at(i, j - 1) + rho_at(i, j));
        const double s2 = (pv(i, j) - pv(i, j - 1)) * rdx;  // rdx, as the forward
        lp -= c * dt * rdx / s1;
        lr += 0.5 * c * dt * s2 / (s1 * s1);
    }
    if (i >= 1 && i <= N && j <= M) {
        const double c = Lvy[frame_off(i, j, P)] / (double)cyb[j + 1];
        const double s1 = 0.5 * (rho_at(i, j) + rho_at(i, j + 1));
        const double s2 = (pv(i, j + 1) - pv(i, j)) * rdx;
        lp += c * dt * rdx / s1;
        lr += 0.5 * c * dt * s2 / (s1 * s1);
    }
    const int64_t o = frame_off(i, j, P);
    Lp0[o] += lp;
    if (!first) LfC[o] -= rho * lr;
}

// Adjoint of k_velocity, part b: the updated nodes' own term vx cxa / cxb.
template <typename F>
__global__ void k_adj_velocity_b(GridParams g, double *Lvx, double *Lvy, const F *cxa,
                                 const F *cxb, const F *cya, const F *cyb) {
    const int i = (int)(blockIdx.x * kAX + threadIdx.x);
    const int j = (int)(blockIdx.y * kAY + threadIdx.y);
    if (i > g.N || j > g.M) return;
    const int64_t o = frame_off(i, j, g.pitch);
    if (j >= 1) Lvx[o] *= (double)cxa[i + 1] / (double)cxb[i + 1];
    if (i >= 1) Lvy[o] *= (double)cya[j + 1] / (double)cyb[j + 1];
}

// Adjoint of the scatter of iteration k: lRt = dJ/dR_t, and the position
// adjoints of the weights.
__global__ __launch_bounds__(256) void k_adj_scatter(GridParams g, BubbleParams bp,
                                                      BubbleArrays b, const double *Rt,
                                                      const double *__restrict__ Lf, double *lRt,
                                                      double *Lx, double *Ly) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= b.n) return;
    const double x = b.x[idx], y = b.y[idx], R = Rt[idx];
    const int ibnx = bin_idx(x, g.rdx, 0.5), ibny = bin_idx(y, g.rdy, 0.5);
    const int nb = bp.mbe - bp.mbs + 1;
    double s = 0.0, sx = 0.0, sy = 0.0;
    for (int a = 0; a < nb; a++) {
        const int i = ibnx + bp.mbs + a;
        const double wx = wdelta(bp.dcx, bp.dcy, i, g.dx, x);
        const double dwx = wdelta_db(bp.dcx, bp.dcy, i, g.dx, x);
        for (int c = 0; c < nb; c++) {
            const int j = ibny + bp.mbs + c;
            if (i >= 0 && i <= g.N + 1 && j >= 0 && j <= g.M + 1) {
                const double wy = wdelta(bp.dcx, bp.dcz, j, g.dy, y);
                const double dwy = wdelta_db(bp.dcx, bp.dcz, j, g.dy, y);
                const double l = Lf[frame_off(i, j, g.pitch)];
                s += l * wx * wy;
                sx += l * dwx * wy;
                sy += l * wx * dwy;
            }
        }
    }
    const double base = bp.pi4r3 * b.n_B[idx] * g.rdx * g.rdy;  // fg_temp = base R^3
    lRt[idx] = s * 3.0 * base * R * R;
    const double ft = base * R * R * R;
    Lx[idx] += ft * sx;
    Ly[idx] += ft * sy;
}

// FP64 contributions of a workgroup's bubble stencils to a frame array,
// summed in an LDS tile over the workgroup's bounding box and flushed with
// one global atomic per touched cell: bubbles are cell-sorted, so a
// workgroup's stencils cover a small patch, and in dense clouds (hundreds
// of bubbles per cell) per-contribution global atomics on the same
// addresses serialise.  A patch larger than the tile falls back to global
// atomics per contribution.  Every thread of the workgroup calls begin and
// end (they synchronise); `any` says whether the thread has a stencil.
constexpr int kAdjTile = 4096;  // cells (32 KB of LDS)
struct TileAcc {
    double *tile;
    int bx0, by0, bw;
    bool binned;
};
__device__ inline TileAcc tile_begin(double *tile, int *box, bool any, int i0, int i1, int j0,
                                     int j1) {
    __syncthreads();  // the previous use of box / tile has finished
    if (threadIdx.x == 0) {
        box[0] = 0x7fffffff;
        box[1] = -0x7fffffff;
        box[2] = 0x7fffffff;
        box[3] = -0x7fffffff;
    }
    __syncthreads();
    if (any) {
        atomicMin(&box[0], i0);
        atomicMax(&box[1], i1);
        atomicMin(&box[2], j0);
        atomicMax(&box[3], j1);
    }
    __syncthreads();
    TileAcc t{tile, box[0], box[2], box[1] - box[0] + 1, false};
    const int bh = box[3] - box[2] + 1;
    t.binned = t.bw > 0 && bh > 0 && (int64_t)t.bw * bh <= kAdjTile;  // workgroup-uniform
    if (t.binned) {
        for (int k = threadIdx.x; k < t.bw * bh; k += blockDim.x) tile[k] = 0.0;
        __syncthreads();
    }
    return t;
}
__device__ inline void tile_add(const TileAcc &t, double *glob, int pitch, int i, int j, double v) {
    if (t.binned)
        atomicAdd(&t.tile[(j - t.by0) * t.bw + (i - t.bx0)], v);
    else
        atomicAdd(&glob[frame_off(i, j, pitch)], v);
}
__device__ inline void tile_end(const TileAcc &t, const int *box, double *glob, int pitch) {
    if (!t.binned) return;
    __syncthreads();
    const int n = t.bw * (box[3] - box[2] + 1);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const double v = t.tile[k];
        if (v != 0.0) atomicAdd(&glob[frame_off(t.bx0 + k % t.bw, t.by0 + k / t.bw, pitch)], v);
    }
}

// Adjoint of the gather PL_p = sum p0 w of iteration k (p0 = the previous
// evaluation's pressure): FP64 atomics into Lp0 (through the workgroup's
// LDS tile), position adjoints.
template <typename F>
__global__ __launch_bounds__(256) void k_adj_gather(GridParams g, BubbleParams bp,
                                                     BubbleArrays b, Rect<F> p0,
                                                     const double *lPL, const double *lPL_extra,
                                                     double *Lp0, double *Lx, double *Ly) {
    __shared__ double tile[kAdjTile];
    __shared__ int box[4];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = idx < b.n;
    double lam = 0.0, x = 0.0, y = 0.0;
    int ibnx = 0, ibny = 0;
    if (valid) {
        lam = lPL[idx] + (lPL_extra ? lPL_extra[idx] : 0.0);
        x = b.x[idx];
        y = b.y[idx];
        ibnx = bin_idx(x, g.rdx, 0.5);
        ibny = bin_idx(y, g.rdy, 0.5);
    }
    const int i0 = max(ibnx + bp.mbs, 0), i1 = min(ibnx + bp.mbe, g.N + 1);
    const int j0 = max(ibny + bp.mbs, 0), j1 = min(ibny + bp.mbe, g.M + 1);
    const bool any = valid && lam != 0.0 && i0 <= i1 && j0 <= j1;
    const TileAcc t = tile_begin(tile, box, any, i0, i1, j0, j1);
    double sx = 0.0, sy = 0.0;
    if (valid) {
        for (int i = i0; i <= i1; i++) {
            const double wx = wdelta(bp.dcx, bp.dcy, i, g.dx, x);
            const double dwx = wdelta_db(bp.dcx, bp.dcy, i, g.dx, x);
            for (int j = j0; j <= j1; j++) {
                const double wy = wdelta(bp.dcx, bp.dcz, j, g.dy, y);
                const double dwy = wdelta_db(bp.dcx, bp.dcz, j, g.dy, y);
                if (any) tile_add(t, Lp0, g.pitch, i, j, lam * wx * wy);
                const double p = (double)p0.at(i, j);
                sx += p * dwx * wy;
                sy += p * wx * dwy;
            }
        }
        Lx[idx] += lam * sx;
        Ly[idx] += lam * sy;
    }
    tile_end(t, box, Lp0, g.pitch);
}

// Adjoint of k_motion: x' = x + dt vLx(x, y; vx), y' = y + dt vLy(x, y; vy)
// at the step's start positions (xs, ys); Lx, Ly hold dJ/dx', dJ/dy' and
// become dJ/dx, dJ/dy.  The node adjoints Lvx, Lvy are summed through the
// workgroup's LDS tile, one array after the other.
template <typename F>
__global__ __launch_bounds__(256) void k_adj_motion(GridParams g, BubbleParams bp, int n,
                                                     const double *xs, const double *ys,
                                                     Rect<F> vx, Rect<F> vy, double *Lvx,
                                                     double *Lvy, double *Lx, double *Ly) {
    __shared__ double tile[kAdjTile];
    __shared__ int box[4];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = idx < n;
    double x = 0.0, y = 0.0, lx = 0.0, ly = 0.0;
    int ibmx = 0, ibmy = 0, ibnx = 0, ibny = 0;
    const double dt = g.dt;
    if (valid) {
        x = xs[idx];
        y = ys[idx];
        lx = Lx[idx];
        ly = Ly[idx];
        ibmx = bin_idx(x, g.rdx, 1.0);
        ibmy = bin_idx(y, g.rdy, 1.0);
        ibnx = bin_idx(x, g.rdx, 0.5);
        ibny = bin_idx(y, g.rdy, 0.5);
    }
    double dvx_dx = 0.0, dvx_dy = 0.0, dvy_dx = 0.0, dvy_dy = 0.0;
    {  // vx nodes: columns ibm + [nbs, nbe] in [-1, N+1], rows ibn + [mbs, mbe] in [0, M+1]
        const int i0 = max(ibmx + bp.nbs, -1), i1 = min(ibmx + bp.nbe, g.N + 1);
        const int j0 = max(ibny + bp.mbs, 0), j1 = min(ibny + bp.mbe, g.M + 1);
        const bool any = valid && lx != 0.0 && i0 <= i1 && j0 <= j1;
        const TileAcc t = tile_begin(tile, box, any, i0, i1, j0, j1);
        if (valid)
            for (int i = i0; i <= i1; i++) {
                const double Dx = wdelta(bp.dcx, bp.dcy, i, g.dx, x);
                const double dDx = wdelta_db(bp.dcx, bp.dcy, i, g.dx, x);
                for (int j = j0; j <= j1; j++) {
                    const double Dy = wdelta(bp.dcx, bp.dcz, j, g.dy, y);
                    const double dDy = wdelta_db(bp.dcx, bp.dcz, j, g.dy, y);
                    if (any) tile_add(t, Lvx, g.pitch, i, j, dt * lx * Dx * Dy);
                    const double v = (double)vx.at(i, j);
                    dvx_dx += v * dDx * Dy;
                    dvx_dy += v * Dx * dDy;
                }
            }
        tile_end(t, box, Lvx, g.pitch);
    }
    {  // vy nodes: columns ibn + [mbs, mbe] in [0, N+1], rows ibm + [nbs, nbe] in [-1, M+1]
        const int i0 = max(ibnx + bp.mbs, 0), i1 = min(ibnx + bp.mbe, g.N + 1);
        const int j0 = max(ibmy + bp.nbs, -1), j1 = min(ibmy + bp.nbe, g.M + 1);
        const bool any = valid && ly != 0.0 && i0 <= i1 && j0 <= j1;
        const TileAcc t = tile_begin(tile, box, any, i0, i1, j0, j1);
        if (valid)
            for (int i = i0; i <= i1; i++) {
                const double Dx = wdelta(bp.dcx, bp.dcy, i, g.dx, x);
                const double dDx = wdelta_db(bp.dcx, bp.dcy, i, g.dx, x);
                for (int j = j0; j <= j1; j++) {
                    const double Dy = wdelta(bp.dcx, bp.dcz, j, g.dy, y);
                    const double dDy = wdelta_db(bp.dcx, bp.dcz, j, g.dy, y);
                    if (any) tile_add(t, Lvy, g.pitch, i, j, dt * ly * Dx * Dy);
                    const double v = (double)vy.at(i, j);
                    dvy_dx += v * dDx * Dy;
                    dvy_dy += v * Dx * dDy;
                }
            }
        tile_end(t, box, Lvy, g.pitch);
    }
    if (valid) {
        Lx[idx] = lx + dt * (lx * dvx_dx + ly * dvy_dx);
        Ly[idx] = ly + dt * (lx * dvx_dy + ly * dvy_dy);
    }
}

// Adjoint of the RP calls (k_bubble_persistent's BubbleRadiusKernel part).
// The RP map of one Picard iteration, (R, Rn, d1R, PG, PL_p, PL_n, PL_m) ->
// (R', Rn', d1R', PG', R_t), depends only on the forward's taped inputs, so
// its Jacobian is computed for every Picard iteration of a run of steps in
// one launch (k_adj_rp_jac, after the items are ordered by k_jac_rank), and
// each reverse step applies its transpose to the step's adjoints
// (k_adj_rp_apply).  Without it every iteration's integration would be a
// launch of its own, bounded by that iteration's longest chain.

// Keys of the item order (k_jac_rank): first the items of more than
// kJacLong sub-steps, longest first (counts from kJacBins on share one key),
// whose Jacobians k_adj_rp_jac takes in forward mode, seven tangent lanes
// each, so that a long chain takes one pass; then the rest, in reverse mode
// on one lane each -- the items of a step's last Picard iteration (five
// outputs) before the others (one output, R_t), each longest first, so that
// a wavefront holds items of one kind.  kJacKeys counts the keys; the
// prefix sum at key kJacBins (the number of long items) is kept after them.
constexpr int kJacBins = 4096;
constexpr int kJacLong = 1024;
constexpr int kJacKeys = 2 * kJacBins;
constexpr int kJacTile = 4096;  // items per workgroup of k_jac_rank
// Reverse-mode blocks: every state of an item of at most kJacRevM
// sub-steps, else every kJacRevM-th and the current block's
// (kJacRevM * kJacRevC >= kJacLong).
constexpr int kJacRevM = 32, kJacRevC = kJacLong / kJacRevM;

__device__ inline int jac_key(int sub, bool last) {
    sub = max(sub, 0);
    if (sub > kJacLong) return kJacBins - 1 - min(sub, kJacBins - 1);
    return kJacBins + (last ? 0 : kJacLong + 1) + (kJacLong - sub);
}

// Two passes over tiles of kJacTile items.  PLACE = false: global counts per
// bin.  PLACE = true (after k_jac_scan turned the counts into offsets): each
// tile reserves its range of every bin with one atomic and writes its items'
// indices there.  The order inside a bin depends on timing; every item's
// result does not.
template <bool PLACE>
__global__ __launch_bounds__(256) void k_jac_rank(const JacSlot *slots, int nb,
                                                  unsigned n_items, int *bins,
                                                  unsigned *ord) {
    __shared__ int cnt[kJacKeys];
    __shared__ int base[PLACE ? kJacKeys : 1];
    for (int i = threadIdx.x; i < kJacKeys; i += 256) cnt[i] = 0;
    __syncthreads();
    const unsigned t0 = blockIdx.x * (unsigned)kJacTile;
    constexpr int kPer = kJacTile / 256;
    int bin[kPer];
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        const unsigned q = t0 + r * 256 + threadIdx.x;
        bin[r] = -1;
        if (q < n_items) {
            const unsigned sl = q / (unsigned)nb, idx = q - sl * (unsigned)nb;
            bin[r] = jac_key(slots[sl].sub[idx], slots[sl].last);
            if (!PLACE) atomicAdd(&cnt[bin[r]], 1);
        }
    }
    if constexpr (!PLACE) {
        __syncthreads();
        for (int i = threadIdx.x; i < kJacKeys; i += 256)
            if (cnt[i]) atomicAdd(&bins[i], cnt[i]);
    } else {
#pragma unroll
        for (int r = 0; r < kPer; r++)
            if (bin[r] >= 0) atomicAdd(&cnt[bin[r]], 1);
        __syncthreads();
        for (int i = threadIdx.x; i < kJacKeys; i += 256) {
            base[i] = cnt[i] ? atomicAdd(&bins[i], cnt[i]) : 0;
            cnt[i] = 0;
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kPer; r++)
            if (bin[r] >= 0)
                ord[base[bin[r]] + atomicAdd(&cnt[bin[r]], 1)] = t0 + r * 256 + threadIdx.x;
    }
}

// Exclusive prefix sum of the kJacKeys counts, in place (one workgroup);
// bins[kJacKeys] = the number of long items.
__global__ __launch_bounds__(1024) void k_jac_scan(int *bins) {
    constexpr int kPer = kJacKeys / 1024;
    __shared__ int part[1024];
    int v[kPer], sum = 0;
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        v[r] = bins[threadIdx.x * kPer + r];
        sum += v[r];
    }
    part[threadIdx.x] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int add = threadIdx.x >= off ? part[threadIdx.x - off] : 0;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    int run = part[threadIdx.x] - sum;
#pragma unroll
    for (int r = 0; r < kPer; r++) {
        bins[threadIdx.x * kPer + r] = run;
        if (threadIdx.x * kPer + r == kJacBins) bins[kJacKeys] = run;
        run += v[r];
    }
}

// One item of k_adj_rp_jac on lane d of its seven.
__device__ inline void jac_item(const BubbleParams &bp, const BubbleArrays &b, double dt_mix,
                                const JacSlot *slots, unsigned item, int d,
                                const RPFastShared &fs, unsigned *mismatch) {
    const unsigned nb = (unsigned)b.n;
    const unsigned sl = item / nb, idx = item - sl * nb;
    const JacSlot &S = slots[sl];
    const RPFastConst fk = rp_fast_const(b.R0[idx], b.PG0[idx], b.calpha[idx], bp.T0, fs);
    auto seed = [&](double v, int k) { return dd(v, d == k ? 1.0 : 0.0); };
    RPLaneF<dd> L;
    rp_begin_g<dd>(L, seed(S.rp[0][idx], 0), seed(S.rp[1][idx], 1), seed(S.rp[2][idx], 2),
                   seed(S.rp[3][idx], 3), S.rp[4][idx], S.rp[5][idx], seed(S.PLm[idx], 6),
                   seed(S.PLn[idx], 5), seed(S.PLp[idx], 4), dt_mix, bp.PL0);
    while (L.remain > 0.0 && L.counter < bp.max_substeps) rp_substep_g(L, fk, fs);
    double *J = S.jac + idx;
    J[(size_t)d * nb] = L.Rt.d;
    if (S.last) {
        J[(size_t)(7 + d) * nb] = L.Rp.d;
        J[(size_t)(14 + d) * nb] = L.Rn.d;
        J[(size_t)(21 + d) * nb] = L.d1R.d;
        J[(size_t)(28 + d) * nb] = L.PG.d;
    }
    if (d == 0 && S.sub[idx] != L.counter) atomicAdd(mismatch, 1u);
}

// One item of k_adj_rp_jac in reverse mode, on one lane (rp_jac_rev).
template <int NOUT>
__device__ inline void jac_item_rev(const BubbleParams &bp, const BubbleArrays &b, double dt_mix,
                                    const JacSlot &S, unsigned idx, const RPFastShared &fs,
                                    const RPRevBuf &B, unsigned *mismatch) {
    const RPFastConst fk = rp_fast_const(b.R0[idx], b.PG0[idx], b.calpha[idx], bp.T0, fs);
    RPLaneF<double> L0;
    rp_begin_g(L0, S.rp[0][idx], S.rp[1][idx], S.rp[2][idx], S.rp[3][idx], S.rp[4][idx],
               S.rp[5][idx], S.PLm[idx], S.PLn[idx], S.PLp[idx], dt_mix, bp.PL0);
    if (!rp_jac_rev<NOUT>(L0, S.sub[idx], S.last ? 5 : 1, bp.max_substeps, dt_mix, fk, fs, B,
                          S.jac + idx, (size_t)b.n))
        atomicAdd(mismatch, 1u);
}

// Jacobian columns of the RP map: jac[c * nb + idx] for c = 7 o + d, o = 0
// (R_t) and, in the step's last iteration, o = 1..4 (R', Rn', d1R', PG'),
// d = 0..6 (inputs R, Rn, d1R, PG, PL_p, PL_n, PL_m).  Persistent, over the
// order of k_jac_rank: first the long items (more than kJacLong sub-steps),
// seven lanes each, nine per wavefront (lane 63 idles): lane d runs the
// sub-steps in (value, tangent) arithmetic with the tangent seeded on input
// d, so a long chain takes one pass; its value part repeats the forward's
// operations exactly, so each lane takes the forward's sub-steps, and lane 0
// counts a difference from the taped count in *mismatch.  Then the other
// items, 64 per wavefront, in reverse mode (rp_jac_rev): one forward pass
// keeping the states, one reverse pass with all of the item's output
// adjoints (the wavefront's items are of one kind but at the boundary, where
// it carries five).  Per item and sub-step the reverse mode issues about a
// fifth of the forward mode's instructions.  buf holds kJacRevC + kJacRevM
// seven-double records per lane of the grid.
constexpr int kJacPerWave = 9;
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_adj_rp_jac(
    BubbleParams bp, BubbleArrays b, double dt_mix, const JacSlot *slots, const unsigned *ord,
    unsigned n_items, const int *bins, unsigned *head, unsigned *mismatch, double *buf) {
    __shared__ double vtab[kRpVSize];  // the V and Y tables
    rp_vtab_stage(bp.rp_shared, vtab);
    __syncthreads();
    const RPFastShared fs = rp_shared_load(bp.rp_shared, vtab);
    const unsigned n_long = (unsigned)bins[kJacKeys];
    const int lane = threadIdx.x & 63, slot = lane / 7, d = lane - 7 * slot;
    for (;;) {
        unsigned first = 0;
        if (lane == 0) first = atomicAdd(&head[0], (unsigned)kJacPerWave);
        first = __shfl(first, 0);
        if (first >= n_long) break;
        const unsigned q = first + slot;
        if (slot < kJacPerWave && q < n_long)
            jac_item(bp, b, dt_mix, slots, ord[q], d, fs, mismatch);
    }
    const size_t nl = (size_t)gridDim.x * blockDim.x;
    const size_t gl = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const RPRevBuf B{buf + gl, buf + (size_t)kJacRevC * 7 * nl + gl, nl, kJacRevM, kJacRevC};
    const unsigned nb = (unsigned)b.n;
    for (;;) {
        unsigned first = 0;
        if (lane == 0) first = atomicAdd(&head[1], 64u);
        first = __shfl(first, 0);
        if (first >= n_items - n_long) break;
        const unsigned q = n_long + first + lane;
        const bool has = q < n_items;
        unsigned sl = 0, idx = 0;
        if (has) {
            const unsigned item = ord[q];
            sl = item / nb;
            idx = item - sl * nb;
        }
        const bool last = has && slots[sl].last;
        if (__ballot(last) != 0ull) {
            if (has) jac_item_rev<5>(bp, b, dt_mix, slots[sl], idx, fs, B, mismatch);
        } else if (has) {
            jac_item_rev<1>(bp, b, dt_mix, slots[sl], idx, fs, B, mismatch);
        }
    }
}

// Transpose of one Picard iteration's RP map applied to its adjoints:
//   lam_d = lRt dR_t/dIn_d (+ sum_o LRe_o dOut_o/dIn_d in the last iteration)
// then LRs[d] += lam_d (d < 4), lPLp = lam_4, lPLn += lam_5, lPLm += lam_6.
__global__ __launch_bounds__(256) void k_adj_rp_apply(int nb, const double *jac, const double *lRt,
                                                       const double *LRe0, const double *LRe1,
                                                       const double *LRe2, const double *LRe3,
                                                       double *LRs0, double *LRs1, double *LRs2,
                                                       double *LRs3, double *lPLp, double *lPLn,
                                                       double *lPLm) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nb) return;
    const double *J = jac + idx;
    const size_t n = (size_t)nb;
    const double lr = lRt[idx];
    double lam[7];
#pragma unroll
    for (int d = 0; d < 7; d++) lam[d] = lr * J[d * n];
    if (LRe0) {
        const double e0 = LRe0[idx], e1 = LRe1[idx], e2 = LRe2[idx], e3 = LRe3[idx];
#pragma unroll
        for (int d = 0; d < 7; d++)
            lam[d] += e0 * J[(7 + d) * n] + e1 * J[(14 + d) * n] + e2 * J[(21 + d) * n] +
                      e3 * J[(28 + d) * n];
    }
    LRs0[idx] += lam[0];
    LRs1[idx] += lam[1];
    LRs2[idx] += lam[2];
    LRs3[idx] += lam[3];
    lPLp[idx] = lam[4];
    lPLn[idx] += lam[5];
    lPLm[idx] += lam[6];
}

// Liquid-pressure history adjoints across the step boundary:
//   Picard ran:  LA <- lPLn + LB,  LB <- lPLm
//   no Picard:   LA <- LA + LB,    LB <- 0        (PL_p was copied from PL_n)
__global__ void k_adj_pl_history(int n, int picard, double *LA, double *LB, double *lPLn,
                                 double *lPLm) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n) return;
    if (picard) {
        LA[idx] = lPLn[idx] + LB[idx];
        LB[idx] = lPLm[idx];
    } else {
        LA[idx] = LA[idx] + LB[idx];
        LB[idx] = 0.0;
    }
    lPLn[idx] = 0.0;
    lPLm[idx] = 0.0;
}

__global__ void k_fill_d(double *p, size_t n, double v) {
    const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = v;
}

__global__ void k_fill_i(int *p, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0;
}

}  // namespace

// ---- launchers ----------------------------------------------------------

template <typename F>
void launch_box_sumsq(hipStream_t s, const GridParams &g, const F *p0, const int box[4],
                      double *out) {
    hipLaunchKernelGGL(k_box_sumsq<F>, dim3(1), dim3(256), 0, s, g, p0, box[0], box[1], box[2],
                       box[3], out);
}

template <typename F>
void launch_adj_seed(hipStream_t s, const GridParams &g, Rect<F> p0, double *Lp0,
                     const int box[4], double c) {
    const int ni = box[1] - box[0] + 1, nj = box[3] - box[2] + 1;
    hipLaunchKernelGGL(k_adj_seed<F>, dim3((ni + 63) / 64, nj), dim3(64), 0, s, g, p0, Lp0, box[0],
                       box[1], box[2], box[3], c);
}

void launch_adj_pressure_ghosts(hipStream_t s, const GridParams &g, double *L,
                                const double *el_amp, const double *el_phase, double t,
                                double *gA, double *gP) {
    hipLaunchKernelGGL(k_adj_pressure_ghosts, dim3(1), dim3(1024), 0, s, g, L, el_amp, el_phase,
                       t, gA, gP);
}

template <typename F>
void launch_adj_pressure(hipStream_t s, const GridParams &g, bool first, const AdjField &a,
                         Rect<int64_t> fk, Rect<int64_t> gC, const double *extra, double *Lf,
                         double *LgC, double *LfN, double *LfM, const F *ax, const F *bx,
                         const F *ay, const F *by, const Window *win) {
    Window w{-1, g.N + 1, -1, g.M + 1};
    if (win) {
        w = Window{std::max(win->i0, -1), std::min(win->i1, g.N + 1), std::max(win->j0, -1),
                   std::min(win->j1, g.M + 1)};
        if (w.i0 > w.i1 || w.j0 > w.j1) return;
    }
    const dim3 blk(kAX, kAY),
        grid((w.i1 - w.i0 + 1 + kAX - 1) / kAX, (w.j1 - w.j0 + 1 + kAY - 1) / kAY);
    if (first)
        hipLaunchKernelGGL((k_adj_pressure<F, true>), grid, blk, 0, s, g, a.p0, a.px, a.py, a.pnx,
                           a.pny, a.vx, a.vy, fk, gC, extra, Lf, LgC, LfN, LfM, ax, bx, ay, by, w);
    else
        hipLaunchKernelGGL((k_adj_pressure<F, false>), grid, blk, 0, s, g, a.p0, a.px, a.py,
                           a.pnx, a.pny, a.vx, a.vy, fk, gC, extra, Lf, LgC, LfN, LfM, ax, bx, ay,
                           by, w);
    const Window wi{std::max(w.i0, 1), std::min(w.i1, g.N), std::max(w.j0, 1), std::min(w.j1, g.M)};
    if (wi.i0 > wi.i1 || wi.j0 > wi.j1) return;
    const dim3 gi((wi.i1 - wi.i0 + 1 + kAX - 1) / kAX, (wi.j1 - wi.j0 + 1 + kAY - 1) / kAY);
    hipLaunchKernelGGL(k_adj_zero_interior, gi, blk, 0, s, g, a.p0, a.px, a.py, wi);
}

void launch_adj_velocity_ghosts(hipStream_t s, const GridParams &g, double *Lvx, double *Lvy) {
    hipLaunchKernelGGL(k_adj_velocity_ghosts, dim3(1), dim3(1024), 0, s, g, Lvx, Lvy);
}

template <typename F>
void launch_adj_velocity(hipStream_t s, const GridParams &g, double *Lvx, double *Lvy,
                         const F *p0, Rect<int64_t> fgn, int first, const F *cxa, const F *cxb,
                         const F *cya, const F *cyb, double *Lp0, double *LfC) {
    const dim3 blk(kAX, kAY);
    const dim3 ga((g.N + 2 + kAX - 1) / kAX, (g.M + 2 + kAY - 1) / kAY);
    hipLaunchKernelGGL(k_adj_velocity_a<F>, ga, blk, 0, s, g, Lvx, Lvy, p0, fgn, first, cxb, cyb,
                       Lp0, LfC);
    const dim3 gb((g.N + 1 + kAX - 1) / kAX, (g.M + 1 + kAY - 1) / kAY);
    hipLaunchKernelGGL(k_adj_velocity_b<F>, gb, blk, 0, s, g, Lvx, Lvy, cxa, cxb, cya, cyb);
}

void launch_adj_scatter(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                        const BubbleArrays &b, const double *Rt, const double *Lf, double *lRt,
                        double *Lx, double *Ly) {
    if (b.n <= 0) return;
    hipLaunchKernelGGL(k_adj_scatter, dim3((b.n + 255) / 256), dim3(256), 0, s, g, bp, b, Rt, Lf,
                       lRt, Lx, Ly);
}

template <typename F>
void launch_adj_gather(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                       const BubbleArrays &b, Rect<F> p0, const double *lPL,
                       const double *lPL_extra, double *Lp0, double *Lx, double *Ly) {
    if (b.n <= 0) return;
    hipLaunchKernelGGL(k_adj_gather<F>, dim3((b.n + 255) / 256), dim3(256), 0, s, g, bp, b, p0,
                       lPL, lPL_extra, Lp0, Lx, Ly);
}

template <typename F>
void launch_adj_motion(hipStream_t s, const GridParams &g, const BubbleParams &bp, int n,
                       const double *xs, const double *ys, Rect<F> vx, Rect<F> vy, double *Lvx,
                       double *Lvy, double *Lx, double *Ly) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_adj_motion<F>, dim3((n + 255) / 256), dim3(256), 0, s, g, bp, n, xs, ys,
                       vx, vy, Lvx, Lvy, Lx, Ly);
}

size_t adj_rp_jac_buf_doubles(int blocks) {
    return (size_t)blocks * 256 * (kJacRevC + kJacRevM) * 7;
}

int adj_rp_jac_per_cu() {
    int nblk = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nblk, k_adj_rp_jac, 256, 0) != hipSuccess ||
        nblk < 1)
        nblk = 1;
    return nblk;
}

void launch_adj_rp_jacobians(hipStream_t s, const BubbleParams &bp, const BubbleArrays &b,
                             double dt_mix, const JacSlot *slots, int n_slots, unsigned *ord,
                             int *bins, unsigned *head, unsigned *mismatch, double *buf,
                             int max_blocks) {
    if (b.n <= 0 || n_slots <= 0) return;
    const unsigned n_items = (unsigned)((long)n_slots * b.n);
    const unsigned tiles = (n_items + kJacTile - 1) / kJacTile;
    hipLaunchKernelGGL(k_fill_i, dim3((kJacKeys + 1 + 255) / 256), dim3(256), 0, s, bins,
                       kJacKeys + 1);
    hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(256), 0, s, (int *)head, 2);
    hipLaunchKernelGGL(k_jac_rank<false>, dim3(tiles), dim3(256), 0, s, slots, b.n, n_items, bins,
                       ord);
    hipLaunchKernelGGL(k_jac_scan, dim3(1), dim3(1024), 0, s, bins);
    hipLaunchKernelGGL(k_jac_rank<true>, dim3(tiles), dim3(256), 0, s, slots, b.n, n_items, bins,
                       ord);
    const long blocks = std::max<long>(1, std::min<long>(max_blocks, ((long)n_items + 255) / 256));
    hipLaunchKernelGGL(k_adj_rp_jac, dim3((unsigned)blocks), dim3(256), 0, s, bp, b, dt_mix, slots,
                       ord, n_items, bins, head, mismatch, buf);
}

void launch_adj_rp_apply(hipStream_t s, int nb, const double *jac, const double *lRt,
                         double *const LRe[4], double *const LRs[4], double *lPLp, double *lPLn,
                         double *lPLm) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_adj_rp_apply, dim3((nb + 255) / 256), dim3(256), 0, s, nb, jac, lRt,
                       LRe[0], LRe[1], LRe[2], LRe[3], LRs[0], LRs[1], LRs[2], LRs[3], lPLp, lPLn,
                       lPLm);
}

void launch_adj_pl_history(hipStream_t s, int n, int picard, double *LA, double *LB,
                           double *lPLn, double *lPLm) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_adj_pl_history, dim3((n + 255) / 256), dim3(256), 0, s, n, picard, LA, LB,
                       lPLn, lPLm);
}

void launch_fill_d(hipStream_t s, double *p, size_t n, double v) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
}

#define BRP_ADJ_INSTANTIATE(F)                                                                  \
    template void launch_box_sumsq<F>(hipStream_t, const GridParams &, const F *, const int[4], \
                                      double *);                                                \
    template void launch_adj_seed<F>(hipStream_t, const GridParams &, Rect<F>, double *,        \
                                     const int[4], double);                                     \
    template void launch_adj_pressure<F>(hipStream_t, const GridParams &, bool, const AdjField &, \
                                         Rect<int64_t>, Rect<int64_t>, const double *,          \
                                         double *, double *, double *, double *, const F *,     \
                                         const F *, const F *, const F *, const Window *);      \
    template void launch_adj_velocity<F>(hipStream_t, const GridParams &, double *, double *,   \
                                         const F *, Rect<int64_t>, int, const F *, const F *,   \
                                         const F *, const F *, double *, double *);             \
    template void launch_adj_gather<F>(hipStream_t, const GridParams &, const BubbleParams &,   \
                                       const BubbleArrays &, Rect<F>, const double *,           \
                                       const double *, double *, double *, double *);           \
    template void launch_adj_motion<F>(hipStream_t, const GridParams &, const BubbleParams &,   \
                                       int, const double *, const double *, Rect<F>, Rect<F>,   \
                                       double *, double *, double *, double *);
BRP_ADJ_INSTANTIATE(float)
BRP_ADJ_INSTANTIATE(double)

}  // namespace brp
