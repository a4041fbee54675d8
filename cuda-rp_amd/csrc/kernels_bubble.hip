// kernels_bubble.hip -- Lagrangian bubble kernels for gfx950.
//
// One bubble per lane.  Bubbles are stored sorted by cell, so the B x B
// scatter stencils of neighbouring lanes hit the same few cells; the
// void-fraction accumulation is int64 fixed point (value * 2^60): integer
// atomics are associative, so the scatter is bit-reproducible in any order
// (the FP64 CAS loop of the reference, bubble_CUDA_kernel.cuh:95-114, is not).
#include "brp_internal.h"
#include "rp_device.h"
#include "rp_fast.h"

#include <math.h>

#include <algorithm>

namespace brp {
namespace {

constexpr int kMaxBand = 8;

// smooth_delta_x / smooth_delta_y (bubble_CUDA_kernel.cuh:214-222)
__device__ inline double delta_w(double coef_x, double coef_k, int pos, double h, double b) {
#pragma clang fp contract(off)
    return coef_x * (1.0 + cos(coef_k * (((double)pos - 0.5) * h - b)));
}

__device__ inline int64_t to_fixed(double v, bool *bad) {
    double s = v * kFgScale;
    if (!(fabs(s) < 4.611686018427388e18)) {  // |v| < 4, NaN rejected
        *bad = true;
        return 0;
    }
    return (int64_t)__double2ll_rn(s);
}

// Block-wide reduction of (min, max) pairs into this block's slot of out
// (4 ints); k_bbox_reduce folds the slots into sc->bbox.  One plain store per
// block instead of device-scope atomics on four shared addresses, which
// serialise across the whole grid.
__device__ inline void block_bbox(int imin, int imax, int jmin, int jmax, int *out) {
    for (int off = 32; off > 0; off >>= 1) {
        imin = min(imin, __shfl_xor(imin, off, 64));
        imax = max(imax, __shfl_xor(imax, off, 64));
        jmin = min(jmin, __shfl_xor(jmin, off, 64));
        jmax = max(jmax, __shfl_xor(jmax, off, 64));
    }
    __shared__ int red[4][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[0][w] = imin;
        red[1][w] = imax;
        red[2][w] = jmin;
        red[3][w] = jmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = blockDim.x >> 6;
        for (int k = 1; k < nw; k++) {
            imin = min(imin, red[0][k]);
            imax = max(imax, red[1][k]);
            jmin = min(jmin, red[2][k]);
            jmax = max(jmax, red[3][k]);
        }
        out[4 * blockIdx.x + 0] = imin;
        out[4 * blockIdx.x + 1] = imax;
        out[4 * blockIdx.x + 2] = jmin;
        out[4 * blockIdx.x + 3] = jmax;
    }
}

__global__ __launch_bounds__(256) void k_bbox_reduce(const int *blk, int nblk, StepScalars *sc) {
    int imin = 0x7fffffff, imax = -0x7fffffff, jmin = 0x7fffffff, jmax = -0x7fffffff;
    for (int k = threadIdx.x; k < nblk; k += blockDim.x) {
        imin = min(imin, blk[4 * k + 0]);
        imax = max(imax, blk[4 * k + 1]);
        jmin = min(jmin, blk[4 * k + 2]);
        jmax = max(jmax, blk[4 * k + 3]);
    }
    __shared__ int red[4][256];
    red[0][threadIdx.x] = imin;
    red[1][threadIdx.x] = imax;
    red[2][threadIdx.x] = jmin;
    red[3][threadIdx.x] = jmax;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] = min(red[0][threadIdx.x], red[0][threadIdx.x + s]);
            red[1][threadIdx.x] = max(red[1][threadIdx.x], red[1][threadIdx.x + s]);
            red[2][threadIdx.x] = min(red[2][threadIdx.x], red[2][threadIdx.x + s]);
            red[3][threadIdx.x] = max(red[3][threadIdx.x], red[3][threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0][0] <= red[1][0]) {
        sc->bbox[0] = red[0][0];
        sc->bbox[1] = red[1][0];
        sc->bbox[2] = red[2][0];
        sc->bbox[3] = red[3][0];
    }
}

// BubbleInterpolationVelocityKernel + v_B <- v_L + BubbleMotionKernel
// (bubble_CUDA_kernel.cuh:398-468, 1282-1297), fused per bubble; also the
// bounding box of the new scatter stencils for the Picard window.
template <typename F>
__global__ __launch_bounds__(256) void k_motion(GridParams g, BubbleParams bp, BubbleArrays b,
                                                 const F *__restrict__ vx,
                                                 const F *__restrict__ vy, StepScalars *sc) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int imin = 0x7fffffff, imax = -0x7fffffff, jmin = 0x7fffffff, jmax = -0x7fffffff;
    if (idx < b.n) {
        const double x = b.x[idx], y = b.y[idx];
        const int ibmx = bin_index(x, g.rdx, 1.0), ibmy = bin_index(y, g.rdy, 1.0);
        const int ibnx = bin_index(x, g.rdx, 0.5), ibny = bin_index(y, g.rdy, 0.5);
        // the four weight sets (x at ibm / ibn columns, y at ibn / ibm rows),
        // each evaluated once
        const int nb = bp.mbe - bp.mbs + 1;
        double wxm[kMaxBand], wyn[kMaxBand], wxn[kMaxBand], wym[kMaxBand];
        for (int a = 0; a < nb; a++) {
            wxm[a] = delta_w(bp.dcx, bp.dcy, ibmx + bp.nbs + a, g.dx, x);
            wxn[a] = delta_w(bp.dcx, bp.dcy, ibnx + bp.mbs + a, g.dx, x);
            wyn[a] = delta_w(bp.dcx, bp.dcz, ibny + bp.mbs + a, g.dy, y);
            wym[a] = delta_w(bp.dcx, bp.dcz, ibmy + bp.nbs + a, g.dy, y);
        }
        double vLx = 0.0, vLy = 0.0;
        for (int a = 0; a < nb; a++) {
            const int i = ibmx + bp.nbs + a;
            for (int c = 0; c < nb; c++) {
                const int j = ibny + bp.mbs + c;
                if (i >= -1 && i <= g.N + 1 && j >= 0 && j <= g.M + 1)
                    vLx += (double)vx[frame_off(i, j, g.pitch)] * wxm[a] * wyn[c];
            }
        }
        for (int a = 0; a < nb; a++) {
            const int i = ibnx + bp.mbs + a;
            for (int c = 0; c < nb; c++) {
                const int j = ibmy + bp.nbs + c;
                if (i >= 0 && i <= g.N + 1 && j >= -1 && j <= g.M + 1)
                    vLy += (double)vy[frame_off(i, j, g.pitch)] * wxn[a] * wym[c];
            }
        }
        b.vx[idx] = vLx;
        b.vy[idx] = vLy;
        double nx, ny;
        {
#pragma clang fp contract(off)
            nx = x + vLx * g.dt;
            ny = y + vLy * g.dt;
        }
        b.x[idx] = nx;
        b.y[idx] = ny;
        const int kx = bin_index(nx, g.rdx, 0.5), ky = bin_index(ny, g.rdy, 0.5);
        imin = max(kx + bp.mbs, 0);
        imax = min(kx + bp.mbe, g.N + 1);
        jmin = max(ky + bp.mbs, 0);
        jmax = min(ky + bp.mbe, g.M + 1);
        if (imin > imax || jmin > jmax) {
            imin = 0x7fffffff;
            imax = -0x7fffffff;
            jmin = 0x7fffffff;
            jmax = -0x7fffffff;
        }
    }
    block_bbox(imin, imax, jmin, jmax, b.bbox_blk);
}

// A bubble's volume per cell area, the quantity the scatter spreads
// (VoidFractionReverseLookupKernel, bubble_CUDA_kernel.cuh:576-619).
__device__ inline double fg_temp_of(const GridParams &g, const BubbleParams &bp, double n_B,
                                    double R, double x) {
#pragma clang fp contract(off)
    double v = bp.pi4r3 * n_B * (R * R * R) * g.rdx * g.rdy;
    if (bp.cylindrical) v = bp.pi4r3 * n_B * (R * R * R) * g.rdx * g.rdy / x;
    return v;
}

// Gather of p0 at a bubble (BubbleInterpolationScalarKernel,
// bubble_CUDA_kernel.cuh:361-395) and the bubble's B x B stencil weights.
template <typename F>
__device__ inline double gather_p0(const GridParams &g, const BubbleParams &bp, const F *p0,
                                   double x, double y, int ibnx, int ibny) {
    const int nb = bp.mbe - bp.mbs + 1;
    double PL = 0.0;
    for (int a = 0; a < nb; a++) {
        const int i = ibnx + bp.mbs + a;
        const double Dx = delta_w(bp.dcx, bp.dcy, i, g.dx, x);
        for (int c = 0; c < nb; c++) {
            const int j = ibny + bp.mbs + c;
            const double Dy = delta_w(bp.dcx, bp.dcz, j, g.dy, y);
            if (i >= 0 && i <= g.N + 1 && j >= 0 && j <= g.M + 1) {
#pragma clang fp contract(off)
                PL += (double)p0[frame_off(i, j, g.pitch)] * Dx * Dy;
            }
        }
    }
    return PL;
}

// Per-wavefront aggregation of the void-fraction scatter in LDS: an
// open-addressing table of (frame offset, int64 sum) pairs.  A lane adds its
// stencil's contributions to the table; the wavefront flushes the table to
// the grid with one global atomic per distinct cell when it is half full and
// when the wavefront ends.  Bubbles reach a wavefront in batches of
// consecutive cell-sorted indices, so its table sees few distinct cells: with
// clustered bubbles (thousands per cell) this replaces thousands of
// same-address device atomics by one, and int64 addition keeps the result
// identical in any order.  A contribution that finds no slot in kAggProbe
// probes goes to the grid directly.
constexpr int kAggSlots = 512;  // per wavefront (a power of two)
constexpr int kAggProbe = 16;
struct WaveAgg {
    unsigned long long val[kAggSlots];
    int key[kAggSlots];  // frame offset, -1 when empty
    int count;           // occupied slots
};

__device__ inline void agg_clear(WaveAgg &A, int lane) {
    for (int k = lane; k < kAggSlots; k += 64) {
        A.key[k] = -1;
        A.val[k] = 0ull;
    }
    if (lane == 0) A.count = 0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
}

#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) unsigned long long lds_u64_t;
#else
typedef unsigned long long lds_u64_t;  // the host pass only parses kernel bodies
#endif
__device__ inline lds_u64_t *lds_u64(unsigned long long *p) { return (lds_u64_t *)p; }

__device__ inline void agg_add(WaveAgg &A, int cell, unsigned long long q, int64_t *fg) {
    int h = (int)(((unsigned)cell * 2654435761u) >> 23) & (kAggSlots - 1);
    for (int probe = 0; probe < kAggProbe; probe++) {
        const int k = atomicCAS(&A.key[h], -1, cell);
        if (k == -1) atomicAdd(&A.count, 1);
        if (k == -1 || k == cell) {
            __hip_atomic_fetch_add(lds_u64(&A.val[h]), q, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_WAVEFRONT);
            return;
        }
        h = (h + 1) & (kAggSlots - 1);
    }
    atomicAdd((unsigned long long *)&fg[cell], q);
}

// Whole wavefront: the table's sums to the grid, then an empty table.
__device__ inline void agg_flush(WaveAgg &A, int lane, int64_t *fg) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (int k = lane; k < kAggSlots; k += 64) {
        const int c = A.key[k];
        if (c >= 0) {
            const unsigned long long v = A.val[k];
            if (v) atomicAdd((unsigned long long *)&fg[c], v);
        }
    }
    agg_clear(A, lane);
}

// VoidFractionReverseLookupKernel's per-bubble scatter (bubble_CUDA_kernel.cuh:
// 576-619): fg_temp times the B x B smooth-delta weights, as int64 fixed
// point, added to the wavefront's table.
__device__ inline void scatter_bubble_agg(const GridParams &g, const BubbleParams &bp, double x,
                                          double y, int ibnx, int ibny, double fg_temp,
                                          WaveAgg &A, int64_t *fg, bool *bad) {
    const int nb = bp.mbe - bp.mbs + 1;
    double wx[kMaxBand], wy[kMaxBand];
    for (int a = 0; a < nb; a++) {
        wx[a] = delta_w(bp.dcx, bp.dcy, ibnx + bp.mbs + a, g.dx, x);
        wy[a] = delta_w(bp.dcx, bp.dcz, ibny + bp.mbs + a, g.dy, y);
    }
    for (int a = 0; a < nb; a++) {
        const int i = ibnx + bp.mbs + a;
        for (int c = 0; c < nb; c++) {
            const int j = ibny + bp.mbs + c;
            if (i >= 0 && i <= g.N + 1 && j >= 0 && j <= g.M + 1) {
                double v;
                {
#pragma clang fp contract(off)
                    v = fg_temp * wx[a] * wy[c];
                }
                const int64_t q = to_fixed(v, bad);
                if (q != 0) agg_add(A, (int)frame_off(i, j, g.pitch), (unsigned long long)q, fg);
            }
        }
    }
}

__device__ inline unsigned long long lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// One Picard iteration for every bubble, fused (after k_queue_keys gathered
// PL_p, BubbleInterpolationScalarKernel, kernel.cuh:361-395):
//   BubbleRadiusKernel              (RP from the step's start state, 471-554)
//   VoidFractionReverseLookupKernel (scatter R_t into fg_target, 576-619)
// A bubble whose key is -1 keeps the previous iteration's results (the same
// inputs) and is only scattered.
//
// Persistent kernel with lane-level work refill.  The number of RP sub-steps
// differs by orders of magnitude between bubbles (about 11 at rest, thousands
// through an inertial collapse), so one bubble per lane would run each
// wavefront at the pace of its slowest lane.  Here every lane integrates one
// sub-step per loop pass; lanes whose bubbles reached the sync time are
// written back and refilled with the next bubbles once kRefillBatch of them
// have finished.  Bubbles are handed to wavefronts in
// batches of 64 consecutive (cell-sorted) indices through one atomic counter
// per Picard iteration, so a wavefront costs the sum, not the max, of its
// lanes' sub-steps.  Per bubble the arithmetic is exactly that of the
// reference kernel, so results do not depend on the schedule.
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(4))) const RPLaunch crpl;  // scalar-loadable
#else
typedef const RPLaunch crpl;  // the host pass only parses kernel bodies
#endif
// The launch parameters, through a pointer opaque to the optimiser so that
// each use site loads what it needs (scalar loads) rather than the kernel
// keeping every field live across the sub-step loop.
__device__ inline crpl *rpl(const RPLaunch *p) {
    crpl *q = (crpl *)p;
    asm volatile("" : "+s"(q));
    return q;
}

// Finished lanes a wavefront collects before it leaves the sub-step loop to
// write them back and refill them (4096^2 clustered on the lean grid, ms per
// step: 16 24.5, 24 23.3, 32 23.0, 40 23.2, 48 24.2; 2048^2 equal within 1 %
// from 24 to 48).
constexpr int kRefillBatch = 32;

template <typename F, bool FAST>
__global__ __launch_bounds__(256) void k_bubble_persistent(const RPLaunch *kp, int max_substeps,
                                                            const double *rp_shared, int cur,
                                                            int plm, int pln, int plp,
                                                            int64_t *__restrict__ fg_target,
                                                            StepScalars *sc, int slot,
                                                            int solo_cap) {
    const int lane = threadIdx.x & 63;
    const int nxt = cur ^ 1;
    __shared__ WaveAgg agg_all[4];  // 25 KB: one table per wavefront
    __shared__ double vtab[FAST ? kRpVSize : 1];  // 50 KB: the V and Y tables
    WaveAgg &agg = agg_all[threadIdx.x >> 6];
    agg_clear(agg, lane);
    // The shared RP constants go through LDS into vector registers, held for
    // the whole kernel (247 VGPRs, two wavefronts per SIMD still fit): read
    // from the constant table instead, each sub-step waited on its scalar
    // loads (no scalar registers are free to hold them).
    __shared__ double shc[16];
    if (FAST) {
        rp_vtab_stage(rp_shared, vtab);
        if (threadIdx.x < 15) shc[threadIdx.x] = rp_shared[threadIdx.x];
        __syncthreads();
    }
    RPFastShared fsh{};
    if (FAST) {
        const ldouble *t = (const ldouble *)shc;
        fsh = RPFastShared{t[0], t[1], t[2],  t[3],  t[4],  t[5],  t[6], t[7],
                           t[8], t[9], t[10], t[11], t[12], t[13], t[14], vtab};
    }
    RPShared sh{};
    if (!FAST) {
        const BubbleParams &bp = rpl(kp)->bp;
        sh = RPShared{bp.PL0, bp.T0, bp.gam, bp.sig, bp.mu, bp.rho, bp.K0, bp.dt0, bp.pi};
    }
    RPFastConst fk{0, 0, 0, 0, 0, 0, 0, 0};
    // The E longest bubbles (the queue's first entries) run one per workgroup
    // in lane 0 of its first wavefront, alone: a wavefront issues the same
    // instructions whether one lane or 64 are active, so a lone lane skips
    // the paths the other lanes would need (the thermal impedance is a
    // closed form instead of a table lookup for s > 20), and it never waits
    // for the other lanes' refills.  The wavefront runs at raised issue priority
    // until that bubble is done, then joins the shared queue, which starts
    // after the E dedicated entries.
    int E, nb;
    {
        const BubbleArrays &b = rpl(kp)->b;
        nb = b.n;
        E = min(min(b.part[(b.n + 255) / 256], (int)gridDim.x), solo_cap);
    }
    int idx = -1;                  // bubble of this lane, -1 when idle
    bool reuse = false;            // idx keeps its previous iteration's results
    bool pre = false;
    if (threadIdx.x == 0 && (int)blockIdx.x < E) {
        idx = rpl(kp)->b.order[blockIdx.x];
        pre = true;
    }
    bool solo = threadIdx.x < 64 && (int)blockIdx.x < E;  // wave-uniform
    if (solo) __builtin_amdgcn_s_setprio(3);
    int batch_next = 0, batch_end = 0;  // wave-uniform
    bool exhausted = false;             // wave-uniform
    RPLane L;              // operation-by-operation path
    RPLaneF<double> LF;    // shortened-chain path
    RPConst k{0.0, 0.0, 0.0, 0.0};
    double x = 0.0, y = 0.0;
    int ibnx = 0, ibny = 0;
    int sub_max = 0;
    unsigned long long sub_sum = 0;
    unsigned int done = 0;  // bubbles this lane wrote back
    unsigned int kept = 0;  // of them, with the previous iteration's results
    bool bad = false, capped_any = false;
    while (true) {
        if (solo && __ballot(idx >= 0) == 0ull) {
            solo = false;
            __builtin_amdgcn_s_setprio(0);
        }
        // ---- refill idle lanes from the wave's batch / the global queue
        bool need = idx < 0 && !solo;
        {
        crpl *P = rpl(kp);
        const GridParams &g = P->g;
        const BubbleParams &bp = P->bp;
        const BubbleArrays &b = P->b;
        unsigned long long m = __ballot(need);
        bool fresh = pre;
        pre = false;
        while (m != 0ull && !exhausted) {
            if (batch_next >= batch_end) {
                const int leader = __ffsll((long long)m) - 1;
                unsigned int base = 0;
                if (lane == leader) base = atomicAdd(&sc->work[slot], 64u);
                base = __shfl(base, leader, 64);
                if ((int)base >= nb - E) {
                    exhausted = true;
                    break;
                }
                batch_next = (int)base + E;
                batch_end = min((int)base + E + 64, nb);
            }
            const int avail = batch_end - batch_next;
            const int rank = __popcll(m & lanemask_lt(lane));
            if (need && rank < avail) {
                idx = b.order[batch_next + rank];
                need = false;
                fresh = true;
            }
            batch_next += min(avail, __popcll(m));
            m = __ballot(need);
        }
        if (fresh) {
            x = b.x[idx];
            y = b.y[idx];
            ibnx = bin_index(x, g.rdx, 0.5);
            ibny = bin_index(y, g.rdy, 0.5);
            const double PL = b.PL[plp][idx];  // gathered by k_queue_keys
            reuse = b.key[idx] < 0;            // the previous iteration's results stand
            const double R0 = b.R0[idx];
            k = RPConst{R0, R0 * R0 * R0, b.PG0[idx], b.calpha[idx]};
            if (FAST) fk = rp_fast_const(R0, k.PG0, k.calpha, bp.T0, fsh);
            if (reuse) {
                LF.remain = 0.0;
                LF.counter = 0;
                L.remain = 0.0;
                L.counter = 0;
            } else if (FAST) {
                rp_begin_g(LF, b.R[cur][idx], b.Rn[cur][idx], b.d1R[cur][idx], b.PG[cur][idx],
                           b.bdt[cur][idx], b.re[cur][idx], b.PL[plm][idx], b.PL[pln][idx], PL,
                           g.dt, bp.PL0);
            } else {
                RPState in{b.R[cur][idx], b.Rn[cur][idx], b.d1R[cur][idx],
                           b.PG[cur][idx], b.bdt[cur][idx], b.re[cur][idx]};
                rp_begin(L, in, b.PL[plm][idx], b.PL[pln][idx], PL, g.dt, sh);
            }
        }
        }
        const bool active = idx >= 0;
        if (__ballot(active) == 0ull) break;
        // ---- sub-steps until kRefillBatch lanes' bubbles have reached the
        // sync time (or the sub-step cap), or no active lane is left
        // running; finished lanes idle meanwhile.  The refill and write-back
        // code stays outside this loop, and with it the values only that
        // code needs.  Refilling as soon as one lane finished made every
        // wave with a long chain among short bubbles pay the write-back,
        // scatter, pressure gather and set-up nearly every iteration
        // (measured, as ms per step: 2048^2 16.9 -> 16.0, 4096^2 clustered
        // 46.1 -> 29.2 with a batch of 24 on the full grid).
        bool fin, capped;
        double remain;
        if (FAST && solo) {
            // a dedicated wavefront: lane 0 runs its bubble to the sync time
            // without the batch bookkeeping of the shared loop below
            // (lane 0 alone: its conditions are wave-uniform)
            if (active)
                while (cond<true>(LF.remain > 0.0 && LF.counter < max_substeps))
                    rp_substep_g<double, false, true, true>(LF, fk, fsh);
            capped = active && LF.counter >= max_substeps;
            remain = active ? LF.remain : 0.0;
            fin = active;
        } else
        while (true) {
            capped = false;
            remain = 0.0;
            if (active) {
                capped = (FAST ? LF.counter : L.counter) >= max_substeps;
                remain = FAST ? LF.remain : L.remain;
                if (remain > 0.0 && !capped) {
                    if (FAST)
                        rp_substep_g(LF, fk, fsh);
                    else
                        rp_substep(L, k, sh);
                    remain = FAST ? LF.remain : L.remain;
                }
            }
            fin = active && (!(remain > 0.0) || capped);
            const unsigned long long fm = __ballot(fin);
            if (fm != 0ull && (__popcll(fm) >= kRefillBatch || __ballot(active && !fin) == 0ull))
                break;
        }
        // ---- write back the bubbles that reached the sync time
        if (active) {
            crpl *P = rpl(kp);
            const GridParams &g = P->g;
            const BubbleParams &bp = P->bp;
            const BubbleArrays &b = P->b;
            if (fin) {
                double R_t;
                if (reuse) {  // the previous iteration's results are in place
                    R_t = b.R_t[idx];
                    kept++;
                } else {
                    RPState out;
                    double Q_B;
                    if (FAST)
                        rp_end_g(LF, g.dt, &out, &R_t, &Q_B);
                    else
                        rp_end(L, g.dt, &out, &R_t, &Q_B);
                    const int cnt = FAST ? LF.counter : L.counter;
                    b.R[nxt][idx] = out.Rp;
                    b.Rn[nxt][idx] = out.Rn;
                    b.d1R[nxt][idx] = out.d1R;
                    b.PG[nxt][idx] = out.PG;
                    b.bdt[nxt][idx] = out.dt;
                    b.re[nxt][idx] = out.re;
                    b.R_t[idx] = R_t;
                    b.Q_B[idx] = Q_B;
                    b.sub[idx] = cnt;
                    sub_max = max(sub_max, cnt);
                    sub_sum += (unsigned long long)cnt;
                    capped_any |= capped && remain > 0.0;
                }
                done++;
                bad |= !isfinite(R_t);
                scatter_bubble_agg(g, bp, x, y, ibnx, ibny,
                                   fg_temp_of(g, bp, b.n_B[idx], R_t, x), agg, fg_target, &bad);
                idx = -1;
                reuse = false;
            }
        }
        if (__builtin_amdgcn_readfirstlane(agg.count) > kAggSlots / 2) agg_flush(agg, lane, fg_target);
    }
    agg_flush(agg, lane, fg_target);
    for (int off = 32; off > 0; off >>= 1) {
        sub_max = max(sub_max, __shfl_xor(sub_max, off, 64));
        sub_sum += __shfl_xor(sub_sum, off, 64);
        done += __shfl_xor(done, off, 64);
        kept += __shfl_xor(kept, off, 64);
    }
    if (lane == 0) {
        if (done) atomicAdd(&sc->done, done);
        if (kept) atomicAdd(&sc->reused, kept);
        if (sub_sum) atomicAdd(&sc->sub_sum, sub_sum);
        if (sub_max > __hip_atomic_load(&sc->sub_max, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(&sc->sub_max, sub_max);
    }
    if (bad) atomicOr(&sc->flags, 2);
    if (capped_any) atomicOr(&sc->flags, 4);
}

// ---- frozen schedule (finite differences of the frozen-schedule map) ------
// Record: every bubble's RP integration of this Picard iteration again, from
// the same inputs (the RP start state, the pressure history and the PL_p the
// RP launch gathered), with the same sub-step function, writing each accepted
// step size to dts[off[idx] + k]; the count must equal the launch's, and the
// results the launch left (computed, or kept from the previous iteration)
// must equal these bit for bit (a difference is counted in *mismatch, and a
// count's extra sizes are dropped).
__global__ __launch_bounds__(256) void k_rp_record(const RPLaunch *kp, int cur, int plm, int pln,
                                                    int plp, const int64_t *off, double *dts,
                                                    unsigned int *mismatch) {
    const RPLaunch &P = *kp;
    __shared__ double vtab[kRpVSize];
    rp_vtab_stage(P.bp.rp_shared, vtab);
    __syncthreads();
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const BubbleArrays &b = P.b;
    if (idx >= b.n) return;
    const RPFastShared fs = rp_shared_load(P.bp.rp_shared, vtab);
    const RPFastConst fk = rp_fast_const(b.R0[idx], b.PG0[idx], b.calpha[idx], P.bp.T0, fs);
    RPLaneF<double> L;
    rp_begin_g(L, b.R[cur][idx], b.Rn[cur][idx], b.d1R[cur][idx], b.PG[cur][idx],
               b.bdt[cur][idx], b.re[cur][idx], b.PL[plm][idx], b.PL[pln][idx], b.PL[plp][idx],
               P.g.dt, P.bp.PL0);
    const int64_t o = off[idx], cnt = off[idx + 1] - o;
    while (L.remain > 0.0 && L.counter < P.bp.max_substeps) {
        rp_substep_g(L, fk, fs);
        if (L.counter <= cnt) dts[o + L.counter - 1] = L.dtL;
    }
    RPState out;
    double R_t, Q_B;
    rp_end_g(L, P.g.dt, &out, &R_t, &Q_B);
    const int nxt = cur ^ 1;
    const bool same = R_t == b.R_t[idx] && Q_B == b.Q_B[idx] && out.Rp == b.R[nxt][idx] &&
                      out.Rn == b.Rn[nxt][idx] && out.d1R == b.d1R[nxt][idx] &&
                      out.PG == b.PG[nxt][idx] && out.dt == b.bdt[nxt][idx] &&
                      out.re == b.re[nxt][idx];
    if (L.counter != cnt || !same) atomicAdd(mismatch, 1u);
}

// Replay: one Picard iteration of every bubble with the recorded step sizes
// dts[off[idx] .. off[idx + 1]) instead of the step-size control: the gather,
// the RP integration and the aggregated void-fraction scatter of the
// persistent kernel, one bubble per lane.  A bubble that has not reached the
// sync time when its recorded sizes run out sets the sub-step cap flag.
template <typename F>
__global__ __launch_bounds__(256) void k_rp_replay(const RPLaunch *kp, int cur, int plm, int pln,
                                                    int plp, const F *__restrict__ p0,
                                                    int64_t *__restrict__ fg_target,
                                                    StepScalars *sc, const int64_t *off,
                                                    const double *dts) {
    const RPLaunch &P = *kp;
    const GridParams &g = P.g;
    const BubbleParams &bp = P.bp;
    const BubbleArrays &b = P.b;
    const int lane = threadIdx.x & 63;
    __shared__ WaveAgg agg_all[4];
    __shared__ double vtab[kRpVSize];
    WaveAgg &agg = agg_all[threadIdx.x >> 6];
    agg_clear(agg, lane);
    rp_vtab_stage(bp.rp_shared, vtab);
    __syncthreads();
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    const int nxt = cur ^ 1;
    bool bad = false, capped = false;
    int cnt = 0;
    unsigned int done = 0;
    if (idx < b.n) {
        const double x = b.x[idx], y = b.y[idx];
        const int ibnx = bin_index(x, g.rdx, 0.5), ibny = bin_index(y, g.rdy, 0.5);
        const double PL = gather_p0(g, bp, p0, x, y, ibnx, ibny);
        b.PL[plp][idx] = PL;
        const RPFastShared fs = rp_shared_load(bp.rp_shared, vtab);
        const RPFastConst fk = rp_fast_const(b.R0[idx], b.PG0[idx], b.calpha[idx], bp.T0, fs);
        RPLaneF<double> L;
        rp_begin_g(L, b.R[cur][idx], b.Rn[cur][idx], b.d1R[cur][idx], b.PG[cur][idx],
                   b.bdt[cur][idx], b.re[cur][idx], b.PL[plm][idx], b.PL[pln][idx], PL, g.dt,
                   bp.PL0);
        const int64_t o = off[idx], n = off[idx + 1] - o;
        for (int64_t k = 0; k < n && L.remain > 0.0; k++)
            rp_substep_g<double, true>(L, fk, fs, dts[o + k]);
        capped = L.remain > 0.0;
        RPState out;
        double R_t, Q_B;
        rp_end_g(L, g.dt, &out, &R_t, &Q_B);
        b.R[nxt][idx] = out.Rp;
        b.Rn[nxt][idx] = out.Rn;
        b.d1R[nxt][idx] = out.d1R;
        b.PG[nxt][idx] = out.PG;
        b.bdt[nxt][idx] = out.dt;
        b.re[nxt][idx] = out.re;
        b.R_t[idx] = R_t;
        b.Q_B[idx] = Q_B;
        b.sub[idx] = L.counter;
        cnt = L.counter;
        done = 1;
        bad |= !isfinite(R_t);
        scatter_bubble_agg(g, bp, x, y, ibnx, ibny, fg_temp_of(g, bp, b.n_B[idx], R_t, x), agg,
                           fg_target, &bad);
    }
    agg_flush(agg, lane, fg_target);
    unsigned long long sum = (unsigned long long)cnt;
    for (int o2 = 32; o2 > 0; o2 >>= 1) {
        cnt = max(cnt, __shfl_xor(cnt, o2, 64));
        sum += __shfl_xor(sum, o2, 64);
        done += __shfl_xor(done, o2, 64);
    }
    if (lane == 0) {
        if (done) atomicAdd(&sc->done, done);
        if (sum) atomicAdd(&sc->sub_sum, sum);
        atomicMax(&sc->sub_max, cnt);
    }
    if (bad) atomicOr(&sc->flags, 2);
    if (capped) atomicOr(&sc->flags, 4);
}

// The initial void fraction (calculate_void_fraction at set-up,
// bubble_CUDA.cu:180-185): scatter of R_t = R0.
__global__ __launch_bounds__(256) void k_scatter_initial(GridParams g, BubbleParams bp,
                                                          BubbleArrays b, int64_t *fg,
                                                          StepScalars *sc) {
    __shared__ WaveAgg agg_all[4];
    WaveAgg &agg = agg_all[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    agg_clear(agg, lane);
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (idx < b.n) {
        const double x = b.x[idx], y = b.y[idx];
        const int ibnx = bin_index(x, g.rdx, 0.5), ibny = bin_index(y, g.rdy, 0.5);
        scatter_bubble_agg(g, bp, x, y, ibnx, ibny, fg_temp_of(g, bp, b.n_B[idx], b.R_t[idx], x),
                           agg, fg, &bad);
    }
    agg_flush(agg, lane, fg);
    if (bad) atomicOr(&sc->flags, 2);
}

// The scatter of given clamped radii (the persistent kernel's scatter without
// its RP integration): every bubble's B x B stencil at its current position.
__global__ __launch_bounds__(256) void k_scatter_radii(GridParams g, BubbleParams bp,
                                                        BubbleArrays b, const double *Rt,
                                                        int64_t *fg, StepScalars *sc) {
    __shared__ WaveAgg agg_all[4];
    WaveAgg &agg = agg_all[threadIdx.x >> 6];
    const int lane = threadIdx.x & 63;
    agg_clear(agg, lane);
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    if (idx < b.n) {
        const double x = b.x[idx], y = b.y[idx];
        const int ibnx = bin_index(x, g.rdx, 0.5), ibny = bin_index(y, g.rdy, 0.5);
        scatter_bubble_agg(g, bp, x, y, ibnx, ibny, fg_temp_of(g, bp, b.n_B[idx], Rt[idx], x),
                           agg, fg, &bad);
    }
    agg_flush(agg, lane, fg);
    if (bad) atomicOr(&sc->flags, 2);
}

// The scatter of the current clamped radii with another accumulation than
// the int64 fixed point, for comparing accumulation schemes (diagnostic):
//   mode 1: FP64 atomic adds into out (double), the arithmetic of the
//           reference's CAS loop (bubble_CUDA_kernel.cuh:95-114) -- the
//           result depends on the order the adds land in;
//   mode 2: compensated FP32: atomic adds into a float sum; the add returns
//           the old value, so its rounding error is recovered exactly
//           (TwoSum) and added, with the contribution's own rounding to
//           FP32, into a second float array comp; the value is sum + comp;
//   mode 3: plain FP32 atomic adds into the float sum.
// out holds `frame` doubles (mode 1) or 2 * frame floats (modes 2, 3: sum,
// then comp) and must be zero on entry.
template <int MODE>
__global__ __launch_bounds__(256) void k_scatter_variant(GridParams g, BubbleParams bp,
                                                         BubbleArrays b, void *out,
                                                         int64_t frame) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= b.n) return;
    const double x = b.x[idx], y = b.y[idx];
    const int ibnx = bin_index(x, g.rdx, 0.5), ibny = bin_index(y, g.rdy, 0.5);
    const double fg_temp = fg_temp_of(g, bp, b.n_B[idx], b.R_t[idx], x);
    const int nb = bp.mbe - bp.mbs + 1;
    for (int a = 0; a < nb; a++) {
        const int i = ibnx + bp.mbs + a;
        const double wx = delta_w(bp.dcx, bp.dcy, i, g.dx, x);
        for (int c = 0; c < nb; c++) {
            const int j = ibny + bp.mbs + c;
            if (i < 0 || i > g.N + 1 || j < 0 || j > g.M + 1) continue;
            double v;
            {
#pragma clang fp contract(off)
                v = fg_temp * wx * delta_w(bp.dcx, bp.dcz, j, g.dy, y);
            }
            const int64_t o = frame_off(i, j, g.pitch);
            if (MODE == 1) {
                atomicAdd((double *)out + o, v);
            } else {
                float *sum = (float *)out, *comp = (float *)out + frame;
                const float vf = (float)v;
                const float old = atomicAdd(sum + o, vf);
                if (MODE == 2) {
#pragma clang fp contract(off)
                    const float s = old + vf;  // the value the atomic stored
                    const float bv = s - old;
                    const float av = s - bv;
                    const float err = (old - av) + (vf - bv);
                    const float ev = (float)(v - (double)vf);
                    atomicAdd(comp + o, err + ev);
                }
            }
        }
    }
}

// VoidFractionCylinderKernel (bubble_CUDA_kernel.cuh:559-573) for order 2:
// the axis fold f_g(0, j) += f_g(1, j); f_g(1, j) = f_g(0, j), exact in int64.
__global__ void k_cyl_fold(GridParams g, int64_t *fg) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j > g.M + 1) return;
    const int64_t o0 = frame_off(0, j, g.pitch), o1 = frame_off(1, j, g.pitch);
    const int64_t v = fg[o0] + fg[o1];
    fg[o0] = v;
    fg[o1] = v;
}

__global__ void k_copy_d(double *dst, const double *src, int n) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}

// Exports ibm / ibn (which = 0 ibm.x, 1 ibm.y, 2 ibn.x, 3 ibn.y) as doubles.
__global__ void k_bubble_index(GridParams g, BubbleArrays b, double *out, int which) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= b.n) return;
    const double v = (which & 1) ? b.y[idx] : b.x[idx];
    const double r = (which & 1) ? g.rdy : g.rdx;
    out[idx] = (double)bin_index(v, r, which < 2 ? 1.0 : 0.5);
}

}  // namespace

template <typename F>
void launch_motion(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                   const BubbleArrays &b, const F *vx, const F *vy, StepScalars *sc) {
    if (b.n <= 0) return;
    const int nblk = (b.n + 255) / 256;
    hipLaunchKernelGGL(k_motion<F>, dim3(nblk), dim3(256), 0, s, g, bp, b, vx, vy, sc);
    hipLaunchKernelGGL(k_bbox_reduce, dim3(1), dim3(256), 0, s, b.bbox_blk, nblk, sc);
}

// Workgroups of the persistent RP kernel: as many as fit (two per CU) when
// the launch is throughput-bound; one per CU ("lean") when a few long
// sub-step chains bound it, so that each dedicated wavefront has a SIMD of
// its own and fewer shared-queue wavefronts compete with it.  Measured with
// the void-fraction scatter aggregated in LDS: 2048^2, 2^20 bubbles during
// cavitation (a 3500-sub-step chain, 430 sub-steps per lean-grid lane):
// lean 23.0 vs full 26.2 ms of RP per step; 4096^2, 2^22 clustered bubbles
// at 3 MPa (a 4000-sub-step chain, several thousand per lane): lean 84 vs
// full 60 ms.  bubble_step_lean decides from the previous launch.
void rp_grid_query(int device, int *cus, int *per_cu) {
    int c = 0, p = 0;
    (void)hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, device);
    (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&p, k_bubble_persistent<float, true>, 256,
                                                       0);
    *cus = std::max(1, c);
    *per_cu = std::max(1, p);
}

bool bubble_step_lean(int sub_max, double sub_per_launch, int extreme_threshold, int cus) {
    const double lanes = 256.0 * cus;  // lanes of the lean grid
    return sub_max >= 2 * extreme_threshold && sub_per_launch < 0.6 * lanes * sub_max;
}

int bubble_step_blocks(int n_bubbles, bool lean, int cus, int per_cu) {
    return std::max(1, std::min(cus * (lean ? 1 : per_cu), (n_bubbles + 255) / 256));
}

template <typename F>
void launch_bubble_step(hipStream_t s, const RPLaunch &host, const RPLaunch *dev, int cur,
                        int plm, int pln, int plp, int64_t *fg_target,
                        StepScalars *sc, int slot, bool lean, int cus, int per_cu) {
    if (host.b.n <= 0) return;
    const dim3 grid(bubble_step_blocks(host.b.n, lean, cus, per_cu));
    const int cap = host.bp.max_substeps;
    // Dedicated wavefronts: up to one per workgroup on the lean grid, where
    // the longest chains bound the launch; CUs / 8 on the full grid, which
    // is chosen when the total work bounds it: a wavefront with one active
    // lane still takes its SIMD's issue slots, so every dedicated wavefront
    // costs a SIMD of throughput there.  4096^2, 2^22 clustered bubbles
    // (full grid): 512 dedicated 49.8, 128 46.8, 32 46.4, 0 46.2 ms/step;
    // with refills batched by kRefillBatch: 128 29.6, 64 29.2, 32 29.3,
    // 0 30.8; 2048^2 (lean grid): 256 17.2, 32 17.4, 0 19.1 ms/step.
    const int solo_cap = lean ? (int)grid.x : std::max(1, cus / 8);
    if (host.bp.rp_exact)
        hipLaunchKernelGGL((k_bubble_persistent<F, false>), grid, dim3(256), 0, s, dev, cap,
                           host.bp.rp_shared, cur, plm, pln, plp, fg_target, sc, slot,
                           solo_cap);
    else
        hipLaunchKernelGGL((k_bubble_persistent<F, true>), grid, dim3(256), 0, s, dev, cap,
                           host.bp.rp_shared, cur, plm, pln, plp, fg_target, sc, slot,
                           solo_cap);
}

// Work-queue order (longest-first scheduling: long sequential sub-step
// chains start at once and overlap everything else): bubbles whose last RP
// call took >= ext sub-steps ("extreme"), sorted by descending count, then
// those with >= thr ("heavy"), then the rest, the last two in index
// (cell-sorted) order.  Passes: per-block class counts, one-block exclusive
// scans, scatter (the extreme class into a scratch array), then the sort of
// the extreme class.
constexpr int kSortBins = 2048;  // bins of the extreme-class sort

__global__ __launch_bounds__(256) void k_part_count(const int *sub, int n, int ext, int thr,
                                                    int *cntE, int *cntH, int *sort_hist) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0)
        for (int k = threadIdx.x; k < kSortBins; k += 256) sort_hist[k] = 0;
    const int v = idx < n ? sub[idx] : 0;
    const int ce = __syncthreads_count(idx < n && v >= ext);
    const int ch = __syncthreads_count(idx < n && v >= thr && v < ext);
    if (threadIdx.x == 0) {
        cntE[blockIdx.x] = ce;
        cntH[blockIdx.x] = ch;
    }
}

// Exclusive prefix sums of the per-block class counts, in place, one
// workgroup per class (cntE, then cntH = cntE + nblk + 1): each thread sums a
// run of consecutive blocks, the workgroup scans the runs, each thread
// writes its run's prefixes; cnt[nblk] = the class total.
__global__ __launch_bounds__(1024) void k_part_scan(int *cnt0, int nblk) {
    int *cnt = cnt0 + (size_t)blockIdx.x * (nblk + 1);
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int per = (nblk + 1023) / 1024;
    const int b0 = min(nblk, (int)threadIdx.x * per), b1 = min(nblk, b0 + per);
    int sum = 0;
    for (int k = b0; k < b1; k++) sum += cnt[k];
    int x = sum;  // inclusive wave scan of the runs
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int run = x - sum;
    for (int q = 0; q < w; q++) run += wsum[q];
    for (int k = b0; k < b1; k++) {
        const int v = cnt[k];
        cnt[k] = run;
        run += v;
    }
    if (threadIdx.x == 1023) cnt[nblk] = run;
}

__device__ inline int block_rank(bool f, int *wcnt, int *total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const unsigned long long m = __ballot(f);
    __syncthreads();
    if (lane == 0) wcnt[w] = __popcll(m);
    __syncthreads();
    int before = 0, all = 0;
    for (int q = 0; q < 4; q++) {
        if (q < w) before += wcnt[q];
        all += wcnt[q];
    }
    *total = all;
    return before + __popcll(m & lanemask_lt(lane));
}

__global__ __launch_bounds__(256) void k_part_scatter(const int *sub, int n, int ext, int thr,
                                                      const int *offE, const int *offH, int nblk,
                                                      int *order, int *ext_out) {
    __shared__ int wcnt[4];
    const int idx = blockIdx.x * 256 + threadIdx.x;
    const bool valid = idx < n;
    const int v = valid ? sub[idx] : 0;
    const bool e = valid && v >= ext;
    const bool h = valid && v >= thr && v < ext;
    int ne, nh;
    const int re = block_rank(e, wcnt, &ne);
    const int rh = block_rank(h, wcnt, &nh);
    const int totE = offE[nblk], totH = offH[nblk];
    const int bE = offE[blockIdx.x], bH = offH[blockIdx.x];
    if (valid) {
        if (e)
            ext_out[bE + re] = idx;
        else if (h)
            order[totE + bH + rh] = idx;
        else  // light rank = position among the block's light bubbles
            order[totE + totH + (blockIdx.x * 256 - bE - bH) + (threadIdx.x - re - rh)] = idx;
    }
}

// The extreme class, longest first: ext_in[0, E) (E = *totE) scattered into
// order[0, E) by descending sub-step count, in bins of 8 sub-steps (order
// within a bin unspecified; results do not depend on the queue order).  A
// counting sort over many workgroups: bin counts (LDS histograms merged with
// one atomic per non-empty bin and workgroup), one-workgroup exclusive scan,
// then per 256-element chunk a range reserved per non-empty bin and LDS ranks
// within it.  hist (kSortBins ints) is cleared by k_part_count.
__device__ inline int sort_bin(int v) {  // ascending bin = descending count
    return kSortBins - 1 - min(v >> 3, kSortBins - 1);
}

__global__ __launch_bounds__(256) void k_sort_hist(const int *sub, const int *totE,
                                                   const int *ext_in, int *hist) {
    __shared__ int h[kSortBins];
    const int E = *totE;
    if ((int)blockIdx.x * 256 >= E) return;  // workgroup-uniform
    for (int k = threadIdx.x; k < kSortBins; k += 256) h[k] = 0;
    __syncthreads();
    for (int i = blockIdx.x * 256 + threadIdx.x; i < E; i += gridDim.x * 256)
        atomicAdd(&h[sort_bin(sub[ext_in[i]])], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < kSortBins; k += 256)
        if (h[k]) atomicAdd(&hist[k], h[k]);
}

__global__ __launch_bounds__(1024) void k_sort_scan(int *hist) {
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    // exclusive scan in place: thread t owns bins 2t and 2t + 1
    const int h0 = hist[2 * threadIdx.x], h1 = hist[2 * threadIdx.x + 1];
    int x = h0 + h1;  // inclusive wave scan of the pair sums
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off, 64);
        if (lane >= off) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = 0;
    for (int q = 0; q < w; q++) before += wsum[q];
    const int excl = before + x - (h0 + h1);
    hist[2 * threadIdx.x] = excl;
    hist[2 * threadIdx.x + 1] = excl + h0;
}

__global__ __launch_bounds__(256) void k_sort_scatter(const int *sub, const int *totE,
                                                      const int *ext_in, int *off, int *order) {
    __shared__ int cnt[kSortBins];  // chunk counts, then reserved bases
    __shared__ int rank[kSortBins];
    const int E = *totE;
    for (int c0 = blockIdx.x * 256; c0 < E; c0 += gridDim.x * 256) {  // workgroup-uniform
        for (int k = threadIdx.x; k < kSortBins; k += 256) {
            cnt[k] = 0;
            rank[k] = 0;
        }
        __syncthreads();
        const int i = c0 + threadIdx.x;
        int idx = -1, bin = 0;
        if (i < E) {
            idx = ext_in[i];
            bin = sort_bin(sub[idx]);
            atomicAdd(&cnt[bin], 1);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kSortBins; k += 256)
            if (cnt[k]) cnt[k] = atomicAdd(&off[k], cnt[k]);
        __syncthreads();
        if (idx >= 0) order[cnt[bin] + atomicAdd(&rank[bin], 1)] = idx;
        __syncthreads();
    }
}

// Queue keys of a Picard iteration, with the pressure gather of its bubbles
// (BubbleInterpolationScalarKernel, kernel.cuh:361-395) done here, before
// the queue is ordered; PL_p goes to PL[plp] for k_bubble_persistent.
//  * First iteration of a step: in a cavitating step the longest RP chains
//    belong to bubbles that are already collapsing fast when the step
//    starts, and they are other bubbles every step, so the previous step's
//    counts do not find them.  A bubble shrinking faster than
//    kCollapseSpeed gets the key 10 (-dR/dt) (R / R0) sqrt(kKeyRadius / R0),
//    an estimate of its sub-steps (on the 4096^2 clustered workload the
//    median count is about ten times (-dR/dt) R / R0 up to 200 m/s; ordered
//    by the key, the 16 longest chains of a step sit within its first 150
//    queue entries, 94 % of the 64 longest within the first 256, and the
//    4096 longest within the first 26 000 of 4.2 million); every other
//    bubble keeps its previous count, capped below the extreme class.
//  * Later iterations: the previous iteration's sub-step count, or -1 where
//    the gathered pressure equals the previous iteration's bit for bit.
//    That iteration integrated the same inputs (start state, pressure
//    history, PL_p), so its results stand: k_bubble_persistent only
//    scatters the bubble, from the light end of the queue.  In mixed
//    precision most cells' FP32 pressure stops changing once the Picard
//    residual is small (4096^2 clustered, a step's fifth iteration: 92-97 %
//    of the bubbles; 2048^2, its fourth: 97 %).
constexpr double kCollapseSpeed = 25.0;  // m/s
constexpr double kKeyRadius = 3.0e-6;    // m

template <typename F>
__global__ __launch_bounds__(256) void k_queue_keys(GridParams g, BubbleParams bp, BubbleArrays b,
                                                    const F *__restrict__ p0, int cur, int plp,
                                                    int slot, int ext) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= b.n) return;
    const double x = b.x[idx], y = b.y[idx];
    const double PL =
        gather_p0(g, bp, p0, x, y, bin_index(x, g.rdx, 0.5), bin_index(y, g.rdy, 0.5));
    int k;
    if (slot == 1) {
        const double v = b.d1R[cur][idx];
        k = min(b.sub[idx], ext - 1);
        if (v < -kCollapseSpeed) {
            const double R0 = b.R0[idx];
            k = (int)fmin(-10.0 * v * (b.R[cur][idx] / R0) * sqrt(kKeyRadius / R0), 1.0e6);
        }
    } else {
        k = PL == b.PL[plp][idx] ? -1 : b.sub[idx];
    }
    b.PL[plp][idx] = PL;
    b.key[idx] = k;
}

template <typename F>
void launch_queue_keys(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                       const BubbleArrays &b, const F *p0, int cur, int plp, int slot, int ext) {
    if (b.n <= 0) return;
    hipLaunchKernelGGL(k_queue_keys<F>, dim3((b.n + 255) / 256), dim3(256), 0, s, g, bp, b, p0,
                       cur, plp, slot, ext);
}
template void launch_queue_keys<double>(hipStream_t, const GridParams &, const BubbleParams &,
                                        const BubbleArrays &, const double *, int, int, int, int);
template void launch_queue_keys<float>(hipStream_t, const GridParams &, const BubbleParams &,
                                       const BubbleArrays &, const float *, int, int, int, int);

void launch_order_partition(hipStream_t s, const BubbleArrays &b, const int *key, int threshold,
                            int ext) {
    if (b.n <= 0) return;
    const int nblk = (b.n + 255) / 256;
    int *cntE = b.part, *cntH = b.part + (nblk + 1);
    hipLaunchKernelGGL(k_part_count, dim3(nblk), dim3(256), 0, s, key, b.n, ext, threshold,
                       cntE, cntH, b.sort_hist);
    hipLaunchKernelGGL(k_part_scan, dim3(2), dim3(1024), 0, s, cntE, nblk);
    hipLaunchKernelGGL(k_part_scatter, dim3(nblk), dim3(256), 0, s, key, b.n, ext, threshold,
                       cntE, cntH, nblk, b.order, b.order_tmp);
    const int sblk = std::min(1024, nblk);
    hipLaunchKernelGGL(k_sort_hist, dim3(sblk), dim3(256), 0, s, key, cntE + nblk, b.order_tmp,
                       b.sort_hist);
    hipLaunchKernelGGL(k_sort_scan, dim3(1), dim3(1024), 0, s, b.sort_hist);
    hipLaunchKernelGGL(k_sort_scatter, dim3(sblk), dim3(256), 0, s, key, cntE + nblk,
                       b.order_tmp, b.sort_hist, b.order);
}

void launch_scatter_initial(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                            const BubbleArrays &b, int64_t *fg_target, StepScalars *sc) {
    if (b.n <= 0) return;
    hipLaunchKernelGGL(k_scatter_initial, dim3((b.n + 255) / 256), dim3(256), 0, s, g, bp, b,
                       fg_target, sc);
}

void launch_scatter_radii(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                          const BubbleArrays &b, const double *Rt, int64_t *fg, StepScalars *sc) {
    if (b.n <= 0) return;
    hipLaunchKernelGGL(k_scatter_radii, dim3((b.n + 255) / 256), dim3(256), 0, s, g, bp, b, Rt, fg,
                       sc);
}

void launch_scatter_variant(hipStream_t s, const GridParams &g, const BubbleParams &bp,
                            const BubbleArrays &b, int mode, void *out, int64_t frame) {
    if (b.n <= 0) return;
    const dim3 grid((b.n + 255) / 256), blk(256);
    if (mode == 1)
        hipLaunchKernelGGL(k_scatter_variant<1>, grid, blk, 0, s, g, bp, b, out, frame);
    else if (mode == 2)
        hipLaunchKernelGGL(k_scatter_variant<2>, grid, blk, 0, s, g, bp, b, out, frame);
    else
        hipLaunchKernelGGL(k_scatter_variant<3>, grid, blk, 0, s, g, bp, b, out, frame);
}

void launch_rp_record(hipStream_t s, const RPLaunch *dev, int n, int cur, int plm, int pln,
                      int plp, const int64_t *off, double *dts, unsigned int *mismatch) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_rp_record, dim3((n + 255) / 256), dim3(256), 0, s, dev, cur, plm, pln,
                       plp, off, dts, mismatch);
}

template <typename F>
void launch_rp_replay(hipStream_t s, const RPLaunch *dev, int n, int cur, int plm, int pln,
                      int plp, const F *p0, int64_t *fg_target, StepScalars *sc,
                      const int64_t *off, const double *dts) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_rp_replay<F>, dim3((n + 255) / 256), dim3(256), 0, s, dev, cur, plm, pln,
                       plp, p0, fg_target, sc, off, dts);
}
template void launch_rp_replay<double>(hipStream_t, const RPLaunch *, int, int, int, int, int,
                                       const double *, int64_t *, StepScalars *, const int64_t *,
                                       const double *);
template void launch_rp_replay<float>(hipStream_t, const RPLaunch *, int, int, int, int, int,
                                      const float *, int64_t *, StepScalars *, const int64_t *,
                                      const double *);

void launch_cyl_fold(hipStream_t s, const GridParams &g, int64_t *fg) {
    hipLaunchKernelGGL(k_cyl_fold, dim3((g.M + 2 + 255) / 256), dim3(256), 0, s, g, fg);
}

void launch_copy_d(hipStream_t s, double *dst, const double *src, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_d, dim3((n + 255) / 256), dim3(256), 0, s, dst, src, n);
}

void launch_bubble_index(hipStream_t s, const GridParams &g, const BubbleArrays &b, double *out,
                         int which) {
    if (b.n <= 0) return;
    hipLaunchKernelGGL(k_bubble_index, dim3((b.n + 255) / 256), dim3(256), 0, s, g, b, out, which);
}

template void launch_motion<double>(hipStream_t, const GridParams &, const BubbleParams &,
                                    const BubbleArrays &, const double *, const double *,
                                    StepScalars *);
template void launch_motion<float>(hipStream_t, const GridParams &, const BubbleParams &,
                                   const BubbleArrays &, const float *, const float *,
                                   StepScalars *);
template void launch_bubble_step<double>(hipStream_t, const RPLaunch &, const RPLaunch *, int, int,
                                         int, int, int64_t *, StepScalars *, int, bool, int,
                                         int);
template void launch_bubble_step<float>(hipStream_t, const RPLaunch &, const RPLaunch *, int, int,
                                         int, int, int64_t *, StepScalars *, int, bool, int,
                                         int);

}  // namespace brp
