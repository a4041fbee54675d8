// api.hip -- C ABI of libbubbles_rp (include/bubbles_rp.h): set-up, the
// step scheduler and data movement.  Host code; kernels live in
// kernels_grid.hip and kernels_bubble.hip.
//
// The step is driven from the host: each pressure evaluation ends with one
// small device-to-host read of the residual (and, after the first
// evaluation, the bubble bounding box), which decides the Picard sub-loop
// exactly as the reference's host loop does (bubble_CUDA.cu:240-276).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "brp_internal.h"

using namespace brp;

// Staging array of brp_get_bubbles_noalloc, for callers that cannot allocate
// device memory: one bubble array of up to kScratchBubbles doubles (32 MiB),
// resident on every GPU the library runs on until the process ends.  Host
// code never passes `scratch` itself (in host code the name is a host-side
// stand-in, not a device address): brp_create looks up each device's address
// once, with that device current, and every use goes through it.
constexpr int kScratchBubbles = 1 << 22;
static __device__ __align__(256) double scratch[kScratchBubbles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};  // device address of `scratch`, per device

// The address of `scratch` on the current device `dev` (looked up once per
// device), or nullptr with the HIP error in *e.
double *scratch_address(int dev, hipError_t *e) {
    *e = hipSuccess;
    if (dev < 0 || dev >= kMaxDevices) {
        *e = hipErrorInvalidDevice;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        *e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (*e != hipSuccess) return nullptr;
        g_scratch_dev[dev] = (double *)p;
    }
    return g_scratch_dev[dev];
}

// out[perm[k]] = the bubble value of sorted bubble k: a bubble array
// (src != nullptr), or ibm / ibn (which_index 0..3) from the positions.
__global__ void k_bubbles_to_caller(int n, const int *perm, const double *src, const int *isrc,
                                    const double *x, const double *y, double rdx, double rdy,
                                    int which_index, double *out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double v;
    if (src)
        v = src[k];
    else if (isrc)
        v = (double)isrc[k];
    else if (which_index & 1)
        v = (double)bin_index(y[k], rdy, which_index < 2 ? 1.0 : 0.5);
    else
        v = (double)bin_index(x[k], rdx, which_index < 2 ? 1.0 : 0.5);
    out[perm[k]] = v;
}

thread_local std::string g_create_error;

struct Snapshot {
    bool valid = false;
    std::vector<void *> bufs;   // device copies, same order as the handle's buffer list
    std::vector<size_t> sizes;
    // scalar state
    int pc, fcur, fn, fm, fz, rp_cur, pl_m, pl_n, pl_p, first_step;
    Window W;
    double ctrl_x, ctrl_y, ctrl_dist;  // source focus (moved by brp_set_focus)
    int64_t nstep;
    double t, tx;
};

// Device memory carved sequentially from large chunks (the adjoint's bubble
// tape holds thousands of small per-step arrays).
struct Arena {
    std::vector<void *> chunks;
    size_t chunk_bytes = size_t(1) << 29, used = 0;
    ~Arena() {
        for (void *c : chunks) (void)hipFree(c);
    }
    size_t bytes() const { return chunks.size() * chunk_bytes; }
    void *get(size_t bytes) {
        bytes = (std::max<size_t>(bytes, 1) + 255) & ~size_t(255);
        if (bytes > chunk_bytes) return nullptr;
        if (chunks.empty() || used + bytes > chunk_bytes) {
            void *c = nullptr;
            if (hipMalloc(&c, chunk_bytes) != hipSuccess) return nullptr;
            chunks.push_back(c);
            used = 0;
        }
        void *q = (char *)chunks.back() + used;
        used += bytes;
        return q;
    }
};

// Intermediates of one step, kept for its adjoint.  Grid part: the
// start-of-step pressure (whole frame) and, on the rectangle `rect` only,
// the start-of-step f_g, the velocities after the update, every pressure
// evaluation's p0 and every void-fraction iterate (k = 0 is the
// prediction).  rect is the Picard window (which holds every nonzero f_g
// cell and every bubble stencil) grown by kRecMargin cells, joined with the
// objective box: the adjoint reads f_g everywhere (zero outside the window)
// but p0 and the velocities only at bubble stencils and in the box.  Bubble
// part: the RP start state and pressure history, the positions before and
// after the motion, and each Picard iteration's gathered pressure, clamped
// radius and sub-step count.
constexpr int kRecMargin = 2;
struct StepRecord {
    bool want_grid = true, want_bub = true;
    Arena *arena = nullptr;  // bubble part from here when set
    int box[4] = {0, -1, 0, -1};  // objective box (inclusive cells) the rectangle covers
    // grid part
    void *p0_start = nullptr;                    // whole frame
    Window rect{0, -1, 0, -1};
    Window win{0, -1, 0, -1};                    // the step's Picard window
    size_t rcap = 0;                             // cells of every rectangle buffer
    std::vector<void *> rect_owned;
    int64_t *fgC = nullptr;                      // f_g at the start (the step's f_gn)
    void *vx_end = nullptr, *vy_end = nullptr;   // velocities after the update
    std::vector<void *> p0_it;       // [k] after evaluation k (k = 0..n_pic)
    std::vector<int64_t *> fg_it;    // [k] void-fraction iterate k
    // bubble part
    double *x_start = nullptr, *y_start = nullptr;
    double *rp[6] = {};                     // RP start state R, Rn, d1R, PG, dt, re
    double *PLm = nullptr, *PLn = nullptr;  // liquid-pressure history of the step
    double *x_end = nullptr, *y_end = nullptr;   // positions after the motion
    std::vector<double *> PLp_it;    // [k] gathered PL_p of iteration k >= 1
    std::vector<double *> Rt_it;     // [k] R_t of iteration k >= 1
    std::vector<int *> sub_it;       // [k] RP sub-steps of iteration k >= 1, per bubble
    std::vector<void *> owned;
    bool started = false;
    // scalars of the recorded step
    int n_pic = 0, first = 0;
    double t = 0.0;
};

// A recorded frozen schedule: each step's Picard count and, for every RP
// launch in order, each bubble's accepted step sizes (dts[off[i] .. off[i+1])
// for bubble i in the library's cell-sorted order).  Immutable once recorded;
// handles replaying it share it.
struct Schedule {
    struct Call {
        int64_t *off = nullptr;
        double *dts = nullptr;
    };
    int nb = 0;
    int device = 0;
    std::vector<int> npic;
    std::vector<Call> calls;
    std::vector<void *> bufs;
    unsigned int *mismatch = nullptr;  // device counter of recordings that took another count
    ~Schedule() {
        for (void *q : bufs) (void)hipFree(q);
    }
    void *alloc(size_t bytes) {
        void *q = nullptr;
        if (hipMalloc(&q, std::max<size_t>(bytes, 256)) != hipSuccess) return nullptr;
        bufs.push_back(q);
        return q;
    }
};

// Optional control of run_steps: a forced Picard count per step (the frozen
// schedule of a replay), a record of one step's intermediates, the Picard
// counts taken, and a bubble tape to replay instead of integrating RP (the
// void-fraction iterates are then scattered from the taped radii).
struct StepCtl {
    const int *force_npic = nullptr;
    StepRecord *rec = nullptr;
    std::vector<int> *npic_out = nullptr;
    const StepRecord *tape_in = nullptr;
};

}  // namespace

struct brp_handle {
    brp_config cfg{};
    std::string err;
    int prec = BRP_PREC_FP64;
    size_t fsz = 8;
    GridParams g{};
    BubbleParams bp{};
    int N = 0, M = 0, pitch = 0, rows = 0;
    size_t frame = 0;  // elements per frame
    double T_inf = 0, P_inf = 0, rho_inf = 0, cs_inf = 0, dt = 0, dt0_rp = 0;
    int device = 0;                   // HIP device of the handle (current at brp_create)
    int cus = 1, per_cu = 1;          // compute units; RP kernel workgroups per CU
    int last_sub_max = 0;             // longest RP sub-step chain so far in the last step
    double last_sub_per_launch = 0.0;  // RP sub-steps per launch in the last step (grid choice)
    double fp_x = 0, fp_y = 0, omega = 0, ctrl_x = 0, ctrl_y = 0, ctrl_dist = 0;
    std::vector<double> sig_mx, sig_my, sig_nx, sig_ny;

    // device grid state
    void *vx = nullptr, *vy = nullptr, *p0 = nullptr;
    void *pb[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [buffer][x|y]
    int pc = 0;                                                  // buffer holding p
    int64_t *fgb[4] = {nullptr, nullptr, nullptr, nullptr};
    int fcur = 0, fn = 1, fm = 2, fz = 3;
    void *cxa = nullptr, *cxb = nullptr, *cya = nullptr, *cyb = nullptr;
    void *ax = nullptr, *bx = nullptr, *ay = nullptr, *by = nullptr;
    double *xu_d = nullptr, *rxp_d = nullptr;
    double *el_amp_d = nullptr, *el_phase_d = nullptr;
    int *perm_d = nullptr;           // perm on the device (sorted -> caller index)
    double *scratch_d = nullptr;     // this device's `scratch` (nullptr: lookup failed)
    std::string scratch_err;
    double *T_d = nullptr, *dT_d = nullptr, *heat_d = nullptr;  // temperature solve
    int n_elem = 0;

    // bubbles
    BubbleArrays b{};
    int nb = 0;
    int rp_cur = 0, pl_m = 0, pl_n = 1, pl_p = 2;
    std::vector<int> perm;  // perm[k] = caller index of sorted bubble k
    std::vector<double> pending_x, pending_y, pending_R0, pending_nB;  // host-side set-up

    StepScalars *sc_d = nullptr;
    StepScalars *sc_h = nullptr;
    RPLaunch *rpl_h = nullptr, *rpl_d = nullptr;  // the RP kernel's fixed parameters
    Window W{1, 0, 1, 0};
    int first_step = 1;
    int64_t nstep = 0;
    double t = 0.0, tx = 0.0;
    bool prepared = false;

    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> ev;
    double phase_ms[5] = {0, 0, 0, 0, 0};

    std::vector<void *> allocs;
    std::vector<size_t> alloc_sizes;
    size_t dev_bytes = 0;
    Snapshot snaps[16];

    // frozen schedule (brp_schedule_mode): 1 record, 2 replay
    int sched_mode = 0;
    std::shared_ptr<Schedule> sched;
    size_t sched_step = 0, sched_call = 0;
};

namespace {

int set_err(brp_handle *h, int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (h)
        h->err = buf;
    else
        g_create_error = buf;
    return code;
}

#define HIPCHK(h, call)                                                                  \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return set_err(h, BRP_ERR_DEVICE, "%s: %s (%s:%d)", #call, hipGetErrorString(e_), \
                           __FILE__, __LINE__);                                          \
    } while (0)

int dalloc(brp_handle *h, void **p, size_t bytes) {
    *p = nullptr;
    if (bytes == 0) bytes = 256;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess)
        return set_err(h, BRP_ERR_DEVICE, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    e = hipMemsetAsync(*p, 0, bytes, h->stream);
    if (e != hipSuccess)
        return set_err(h, BRP_ERR_DEVICE, "hipMemset: %s", hipGetErrorString(e));
    h->allocs.push_back(*p);
    h->alloc_sizes.push_back(bytes);
    h->dev_bytes += bytes;
    return BRP_OK;
}

// water properties used by the grid (bubble_CUDA_kernel.cuh:153-169)
double density_water() { return 1000.0; }
double sound_speed_water() { return 1500.0; }

// init_sigma (bubble_CUDA.cu:946-1205) for order 2: cell profiles mx, my
// indexed by cell i (0..N+1) and node profiles nx, ny indexed by node + 1.
void init_sigma(brp_handle *h) {
    const brp_config &c = h->cfg;
    const int N = c.grid_x, M = c.grid_y;
    h->sig_mx.assign(N + 2, 0.0);
    h->sig_my.assign(M + 2, 0.0);
    h->sig_nx.assign(N + 3, 0.0);
    h->sig_ny.assign(M + 3, 0.0);
    const int npml = c.pml_grid_number;
    const double order = c.pml_order;
    auto smax = [&](double d) {
        return -std::log(c.pml_sigma) * 0.5 * ((double)order + 1.0) / (d * (double)npml);
    };
    auto cell = [&](double s, int n) { return s * std::pow(((double)n - 0.5) / ((double)npml), order); };
    auto node = [&](double s, int n) { return s * std::pow(((double)n) / ((double)npml), order); };
    // ms = 0, me = 1, ns = -1, ne = 0 for order 2
    if (c.pml_x0)
        for (int i = 1; i <= std::min(N, npml); i++) h->sig_mx[i] = cell(smax(h->g.dx), npml - i + 1);
    if (c.pml_x1 && N - npml + 1 <= N)
        for (int i = std::max(1, N - npml + 1); i <= N; i++) h->sig_mx[i] = cell(smax(h->g.dx), i - N + npml);
    if (c.pml_y0)
        for (int j = 1; j <= std::min(M, npml); j++) h->sig_my[j] = cell(smax(h->g.dy), npml - j + 1);
    if (c.pml_y1 && M - npml + 1 <= M)
        for (int j = std::max(1, M - npml + 1); j <= M; j++) h->sig_my[j] = cell(smax(h->g.dy), j - M + npml);
    if (c.pml_x0)
        for (int i = 0; i <= std::min(N, npml - 1); i++) h->sig_nx[i + 1] = node(smax(h->g.dx), npml - i);
    if (c.pml_x1 && N - npml + 1 <= N)
        for (int i = std::max(0, N - npml + 1); i <= std::min(N, N + 2); i++)
            h->sig_nx[i + 1] = node(smax(h->g.dx), i - N + npml);
    if (c.pml_y0)
        for (int j = 0; j <= std::min(M, npml - 1); j++) h->sig_ny[j + 1] = node(smax(h->g.dy), npml - j);
    if (c.pml_y1 && M - npml + 1 <= M)
        for (int j = std::max(0, M - npml + 1); j <= std::min(M, M + 2); j++)
            h->sig_ny[j + 1] = node(smax(h->g.dy), j - M + npml);
}

template <typename F>
int upload_coef(brp_handle *h, void **dst, const std::vector<double> &v) {
    std::vector<F> tmp(v.size());
    for (size_t k = 0; k < v.size(); k++) tmp[k] = (F)v[k];
    int rc = dalloc(h, dst, sizeof(F) * v.size());
    if (rc) return rc;
    HIPCHK(h, hipMemcpy(*dst, tmp.data(), sizeof(F) * v.size(), hipMemcpyHostToDevice));
    return BRP_OK;
}

int upload_d(brp_handle *h, double **dst, const std::vector<double> &v) {
    int rc = dalloc(h, (void **)dst, sizeof(double) * std::max<size_t>(v.size(), 1));
    if (rc) return rc;
    if (!v.empty())
        HIPCHK(h, hipMemcpy(*dst, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
    return BRP_OK;
}

// Coefficient arrays of the velocity and pressure updates, computed in FP64
// with the reference's operation order so the FP64 mode reproduces it:
//   velocity s3 = ((c + c) * 0.5) * 0.5 * dt * sigma_n  (kernel.cuh:662-664)
//   pressure s2 = (0.5 * dt * c) * sigma_m              (kernel.cuh:853-856)
template <typename F>
int init_coefficients(brp_handle *h) {
    const double c = h->cs_inf, dt = h->dt;
    auto mk = [&](const std::vector<double> &sig, bool vel, std::vector<double> &a,
                  std::vector<double> &b) {
        a.resize(sig.size());
        b.resize(sig.size());
        for (size_t k = 0; k < sig.size(); k++) {
            double s;
            if (vel) {
                s = (c + c) * 0.5;
                s = s * 0.5 * dt * sig[k];
            } else {
                s = (0.5 * dt * c) * sig[k];
            }
            a[k] = 1.0 - s;
            b[k] = 1.0 + s;
        }
    };
    std::vector<double> a, b;
    int rc;
    mk(h->sig_nx, true, a, b);
    if ((rc = upload_coef<F>(h, &h->cxa, a)) || (rc = upload_coef<F>(h, &h->cxb, b))) return rc;
    mk(h->sig_ny, true, a, b);
    if ((rc = upload_coef<F>(h, &h->cya, a)) || (rc = upload_coef<F>(h, &h->cyb, b))) return rc;
    // cell profiles are stored at index i + 1 like every 1-D array
    std::vector<double> mx(h->N + 3, 0.0), my(h->M + 3, 0.0);
    for (int i = 0; i <= h->N + 1; i++) mx[i + 1] = h->sig_mx[i];
    for (int j = 0; j <= h->M + 1; j++) my[j + 1] = h->sig_my[j];
    mk(mx, false, a, b);
    if ((rc = upload_coef<F>(h, &h->ax, a)) || (rc = upload_coef<F>(h, &h->bx, b))) return rc;
    mk(my, false, a, b);
    if ((rc = upload_coef<F>(h, &h->ay, a)) || (rc = upload_coef<F>(h, &h->by, b))) return rc;
    // init_grid_vector (bubble_CUDA.cu:690-713): xu at nodes, rxp at cells, index + 1
    std::vector<double> xu(h->N + 3, 0.0), rxp(h->N + 3, 0.0);
    for (int i = -1; i <= h->N + 1; i++) xu[i + 1] = ((double)i) * h->g.dx;
    for (int i = 0; i <= h->N + 1; i++) rxp[i + 1] = 1.0 / (h->g.dx * ((double)i - 0.5));
    if ((rc = upload_d(h, &h->xu_d, xu)) || (rc = upload_d(h, &h->rxp_d, rxp))) return rc;
    return BRP_OK;
}

// Aperture of the focused source: row-0 cells with |fp - (xm, -dy/2)| <= f_dist.
void find_aperture(brp_handle *h) {
    h->g.ap_lo = 1;
    h->g.ap_hi = 0;
    const double ym = -0.5 * h->g.dy;
    for (int i = 1; i <= h->N; i++) {
        double xm = (i - 0.5) * h->g.dx;
        double s1 = std::sqrt((h->ctrl_x - xm) * (h->ctrl_x - xm) + (h->ctrl_y - ym) * (h->ctrl_y - ym));
        if (s1 <= h->ctrl_dist) {
            if (h->g.ap_hi < h->g.ap_lo) h->g.ap_lo = i;
            h->g.ap_hi = i;
        }
    }
}

inline int host_bin(double x, double r, double c) {
    double v = std::floor(x * r + c);
    return (v >= -1.0e9 && v <= 1.0e9) ? (int)v : -1000000000;
}

int alloc_bubbles(brp_handle *h, int n) {
    h->nb = n;
    h->b.n = n;
    const size_t bytes = sizeof(double) * (size_t)std::max(n, 1);
    int rc;
    double **arrs[] = {&h->b.x,      &h->b.y,      &h->b.R[0],   &h->b.R[1],   &h->b.Rn[0],
                       &h->b.Rn[1],  &h->b.d1R[0], &h->b.d1R[1], &h->b.PG[0],  &h->b.PG[1],
                       &h->b.bdt[0], &h->b.bdt[1], &h->b.re[0],  &h->b.re[1],  &h->b.R_t,
                       &h->b.Q_B,    &h->b.PL[0],  &h->b.PL[1],  &h->b.PL[2],  &h->b.n_B,
                       &h->b.R0,     &h->b.PG0,    &h->b.calpha, &h->b.vx,     &h->b.vy};
    for (double **a : arrs)
        if ((rc = dalloc(h, (void **)a, bytes))) return rc;
    if ((rc = dalloc(h, (void **)&h->b.sub, sizeof(int) * (size_t)std::max(n, 1)))) return rc;
    if ((rc = dalloc(h, (void **)&h->b.order, sizeof(int) * (size_t)std::max(n, 1)))) return rc;
    if ((rc = dalloc(h, (void **)&h->b.key, sizeof(int) * (size_t)std::max(n, 1)))) return rc;
    if ((rc = dalloc(h, (void **)&h->b.order_tmp, sizeof(int) * (size_t)std::max(n, 1))) ||
        (rc = dalloc(h, (void **)&h->b.sort_hist, sizeof(int) * 2048)))
        return rc;
    if ((rc = dalloc(h, (void **)&h->b.part, sizeof(int) * 2 * (size_t)((n + 255) / 256 + 1))))
        return rc;
    if ((rc = dalloc(h, (void **)&h->b.bbox_blk, sizeof(int) * 4 * (size_t)((n + 255) / 256 + 1))))
        return rc;
    return BRP_OK;
}

// bubble_input (bubble_CUDA.cu:900-943) for every bubble, in cell-sorted
// order, with per-bubble R0 (extension E3).  Bubbles must not be re-placed
// after brp_prepare.
int place_bubbles(brp_handle *h, const std::vector<double> &x, const std::vector<double> &y,
                  const std::vector<double> &R0v, const std::vector<double> *nBv, double fg_in) {
    const int n = (int)x.size();
    if (h->b.x) return set_err(h, BRP_ERR_STATE, "bubbles already placed");
    int rc = alloc_bubbles(h, n);
    if (rc) return rc;
    // sort by (ibn.y, ibn.x) so scatter stencils of neighbouring lanes coincide
    h->perm.resize(n);
    std::iota(h->perm.begin(), h->perm.end(), 0);
    std::vector<int64_t> key(n);
    for (int k = 0; k < n; k++) {
        int64_t kx = host_bin(x[k], h->g.rdx, 0.5), ky = host_bin(y[k], h->g.rdy, 0.5);
        key[k] = (ky + 2000000000LL) * 4000000000LL + (kx + 2000000000LL);
    }
    std::stable_sort(h->perm.begin(), h->perm.end(), [&](int a, int b) { return key[a] < key[b]; });
    const double Pi = std::acos(-1.0);
    const brp_config &c = h->cfg;
    std::vector<double> hx(n), hy(n), R0(n), PG0(n), ca(n), nB(n), dt0(n), zero(n, 0.0);
    for (int k = 0; k < n; k++) {
        const int q = h->perm[k];
        hx[k] = x[q];
        hy[k] = y[q];
        R0[k] = R0v[q];
        const double R03 = R0[k] * R0[k] * R0[k];
        PG0[k] = c.bub_P0 + 2.0 * c.bub_sigma / R0[k];
        ca[k] = c.bub_gamma * PG0[k] * R03 / (2.0 * (c.bub_gamma - 1.0) * c.bub_T0 * c.bub_K0);
        if (nBv)
            nB[k] = (*nBv)[q];
        else if (c.tr_cylindrical)
            nB[k] = fg_in * (hx[k] * h->g.dx * h->g.dy) / (4.0 / 3.0 * Pi * std::pow(R0[k], 3));
        else
            nB[k] = fg_in * (h->g.dx * h->g.dy) / (4.0 / 3.0 * Pi * std::pow(R0[k], 3));
        dt0[k] = h->dt0_rp;
    }
    auto up = [&](double *d, const std::vector<double> &v) -> int {
        if (n == 0) return BRP_OK;
        HIPCHK(h, hipMemcpy(d, v.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        return BRP_OK;
    };
    if ((rc = dalloc(h, (void **)&h->perm_d, sizeof(int) * (size_t)std::max(n, 1)))) return rc;
    if (n > 0)
        HIPCHK(h, hipMemcpy(h->perm_d, h->perm.data(), sizeof(int) * n, hipMemcpyHostToDevice));
    if ((rc = up(h->b.x, hx)) || (rc = up(h->b.y, hy)) || (rc = up(h->b.R0, R0)) ||
        (rc = up(h->b.PG0, PG0)) || (rc = up(h->b.calpha, ca)) || (rc = up(h->b.n_B, nB)) ||
        (rc = up(h->b.R_t, R0)) || (rc = up(h->b.Q_B, zero)) || (rc = up(h->b.vx, zero)) ||
        (rc = up(h->b.vy, zero)))
        return rc;
    for (int s = 0; s < 2; s++) {
        if ((rc = up(h->b.R[s], R0)) || (rc = up(h->b.Rn[s], R0)) || (rc = up(h->b.d1R[s], zero)) ||
            (rc = up(h->b.PG[s], PG0)) || (rc = up(h->b.bdt[s], dt0)) || (rc = up(h->b.re[s], zero)))
            return rc;
    }
    for (int s = 0; s < 3; s++)
        if ((rc = up(h->b.PL[s], zero))) return rc;
    h->rp_cur = 0;
    h->pl_m = 0;
    h->pl_n = 1;
    h->pl_p = 2;
    return BRP_OK;
}

void window_union(Window &w, int i0, int i1, int j0, int j1) {
    if (i0 > i1 || j0 > j1) return;
    if (w.i0 > w.i1 || w.j0 > w.j1) {
        w = Window{i0, i1, j0, j1};
        return;
    }
    w.i0 = std::min(w.i0, i0);
    w.i1 = std::max(w.i1, i1);
    w.j0 = std::min(w.j0, j0);
    w.j1 = std::max(w.j1, j1);
}

template <typename F>
int read_frame(brp_handle *h, const void *src, std::vector<double> &out) {
    std::vector<F> tmp(h->frame);
    HIPCHK(h, hipMemcpyAsync(tmp.data(), src, sizeof(F) * h->frame, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    out.resize(h->frame);
    for (size_t k = 0; k < h->frame; k++) out[k] = (double)tmp[k];
    return BRP_OK;
}

int read_frame_fg(brp_handle *h, const int64_t *src, std::vector<double> &out) {
    std::vector<int64_t> tmp(h->frame);
    HIPCHK(h, hipMemcpyAsync(tmp.data(), src, sizeof(int64_t) * h->frame, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    out.resize(h->frame);
    for (size_t k = 0; k < h->frame; k++) out[k] = (double)tmp[k] * kFgInvScale;
    return BRP_OK;
}

void field_extent(const brp_handle *h, int field, int *nx, int *ny, int *i0, int *j0) {
    *i0 = 0;
    *j0 = 0;
    *nx = h->N + 2;
    *ny = h->M + 2;
    if (field == BRP_FIELD_VX) {
        *i0 = -1;
        *nx = h->N + 3;
    } else if (field == BRP_FIELD_VY) {
        *j0 = -1;
        *ny = h->M + 3;
    }
}


// ---- state snapshots (save/restore slots, adjoint checkpoints and tape) ----
int snap_save(brp_handle *h, Snapshot &sn) {
    if (!sn.valid && sn.bufs.empty()) {
        for (size_t k = 0; k < h->allocs.size(); k++) {
            void *p = nullptr;
            if (hipMalloc(&p, h->alloc_sizes[k]) != hipSuccess) {
                for (void *q : sn.bufs) (void)hipFree(q);
                sn.bufs.clear();
                sn.sizes.clear();
                return set_err(h, BRP_ERR_DEVICE, "snapshot allocation failed (%zu bytes)",
                               h->dev_bytes);
            }
            sn.bufs.push_back(p);
            sn.sizes.push_back(h->alloc_sizes[k]);
        }
    }
    if (sn.bufs.size() != h->allocs.size()) return set_err(h, BRP_ERR_STATE, "snapshot layout changed");
    for (size_t k = 0; k < h->allocs.size(); k++)
        HIPCHK(h, hipMemcpyAsync(sn.bufs[k], h->allocs[k], sn.sizes[k], hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    sn.pc = h->pc;
    sn.fcur = h->fcur;
    sn.fn = h->fn;
    sn.fm = h->fm;
    sn.fz = h->fz;
    sn.rp_cur = h->rp_cur;
    sn.pl_m = h->pl_m;
    sn.pl_n = h->pl_n;
    sn.pl_p = h->pl_p;
    sn.first_step = h->first_step;
    sn.W = h->W;
    sn.ctrl_x = h->ctrl_x;
    sn.ctrl_y = h->ctrl_y;
    sn.ctrl_dist = h->ctrl_dist;
    sn.nstep = h->nstep;
    sn.t = h->t;
    sn.tx = h->tx;
    sn.valid = true;
    return BRP_OK;
}

int snap_restore(brp_handle *h, const Snapshot &sn) {
    if (sn.bufs.size() != h->allocs.size()) return set_err(h, BRP_ERR_STATE, "snapshot layout changed");
    for (size_t k = 0; k < h->allocs.size(); k++)
        HIPCHK(h, hipMemcpyAsync(h->allocs[k], sn.bufs[k], sn.sizes[k], hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->pc = sn.pc;
    h->fcur = sn.fcur;
    h->fn = sn.fn;
    h->fm = sn.fm;
    h->fz = sn.fz;
    h->rp_cur = sn.rp_cur;
    h->pl_m = sn.pl_m;
    h->pl_n = sn.pl_n;
    h->pl_p = sn.pl_p;
    h->first_step = sn.first_step;
    h->W = sn.W;
    if (h->ctrl_x != sn.ctrl_x || h->ctrl_y != sn.ctrl_y || h->ctrl_dist != sn.ctrl_dist) {
        h->ctrl_x = h->g.fp_x = sn.ctrl_x;
        h->ctrl_y = h->g.fp_y = sn.ctrl_y;
        h->ctrl_dist = h->g.fdist = sn.ctrl_dist;
        find_aperture(h);
    }
    h->nstep = sn.nstep;
    h->t = sn.t;
    h->tx = sn.tx;
    return BRP_OK;
}

void snap_free(Snapshot &sn) {
    for (void *q : sn.bufs) (void)hipFree(q);
    sn.bufs.clear();
    sn.sizes.clear();
    sn.valid = false;
}

size_t snap_bytes(const brp_handle *h) {
    size_t b = 0;
    for (size_t z : h->alloc_sizes) b += z;
    return b;
}

// Device memory owned by one adjoint call (freed at its end).
struct AdjMem {
    std::vector<void *> p;
    size_t bytes = 0;
    ~AdjMem() {
        for (void *q : p) (void)hipFree(q);
    }
    template <typename T>
    int get(brp_handle *h, T **out, size_t n) {
        void *q = nullptr;
        const size_t z = std::max<size_t>(n * sizeof(T), 256);
        if (hipMalloc(&q, z) != hipSuccess)
            return set_err(h, BRP_ERR_DEVICE, "adjoint allocation of %zu bytes failed", z);
        if (hipMemsetAsync(q, 0, z, h->stream) != hipSuccess)
            return set_err(h, BRP_ERR_DEVICE, "hipMemsetAsync failed");
        p.push_back(q);
        bytes += z;
        *out = (T *)q;
        return BRP_OK;
    }
};

// Cells of a record rectangle.
size_t rect_cells(const Window &w) {
    return w.i0 > w.i1 || w.j0 > w.j1 ? 0 : (size_t)(w.i1 - w.i0 + 1) * (size_t)(w.j1 - w.j0 + 1);
}

// Copies rectangle w of a frame array (elements of `elem` bytes) into a
// dense buffer.
int copy_rect(brp_handle *h, void *dst, const void *src, size_t elem, const Window &w) {
    const size_t cols = (size_t)(w.i1 - w.i0 + 1), rows = (size_t)(w.j1 - w.j0 + 1);
    if (!rect_cells(w)) return BRP_OK;
    const char *s0 = (const char *)src + frame_off(w.i0, w.j0, h->pitch) * elem;
    HIPCHK(h, hipMemcpy2DAsync(dst, cols * elem, s0, (size_t)h->pitch * elem, cols * elem, rows,
                               hipMemcpyDeviceToDevice, h->stream));
    return BRP_OK;
}

int rec_alloc(brp_handle *h, StepRecord *r, size_t bytes, bool bub, void **out) {
    void *q = nullptr;
    if (bub && r->arena) {
        q = r->arena->get(bytes);
        if (!q) return set_err(h, BRP_ERR_DEVICE, "bubble tape allocation failed");
    } else {
        if (hipMalloc(&q, std::max<size_t>(bytes, 256)) != hipSuccess)
            return set_err(h, BRP_ERR_DEVICE, "step record allocation failed");
        r->owned.push_back(q);
    }
    *out = q;
    return BRP_OK;
}

int rect_alloc(brp_handle *h, StepRecord *r, size_t elem, void **out) {
    void *q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(r->rcap * elem, 256)) != hipSuccess)
        return set_err(h, BRP_ERR_DEVICE, "step record allocation failed");
    r->rect_owned.push_back(q);
    *out = q;
    return BRP_OK;
}

// Sets the record's rectangle from the current Picard window and the box,
// (re)allocating the rectangle buffers when they are too small; called once
// per step, after the window update.
int rec_set_rect(brp_handle *h, StepRecord *r) {
    Window w{0, -1, 0, -1};
    if (h->W.i0 <= h->W.i1 && h->W.j0 <= h->W.j1)
        w = Window{std::max(h->W.i0 - kRecMargin, -1), std::min(h->W.i1 + kRecMargin, h->N + 1),
                   std::max(h->W.j0 - kRecMargin, -1), std::min(h->W.j1 + kRecMargin, h->M + 1)};
    window_union(w, r->box[0], r->box[1], r->box[2], r->box[3]);
    r->rect = w;
    r->win = h->W;
    const size_t need = std::max<size_t>(rect_cells(w), 1);
    if (need <= r->rcap && r->fgC) return BRP_OK;
    for (void *q : r->rect_owned) (void)hipFree(q);
    r->rect_owned.clear();
    r->rcap = std::max(need, r->rcap);
    int rc;
    void *q[3];
    const size_t el[3] = {sizeof(int64_t), h->fsz, h->fsz};
    for (int j = 0; j < 3; j++)
        if ((rc = rect_alloc(h, r, el[j], &q[j]))) return rc;
    r->fgC = (int64_t *)q[0];
    r->vx_end = q[1];
    r->vy_end = q[2];
    for (size_t k = 0; k < r->p0_it.size(); k++) {
        void *a, *b;
        if ((rc = rect_alloc(h, r, h->fsz, &a)) || (rc = rect_alloc(h, r, sizeof(int64_t), &b)))
            return rc;
        r->p0_it[k] = a;
        r->fg_it[k] = (int64_t *)b;
    }
    return BRP_OK;
}

int rec_grow(brp_handle *h, StepRecord *r, int k) {
    // k evaluations recorded so far need p0_it/fg_it[0..k-1] and PLp_it/Rt_it[1..k-1]
    const size_t fb = h->fsz * h->frame;
    const size_t nbb = sizeof(double) * std::max(h->nb, 1);
    int rc;
    if (!r->started) {
        if (r->want_grid && (rc = rec_alloc(h, r, fb, false, &r->p0_start))) return rc;
        if (r->want_bub) {
            void *q[12];
            for (int j = 0; j < 12; j++)
                if ((rc = rec_alloc(h, r, nbb, true, &q[j]))) return rc;
            r->x_start = (double *)q[0];
            r->y_start = (double *)q[1];
            for (int j = 0; j < 6; j++) r->rp[j] = (double *)q[2 + j];
            r->PLm = (double *)q[8];
            r->PLn = (double *)q[9];
            r->x_end = (double *)q[10];
            r->y_end = (double *)q[11];
        }
        r->started = true;
    }
    while ((int)std::max(r->p0_it.size(), r->PLp_it.size()) < k) {
        void *a = nullptr, *b = nullptr, *c = nullptr, *d = nullptr, *e = nullptr;
        const bool first = std::max(r->p0_it.size(), r->PLp_it.size()) == 0;
        if (r->want_grid && r->rcap &&
            ((rc = rect_alloc(h, r, h->fsz, &a)) || (rc = rect_alloc(h, r, sizeof(int64_t), &b))))
            return rc;
        if (r->want_bub && !first &&
            ((rc = rec_alloc(h, r, nbb, true, &c)) || (rc = rec_alloc(h, r, nbb, true, &d)) ||
             (rc = rec_alloc(h, r, sizeof(int) * std::max(h->nb, 1), true, &e))))
            return rc;
        r->p0_it.push_back(a);
        r->fg_it.push_back((int64_t *)b);
        r->PLp_it.push_back((double *)c);
        r->Rt_it.push_back((double *)d);
        r->sub_it.push_back((int *)e);
    }
    return BRP_OK;
}

// Device bytes a record holds.
size_t rec_bytes(const brp_handle *h, const StepRecord &r) {
    size_t b = 0;
    if (!r.started) return 0;
    const size_t nbb = sizeof(double) * std::max(h->nb, 1);
    if (r.want_grid)
        b += h->fsz * h->frame + r.rcap * (sizeof(int64_t) + 2 * h->fsz) +
             r.p0_it.size() * r.rcap * (h->fsz + sizeof(int64_t));
    if (r.want_bub && !r.arena && !r.PLp_it.empty())
        b += 12 * nbb + (r.PLp_it.size() - 1) * (2 * nbb + nbb / 2);
    return b;
}

void rec_free(StepRecord &r) {
    for (void *q : r.owned) (void)hipFree(q);
    for (void *q : r.rect_owned) (void)hipFree(q);
    const bool g = r.want_grid, b = r.want_bub;
    Arena *ar = r.arena;
    int bx[4];
    std::copy(r.box, r.box + 4, bx);
    r = StepRecord{};
    r.want_grid = g;
    r.want_bub = b;
    r.arena = ar;
    std::copy(bx, bx + 4, r.box);
}

// ---- the coupled step ----------------------------------------------------
int rec_grow(brp_handle *h, StepRecord *r, int k);

// Frozen-schedule recording of the RP launch just queued: its per-bubble
// sub-step counts give the offsets, then k_rp_record recomputes the same
// integrations and stores every accepted step size.
int schedule_record_call(brp_handle *h) {
    Schedule &S = *h->sched;
    const int n = h->nb;
    std::vector<int> sub(n);
    HIPCHK(h, hipMemcpyAsync(sub.data(), h->b.sub, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::vector<int64_t> off(n + 1, 0);
    for (int k = 0; k < n; k++) off[k + 1] = off[k] + sub[k];
    Schedule::Call cl;
    cl.off = (int64_t *)S.alloc(sizeof(int64_t) * (n + 1));
    cl.dts = (double *)S.alloc(sizeof(double) * std::max<int64_t>(off[n], 1));
    if (!cl.off || !cl.dts)
        return set_err(h, BRP_ERR_DEVICE, "frozen schedule: allocation of %lld step sizes failed",
                       (long long)off[n]);
    HIPCHK(h, hipMemcpyAsync(cl.off, off.data(), sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice,
                             h->stream));
    launch_rp_record(h->stream, h->rpl_d, n, h->rp_cur, h->pl_m, h->pl_n, h->pl_p, cl.off, cl.dts,
                     S.mismatch);
    HIPCHK(h, hipGetLastError());
    S.calls.push_back(cl);
    return BRP_OK;
}

template <typename F>
int run_steps(brp_handle *h, int k, brp_step_stats *stats, StepCtl *ctl = nullptr) {
    const double thr = (double)1.0e-7f;  // bubble_CUDA.cu:243
    const int max_pic = std::min(h->cfg.max_picard > 0 ? h->cfg.max_picard : 1000, kMaxEval - 1);
    const bool bub = h->cfg.bub_enable && h->nb > 0;
    hipStream_t s = h->stream;
    const size_t hdr = offsetof(StepScalars, resid);
    if (bub) {  // the RP kernel's parameters as of this call (no step is in flight)
        *h->rpl_h = RPLaunch{h->g, h->bp, h->b};
        HIPCHK(h, hipMemcpyAsync(h->rpl_d, h->rpl_h, sizeof(RPLaunch), hipMemcpyHostToDevice, s));
    }
    F *vx = (F *)h->vx, *vy = (F *)h->vy, *p0 = (F *)h->p0;
    const F *cxa = (const F *)h->cxa, *cxb = (const F *)h->cxb, *cya = (const F *)h->cya,
            *cyb = (const F *)h->cyb, *ax = (const F *)h->ax, *bx = (const F *)h->bx,
            *ay = (const F *)h->ay, *by = (const F *)h->by;
    for (int step = 0; step < k; step++) {
        brp_step_stats st{};
        // clock: compensated sum (bubble_CUDA.cu:208-212)
        h->nstep++;
        const double s1 = h->t;
        h->t = h->t + h->dt + h->tx;
        const double s2 = h->t - s1;
        h->tx = h->dt + h->tx - s2;
        // store_variables (kernel.cuh:1229-1262) as renamings: pn <- p,
        // (f_gm, f_gn, f_g) <- (f_gn, f_g, prediction), PL history shift.
        const int pn_i = h->pc, p_i = h->pc ^ 1;
        const int old_cur = h->fcur, old_n = h->fn, old_m = h->fm;
        {
            const int m = h->pl_m, n = h->pl_n, p = h->pl_p;
            h->pl_m = n;
            h->pl_n = p;
            h->pl_p = m;
        }
        launch_zero_scalars(s, h->sc_d, max_pic + 1);
        StepRecord *rec = ctl ? ctl->rec : nullptr;
        int forced = (ctl && ctl->force_npic) ? ctl->force_npic[step] : -1;
        if (h->sched_mode == 2) {
            if (h->sched_step >= h->sched->npic.size())
                return set_err(h, BRP_ERR_STATE, "step %lld: the frozen schedule is exhausted",
                               (long long)h->nstep);
            forced = h->sched->npic[h->sched_step++];
        }
        if (rec) {
            int rc = rec_grow(h, rec, 1);
            if (rc) return rc;
            if (rec->want_grid)
                HIPCHK(h, hipMemcpyAsync(rec->p0_start, h->p0, sizeof(F) * h->frame, hipMemcpyDeviceToDevice, s));
            if (rec->want_bub && h->nb > 0) {
                const size_t nbb = sizeof(double) * h->nb;
                const int c = h->rp_cur;
                double *src[10] = {h->b.x, h->b.y, h->b.R[c], h->b.Rn[c], h->b.d1R[c], h->b.PG[c],
                                   h->b.bdt[c], h->b.re[c], h->b.PL[h->pl_m], h->b.PL[h->pl_n]};
                double *dst[10] = {rec->x_start, rec->y_start, rec->rp[0], rec->rp[1], rec->rp[2],
                                   rec->rp[3], rec->rp[4], rec->rp[5], rec->PLm, rec->PLn};
                for (int q = 0; q < 10; q++)
                    HIPCHK(h, hipMemcpyAsync(dst[q], src[q], nbb, hipMemcpyDeviceToDevice, s));
            }
            rec->first = h->first_step;
            rec->t = h->t;
        }
        HIPCHK(h, hipEventRecord(h->ev[0], s));
        launch_velocity<F>(s, h->g, vx, vy, p0, h->fgb[old_cur], h->W, h->first_step, cxa, cxb,
                           cya, cyb);
        launch_velocity_ghosts<F>(s, h->g, vx, vy, h->t);
        HIPCHK(h, hipEventRecord(h->ev[1], s));
        if (bub) launch_motion<F>(s, h->g, h->bp, h->b, vx, vy, h->sc_d);
        HIPCHK(h, hipEventRecord(h->ev[2], s));
        launch_pressure_first<F>(s, h->g, vx, vy, (F *)h->pb[pn_i][0], (F *)h->pb[pn_i][1],
                                 (F *)h->pb[p_i][0], (F *)h->pb[p_i][1], p0, h->fgb[old_cur],
                                 h->fgb[old_n], h->fgb[old_m], h->W, ax, bx, ay, by, h->xu_d,
                                 h->rxp_d, h->sc_d);
        launch_pressure_ghosts<F>(s, h->g, p0, h->el_amp_d, h->el_phase_d, h->t);
        HIPCHK(h, hipEventRecord(h->ev[3], s));
        HIPCHK(h, hipMemcpyAsync(h->sc_h, h->sc_d, hdr + sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        HIPCHK(h, hipStreamSynchronize(s));
        double r0;
        std::memcpy(&r0, &h->sc_h->resid[0], 8);
        double resimax = r0 / h->P_inf;
        st.resimax0 = resimax;
        if (bub) {
            window_union(h->W, h->sc_h->bbox[0], h->sc_h->bbox[1], h->sc_h->bbox[2], h->sc_h->bbox[3]);
            if (h->cfg.tr_cylindrical && h->W.i0 <= 1 && h->W.i0 <= h->W.i1) h->W.i0 = 0;  // axis fold target
        }
        if (rec && rec->want_grid) {  // the window is final for this step
            int rc = rec_set_rect(h, rec);
            if (rc) return rc;
            if ((rc = copy_rect(h, rec->fgC, h->fgb[old_cur], sizeof(int64_t), rec->rect)) ||
                (rc = copy_rect(h, rec->p0_it[0], p0, sizeof(F), rec->rect)) ||
                (rc = copy_rect(h, rec->fg_it[0], h->fgb[old_m], sizeof(int64_t), rec->rect)))
                return rc;
        }
        int status = BRP_OK;
        long long lost = 0;  // bubbles the RP kernel did not write back
        int loop = 0;
        int target = h->fz, prev = old_m;
        double bub_ms = 0.0, pic_p_ms = 0.0, other_ms = 0.0;
        if (bub) {
            while (forced >= 0 ? loop < forced : resimax > thr) {
                if (loop >= max_pic) {
                    status = BRP_ERR_PICARD_CAP;
                    break;
                }
                loop++;
                if (rec) {
                    int rc = rec_grow(h, rec, loop + 1);
                    if (rc) return rc;
                }
                HIPCHK(h, hipEventRecord(h->ev[7], s));
                const StepRecord *tin = ctl ? ctl->tape_in : nullptr;
                const bool replay = h->sched_mode == 2;
                if (!tin && !replay) {
                    // the pressure gather, then the queue: the first iteration's
                    // from the start state (the previous step's counts do not
                    // find this step's longest chains), the later ones from the
                    // previous iteration's counts, bubbles whose gathered
                    // pressure did not change last
                    launch_queue_keys<F>(s, h->g, h->bp, h->b, p0, h->rp_cur, h->pl_p, loop,
                                         h->bp.extreme_threshold);
                    launch_order_partition(s, h->b, h->b.key, h->bp.heavy_threshold,
                                           h->bp.extreme_threshold);
                }
                HIPCHK(h, hipEventRecord(h->ev[4], s));
                if (tin) {  // replay: the void fraction from the taped radii, no RP
                    launch_scatter_radii(s, h->g, h->bp, h->b, tin->Rt_it[loop], h->fgb[target], h->sc_d);
                } else if (replay) {  // the recorded step sizes of this RP call
                    if (h->sched_call >= h->sched->calls.size())
                        return set_err(h, BRP_ERR_STATE, "step %lld: the frozen schedule has no "
                                       "RP call %zu", (long long)h->nstep, h->sched_call);
                    const Schedule::Call &cl = h->sched->calls[h->sched_call++];
                    launch_rp_replay<F>(s, h->rpl_d, h->nb, h->rp_cur, h->pl_m, h->pl_n, h->pl_p,
                                        p0, h->fgb[target], h->sc_d, cl.off, cl.dts);
                } else {
                    const bool lean = bubble_step_lean(h->last_sub_max, h->last_sub_per_launch,
                                                       h->bp.extreme_threshold, h->cus);
                    st.rp_lean_launches += lean ? 1 : 0;
                    launch_bubble_step<F>(s, *h->rpl_h, h->rpl_d, h->rp_cur, h->pl_m, h->pl_n,
                                          h->pl_p, h->fgb[target], h->sc_d, loop, lean, h->cus,
                                          h->per_cu);
                    if (h->sched_mode == 1) {
                        int rc = schedule_record_call(h);
                        if (rc) return rc;
                    }
                }
                if (h->cfg.tr_cylindrical) launch_cyl_fold(s, h->g, h->fgb[target]);
                if (rec && rec->want_grid) {
                    int rc = copy_rect(h, rec->fg_it[loop], h->fgb[target], sizeof(int64_t), rec->rect);
                    if (rc) return rc;
                }
                if (rec && rec->want_bub) {
                    HIPCHK(h, hipMemcpyAsync(rec->PLp_it[loop], h->b.PL[h->pl_p], sizeof(double) * h->nb, hipMemcpyDeviceToDevice, s));
                    HIPCHK(h, hipMemcpyAsync(rec->Rt_it[loop], h->b.R_t, sizeof(double) * h->nb, hipMemcpyDeviceToDevice, s));
                    HIPCHK(h, hipMemcpyAsync(rec->sub_it[loop], h->b.sub, sizeof(int) * h->nb, hipMemcpyDeviceToDevice, s));
                }
                HIPCHK(h, hipEventRecord(h->ev[5], s));
                launch_pressure_picard<F>(s, h->g, vx, vy, (F *)h->pb[pn_i][0], (F *)h->pb[pn_i][1],
                                          (F *)h->pb[p_i][0], (F *)h->pb[p_i][1], p0,
                                          h->fgb[target], h->fgb[old_cur], h->fgb[prev], h->W, ax,
                                          bx, ay, by, h->xu_d, h->rxp_d, h->sc_d, loop);
                launch_pressure_ghosts<F>(s, h->g, p0, h->el_amp_d, h->el_phase_d, h->t);
                if (rec && rec->want_grid) {
                    int rc = copy_rect(h, rec->p0_it[loop], p0, sizeof(F), rec->rect);
                    if (rc) return rc;
                }
                HIPCHK(h, hipEventRecord(h->ev[6], s));
                HIPCHK(h, hipMemcpyAsync(h->sc_h, h->sc_d, hdr, hipMemcpyDeviceToHost, s));
                HIPCHK(h, hipMemcpyAsync(&h->sc_h->resid[loop], &h->sc_d->resid[loop], 8,
                                         hipMemcpyDeviceToHost, s));
                HIPCHK(h, hipStreamSynchronize(s));
                double r;
                std::memcpy(&r, &h->sc_h->resid[loop], 8);
                resimax = r / h->P_inf;
                h->last_sub_max = h->sc_h->sub_max;
                h->last_sub_per_launch = (double)h->sc_h->sub_sum / loop;
                // every bubble is written back exactly once per RP launch
                if (!tin && h->sc_h->done != (unsigned)h->nb * (unsigned)loop) {
                    status = BRP_ERR_STATE;
                    lost = (long long)h->nb * loop - (long long)h->sc_h->done;
                    break;
                }
                std::swap(target, prev);
                float a = 0.f, c = 0.f, o = 0.f;
                (void)hipEventElapsedTime(&o, h->ev[7], h->ev[4]);
                (void)hipEventElapsedTime(&a, h->ev[4], h->ev[5]);
                (void)hipEventElapsedTime(&c, h->ev[5], h->ev[6]);
                other_ms += o;
                bub_ms += a;
                pic_p_ms += c;
                if (h->sc_h->flags & 4) {
                    status = BRP_ERR_SUBSTEP_CAP;
                    break;
                }
            }
        }
        // calculate_temperature (bubble_CUDA.cu:279-282): the final f_g and p,
        // the previous step's properties
        if (h->cfg.thermal) {
            HIPCHK(h, hipEventRecord(h->ev[7], s));
            launch_temperature<F>(s, h->g, h->bp, h->b, bub, h->W, h->T_inf, h->T_d, h->dT_d,
                                  h->heat_d, h->fgb[prev], h->fgb[old_cur], h->first_step,
                                  (const F *)h->pb[p_i][0], (const F *)h->pb[p_i][1],
                                  (const F *)h->pb[pn_i][0], (const F *)h->pb[pn_i][1]);
            launch_focal_point(s, h->g, h->dT_d, h->heat_d, h->sc_d);
            HIPCHK(h, hipMemcpyAsync(h->sc_h->focal, h->sc_d->focal, sizeof(double) * 2,
                                     hipMemcpyDeviceToHost, s));
            HIPCHK(h, hipEventRecord(h->ev[4], s));
        }
        // rotate the void-fraction buffers: final iterate in `prev`, zero buffer in `target`
        h->fcur = prev;
        h->fz = target;
        h->fn = old_cur;
        h->fm = old_n;
        if (loop > 0) h->rp_cur ^= 1;
        if (bub && loop == 0) launch_copy_d(s, h->b.PL[h->pl_p], h->b.PL[h->pl_n], h->nb);
        h->pc = p_i;
        h->first_step = 0;
        if (!bub) HIPCHK(h, hipMemcpyAsync(h->sc_h, h->sc_d, hdr, hipMemcpyDeviceToHost, s));
        HIPCHK(h, hipStreamSynchronize(s));
        {
            float a = 0.f, b2 = 0.f, c = 0.f;
            (void)hipEventElapsedTime(&a, h->ev[0], h->ev[1]);
            (void)hipEventElapsedTime(&b2, h->ev[1], h->ev[2]);
            (void)hipEventElapsedTime(&c, h->ev[2], h->ev[3]);
            h->phase_ms[0] += a;
            h->phase_ms[1] += b2;
            h->phase_ms[2] += c + pic_p_ms;
            h->phase_ms[3] += bub_ms;
            if (h->cfg.thermal) {
                float th = 0.f;
                (void)hipEventElapsedTime(&th, h->ev[7], h->ev[4]);
                other_ms += th;
            }
            h->phase_ms[4] += other_ms;
        }
        const int flags = h->sc_h->flags;
        if (status == BRP_OK && (flags & 3)) status = BRP_ERR_NONFINITE;
        if (status == BRP_OK && (flags & 4)) status = BRP_ERR_SUBSTEP_CAP;
        st.nstep = h->nstep;
        st.time = h->t;
        st.resimax = resimax;
        st.n_pic = loop;
        if (h->cfg.thermal) {
            st.focal_x = h->sc_h->focal[0];
            st.focal_y = h->sc_h->focal[1];
        }
        if (rec) {
            rec->n_pic = loop;
            if (rec->want_grid) {
                int rc;
                if ((rc = copy_rect(h, rec->vx_end, vx, sizeof(F), rec->rect)) ||
                    (rc = copy_rect(h, rec->vy_end, vy, sizeof(F), rec->rect)))
                    return rc;
            }
            if (rec->want_bub && h->nb > 0) {
                HIPCHK(h, hipMemcpyAsync(rec->x_end, h->b.x, sizeof(double) * h->nb, hipMemcpyDeviceToDevice, s));
                HIPCHK(h, hipMemcpyAsync(rec->y_end, h->b.y, sizeof(double) * h->nb, hipMemcpyDeviceToDevice, s));
            }
        }
        if (ctl && ctl->npic_out) ctl->npic_out->push_back(loop);
        if (h->sched_mode == 1) {
            h->sched->npic.push_back(loop);
            unsigned int mm = 0;
            HIPCHK(h, hipMemcpy(&mm, h->sched->mismatch, sizeof mm, hipMemcpyDeviceToHost));
            if (mm) status = BRP_ERR_STATE;
        }
        st.rp_max_substeps = h->sc_h->sub_max;
        st.rp_reused = h->sc_h->reused;
        st.rp_mean_substeps = (loop > 0 && h->nb > 0)
                                  ? (double)h->sc_h->sub_sum / ((double)h->nb * loop)
                                  : 0.0;
        st.status = status;
        if (stats) stats[step] = st;
        if (status == BRP_ERR_STATE && lost)
            return set_err(h, status, "step %lld: the RP launch wrote back %lld bubbles fewer than "
                           "it was given", (long long)h->nstep, lost);
        if (status == BRP_ERR_STATE)
            return set_err(h, status, "step %lld: recording the schedule took another sub-step "
                           "count than the RP launch", (long long)h->nstep);
        if (status != BRP_OK)
            return set_err(h, status, "step %lld: %s (flags %d, Picard %d)", (long long)h->nstep,
                           status == BRP_ERR_PICARD_CAP ? "Picard cap reached"
                           : status == BRP_ERR_SUBSTEP_CAP ? "RP sub-step cap reached"
                                                           : "non-finite value or void fraction out of range",
                           flags, loop);
    }
    return BRP_OK;
}

template <typename F>
int alloc_grid(brp_handle *h) {
    int rc;
    const size_t fb = sizeof(F) * h->frame;
    if ((rc = dalloc(h, &h->vx, fb)) || (rc = dalloc(h, &h->vy, fb)) || (rc = dalloc(h, &h->p0, fb)))
        return rc;
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++)
            if ((rc = dalloc(h, &h->pb[a][c], fb))) return rc;
    for (int a = 0; a < 4; a++)
        if ((rc = dalloc(h, (void **)&h->fgb[a], sizeof(int64_t) * h->frame))) return rc;
    return init_coefficients<F>(h);
}


// ---- discrete adjoint ------------------------------------------------------
struct AdjState {
    AdjField f{};
    double *LfC = nullptr, *LfN = nullptr, *LfM = nullptr, *Lf = nullptr;  // f_g adjoints
    double *gA = nullptr, *gP = nullptr;       // per source column
    double *LR[4] = {}, *LRs[4] = {};          // RP state adjoints: step end / start
    double *LA = nullptr, *LB = nullptr;       // PL history (PL_p, PL_n of the step end)
    double *Lx = nullptr, *Ly = nullptr;
    double *lRt = nullptr, *lPLp = nullptr, *lPLn = nullptr, *lPLm = nullptr;
    unsigned int *mismatch = nullptr;  // RP replays whose sub-step count differed
    // RP Jacobians of a run of steps (k_adj_rp_jac): columns, iteration
    // descriptors, longest-first order and its bins and queue head
    double *jac = nullptr;
    double *jac_buf = nullptr;  // reverse-mode lane states of k_adj_rp_jac
    JacSlot *slots = nullptr;
    unsigned *ord = nullptr, *head = nullptr;
    int *bins = nullptr;
};

// Reverse of one replayed step.  On entry `a` holds the adjoint of the state
// after the step and the handle holds that state with `rec` its
// intermediates; jac[k] holds the RP Jacobian of Picard iteration k; on exit
// `a` holds the adjoint of the state before it.
template <typename F>
void adjoint_step(brp_handle *h, const StepRecord &rec, const StepRecord &brec, bool seed,
                  double cJ, const int box[4], const double *const *jac, AdjState &a) {
    hipStream_t s = h->stream;
    const GridParams &g = h->g;
    const bool bub = h->cfg.bub_enable && h->nb > 0;
    const int n = rec.n_pic;
    const Window &w = rec.rect;
    const int rw = w.i1 - w.i0 + 1, rh = w.j1 - w.j0 + 1;
    auto fr = [&](const void *p) { return Rect<F>{(const F *)p, w.i0, w.j0, rw, rh}; };
    auto gr = [&](const int64_t *p) { return Rect<int64_t>{p, w.i0, w.j0, rw, rh}; };
    const Rect<int64_t> gC = gr(rec.fgC);
    const Window pw{rec.win.i0 - 2, rec.win.i1 + 2, rec.win.j0 - 2, rec.win.j1 + 2};
    const F *ax = (const F *)h->ax, *bx = (const F *)h->bx, *ay = (const F *)h->ay,
            *by = (const F *)h->by;
    // the bubble arrays as the step saw them: positions after the motion, RP
    // start state in slot 0, PL history in slots 0 (m) and 1 (n)
    BubbleArrays bv = h->b;
    bv.x = brec.x_end;
    bv.y = brec.y_end;
    bv.R[0] = brec.rp[0];
    bv.Rn[0] = brec.rp[1];
    bv.d1R[0] = brec.rp[2];
    bv.PG[0] = brec.rp[3];
    bv.bdt[0] = brec.rp[4];
    bv.re[0] = brec.rp[5];
    bv.PL[0] = brec.PLm;
    bv.PL[1] = brec.PLn;
    if (seed) launch_adj_seed<F>(s, g, fr(rec.p0_it[n]), a.f.p0, box, cJ);
    // start-of-step f_g adjoints: f_gn' = f_g and f_gm' = f_gn of the start
    double *LgC = a.LfN, *LfNs = a.LfM, *LfMs = a.LfC;
    double *const none[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int k = n; k >= 1; k--) {
        launch_adj_pressure_ghosts(s, g, a.f.p0, h->el_amp_d, h->el_phase_d, rec.t, a.gA, a.gP);
        // before the step's last evaluation, the adjoint inputs are the gather
        // stencils' (inside the Picard window): that evaluation's transpose is
        // zero outside the window grown by the stencil's reach
        launch_adj_pressure<F>(s, g, false, a.f, gr(rec.fg_it[k]), gC, k == n ? a.LfC : nullptr, a.Lf,
                               LgC, nullptr, nullptr, ax, bx, ay, by, k == n ? nullptr : &pw);
        launch_adj_scatter(s, g, h->bp, bv, brec.Rt_it[k], a.Lf, a.lRt, a.Lx, a.Ly);
        launch_adj_rp_apply(s, bv.n, jac[k], a.lRt, k == n ? a.LR : none, a.LRs, a.lPLp, a.lPLn,
                            a.lPLm);
        launch_adj_gather<F>(s, g, h->bp, bv, fr(rec.p0_it[k - 1]), a.lPLp,
                             k == n ? a.LA : nullptr, a.f.p0, a.Lx, a.Ly);
    }
    launch_adj_pressure_ghosts(s, g, a.f.p0, h->el_amp_d, h->el_phase_d, rec.t, a.gA, a.gP);
    launch_adj_pressure<F>(s, g, true, a.f, gr(rec.fg_it[0]), gC, n == 0 ? a.LfC : nullptr, nullptr,
                           LgC, LfNs, LfMs, ax, bx, ay, by, nullptr);
    if (bub)
        launch_adj_motion<F>(s, g, h->bp, h->nb, brec.x_start, brec.y_start, fr(rec.vx_end),
                             fr(rec.vy_end), a.f.vx, a.f.vy, a.Lx, a.Ly);
    launch_adj_velocity_ghosts(s, g, a.f.vx, a.f.vy);
    launch_adj_velocity<F>(s, g, a.f.vx, a.f.vy, (const F *)rec.p0_start, gC, rec.first,
                           (const F *)h->cxa, (const F *)h->cxb, (const F *)h->cya,
                           (const F *)h->cyb, a.f.p0, LgC);
    if (bub) {
        launch_adj_pl_history(s, h->nb, n > 0, a.LA, a.LB, a.lPLn, a.lPLm);
        if (n > 0)
            for (int q = 0; q < 4; q++) {
                std::swap(a.LR[q], a.LRs[q]);
                launch_fill_d(s, a.LRs[q], (size_t)h->nb, 0.0);
            }
    }
    std::swap(a.f.px, a.f.pnx);
    std::swap(a.f.py, a.f.pny);
    launch_fill_d(s, a.f.pnx, h->frame, 0.0);
    launch_fill_d(s, a.f.pny, h->frame, 0.0);
    a.LfC = LgC;
    a.LfN = LfNs;
    a.LfM = LfMs;
}

template <typename F>
int adjoint_impl(brp_handle *h, const brp_objective &ob, double *grad_amp, double *grad_phase,
                 double *grad_fg0, int64_t ld, brp_adjoint_stats *stats) {
    using clk = std::chrono::steady_clock;
    const int n = ob.n_steps;
    const int K = ob.checkpoint_every > 0 ? std::min(ob.checkpoint_every, n)
                                          : (int)std::ceil(std::sqrt((double)n));
    const int nseg = (n + K - 1) / K;
    const int box[4] = {ob.box_i0, ob.box_i1, ob.box_j0, ob.box_j1};
    const double cJ =
        1.0 / ((double)ob.window * (box[1] - box[0] + 1) * (double)(box[3] - box[2] + 1));
    hipStream_t s = h->stream;
    AdjMem mem;
    std::vector<Snapshot> ck(nseg);
    Snapshot fin;
    std::vector<StepRecord> recs(K);  // grid part (and, without the tape, bubble part)
    // Bubble tape: the bubble part of every step, recorded by the forward
    // sweep, so that recomputing a segment replays the grid only (the RP
    // integrations, the dominant cost, are not repeated).  Used while it
    // fits in tape_budget bytes; otherwise each segment's recomputation
    // integrates RP again and records the bubble part itself.
    const bool bub = h->cfg.bub_enable && h->nb > 0;
    Arena arena;
    std::vector<StepRecord> tape;
    const size_t tape_budget = size_t(96) << 30;
    bool use_tape = bub && !(ob.flags & BRP_ADJ_NO_TAPE) &&
                    (size_t)n * (size_t)h->nb * 8 * 24 < tape_budget;
    if (use_tape) {
        tape.resize(n);
        for (auto &t : tape) {
            t.want_grid = false;
            t.arena = &arena;
        }
        for (auto &r : recs) r.want_bub = false;
    }
    for (auto &r : recs) std::copy(box, box + 4, r.box);
    struct Cleanup {
        std::vector<Snapshot> *a;
        Snapshot *c;
        std::vector<StepRecord> *r, *t;
        ~Cleanup() {
            for (auto &x : *a) snap_free(x);
            snap_free(*c);
            for (auto &x : *r) rec_free(x);
            for (auto &x : *t) rec_free(x);
        }
    } cleanup{&ck, &fin, &recs, &tape};
    int rc;
    double *jacc = nullptr;
    if ((rc = mem.get(h, &jacc, (size_t)n))) return rc;

    // forward sweep with checkpoints; Picard counts recorded
    const auto t0 = clk::now();
    std::vector<int> npic;
    npic.reserve(n);
    StepCtl fctl;
    fctl.npic_out = &npic;
    for (int sg = 0; sg < nseg; sg++) {
        if ((rc = snap_save(h, ck[sg]))) return rc;
        const int s_end = std::min(n, (sg + 1) * K);
        for (int st = sg * K + 1; st <= s_end; st++) {
            fctl.rec = use_tape ? &tape[st - 1] : nullptr;
            if ((rc = run_steps<F>(h, 1, nullptr, &fctl))) return rc;
            if (st > n - ob.window) launch_box_sumsq<F>(s, h->g, (const F *)h->p0, box, jacc + st - 1);
        }
    }
    if ((rc = snap_save(h, fin))) return rc;
    std::vector<double> jh(n);
    HIPCHK(h, hipMemcpy(jh.data(), jacc, sizeof(double) * n, hipMemcpyDeviceToHost));
    double J = 0.0;
    for (int st = n - ob.window + 1; st <= n; st++) J += jh[st - 1];
    J *= cJ;
    const auto t1 = clk::now();

    // adjoint state
    AdjState a;
    const size_t fr = h->frame, nb = (size_t)std::max(h->nb, 1);
    double **fields[] = {&a.f.p0, &a.f.px, &a.f.py, &a.f.pnx, &a.f.pny, &a.f.vx, &a.f.vy,
                         &a.LfC, &a.LfN, &a.LfM, &a.Lf};
    for (double **q : fields)
        if ((rc = mem.get(h, q, fr))) return rc;
    if ((rc = mem.get(h, &a.gA, (size_t)h->N + 2)) || (rc = mem.get(h, &a.gP, (size_t)h->N + 2)))
        return rc;
    double **bubs[] = {&a.LR[0], &a.LR[1], &a.LR[2], &a.LR[3], &a.LRs[0], &a.LRs[1], &a.LRs[2],
                       &a.LRs[3], &a.LA, &a.LB, &a.Lx, &a.Ly, &a.lRt, &a.lPLp, &a.lPLn, &a.lPLm};
    for (double **q : bubs)
        if ((rc = mem.get(h, q, nb))) return rc;
    if ((rc = mem.get(h, &a.mismatch, 1))) return rc;
    // RP Jacobians: 7 columns per bubble and Picard iteration, 35 in a
    // step's last; computed for runs of steps that fit jac_budget bytes
    // (at least one step) and at most 2^31 items per launch
    auto jac_doubles = [&](int np) -> size_t {
        return np > 0 ? nb * (size_t)(kJacCols * (np - 1) + kJacColsLast) : 0;
    };
    const size_t jac_budget = size_t(24) << 30;
    size_t jac_cap = 0, item_cap = 0, slot_cap = 0;
    if (bub) {
        for (int sg = 0; sg < nseg; sg++) {  // the largest run of steps any segment needs
            size_t dsum = 0, isum = 0, ssum = 0;
            for (int st = std::min(n, (sg + 1) * K); st >= sg * K + 1; st--) {
                const size_t dj = jac_doubles(npic[st - 1]);
                const size_t ij = nb * (size_t)npic[st - 1];
                if (ssum && ((dsum + dj) * 8 > jac_budget || isum + ij >= (size_t(1) << 31))) {
                    dsum = isum = ssum = 0;
                }
                dsum += dj;
                isum += ij;
                ssum += npic[st - 1];
                jac_cap = std::max(jac_cap, dsum);
                item_cap = std::max(item_cap, isum);
                slot_cap = std::max(slot_cap, ssum);
            }
        }
        if (item_cap >= (size_t(1) << 31))
            return set_err(h, BRP_ERR_INVALID, "one step's RP Jacobians exceed 2^31 items");
        if ((rc = mem.get(h, &a.jac, std::max<size_t>(jac_cap, 1))) ||
            (rc = mem.get(h, &a.slots, std::max<size_t>(slot_cap, 1))) ||
            (rc = mem.get(h, &a.ord, std::max<size_t>(item_cap, 1))) ||
            (rc = mem.get(h, &a.head, 2)) || (rc = mem.get(h, &a.bins, kAdjJacBins)))
            return rc;
    }
    std::vector<JacSlot> slots_h;
    std::vector<std::vector<const double *>> jk(K);
    const int jac_blocks = bub ? adj_rp_jac_per_cu() * h->cus : 1;
    if (bub && (rc = mem.get(h, &a.jac_buf, adj_rp_jac_buf_doubles(jac_blocks)))) return rc;

    // reverse sweep: per segment, the steps are recomputed from the segment's
    // checkpoint with their intermediates recorded, then reversed
    for (int sg = nseg - 1; sg >= 0; sg--) {
        if ((rc = snap_restore(h, ck[sg]))) return rc;
        const int s0 = sg * K + 1, s1 = std::min(n, (sg + 1) * K);
        for (int st = s0; st <= s1; st++) {
            StepCtl c;
            c.force_npic = &npic[st - 1];
            c.rec = &recs[st - s0];
            c.tape_in = use_tape ? &tape[st - 1] : nullptr;
            if ((rc = run_steps<F>(h, 1, nullptr, &c))) return rc;
            if (recs[st - s0].n_pic != npic[st - 1])
                return set_err(h, BRP_ERR_STATE, "replay of step %d took %d Picard iterations, not %d",
                               st, recs[st - s0].n_pic, npic[st - 1]);
        }
        // reversed in runs [lo, hi] whose RP Jacobians are computed together
        for (int hi = s1; hi >= s0;) {
            int lo = hi;
            size_t dsum = 0, isum = 0, ssum = 0;
            for (; lo >= s0; lo--) {
                const int np = npic[lo - 1];
                const size_t dj = jac_doubles(np), ij = nb * (size_t)np;
                if (lo < hi &&
                    (dsum + dj > jac_cap || isum + ij > item_cap || ssum + np > slot_cap))
                    break;
                dsum += dj;
                isum += ij;
                ssum += np;
            }
            lo++;
            slots_h.clear();
            double *jp = a.jac;
            for (int st = lo; st <= hi; st++) {
                const StepRecord &br = use_tape ? tape[st - 1] : recs[st - s0];
                const int np = npic[st - 1];
                std::vector<const double *> &v = jk[st - s0];
                v.assign(np + 1, nullptr);
                for (int k = 1; k <= np; k++) {
                    JacSlot sl{};
                    for (int q = 0; q < 6; q++) sl.rp[q] = br.rp[q];
                    sl.PLm = br.PLm;
                    sl.PLn = br.PLn;
                    sl.PLp = br.PLp_it[k];
                    sl.sub = br.sub_it[k];
                    sl.jac = jp;
                    sl.last = k == np;
                    v[k] = jp;
                    if (jp) jp += nb * (size_t)(k == np ? kJacColsLast : kJacCols);
                    slots_h.push_back(sl);
                }
            }
            if (bub && !slots_h.empty()) {
                HIPCHK(h, hipMemcpyAsync(a.slots, slots_h.data(), sizeof(JacSlot) * slots_h.size(),
                                         hipMemcpyHostToDevice, s));
                launch_adj_rp_jacobians(s, h->bp, h->b, h->g.dt, a.slots, (int)slots_h.size(), a.ord,
                                        a.bins, a.head, a.mismatch, a.jac_buf, jac_blocks);
            }
            for (int st = hi; st >= lo; st--)
                adjoint_step<F>(h, recs[st - s0], use_tape ? tape[st - 1] : recs[st - s0],
                                st > n - ob.window, cJ, box, jk[st - s0].data(), a);
            // slots_h is rewritten by the next run: its copy must have completed
            HIPCHK(h, hipStreamSynchronize(s));
            hi = lo - 1;
        }
    }
    HIPCHK(h, hipStreamSynchronize(s));
    const auto t2 = clk::now();

    // gradients
    if (grad_fg0) {
        std::vector<double> c(fr), nn(fr), m(fr);
        HIPCHK(h, hipMemcpy(c.data(), a.LfC, sizeof(double) * fr, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(nn.data(), a.LfN, sizeof(double) * fr, hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(m.data(), a.LfM, sizeof(double) * fr, hipMemcpyDeviceToHost));
        for (int j = 0; j <= h->M + 1; j++)
            for (int i = 0; i <= h->N + 1; i++) {
                const int64_t o = frame_off(i, j, h->pitch);
                grad_fg0[(int64_t)j * ld + i] = c[o] + nn[o] + m[o];
            }
    }
    if ((grad_amp || grad_phase) && h->n_elem > 0) {
        std::vector<double> gA(h->N + 2), gP(h->N + 2);
        HIPCHK(h, hipMemcpy(gA.data(), a.gA, sizeof(double) * gA.size(), hipMemcpyDeviceToHost));
        HIPCHK(h, hipMemcpy(gP.data(), a.gP, sizeof(double) * gP.size(), hipMemcpyDeviceToHost));
        std::vector<double> ea(h->n_elem, 0.0), ep(h->n_elem, 0.0);
        const long n_ap = h->g.ap_hi - h->g.ap_lo + 1;
        for (int i = std::max(1, h->g.ap_lo); i <= std::min(h->N, h->g.ap_hi); i++) {
            const int e = (int)((long)(i - h->g.ap_lo) * h->n_elem / n_ap);
            ea[e] += gA[i];
            ep[e] += gP[i];
        }
        if (grad_amp) std::copy(ea.begin(), ea.end(), grad_amp);
        if (grad_phase) std::copy(ep.begin(), ep.end(), grad_phase);
    }
    unsigned int mism = 0;
    HIPCHK(h, hipMemcpy(&mism, a.mismatch, sizeof mism, hipMemcpyDeviceToHost));
    if (stats) {
        stats->schedule_mismatch = mism;
        stats->J = J;
        stats->forward_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
        stats->reverse_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        // step records: per evaluation p0 and f_g, per iteration PL_p, R_t and
        // sub-step counts, per step start/end fields and bubble state
        size_t recb = arena.bytes();
        for (const StepRecord &r : recs) recb += rec_bytes(h, r);
        stats->memory_bytes = (int64_t)((nseg + 1) * snap_bytes(h) + mem.bytes + recb);
        stats->n_checkpoints = nseg;
        stats->segment = K;
    }
    return snap_restore(h, fin);
}
}  // namespace

extern "C" {

void brp_config_default(brp_config *cfg) {
    if (!cfg) return;
    std::memset(cfg, 0, sizeof *cfg);
    cfg->sim_order = 2;
    cfg->sim_smooth_delta_band = 2;
    cfg->pml_order = 2;
    cfg->pml_sigma = 0.1;
    cfg->sim_save_interval = 100;
    cfg->tr_pid_start_step = 1;
    cfg->tr_pid_iterate = 1;
    cfg->precision = BRP_PREC_FP64;
}

const char *brp_last_error(const brp_handle *h) {
    return h ? h->err.c_str() : g_create_error.c_str();
}

int brp_create(const brp_config *cfg, brp_handle **out) {
    if (!out) return set_err(nullptr, BRP_ERR_INVALID, "out is NULL");
    *out = nullptr;
    if (!cfg) return set_err(nullptr, BRP_ERR_INVALID, "cfg is NULL");
    const brp_config &c = *cfg;
    if (c.grid_x < 4 || c.grid_y < 4 || !(c.grid_lx > 0) || !(c.grid_ly > 0))
        return set_err(nullptr, BRP_ERR_INVALID, "grid.x, grid.y >= 4 and grid.lx, grid.ly > 0 required");
    if (c.grid_x > 65536 || c.grid_y > 65536)
        return set_err(nullptr, BRP_ERR_INVALID, "grid.x and grid.y must be <= 65536");
    if (c.sim_order != 2) return set_err(nullptr, BRP_ERR_INVALID, "simulation.order must be 2");
    if (c.sim_smooth_delta_band < 1 || c.sim_smooth_delta_band > 8)
        return set_err(nullptr, BRP_ERR_INVALID, "simulation.smooth_delta_band must be in [1, 8]");
    if ((c.pml_x0 || c.pml_x1 || c.pml_y0 || c.pml_y1) &&
        (c.pml_grid_number < 1 || !(c.pml_sigma > 0.0)))
        return set_err(nullptr, BRP_ERR_INVALID, "pml.grid_number >= 1 and pml.sigma > 0 required with PML");
    if (!(c.sim_cfl > 0)) return set_err(nullptr, BRP_ERR_INVALID, "simulation.cfl must be > 0");
    if (c.precision != BRP_PREC_FP64 && c.precision != BRP_PREC_MIXED)
        return set_err(nullptr, BRP_ERR_INVALID, "precision must be 0 (FP64) or 1 (mixed)");

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_err(nullptr, BRP_ERR_NO_DEVICE, "no HIP device visible");

    brp_handle *h = new brp_handle();
    h->cfg = c;
    h->prec = c.precision;
    h->fsz = h->prec == BRP_PREC_FP64 ? 8 : 4;
    h->N = c.grid_x;
    h->M = c.grid_y;
    // columns -1..N+1 from offset kCol0 - 1, plus room for the grid kernels'
    // last 4-cell group of a row and its right neighbour (columns up to N + 4)
    h->pitch = ((h->N + kCol0 + 5 + 63) / 64) * 64;
    h->rows = h->M + 3;
    h->frame = (size_t)h->pitch * h->rows;
    auto fail = [&](int rc) {
        g_create_error = h->err;
        brp_destroy(h);
        return rc;
    };
    if (h->frame > (size_t)INT32_MAX)  // the RP scatter keys cells by 32-bit frame offset
        return fail(set_err(h, BRP_ERR_INVALID, "grid of %d x %d cells exceeds 2^31 frame elements",
                            h->N, h->M));
    if (hipGetDevice(&h->device) != hipSuccess)
        return fail(set_err(h, BRP_ERR_DEVICE, "hipGetDevice failed"));
    rp_grid_query(h->device, &h->cus, &h->per_cu);
    {
        hipError_t e;
        h->scratch_d = scratch_address(h->device, &e);
        if (!h->scratch_d)
            h->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    }
    if (hipStreamCreate(&h->stream) != hipSuccess)
        return fail(set_err(h, BRP_ERR_DEVICE, "hipStreamCreate failed"));
    h->ev.resize(8);
    for (auto &e : h->ev)
        if (hipEventCreate(&e) != hipSuccess) return fail(set_err(h, BRP_ERR_DEVICE, "hipEventCreate failed"));

    // init_grid_size (bubble_CUDA.cu:524-540)
    GridParams &g = h->g;
    g.N = h->N;
    g.M = h->M;
    g.pitch = h->pitch;
    g.dx = (double)c.grid_lx / (double)c.grid_x;
    g.dy = (double)c.grid_ly / (double)c.grid_y;
    g.rdx = (double)1.0 / (double)g.dx;
    g.rdy = (double)1.0 / (double)g.dy;
    // init_transducer (bubble_CUDA.cu:543-558), E2 focus x
    h->fp_x = 0.0;
    h->fp_y = c.tr_focal_dist ? c.tr_focal_dist * 0.5 * std::sqrt(3.0) : c.grid_ly * 0.5;
    if (c.ext_focus_x_set) h->fp_x = c.ext_focus_x;
    h->omega = 2.0 * std::acos(-1.0) * c.tr_freq;
    // control point (bubble_CUDA.cu:132-153, 214-215)
    if (c.tr_pid && c.tr_init_control) {
        h->ctrl_x = c.ext_focus_x_set ? h->fp_x : 0.0;
        h->ctrl_y = c.tr_init_control;
    } else {
        h->ctrl_x = h->fp_x;
        h->ctrl_y = h->fp_y;
    }
    h->ctrl_dist = h->ctrl_y / 0.5 / std::sqrt(3.0);
    // init_mix, mix_set_time_increment (bubble_CUDA.cu:636-658)
    h->T_inf = 293.15;
    h->P_inf = 0.1e6;
    h->rho_inf = density_water();
    h->cs_inf = sound_speed_water();
    h->dt = c.sim_cfl * std::min(g.dx, g.dy) / h->cs_inf;
    h->dt0_rp = 0.1 * h->dt;
    g.dt = h->dt;
    g.rho_inf = h->rho_inf;
    g.cs_inf = h->cs_inf;
    g.P_inf = h->P_inf;
    g.pml_x0 = c.pml_x0;
    g.pml_x1 = c.pml_x1;
    g.pml_y0 = c.pml_y0;
    g.pml_y1 = c.pml_y1;
    g.plane_v = c.tr_plane_v;
    g.plane_p = c.tr_plane_p;
    g.focused_v = c.tr_focused_v;
    g.focused_p = c.tr_focused_p;
    g.src_row = (!c.pml_y0 || c.ext_source_in_pml) ? 1 : 0;
    g.cylindrical = c.tr_cylindrical;
    g.amp = c.tr_amp;
    g.freq = c.tr_freq;
    g.omega = h->omega;
    g.on_wave = c.tr_on_wave;
    g.off_wave = c.tr_off_wave;
    g.fp_x = h->ctrl_x;
    g.fp_y = h->ctrl_y;
    g.fdist = h->ctrl_dist;
    g.n_elem = 0;
    // MixturePressureKernel s3 = dt * rho_l * c * c (kernel.cuh:852)
    g.s3 = h->dt * h->rho_inf * h->cs_inf * h->cs_inf;
    find_aperture(h);
    init_sigma(h);
    // bubble parameters (init_bub_params, bubble_CUDA.cu:661-675; delta_coef 1467-1471)
    BubbleParams &bp = h->bp;
    bp.PL0 = c.bub_P0;
    bp.T0 = c.bub_T0;
    bp.gam = c.bub_gamma;
    bp.sig = c.bub_sigma;
    bp.mu = c.bub_mu;
    bp.rho = c.bub_rho;
    bp.K0 = c.bub_K0;
    bp.dt0 = h->dt0_rp;
    bp.mbs = -c.sim_smooth_delta_band / 2 + 1;
    bp.mbe = c.sim_smooth_delta_band + bp.mbs - 1;
    bp.nbs = -c.sim_smooth_delta_band / 2;
    bp.nbe = c.sim_smooth_delta_band + bp.nbs - 1;
    bp.dcx = 1.0 / ((double)c.sim_smooth_delta_band);
    bp.dcy = 2.0 * std::acos(-1.0) / ((double)c.sim_smooth_delta_band) * g.rdx;
    bp.dcz = 2.0 * std::acos(-1.0) / ((double)c.sim_smooth_delta_band) * g.rdy;
    bp.pi = std::acos(-1.0);
    bp.pi4r3 = bp.pi * 4.0 / 3.0;
    bp.max_substeps = c.max_substeps > 0 ? c.max_substeps : 1000000;
    bp.cylindrical = c.tr_cylindrical;
    bp.rp_exact = c.rp_exact ? 1 : 0;
    bp.heavy_threshold = 48;
    bp.extreme_threshold = 256;

    int rc = h->prec == BRP_PREC_FP64 ? alloc_grid<double>(h) : alloc_grid<float>(h);
    if (rc) return fail(rc);
    if (c.thermal &&
        ((rc = dalloc(h, (void **)&h->T_d, sizeof(double) * h->frame)) ||
         (rc = dalloc(h, (void **)&h->dT_d, sizeof(double) * h->frame)) ||
         (rc = dalloc(h, (void **)&h->heat_d, sizeof(double) * h->frame))))
        return fail(rc);
    if ((rc = dalloc(h, (void **)&h->sc_d, sizeof(StepScalars)))) return fail(rc);
    {
        double v[kRpTable] = {};
        rp_table_values(bp, v);
        double *tab = nullptr;
        if ((rc = dalloc(h, (void **)&tab, sizeof v))) return fail(rc);
        if (hipMemcpy(tab, v, sizeof v, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_err(h, BRP_ERR_DEVICE, "hipMemcpy failed"));
        bp.rp_shared = tab;
    }
    if (hipHostMalloc((void **)&h->sc_h, sizeof(StepScalars), hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void **)&h->rpl_h, sizeof(RPLaunch), hipHostMallocDefault) != hipSuccess)
        return fail(set_err(h, BRP_ERR_DEVICE, "hipHostMalloc failed"));
    if ((rc = dalloc(h, (void **)&h->rpl_d, sizeof(RPLaunch)))) return fail(rc);
    // init_bub_array (bubble_CUDA.cu:797-897): one bubble per cell centre in the box
    if (c.bub_enable) {
        for (int i = 1; i <= h->N; i++) {
            const double px = ((double)i - 0.5) * g.dx;
            for (int j = 1; j <= h->M; j++) {
                const double py = ((double)j - 0.5) * g.dy;
                if (c.tr_box_size ? (std::fabs(px - h->fp_x) < 0.5 * c.tr_box_size &&
                                     std::fabs(py - h->fp_y) < 0.5 * c.tr_box_size)
                                  : true) {
                    h->pending_x.push_back(px);
                    h->pending_y.push_back(py);
                }
            }
        }
        h->pending_R0.assign(h->pending_x.size(), c.bub_R0);
    }
    if (hipStreamSynchronize(h->stream) != hipSuccess)
        return fail(set_err(h, BRP_ERR_DEVICE, "device initialisation failed"));
    *out = h;
    return BRP_OK;
}

void brp_destroy(brp_handle *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (auto &sn : h->snaps)
        for (void *p : sn.bufs) (void)hipFree(p);
    for (void *p : h->allocs) (void)hipFree(p);
    if (h->sc_h) (void)hipHostFree(h->sc_h);
    if (h->rpl_h) (void)hipHostFree(h->rpl_h);
    for (auto &e : h->ev)
        if (e) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

int brp_get_info(const brp_handle *h, brp_info *info) {
    if (!h || !info) return BRP_ERR_INVALID;
    std::memset(info, 0, sizeof *info);
    info->nx = h->N;
    info->ny = h->M;
    info->n_bubbles = h->b.x ? h->nb : (int)h->pending_x.size();
    info->precision = h->prec;
    info->dx = h->g.dx;
    info->dy = h->g.dy;
    info->rdx = h->g.rdx;
    info->rdy = h->g.rdy;
    info->dt = h->dt;
    info->dt0_rp = h->dt0_rp;
    info->fp_x = h->fp_x;
    info->fp_y = h->fp_y;
    info->omega = h->omega;
    info->time = h->t;
    info->nstep = h->nstep;
    info->device_bytes = (int64_t)h->dev_bytes;
    return BRP_OK;
}

int brp_set_bubbles(brp_handle *h, int32_t n, const double *x, const double *y, const double *R0,
                    const double *n_B, double fg_in) {
    if (!h) return BRP_ERR_INVALID;
    if (n < 0 || (n > 0 && (!x || !y || !R0))) return set_err(h, BRP_ERR_INVALID, "bad bubble arrays");
    if (h->prepared) return set_err(h, BRP_ERR_STATE, "set_bubbles after prepare");
    if (!h->cfg.bub_enable) return set_err(h, BRP_ERR_INVALID, "bubbles.enable is 0");
    for (int k = 0; k < n; k++)
        if (!(R0[k] > 0) || !std::isfinite(x[k]) || !std::isfinite(y[k]))
            return set_err(h, BRP_ERR_INVALID, "bubble %d: R0 must be > 0, x and y finite", k);
    h->pending_x.assign(x, x + n);
    h->pending_y.assign(y, y + n);
    h->pending_R0.assign(R0, R0 + n);
    if (n_B)
        h->pending_nB.assign(n_B, n_B + n);
    else
        h->pending_nB.clear();
    h->cfg.bub_fg0 = fg_in;  // fg_in of bubble_input for these bubbles
    return BRP_OK;
}

int brp_set_source(brp_handle *h, int32_t n_elem, const double *amp, const double *phase) {
    if (!h) return BRP_ERR_INVALID;
    if (n_elem < 0 || (n_elem > 0 && (!amp || !phase)))
        return set_err(h, BRP_ERR_INVALID, "bad source arrays");
    std::vector<double> a(amp, amp + n_elem), p(phase, phase + n_elem);
    if (!h->el_amp_d) {
        int rc;
        if ((rc = dalloc(h, (void **)&h->el_amp_d, sizeof(double) * 4096)) ||
            (rc = dalloc(h, (void **)&h->el_phase_d, sizeof(double) * 4096)))
            return rc;
    }
    if (n_elem > 4096) return set_err(h, BRP_ERR_INVALID, "at most 4096 source elements");
    if (n_elem > 0) {
        HIPCHK(h, hipMemcpy(h->el_amp_d, a.data(), sizeof(double) * n_elem, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->el_phase_d, p.data(), sizeof(double) * n_elem, hipMemcpyHostToDevice));
    }
    h->n_elem = n_elem;
    h->g.n_elem = n_elem;
    return BRP_OK;
}

int brp_prepare(brp_handle *h) {
    if (!h) return BRP_ERR_INVALID;
    if (h->prepared) return BRP_OK;
    int rc;
    if (h->cfg.bub_enable) {
        const std::vector<double> *nB = h->pending_nB.empty() ? nullptr : &h->pending_nB;
        if ((rc = place_bubbles(h, h->pending_x, h->pending_y, h->pending_R0, nB, h->cfg.bub_fg0)))
            return rc;
        // initial window: every bubble's scatter stencil
        for (int k = 0; k < h->nb; k++) {
            const int ix = host_bin(h->pending_x[k], h->g.rdx, 0.5);
            const int iy = host_bin(h->pending_y[k], h->g.rdy, 0.5);
            window_union(h->W, std::max(ix + h->bp.mbs, 0), std::min(ix + h->bp.mbe, h->N + 1),
                         std::max(iy + h->bp.mbs, 0), std::min(iy + h->bp.mbe, h->M + 1));
        }
        launch_zero_scalars(h->stream, h->sc_d, 1);
        if (h->cfg.tr_cylindrical && h->W.i0 <= 1 && h->W.i0 <= h->W.i1) h->W.i0 = 0;
        launch_scatter_initial(h->stream, h->g, h->bp, h->b, h->fgb[0], h->sc_d);
        if (h->cfg.tr_cylindrical) launch_cyl_fold(h->stream, h->g, h->fgb[0]);
        HIPCHK(h, hipMemcpyAsync(h->fgb[1], h->fgb[0], sizeof(int64_t) * h->frame, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->fgb[2], h->fgb[0], sizeof(int64_t) * h->frame, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(h, hipMemcpyAsync(h->sc_h, h->sc_d, offsetof(StepScalars, resid), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (h->sc_h->flags) return set_err(h, BRP_ERR_NONFINITE, "initial void fraction out of range");
    }
    h->fcur = 0;
    h->fn = 1;
    h->fm = 2;
    h->fz = 3;
    h->prepared = true;
    return BRP_OK;
}

int brp_step(brp_handle *h, int32_t k, brp_step_stats *stats) {
    if (!h || k < 0) return BRP_ERR_INVALID;
    int rc = brp_prepare(h);
    if (rc) return rc;
    for (double &v : h->phase_ms) v = 0.0;
    return h->prec == BRP_PREC_FP64 ? run_steps<double>(h, k, stats) : run_steps<float>(h, k, stats);
}

int brp_field_shape(const brp_handle *h, int32_t field, int32_t *nx, int32_t *ny, int32_t *i0,
                    int32_t *j0) {
    if (!h || field < 0 || field >= BRP_FIELD_COUNT) return BRP_ERR_INVALID;
    int a, b2, c, d;
    field_extent(h, field, &a, &b2, &c, &d);
    if (nx) *nx = a;
    if (ny) *ny = b2;
    if (i0) *i0 = c;
    if (j0) *j0 = d;
    return BRP_OK;
}

int brp_get_field(brp_handle *h, int32_t field, double *dst, int64_t ld) {
    if (!h || !dst || field < 0 || field >= BRP_FIELD_COUNT) return BRP_ERR_INVALID;
    int nx, ny, i0, j0;
    field_extent(h, field, &nx, &ny, &i0, &j0);
    if (ld < nx) return set_err(h, BRP_ERR_INVALID, "ld %lld < %d", (long long)ld, nx);
    int rc = brp_prepare(h);
    if (rc) return rc;
    std::vector<double> fr;
    const bool dbl = h->prec == BRP_PREC_FP64;
    auto rd = [&](const void *src) { return dbl ? read_frame<double>(h, src, fr) : read_frame<float>(h, src, fr); };
    switch (field) {
        case BRP_FIELD_P0: rc = rd(h->p0); break;
        case BRP_FIELD_PX: rc = rd(h->pb[h->pc][0]); break;
        case BRP_FIELD_PY: rc = rd(h->pb[h->pc][1]); break;
        case BRP_FIELD_VX: rc = rd(h->vx); break;
        case BRP_FIELD_VY: rc = rd(h->vy); break;
        case BRP_FIELD_FG: rc = read_frame_fg(h, h->fgb[h->fcur], fr); break;
        case BRP_FIELD_FGN: rc = read_frame_fg(h, h->fgb[h->fn], fr); break;
        case BRP_FIELD_FGM: rc = read_frame_fg(h, h->fgb[h->fm], fr); break;
        case BRP_FIELD_RHO_M:
            if (h->first_step) {
                fr.assign(h->frame, h->rho_inf);
            } else {
                rc = read_frame_fg(h, h->fgb[h->fcur], fr);
                for (double &v : fr) v = (1.0 - v) * h->rho_inf;
            }
            break;
        case BRP_FIELD_T:
            if (h->T_d) {
                fr.resize(h->frame);
                HIPCHK(h, hipMemcpy(fr.data(), h->T_d, sizeof(double) * h->frame, hipMemcpyDeviceToHost));
            } else {
                fr.assign(h->frame, 0.0);
            }
            break;
        default: return BRP_ERR_INVALID;
    }
    if (rc) return rc;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) dst[(size_t)j * ld + i] = fr[frame_off(i + i0, j + j0, h->pitch)];
    return BRP_OK;
}

int brp_set_field(brp_handle *h, int32_t field, const double *src, int64_t ld) {
    if (!h || !src || field < 0 || field >= BRP_FIELD_COUNT) return BRP_ERR_INVALID;
    int rc = brp_prepare(h);
    if (rc) return rc;
    int nx, ny, i0, j0;
    field_extent(h, field, &nx, &ny, &i0, &j0);
    if (ld < nx) return set_err(h, BRP_ERR_INVALID, "ld %lld < %d", (long long)ld, nx);
    if (field == BRP_FIELD_FG || field == BRP_FIELD_FGN || field == BRP_FIELD_FGM) {
        std::vector<int64_t> fr(h->frame, 0);
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                const double v = src[(size_t)j * ld + i];
                if (!(std::fabs(v) < 1.0)) return set_err(h, BRP_ERR_INVALID, "void fraction must be in (-1, 1)");
                const int64_t q = (int64_t)std::llrint(v * kFgScale);
                fr[frame_off(i + i0, j + j0, h->pitch)] = q;
                if (q != 0) window_union(h->W, i + i0, i + i0, j + j0, j + j0);
            }
        if (field == BRP_FIELD_FG) {
            for (int a : {h->fcur, h->fn, h->fm})
                HIPCHK(h, hipMemcpy(h->fgb[a], fr.data(), sizeof(int64_t) * h->frame, hipMemcpyHostToDevice));
        } else {
            const int a = field == BRP_FIELD_FGN ? h->fn : h->fm;
            HIPCHK(h, hipMemcpy(h->fgb[a], fr.data(), sizeof(int64_t) * h->frame, hipMemcpyHostToDevice));
        }
        HIPCHK(h, hipMemset(h->fgb[h->fz], 0, sizeof(int64_t) * h->frame));
        return BRP_OK;
    }
    if (field == BRP_FIELD_T) {
        if (!h->T_d) return set_err(h, BRP_ERR_INVALID, "T is not settable with thermal = 0");
        std::vector<double> fr(h->frame);
        HIPCHK(h, hipMemcpy(fr.data(), h->T_d, sizeof(double) * h->frame, hipMemcpyDeviceToHost));
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) fr[frame_off(i + i0, j + j0, h->pitch)] = src[(size_t)j * ld + i];
        HIPCHK(h, hipMemcpy(h->T_d, fr.data(), sizeof(double) * h->frame, hipMemcpyHostToDevice));
        return BRP_OK;
    }
    void *dstp = nullptr;
    switch (field) {
        case BRP_FIELD_P0: dstp = h->p0; break;
        case BRP_FIELD_PX: dstp = h->pb[h->pc][0]; break;
        case BRP_FIELD_PY: dstp = h->pb[h->pc][1]; break;
        case BRP_FIELD_VX: dstp = h->vx; break;
        case BRP_FIELD_VY: dstp = h->vy; break;
        default: return set_err(h, BRP_ERR_INVALID, "field %d is not settable", field);
    }
    const bool dbl = h->prec == BRP_PREC_FP64;
    std::vector<double> fr;
    if ((rc = dbl ? read_frame<double>(h, dstp, fr) : read_frame<float>(h, dstp, fr))) return rc;
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) fr[frame_off(i + i0, j + j0, h->pitch)] = src[(size_t)j * ld + i];
    if (dbl) {
        HIPCHK(h, hipMemcpy(dstp, fr.data(), sizeof(double) * h->frame, hipMemcpyHostToDevice));
    } else {
        std::vector<float> tmp(fr.begin(), fr.end());
        HIPCHK(h, hipMemcpy(dstp, tmp.data(), sizeof(float) * h->frame, hipMemcpyHostToDevice));
    }
    return BRP_OK;
}

int brp_get_bubbles(brp_handle *h, int32_t which, double *dst) {
    if (!h || !dst || which < 0 || which >= BRP_BUB_COUNT) return BRP_ERR_INVALID;
    int rc = brp_prepare(h);
    if (rc) return rc;
    const int n = h->nb;
    if (n == 0) return BRP_OK;
    const double *src = nullptr;
    double *tmpd = nullptr;
    const int c = h->rp_cur;
    switch (which) {
        case BRP_BUB_X: src = h->b.x; break;
        case BRP_BUB_Y: src = h->b.y; break;
        case BRP_BUB_R_T: src = h->b.R_t; break;
        case BRP_BUB_R_P: src = h->b.R[c]; break;
        case BRP_BUB_R_N: src = h->b.Rn[c]; break;
        case BRP_BUB_D1R: src = h->b.d1R[c]; break;
        case BRP_BUB_PG: src = h->b.PG[c]; break;
        case BRP_BUB_PL_P: src = h->b.PL[h->pl_p]; break;
        case BRP_BUB_PL_N: src = h->b.PL[h->pl_n]; break;
        case BRP_BUB_PL_M: src = h->b.PL[h->pl_m]; break;
        case BRP_BUB_Q_B: src = h->b.Q_B; break;
        case BRP_BUB_N_B: src = h->b.n_B; break;
        case BRP_BUB_DT: src = h->b.bdt[c]; break;
        case BRP_BUB_RE: src = h->b.re[c]; break;
        case BRP_BUB_VX: src = h->b.vx; break;
        case BRP_BUB_VY: src = h->b.vy; break;
        case BRP_BUB_R0: src = h->b.R0; break;
        case BRP_BUB_SUB: {
            std::vector<int> iv(n);
            HIPCHK(h, hipMemcpy(iv.data(), h->b.sub, sizeof(int) * n, hipMemcpyDeviceToHost));
            for (int k = 0; k < n; k++) dst[h->perm[k]] = (double)iv[k];
            return BRP_OK;
        }
        default: {
            HIPCHK(h, hipMalloc((void **)&tmpd, sizeof(double) * n));
            launch_bubble_index(h->stream, h->g, h->b, tmpd, which - BRP_BUB_IBM_X);
            src = tmpd;
        }
    }
    std::vector<double> v(n);
    hipError_t e = hipMemcpyAsync(v.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (tmpd) (void)hipFree(tmpd);
    if (e != hipSuccess) return set_err(h, BRP_ERR_DEVICE, "get_bubbles: %s", hipGetErrorString(e));
    for (int k = 0; k < n; k++) dst[h->perm[k]] = v[k];
    return BRP_OK;
}

int brp_set_bubble_array(brp_handle *h, int32_t which, const double *src) {
    if (!h || !src || which < 0 || which >= BRP_BUB_COUNT) return BRP_ERR_INVALID;
    int rc = brp_prepare(h);
    if (rc) return rc;
    const int n = h->nb;
    if (n == 0) return BRP_OK;
    const int c = h->rp_cur;
    double *dst = nullptr;
    switch (which) {
        case BRP_BUB_X: dst = h->b.x; break;
        case BRP_BUB_Y: dst = h->b.y; break;
        case BRP_BUB_R_T: dst = h->b.R_t; break;
        case BRP_BUB_R_P: dst = h->b.R[c]; break;
        case BRP_BUB_R_N: dst = h->b.Rn[c]; break;
        case BRP_BUB_D1R: dst = h->b.d1R[c]; break;
        case BRP_BUB_PG: dst = h->b.PG[c]; break;
        case BRP_BUB_PL_P: dst = h->b.PL[h->pl_p]; break;
        case BRP_BUB_PL_N: dst = h->b.PL[h->pl_n]; break;
        case BRP_BUB_PL_M: dst = h->b.PL[h->pl_m]; break;
        case BRP_BUB_Q_B: dst = h->b.Q_B; break;
        case BRP_BUB_DT: dst = h->b.bdt[c]; break;
        case BRP_BUB_RE: dst = h->b.re[c]; break;
        case BRP_BUB_VX: dst = h->b.vx; break;
        case BRP_BUB_VY: dst = h->b.vy; break;
        default: return set_err(h, BRP_ERR_INVALID, "bubble array %d is not settable", which);
    }
    std::vector<double> v(n);
    for (int k = 0; k < n; k++) {
        v[k] = src[h->perm[k]];
        if (!std::isfinite(v[k])) return set_err(h, BRP_ERR_INVALID, "bubble %d: non-finite value", h->perm[k]);
    }
    if (which == BRP_BUB_X || which == BRP_BUB_Y) {
        // the moved stencils join the void-fraction window
        std::vector<double> other(n);
        HIPCHK(h, hipMemcpy(other.data(), which == BRP_BUB_X ? h->b.y : h->b.x, sizeof(double) * n,
                            hipMemcpyDeviceToHost));
        for (int k = 0; k < n; k++) {
            const double px = which == BRP_BUB_X ? v[k] : other[k];
            const double py = which == BRP_BUB_X ? other[k] : v[k];
            const int ix = host_bin(px, h->g.rdx, 0.5), iy = host_bin(py, h->g.rdy, 0.5);
            window_union(h->W, std::max(ix + h->bp.mbs, 0), std::min(ix + h->bp.mbe, h->N + 1),
                         std::max(iy + h->bp.mbs, 0), std::min(iy + h->bp.mbe, h->M + 1));
        }
        if (h->cfg.tr_cylindrical && h->W.i0 <= 1 && h->W.i0 <= h->W.i1) h->W.i0 = 0;
    }
    HIPCHK(h, hipMemcpy(dst, v.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    return BRP_OK;
}

int brp_scatter_variant(brp_handle *h, int32_t mode, double *dst, int64_t ld) {
    if (!h || !dst || mode < BRP_VF_INT64 || mode > BRP_VF_FP32_ATOMIC) return BRP_ERR_INVALID;
    int rc = brp_prepare(h);
    if (rc) return rc;
    int nx, ny, i0, j0;
    field_extent(h, BRP_FIELD_FG, &nx, &ny, &i0, &j0);
    if (ld < nx) return set_err(h, BRP_ERR_INVALID, "ld %lld < %d", (long long)ld, nx);
    // The spare void-fraction buffer (zero between steps) holds the result:
    // one int64 / double per cell, or two floats (sum, compensation).
    int64_t *spare = h->fgb[h->fz];
    hipStream_t s = h->stream;
    if (mode == BRP_VF_INT64)
        launch_scatter_radii(s, h->g, h->bp, h->b, h->b.R_t, spare, h->sc_d);
    else
        launch_scatter_variant(s, h->g, h->bp, h->b, mode, spare, (int64_t)h->frame);
    if (h->cfg.tr_cylindrical && mode == BRP_VF_INT64) launch_cyl_fold(s, h->g, spare);
    HIPCHK(h, hipGetLastError());
    std::vector<double> fr(h->frame);
    if (mode == BRP_VF_INT64) {
        rc = read_frame_fg(h, spare, fr);
    } else if (mode == BRP_VF_FP64_ATOMIC) {
        rc = read_frame<double>(h, spare, fr);
    } else {
        std::vector<float> t(2 * h->frame);
        HIPCHK(h, hipMemcpyAsync(t.data(), spare, sizeof(float) * 2 * h->frame,
                                 hipMemcpyDeviceToHost, s));
        HIPCHK(h, hipStreamSynchronize(s));
        for (size_t k = 0; k < h->frame; k++) fr[k] = (double)t[k] + (double)t[h->frame + k];
    }
    HIPCHK(h, hipMemsetAsync(spare, 0, sizeof(int64_t) * h->frame, s));
    HIPCHK(h, hipStreamSynchronize(s));
    if (rc) return rc;
    if (h->cfg.tr_cylindrical && mode != BRP_VF_INT64)
        for (int j = 0; j <= h->M + 1; j++) {  // the axis fold of the other modes, on the host
            const int64_t a = frame_off(0, j, h->pitch), b = frame_off(1, j, h->pitch);
            fr[a] = fr[b] = fr[a] + fr[b];
        }
    for (int j = 0; j < ny; j++)
        for (int i = 0; i < nx; i++) dst[(size_t)j * ld + i] = fr[frame_off(i + i0, j + j0, h->pitch)];
    return BRP_OK;
}

int brp_schedule_mode(brp_handle *h, int32_t mode) {
    if (!h || mode < 0 || mode > 2) return BRP_ERR_INVALID;
    if (mode != 0 && h->cfg.rp_exact && h->cfg.bub_enable)
        return set_err(h, BRP_ERR_INVALID, "schedule: rp_exact = 1 is not supported");
    if (mode == 1) {
        auto S = std::make_shared<Schedule>();
        S->nb = h->nb;
        S->device = h->device;
        S->mismatch = (unsigned int *)S->alloc(sizeof(unsigned int));
        if (!S->mismatch || hipMemset(S->mismatch, 0, sizeof(unsigned int)) != hipSuccess)
            return set_err(h, BRP_ERR_DEVICE, "schedule: allocation failed");
        h->sched = S;
    } else if (mode == 2 && !h->sched) {
        return set_err(h, BRP_ERR_STATE, "schedule: nothing recorded to replay");
    }
    if (mode == 2 && h->sched->nb != h->nb)
        return set_err(h, BRP_ERR_STATE, "schedule: recorded with %d bubbles, handle has %d",
                       h->sched->nb, h->nb);
    h->sched_mode = mode;
    h->sched_step = 0;
    h->sched_call = 0;
    return BRP_OK;
}

int brp_schedule_copy(brp_handle *dst, const brp_handle *src) {
    if (!dst || !src || !src->sched) return BRP_ERR_INVALID;
    if (dst->device != src->device || (dst->prepared ? dst->nb : (int)dst->pending_x.size()) !=
                                          src->sched->nb)
        return set_err(dst, BRP_ERR_INVALID, "schedule_copy: needs the same device and bubble count");
    dst->sched = src->sched;
    dst->sched_step = 0;
    dst->sched_call = 0;
    return BRP_OK;
}

int brp_get_clock(const brp_handle *h, int64_t *nstep, double *time, double *carry) {
    if (!h) return BRP_ERR_INVALID;
    if (nstep) *nstep = h->nstep;
    if (time) *time = h->t;
    if (carry) *carry = h->tx;
    return BRP_OK;
}

int brp_set_clock(brp_handle *h, int64_t nstep, double time, double carry) {
    if (!h || nstep < 0 || !std::isfinite(time) || !std::isfinite(carry)) return BRP_ERR_INVALID;
    int rc = brp_prepare(h);
    if (rc) return rc;
    h->nstep = nstep;
    h->t = time;
    h->tx = carry;
    h->first_step = nstep == 0 ? 1 : 0;
    return BRP_OK;
}

int brp_get_bubbles_noalloc(brp_handle *h, int32_t which, double *dst) {
    if (!h || !dst || which < 0 || which >= BRP_BUB_COUNT) return BRP_ERR_INVALID;
    if (!h->prepared) return set_err(h, BRP_ERR_STATE, "get_bubbles_noalloc: call brp_prepare first");
    const int n = h->nb;
    if (n == 0) return BRP_OK;
    if (n > kScratchBubbles)
        return set_err(h, BRP_ERR_INVALID, "get_bubbles_noalloc: %d bubbles exceed its staging "
                       "array of %d", n, kScratchBubbles);
    if (!h->scratch_d) return set_err(h, BRP_ERR_DEVICE, "%s", h->scratch_err.c_str());
    const int c = h->rp_cur;
    const double *src = nullptr;
    const int *isrc = nullptr;
    int wi = -1;
    switch (which) {
        case BRP_BUB_X: src = h->b.x; break;
        case BRP_BUB_Y: src = h->b.y; break;
        case BRP_BUB_R_T: src = h->b.R_t; break;
        case BRP_BUB_R_P: src = h->b.R[c]; break;
        case BRP_BUB_R_N: src = h->b.Rn[c]; break;
        case BRP_BUB_D1R: src = h->b.d1R[c]; break;
        case BRP_BUB_PG: src = h->b.PG[c]; break;
        case BRP_BUB_PL_P: src = h->b.PL[h->pl_p]; break;
        case BRP_BUB_PL_N: src = h->b.PL[h->pl_n]; break;
        case BRP_BUB_PL_M: src = h->b.PL[h->pl_m]; break;
        case BRP_BUB_Q_B: src = h->b.Q_B; break;
        case BRP_BUB_N_B: src = h->b.n_B; break;
        case BRP_BUB_DT: src = h->b.bdt[c]; break;
        case BRP_BUB_RE: src = h->b.re[c]; break;
        case BRP_BUB_VX: src = h->b.vx; break;
        case BRP_BUB_VY: src = h->b.vy; break;
        case BRP_BUB_R0: src = h->b.R0; break;
        case BRP_BUB_SUB: isrc = h->b.sub; break;
        default: wi = which - BRP_BUB_IBM_X;
    }
    hipLaunchKernelGGL(k_bubbles_to_caller, dim3((n + 255) / 256), dim3(256), 0, h->stream, n,
                       h->perm_d, src, isrc, h->b.x, h->b.y, h->g.rdx, h->g.rdy, wi, h->scratch_d);
    HIPCHK(h, hipGetLastError());
    HIPCHK(h, hipMemcpyAsync(dst, h->scratch_d, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return BRP_OK;
}

int brp_save_state(brp_handle *h, int32_t slot) {
    if (!h || slot < 0 || slot >= 16) return BRP_ERR_INVALID;
    int rc = brp_prepare(h);
    if (rc) return rc;
    return snap_save(h, h->snaps[slot]);
}

int brp_restore_state(brp_handle *h, int32_t slot) {
    if (!h || slot < 0 || slot >= 16) return BRP_ERR_INVALID;
    Snapshot &sn = h->snaps[slot];
    if (!sn.valid) return set_err(h, BRP_ERR_STATE, "slot %d holds no state", slot);
    return snap_restore(h, sn);
}

int brp_adjoint(brp_handle *h, const brp_objective *obj, double *grad_amp, double *grad_phase,
                double *grad_fg0, int64_t ld, brp_adjoint_stats *stats) {
    if (!h || !obj) return BRP_ERR_INVALID;
    const brp_objective &o = *obj;
    if (o.n_steps < 1 || o.window < 1 || o.window > o.n_steps || o.checkpoint_every < 0)
        return set_err(h, BRP_ERR_INVALID, "adjoint: need 1 <= window <= n_steps");
    if (o.box_i0 < 1 || o.box_i1 > h->N || o.box_i0 > o.box_i1 || o.box_j0 < 1 ||
        o.box_j1 > h->M || o.box_j0 > o.box_j1)
        return set_err(h, BRP_ERR_INVALID, "adjoint: focal box outside the interior");
    if (grad_fg0 && ld < h->N + 2) return set_err(h, BRP_ERR_INVALID, "adjoint: ld < N + 2");
    if (h->cfg.tr_cylindrical || h->cfg.sim_order != 2)
        return set_err(h, BRP_ERR_INVALID, "adjoint: Cartesian, simulation.order = 2 only");
    // The reverse RP kernel linearises the shortened-chain sub-step; with
    // rp_exact = 1 the forward integrates the operation-by-operation form,
    // whose sub-step counts can differ, so the gradient would follow another
    // trajectory than the one run.
    if (h->cfg.rp_exact && h->cfg.bub_enable)
        return set_err(h, BRP_ERR_INVALID, "adjoint: rp_exact = 1 is not differentiated "
                       "(use rp_exact = 0)");
    if (h->sched_mode != 0)
        return set_err(h, BRP_ERR_STATE, "adjoint: set schedule mode 0 first");
    int rc = brp_prepare(h);
    if (rc) return rc;
    brp_adjoint_stats local{};
    brp_adjoint_stats *st = stats ? stats : &local;
    rc = h->prec == BRP_PREC_FP64
             ? adjoint_impl<double>(h, o, grad_amp, grad_phase, grad_fg0, ld, st)
             : adjoint_impl<float>(h, o, grad_amp, grad_phase, grad_fg0, ld, st);
    if (rc == BRP_OK && st->schedule_mismatch != 0)
        return set_err(h, BRP_ERR_STATE, "adjoint: %lld replayed RP calls took another sub-step "
                       "count than the forward (the gradient is not that of the run)",
                       (long long)st->schedule_mismatch);
    return rc;
}

int brp_set_focus(brp_handle *h, double fp_x, double fp_y, double f_dist) {
    if (!h) return BRP_ERR_INVALID;
    if (!std::isfinite(fp_x) || !std::isfinite(fp_y) || !std::isfinite(f_dist) || f_dist < 0)
        return set_err(h, BRP_ERR_INVALID, "set_focus: finite point and f_dist >= 0 required");
    h->ctrl_x = h->g.fp_x = fp_x;
    h->ctrl_y = h->g.fp_y = fp_y;
    h->ctrl_dist = h->g.fdist = f_dist;
    find_aperture(h);
    return BRP_OK;
}

int brp_get_phase_times(const brp_handle *h, double ms[5]) {
    if (!h || !ms) return BRP_ERR_INVALID;
    for (int k = 0; k < 5; k++) ms[k] = h->phase_ms[k];
    return BRP_OK;
}

}  // extern "C"
