// mdx_internal.h -- device state and kernel launch interfaces of libmobydiff_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mobydiff_hip.h"

namespace mdx {

// Ghosted q(0:nx+1,0:ny+1,0:nz+1) and interior r(1:nx,1:ny,1:nz) indexing, column-major.
struct Dims {
    int nx, ny, nz;
    int sx, sy;          // nx+2, ny+2
    __host__ __device__ inline size_t g(int i, int j, int k) const {
        return (size_t)i + (size_t)sx * ((size_t)j + (size_t)sy * (size_t)k);
    }
    __host__ __device__ inline size_t in(int i, int j, int k) const {
        return (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
    }
};

// Grid scalars passed by value to kernels.
struct GridK {
    Dims d;
    double dx, dy, dz;
    double re;
    double fx, fy, fz;
    int per[3];
    const double *ihc;   // stretched-y metrics (device, ny+2) or nullptr
    const double *ihv;
};

// BC table by value (boundary.f90:6-10)
struct BcK {
    int type[24];
    double value[24];
};
__host__ __device__ inline int bci(int dir, int side, int var) { return (dir - 1) + 3 * side + 6 * var; }

// Face coefficients mu = 1/(1+dt_gamma*coef) of the red-black solver, precomputed per stage.
// plane = 1: the IBM geometry is invariant in z and only the k=0 plane is stored.
template <typename T>
struct MuK {
    const T *u, *v, *w;  // nullptr => IBM off (mu = 1)
    int plane;
    int sx, sy;
    __device__ inline size_t at(int i, int j, int k) const {
        return (size_t)i + (size_t)sx * ((size_t)j + (size_t)(plane ? 0 : (size_t)sy * (size_t)k));
    }
};

// The tables of a stage: entry q (of n) of mu, mv, mw (FP64) and fu, fv, fw (FP32) from the
// coefficients; n = 0: nothing to build.  build_mu_kernel and the flow-rate fold both use mu_entry.
struct MuBuild {
    size_t n = 0;
    double dtg = 0.0;
    const double *cu = nullptr, *cv = nullptr, *cw = nullptr;
    double *mu = nullptr, *mv = nullptr, *mw = nullptr;
    float *fu = nullptr, *fv = nullptr, *fw = nullptr;
};
__device__ inline void mu_entry(const MuBuild &b, size_t q) {
    const double a = 1.0 / (1.0 + b.dtg * b.cu[q]), c = 1.0 / (1.0 + b.dtg * b.cv[q]),
                 e = 1.0 / (1.0 + b.dtg * b.cw[q]);
    b.mu[q] = a;
    b.mv[q] = c;
    b.mw[q] = e;
    b.fu[q] = (float)a;
    b.fv[q] = (float)c;
    b.fw[q] = (float)e;
}

// XCD-aware tile order for the 3-D stencil kernels.  Workgroups are dispatched round-robin over the
// 8 XCDs (blockIdx.x % 8); XCD x is given the contiguous tile range [x*per, (x+1)*per), tiles ordered
// x fastest, then y, then z, so the neighbouring z planes a stencil reads stay in that XCD's L2.
// Launch with xcd_grid(); returns false for the padding workgroups.
__host__ inline unsigned xcd_grid(int items_x, int items_y, int nz, int bxx, int byy) {
    const long nt = (long)((items_x + bxx - 1) / bxx) * ((items_y + byy - 1) / byy) * nz;
    return (unsigned)(8 * ((nt + 7) / 8));
}
__device__ inline bool xcd_tile(int items_x, int items_y, int nz, int bxx, int byy, int &tx, int &ty, int &tz) {
    const int ntx = (items_x + bxx - 1) / bxx, nty = (items_y + byy - 1) / byy;
    const int nt = ntx * nty * nz;
    const int per = (nt + 7) / 8;
    const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
    if (t >= nt) return false;
    tx = t % ntx;
    ty = (t / ntx) % nty;
    tz = t / (ntx * nty);
    return true;
}

// per-workgroup maximum of non-negative doubles into part[workgroup]; NaN propagates
__device__ inline double nanmax_any(double a, double b) { return (a != a || a > b) ? a : b; }
__device__ inline void block_max_part(double v, double *part) {
    __shared__ double wmax[16];
    const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nt = blockDim.x * blockDim.y * blockDim.z;
    for (int off = 32; off > 0; off >>= 1) v = nanmax_any(v, __shfl_xor(v, off, 64));
    if ((t & 63) == 0) wmax[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double m = wmax[0];
        for (int w = 1; w < (nt + 63) / 64; ++w) m = nanmax_any(m, wmax[w]);
        part[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] = m;
    }
}

// decomposition of the LDS-resident red-black solver (mdx_sor_lds.hip)
struct RbPlanOut {
    int px, py, pz, bx, by, bz, maxp;
    size_t lds;
    int total;   // halo values exchanged per workgroup and iteration
};

}  // namespace mdx

struct mdx_handle {
    mdx_config cfg;
    mdx::Dims d;
    double dx, dy, dz;
    size_t ng, ni;                 // ghosted / interior element counts
    hipStream_t stream = nullptr;
    int device = 0;

    // field_type (init.f90:14-19), device
    double *un = nullptr, *vn = nullptr, *wn = nullptr, *pn = nullptr;
    double *us = nullptr, *vs = nullptr, *ws = nullptr;
    double *oru = nullptr, *orv = nullptr, *orw = nullptr;
    // ibm_type coefficients (ibm.f90:27-34), device, ghosted; nullptr when IBM off
    double *cu = nullptr, *cv = nullptr, *cw = nullptr;
    bool ibm_on = false;
    bool coef_plane = false;       // coefficients invariant in z
    std::vector<double> h_cu, h_cv, h_cw;   // host copies (geometry masks for the initial field)

    // wall-normal metrics
    bool stretched = false;
    double *d_ihc = nullptr, *d_ihv = nullptr, *d_yc = nullptr;
    std::vector<double> yc, yv, ihc, ihv;   // host, length ny+2

    // red-black SOR in pressure-correction form
    double *phi = nullptr, *phi2 = nullptr, *rhs = nullptr;   // FP64 interior
    float *e32 = nullptr, *r32 = nullptr;                     // FP32 interior
    float *e32b = nullptr;                                    // second FP32 iterate (fused iterations)
    double *mu64[3] = {nullptr, nullptr, nullptr};
    float *mu32[3] = {nullptr, nullptr, nullptr};
    double mu_dtg = -1.0;          // dt_gamma the mu tables were built for

    // FFT path
    double *fft_rhs = nullptr, *hat = nullptr, *cp = nullptr, *den_inv = nullptr, *arow = nullptr;
    double *tw_x = nullptr, *tw_z = nullptr;
    // mixed-precision FFT solve: FP64 rhs in fft_rhs, FP64 phi and residual, FP32 spectrum and tables
    double *fft_phi = nullptr, *fft_res = nullptr, *fft_mean = nullptr;   // fft_mean: 16 ny partial plane sums + ny mean phi
    float *hat32 = nullptr, *cp32 = nullptr, *den32 = nullptr, *arow32 = nullptr, *tw_x32 = nullptr, *tw_z32 = nullptr;

    // reductions
    double *red_part = nullptr;    // per-block partial results
    unsigned long long *red_max = nullptr;   // [0] max result (double bits), [1] LDS-solver error word
    double *max_part = nullptr;    // per-workgroup maxima (first pass of the max reductions)
    double *h_pinned = nullptr;    // pinned host staging (64 doubles)
    double *d_pinned = nullptr;    // the same memory as the device addresses it
    int red_blocks = 0;

    // timing
    bool timing = false;
    hipEvent_t ev[10] = {};         // [8], [9]: the learned guess
    // per-stage phase events (0 stage start, 1 Poisson start, 2 projection start, 3 projection end,
    // 4 stage end, 5 momentum kernel end) in a ring of PHQ stages, added to `times` once complete
    // (no stall per stage); set PHQ is scratch for phase events recorded outside mdx_step
    static constexpr int PHQ = 8, PH_EVERY = 4, PH_EVENTS = 6;
    long long ph_seen = 0;   // stages stepped so far
    bool ph_skip = false;    // this stage is not timed
    hipEvent_t ev_ph[PHQ + 1][PH_EVENTS] = {};
    int ph_head = 0, ph_count = 0, ph_cur = PHQ;
    mdx_phase_times times{};
    // sweep batches not yet added to the counters: start/end events, iterations and launches
    // (a speculative batch can be in flight while the previous one is accounted).  Every
    // SW_EVERY-th batch is timed: the two events around a batch hold the stream for several
    // microseconds, so timing every batch would slow the solve it measures.
    static constexpr int SW_Q = 4, SW_EVERY = 4;
    long long sw_seen = 0;   // batches so far
    int sw_last = -1;        // accounting entry of the last batch (-1: not timed)
    hipEvent_t ev_sw[SW_Q][2] = {};
    int sw_its[SW_Q] = {}, sw_launches[SW_Q] = {};
    int sw_head = 0, sw_count = 0;
    bool sw_void[SW_Q] = {};   // a batch that found its input converged and did nothing
    // sequence number of the last residual check posted to a pinned slot: a check writes
    // [max, error word] and then its number, which the host polls for
    unsigned long long chk_seq = 0;

    // captured sweep batches: [0] FP64, [1] FP32
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int m = 0;
        bool ibm = false;
        const void *iterate = nullptr;   // the iterate buffer the batch was captured on
    } graphs[2];
    bool use_graphs = true;
    // momentum through the LDS-tiled kernel (MOBYDIFF_MOMENTUM_DIRECT=1: one thread per cell)
    bool tiled_momentum = true;
    // in a step, the tiled momentum kernel forms the flow-rate partial sums (fr_part, 2 doubles per
    // tile) and one launch folds them and shifts u* (MOBYDIFF_FUSED_FLOWRATE=0: the separate sum pass)
    bool fused_flowrate = true;
    double *fr_part = nullptr;
    // inside a step (step_bc, set around its projection, whose fill follows at once) the SOR
    // projection kernel writes the channel's boundary conditions itself (apply_writes_bc;
    // MOBYDIFF_FUSED_BC=0: the separate fills); apply_wrote_bc: the last projection did.  The
    // projection called on its own leaves the ghost shell as the reference's projection does.
    bool fused_bc = true;
    bool step_bc = false;
    // in a step with constant flow rate the fold-and-shift launch also builds the stage's face
    // coefficient tables (MOBYDIFF_FUSED_MU=0: build_mu_kernel before the solve)
    bool fused_mu = true;
    bool apply_wrote_bc = false;
    // streaming red-black sweeps: z planes marched per thread (0: one cell per thread)
    int sweep_kz = 4;
    // one launch per red+black iteration (halo-recomputing tiles) instead of one per colour
    bool sweep_fused = false;

    // LDS-resident red-black solver (MOBYDIFF_LDS_SOR): 0 off, 1 FP32 sweep batches (the default),
    // 2 FP32 and FP64 batches; plans per precision and the tagged face buffers
    int lds_sor_mode = 1;
    // launch the LDS-resident kernel through hipLaunchCooperativeKernel (MOBYDIFF_LDS_COOP=1)
    // instead of an ordinary launch of a grid that the plan already sized to be co-resident
    bool rb_coop = false;
    // thread layout of the LDS-resident kernel (MOBYDIFF_LDS_FORM): 0 half a z-column per thread,
    // in workgroups of 1024 threads, where the subdomain and its face ring fit; 1 x-pairs anywhere
    // in the tile; 2 one z-column per thread
    int rb_form = 0;
    // half-column form: the interior columns relax the halo-free part of each colour while the
    // exchange is in flight (MOBYDIFF_LDS_OVERLAP=0 keeps the plain colour-by-colour order)
    bool rb_overlap = true;
    // relaxation factor of the FP32 inner sweeps of the mixed solve (the configured omega unless
    // MOBYDIFF_INNER_OMEGA sets another); the FP64 sweeps always use the configured one
    double inner_omega = 1.0;
    // the learned start's input pass fused with the right-hand side (MOBYDIFF_FUSED_INPUT=0: the
    // separate right-hand side, residual and restriction kernels)
    bool fused_input = true;
    // mdx_step reports the bulk velocity (a wait for the step's end); mdx_set_step_report
    bool step_ub = true;
    // z-planes each workgroup of the tiled momentum kernel marches (momentum_default_planes, 8 or
    // 12; MOBYDIFF_MOM_KZ)
    int mom_kz = 8;
    int rb_state = 0;              // 0 not planned, 1 ready, -1 unavailable
    // polls of a neighbour wait before it gives up (MOBYDIFF_LDS_SPIN_LIMIT; a test hook); a wait
    // that gives up sets the error word, the batch drains, and the solve restarts on the
    // streaming sweeps (lds_suspended for the rest of that solve; after LDS_TIMEOUT_MAX such
    // restarts the LDS-resident solver stays off)
    int rb_spin_limit = 1 << 20;
    int lds_timeouts = 0;
    bool lds_suspended = false;
    mdx::RbPlanOut rb_plan64{}, rb_plan32{};
    void *rb_faces = nullptr;      // tagged face words of the LDS-resident sweeps
    int *rb_xl32 = nullptr, *rb_xl64 = nullptr;   // exchange lists of the two plans
    size_t rb_faces_bytes = 0;
    int *rb_err = nullptr;         // points into red_max[1]
    unsigned long long rb_tag = 0; // half-sweeps published so far (the tag of the last one)
    std::string rb_note;           // why the streaming sweeps are in use, or the chosen plan

    // learned guess (attached network) and training-pair capture
    struct mdx_net *net = nullptr;
    bool net_capture = false;
    // converged Phi of each RK stage at the previous step (initial-guess history), its dt_gamma
    // (0: no history yet) and the count of SOR projections (stage = count mod 3)
    double *phi_hist = nullptr;
    double *res64 = nullptr;       // residual field of the previous-Phi guess (network input)
    double *cvel = nullptr;        // cell-centred velocity for output (3 interior arrays, allocated on first use)
    // one-wave-per-line tridiagonal solve: modified wavenumbers (nxh, nz) and row coefficients (ny)
    double *lamx = nullptr, *lamz = nullptr, *tw_a = nullptr, *bdrow = nullptr, *crow = nullptr;
    float *lamx32 = nullptr, *lamz32 = nullptr, *tw_a32 = nullptr, *bdrow32 = nullptr, *crow32 = nullptr;
    bool wave_tridiag = true;      // MOBYDIFF_WAVE_TRIDIAG=0: the thread-per-line Thomas recurrence
    bool part_tridiag = true;      // MOBYDIFF_PART_TRIDIAG=0: the wave-per-line PCR solve instead of the
                                   // partitioned Thomas one (64 <= ny <= 256)
    double hist_dtg[3] = {0.0, 0.0, 0.0};
    // learned-guess gate per stage: consecutive solves the network's start lost, solves left to
    // skip, whether the current solve starts from the network's guess, and the sweeps of the last
    // solve that did not (the yardstick of a network start)
    int net_losses[3] = {0, 0, 0}, net_skip[3] = {0, 0, 0};
    // the network's start is taken when its residual max-norm is below net_margin times the other
    // start's (MOBYDIFF_NET_MARGIN)
    double net_margin = 0.8;
    // the time gate of the learned start (MOBYDIFF_NET_TIME_GATE=0 switches it off: the network is
    // then evaluated for every solve and judged by the residual test alone)
    bool net_time_gate = true;
    // the learned mode's start when the network's is not taken: Phi = 0, the previous pressure
    // (false, the default), or Phi_0, the previous-step guess (MOBYDIFF_NET_FALLBACK=prev).  Warm
    // starts carried from step to step gather slowly converging error: from Phi_0 the solves of the
    // 192x128x96 channel went from 40 to 170 iterations within 20 steps, while the plain start
    // holds its count; plain solves between network starts keep the network's input fresh.
    bool net_fallback_prev = false;
    // mixed solve: at most this many inner_iters batches run as one while the checks between them
    // are predicted to fail (MOBYDIFF_MERGE_MAX; 1 keeps every check)
    int merge_max = 8;
    // iterations of the last converged mixed solve of each RK stage and start kind (the merge
    // prior; index stage + 3 * (start from Phi_0); 0: none)
    int mix_prev_its[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // [stage + 3 * start kind (0 plain, 1 Phi_0, 2 network)]
    bool net_started[3] = {false, false, false};
    // the previous solve of each stage was network-started: the next learned start queues its first
    // batch before the device's choice has been read (MOBYDIFF_DEFER_CHOICE=0: it always waits)
    bool net_prev_started[3] = {false, false, false};
    bool defer_choice = true;
    double base_sec[3] = {0.0, 0.0, 0.0};   // host time of the last solve of each stage not started by the network
    long proj_count = 0;
};

// kernel launchers (defined in the .hip files)
namespace mdx {
GridK grid_k(const mdx_handle *h);
BcK bc_k(const mdx_handle *h);

// frpart: the tiled kernel also writes the flow-rate partial sums of its momentum_tiles(h) tiles
hipError_t launch_momentum(mdx_handle *h, double dta, double dtb, double dtg, double *frpart = nullptr);
int momentum_tiles(const mdx_handle *h);
int momentum_default_planes(const Dims &d);
// array sets of the FFT build's apply_bc: the current fields and the predictor's
enum { BC_N = 1, BC_S = 2, BC_ALL = 3 };
hipError_t launch_apply_bc_fft(mdx_handle *h, int fields = BC_ALL);
hipError_t launch_apply_bc_redblack(mdx_handle *h);
hipError_t launch_div_max(mdx_handle *h, const double *u, const double *v, const double *w);
hipError_t launch_cfl_max(mdx_handle *h);
hipError_t launch_center_vel(mdx_handle *h, double *uc, double *vc, double *wc);
hipError_t launch_bulk_sum(mdx_handle *h, const double *u, double dtg, double *out_dev);
// constant flow rate: the bulk sums of u*, and the shift u* += du * mu_u (every workgroup of the
// second launch folds the sums itself); ub, mean mu_u, du to out_dev
hipError_t launch_flowrate(mdx_handle *h, double dtg, double *out_dev);
// fold nparts (sum u* h_j, sum mu_u h_j) pairs from part, then u* += du mu_u; ub, mean mu_u, du to out_dev
// mb: also build those tables (mu_build), once the fold is done
hipError_t launch_flow_fold(mdx_handle *h, double dtg, const double *part, int nparts, double *out_dev,
                            const MuBuild &mb = MuBuild{});
hipError_t launch_initial_field(mdx_handle *h, uint32_t seed, double amp, int use_geo,
                                const double *hx_u, double ymid);
hipError_t launch_reset_max(mdx_handle *h);
// max_part[0..n) -> red_max[0]; with host_slot (device address of pinned words) also the max and
// the LDS-solver error word to host_slot[0..1], then `seq` to host_slot[2]
hipError_t launch_max_final(mdx_handle *h, unsigned nparts, double *host_slot = nullptr,
                            unsigned long long seq = 0);
static constexpr unsigned MAX_PARTS = 1u << 17;

// red-black solver
hipError_t launch_sola_sweep(mdx_handle *h, int parity, double sor, double dtg);
hipError_t launch_build_mu(mdx_handle *h, double dtg);
// the stage's tables for launch_build_mu's work (n = 0 without IBM)
MuBuild mu_build(const mdx_handle *h, double dtg);
hipError_t launch_sor_prepare(mdx_handle *h, double *rhs_max_slot = nullptr, unsigned long long seq = 0);
// sor_prepare + Phi = fac * prev (prev null: 0) + the first defect check of a mixed solve, fused
hipError_t launch_sor_start(mdx_handle *h, const double *prev, double fac, double *host_slot,
                            unsigned long long seq);
hipError_t launch_sor_sweep64(mdx_handle *h, int color, double omega);
hipError_t launch_sor_residual64(mdx_handle *h, const double *phi);
// the projection; with hist != nullptr Phi is also copied there (the stage's history)
// (inside a step: sets h->apply_wrote_bc, whether the launch also wrote the boundary conditions)
hipError_t launch_sor_apply(mdx_handle *h, const double *phi, double dtg, double *hist = nullptr);
bool apply_writes_bc(const mdx_handle *h);
// parts_only != nullptr: the workgroup maxima stay in max_part (their count to *parts_only) and
// no fold follows (the caller's next kernel folds them)
hipError_t launch_defect64(mdx_handle *h, const double *phi_in, const float *e, double *phi_out, float *r32,
                           double *host_slot = nullptr, unsigned long long seq = 0, unsigned *parts_only = nullptr);
hipError_t launch_sor_sweep32(mdx_handle *h, int color, float omega);
hipError_t launch_sor_iters64(mdx_handle *h, int m, double omega);
hipError_t launch_sor_iters32(mdx_handle *h, int m, float omega);
hipError_t launch_zero32(mdx_handle *h, float *p, size_t n);
hipError_t launch_zero64(mdx_handle *h, double *p, size_t n);
hipError_t launch_scale_copy64(mdx_handle *h, double *dst, const double *src, double fac, size_t n);
hipError_t launch_residual_field64(mdx_handle *h, const double *phi, double *field, double *host_slot = nullptr);
hipError_t launch_phi0_residual(mdx_handle *h, const double *prev, double fac, double *phi0, double *res,
                                double *host_slot, double *stat);
// the learned start's input pass fused with the right-hand side (learned_input_kernel): rhs, r0 and
// Phi_0 (into phi0 / res when given), voxel sums at the network's stride into vsum, and
// 4 * nparts workgroup partials in max_part
bool learned_input_fits(const mdx_handle *h, int stride);
hipError_t launch_learned_input(mdx_handle *h, const double *prev, double fac, double *phi0, double *res,
                                double *vsum, int stride, unsigned *nparts);
// the learned start's choice (select_start_kernel): rg folded from the nb maxima of its first
// check in max_part, posted to rg_slot
hipError_t launch_select_start(mdx_handle *h, double margin, unsigned nb, double *rg_slot, const double *other_slot,
                               const double *phi0, const double *res0, double *host_slot, unsigned long long seq);
bool rb_plan(const mdx_handle *h, bool fp32, RbPlanOut &out, std::string &why);
// defect epilogue of an FP32 LDS batch: phi_out = phi_in + e, r32 = defect, max_part = its maxima
struct RbDefect {
    const double *phi_in;
    double *phi_out;
    double *host_slot;   // device address of the pinned words receiving [max, error word, seq]
    unsigned long long seq;
    // skip: the batch does nothing when the previous check (red_max[0]) is not >= skip_tol, the
    // solve's own stopping test, so a batch queued before that check was read costs one launch
    int skip;
    double skip_tol;
};
hipError_t launch_rb_lds(mdx_handle *h, bool fp32, int iters, int zero_init, const RbDefect *df = nullptr);
size_t rb_faces_bytes(const RbPlanOut &p, bool fp32);
size_t rb_xlist_bytes(const RbPlanOut &p);
hipError_t rb_build_xlists(mdx_handle *h, const RbPlanOut &p, int *xl);

// FFT path
hipError_t launch_fft_forward(mdx_handle *h, double dtg);
hipError_t launch_tridiag(mdx_handle *h);
hipError_t launch_fft_backward_correct(mdx_handle *h, double dtg);
hipError_t launch_fft_mixed_setup(mdx_handle *h, double dtg);
hipError_t launch_fft_residual(mdx_handle *h);
hipError_t launch_fft32_solve_add(mdx_handle *h);
hipError_t launch_fft_pin(mdx_handle *h);
hipError_t launch_fft_correct(mdx_handle *h, const double *phi, double dtg);
}  // namespace mdx

// error helpers
void mdx_set_error(const std::string &msg);
#define MDX_HIP(call)                                                                    \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            mdx_set_error(std::string(#call) + ": " + hipGetErrorString(e_));            \
            return MDX_ERR_HIP;                                                          \
        }                                                                                \
    } while (0)
