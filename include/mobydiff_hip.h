/*
 * mobydiff_hip.h -- C ABI of the MI355X time-step library (libmobydiff_hip.so).
 *
 * This is the drop-in boundary.  Each entry point replaces one module procedure that the
 * reference driver (main.f90) calls, so a thin Fortran wrapper with bind(C) interfaces can keep
 * the reference call surface unchanged.  All functions return 0 on success and a non-zero code
 * on failure; mdx_last_error() then returns a message (the Fortran wrappers turn it into
 * `error stop`, as cufft_bindings.f90:56-74 does for cuFFT/CUDA codes).
 *
 * Arrays cross the boundary as plain host pointers in the reference's layout:
 *   ghosted field   q(0:nx+1,0:ny+1,0:nz+1)  -> q[i + (nx+2)*(j + (ny+2)*k)]   (FP64)
 *   interior field  r(1:nx,1:ny,1:nz)        -> r[(i-1) + nx*((j-1) + ny*(k-1))] (FP64)
 *   BC table bcType(1:3,0:1,0:3) / bcValue   -> t[(dir-1) + 3*side + 6*var]
 * The library owns all device memory from mdx_create to mdx_destroy (the persistent OpenMP
 * mapping of gpu_runtime.f90:22-48 in the reference).  Calls are ordered on one HIP stream; any
 * call that returns host-visible data synchronises that stream first.
 */
#ifndef MOBYDIFF_HIP_H
#define MOBYDIFF_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdx_handle mdx_handle;

enum { MDX_OK = 0, MDX_ERR_ARG = 1, MDX_ERR_HIP = 2, MDX_ERR_UNSUPPORTED = 3, MDX_ERR_STATE = 4,
       MDX_ERR_NOT_CONVERGED = 5 };

enum { MDX_SOLVER_FFT = 0,   /* FFT(x,z) + Thomas(y), pressure_solver.f90:48-63 (FFT build)      */
       MDX_SOLVER_SOLA = 1,  /* red-black SOLA, fixed niter, pressure_redblack.f90:27-254        */
       MDX_SOLVER_SOR = 2 }; /* same red-black SOR run to ||div||_inf < tol (pressure form)      */

enum { MDX_PREC_FP64 = 0, MDX_PREC_MIXED = 1 };   /* MIXED: FP32 sweeps in FP64 defect correction */

/* initial guess of the SOR solve: PREVIOUS = Phi 0 (start from the previous pressure, the plain
 * method); LEARNED = the attached network on (div u*, Phi of the same RK stage at the previous
 * step); PHI_PREV = that previous Phi alone.  LEARNED and PHI_PREV fall back to PREVIOUS whenever
 * their initial residual is not smaller. */
enum { MDX_GUESS_PREVIOUS = 0, MDX_GUESS_LEARNED = 1, MDX_GUESS_PHI_PREV = 2 };

typedef struct {
    /* grid_type (init.f90:5-12) */
    int32_t nx, ny, nz;
    double lx, ly, lz;
    double re;
    double forcing[3];
    /* boundary_type (boundary.f90:6-10); periodic as 0/1 */
    int32_t periodic[3];
    int32_t bc_type[24];
    double bc_value[24];
    /* wall-normal grid: 0 = uniform dy (reference); >0 = tanh stretching parameter beta */
    double stretch_beta;
    /* pressure solver (pressure_workspace.f90:5-24) */
    int32_t solver;          /* MDX_SOLVER_* */
    int32_t niter;           /* SOLA iterations (reference default 3) */
    double omega;            /* SOR relaxation (reference default 1.5) */
    double tol;              /* MDX_SOLVER_SOR: ||div||_inf tolerance */
    int32_t max_iter;        /* MDX_SOLVER_SOR: iteration cap (then MDX_ERR_NOT_CONVERGED unless 0) */
    int32_t check_every;     /* MDX_SOLVER_SOR: iterations between residual checks */
    int32_t precision;       /* MDX_PREC_* */
    int32_t inner_iters;     /* MDX_PREC_MIXED: FP32 iterations per defect-correction cycle */
    int32_t guess;           /* MDX_GUESS_* */
    /* constant flow rate: bulk velocity held at ub_target over a channel of height flow_height */
    int32_t flowrate;
    double ub_target, flow_height;
    /* collect per-phase GPU times (stencil / Poisson / projection) with HIP events */
    int32_t timing;
} mdx_config;

typedef struct {
    int32_t kind;            /* 0: reference bottom sine bump (ibm.f90:84-99); 1: two-sided wavy channel */
    int32_t second_order;    /* USE_IBM_SECONDORDER (CMakeLists.txt:11, default ON) */
    int32_t n_wave_x;        /* kind 0 */
    double amp_x, phase_x;   /* kind 0; amp_x <= 0 selects the reference's 5*dy (ibm.f90:47) */
    double A, lambda, half_height;   /* kind 1: walls at ly/2 +- (half_height - A cos(2 pi x / lambda)) */
} mdx_geometry;

typedef struct {
    int32_t iterations;      /* full red+black iterations (SOLA: niter) */
    int32_t converged;
    int32_t fallback;        /* learned mode: the solve started from the previous pressure (Phi = 0), the
                              * network's start being rejected (residual not below 0.8x) or skipped by
                              * the gate */
    int32_t checks;          /* residual evaluations */
    double div0_inf;         /* ||div||_inf of the initial guess */
    double div_inf;          /* ||div||_inf after the solve (from the final residual) */
    double guess_div_inf;    /* ||div||_inf of the learned guess (when used) */
    int32_t lds_sweeps;      /* 1: sweeps ran in the LDS-resident kernel, 0: streaming kernels */
    int32_t guess_skipped;   /* learned mode: the network was not evaluated (its start lost the
                                previous solves of this stage; it is re-tried after 2..16 solves) */
    int32_t lds_restarts;    /* 1: a neighbour wait of the LDS-resident sweeps gave up (another stream
                                or process held CUs); the solve was redone from Phi = 0 on the
                                streaming sweeps and the statistics describe that solve */
} mdx_solve_stats;

typedef struct {
    mdx_solve_stats stage[3];
    double ub;               /* bulk velocity after the step (flowrate on) */
} mdx_step_stats;

/* stencil, Poisson and projection times cover a sample of the stages mdx_step runs (every fourth
 * stage is event-bracketed) and `stages` counts that sample; guess_ms covers every network start */
typedef struct {
    double stencil_ms;       /* momentum + flow-rate shift + boundary conditions */
    double poisson_ms;       /* divergence RHS + initial guess + pressure solve (FFT/tridiag or SOR sweeps + checks) */
    double projection_ms;    /* velocity/pressure correction */
    double guess_ms;         /* learned initial guess (network input with the divergence RHS pass, inference,
                                prolongation; part of poisson_ms) */
    int64_t stages;
    double sweep_ms;         /* red-black sweep kernels alone, over a sample of the batches (every
                                fourth batch is event-bracketed); the fields below count the same
                                sample */
    int64_t sweep_launches;  /* kernel launches in those batches: one per colour (streaming), one
                                per red+black iteration (fused), or one per batch (LDS-resident) */
    int64_t sweep_cells;     /* colour passes over the grid in those batches times nx ny nz, i.e.
                                2 x iterations x nx ny nz (a one-colour launch covers nx ny nz of it,
                                a fused launch 2 nx ny nz, an LDS-resident batch 2 x its iterations
                                x nx ny nz) */
    double momentum_ms;      /* the momentum kernel alone (its launch inside the step, between events
                                on the handle's stream), over the `stages` sample */
} mdx_phase_times;

/* error text of the last failing call on this thread */
const char *mdx_last_error(void);

/* lifecycle: enter_field_data / init_pressure_solver (gpu_runtime.f90:22-34, pressure_solver.f90:35-46) */
int mdx_create(const mdx_config *cfg, mdx_handle **out);
/* exit_field_data / destroy_pressure_solver (gpu_runtime.f90:36-48, pressure_solver.f90:65-74) */
int mdx_destroy(mdx_handle *h);
int mdx_synchronize(mdx_handle *h);

/* `target update to/from` of the field_type arrays (io.f90:124-127,161-164); NULL skips an array */
int mdx_upload_fields(mdx_handle *h, const double *un, const double *vn, const double *wn,
                      const double *pn, const double *us, const double *vs, const double *ws,
                      const double *oldrhsu, const double *oldrhsv, const double *oldrhsw);
int mdx_download_fields(mdx_handle *h, double *un, double *vn, double *wn, double *pn,
                        double *us, double *vs, double *ws,
                        double *oldrhsu, double *oldrhsv, double *oldrhsw);

/* enter_ibm_data (ibm.f90:57-68): upload coef_u/v/w (ghosted); all NULL disables the IBM */
int mdx_set_ibm(mdx_handle *h, const double *coef_u, const double *coef_v, const double *coef_w);
/* init_ibm + set_ibm_coeff x3 + enter_ibm_data (ibm.f90:41-55,133-187; main.f90:57-61),
 * computed by the library on the host and uploaded; the arrays are also returned when non-NULL */
int mdx_compute_ibm(mdx_handle *h, const mdx_geometry *geo, double *coef_u, double *coef_v,
                    double *coef_w);
/* set_ibm_coeff(g, ibm, coeff, dix, diy, diz) (ibm.f90:133): one ghosted coefficient array for the
 * staggering (dix, diy, diz), computed on the host into coeff; nothing is uploaded */
int mdx_ibm_coefficient(mdx_handle *h, const mdx_geometry *geo, int32_t dix, int32_t diy, int32_t diz,
                        double *coeff);

/* step.f90:27 momentum(f, g, dt_alpha, dt_beta, dt_gamma, ibm, bc) */
int mdx_momentum(mdx_handle *h, double dt_alpha, double dt_beta, double dt_gamma);
/* boundary.f90:23 apply_bc(f, g) -- FFT build (n- and *-fields) */
int mdx_apply_bc(mdx_handle *h);
/* boundary.f90:173 apply_bc_redblack(f, g, bc) */
int mdx_apply_bc_redblack(mdx_handle *h);
/* pressure_solver.f90:48 pressure_projection(ps, f, g, dt_gamma, ibm, bc); stats may be NULL */
int mdx_pressure_projection(mdx_handle *h, double dt_gamma, mdx_solve_stats *stats);
/* step.f90:210 get_cfl(f, g) */
int mdx_get_cfl(mdx_handle *h, double *cfl);

/* io.f90:184-222 center_vel: cell-centred velocity uc = (u(i)+u(i+1))/2, vc, wc likewise, computed
 * on the device from the current fields and copied to the host arrays uc/vc/wc(1:nx,1:ny,1:nz)
 * (interior layout; NULL skips an array).  The VTK writer of the driver (io.f90:224-259) formats
 * them.  Allocates a device buffer of 3 nx ny nz doubles on its first call. */
int mdx_center_velocity(mdx_handle *h, double *uc, double *vc, double *wc);

/* ||div u||_inf over interior cells (FP64) */
int mdx_div_norm(mdx_handle *h, double *div_inf);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): mdx_div_norm without a
 * handle and without any allocation.  un/vn/wn are device arrays of (nz+2) x (ny+2) x (nx+2)
 * doubles in the handle's layout (one ghost layer per side, x fastest); *div_inf is a device
 * double that receives ||div u||_inf over the interior cells, bit for bit what mdx_div_norm
 * returns for the same fields, after the work queued on `stream` (NULL: the null stream) is done.
 * periodic[3] flags x, y, z as in mdx_config.  The per-workgroup partial maxima live in a
 * file-scope device array of the library (131072 doubles, 1 MiB, per GPU, held until the process
 * ends), so grids needing more than 131072 workgroups of 64 x 4 cells (about 33.5 M cells) are
 * rejected with MDX_ERR_ARG before anything is launched.  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other streams, other threads) race. */
int mdx_div_norm_noalloc(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz, const int32_t *periodic,
                         const double *un, const double *vn, const double *wn, double *div_inf, void *stream);
/* constant flow rate: bulk-velocity correction of u* for the stage with dt_gamma; returns du */
int mdx_flowrate_shift(mdx_handle *h, double dt_gamma, double *du);
/* bulk velocity of u over the channel height */
int mdx_bulk_velocity(mdx_handle *h, double *ub);
/* Poiseuille + seeded uniform perturbation of amplitude amp, projected divergence-free, p = 0.
 * geo selects the wall geometry for the profile and masks (NULL: flat channel). */
int mdx_init_fields(mdx_handle *h, uint32_t seed, double amp, const mdx_geometry *geo, double dt);
/* select the initial guess of later SOR solves (MDX_GUESS_*); the per-stage history is kept */
int mdx_set_guess(mdx_handle *h, int32_t guess);
/* one RK3 step (main.f90:66-89): 3 x {momentum, [flow rate], [apply_bc], projection, bc} */
int mdx_step(mdx_handle *h, double dt, mdx_step_stats *stats);
/* whether mdx_step reports the bulk velocity in stats->ub (flowrate on; the default 1).  The report
 * waits for the step's last kernel before mdx_step returns; with 0, stats->ub is NaN and mdx_step
 * returns once the step is queued, so the next step's kernels queue behind it without a gap.  No
 * counterpart in the reference, whose time loop computes no bulk velocity (main.f90:66-104). */
int mdx_set_step_report(mdx_handle *h, int32_t bulk_velocity);

/* one RK3 step of n independent channels, driven concurrently (one host thread and stream per
 * handle); stats may be NULL or hold n entries.  No counterpart in the reference. */
int mdx_step_batch(mdx_handle *const *hs, int32_t n, double dt, mdx_step_stats *stats);

/* per-phase GPU time accumulated since the last reset (cfg.timing must be set) */
int mdx_get_phase_times(mdx_handle *h, mdx_phase_times *t);
int mdx_reset_phase_times(mdx_handle *h);
/* device time of the momentum kernel alone: reps launches of mdx_momentum(dta, dtb, dtg) from the
 * current fields between HIP events on the handle's stream (after one untimed launch); the old
 * right-hand sides are restored afterwards, the predictor fields hold the last launch's result.
 * Allocates a temporary copy of the old right-hand sides.  No counterpart in the reference. */
int mdx_time_momentum(mdx_handle *h, double dta, double dtb, double dtg, int32_t reps, float *ms_per_launch);

/* sustained LDS read bandwidth of the current device (16-byte reads, 8 workgroups per CU), the
 * ceiling the LDS-resident sweeps are measured against; bytes per clock and CU optional.  Allocates
 * a small device buffer.  No counterpart in the reference. */
int mdx_lds_read_rate(double *bytes_per_s, double *bytes_per_clock_per_cu);

/* wall-normal coordinates of cell centres and v-faces (box coordinates, length ny+2 each) */
int mdx_get_y_coordinates(mdx_handle *h, double *yc, double *yv);

/* one line of text describing the active kernels (which red-black sweep kernel, its decomposition,
 * or why the streaming sweeps are used); truncated to len-1 characters */
int mdx_get_info(mdx_handle *h, char *buf, int32_t len);

/* ---------------------------------------------------------------- learned initial guess
 * A 3-level 3-D U-Net (16/32/64 channels, 3x3x3 convolutions, BF16 weights and activations,
 * FP32 accumulation on MFMA) maps (div u*, previous pressure) to a pressure-correction guess for
 * the red-black SOR.  No counterpart exists in the reference solver.  Parameters are FP32 masters
 * in a flat layout: per layer W(tap, cin, cout) then b(cout), layers e1a e1b e2a e2b b3a b3b d2a
 * d2b d1a d1b head. */
typedef struct mdx_net mdx_net;
int64_t mdx_net_param_count(void);
int mdx_net_create(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, mdx_net **out);
/* a network serving the nx x ny x nz solver grid that runs on that grid coarsened by `stride`
 * (1, 2 or 4) in every direction: its inputs are the means over stride^3 cells, its output is
 * prolonged trilinearly (cell-centred, the outermost cells clamped to the nearest voxel) and added
 * to Phi_0.  The network grid needs dimensions in multiples of 4 and (voxels / 64) a multiple of 64.
 * mdx_net_create is stride 1.  The explicit-input calls (forward_host, accumulate_grad_host) take
 * arrays of the network grid. */
int mdx_net_create_strided(int32_t nx, int32_t ny, int32_t nz, int32_t stride, uint64_t seed, mdx_net **out);
/* the network grid (level 0) and its stride */
int mdx_net_grid(mdx_net *net, int32_t *w, int32_t *h, int32_t *d, int32_t *stride);
/* the transfer operators between the solver grid (fine, FP64, interior layout) and the network grid
 * (coarse, FP32), run by the kernels of the solver coupling with unit scale: restriction = mean over
 * each voxel's stride^3 cells; prolongation = the trilinear interpolation the guess uses.  Allocate
 * a temporary device copy of the fine field. */
int mdx_net_restrict_host(mdx_net *net, const double *fine, float *coarse);
int mdx_net_prolong_host(mdx_net *net, const float *coarse, double *fine);
/* detaches the network from every handle it is attached to, then frees it */
int mdx_net_destroy(mdx_net *net);
/* c0: target scale (Phi = output * rms(div u*) * c0); get may pass NULL, set ignores c0 <= 0 */
int mdx_net_get_params(mdx_net *net, float *params, double *c0);
int mdx_net_set_params(mdx_net *net, const float *params, double c0);
int mdx_net_get_grads(mdx_net *net, float *grads);
int mdx_net_zero_grads(mdx_net *net);
/* explicit inputs x2[v][2] (the two input channels, already scaled), output out[v] */
int mdx_net_forward_host(mdx_net *net, const float *x2, float *out);
/* MSE against target[v]; adds this sample's gradients / batch to the accumulated gradients */
int mdx_net_accumulate_grad_host(mdx_net *net, const float *x2, const float *target, int32_t batch,
                                 float *loss);
/* Adam (lr, beta1 0.9, beta2 0.999, eps 1e-8) on the accumulated gradients, which it then zeroes */
int mdx_net_adam_step(mdx_net *net, float lr);
/* replay buffer of solver-generated (input, converged Phi) pairs, filled by attached handles */
int mdx_net_replay_reserve(mdx_net *net, int32_t capacity);
int mdx_net_replay_count(mdx_net *net, int32_t *count);
int mdx_net_train(mdx_net *net, int32_t steps, int32_t batch, float lr, uint64_t seed, float *mean_loss);
int mdx_net_eval_replay(mdx_net *net, int32_t count, float *mean_loss);
/* device time of one forward pass over replay sample 0, or the last input (HIP events around reps passes), and the
 * MFMA flops of one pass (as the convolution kernels issue them) */
int mdx_net_time_forward(mdx_net *net, int32_t reps, float *ms_per_forward, double *flops);
/* attach (or detach with net = NULL) a network to a solver handle: cfg.guess = MDX_GUESS_LEARNED
 * uses it for the initial guess; capture != 0 appends a pair per converged SOR solve.  One network
 * may be attached to several handles that are stepped one after another; mdx_step_batch rejects
 * handles that share a network. */
int mdx_attach_net(mdx_handle *h, mdx_net *net, int32_t capture);

#ifdef __cplusplus
}
#endif
#endif
