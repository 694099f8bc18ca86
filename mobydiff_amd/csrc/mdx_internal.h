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

struct Timer {
    hipEvent_t ev[8];
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
    double *mu64[3] = {nullptr, nullptr, nullptr};
    float *mu32[3] = {nullptr, nullptr, nullptr};
    double mu_dtg = -1.0;          // dt_gamma the mu tables were built for

    // FFT path
    double *fft_rhs = nullptr, *hat = nullptr, *cp = nullptr, *den_inv = nullptr, *arow = nullptr;
    double *tw_x = nullptr, *tw_z = nullptr;

    // reductions
    double *red_part = nullptr;    // per-block partial results
    unsigned long long *red_max = nullptr;   // atomic max of non-negative doubles (bit patterns)
    double *h_pinned = nullptr;    // pinned host staging (8 doubles)
    int red_blocks = 0;

    // timing
    bool timing = false;
    hipEvent_t ev[8];
    mdx_phase_times times{};
    int sweep_pending = 0;         // iterations of the last (unaccounted) sweep batch

    // captured sweep batches: [0] FP64, [1] FP32
    struct GraphSlot {
        hipGraphExec_t exec = nullptr;
        int m = 0;
        bool ibm = false;
    } graphs[2];
    bool use_graphs = true;

    // learned guess (attached network), opaque
    void *net = nullptr;
};

// kernel launchers (defined in the .hip files)
namespace mdx {
GridK grid_k(const mdx_handle *h);
BcK bc_k(const mdx_handle *h);

hipError_t launch_momentum(mdx_handle *h, double dta, double dtb, double dtg);
hipError_t launch_apply_bc_fft(mdx_handle *h);
hipError_t launch_apply_bc_redblack(mdx_handle *h);
hipError_t launch_div_max(mdx_handle *h, const double *u, const double *v, const double *w);
hipError_t launch_cfl_max(mdx_handle *h);
hipError_t launch_bulk_sum(mdx_handle *h, const double *u, double dtg, double *out_dev);
hipError_t launch_flow_shift(mdx_handle *h, double dtg, const double *du_dev);
hipError_t launch_initial_field(mdx_handle *h, uint32_t seed, double amp, int use_geo,
                                const double *hx_u, double ymid);
hipError_t launch_reset_max(mdx_handle *h);

// red-black solver
hipError_t launch_sola_sweep(mdx_handle *h, int parity, double sor, double dtg);
hipError_t launch_build_mu(mdx_handle *h, double dtg);
hipError_t launch_sor_prepare(mdx_handle *h);
hipError_t launch_sor_sweep64(mdx_handle *h, int color, double omega);
hipError_t launch_sor_residual64(mdx_handle *h, const double *phi);
hipError_t launch_sor_apply(mdx_handle *h, const double *phi, double dtg);
hipError_t launch_defect64(mdx_handle *h, const double *phi_in, const float *e, double *phi_out, float *r32);
hipError_t launch_sor_sweep32(mdx_handle *h, int color, float omega);
hipError_t launch_zero32(mdx_handle *h, float *p, size_t n);
hipError_t launch_zero64(mdx_handle *h, double *p, size_t n);

// FFT path
hipError_t launch_fft_forward(mdx_handle *h, double dtg);
hipError_t launch_tridiag(mdx_handle *h);
hipError_t launch_fft_backward_correct(mdx_handle *h, double dtg);
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
