// mdx_unet.h -- state of the pressure-guess U-Net and its launch interface.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

typedef __bf16 bf16;

struct mdx_net {
    int W[3], H[3], D[3];          // per level of the network grid: level 0, /2, /4
    long V[3];
    // the solver grid the network serves: level 0 is that grid coarsened by `stride` in every
    // direction (inputs restricted by the mean over stride^3 cells, the output prolonged
    // trilinearly); stride 1 runs the network on the solver grid itself
    int stride = 1;
    int FW = 0, FH = 0, FD = 0;
    long FV = 0;
    int64_t nparam = 0;
    int64_t woff[11], boff[11];    // flat FP32 parameter layout: per layer W(tap,cin,cout) then b(cout)
    hipStream_t stream = nullptr;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    bf16 *wp[11] = {}, *wd[11] = {};   // forward / input-gradient weights in MFMA fragment order
    bf16 *x_in = nullptr, *a[10] = {}, *p1 = nullptr, *p2 = nullptr, *cat1 = nullptr, *cat2 = nullptr;
    bf16 *g[10] = {}, *gp1 = nullptr, *gp2 = nullptr, *gcat1 = nullptr, *gcat2 = nullptr;
    float *out = nullptr, *dout = nullptr, *target = nullptr, *loss = nullptr, *x2_stage = nullptr;
    double *stats = nullptr, *part = nullptr, *h_pinned = nullptr;
    double *vsum = nullptr;        // per voxel: the sums of r0 and of Phi_0 over its cells (fused input pass)
    std::vector<void *> allocs;
    double c0 = 0.0;               // target scale: Phi = out * rms(div u*) * c0
    int adam_t = 0;
    // replay buffer of captured (input, target) pairs
    int cap = 0, count = 0, next = 0;
    bf16 *rb_x = nullptr;          // [cap][V][2]
    float *rb_t = nullptr;         // [cap][V]
    uint64_t sample_state = 0;
    // inference timing
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool tiled_conv = true;        // LDS-tiled convolutions where the grid allows (bit-identical)
    bool split_conv = true;        // split-K convolutions on small grids the tiled kernel cannot take
    bool fuse_cat = true;          // inference: decoder convolutions read the up-sampling and skip in place
    bool fuse_pool = true;         // inference: the first pooling in the epilogue of the convolution before it
    std::vector<struct mdx_handle *> users;   // solver handles this network is attached to
};

namespace mdx {
namespace unet {

struct LayerDesc {
    const char *name;
    int cin, cout, ks, level;   // level 0 = full resolution
};
static const LayerDesc LAYERS[] = {
    {"e1a", 16, 16, 3, 0}, {"e1b", 16, 16, 3, 0}, {"e2a", 16, 32, 3, 1}, {"e2b", 32, 32, 3, 1},
    {"b3a", 32, 64, 3, 2}, {"b3b", 64, 64, 3, 2}, {"d2a", 96, 32, 3, 1}, {"d2b", 32, 32, 3, 1},
    {"d1a", 48, 16, 3, 0}, {"d1b", 16, 16, 3, 0}, {"head", 16, 1, 1, 0},
};
static constexpr int NL = 11;

hipError_t repack_all(mdx_net *n);
hipError_t upload_input2(mdx_net *n, const float *x2);
// infer: no backward pass follows (decoder concats are not materialised)
hipError_t net_forward(mdx_net *n, bool infer = false);
hipError_t net_backward(mdx_net *n, float scale);
hipError_t net_adam(mdx_net *n, float lr);
// solver coupling.  The solve starts from Phi_0 = fac * Phi_prev (the converged correction of the
// same RK stage at the previous step, fac = dt_gamma / dt_gamma_prev; Phi_0 = 0 without history).
// Inputs: ch0 = r0 / s_r with r0 = rhs - L_mu(Phi_0) computed in FP64 by the solver, ch1 = Phi_0 /
// s_p.  The network predicts the remaining correction: Phi = Phi_0 + out * s_r * c0, so the
// training target is (Phi - Phi_0) / (s_r c0).
hipError_t net_input_from_solver(mdx_net *n, const mdx_handle *h, const double *r0, const double *phi0,
                                 bool stats_ready = false);
hipError_t net_guess_to_phi(mdx_net *n, const double *phi0, double *phi);
// The input from the fused solver pass (launch_learned_input): stats[0..1] = rms(r0), rms(Phi_0)
// from the nb workgroup partials in part (sums of r0^2 at 2 nb, of Phi_0^2 at 3 nb), the network
// input from n->vsum, and the maxima |rhs| (part[0, nb)) and |r0| (part[nb, 2 nb)) into the pinned
// slots rz_slot and rp_slot (value, the error word err[0], then 0 as the slot's sequence number).
hipError_t net_input_from_sums(mdx_net *n, const double *part, unsigned nb, double *rz_slot, double *rp_slot,
                               const unsigned long long *err);
// the guess Phi = fac * prev + out * s_r * c0 (Phi_0 formed from the history, as the input pass does)
hipError_t net_guess_to_phi_scaled(mdx_net *n, const double *prev, double fac, double *phi);
hipError_t net_capture(mdx_net *n, const double *phi, const double *phi_prev, double fac);
hipError_t net_load_replay(mdx_net *n, int slot);
hipError_t net_restrict(mdx_net *n, const double *fine, float *coarse);
hipError_t net_prolong(mdx_net *n, const float *coarse, double *fine);
hipError_t net_mse_only(mdx_net *n);
// drop handle h from the network's attached handles
void net_detach(mdx_net *n, struct mdx_handle *h);
hipError_t net_rms_to_stats3(mdx_net *n, const mdx_handle *h, const double *phi, const double *phi_prev,
                             double fac);
}  // namespace unet
}  // namespace mdx
