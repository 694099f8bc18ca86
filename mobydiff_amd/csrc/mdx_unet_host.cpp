// mdx_unet_host.cpp -- network state, training loop and solver coupling for the pressure-guess
// U-Net (kernels in mdx_unet.hip).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "mdx_internal.h"
#include "mdx_unet.h"

using namespace mdx;
using namespace mdx::unet;

#define NET_CHECK(cond, msg)             \
    do {                                 \
        if (!(cond)) {                   \
            mdx_set_error(msg);          \
            return MDX_ERR_ARG;          \
        }                                \
    } while (0)

static int64_t layer_wcount(int l) { return (int64_t)LAYERS[l].ks * LAYERS[l].ks * LAYERS[l].ks * LAYERS[l].cin * LAYERS[l].cout; }

extern "C" int64_t mdx_net_param_count(void) {
    int64_t n = 0;
    for (int l = 0; l < NL; ++l) n += layer_wcount(l) + LAYERS[l].cout;
    return n;
}

static void free_net(mdx_net *n) {
    for (void *p : n->allocs) (void)hipFree(p);
    n->allocs.clear();
    if (n->h_pinned) (void)hipHostFree(n->h_pinned);
    if (n->stream) (void)hipStreamDestroy(n->stream);
}

template <typename T>
static hipError_t nalloc(mdx_net *n, T **p, size_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n->allocs.push_back((void *)*p);
    return e;
}

extern "C" int mdx_net_create_strided(int32_t nx, int32_t ny, int32_t nz, int32_t stride, uint64_t seed,
                                      mdx_net **out) {
    NET_CHECK(out, "mdx_net_create: null output");
    *out = nullptr;
    NET_CHECK(stride == 1 || stride == 2 || stride == 4, "mdx_net_create: the stride must be 1, 2 or 4");
    NET_CHECK(nx > 0 && ny > 0 && nz > 0 && nx % stride == 0 && ny % stride == 0 && nz % stride == 0,
              "mdx_net_create: the grid must divide by the stride");
    const int w0 = nx / stride, h0 = ny / stride, d0 = nz / stride;
    // two 2x2x2 poolings, and every level's voxel count a multiple of 64 (one conv workgroup = 64
    // voxels; the coarsest level's count divides the others')
    NET_CHECK(w0 % 4 == 0 && h0 % 4 == 0 && d0 % 4 == 0 && ((long)(w0 / 4) * (h0 / 4) * (d0 / 4)) % 64 == 0,
              "mdx_net_create: needs the network grid (solver grid / stride) in multiples of 4 with (its voxel "
              "count / 64) a multiple of 64");
    mdx_net *n = new mdx_net();
    if (const char *t = std::getenv("MOBYDIFF_NO_TILED_CONV")) n->tiled_conv = !(t[0] == '1');
    if (const char *t = std::getenv("MOBYDIFF_NO_SPLIT_CONV")) n->split_conv = !(t[0] == '1');
    if (const char *t = std::getenv("MOBYDIFF_NO_CAT_FUSION")) n->fuse_cat = !(t[0] == '1');
    if (const char *t = std::getenv("MOBYDIFF_NO_POOL_FUSION")) n->fuse_pool = !(t[0] == '1');
    n->stride = stride;
    n->FW = nx; n->FH = ny; n->FD = nz;
    n->FV = (long)nx * ny * nz;
    n->W[0] = w0; n->H[0] = h0; n->D[0] = d0;
    for (int q = 1; q < 3; ++q) { n->W[q] = n->W[q - 1] / 2; n->H[q] = n->H[q - 1] / 2; n->D[q] = n->D[q - 1] / 2; }
    for (int q = 0; q < 3; ++q) n->V[q] = (long)n->W[q] * n->H[q] * n->D[q];
    n->nparam = mdx_net_param_count();
    int64_t off = 0;
    for (int l = 0; l < NL; ++l) {
        n->woff[l] = off;
        off += layer_wcount(l);
        n->boff[l] = off;
        off += LAYERS[l].cout;
    }
    auto fail = [&](hipError_t e) {
        mdx_set_error(std::string("mdx_net_create: ") + hipGetErrorString(e));
        free_net(n);
        delete n;
        return MDX_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e);
    if ((e = nalloc(n, &n->params, n->nparam)) != hipSuccess || (e = nalloc(n, &n->grads, n->nparam)) != hipSuccess ||
        (e = nalloc(n, &n->adam_m, n->nparam)) != hipSuccess || (e = nalloc(n, &n->adam_v, n->nparam)) != hipSuccess)
        return fail(e);
    for (int l = 0; l < NL; ++l) {
        const LayerDesc &L = LAYERS[l];
        if (L.ks != 3) continue;
        const int K = 27 * L.cin, KC = (K + 31) / 32, Kd = 27 * L.cout, KCd = (Kd + 31) / 32;
        if ((e = nalloc(n, &n->wp[l], (size_t)KC * (L.cout / 16) * 512)) != hipSuccess) return fail(e);
        if (L.cin >= 16 && (e = nalloc(n, &n->wd[l], (size_t)KCd * (L.cin / 16) * 512)) != hipSuccess) return fail(e);
    }
    // activations and their gradients (bf16, channels-last), at the network grid
    const long V0 = n->V[0], V1 = n->V[1], V2 = n->V[2];
    struct { bf16 **p; size_t n; } bufs[] = {
        {&n->x_in, (size_t)V0 * 16}, {&n->a[0], (size_t)V0 * 16}, {&n->a[1], (size_t)V0 * 16},
        {&n->p1, (size_t)V1 * 16}, {&n->a[2], (size_t)V1 * 32}, {&n->a[3], (size_t)V1 * 32},
        {&n->p2, (size_t)V2 * 32}, {&n->a[4], (size_t)V2 * 64}, {&n->a[5], (size_t)V2 * 64},
        {&n->cat2, (size_t)V1 * 96}, {&n->a[6], (size_t)V1 * 32}, {&n->a[7], (size_t)V1 * 32},
        {&n->cat1, (size_t)V0 * 48}, {&n->a[8], (size_t)V0 * 16}, {&n->a[9], (size_t)V0 * 16},
        {&n->g[0], (size_t)V0 * 16}, {&n->g[1], (size_t)V0 * 16}, {&n->gp1, (size_t)V1 * 16},
        {&n->g[2], (size_t)V1 * 32}, {&n->g[3], (size_t)V1 * 32}, {&n->gp2, (size_t)V2 * 32},
        {&n->g[4], (size_t)V2 * 64}, {&n->g[5], (size_t)V2 * 64}, {&n->gcat2, (size_t)V1 * 96},
        {&n->g[6], (size_t)V1 * 32}, {&n->g[7], (size_t)V1 * 32}, {&n->gcat1, (size_t)V0 * 48},
        {&n->g[8], (size_t)V0 * 16}, {&n->g[9], (size_t)V0 * 16},
    };
    for (auto &b : bufs)
        if ((e = nalloc(n, b.p, b.n)) != hipSuccess) return fail(e);
    if ((e = nalloc(n, &n->out, V0)) != hipSuccess || (e = nalloc(n, &n->dout, V0)) != hipSuccess ||
        (e = nalloc(n, &n->x2_stage, 2 * V0)) != hipSuccess ||
        (e = nalloc(n, &n->target, V0)) != hipSuccess || (e = nalloc(n, &n->loss, 4)) != hipSuccess ||
        (e = nalloc(n, &n->stats, 8)) != hipSuccess || (e = nalloc(n, &n->part, 3 * 512)) != hipSuccess ||
        (e = nalloc(n, &n->vsum, 2 * V0)) != hipSuccess)
        return fail(e);
    if ((e = hipHostMalloc((void **)&n->h_pinned, 64 * sizeof(double))) != hipSuccess) return fail(e);
    // He-normal initialisation (fan-in), zero biases, deterministic in the seed
    std::vector<float> p(n->nparam, 0.f);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int l = 0; l < NL; ++l) {
        const double std = std::sqrt(2.0 / ((double)LAYERS[l].ks * LAYERS[l].ks * LAYERS[l].ks * LAYERS[l].cin));
        for (int64_t q = 0; q < layer_wcount(l); ++q) p[n->woff[l] + q] = (float)(nd(rng) * std);
    }
    if ((e = hipMemcpy(n->params, p.data(), n->nparam * 4, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = hipMemset(n->grads, 0, n->nparam * 4)) != hipSuccess || (e = hipMemset(n->adam_m, 0, n->nparam * 4)) != hipSuccess ||
        (e = hipMemset(n->adam_v, 0, n->nparam * 4)) != hipSuccess)
        return fail(e);
    if ((e = repack_all(n)) != hipSuccess) return fail(e);
    if ((e = hipStreamSynchronize(n->stream)) != hipSuccess) return fail(e);
    *out = n;
    return MDX_OK;
}

extern "C" int mdx_net_create(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, mdx_net **out) {
    return mdx_net_create_strided(nx, ny, nz, 1, seed, out);
}

extern "C" int mdx_net_grid(mdx_net *n, int32_t *w, int32_t *h, int32_t *d, int32_t *stride) {
    NET_CHECK(n, "mdx_net_grid: null network");
    if (w) *w = n->W[0];
    if (h) *h = n->H[0];
    if (d) *d = n->D[0];
    if (stride) *stride = n->stride;
    return MDX_OK;
}

void mdx::unet::net_detach(mdx_net *n, mdx_handle *h) {
    n->users.erase(std::remove(n->users.begin(), n->users.end(), h), n->users.end());
}

// Destroying a network detaches it from every handle it is attached to (their learned mode then
// falls back to the plain start).
extern "C" int mdx_net_destroy(mdx_net *n) {
    if (!n) return MDX_OK;
    for (mdx_handle *h : n->users) {
        (void)hipStreamSynchronize(h->stream);
        h->net = nullptr;
        h->net_capture = false;
    }
    n->users.clear();
    (void)hipStreamSynchronize(n->stream);
    free_net(n);
    delete n;
    return MDX_OK;
}

extern "C" int mdx_net_get_params(mdx_net *n, float *p, double *c0) {
    NET_CHECK(n && p, "mdx_net_get_params: null argument");
    MDX_HIP(hipMemcpyAsync(p, n->params, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
    MDX_HIP(hipStreamSynchronize(n->stream));
    if (c0) *c0 = n->c0;
    return MDX_OK;
}

extern "C" int mdx_net_set_params(mdx_net *n, const float *p, double c0) {
    NET_CHECK(n && p, "mdx_net_set_params: null argument");
    MDX_HIP(hipMemcpyAsync(n->params, p, n->nparam * 4, hipMemcpyHostToDevice, n->stream));
    if (c0 > 0.0) n->c0 = c0;
    MDX_HIP(repack_all(n));
    MDX_HIP(hipStreamSynchronize(n->stream));
    return MDX_OK;
}

extern "C" int mdx_net_get_grads(mdx_net *n, float *g) {
    NET_CHECK(n && g, "mdx_net_get_grads: null argument");
    MDX_HIP(hipMemcpyAsync(g, n->grads, n->nparam * 4, hipMemcpyDeviceToHost, n->stream));
    MDX_HIP(hipStreamSynchronize(n->stream));
    return MDX_OK;
}

extern "C" int mdx_net_zero_grads(mdx_net *n) {
    NET_CHECK(n, "null network");
    MDX_HIP(hipMemsetAsync(n->grads, 0, n->nparam * 4, n->stream));
    return MDX_OK;
}

// host input [V][2] (div, p) and output [V]: runs the network on explicit inputs
extern "C" int mdx_net_forward_host(mdx_net *n, const float *x2, float *out) {
    NET_CHECK(n && x2 && out, "mdx_net_forward_host: null argument");
    MDX_HIP(upload_input2(n, x2));
    MDX_HIP(net_forward(n, true));
    MDX_HIP(hipMemcpyAsync(out, n->out, n->V[0] * 4, hipMemcpyDeviceToHost, n->stream));
    MDX_HIP(hipStreamSynchronize(n->stream));
    return MDX_OK;
}

// one sample's MSE loss; its parameter gradients are ADDED to the accumulated gradients
// (scaled by 1/batch so that a batch's accumulated gradient is the mean)
extern "C" int mdx_net_accumulate_grad_host(mdx_net *n, const float *x2, const float *target, int32_t batch,
                                            float *loss) {
    NET_CHECK(n && x2 && target && batch > 0, "mdx_net_accumulate_grad_host: bad argument");
    MDX_HIP(upload_input2(n, x2));
    MDX_HIP(hipMemcpyAsync(n->target, target, n->V[0] * 4, hipMemcpyHostToDevice, n->stream));
    MDX_HIP(net_forward(n));
    MDX_HIP(net_backward(n, 1.0f / (float)batch));
    MDX_HIP(hipMemcpyAsync(n->h_pinned, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
    MDX_HIP(hipStreamSynchronize(n->stream));
    if (loss) std::memcpy(loss, n->h_pinned, 4);
    return MDX_OK;
}

extern "C" int mdx_net_adam_step(mdx_net *n, float lr) {
    NET_CHECK(n && lr > 0.f, "mdx_net_adam_step: bad argument");
    MDX_HIP(net_adam(n, lr));
    MDX_HIP(hipStreamSynchronize(n->stream));
    return MDX_OK;
}

// the strided network's transfer operators on host arrays (the device kernels of the solver
// coupling, with unit scale): restriction of a solver-grid field, prolongation of a network-grid one
static int transfer_host(mdx_net *n, const double *fine_in, const float *coarse_in, float *coarse_out,
                         double *fine_out) {
    double *tmp = nullptr;
    MDX_HIP(hipMalloc((void **)&tmp, (size_t)n->FV * 8));
    hipError_t e = hipSuccess;
    if (fine_in) {
        e = hipMemcpyAsync(tmp, fine_in, (size_t)n->FV * 8, hipMemcpyHostToDevice, n->stream);
        if (e == hipSuccess) e = net_restrict(n, tmp, n->target);
        if (e == hipSuccess) e = hipMemcpyAsync(coarse_out, n->target, (size_t)n->V[0] * 4, hipMemcpyDeviceToHost, n->stream);
    } else {
        e = hipMemcpyAsync(n->out, coarse_in, (size_t)n->V[0] * 4, hipMemcpyHostToDevice, n->stream);
        if (e == hipSuccess) e = net_prolong(n, n->out, tmp);
        if (e == hipSuccess) e = hipMemcpyAsync(fine_out, tmp, (size_t)n->FV * 8, hipMemcpyDeviceToHost, n->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(n->stream);
    (void)hipFree(tmp);
    MDX_HIP(e);
    return MDX_OK;
}

extern "C" int mdx_net_restrict_host(mdx_net *n, const double *fine, float *coarse) {
    NET_CHECK(n && fine && coarse, "mdx_net_restrict_host: null argument");
    return transfer_host(n, fine, nullptr, coarse, nullptr);
}

extern "C" int mdx_net_prolong_host(mdx_net *n, const float *coarse, double *fine) {
    NET_CHECK(n && coarse && fine, "mdx_net_prolong_host: null argument");
    return transfer_host(n, nullptr, coarse, nullptr, fine);
}

// ----------------------------------------------------------------------------------- replay / training
extern "C" int mdx_net_replay_reserve(mdx_net *n, int32_t capacity) {
    NET_CHECK(n && capacity > 0, "mdx_net_replay_reserve: capacity must be positive");
    NET_CHECK(n->cap == 0, "mdx_net_replay_reserve: already reserved");
    const long V = n->V[0];
    MDX_HIP(nalloc(n, &n->rb_x, (size_t)capacity * 2 * V));
    MDX_HIP(nalloc(n, &n->rb_t, (size_t)capacity * V));
    n->cap = capacity;
    n->count = n->next = 0;
    return MDX_OK;
}

extern "C" int mdx_net_replay_count(mdx_net *n, int32_t *count) {
    NET_CHECK(n && count, "mdx_net_replay_count: null argument");
    *count = n->count;
    return MDX_OK;
}

// steps Adam updates, each on `batch` samples drawn uniformly from the replay buffer
extern "C" int mdx_net_train(mdx_net *n, int32_t steps, int32_t batch, float lr, uint64_t seed, float *mean_loss) {
    NET_CHECK(n && steps > 0 && batch > 0 && lr > 0.f, "mdx_net_train: bad argument");
    NET_CHECK(n->count > 0, "mdx_net_train: the replay buffer is empty");
    std::mt19937_64 rng(seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(n->adam_t + 1)));
    double total = 0.0;
    MDX_HIP(hipMemsetAsync(n->grads, 0, n->nparam * 4, n->stream));
    for (int s = 0; s < steps; ++s) {
        float lsum = 0.f;
        for (int b = 0; b < batch; ++b) {
            const int slot = (int)(rng() % (uint64_t)n->count);
            MDX_HIP(net_load_replay(n, slot));
            MDX_HIP(net_forward(n));
            MDX_HIP(net_backward(n, 1.0f / (float)batch));
            MDX_HIP(hipMemcpyAsync(n->h_pinned + b, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
        }
        MDX_HIP(net_adam(n, lr));
        MDX_HIP(hipStreamSynchronize(n->stream));
        for (int b = 0; b < batch; ++b) {
            float l;
            std::memcpy(&l, n->h_pinned + b, 4);
            lsum += l;
        }
        total += lsum;
    }
    if (mean_loss) *mean_loss = (float)(total / steps);
    return MDX_OK;
}

// mean MSE (normalised target units) of the network over the replay buffer's first `count` slots
extern "C" int mdx_net_eval_replay(mdx_net *n, int32_t count, float *mean_loss) {
    NET_CHECK(n && mean_loss && count > 0 && count <= n->count, "mdx_net_eval_replay: bad argument");
    double total = 0.0;
    for (int s = 0; s < count; ++s) {
        MDX_HIP(net_load_replay(n, s));
        MDX_HIP(net_forward(n));
        MDX_HIP(net_mse_only(n));
        MDX_HIP(hipMemcpyAsync(n->h_pinned, n->loss, 4, hipMemcpyDeviceToHost, n->stream));
        MDX_HIP(hipStreamSynchronize(n->stream));
        float l;
        std::memcpy(&l, n->h_pinned, 4);
        total += l;
    }
    *mean_loss = (float)(total / count);
    return MDX_OK;
}

// Device time of one forward pass (HIP events on the network's stream around `reps` passes over
// replay sample 0, or the last input without a replay buffer, after one untimed pass) and the MFMA work of one pass: 2 * voxels * 27 * cin *
// cout per convolution as the kernels issue it (the 2-channel input padded to 16), plus the head.
extern "C" int mdx_net_time_forward(mdx_net *n, int32_t reps, float *ms_per_forward, double *flops) {
    NET_CHECK(n && ms_per_forward && flops && reps > 0, "mdx_net_time_forward: bad argument");
    const double v0 = (double)n->V[0], v1 = (double)n->V[1], v2 = (double)n->V[2];
    const double convs = v0 * (16 * 16 + 16 * 16 + 48 * 16 + 16 * 16) + v1 * (16 * 32 + 32 * 32 + 96 * 32 + 32 * 32) +
                         v2 * (32 * 64 + 64 * 64);
    *flops = 2.0 * 27.0 * convs + 2.0 * v0 * 16.0;
    if (n->count > 0) MDX_HIP(net_load_replay(n, 0));   // else the input of the last forward pass
    MDX_HIP(net_forward(n, true));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    MDX_HIP(hipEventCreate(&e0));
    hipError_t e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, n->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = net_forward(n, true);
    if (e == hipSuccess) e = hipEventRecord(e1, n->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    MDX_HIP(e);
    *ms_per_forward = ms / (float)reps;
    return MDX_OK;
}

// attach a network to a solver handle: used for the initial guess when cfg.guess is
// MDX_GUESS_LEARNED, and (capture != 0) every converged SOR solve appends a training pair
extern "C" int mdx_attach_net(mdx_handle *h, mdx_net *n, int32_t capture) {
    if (!h) { mdx_set_error("mdx_attach_net: null handle"); return MDX_ERR_ARG; }
    if (n) {
        NET_CHECK(n->FW == h->d.nx && n->FH == h->d.ny && n->FD == h->d.nz,
                  "mdx_attach_net: the network serves another solver grid");
        NET_CHECK(h->cfg.solver == MDX_SOLVER_SOR, "mdx_attach_net: the learned guess serves the SOR solver");
        NET_CHECK(!capture || n->cap > 0, "mdx_attach_net: capture needs mdx_net_replay_reserve first");
    }
    if (h->net) mdx::unet::net_detach(h->net, h);
    h->net = n;
    h->net_capture = n && capture;
    if (n) n->users.push_back(h);
    return MDX_OK;
}
