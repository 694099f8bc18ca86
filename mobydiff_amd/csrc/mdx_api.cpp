// mdx_api.cpp -- the C ABI of libmobydiff_hip: device state, solver dispatch and host-side setup.
#include <immintrin.h>

#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <functional>
#include <string>
#include <thread>
#include <vector>

#include "mdx_internal.h"
#include "mdx_unet.h"

static thread_local std::string g_last_error;

void mdx_set_error(const std::string &msg) { g_last_error = msg; }

extern "C" const char *mdx_last_error(void) { return g_last_error.c_str(); }

namespace mdx {

GridK grid_k(const mdx_handle *h) {
    GridK g;
    g.d = h->d;
    g.dx = h->dx;
    g.dy = h->dy;
    g.dz = h->dz;
    g.re = h->cfg.re;
    g.fx = h->cfg.forcing[0];
    g.fy = h->cfg.forcing[1];
    g.fz = h->cfg.forcing[2];
    for (int q = 0; q < 3; ++q) g.per[q] = h->cfg.periodic[q] ? 1 : 0;
    g.ihc = h->stretched ? h->d_ihc : nullptr;
    g.ihv = h->stretched ? h->d_ihv : nullptr;
    return g;
}

BcK bc_k(const mdx_handle *h) {
    BcK b;
    for (int q = 0; q < 24; ++q) {
        b.type[q] = h->cfg.bc_type[q];
        b.value[q] = h->cfg.bc_value[q];
    }
    return b;
}

}  // namespace mdx

using namespace mdx;

static const double PI = 3.141592653589793;

#define MDX_CHECK(cond, code, msg)  \
    do {                            \
        if (!(cond)) {              \
            mdx_set_error(msg);     \
            return code;            \
        }                           \
    } while (0)

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

// pressure_fft.f90:63,79 divide one_c by a complex denominator; compiled with the scaled C99 Annex G
// algorithm this rounds as c/(c*c) after scaling c by 2^-logb(c).  Kept identical so the Thomas
// coefficients match the reference bit for bit.
static double recip_complex_div(double c) {
    double lb = std::logb(std::fabs(c));
    int e = 0;
    if (std::isfinite(lb)) {
        e = (int)lb;
        c = std::scalbn(c, -e);
    }
    double denom = c * c;
    return std::scalbn(c / denom, -e);
}

// Modified-wavenumber Thomas factorisation per (kx,kz) line (pressure_fft.f90:20-100), generalised
// to a stretched wall-normal grid.  cp/den_inv layout (nxh, ny, nz), a_row(ny).
static void tridiag_tables(const mdx_handle *h, std::vector<double> &cp, std::vector<double> &di,
                           std::vector<double> &ar) {
    const int nx = h->d.nx, ny = h->d.ny, nz = h->d.nz, nxh = nx / 2 + 1;
    const double dx = h->dx, dy = h->dy, dz = h->dz;
    const double dyi2 = 1.0 / (dy * dy);
    cp.assign((size_t)nxh * ny * nz, 0.0);
    di.assign((size_t)nxh * ny * nz, 0.0);
    ar.assign(ny, 0.0);
    for (int j = 1; j <= ny; ++j) ar[j - 1] = h->stretched ? h->ihv[j] * h->ihc[j] : dyi2;
    for (int ikx = 1; ikx <= nxh; ++ikx)
        for (int ikz = 1; ikz <= nz; ++ikz) {
            const int kx = ikx - 1;
            const double s = std::sin(PI * (double)kx / (double)nx);
            const double kx_ph = (4.0 / (dx * dx)) * (s * s);
            const int kz = (ikz <= nz / 2 + 1) ? ikz - 1 : ikz - nz - 1;
            const double sz = std::sin(PI * (double)kz / (double)nz);
            const double kz_ph = (4.0 / (dz * dz)) * (sz * sz);
            const double k_tot = kx_ph + kz_ph;
            const bool zero_mode = (kx == 0 && kz == 0);
            auto CI = [&](int jj) { return (size_t)(ikx - 1) + (size_t)nxh * ((size_t)(jj - 1) + (size_t)ny * (ikz - 1)); };
            if (!h->stretched) {
                if (zero_mode) {
                    di[CI(1)] = 1.0;
                    cp[CI(1)] = 0.0;
                } else {
                    double b = -dyi2 - k_tot;
                    di[CI(1)] = recip_complex_div(b);
                    cp[CI(1)] = dyi2 * di[CI(1)];
                }
                for (int jj = 2; jj <= ny; ++jj) {
                    double b, c;
                    if (jj < ny) { b = -2.0 * dyi2 - k_tot; c = dyi2; }
                    else { b = -dyi2 - k_tot; c = 0.0; }
                    double denom = b - dyi2 * cp[CI(jj - 1)];
                    di[CI(jj)] = recip_complex_div(denom);
                    cp[CI(jj)] = (jj < ny) ? c * di[CI(jj)] : 0.0;
                }
            } else {
                const double *ihc = h->ihc.data(), *ihv = h->ihv.data();
                if (zero_mode) {
                    di[CI(1)] = 1.0;
                    cp[CI(1)] = 0.0;
                } else {
                    double c1 = ihv[2] * ihc[1];
                    double b = -c1 - k_tot;
                    di[CI(1)] = recip_complex_div(b);
                    cp[CI(1)] = c1 * di[CI(1)];
                }
                for (int jj = 2; jj <= ny; ++jj) {
                    double aj = ihv[jj] * ihc[jj];
                    double cj = (jj < ny) ? ihv[jj + 1] * ihc[jj] : 0.0;
                    double b = -(aj + cj) - k_tot;
                    double denom = b - aj * cp[CI(jj - 1)];
                    di[CI(jj)] = recip_complex_div(denom);
                    cp[CI(jj)] = (jj < ny) ? cj * di[CI(jj)] : 0.0;
                }
            }
        }
}

// Row coefficients and modified wavenumbers of the one-wave-per-line solve (tridiag_wave_kernel):
// row j (0-based) reads a_j x_{j-1} + (bd_j - k) x_j + c_j x_{j+1} with k = lamx[kx] + lamz[kz];
// a_0 = 0 and c_{ny-1} = 0 (the homogeneous Neumann rows), the same expressions as the Thomas
// tables above.
static void tridiag_row_tables(const mdx_handle *h, std::vector<double> &lx, std::vector<double> &lz,
                       std::vector<double> &a, std::vector<double> &bd, std::vector<double> &c) {
    const int nx = h->d.nx, ny = h->d.ny, nz = h->d.nz, nxh = nx / 2 + 1;
    const double dx = h->dx, dy = h->dy, dz = h->dz;
    const double dyi2 = 1.0 / (dy * dy);
    lx.assign(nxh, 0.0);
    lz.assign(nz, 0.0);
    for (int kx = 0; kx < nxh; ++kx) {
        const double s = std::sin(PI * (double)kx / (double)nx);
        lx[kx] = (4.0 / (dx * dx)) * (s * s);
    }
    for (int ikz = 1; ikz <= nz; ++ikz) {
        const int kz = (ikz <= nz / 2 + 1) ? ikz - 1 : ikz - nz - 1;
        const double sz = std::sin(PI * (double)kz / (double)nz);
        lz[ikz - 1] = (4.0 / (dz * dz)) * (sz * sz);
    }
    a.assign(ny, 0.0);
    bd.assign(ny, 0.0);
    c.assign(ny, 0.0);
    for (int jj = 1; jj <= ny; ++jj) {
        double aj, cj;
        if (!h->stretched) {
            aj = jj > 1 ? dyi2 : 0.0;
            cj = jj < ny ? dyi2 : 0.0;
            bd[jj - 1] = (jj == 1 || jj == ny) ? -dyi2 : -2.0 * dyi2;
        } else {
            aj = jj > 1 ? h->ihv[jj] * h->ihc[jj] : 0.0;
            cj = jj < ny ? h->ihv[jj + 1] * h->ihc[jj] : 0.0;
            bd[jj - 1] = jj == 1 ? -cj : -(aj + cj);
        }
        a[jj - 1] = aj;
        c[jj - 1] = cj;
    }
}

static void twiddles(int n, std::vector<double> &tw) {
    tw.assign((size_t)n, 0.0);   // n/2 complex
    for (int m = 0; m < n / 2; ++m) {
        double ang = -2.0 * PI * (double)m / (double)n;
        tw[2 * m] = std::cos(ang);
        tw[2 * m + 1] = std::sin(ang);
    }
}

// tanh-stretched wall-normal grid on [0, ly]: faces yv(j), j=1..ny+1, centres midway, ghost
// centres mirrored about the walls; uniform grid uses the IBM sample-position expressions.
static void y_metrics(mdx_handle *h) {
    const int ny = h->d.ny;
    h->yc.assign(ny + 2, 0.0);
    h->yv.assign(ny + 2, 0.0);
    h->ihc.assign(ny + 2, 0.0);
    h->ihv.assign(ny + 2, 0.0);
    const double ly = h->cfg.ly, beta = h->cfg.stretch_beta;
    if (!h->stretched) {
        for (int j = 0; j <= ny + 1; ++j) {
            h->yc[j] = ((double)(j + 1) - 1.5) * h->dy;
            h->yv[j] = ((double)(j + 1) - 0.5 - 1.5) * h->dy;
            h->ihc[j] = 1.0 / h->dy;
            h->ihv[j] = 1.0 / h->dy;
        }
        return;
    }
    std::vector<double> yv(ny + 2);
    const double tb = std::tanh(beta);
    for (int j = 1; j <= ny + 1; ++j) {
        double eta = (double)(j - 1) / (double)ny;
        yv[j] = 0.5 * ly * (1.0 + std::tanh(beta * (2.0 * eta - 1.0)) / tb);
    }
    yv[1] = 0.0;
    yv[ny + 1] = ly;
    yv[0] = -yv[2];
    for (int j = 1; j <= ny; ++j) h->yc[j] = 0.5 * (yv[j] + yv[j + 1]);
    h->yc[0] = 2.0 * yv[1] - h->yc[1];
    h->yc[ny + 1] = 2.0 * yv[ny + 1] - h->yc[ny];
    for (int j = 1; j <= ny; ++j) h->ihc[j] = 1.0 / (yv[j + 1] - yv[j]);
    h->ihc[0] = h->ihc[1];
    h->ihc[ny + 1] = h->ihc[ny];
    for (int j = 1; j <= ny + 1; ++j) h->ihv[j] = 1.0 / (h->yc[j] - h->yc[j - 1]);
    h->ihv[0] = h->ihv[1];
    for (int j = 0; j <= ny + 1; ++j) h->yv[j] = yv[j];
}

template <typename T>
static hipError_t dalloc(T **p, size_t n) {
    return hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
}

static void free_all(mdx_handle *h) {
    void *ptrs[] = {h->un, h->vn, h->wn, h->pn, h->us, h->vs, h->ws, h->oru, h->orv, h->orw,
                    h->cu, h->cv, h->cw, h->d_ihc, h->d_ihv, h->d_yc, h->phi, h->phi2, h->rhs,
                    h->e32, h->r32, h->e32b, h->mu64[0], h->mu64[1], h->mu64[2], h->mu32[0], h->mu32[1],
                    h->mu32[2], h->fft_rhs, h->hat, h->cp, h->den_inv, h->arow, h->tw_x, h->tw_z,
                    h->red_part, h->red_max, h->max_part, h->rb_faces, h->phi_hist, h->res64,
                    h->fft_phi, h->fft_res, h->fft_mean, h->hat32, h->cp32, h->den32, h->arow32, h->tw_x32,
                    h->tw_z32, h->rb_xl32, h->rb_xl64, h->cvel, h->lamx, h->lamz, h->tw_a, h->bdrow, h->crow, h->lamx32, h->lamz32,
                    h->tw_a32, h->bdrow32, h->crow32, h->fr_part};
    for (auto &s : h->graphs)
        if (s.exec) (void)hipGraphExecDestroy(s.exec);
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (h->h_pinned) (void)hipHostFree(h->h_pinned);
    if (h->timing) {
        for (auto &e : h->ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &q : h->ev_sw)
            for (auto &e : q)
                if (e) (void)hipEventDestroy(e);
        for (auto &q : h->ev_ph)
            for (auto &e : q)
                if (e) (void)hipEventDestroy(e);
    }
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

extern "C" int mdx_create(const mdx_config *cfg, mdx_handle **out) {
    MDX_CHECK(cfg && out, MDX_ERR_ARG, "mdx_create: null argument");
    *out = nullptr;
    const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    MDX_CHECK(nx >= 2 && ny >= 2 && nz >= 2, MDX_ERR_ARG, "mdx_create: nx, ny, nz must be >= 2");
    MDX_CHECK((double)(nx + 2) * (ny + 2) * (nz + 2) < 2.0e9 && xcd_grid(nx, ny, nz, 64, 4) <= MAX_PARTS,
              MDX_ERR_ARG, "mdx_create: grid too large");
    MDX_CHECK(cfg->lx > 0 && cfg->ly > 0 && cfg->lz > 0 && cfg->re > 0, MDX_ERR_ARG,
              "mdx_create: lengths and re must be positive");
    if (cfg->solver == MDX_SOLVER_FFT) {
        MDX_CHECK(is_pow2(nx) && is_pow2(nz) && nx <= 1024 && nz <= 1024, MDX_ERR_UNSUPPORTED,
                  "FFT pressure solver: nx and nz must be powers of two <= 1024");
        MDX_CHECK(cfg->precision != MDX_PREC_MIXED || cfg->tol > 0.0, MDX_ERR_ARG,
                  "mixed-precision FFT solve: tol > 0 required (||div||_inf target of the defect correction)");
        MDX_CHECK(cfg->precision != MDX_PREC_MIXED || ny <= 1024, MDX_ERR_UNSUPPORTED,
                  "mixed-precision FFT solve: ny <= 1024 (the FP64 mean-mode solve stages its line in LDS)");
    } else if (cfg->solver == MDX_SOLVER_SOLA || cfg->solver == MDX_SOLVER_SOR) {
        MDX_CHECK(nx % 2 == 0 && nz % 2 == 0, MDX_ERR_ARG,
                  "red-black pressure solver requires even nx and nz for periodic coloring");
        MDX_CHECK(cfg->omega > 0.0 && cfg->omega < 2.0, MDX_ERR_ARG, "red-black pressure solver: 0 < omega < 2");
        if (cfg->solver == MDX_SOLVER_SOLA) MDX_CHECK(cfg->niter > 0, MDX_ERR_ARG, "pressure nIter must be positive");
        if (cfg->solver == MDX_SOLVER_SOR) {
            MDX_CHECK(cfg->periodic[0] && cfg->periodic[2], MDX_ERR_UNSUPPORTED,
                      "SOR to tolerance: x and z must be periodic");
            if (!cfg->periodic[1]) {
                MDX_CHECK(cfg->bc_type[bci(2, 0, 2)] == 0 && cfg->bc_type[bci(2, 1, 2)] == 0 &&
                              cfg->bc_type[bci(2, 0, 0)] == 1 && cfg->bc_type[bci(2, 1, 0)] == 1,
                          MDX_ERR_UNSUPPORTED,
                          "SOR to tolerance: y walls need Dirichlet v and Neumann p (or periodic y)");
            }
            MDX_CHECK(cfg->tol > 0.0 && cfg->check_every > 0, MDX_ERR_ARG,
                      "SOR to tolerance: tol > 0 and check_every > 0 required");
            MDX_CHECK(cfg->precision != MDX_PREC_MIXED || cfg->inner_iters > 0, MDX_ERR_ARG,
                      "mixed precision: inner_iters > 0 required");
        }
    } else {
        MDX_CHECK(false, MDX_ERR_ARG, "mdx_create: unknown solver");
    }
    MDX_CHECK(cfg->guess == MDX_GUESS_PREVIOUS || (cfg->solver == MDX_SOLVER_SOR &&
              (cfg->guess == MDX_GUESS_LEARNED || cfg->guess == MDX_GUESS_PHI_PREV)),
              MDX_ERR_ARG, "mdx_create: unknown guess mode, or a guess for a solver other than SOR");
    if (cfg->flowrate)
        MDX_CHECK(cfg->periodic[0] && cfg->flow_height > 0.0, MDX_ERR_UNSUPPORTED,
                  "constant flow rate needs periodic x and flow_height > 0");
    MDX_CHECK(cfg->stretch_beta >= 0.0, MDX_ERR_ARG, "stretch_beta must be >= 0");

    mdx_handle *h = new mdx_handle();
    h->cfg = *cfg;
    h->d = Dims{nx, ny, nz, nx + 2, ny + 2};
    h->dx = cfg->lx / (double)nx;   // init.f90:48-55
    h->dy = cfg->ly / (double)ny;
    h->dz = cfg->lz / (double)nz;
    h->ng = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    h->ni = (size_t)nx * ny * nz;
    h->stretched = cfg->stretch_beta > 0.0;
    if (const char *ng = std::getenv("MOBYDIFF_NO_GRAPHS")) h->use_graphs = !(ng[0] == '1');
    if (const char *fi = std::getenv("MOBYDIFF_FUSED_INPUT")) h->fused_input = !(fi[0] == '0');
    if (const char *dc = std::getenv("MOBYDIFF_DEFER_CHOICE")) h->defer_choice = !(dc[0] == '0');
    h->inner_omega = h->cfg.omega;
    if (const char *io = std::getenv("MOBYDIFF_INNER_OMEGA")) {
        const double v = std::atof(io);
        if (v > 0.0 && v < 2.0) h->inner_omega = v;
    }
    y_metrics(h);

    auto fail = [&](hipError_t e, const char *what) {
        mdx_set_error(std::string("mdx_create: ") + what + ": " + hipGetErrorString(e));
        free_all(h);
        delete h;
        return MDX_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipGetDevice(&h->device)) != hipSuccess) return fail(e, "hipGetDevice");
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
    double **fields[] = {&h->un, &h->vn, &h->wn, &h->pn, &h->us, &h->vs, &h->ws};
    for (double **f : fields)
        if ((e = dalloc(f, h->ng)) != hipSuccess || (e = hipMemset(*f, 0, h->ng * 8)) != hipSuccess)
            return fail(e, "field allocation");
    double **olds[] = {&h->oru, &h->orv, &h->orw};
    for (double **f : olds)
        if ((e = dalloc(f, h->ni)) != hipSuccess || (e = hipMemset(*f, 0, h->ni * 8)) != hipSuccess)
            return fail(e, "oldrhs allocation");
    if ((e = dalloc(&h->d_ihc, ny + 2)) != hipSuccess || (e = dalloc(&h->d_ihv, ny + 2)) != hipSuccess ||
        (e = dalloc(&h->d_yc, ny + 2)) != hipSuccess)
        return fail(e, "metric allocation");
    if ((e = hipMemcpy(h->d_ihc, h->ihc.data(), (ny + 2) * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->d_ihv, h->ihv.data(), (ny + 2) * 8, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(h->d_yc, h->yc.data(), (ny + 2) * 8, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "metric upload");
    if ((e = dalloc(&h->red_part, 4096)) != hipSuccess || (e = dalloc(&h->max_part, MAX_PARTS)) != hipSuccess)
        return fail(e, "reduction buffer");
    if ((e = dalloc(&h->red_max, 4)) != hipSuccess || (e = hipMemset(h->red_max, 0, 32)) != hipSuccess)
        return fail(e, "reduction buffer");
    h->rb_err = reinterpret_cast<int *>(h->red_max + 1);
    if (const char *ls = std::getenv("MOBYDIFF_LDS_SOR")) {
        const int v = std::atoi(ls);
        if (v < 0 || v > 2) {
            mdx_set_error("mdx_create: MOBYDIFF_LDS_SOR must be 0, 1 or 2");
            free_all(h);
            delete h;
            return MDX_ERR_ARG;
        }
        h->lds_sor_mode = v;
    }
    if (const char *lc = std::getenv("MOBYDIFF_LDS_COOP")) h->rb_coop = lc[0] == '1';
    if (const char *sl = std::getenv("MOBYDIFF_LDS_SPIN_LIMIT")) h->rb_spin_limit = std::max(0, std::atoi(sl));
    if (const char *nm = std::getenv("MOBYDIFF_NET_MARGIN")) h->net_margin = std::atof(nm);
    if (const char *tg = std::getenv("MOBYDIFF_NET_TIME_GATE")) h->net_time_gate = tg[0] != '0';
    if (const char *nf = std::getenv("MOBYDIFF_NET_FALLBACK")) h->net_fallback_prev = std::strcmp(nf, "prev") == 0;
    if (const char *wt = std::getenv("MOBYDIFF_WAVE_TRIDIAG")) h->wave_tridiag = wt[0] != '0';
    if (const char *pt = std::getenv("MOBYDIFF_PART_TRIDIAG")) h->part_tridiag = pt[0] != '0';
    if (const char *mm = std::getenv("MOBYDIFF_MERGE_MAX")) h->merge_max = std::max(1, std::atoi(mm));
    if (const char *lf = std::getenv("MOBYDIFF_LDS_FORM")) {
        if (std::strcmp(lf, "pairs") == 0) h->rb_form = 1;
        else if (std::strcmp(lf, "columns") == 0) h->rb_form = 2;
        else if (std::strcmp(lf, "halves") != 0) {
            mdx_set_error("mdx_create: MOBYDIFF_LDS_FORM must be halves, columns or pairs");
            free_all(h);
            delete h;
            return MDX_ERR_ARG;
        }
    }
    if (const char *lo = std::getenv("MOBYDIFF_LDS_OVERLAP")) h->rb_overlap = lo[0] != '0';
    h->mom_kz = momentum_default_planes(h->d);
    if (const char *mk = std::getenv("MOBYDIFF_MOM_KZ")) h->mom_kz = std::max(1, std::atoi(mk));
    if (const char *md = std::getenv("MOBYDIFF_MOMENTUM_DIRECT")) h->tiled_momentum = md[0] != '1';
    if (const char *ff = std::getenv("MOBYDIFF_FUSED_FLOWRATE")) h->fused_flowrate = ff[0] != '0';
    if (const char *fb = std::getenv("MOBYDIFF_FUSED_BC")) h->fused_bc = fb[0] != '0';
    if (const char *fm = std::getenv("MOBYDIFF_FUSED_MU")) h->fused_mu = fm[0] != '0';
    if (h->cfg.flowrate && h->tiled_momentum && h->fused_flowrate &&
        (e = dalloc(&h->fr_part, 2 * (size_t)momentum_tiles(h))) != hipSuccess)
        return fail(e, "flow-rate partial buffer");
    if (const char *sf = std::getenv("MOBYDIFF_SWEEP_FUSED")) h->sweep_fused = sf[0] == '1';
    if (const char *kz = std::getenv("MOBYDIFF_SWEEP_KZ")) {
        const int v = std::atoi(kz);
        if (!(v == 0 || v == 1 || v == 2 || v == 4 || v == 8)) {
            mdx_set_error("mdx_create: MOBYDIFF_SWEEP_KZ must be 0, 1, 2, 4 or 8");
            free_all(h);
            delete h;
            return MDX_ERR_ARG;
        }
        h->sweep_kz = v;
    }
    // coherent (fine-grained): the solver's checks write their results here while later work runs
    if ((e = hipHostMalloc((void **)&h->h_pinned, 64 * sizeof(double), hipHostMallocMapped | hipHostMallocCoherent)) !=
        hipSuccess)
        return fail(e, "pinned buffer");
    // the check slots' sequence words must not hold a number this handle will wait for: a block
    // freed by an earlier handle can come back with its last sequence numbers in it
    std::memset(h->h_pinned, 0, 64 * sizeof(double));
    if ((e = hipHostGetDevicePointer((void **)&h->d_pinned, h->h_pinned, 0)) != hipSuccess)
        return fail(e, "pinned buffer device address");

    if (cfg->solver == MDX_SOLVER_SOR) {
        if ((e = dalloc(&h->phi, h->ni)) != hipSuccess || (e = dalloc(&h->phi2, h->ni)) != hipSuccess ||
            (e = dalloc(&h->rhs, h->ni)) != hipSuccess)
            return fail(e, "SOR buffers");
        if (cfg->precision == MDX_PREC_MIXED)
            if ((e = dalloc(&h->e32, h->ni)) != hipSuccess || (e = dalloc(&h->r32, h->ni)) != hipSuccess ||
                (e = dalloc(&h->e32b, h->ni)) != hipSuccess)
                return fail(e, "SOR FP32 buffers");
        if ((e = hipMemset(h->phi, 0, h->ni * 8)) != hipSuccess) return fail(e, "SOR buffers");
        if ((e = dalloc(&h->phi_hist, 3 * h->ni)) != hipSuccess || (e = dalloc(&h->res64, h->ni)) != hipSuccess)
            return fail(e, "SOR guess history");
    }
    if (cfg->solver == MDX_SOLVER_FFT) {
        const int nxh = nx / 2 + 1;
        const size_t nh = (size_t)nxh * ny * nz;
        std::vector<double> cp, di, ar, twx, twz;
        tridiag_tables(h, cp, di, ar);
        twiddles(nx, twx);
        twiddles(nz, twz);
        if ((e = dalloc(&h->fft_rhs, h->ni)) != hipSuccess || (e = dalloc(&h->hat, 2 * nh)) != hipSuccess ||
            (e = dalloc(&h->cp, nh)) != hipSuccess || (e = dalloc(&h->den_inv, nh)) != hipSuccess ||
            (e = dalloc(&h->arow, ny)) != hipSuccess || (e = dalloc(&h->tw_x, nx)) != hipSuccess ||
            (e = dalloc(&h->tw_z, nz)) != hipSuccess)
            return fail(e, "FFT workspace");
        std::vector<double> plx, plz, pa, pbd, pc;
        tridiag_row_tables(h, plx, plz, pa, pbd, pc);
        if ((e = dalloc(&h->lamx, nxh)) != hipSuccess || (e = dalloc(&h->lamz, nz)) != hipSuccess ||
            (e = dalloc(&h->tw_a, ny)) != hipSuccess || (e = dalloc(&h->bdrow, ny)) != hipSuccess ||
            (e = dalloc(&h->crow, ny)) != hipSuccess)
            return fail(e, "FFT workspace");
        if ((e = hipMemcpy(h->lamx, plx.data(), nxh * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->lamz, plz.data(), nz * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->tw_a, pa.data(), ny * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->bdrow, pbd.data(), ny * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->crow, pc.data(), ny * 8, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "FFT table upload");
        if ((e = hipMemcpy(h->cp, cp.data(), nh * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->den_inv, di.data(), nh * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->arow, ar.data(), ny * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->tw_x, twx.data(), nx * 8, hipMemcpyHostToDevice)) != hipSuccess ||
            (e = hipMemcpy(h->tw_z, twz.data(), nz * 8, hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "FFT table upload");
        if (cfg->precision == MDX_PREC_MIXED) {
            auto to32 = [](const std::vector<double> &v) { return std::vector<float>(v.begin(), v.end()); };
            const std::vector<float> cp32 = to32(cp), di32 = to32(di), ar32 = to32(ar), twx32 = to32(twx),
                                     twz32 = to32(twz);
            if ((e = dalloc(&h->fft_phi, h->ni)) != hipSuccess || (e = dalloc(&h->fft_res, h->ni)) != hipSuccess ||
                (e = dalloc(&h->fft_mean, 17 * (size_t)ny)) != hipSuccess ||
                (e = dalloc(&h->hat32, 2 * nh)) != hipSuccess || (e = dalloc(&h->cp32, nh)) != hipSuccess ||
                (e = dalloc(&h->den32, nh)) != hipSuccess || (e = dalloc(&h->arow32, ny)) != hipSuccess ||
                (e = dalloc(&h->tw_x32, nx)) != hipSuccess || (e = dalloc(&h->tw_z32, nz)) != hipSuccess)
                return fail(e, "mixed FFT workspace");
            if ((e = hipMemcpy(h->cp32, cp32.data(), nh * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->den32, di32.data(), nh * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->arow32, ar32.data(), ny * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->tw_x32, twx32.data(), nx * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->tw_z32, twz32.data(), nz * 4, hipMemcpyHostToDevice)) != hipSuccess)
                return fail(e, "mixed FFT table upload");
            const std::vector<float> plx32 = to32(plx), plz32 = to32(plz), pa32 = to32(pa), pbd32 = to32(pbd),
                                     pc32 = to32(pc);
            if ((e = dalloc(&h->lamx32, nxh)) != hipSuccess || (e = dalloc(&h->lamz32, nz)) != hipSuccess ||
                (e = dalloc(&h->tw_a32, ny)) != hipSuccess || (e = dalloc(&h->bdrow32, ny)) != hipSuccess ||
                (e = dalloc(&h->crow32, ny)) != hipSuccess)
                return fail(e, "mixed FFT workspace");
            if ((e = hipMemcpy(h->lamx32, plx32.data(), nxh * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->lamz32, plz32.data(), nz * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->tw_a32, pa32.data(), ny * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->bdrow32, pbd32.data(), ny * 4, hipMemcpyHostToDevice)) != hipSuccess ||
                (e = hipMemcpy(h->crow32, pc32.data(), ny * 4, hipMemcpyHostToDevice)) != hipSuccess)
                return fail(e, "mixed FFT table upload");
        }
    }
    if (cfg->timing) {
        h->timing = true;
        for (auto &ev : h->ev)
            if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(e, "events");
        for (auto &q : h->ev_sw)
            for (auto &ev : q)
                if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(e, "events");
        for (auto &q : h->ev_ph)
            for (auto &ev : q)
                if ((e = hipEventCreate(&ev)) != hipSuccess) return fail(e, "events");
    }
    if ((e = hipDeviceSynchronize()) != hipSuccess) return fail(e, "synchronize");
    *out = h;
    return MDX_OK;
}

extern "C" int mdx_destroy(mdx_handle *h) {
    if (!h) return MDX_OK;
    (void)hipStreamSynchronize(h->stream);
    if (h->net) unet::net_detach(h->net, h);
    free_all(h);
    delete h;
    return MDX_OK;
}

extern "C" int mdx_synchronize(mdx_handle *h) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_HIP(hipStreamSynchronize(h->stream));
    return MDX_OK;
}

static int copy_in(mdx_handle *h, double *dst, const double *src, size_t n) {
    if (!src) return MDX_OK;
    MDX_HIP(hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, h->stream));
    return MDX_OK;
}
static int copy_out(mdx_handle *h, double *dst, const double *src, size_t n) {
    if (!dst) return MDX_OK;
    MDX_HIP(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, h->stream));
    return MDX_OK;
}

extern "C" int mdx_upload_fields(mdx_handle *h, const double *un, const double *vn, const double *wn,
                                 const double *pn, const double *us, const double *vs, const double *ws,
                                 const double *oru, const double *orv, const double *orw) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    int r;
    if ((r = copy_in(h, h->un, un, h->ng)) || (r = copy_in(h, h->vn, vn, h->ng)) || (r = copy_in(h, h->wn, wn, h->ng)) ||
        (r = copy_in(h, h->pn, pn, h->ng)) || (r = copy_in(h, h->us, us, h->ng)) || (r = copy_in(h, h->vs, vs, h->ng)) ||
        (r = copy_in(h, h->ws, ws, h->ng)) || (r = copy_in(h, h->oru, oru, h->ni)) ||
        (r = copy_in(h, h->orv, orv, h->ni)) || (r = copy_in(h, h->orw, orw, h->ni)))
        return r;
    MDX_HIP(hipStreamSynchronize(h->stream));
    return MDX_OK;
}

extern "C" int mdx_download_fields(mdx_handle *h, double *un, double *vn, double *wn, double *pn, double *us,
                                   double *vs, double *ws, double *oru, double *orv, double *orw) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    int r;
    if ((r = copy_out(h, un, h->un, h->ng)) || (r = copy_out(h, vn, h->vn, h->ng)) || (r = copy_out(h, wn, h->wn, h->ng)) ||
        (r = copy_out(h, pn, h->pn, h->ng)) || (r = copy_out(h, us, h->us, h->ng)) || (r = copy_out(h, vs, h->vs, h->ng)) ||
        (r = copy_out(h, ws, h->ws, h->ng)) || (r = copy_out(h, oru, h->oru, h->ni)) ||
        (r = copy_out(h, orv, h->orv, h->ni)) || (r = copy_out(h, orw, h->orw, h->ni)))
        return r;
    MDX_HIP(hipStreamSynchronize(h->stream));
    return MDX_OK;
}

// ----------------------------------------------------------------------------------- IBM
// captured sweep batches hold the mu-table pointers: drop them when the tables change
static void invalidate_graphs(mdx_handle *h) {
    for (auto &s : h->graphs) {
        if (s.exec) (void)hipGraphExecDestroy(s.exec);
        s = mdx_handle::GraphSlot{};
    }
    h->rb_state = 0;   // the LDS solver's plan depends on the IBM state too
}

static int upload_ibm(mdx_handle *h) {
    const size_t ng = h->ng;
    const int sx = h->d.sx, sy = h->d.sy;
    const size_t plane = (size_t)sx * sy;
    // z-invariant geometry: every k plane equals the first
    bool inv = true;
    for (int k = 1; k < h->d.nz + 2 && inv; ++k)
        inv = std::memcmp(h->h_cu.data(), h->h_cu.data() + k * plane, plane * 8) == 0 &&
              std::memcmp(h->h_cv.data(), h->h_cv.data() + k * plane, plane * 8) == 0 &&
              std::memcmp(h->h_cw.data(), h->h_cw.data() + k * plane, plane * 8) == 0;
    h->coef_plane = inv;
    if (!h->cu) {
        MDX_HIP(dalloc(&h->cu, ng));
        MDX_HIP(dalloc(&h->cv, ng));
        MDX_HIP(dalloc(&h->cw, ng));
    }
    MDX_HIP(hipMemcpy(h->cu, h->h_cu.data(), ng * 8, hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(h->cv, h->h_cv.data(), ng * 8, hipMemcpyHostToDevice));
    MDX_HIP(hipMemcpy(h->cw, h->h_cw.data(), ng * 8, hipMemcpyHostToDevice));
    if (h->cfg.solver == MDX_SOLVER_SOR) {
        for (int q = 0; q < 3; ++q) {
            if (h->mu64[q]) { (void)hipFree(h->mu64[q]); h->mu64[q] = nullptr; }
            if (h->mu32[q]) { (void)hipFree(h->mu32[q]); h->mu32[q] = nullptr; }
            size_t n = inv ? plane : ng;
            MDX_HIP(dalloc(&h->mu64[q], n));
            MDX_HIP(dalloc(&h->mu32[q], n));
        }
    }
    h->mu_dtg = -1.0;
    h->ibm_on = true;
    invalidate_graphs(h);
    return MDX_OK;
}

extern "C" int mdx_set_ibm(mdx_handle *h, const double *cu, const double *cv, const double *cw) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_HIP(hipStreamSynchronize(h->stream));
    if (!cu && !cv && !cw) {
        h->ibm_on = false;
        h->mu_dtg = -1.0;
        invalidate_graphs(h);
        return MDX_OK;
    }
    MDX_CHECK(cu && cv && cw, MDX_ERR_ARG, "mdx_set_ibm: give all three coefficient arrays or none");
    h->h_cu.assign(cu, cu + h->ng);
    h->h_cv.assign(cv, cv + h->ng);
    h->h_cw.assign(cw, cw + h->ng);
    return upload_ibm(h);
}

// ibm.f90:84-99 (kind 0) and the two-sided wavy channel (kind 1)
static bool in_body(const mdx_handle *h, const mdx_geometry *geo, const double x[3]) {
    if (geo->kind == 0) {
        double amp = geo->amp_x > 0.0 ? geo->amp_x : 5 * h->dy;
        int nw = geo->n_wave_x > 0 ? geo->n_wave_x : 1;
        double y_body = amp * 0.5 * (1.0 + std::sin(2.0 * PI * (double)nw * x[0] / h->cfg.lx + geo->phase_x)) + 2 * h->dy;
        return x[1] < y_body;
    }
    double hh = geo->half_height - geo->A * std::cos(2.0 * PI * x[0] / geo->lambda);
    return std::fabs(x[1] - 0.5 * h->cfg.ly) > hh;
}

// ibm.f90:101-130: bisection to the interface along the segment xa -> xb
static void bisection(const mdx_handle *h, const mdx_geometry *geo, const double xa_in[3], double xb[3]) {
    double xa[3] = {xa_in[0], xa_in[1], xa_in[2]}, xm[3] = {0, 0, 0};
    for (int it = 1; it <= 200; ++it) {
        for (int q = 0; q < 3; ++q) xm[q] = 0.5 * (xa[q] + xb[q]);
        double e0 = xb[0] - xa[0], e1 = xb[1] - xa[1], e2 = xb[2] - xa[2];
        if (std::sqrt(e0 * e0 + e1 * e1 + e2 * e2) < 1.0e-10) break;
        bool la = in_body(h, geo, xa), lm = in_body(h, geo, xm);
        if (lm == la) std::memcpy(xa, xm, sizeof xa);
        else std::memcpy(xb, xm, sizeof xa);
    }
    std::memcpy(xb, xm, sizeof xm);
}

// ibm.f90:133-187 over the ghosted array; a z-invariant geometry is evaluated on one k plane
static void ibm_coeff(const mdx_handle *h, const mdx_geometry *geo, int dix, int diy, int diz,
                      std::vector<double> &coeff) {
    const int nx = h->d.nx, ny = h->d.ny, nz = h->d.nz;
    static const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
    coeff.assign(h->ng, 0.0);
    const bool zinv = true;   // both geometries depend on x and y only
    const int kmax = zinv ? 1 : nz + 2;
    for (int iz = 1; iz <= kmax; ++iz)

        for (int iy = 1; iy <= ny + 2; ++iy)
            for (int ix = 1; ix <= nx + 2; ++ix) {
                double xa[3] = {((double)ix - (double)dix * 0.5 - 1.5) * h->dx,
                                ((double)iy - (double)diy * 0.5 - 1.5) * h->dy,
                                ((double)iz - (double)diz * 0.5 - 1.5) * h->dz};
                double acc = 0.0;
                if (in_body(h, geo, xa)) {
                    acc = 1.0e30;
                } else if (geo->second_order) {
                    for (int q = 0; q < 6; ++q) {
                        double xb[3] = {((double)(ix + nb[q][0]) - (double)dix * 0.5 - 1.5) * h->dx,
                                        ((double)(iy + nb[q][1]) - (double)diy * 0.5 - 1.5) * h->dy,
                                        ((double)(iz + nb[q][2]) - (double)diz * 0.5 - 1.5) * h->dz};
                        double e0 = xb[0] - xa[0], e1 = xb[1] - xa[1], e2 = xb[2] - xa[2];
                        double d0 = std::sqrt(e0 * e0 + e1 * e1 + e2 * e2);
                        if (in_body(h, geo, xb)) {
                            bisection(h, geo, xa, xb);
                            double f0 = xb[0] - xa[0], f1 = xb[1] - xa[1], f2 = xb[2] - xa[2];
                            double dd = std::sqrt(f0 * f0 + f1 * f1 + f2 * f2);
                            acc = acc + ((d0 - dd) / dd) / (d0 * d0);
                        }
                    }
                }
                coeff[(size_t)(ix - 1) + (size_t)(nx + 2) * ((size_t)(iy - 1) + (size_t)(ny + 2) * (iz - 1))] =
                    acc / h->cfg.re;
            }
    if (zinv) {
        const size_t plane = (size_t)(nx + 2) * (ny + 2);
        for (int k = 1; k < nz + 2; ++k) std::copy(coeff.begin(), coeff.begin() + plane, coeff.begin() + k * plane);
    }
}

extern "C" int mdx_ibm_coefficient(mdx_handle *h, const mdx_geometry *geo, int32_t dix, int32_t diy, int32_t diz,
                                   double *coeff) {
    MDX_CHECK(h && geo && coeff, MDX_ERR_ARG, "mdx_ibm_coefficient: null argument");
    MDX_CHECK(geo->kind == 0 || geo->kind == 1, MDX_ERR_ARG, "mdx_ibm_coefficient: unknown geometry kind");
    MDX_CHECK(!h->stretched, MDX_ERR_UNSUPPORTED, "IBM geometry requires the uniform wall-normal grid");
    std::vector<double> c;
    ibm_coeff(h, geo, dix, diy, diz, c);
    std::memcpy(coeff, c.data(), h->ng * 8);
    return MDX_OK;
}

extern "C" int mdx_compute_ibm(mdx_handle *h, const mdx_geometry *geo, double *cu, double *cv, double *cw) {
    MDX_CHECK(h && geo, MDX_ERR_ARG, "mdx_compute_ibm: null argument");
    MDX_CHECK(geo->kind == 0 || geo->kind == 1, MDX_ERR_ARG, "mdx_compute_ibm: unknown geometry kind");
    MDX_CHECK(!h->stretched, MDX_ERR_UNSUPPORTED, "IBM geometry requires the uniform wall-normal grid");
    if (geo->kind == 1)
        MDX_CHECK(geo->lambda > 0.0 && geo->half_height > 0.0, MDX_ERR_ARG, "wavy geometry: lambda, half_height > 0");
    MDX_HIP(hipStreamSynchronize(h->stream));
    ibm_coeff(h, geo, 1, 0, 0, h->h_cu);
    ibm_coeff(h, geo, 0, 1, 0, h->h_cv);
    ibm_coeff(h, geo, 0, 0, 1, h->h_cw);
    if (cu) std::memcpy(cu, h->h_cu.data(), h->ng * 8);
    if (cv) std::memcpy(cv, h->h_cv.data(), h->ng * 8);
    if (cw) std::memcpy(cw, h->h_cw.data(), h->ng * 8);
    return upload_ibm(h);
}

// ----------------------------------------------------------------------------------- operators
extern "C" int mdx_momentum(mdx_handle *h, double dta, double dtb, double dtg) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_HIP(launch_momentum(h, dta, dtb, dtg));
    return MDX_OK;
}

// Device time of the momentum kernel: `reps` launches of mdx_momentum(dta, dtb, dtg) from the
// current fields between two HIP events on the handle's stream, after one untimed launch.  The old
// right-hand sides are saved before and restored after, so the solver's RK3 history is unchanged;
// the predictor fields hold the last launch's result.
extern "C" int mdx_time_momentum(mdx_handle *h, double dta, double dtb, double dtg, int32_t reps,
                                 float *ms_per_launch) {
    MDX_CHECK(h && ms_per_launch && reps > 0, MDX_ERR_ARG, "mdx_time_momentum: bad argument");
    const size_t bytes = h->ni * sizeof(double);
    double *save = nullptr;
    MDX_HIP(hipMalloc(&save, 3 * bytes));
    double *const old[3] = {h->oru, h->orv, h->orw};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipSuccess;
    for (int c = 0; c < 3 && e == hipSuccess; ++c)
        e = hipMemcpyAsync(save + c * h->ni, old[c], bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = launch_momentum(h, dta, dtb, dtg);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_momentum(h, dta, dtb, dtg);
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    for (int c = 0; c < 3 && e == hipSuccess; ++c)
        e = hipMemcpyAsync(old[c], save + c * h->ni, bytes, hipMemcpyDeviceToDevice, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(save);
    MDX_HIP(e);
    *ms_per_launch = ms / (float)reps;
    return MDX_OK;
}

extern "C" int mdx_apply_bc(mdx_handle *h) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_HIP(launch_apply_bc_fft(h));
    return MDX_OK;
}

extern "C" int mdx_apply_bc_redblack(mdx_handle *h) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_HIP(launch_apply_bc_redblack(h));
    return MDX_OK;
}

// Internal status (never returned through the C ABI): a neighbour wait of the LDS-resident solver
// gave up.  The batch drained, but its iterate and residual are not trustworthy; the SOR solve that
// saw it restarts on the streaming sweeps.
static constexpr int MDX_LDS_TIMEOUT = 100;
// restarts after which the LDS-resident solver stays off for the handle
static constexpr int LDS_TIMEOUT_MAX = 3;

// The wait-timeout word has been seen set: clear it on the device (it is set again only by a later
// wait that gives up) and report the timeout once.
static int lds_timeout_seen(mdx_handle *h) {
    MDX_HIP(hipStreamSynchronize(h->stream));
    MDX_HIP(hipMemsetAsync(h->red_max + 1, 0, sizeof(unsigned long long), h->stream));
    MDX_HIP(hipStreamSynchronize(h->stream));
    mdx_set_error("LDS red-black solver: a neighbour wait exceeded its spin limit");
    return MDX_LDS_TIMEOUT;
}

// Reads the max reduction (red_max[0]) and the LDS solver's wait-timeout word (red_max[1]) in
// one copy.  A set timeout word is cleared and reported as MDX_LDS_TIMEOUT; the maximum itself was
// formed by a kernel of its own and is valid.
static int read_max(mdx_handle *h, double *out) {
    MDX_HIP(hipMemcpyAsync(h->h_pinned, h->red_max, 16, hipMemcpyDeviceToHost, h->stream));
    MDX_HIP(hipStreamSynchronize(h->stream));
    *out = h->h_pinned[0];
    int err;
    std::memcpy(&err, &h->h_pinned[1], sizeof err);
    if (err) return lds_timeout_seen(h);
    return MDX_OK;
}

// read_max for the norms: a timeout word left by a batch whose check was never read (discarded
// after convergence) does not concern them
static int read_norm(mdx_handle *h, double *out) {
    const int rc = read_max(h, out);
    return rc == MDX_LDS_TIMEOUT ? MDX_OK : rc;
}

// Plan and allocate the LDS-resident solver once the IBM state is known (uniform grid, z-invariant
// or absent IBM coefficients); rb_state = -1 leaves the streaming sweeps in charge.
static int rb_prepare(mdx_handle *h) {
    if (h->rb_state != 0) return MDX_OK;
    h->rb_state = -1;
    if (h->lds_sor_mode == 0) {
        h->rb_note = "streaming sweeps: z-marching pairs, " + std::to_string(h->sweep_kz) + " planes per thread";
        if (h->sweep_kz == 0) h->rb_note = "streaming sweeps: one cell per thread";
        if (h->sweep_fused) h->rb_note = "fused red+black iterations: 32x8x8 tiles with a two-cell halo in LDS";
        return MDX_OK;
    }
    if (h->stretched) { h->rb_note = "streaming sweeps: stretched grid"; return MDX_OK; }
    if (h->ibm_on && !h->coef_plane) { h->rb_note = "streaming sweeps: IBM coefficients vary in z"; return MDX_OK; }
    RbPlanOut p64{}, p32{};
    std::string why;
    if (!rb_plan(h, true, p32, why)) {
        h->rb_note = "streaming sweeps: " + why;
        return MDX_OK;
    }
    std::string note64;
    if (h->lds_sor_mode == 2 && !rb_plan(h, false, p64, why)) {
        h->lds_sor_mode = 1;   // FP64 batches stay on the streaming sweeps
        note64 = "; fp64: streaming (" + why + ")";
    }
    auto desc = [](const RbPlanOut &p) {
        return std::to_string(p.px) + "x" + std::to_string(p.py) + "x" + std::to_string(p.pz) + " workgroups of " +
               std::to_string(p.bx) + "x" + std::to_string(p.by) + "x" + std::to_string(p.bz) + " cells (" +
               (p.maxp == -1 ? "a z-column per thread" : (p.maxp == -2 ? "half a z-column per thread" : "x-pairs")) + "), LDS " + std::to_string(p.lds) + " B";
    };
    h->rb_note = "LDS-resident sweeps (two-deep halo, one exchange per iteration): fp32 " + desc(p32);
    if (h->lds_sor_mode == 2) h->rb_note += "; fp64 " + desc(p64);
    else if (!note64.empty()) h->rb_note += note64;
    else h->rb_note += "; fp64: streaming z-marching pairs, " + std::to_string(h->sweep_kz) + " planes per thread";
    size_t need = rb_faces_bytes(p32, true);
    if (h->lds_sor_mode == 2) need = std::max(need, rb_faces_bytes(p64, false));
    if (h->rb_faces) { (void)hipFree(h->rb_faces); h->rb_faces = nullptr; }
    MDX_HIP(hipMalloc(&h->rb_faces, need));
    MDX_HIP(hipMemset(h->rb_faces, 0, need));   // tag 0 is never published
    h->rb_faces_bytes = need;
    for (int **xl : {&h->rb_xl32, &h->rb_xl64})
        if (*xl) { (void)hipFree(*xl); *xl = nullptr; }
    MDX_HIP(hipMalloc(&h->rb_xl32, rb_xlist_bytes(p32)));
    MDX_HIP(rb_build_xlists(h, p32, h->rb_xl32));
    if (h->lds_sor_mode == 2) {
        MDX_HIP(hipMalloc(&h->rb_xl64, rb_xlist_bytes(p64)));
        MDX_HIP(rb_build_xlists(h, p64, h->rb_xl64));
    }
    h->rb_tag = 0;
    h->rb_plan64 = p64;
    h->rb_plan32 = p32;
    h->rb_state = 1;
    return MDX_OK;
}

extern "C" int mdx_div_norm(mdx_handle *h, double *div_inf) {
    MDX_CHECK(h && div_inf, MDX_ERR_ARG, "null argument");
    MDX_HIP(launch_reset_max(h));
    MDX_HIP(launch_div_max(h, h->un, h->vn, h->wn));
    return read_norm(h, div_inf);
}

extern "C" int mdx_center_velocity(mdx_handle *h, double *uc, double *vc, double *wc) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    if (!h->cvel) MDX_HIP(dalloc(&h->cvel, 3 * h->ni));
    double *uc_d = h->cvel, *vc_d = h->cvel + h->ni, *wc_d = h->cvel + 2 * h->ni;
    MDX_HIP(launch_center_vel(h, uc_d, vc_d, wc_d));
    int r;
    if ((r = copy_out(h, uc, uc_d, h->ni)) || (r = copy_out(h, vc, vc_d, h->ni)) || (r = copy_out(h, wc, wc_d, h->ni)))
        return r;
    MDX_HIP(hipStreamSynchronize(h->stream));
    return MDX_OK;
}

extern "C" int mdx_get_cfl(mdx_handle *h, double *cfl) {
    MDX_CHECK(h && cfl, MDX_ERR_ARG, "null argument");
    MDX_HIP(launch_reset_max(h));
    MDX_HIP(launch_cfl_max(h));
    return read_norm(h, cfl);
}

static int ensure_mu(mdx_handle *h, double dtg) {
    if (h->ibm_on && h->mu_dtg != dtg) {
        MDX_HIP(launch_build_mu(h, dtg));
        h->mu_dtg = dtg;
    }
    return MDX_OK;
}

static void ev_record(mdx_handle *h, int q) {
    const bool phase = q < mdx_handle::PH_EVENTS;
    if (!h->timing || (phase && h->ph_skip)) return;
    (void)hipEventRecord(phase ? h->ev_ph[h->ph_cur][q] : h->ev[q], h->stream);
}

// add the phase times of completed stages (`wait`: of all recorded stages) to the counters
static void phase_account(mdx_handle *h, bool wait) {
    while (h->timing && h->ph_count > 0) {
        hipEvent_t *e = h->ev_ph[h->ph_head];
        if (wait) (void)hipEventSynchronize(e[4]);
        else if (hipEventQuery(e[4]) != hipSuccess) break;
        float a = 0, b = 0, c = 0, d = 0, m = 0;
        (void)hipEventElapsedTime(&a, e[0], e[1]);
        (void)hipEventElapsedTime(&b, e[1], e[2]);
        (void)hipEventElapsedTime(&c, e[2], e[3]);
        (void)hipEventElapsedTime(&d, e[3], e[4]);
        (void)hipEventElapsedTime(&m, e[0], e[5]);
        h->times.stencil_ms += a + d;
        h->times.poisson_ms += b;
        h->times.projection_ms += c;
        h->times.momentum_ms += m;
        h->times.stages += 1;
        h->ph_head = (h->ph_head + 1) % mdx_handle::PHQ;
        h->ph_count -= 1;
    }
}

// A batch of m red+black iterations.  The launch sequence of a batch never changes for a given
// (precision, m, iterate buffer), so it is captured once into a HIP graph and replayed; the mu
// tables are rewritten in place, and a batch is captured again if the iterate buffer changed.
static hipError_t sweep_launches(mdx_handle *h, bool fp32, int m) {
    hipError_t e = hipSuccess;
    if (h->sweep_fused) return fp32 ? launch_sor_iters32(h, m, (float)h->inner_omega) : launch_sor_iters64(h, m, h->cfg.omega);
    for (int q = 0; q < m && e == hipSuccess; ++q) {
        if (fp32) {
            if ((e = launch_sor_sweep32(h, 0, (float)h->inner_omega)) == hipSuccess)
                e = launch_sor_sweep32(h, 1, (float)h->inner_omega);
        } else {
            if ((e = launch_sor_sweep64(h, 0, h->cfg.omega)) == hipSuccess)
                e = launch_sor_sweep64(h, 1, h->cfg.omega);
        }
    }
    return e;
}

// whether batches of this precision run on the LDS-resident solver
static bool lds_for(const mdx_handle *h, bool fp32) {
    return h->rb_state == 1 && !h->lds_suspended && (fp32 || h->lds_sor_mode == 2);
}

// sweep counters: account finished batches in order (end event complete; `wait`: block for them)
static void sweep_account(mdx_handle *h, bool wait = false) {
    while (h->timing && h->sw_count > 0) {
        const int q = h->sw_head;
        if (wait) (void)hipEventSynchronize(h->ev_sw[q][1]);
        else if (hipEventQuery(h->ev_sw[q][1]) != hipSuccess) break;
        float ms = 0.f;
        if (h->sw_void[q]) {
            h->sw_void[q] = false;   // a batch that found its input converged: nothing to count
        } else if (hipEventElapsedTime(&ms, h->ev_sw[q][0], h->ev_sw[q][1]) == hipSuccess) {
            h->times.sweep_ms += ms;
            h->times.sweep_launches += h->sw_launches[q];
            h->times.sweep_cells += 2 * (int64_t)h->sw_its[q] * (int64_t)h->ni;
        }
        h->sw_head = (q + 1) % mdx_handle::SW_Q;
        h->sw_count -= 1;
    }
}

static int sweep_batch(mdx_handle *h, bool fp32, int m, const RbDefect *df = nullptr) {
    int qs = -1;   // accounting queue entry of this batch (every SW_EVERY-th batch is timed)
    if (h->timing && h->sw_seen++ % mdx_handle::SW_EVERY == 0) {
        if (h->sw_count == mdx_handle::SW_Q) sweep_account(h, true);
        qs = (h->sw_head + h->sw_count) % mdx_handle::SW_Q;
        MDX_HIP(hipEventRecord(h->ev_sw[qs][0], h->stream));
    }
    int launches = 0;
    auto &slot = h->graphs[fp32 ? 1 : 0];
    if (lds_for(h, fp32)) {
        // LDS-resident sweeps; the FP32 defect solve starts from zero inside the kernel
        hipError_t e = launch_rb_lds(h, fp32, m, fp32 ? 1 : 0, df);
        launches = 1;
        if (e != hipSuccess) {
            // the runtime refused the launch (nothing ran): use the streaming sweeps
            (void)hipGetLastError();
            launches = 0;
            h->rb_state = -1;
            h->rb_note = std::string("streaming sweeps: LDS-resident launch refused (") + hipGetErrorString(e) + ")";
            if (fp32) MDX_HIP(launch_zero32(h, h->e32, h->ni));   // the streaming sweeps start from e = 0
        }
    }
    if (lds_for(h, fp32)) {
    } else if (h->use_graphs && m >= 2 && m == (fp32 ? h->cfg.inner_iters : h->cfg.check_every)) {
        const void *iterate = fp32 ? (const void *)h->e32 : (const void *)h->phi;
        if (!slot.exec || slot.m != m || slot.ibm != h->ibm_on || slot.iterate != iterate) {
            if (slot.exec) { (void)hipGraphExecDestroy(slot.exec); slot.exec = nullptr; }
            hipGraph_t graph = nullptr;
            MDX_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = sweep_launches(h, fp32, m);
            hipError_t e2 = hipStreamEndCapture(h->stream, &graph);
            MDX_HIP(e);
            MDX_HIP(e2);
            hipError_t e3 = hipGraphInstantiate(&slot.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            MDX_HIP(e3);
            slot.m = m;
            slot.ibm = h->ibm_on;
            slot.iterate = iterate;
        }
        MDX_HIP(hipGraphLaunch(slot.exec, h->stream));
    } else {
        MDX_HIP(sweep_launches(h, fp32, m));
    }
    if (!launches) launches = (h->sweep_fused ? 1 : 2) * m;   // streaming: one launch per colour
    // a batch that should have ended with the LDS defect epilogue but ran on the streaming sweeps
    // (launch refused): the separate defect kernel forms the same check
    if (df && !lds_for(h, fp32))
        MDX_HIP(launch_defect64(h, df->phi_in, h->e32, df->phi_out, h->r32, df->host_slot, df->seq));
    h->sw_last = qs;
    if (qs >= 0) {
        MDX_HIP(hipEventRecord(h->ev_sw[qs][1], h->stream));
        h->sw_its[qs] = m;
        h->sw_launches[qs] = launches;
        h->sw_count += 1;
    }
    return MDX_OK;
}

// Plain FP64 red-black SOR to ||div||_inf < tol; Phi starts at zero (the previous-step pressure).
// r_first >= 0: the residual of the initial iterate is already known (learned-guess selection).
static int sor_solve64(mdx_handle *h, mdx_solve_stats *st, double r_first) {
    const mdx_config &c = h->cfg;
    double r = 0.0;
    int it = 0, checks = 0;
    for (;;) {
        int rc;
        if (checks == 0 && r_first >= 0.0) {
            r = r_first;
        } else {
            MDX_HIP(launch_reset_max(h));
            MDX_HIP(launch_sor_residual64(h, h->phi));
            if ((rc = read_max(h, &r))) return rc;
        }
        sweep_account(h);
        if (checks == 0) st->div0_inf = r;
        ++checks;
        if (!(r >= c.tol) || it >= c.max_iter) break;   // NaN stops too and is reported below
        const int m = std::min(c.check_every, c.max_iter - it);
        if ((rc = sweep_batch(h, false, m))) return rc;
        it += m;
    }
    st->iterations = it;
    st->checks = checks;
    st->div_inf = r;
    st->converged = r < c.tol;
    st->lds_sweeps = lds_for(h, false);
    return MDX_OK;
}

// pinned words of the mixed solve's two check slots: [CHECK_PIN + 4 * slot] = max, error word, seq
static constexpr int CHECK_PIN = 16;
// pinned words of the learned start's comparison, each [max, error word, seq]: the residual of
// Phi_0 (RP), max |rhs| (RZ: the residual of Phi = 0) and the learned start's first check (SEL)
static constexpr int RP_PIN = 32;
static constexpr int RZ_PIN = 40;
static constexpr int SEL_PIN = 48;

// Wait for the residual check that posts sequence number `seq` to the pinned words [w, w + 3):
// [max, error word, seq].  The host polls the sequence word (no event between the kernels); now
// and then it queries the stream, so a stream that finished or failed without posting ends the
// wait with an error instead of spinning.
static int wait_check(mdx_handle *h, int w, unsigned long long seq, double *r) {
    volatile unsigned long long *sq = reinterpret_cast<volatile unsigned long long *>(h->h_pinned + w + 2);
    for (unsigned spins = 1; *sq != seq; ++spins) {
        if ((spins & 255) == 0) {
            const hipError_t q = hipStreamQuery(h->stream);
            if (q == hipErrorNotReady) continue;
            if (*sq == seq) break;
            if (q != hipSuccess) {
                mdx_set_error(std::string("SOR check: ") + hipGetErrorString(q));
                return MDX_ERR_HIP;
            }
            mdx_set_error("SOR check: the stream finished without posting its residual");
            return MDX_ERR_HIP;
        }
        _mm_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    *r = h->h_pinned[w];
    int err;
    std::memcpy(&err, &h->h_pinned[w + 1], sizeof err);
    if (err) return lds_timeout_seen(h);
    return MDX_OK;
}

// FP32 red-black sweeps on the defect equation inside an FP64 defect-correction loop.
//
// A check k forms Phi += e and the FP64 residual r_k; the FP64 iterate changes only there, so the
// next batch of FP32 sweeps can be queued before r_k is read: if r_k turns out converged, that
// batch's correction is simply never applied, and Phi, the iteration count and every check are
// what the synchronous loop gives.
// On the LDS-resident solver a batch ends with the next check itself (its defect epilogue), and a
// batch queued ahead reads r_k on the GPU and does nothing when it is converged, so the next batch
// is always queued before r_k is read and the GPU never waits for the host inside a solve.  On the
// streaming sweeps a wasted batch costs its full time: it is queued early at the first check of a
// solve, and later only when the residuals seen so far predict (geometric decay) that check k is
// still at least SPEC_MARGIN times above the tolerance.
// first_seq != 0: the first check is already queued in slot 0 with that sequence number (the fused
// start kernel).  key: the RK stage + 3 * the kind of start (0 Phi = 0, 1 Phi_0, 2 the network's):
// the previous solve with the same key gives the merge prior (none for network starts).
// on_first: called once the first check has been read (a learned start whose choice was deferred:
// it learns which start ran and may change the key).
static int sor_solve_mixed(mdx_handle *h, mdx_solve_stats *st, unsigned long long first_seq, int key,
                           const std::function<int(int &)> &on_first = nullptr) {
    constexpr double SPEC_MARGIN = 4.0, SPEC_FIRST = 1e3;
    const int MERGE_MAX = h->merge_max;   // at most this many inner_iters batches run as one
    constexpr int PIN = CHECK_PIN;   // pinned words of the two check slots: [PIN + 4 * slot]
    const mdx_config &c = h->cfg;
    const bool lds = lds_for(h, true);
    const int prior = h->mix_prev_its[key];
    double r = 0.0, r1 = -1.0, r2 = -1.0;   // r_k, r_{k-1}, r_{k-2}
    int it = 0, checks = 0, slot = 0;
    unsigned long long seq[2] = {0, 0};     // the check each slot waits for
    auto slot_dev = [&](int sl) { return h->d_pinned + PIN + 4 * sl; };
    // separate defect kernel: Phi += e, r32, max; the first check of a solve has no e (Phi += 0.0,
    // in place)
    auto enqueue_check = [&](int sl, bool first) -> int {
        seq[sl] = ++h->chk_seq;
        if (first) {
            MDX_HIP(launch_defect64(h, h->phi, nullptr, h->phi, h->r32, slot_dev(sl), seq[sl]));
            return MDX_OK;
        }
        MDX_HIP(launch_defect64(h, h->phi, h->e32, h->phi2, h->r32, slot_dev(sl), seq[sl]));
        std::swap(h->phi, h->phi2);
        return MDX_OK;
    };
    // m iterations; on LDS also the next check (`ahead`: queued before the previous check is read)
    auto enqueue_batch = [&](int m, int sl, bool ahead) -> int {
        if (lds) {
            seq[sl] = ++h->chk_seq;
            const RbDefect df{h->phi, h->phi2, slot_dev(sl), seq[sl], ahead ? 1 : 0, c.tol};
            int rc_ = sweep_batch(h, true, m, &df);
            if (rc_) return rc_;
            std::swap(h->phi, h->phi2);
            return MDX_OK;
        }
        MDX_HIP(launch_zero32(h, h->e32, h->ni));
        return sweep_batch(h, true, m);
    };
    int rc;
    if (first_seq) seq[slot] = first_seq;
    else if ((rc = enqueue_check(slot, true))) return rc;
    int m_a = 0, m_b = 0;   // iterations between the checks of r2 and r1, and of r1 and the pending one
    for (;;) {
        // Batch size: inner_iters, or two or three times that while the residual decay seen so far
        // (per iteration, from r2 to r1) predicts that many more checks would still fail after the
        // pending one: checks that cannot stop the solve are skipped, saving their launches and
        // FP64 corrections (the FP32 inner solve then simply runs longer on the same defect
        // equation).  Before two residuals have been read, the iterations of the previous solve of
        // this stage stand in for the prediction.  The decision uses only residuals and iteration
        // counts, so every path takes the same batches.
        int m = c.inner_iters;
        if (r1 > 0.0 && r2 > 0.0 && r1 < r2 && m_a > 0) {
            const double q = std::pow(r1 / r2, 1.0 / m_a);   // decay per iteration
            const double rp = r1 * std::pow(q, m_b);           // predicted pending residual
            if (rp > c.tol) {
                const int blocks = (int)(std::log(c.tol / rp) / std::log(q) / c.inner_iters);
                if (blocks >= 3 && MERGE_MAX > 1) m = c.inner_iters * std::min(MERGE_MAX, blocks - 1);
            }
        } else if (r2 < 0.0 && prior > it && MERGE_MAX > 1 && key < 6) {
            // (no prior for network starts: their solves vary too much from one step to the next)
            const int blocks = (prior - it) / c.inner_iters;   // after the pending check
            if (blocks >= 3) m = c.inner_iters * std::min(MERGE_MAX, blocks - 1);
        }
        m = std::min(m, c.max_iter - it);
        // the first check of a solve is taken to be unconverged (a fresh right-hand side); later
        // ones are predicted from the residuals seen so far
        bool spec = m > 0 && (lds || checks == 0);
        if (!lds && m > 0 && r1 > 0.0) {
            const double pred = r2 > 0.0 ? r1 * std::min(1.0, r1 / r2) : r1 / SPEC_FIRST;
            spec = pred > SPEC_MARGIN * c.tol;
        }
        if (spec && (rc = enqueue_batch(m, slot ^ 1, true))) return rc;
        if ((rc = wait_check(h, PIN + 4 * slot, seq[slot], &r))) return rc;
        const bool done = !(r >= c.tol) || it >= c.max_iter;   // NaN stops too and is reported
        // a batch queued ahead of a converged check did nothing on the LDS solver: not counted
        if (done && spec && lds && lds_for(h, true) && h->sw_last >= 0) h->sw_void[h->sw_last] = true;
        sweep_account(h);
        if (checks == 0) {
            st->div0_inf = r;
            if (on_first && (rc = on_first(key))) return rc;
        }
        ++checks;
        if (done) {
            // a batch queued ahead is discarded: on LDS its check would have written the other
            // Phi buffer
            if (spec && lds) std::swap(h->phi, h->phi2);
            break;
        }
        if (!spec && (rc = enqueue_batch(m, slot ^ 1, false))) return rc;
        it += m;
        r2 = r1;
        r1 = r;
        m_a = m_b;
        m_b = m;
        if (!lds && (rc = enqueue_check(slot ^ 1, false))) return rc;
        slot ^= 1;
    }
    st->iterations = it;
    st->checks = checks;
    st->div_inf = r;
    st->converged = r < c.tol;
    st->lds_sweeps = lds;
    h->mix_prev_its[key] = st->converged ? it : 0;
    return MDX_OK;
}

// Learned-guess gate bookkeeping for stage s.  Two losses in a row make the network's start skip
// 4 solves, then 16, 64 and NET_SKIP_MAX between evaluations; a win resets the count.
static constexpr int NET_SKIP_MAX = 256;
static void net_loss(mdx_handle *h, int s) {
    h->net_losses[s] += 1;
    if (h->net_losses[s] >= 2) h->net_skip[s] = std::min(NET_SKIP_MAX, 1 << std::min(2 * (h->net_losses[s] - 1), 8));
}

// After the solve of stage s: a network start wins when the network's evaluation and the solve it
// started took at most NET_TIME of the host time of the last solve of this stage that started
// otherwise (each measured from the plan of the start to the converged check, so both include the
// stage's queued work once): sweeps saved count only when they pay for the inference.
static void net_outcome(mdx_handle *h, int s, double seconds, bool evaluated) {
    constexpr double NET_TIME = 0.9;
    if (!h->net_started[s]) {
        // the yardstick is a solve that ran no inference: a network evaluated and then rejected by
        // the residual test would put its inference time into the plain start's time
        if (!evaluated) h->base_sec[s] = seconds;
        return;
    }
    if (h->base_sec[s] > 0.0 && seconds <= NET_TIME * h->base_sec[s]) h->net_losses[s] = 0;
    else net_loss(h, s);
}

// Initial guess of the SOR solve for stage s; the solve itself, and its stopping rule, are the
// same whatever starts it.  Starts:
//   Phi = 0 (the previous pressure: the plain method, and the start without a history);
//   Phi_0 = fac * Phi of stage s at the previous step (MDX_GUESS_PHI_PREV);
//   Phi_0 + the network's correction, predicted from (rhs - L_mu Phi_0, Phi_0) (MDX_GUESS_LEARNED),
//   taken when its FP64 residual max-norm is below net_margin times that of Phi = 0.
// In learned mode a start other than the network's is a fallback (stats.fallback) to Phi = 0.
// With capture on, the network input is built for every solve so the converged Phi can be stored
// as a training pair.
// The start of stage s's solve, decided before any kernel runs: `simple` when no network input is
// formed (no learned start to evaluate this time, no capture), the start then being fac * prev
// (prev non-null) or zero.  Updates the gate's state and the solve's guess flags.
struct GuessPlan {
    bool simple, learned, build_input, have_prev, use_prev;
    const double *prev;
    double fac;
};
static GuessPlan guess_plan(mdx_handle *h, int s, double dtg, mdx_solve_stats *st) {
    GuessPlan p{};
    mdx_net *n = h->net;
    p.have_prev = h->hist_dtg[s] > 0.0;
    p.fac = p.have_prev ? dtg / h->hist_dtg[s] : 0.0;
    const int mode = h->cfg.guess;
    // the network completes Phi_0: without a history of the stage there is nothing to complete,
    // and the solve starts from the previous pressure
    p.learned = n && mode == MDX_GUESS_LEARNED && n->c0 > 0.0 && p.have_prev;
    h->net_started[s] = false;
    // gate: after the network's start has lost twice in a row for this stage, it is evaluated
    // again only after 4, 16, 64, then every NET_SKIP_MAX skipped solves (previous-Phi start then)
    if (p.learned && h->net_time_gate && h->net_skip[s] > 0) {
        h->net_skip[s] -= 1;
        p.learned = false;
        st->guess_skipped = 1;
        st->fallback = 1;
    }
    // a network start is judged against the host time of a plain solve of the same stage: until
    // one has been timed, the stage runs the plain start
    if (p.learned && h->net_time_gate && h->base_sec[s] <= 0.0) {
        p.learned = false;
        st->guess_skipped = 1;
        st->fallback = 1;
    }
    // the learned mode falls back to the plain start (Phi = 0: the previous pressure), where Phi_0
    // serves only as the network's input and the base of its correction, or with
    // net_fallback_prev to the previous-step guess Phi_0
    p.use_prev = p.have_prev && (mode == MDX_GUESS_PHI_PREV || (mode == MDX_GUESS_LEARNED && h->net_fallback_prev));
    // in the learned mode every solve the network does not start is a fallback
    if (n && mode == MDX_GUESS_LEARNED && !p.learned) st->fallback = 1;
    p.build_input = n && (p.learned || h->net_capture);
    p.simple = !p.learned && !p.build_input;
    p.prev = p.use_prev ? h->phi_hist + (size_t)s * h->ni : nullptr;
    return p;
}

// The initial guess after sor_prepare (rhs, Phi = 0): the previous-Phi copy of a simple start, or
// the network's input, its start and the gate's choice.
static int initial_guess(mdx_handle *h, int s, double dtg, const GuessPlan &gp, mdx_solve_stats *st, double *r_first) {
    mdx_net *n = h->net;
    const bool have_prev = gp.have_prev, learned = gp.learned, use_prev = gp.use_prev;
    const double *prev = have_prev ? h->phi_hist + (size_t)s * h->ni : nullptr;
    const double fac = gp.fac;
    if (gp.simple) {
        // previous-Phi start (or the plain one without a history); the solver's first residual
        // check evaluates it, so no candidate residual is formed here
        if (use_prev) MDX_HIP(launch_scale_copy64(h, h->phi, prev, fac, h->ni));
        return MDX_OK;
    }
    (void)dtg;
    int rc;
    ev_record(h, 8);
    // Phi_0 in phi2 and its residual field (the network's input)
    if (have_prev) MDX_HIP(launch_scale_copy64(h, h->phi2, prev, fac, h->ni));
    // the residual of Phi_0 (the network's input field and its max-norm rp); rp goes to pinned
    // host memory behind the network's kernels and is read after the stream's next synchronisation
    double rp = -1.0;
    MDX_HIP(launch_reset_max(h));
    MDX_HIP(launch_residual_field64(h, have_prev ? h->phi2 : h->phi, h->res64));
    MDX_HIP(hipMemcpyAsync(h->h_pinned + RP_PIN, h->red_max, 16, hipMemcpyDeviceToHost, h->stream));
    auto take_rp = [&]() -> int {
        rp = h->h_pinned[RP_PIN];
        int err;
        std::memcpy(&err, &h->h_pinned[RP_PIN + 1], sizeof err);
        return err ? lds_timeout_seen(h) : MDX_OK;
    };
    {
        hipStream_t saved = n->stream;
        n->stream = h->stream;
        hipError_t e = unet::net_input_from_solver(n, h, h->res64, have_prev ? h->phi2 : nullptr);
        // the learned guess is formed in phi2 (in place over Phi_0, or from zero)
        if (e == hipSuccess && learned) e = unet::net_guess_to_phi(n, have_prev ? h->phi2 : nullptr, h->phi2);
        n->stream = saved;
        MDX_HIP(e);
    }
    double rg = -1.0;
    if (learned) {
        MDX_HIP(launch_reset_max(h));
        MDX_HIP(launch_sor_residual64(h, h->phi2));
        if ((rc = read_max(h, &rg))) return rc;   // synchronises the stream: rp is there too
    } else {
        MDX_HIP(hipStreamSynchronize(h->stream));
    }
    if ((rc = take_rp())) return rc;
    ev_record(h, 9);
    if (h->timing && learned) {
        float ms = 0.f;
        MDX_HIP(hipEventSynchronize(h->ev[9]));
        if (hipEventElapsedTime(&ms, h->ev[8], h->ev[9]) == hipSuccess) h->times.guess_ms += ms;
    }
    // rp is the residual of Phi_0 (with a history), else of the previous pressure (Phi = 0); the
    // learned mode falls back to Phi = 0 (rz below), or with net_fallback_prev to Phi_0
    const bool prev_start = use_prev;
    if (!learned) {   // capture only
        if (prev_start) {
            MDX_HIP(hipMemcpyAsync(h->phi, h->phi2, h->ni * 8, hipMemcpyDeviceToDevice, h->stream));
            *r_first = rp;
        } else if (!have_prev) {
            *r_first = rp;   // Phi = 0 was evaluated
        }
        return MDX_OK;
    }
    st->guess_div_inf = rg;
    // the fallback start Phi = 0 with a history: its residual, max |rhs| (sor_prepare wrote it to
    // the pinned slot; the stream has synchronised since)
    double rz = -1.0;
    if (have_prev && !prev_start) rz = h->h_pinned[RZ_PIN];
    // the network's start is taken only when it clearly beats the other start (residual below
    // net_margin times it): a marginally smaller start residual saves no sweeps but costs an
    // inference per solve.  A start not taken counts as a loss; a start taken is judged by the
    // sweeps its solve needs (net_outcome).
    const double other = prev_start || !have_prev ? rp : rz;
    // the chosen start is copied into phi (captured sweep graphs hold that buffer's address)
    if (other < 0.0 || rg < h->net_margin * other) {
        MDX_HIP(hipMemcpyAsync(h->phi, h->phi2, h->ni * 8, hipMemcpyDeviceToDevice, h->stream));
        *r_first = rg;
        h->net_started[s] = true;
        return MDX_OK;
    }
    net_loss(h, s);
    st->fallback = 1;
    if (prev_start) MDX_HIP(launch_scale_copy64(h, h->phi, prev, fac, h->ni));   // phi2 holds the network's
    *r_first = other;
    return MDX_OK;
}

// The learned start of a mixed-precision solve: formed in phi, its first check and the choice
// against the fallback start (Phi = 0, or Phi_0 with net_fallback_prev) made on the device by
// select_start_kernel, so the only host round trip is the wait for that first check, which the
// solve needs anyway (no residual pass of its own, no copy of the start).  *first_seq: the
// sequence number of the first check (slot 0); resolve_learned_start then reads which start won.
// fused: the right-hand side is formed here too (learned_input_kernel, with the network's input at
// its stride); otherwise sor_prepare has formed it.
static int learned_start_mixed(mdx_handle *h, int s, const GuessPlan &gp, bool fused, unsigned long long *first_seq) {
    mdx_net *n = h->net;
    const double *prev = h->phi_hist + (size_t)s * h->ni;   // a learned start has a history
    ev_record(h, 8);
    if (fused) {
        // rhs, Phi_0 and r0 (kept in phi2 / res64 only for a fallback to Phi_0), the voxel sums
        // and the partials in one pass; then the input, its statistics and the RZ / RP maxima in
        // one small kernel, the network, and Phi = fac * prev + guess
        unsigned nb = 0;
        MDX_HIP(launch_learned_input(h, prev, gp.fac, gp.use_prev ? h->phi2 : nullptr,
                                     gp.use_prev ? h->res64 : nullptr, n->vsum, n->stride, &nb));
        hipStream_t saved = n->stream;
        n->stream = h->stream;
        hipError_t e = unet::net_input_from_sums(n, h->max_part, nb, h->d_pinned + RZ_PIN, h->d_pinned + RP_PIN,
                                                 h->red_max + 1);
        if (e == hipSuccess) e = unet::net_guess_to_phi_scaled(n, prev, gp.fac, h->phi);
        n->stream = saved;
        MDX_HIP(e);
    } else {
        // Phi_0 in phi2, its residual (the network's input field) in res64, max-norm to the RP
        // slot, and the network's input statistics, in one pass
        MDX_HIP(launch_phi0_residual(h, prev, gp.fac, h->phi2, h->res64, h->d_pinned + RP_PIN, n->stats));
        hipStream_t saved = n->stream;
        n->stream = h->stream;
        hipError_t e = unet::net_input_from_solver(n, h, h->res64, h->phi2, true);
        if (e == hipSuccess) e = unet::net_guess_to_phi(n, h->phi2, h->phi);
        n->stream = saved;
        MDX_HIP(e);
    }
    ev_record(h, 9);
    unsigned nb = 0;   // the check's workgroup maxima, folded by the select kernel
    MDX_HIP(launch_defect64(h, h->phi, nullptr, h->phi, h->r32, nullptr, 0, &nb));
    const unsigned long long seq = ++h->chk_seq;
    const bool prev_fb = gp.use_prev;
    MDX_HIP(launch_select_start(h, h->net_margin, nb, h->d_pinned + SEL_PIN, h->d_pinned + (prev_fb ? RP_PIN : RZ_PIN),
                                prev_fb ? h->phi2 : nullptr, h->res64, h->d_pinned + CHECK_PIN, seq));
    *first_seq = seq;
    return MDX_OK;
}

// Which start the device chose in learned_start_mixed: the same comparison on the same pinned
// values, read once its first check has posted.  The choice sets the kind of start, and with it
// the solve's batch schedule, before the first batch is queued: a fallback solve is the other
// start's solve exactly.
static int resolve_learned_start(mdx_handle *h, int s, const GuessPlan &gp, mdx_solve_stats *st, double *r_first) {
    const double rg = h->h_pinned[SEL_PIN], other = h->h_pinned[gp.use_prev ? RP_PIN : RZ_PIN];
    st->guess_div_inf = rg;
    if (rg < h->net_margin * other) {
        h->net_started[s] = true;
        *r_first = rg;
    } else {
        net_loss(h, s);
        st->fallback = 1;
        *r_first = other;
    }
    if (h->timing) {
        float ms = 0.f;
        MDX_HIP(hipEventSynchronize(h->ev[9]));
        if (hipEventElapsedTime(&ms, h->ev[8], h->ev[9]) == hipSuccess) h->times.guess_ms += ms;
    }
    return MDX_OK;
}

// Append (network input, converged Phi) of stage s to the attached network's replay buffer.  The
// target scale c0 = rms(Phi - Phi_0) / rms(r0) is fixed by the first captured pair.
static int capture_pair(mdx_handle *h, int s, double dtg) {
    mdx_net *n = h->net;
    const bool have_prev = h->hist_dtg[s] > 0.0;
    const double *prev = have_prev ? h->phi_hist + (size_t)s * h->ni : nullptr;
    const double fac = have_prev ? dtg / h->hist_dtg[s] : 0.0;
    hipStream_t saved = n->stream;
    n->stream = h->stream;
    hipError_t e = hipSuccess;
    if (n->c0 <= 0.0) {
        e = unet::net_rms_to_stats3(n, h, h->phi, prev, fac);
        if (e == hipSuccess) e = hipMemcpyAsync(h->h_pinned + 8, n->stats, 4 * 8, hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e == hipSuccess) {
            const double sr = h->h_pinned[8], se = h->h_pinned[11];
            n->c0 = (se > 0.0 && sr > 0.0) ? se / sr : 1.0;
        }
    }
    if (e == hipSuccess) e = unet::net_capture(n, h->phi, prev, fac);
    n->stream = saved;
    MDX_HIP(e);
    return MDX_OK;
}

extern "C" int mdx_pressure_projection(mdx_handle *h, double dtg, mdx_solve_stats *stats) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_CHECK(dtg > 0.0, MDX_ERR_ARG, "pressure_projection: dt_gamma must be positive");
    mdx_solve_stats st{};
    const mdx_config &c = h->cfg;
    if (c.solver == MDX_SOLVER_FFT && c.precision == MDX_PREC_MIXED) {
        // FP32 transforms and Thomas solves in an FP64 defect correction: from phi = 0, each pass
        // solves L e = rhs - L phi (mean mode in FP64, the rest in FP32) and adds e, until
        // ||div||_inf = dt_gamma ||rhs - L phi||_inf < tol (at most MAX_FFT_PASSES passes)
        constexpr int MAX_FFT_PASSES = 4;
        ev_record(h, 1);
        MDX_HIP(launch_fft_mixed_setup(h, dtg));
        int passes = 0, rc;
        double r = 0.0;
        for (;;) {
            MDX_HIP(launch_reset_max(h));
            MDX_HIP(launch_fft_residual(h));
            if ((rc = read_norm(h, &r))) return rc;
            if (passes == 0) st.div0_inf = r * dtg;
            ++st.checks;
            if (!(r * dtg >= c.tol) || passes >= MAX_FFT_PASSES) break;
            MDX_HIP(launch_fft32_solve_add(h));
            ++passes;
        }
        st.iterations = passes;
        st.div_inf = r * dtg;
        st.converged = r * dtg < c.tol;
        MDX_HIP(launch_fft_pin(h));
        ev_record(h, 2);
        MDX_HIP(launch_fft_correct(h, h->fft_phi, dtg));
        ev_record(h, 3);
    } else if (c.solver == MDX_SOLVER_FFT) {
        ev_record(h, 1);
        MDX_HIP(launch_fft_forward(h, dtg));
        MDX_HIP(launch_tridiag(h));
        ev_record(h, 2);
        MDX_HIP(launch_fft_backward_correct(h, dtg));
        ev_record(h, 3);
        st.converged = 1;
    } else if (c.solver == MDX_SOLVER_SOLA) {
        ev_record(h, 1);
        const size_t bytes = h->ng * 8;
        MDX_HIP(hipMemcpyAsync(h->un, h->us, bytes, hipMemcpyDeviceToDevice, h->stream));
        MDX_HIP(hipMemcpyAsync(h->vn, h->vs, bytes, hipMemcpyDeviceToDevice, h->stream));
        MDX_HIP(hipMemcpyAsync(h->wn, h->ws, bytes, hipMemcpyDeviceToDevice, h->stream));
        for (int it = 0; it < c.niter; ++it)
            for (int parity = 0; parity <= 1; ++parity) MDX_HIP(launch_sola_sweep(h, parity, c.omega, dtg));
        ev_record(h, 2);
        ev_record(h, 3);
        st.iterations = c.niter;
    } else {
        ev_record(h, 1);
        int rc = ensure_mu(h, dtg);
        if (rc) return rc;
        if ((rc = rb_prepare(h))) return rc;
        double r_first = -1.0;
        const int stage = (int)(h->proj_count % 3);
        const auto t_plan = std::chrono::steady_clock::now();
        const GuessPlan gp = guess_plan(h, stage, dtg, &st);
        // a mixed-precision learned start is chosen on the device (no host round trip)
        const bool dev_select = gp.learned && c.precision == MDX_PREC_MIXED;
        unsigned long long first_seq = 0;
        bool defer = false;
        if (gp.simple && c.precision == MDX_PREC_MIXED) {
            // rhs, the start and the solve's first check in one kernel
            first_seq = ++h->chk_seq;
            MDX_HIP(launch_sor_start(h, gp.prev, gp.fac, h->d_pinned + CHECK_PIN, first_seq));
        } else {
            // the learned start falling back to Phi = 0 compares against that start's residual,
            // max |rhs|, formed here on the way
            const bool want_rz = gp.learned && gp.have_prev && !gp.use_prev;
            // the fused input pass also forms rhs and max |rhs|
            const bool fused = dev_select && h->fused_input && learned_input_fits(h, h->net->stride);
            if (!fused) MDX_HIP(launch_sor_prepare(h, want_rz ? h->d_pinned + RZ_PIN : nullptr, 0));
            if (dev_select) {
                rc = learned_start_mixed(h, stage, gp, fused, &first_seq);
                // After a network-started solve of this stage the network's start is expected
                // again: the solve queues its first batch (inner_iters, the network start's first
                // batch) behind the choice and learns which start ran from the first check.  Else
                // the choice is read first, so the batches are those of the start that runs (a
                // fallback solve is then the other mode's solve exactly).
                defer = rc == MDX_OK && h->defer_choice && h->net_prev_started[stage] && lds_for(h, true);
                double r0 = 0.0;
                if (rc == MDX_OK && !defer) rc = wait_check(h, CHECK_PIN, first_seq, &r0);
                if (rc == MDX_OK && !defer) rc = resolve_learned_start(h, stage, gp, &st, &r_first);
            } else {
                rc = initial_guess(h, stage, dtg, gp, &st, &r_first);
            }
            if (rc && rc != MDX_LDS_TIMEOUT) return rc;
        }
        if (rc == MDX_OK) {
            // the merge prior is the previous solve of this stage from the same kind of start
            const int kind = (defer || h->net_started[stage]) ? 2 : (gp.use_prev ? 1 : 0);
            std::function<int(int &)> on_first;
            if (defer)
                on_first = [&](int &key) {
                    const int r_ = resolve_learned_start(h, stage, gp, &st, &r_first);
                    if (!h->net_started[stage]) key = stage + 3 * (gp.use_prev ? 1 : 0);
                    return r_;
                };
            rc = c.precision == MDX_PREC_MIXED ? sor_solve_mixed(h, &st, first_seq, stage + 3 * kind, on_first)
                                               : sor_solve64(h, &st, r_first);
        }
        if (rc == MDX_LDS_TIMEOUT) {
            // a neighbour wait of the LDS-resident sweeps gave up (another stream or process held
            // CUs): the solve restarts from Phi = 0 on the streaming sweeps, and the next solve
            // tries the LDS-resident sweeps again (for good off after LDS_TIMEOUT_MAX restarts)
            h->lds_timeouts += 1;
            h->lds_suspended = true;
            if (h->lds_timeouts >= LDS_TIMEOUT_MAX) {
                h->rb_state = -1;
                h->rb_note = "streaming sweeps: neighbour waits of the LDS-resident sweeps gave up " +
                             std::to_string(h->lds_timeouts) + " times";
            }
            const int learned_mode = c.guess == MDX_GUESS_LEARNED;
            st = mdx_solve_stats{};
            st.lds_restarts = 1;
            st.fallback = learned_mode;
            h->net_started[stage] = false;
            r_first = -1.0;
            MDX_HIP(launch_zero64(h, h->phi, h->ni));
            rc = c.precision == MDX_PREC_MIXED ? sor_solve_mixed(h, &st, 0, stage) : sor_solve64(h, &st, -1.0);
            if (rc == MDX_LDS_TIMEOUT) rc = MDX_ERR_HIP;   // cannot come from the streaming sweeps
        }
        h->lds_suspended = false;
        if (rc) return rc;
        if (st.fallback == 0 && r_first >= 0.0) st.div0_inf = r_first;
        net_outcome(h, stage, std::chrono::duration<double>(std::chrono::steady_clock::now() - t_plan).count(),
                    gp.learned);
        h->net_prev_started[stage] = h->net_started[stage];
        if (h->net && h->net_capture && st.converged && (rc = capture_pair(h, stage, dtg))) return rc;
        // the converged Phi becomes the history of this stage for the next step (copied by the
        // projection kernel, which reads every Phi anyway)
        h->hist_dtg[stage] = dtg;
        h->proj_count += 1;
        ev_record(h, 2);
        MDX_HIP(launch_sor_apply(h, h->phi, dtg, h->phi_hist + (size_t)stage * h->ni));
        ev_record(h, 3);
        // Reaching max_iter is reported through stats.converged (the caller decides); a
        // non-finite residual is an error.
        if (!(st.div_inf < 1e300)) {
            mdx_set_error("pressure_projection: non-finite divergence in the SOR solve");
            if (stats) *stats = st;
            return MDX_ERR_NOT_CONVERGED;
        }
    }
    if (stats) *stats = st;
    return MDX_OK;
}

extern "C" int mdx_set_guess(mdx_handle *h, int32_t guess) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    for (int q = 0; q < 3; ++q) {
        h->net_losses[q] = h->net_skip[q] = 0;
        h->base_sec[q] = 0.0;
        h->net_started[q] = h->net_prev_started[q] = false;
    }
    MDX_CHECK(guess == MDX_GUESS_PREVIOUS || guess == MDX_GUESS_LEARNED || guess == MDX_GUESS_PHI_PREV,
              MDX_ERR_ARG, "mdx_set_guess: unknown guess mode");
    MDX_CHECK(guess == MDX_GUESS_PREVIOUS || h->cfg.solver == MDX_SOLVER_SOR, MDX_ERR_UNSUPPORTED,
              "mdx_set_guess: initial guesses serve the SOR solver");
    h->cfg.guess = guess;
    return MDX_OK;
}

extern "C" int mdx_flowrate_shift(mdx_handle *h, double dtg, double *du) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_CHECK(h->cfg.flow_height > 0.0, MDX_ERR_ARG, "flow_height must be positive");
    double *dev = h->red_part + 2048;
    MDX_HIP(launch_flowrate(h, dtg, dev));
    if (du) {
        MDX_HIP(hipMemcpyAsync(h->h_pinned, dev, 3 * 8, hipMemcpyDeviceToHost, h->stream));
        MDX_HIP(hipStreamSynchronize(h->stream));
        *du = h->h_pinned[2];
    }
    return MDX_OK;
}

extern "C" int mdx_bulk_velocity(mdx_handle *h, double *ub) {
    MDX_CHECK(h && ub, MDX_ERR_ARG, "null argument");
    MDX_CHECK(h->cfg.flow_height > 0.0, MDX_ERR_ARG, "flow_height must be positive");
    double *dev = h->red_part + 2048;
    MDX_HIP(launch_bulk_sum(h, h->un, 1.0, dev));
    MDX_HIP(hipMemcpyAsync(h->h_pinned, dev, 3 * 8, hipMemcpyDeviceToHost, h->stream));
    MDX_HIP(hipStreamSynchronize(h->stream));
    *ub = h->h_pinned[0];
    return MDX_OK;
}

static const double RK_ALPHA[3] = {64.0 / 120.0, 50.0 / 120.0, 90.0 / 120.0};   // step.f90:21-23
static const double RK_BETA[3] = {0.0, -34.0 / 120.0, -50.0 / 120.0};
static const double RK_GAMMA[3] = {64.0 / 120.0, 16.0 / 120.0, 40.0 / 120.0};

// after the projection: the current fields are the only arrays it wrote
static int post_bc(mdx_handle *h) {
    if (h->cfg.solver == MDX_SOLVER_FFT) {
        MDX_HIP(launch_apply_bc_fft(h, BC_N));
        return MDX_OK;
    }
    return mdx_apply_bc_redblack(h);
}

extern "C" int mdx_init_fields(mdx_handle *h, uint32_t seed, double amp, const mdx_geometry *geo, double dt) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_CHECK(dt > 0.0, MDX_ERR_ARG, "mdx_init_fields: dt must be positive");
    const int use_geo = geo != nullptr;
    if (use_geo) MDX_CHECK(h->ibm_on, MDX_ERR_STATE, "mdx_init_fields: geometry given but no IBM coefficients set");
    double *fields[] = {h->un, h->vn, h->wn, h->pn, h->us, h->vs, h->ws};
    for (double *f : fields) MDX_HIP(launch_zero64(h, f, h->ng));
    double *olds[] = {h->oru, h->orv, h->orw};
    for (double *f : olds) MDX_HIP(launch_zero64(h, f, h->ni));
    std::vector<double> hx(h->d.nx + 2, 1.0);
    if (use_geo)
        for (int i = 0; i <= h->d.nx + 1; ++i) {
            double x = ((double)(i + 1) - 0.5 - 1.5) * h->dx;
            hx[i] = geo->half_height - geo->A * std::cos(2.0 * PI * x / geo->lambda);
        }
    double *d_hx = h->red_part + 3072;
    MDX_CHECK(h->d.nx + 2 <= 1024, MDX_ERR_UNSUPPORTED, "mdx_init_fields: nx <= 1022");
    MDX_HIP(hipMemcpyAsync(d_hx, hx.data(), hx.size() * 8, hipMemcpyHostToDevice, h->stream));
    MDX_HIP(launch_initial_field(h, seed, amp, use_geo, d_hx, 0.5 * h->cfg.ly));
    int rc;
    if (h->cfg.solver == MDX_SOLVER_FFT && (rc = mdx_apply_bc(h))) return rc;
    mdx_solve_stats st;
    if ((rc = mdx_pressure_projection(h, dt * RK_GAMMA[0], &st))) return rc;
    if ((rc = post_bc(h))) return rc;
    MDX_HIP(launch_zero64(h, h->pn, h->ng));
    MDX_HIP(hipStreamSynchronize(h->stream));
    // the projection of the initial field is not an RK stage: start the guess history afresh
    h->proj_count = 0;
    for (double &t : h->hist_dtg) t = 0.0;
    return MDX_OK;
}

extern "C" int mdx_step(mdx_handle *h, double dt, mdx_step_stats *stats) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    MDX_CHECK(dt > 0.0, MDX_ERR_ARG, "mdx_step: dt must be positive");
    int rc;
    for (int s = 0; s < 3; ++s) {
        const double dta = dt * RK_ALPHA[s], dtb = dt * RK_BETA[s], dtg = dt * RK_GAMMA[s];
        // every PH_EVERY-th stage is timed, in an event set of the ring (the oldest is folded if
        // it is full); the events held the stream for microseconds each
        const bool timed = h->timing && h->ph_seen++ % mdx_handle::PH_EVERY == 0;
        h->ph_skip = h->timing && !timed;
        if (timed) {
            if (h->ph_count == mdx_handle::PHQ) phase_account(h, true);
            h->ph_cur = (h->ph_head + h->ph_count) % mdx_handle::PHQ;
        }
        ev_record(h, 0);
        if (h->cfg.flowrate && h->fr_part) {
            // momentum forms the flow-rate partial sums of u*; one launch folds them, shifts u* and
            // builds the stage's face-coefficient tables if the solve needs new ones
            MDX_HIP(launch_momentum(h, dta, dtb, dtg, h->fr_part));
            ev_record(h, 5);
            // (the tables exist for the SOR solver only)
            const bool tables = h->fused_mu && h->ibm_on && h->mu_dtg != dtg && h->cfg.solver == MDX_SOLVER_SOR &&
                                h->mu64[0] && h->mu32[0];
            MDX_HIP(launch_flow_fold(h, dtg, h->fr_part, momentum_tiles(h), h->red_part + 2048,
                                     tables ? mu_build(h, dtg) : MuBuild{}));
            if (tables) h->mu_dtg = dtg;
        } else {
            if ((rc = mdx_momentum(h, dta, dtb, dtg))) return rc;
            ev_record(h, 5);
            if (h->cfg.flowrate && (rc = mdx_flowrate_shift(h, dtg, nullptr))) return rc;
        }
        // the predictor's arrays are the only ones written since the last call (the current
        // fields' ghosts are those of the previous stage's projection or of the upload)
        if (h->cfg.solver == MDX_SOLVER_FFT) MDX_HIP(launch_apply_bc_fft(h, BC_S));
        mdx_solve_stats st{};
        h->proj_count = s;   // RK stage of this projection (initial-guess history)
        // the fill follows the projection at once: its kernel may write the boundary conditions
        h->step_bc = true;
        h->apply_wrote_bc = false;
        rc = mdx_pressure_projection(h, dtg, &st);
        h->step_bc = false;
        if (rc) {
            if (stats) stats->stage[s] = st;
            return rc;
        }
        if (h->apply_wrote_bc) h->apply_wrote_bc = false;
        else if ((rc = post_bc(h))) return rc;
        ev_record(h, 4);
        if (stats) stats->stage[s] = st;
        h->ph_skip = false;
        if (timed) {
            h->ph_count += 1;
            h->ph_cur = mdx_handle::PHQ;   // later phase events outside a step go to the scratch set
            phase_account(h, false);
        }
    }
    if (stats && h->cfg.flowrate) {
        if (!h->step_ub) stats->ub = std::nan("");
        else if ((rc = mdx_bulk_velocity(h, &stats->ub))) return rc;
    }
    return MDX_OK;
}

extern "C" int mdx_set_step_report(mdx_handle *h, int32_t bulk_velocity) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    h->step_ub = bulk_velocity != 0;
    return MDX_OK;
}

// A batch of independent channels (for example wavy channels with different amplitudes and
// seeds) stepped concurrently: one host thread per handle drives that handle's own stream, so the
// GPU overlaps the instances' kernels and each instance's residual checks wait only on its own
// work.  Each instance computes exactly what mdx_step computes for it alone.  The LDS-resident
// sweeps need every CU for one launch and are switched off for batched handles.
extern "C" int mdx_step_batch(mdx_handle *const *hs, int32_t n, double dt, mdx_step_stats *stats) {
    MDX_CHECK(hs && n > 0, MDX_ERR_ARG, "mdx_step_batch: need n > 0 handles");
    for (int q = 0; q < n; ++q) {
        MDX_CHECK(hs[q], MDX_ERR_ARG, "mdx_step_batch: null handle");
        for (int r = 0; r < q; ++r) {
            MDX_CHECK(hs[r] != hs[q], MDX_ERR_ARG, "mdx_step_batch: a handle appears twice");
            // a network's activations, stream and replay cursor serve one solve at a time
            MDX_CHECK(!hs[q]->net || hs[r]->net != hs[q]->net, MDX_ERR_ARG,
                      "mdx_step_batch: two handles share one attached network");
        }
    }
    for (int q = 0; q < n; ++q) {
        mdx_handle *h = hs[q];
        if (h->lds_sor_mode != 0) {
            h->lds_sor_mode = 0;
            h->rb_state = -1;
            h->rb_note = "streaming sweeps: concurrent batch";
        }
    }
    std::vector<int> rc(n, MDX_OK);
    std::vector<std::string> err(n);
    std::vector<std::thread> th;
    th.reserve(n);
    for (int q = 0; q < n; ++q)
        th.emplace_back([&, q] {
            if (hipSetDevice(hs[q]->device) != hipSuccess) {
                rc[q] = MDX_ERR_HIP;
                err[q] = "hipSetDevice failed";
                return;
            }
            rc[q] = mdx_step(hs[q], dt, stats ? &stats[q] : nullptr);
            if (rc[q]) err[q] = mdx_last_error();
        });
    for (auto &t : th) t.join();
    for (int q = 0; q < n; ++q)
        if (rc[q]) {
            mdx_set_error("mdx_step_batch: instance " + std::to_string(q) + ": " + err[q]);
            return rc[q];
        }
    return MDX_OK;
}

extern "C" int mdx_get_phase_times(mdx_handle *h, mdx_phase_times *t) {
    MDX_CHECK(h && t, MDX_ERR_ARG, "null argument");
    sweep_account(h, true);   // batches still queued (a discarded speculative one included)
    phase_account(h, true);
    *t = h->times;
    return MDX_OK;
}

extern "C" int mdx_reset_phase_times(mdx_handle *h) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    sweep_account(h, true);
    phase_account(h, true);
    h->times = mdx_phase_times{};
    return MDX_OK;
}

extern "C" int mdx_get_y_coordinates(mdx_handle *h, double *yc, double *yv) {
    MDX_CHECK(h, MDX_ERR_ARG, "null handle");
    if (yc) std::memcpy(yc, h->yc.data(), h->yc.size() * 8);
    if (yv) std::memcpy(yv, h->yv.data(), h->yv.size() * 8);
    return MDX_OK;
}

extern "C" int mdx_get_info(mdx_handle *h, char *buf, int32_t len) {
    MDX_CHECK(h && buf && len > 0, MDX_ERR_ARG, "mdx_get_info: bad argument");
    std::string s = h->cfg.solver == MDX_SOLVER_FFT ? "FFT/tridiagonal pressure solve"
                    : h->cfg.solver == MDX_SOLVER_SOLA ? "SOLA red-black sweeps (fixed niter)"
                    : (h->rb_note.empty() ? "red-black SOR (not yet planned)" : h->rb_note);
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
    return MDX_OK;
}
