/*
 * oracle.c -- CPU restatement of mobydiff's time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library;
 * it is never linked into, loaded by or called from the product.  It is
 * compiled with -ffp-contract=off so its FP64 arithmetic follows the
 * reference's expression order operation by operation.
 *
 * Function -> reference (file:line) -> how it is pinned:
 *   mdo_momentum            step.f90:27-207            reference build (oracle/_ref), random fields
 *   mdo_get_cfl             step.f90:210-258           brute-force max
 *   mdo_apply_bc_fft        boundary.f90:23-170        reference build (FFT variant)
 *   mdo_apply_bc_redblack   boundary.f90:173-368       reference build (RB variant), all BC types
 *   mdo_sola_redblack       pressure_redblack.f90:27-254  reference build, fixed nIter, all BC types
 *   mdo_set_ibm_coeff       ibm.f90:84-187             reference build (2nd-order IBM), kind 0 geometry
 *   mdo_tridiag_coefficients/solve  pressure_fft.f90:20-148  reference build (FFT variant)
 *   mdo_fft_r2c_xz/c2r_xz   pressure_backend_fftw.f90:22-93  DFT definition: numpy.fft fixtures
 *   mdo_build_fft_rhs       pressure_solver.f90:77-115  reference unbuildable (needs FFTW): pinned by
 *                                                       the closed-form divergence of analytic fields
 *   mdo_fft_pressure_correction pressure_solver.f90:117-173  pinned by ||div||_inf <= 1e-12 after projection
 *   mdo_sor_* (pressure-correction SOR)  NEW; equals the SOLA iteration in exact arithmetic:
 *                                         pinned against mdo_sola_redblack at fixed iteration counts
 *   mdo_div_inf, mdo_bulk_u NEW: closed forms on analytic fields
 *   mdo_stretched_metrics   NEW: reduces to uniform dy at beta -> 0; closed form tanh map
 *   mdo_initial_field       NEW: golden RNG values (SplitMix64 known-answer vectors)
 *   mdo_flowrate_shift      NEW: bulk velocity equals the target to round-off
 *   mdo_bulk_u_step         NEW: the step's order of the same sums; closed forms and an exactly
 *                                rounded sum (math.fsum) on analytic and random fields
 *   mdo_rk3_step            main.f90:66-89 stage order; RK constants step.f90:21-23
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define GI(i, j, k) ((size_t)(i) + (size_t)(nx + 2) * ((size_t)(j) + (size_t)(ny + 2) * (size_t)(k)))
#define II(i, j, k) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((j) - 1) + (size_t)ny * (size_t)((k) - 1)))

static const double PI = 3.141592653589793;

/* step.f90:21-23 */
const double mdo_rk_alpha[3] = {64.0 / 120.0, 50.0 / 120.0, 90.0 / 120.0};
const double mdo_rk_beta[3] = {0.0, -34.0 / 120.0, -50.0 / 120.0};
const double mdo_rk_gamma[3] = {64.0 / 120.0, 16.0 / 120.0, 40.0 / 120.0};

/* init.f90:48-55 */
void mdo_finalize_grid(mdo_grid *g) {
    g->dx = g->lx / (double)g->nx;
    g->dy = g->ly / (double)g->ny;
    g->dz = g->lz / (double)g->nz;
}

/* tanh-stretched wall-normal grid on [0, ly]: v-faces yv(j), j=1..ny+1, cell centres
 * yc(j) midway between faces, ghost centres mirrored about the walls. */
void mdo_stretched_metrics(int ny, double ly, double beta, double *yc, double *yv,
                           double *ihc, double *ihv) {
    double tb = tanh(beta);
    for (int j = 1; j <= ny + 1; ++j) {
        double eta = (double)(j - 1) / (double)ny;
        yv[j] = 0.5 * ly * (1.0 + tanh(beta * (2.0 * eta - 1.0)) / tb);
    }
    yv[1] = 0.0;
    yv[ny + 1] = ly;
    yv[0] = -yv[2];
    for (int j = 1; j <= ny; ++j) yc[j] = 0.5 * (yv[j] + yv[j + 1]);
    yc[0] = 2.0 * yv[1] - yc[1];
    yc[ny + 1] = 2.0 * yv[ny + 1] - yc[ny];
    for (int j = 1; j <= ny; ++j) ihc[j] = 1.0 / (yv[j + 1] - yv[j]);
    ihc[0] = ihc[1];
    ihc[ny + 1] = ihc[ny];
    for (int j = 1; j <= ny + 1; ++j) ihv[j] = 1.0 / (yc[j] - yc[j - 1]);
    ihv[0] = ihv[1];
}

/* ------------------------------------------------------------------ momentum */
/* step.f90:27-207.  Uniform-dy branch keeps the reference's expression order exactly;
 * the stretched branch replaces each y difference by its non-uniform counterpart. */
void mdo_momentum(const mdo_grid *g,
                  const double *un, const double *vn, const double *wn, const double *pn,
                  double *us, double *vs, double *ws,
                  double *oru, double *orv, double *orw,
                  const double *cu, const double *cv, const double *cw,
                  double dta, double dtb, double dtg) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const double idx = 1.0 / dx, idy = 1.0 / dy, idz = 1.0 / dz;
    const double idx2 = 1.0 / dx2, idy2 = 1.0 / dy2, idz2 = 1.0 / dz2;
    const double ire = 1.0 / g->re;
    const double fx = g->forcing[0], fy = g->forcing[1], fz = g->forcing[2];
    const int nx0 = g->periodic[0] ? 1 : 2;
    const int ny0 = g->periodic[1] ? 1 : 2;
    const int nz0 = g->periodic[2] ? 1 : 2;
    const double *ihc = g->ihc, *ihv = g->ihv;
    const int stretched = (ihc != NULL);

#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                const int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1, km = k - 1;
                const size_t c = GI(i, j, k);
                const size_t o = II(i, j, k);
                if (i >= nx0) {
                    double uu_p = 0.25 * ((un[GI(ip, j, k)] + un[c]) * (un[GI(ip, j, k)] + un[c]));
                    double uu_m = 0.25 * ((un[c] + un[GI(im, j, k)]) * (un[c] + un[GI(im, j, k)]));
                    double uv_p = 0.25 * (un[GI(i, jp, k)] + un[c]) * (vn[GI(i, jp, k)] + vn[GI(im, jp, k)]);
                    double uv_m = 0.25 * (un[c] + un[GI(i, jm, k)]) * (vn[c] + vn[GI(im, j, k)]);
                    double uw_p = 0.25 * (un[c] + un[GI(i, j, kp)]) * (wn[GI(i, j, kp)] + wn[GI(im, j, kp)]);
                    double uw_m = 0.25 * (un[c] + un[GI(i, j, km)]) * (wn[c] + wn[GI(im, j, k)]);
                    double diff_ux = (un[GI(im, j, k)] - 2.0 * un[c] + un[GI(ip, j, k)]) * idx2;
                    double diff_uy, cy;
                    if (!stretched) {
                        diff_uy = (un[GI(i, jm, k)] - 2.0 * un[c] + un[GI(i, jp, k)]) * idy2;
                        cy = (uv_p - uv_m) * idy;
                    } else {
                        diff_uy = ((un[GI(i, jp, k)] - un[c]) * ihv[jp] - (un[c] - un[GI(i, jm, k)]) * ihv[j]) * ihc[j];
                        cy = (uv_p - uv_m) * ihc[j];
                    }
                    double diff_uz = (un[GI(i, j, km)] - 2.0 * un[c] + un[GI(i, j, kp)]) * idz2;
                    double dpx = (pn[c] - pn[GI(im, j, k)]) * idx;
                    double rhsu = -(uu_p - uu_m) * idx - cy - (uw_p - uw_m) * idz + fx
                                  + ire * (diff_ux + diff_uy + diff_uz);
                    double s = un[c] + dta * rhsu + dtb * oru[o] - dtg * dpx;
                    double mu = 1.0 / (1.0 + dtg * (cu ? cu[c] : 0.0));
                    us[c] = s * mu;
                    oru[o] = rhsu;
                }
                if (j >= ny0) {
                    double vu_p = 0.25 * (vn[c] + vn[GI(ip, j, k)]) * (un[GI(ip, j, k)] + un[GI(ip, jm, k)]);
                    double vu_m = 0.25 * (vn[c] + vn[GI(im, j, k)]) * (un[c] + un[GI(i, jm, k)]);
                    double vv_p = 0.25 * ((vn[c] + vn[GI(i, jp, k)]) * (vn[c] + vn[GI(i, jp, k)]));
                    double vv_m = 0.25 * ((vn[c] + vn[GI(i, jm, k)]) * (vn[c] + vn[GI(i, jm, k)]));
                    double vw_p = 0.25 * (vn[GI(i, j, kp)] + vn[c]) * (wn[GI(i, j, kp)] + wn[GI(i, jm, kp)]);
                    double vw_m = 0.25 * (vn[GI(i, j, km)] + vn[c]) * (wn[c] + wn[GI(i, jm, k)]);
                    double diff_vx = (vn[GI(im, j, k)] - 2.0 * vn[c] + vn[GI(ip, j, k)]) * idx2;
                    double diff_vy, cy, dpy;
                    if (!stretched) {
                        diff_vy = (vn[GI(i, jm, k)] - 2.0 * vn[c] + vn[GI(i, jp, k)]) * idy2;
                        cy = (vv_p - vv_m) * idy;
                        dpy = (pn[c] - pn[GI(i, jm, k)]) * idy;
                    } else {
                        diff_vy = ((vn[GI(i, jp, k)] - vn[c]) * ihc[j] - (vn[c] - vn[GI(i, jm, k)]) * ihc[jm]) * ihv[j];
                        cy = (vv_p - vv_m) * ihv[j];
                        dpy = (pn[c] - pn[GI(i, jm, k)]) * ihv[j];
                    }
                    double diff_vz = (vn[GI(i, j, km)] - 2.0 * vn[c] + vn[GI(i, j, kp)]) * idz2;
                    double rhsv = -(vu_p - vu_m) * idx - cy - (vw_p - vw_m) * idz + fy
                                  + ire * (diff_vx + diff_vy + diff_vz);
                    double s = vn[c] + dta * rhsv + dtb * orv[o] - dtg * dpy;
                    double mu = 1.0 / (1.0 + dtg * (cv ? cv[c] : 0.0));
                    vs[c] = s * mu;
                    orv[o] = rhsv;
                }
                if (k >= nz0) {
                    double wu_p = 0.25 * (wn[c] + wn[GI(ip, j, k)]) * (un[GI(ip, j, k)] + un[GI(ip, j, km)]);
                    double wu_m = 0.25 * (wn[c] + wn[GI(im, j, k)]) * (un[c] + un[GI(i, j, km)]);
                    double ww_p = 0.25 * ((wn[c] + wn[GI(i, j, kp)]) * (wn[c] + wn[GI(i, j, kp)]));
                    double ww_m = 0.25 * ((wn[c] + wn[GI(i, j, km)]) * (wn[c] + wn[GI(i, j, km)]));
                    double wv_p = 0.25 * (wn[c] + wn[GI(i, jp, k)]) * (vn[GI(i, jp, k)] + vn[GI(i, jp, km)]);
                    double wv_m = 0.25 * (wn[c] + wn[GI(i, jm, k)]) * (vn[c] + vn[GI(i, j, km)]);
                    double diff_wx = (wn[GI(im, j, k)] - 2.0 * wn[c] + wn[GI(ip, j, k)]) * idx2;
                    double diff_wy, cy;
                    if (!stretched) {
                        diff_wy = (wn[GI(i, jm, k)] - 2.0 * wn[c] + wn[GI(i, jp, k)]) * idy2;
                        cy = (wv_p - wv_m) * idy;
                    } else {
                        diff_wy = ((wn[GI(i, jp, k)] - wn[c]) * ihv[jp] - (wn[c] - wn[GI(i, jm, k)]) * ihv[j]) * ihc[j];
                        cy = (wv_p - wv_m) * ihc[j];
                    }
                    double diff_wz = (wn[GI(i, j, km)] - 2.0 * wn[c] + wn[GI(i, j, kp)]) * idz2;
                    double dpz = (pn[c] - pn[GI(i, j, km)]) * idz;
                    double rhsw = -(wu_p - wu_m) * idx - cy - (ww_p - ww_m) * idz + fz
                                  + ire * (diff_wx + diff_wy + diff_wz);
                    double s = wn[c] + dta * rhsw + dtb * orw[o] - dtg * dpz;
                    double mu = 1.0 / (1.0 + dtg * (cw ? cw[c] : 0.0));
                    ws[c] = s * mu;
                    orw[o] = rhsw;
                }
            }
}

/* step.f90:210-258 */
double mdo_get_cfl(const mdo_grid *g, const double *un, const double *vn, const double *wn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double cfl = 0.0;
    size_t n = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    for (size_t q = 0; q < n; ++q) {
        double a = fabs(un[q] / g->dx);
        if (a > cfl) cfl = a;
    }
    for (size_t q = 0; q < n; ++q) {
        size_t j = (q / (size_t)(nx + 2)) % (size_t)(ny + 2);
        double a = g->ihv ? fabs(vn[q] * g->ihv[j]) : fabs(vn[q] / g->dy);
        if (a > cfl) cfl = a;
    }
    for (size_t q = 0; q < n; ++q) {
        double a = fabs(wn[q] / g->dz);
        if (a > cfl) cfl = a;
    }
    return cfl;
}

/* ------------------------------------------------------------------ boundaries */
/* boundary.f90:23-170 (FFT build: n- and *-fields, edges included) */
void mdo_apply_bc_fft(const mdo_grid *g, double *un, double *vn, double *wn,
                      double *us, double *vs, double *ws, double *pn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    for (int i = 0; i <= nx + 1; ++i)
        for (int k = 0; k <= nz + 1; ++k) {
            un[GI(i, ny + 1, k)] = -un[GI(i, ny, k)];
            vn[GI(i, ny + 1, k)] = 0.0;
            wn[GI(i, ny + 1, k)] = -wn[GI(i, ny, k)];
            un[GI(i, 0, k)] = -un[GI(i, 1, k)];
            vn[GI(i, 1, k)] = 0.0;
            wn[GI(i, 0, k)] = -wn[GI(i, 1, k)];
            us[GI(i, ny + 1, k)] = -us[GI(i, ny, k)];
            vs[GI(i, ny + 1, k)] = 0.0;
            ws[GI(i, ny + 1, k)] = -ws[GI(i, ny, k)];
            us[GI(i, 0, k)] = -us[GI(i, 1, k)];
            vs[GI(i, 1, k)] = 0.0;
            ws[GI(i, 0, k)] = -ws[GI(i, 1, k)];
        }
    for (int j = 0; j <= ny + 1; ++j)
        for (int k = 0; k <= nz + 1; ++k) {
            un[GI(0, j, k)] = un[GI(nx, j, k)];
            wn[GI(0, j, k)] = wn[GI(nx, j, k)];
            un[GI(nx + 1, j, k)] = un[GI(1, j, k)];
            wn[GI(nx + 1, j, k)] = wn[GI(1, j, k)];
            us[GI(0, j, k)] = us[GI(nx, j, k)];
            ws[GI(0, j, k)] = ws[GI(nx, j, k)];
            us[GI(nx + 1, j, k)] = us[GI(1, j, k)];
            ws[GI(nx + 1, j, k)] = ws[GI(1, j, k)];
            if (j >= 1) {
                vn[GI(0, j, k)] = vn[GI(nx, j, k)];
                vn[GI(nx + 1, j, k)] = vn[GI(1, j, k)];
                vs[GI(0, j, k)] = vs[GI(nx, j, k)];
                vs[GI(nx + 1, j, k)] = vs[GI(1, j, k)];
            }
            if (j >= 1 && j <= ny) {
                pn[GI(0, j, k)] = pn[GI(nx, j, k)];
                pn[GI(nx + 1, j, k)] = pn[GI(1, j, k)];
            }
        }
    for (int i = 0; i <= nx + 1; ++i)
        for (int j = 0; j <= ny + 1; ++j) {
            un[GI(i, j, 0)] = un[GI(i, j, nz)];
            wn[GI(i, j, 0)] = wn[GI(i, j, nz)];
            un[GI(i, j, nz + 1)] = un[GI(i, j, 1)];
            wn[GI(i, j, nz + 1)] = wn[GI(i, j, 1)];
            us[GI(i, j, 0)] = us[GI(i, j, nz)];
            ws[GI(i, j, 0)] = ws[GI(i, j, nz)];
            us[GI(i, j, nz + 1)] = us[GI(i, j, 1)];
            ws[GI(i, j, nz + 1)] = ws[GI(i, j, 1)];
            if (j >= 1) {
                vn[GI(i, j, 0)] = vn[GI(i, j, nz)];
                vn[GI(i, j, nz + 1)] = vn[GI(i, j, 1)];
                vs[GI(i, j, 0)] = vs[GI(i, j, nz)];
                vs[GI(i, j, nz + 1)] = vs[GI(i, j, 1)];
            }
            if (j >= 1 && j <= ny) {
                pn[GI(i, j, 0)] = pn[GI(i, j, nz)];
                pn[GI(i, j, nz + 1)] = pn[GI(i, j, 1)];
            }
        }
}

/* boundary.f90:173-368: table-driven ghosts on the six faces, edges and corners untouched.
 * Executed face by face in the reference's loop order (bottom, top, left, right, front, back). */
void mdo_apply_bc_redblack(const mdo_grid *g, double *un, double *vn, double *wn, double *pn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
    const int *t = g->bc_type;
    const double *v = g->bc_value;
#define T(d, s, q) t[MDO_BCI(d, s, q)]
#define V(d, s, q) v[MDO_BCI(d, s, q)]
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) { /* bottom */
            if (g->periodic[1]) {
                un[GI(i, 0, k)] = un[GI(i, ny, k)];
                vn[GI(i, 0, k)] = vn[GI(i, ny, k)];
                wn[GI(i, 0, k)] = wn[GI(i, ny, k)];
                pn[GI(i, 0, k)] = pn[GI(i, ny, k)];
            } else {
                un[GI(i, 0, k)] = T(2, 0, 1) == 0 ? 2 * V(2, 0, 1) - un[GI(i, 1, k)] : un[GI(i, 1, k)] - V(2, 0, 1) * dy;
                vn[GI(i, 1, k)] = T(2, 0, 2) == 0 ? V(2, 0, 2) : vn[GI(i, 2, k)] - V(2, 0, 2) * dy;
                wn[GI(i, 0, k)] = T(2, 0, 3) == 0 ? 2 * V(2, 0, 3) - wn[GI(i, 1, k)] : wn[GI(i, 1, k)] - V(2, 0, 3) * dy;
            }
        }
    for (int k = 1; k <= nz; ++k)
        for (int i = 1; i <= nx; ++i) { /* top */
            if (g->periodic[1]) {
                un[GI(i, ny + 1, k)] = un[GI(i, 1, k)];
                vn[GI(i, ny + 1, k)] = vn[GI(i, 1, k)];
                wn[GI(i, ny + 1, k)] = wn[GI(i, 1, k)];
                pn[GI(i, ny + 1, k)] = pn[GI(i, 1, k)];
            } else {
                un[GI(i, ny + 1, k)] = T(2, 1, 1) == 0 ? 2 * V(2, 1, 1) - un[GI(i, ny, k)] : un[GI(i, ny, k)] + V(2, 1, 1) * dy;
                vn[GI(i, ny + 1, k)] = T(2, 1, 2) == 0 ? V(2, 1, 2) : vn[GI(i, ny, k)] + V(2, 1, 2) * dy;
                wn[GI(i, ny + 1, k)] = T(2, 1, 3) == 0 ? 2 * V(2, 1, 3) - wn[GI(i, ny, k)] : wn[GI(i, ny, k)] + V(2, 1, 3) * dy;
            }
        }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) { /* left */
            if (g->periodic[0]) {
                un[GI(0, j, k)] = un[GI(nx, j, k)];
                vn[GI(0, j, k)] = vn[GI(nx, j, k)];
                wn[GI(0, j, k)] = wn[GI(nx, j, k)];
                pn[GI(0, j, k)] = pn[GI(nx, j, k)];
            } else {
                un[GI(1, j, k)] = T(1, 0, 1) == 0 ? V(1, 0, 1) : un[GI(2, j, k)] - V(1, 0, 1) * dx;
                vn[GI(0, j, k)] = T(1, 0, 2) == 0 ? 2 * V(1, 0, 2) - vn[GI(1, j, k)] : vn[GI(1, j, k)] - V(1, 0, 2) * dx;
                wn[GI(0, j, k)] = T(1, 0, 3) == 0 ? 2 * V(1, 0, 3) - wn[GI(1, j, k)] : wn[GI(1, j, k)] - V(1, 0, 3) * dx;
            }
        }
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) { /* right */
            if (g->periodic[0]) {
                un[GI(nx + 1, j, k)] = un[GI(1, j, k)];
                vn[GI(nx + 1, j, k)] = vn[GI(1, j, k)];
                wn[GI(nx + 1, j, k)] = wn[GI(1, j, k)];
                pn[GI(nx + 1, j, k)] = pn[GI(1, j, k)];
            } else {
                un[GI(nx + 1, j, k)] = T(1, 1, 1) == 0 ? V(1, 1, 1) : un[GI(nx, j, k)] + V(1, 1, 1) * dx;
                vn[GI(nx + 1, j, k)] = T(1, 1, 2) == 0 ? 2 * V(1, 1, 2) - vn[GI(nx, j, k)] : vn[GI(nx, j, k)] + V(1, 1, 2) * dx;
                wn[GI(nx + 1, j, k)] = T(1, 1, 3) == 0 ? 2 * V(1, 1, 3) - wn[GI(nx, j, k)] : wn[GI(nx, j, k)] + V(1, 1, 3) * dx;
            }
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) { /* front */
            if (g->periodic[2]) {
                un[GI(i, j, 0)] = un[GI(i, j, nz)];
                vn[GI(i, j, 0)] = vn[GI(i, j, nz)];
                wn[GI(i, j, 0)] = wn[GI(i, j, nz)];
                pn[GI(i, j, 0)] = pn[GI(i, j, nz)];
            } else {
                un[GI(i, j, 0)] = T(3, 0, 1) == 0 ? 2 * V(3, 0, 1) - un[GI(i, j, 1)] : un[GI(i, j, 1)] - V(3, 0, 1) * dz;
                vn[GI(i, j, 0)] = T(3, 0, 2) == 0 ? 2 * V(3, 0, 2) - vn[GI(i, j, 1)] : vn[GI(i, j, 1)] - V(3, 0, 2) * dz;
                wn[GI(i, j, 1)] = T(3, 0, 3) == 0 ? V(3, 0, 3) : wn[GI(i, j, 2)] - V(3, 0, 3) * dz;
            }
        }
    for (int j = 1; j <= ny; ++j)
        for (int i = 1; i <= nx; ++i) { /* back */
            if (g->periodic[2]) {
                un[GI(i, j, nz + 1)] = un[GI(i, j, 1)];
                vn[GI(i, j, nz + 1)] = vn[GI(i, j, 1)];
                wn[GI(i, j, nz + 1)] = wn[GI(i, j, 1)];
                pn[GI(i, j, nz + 1)] = pn[GI(i, j, 1)];
            } else {
                un[GI(i, j, nz + 1)] = T(3, 1, 1) == 0 ? 2 * V(3, 1, 1) - un[GI(i, j, nz)] : un[GI(i, j, nz)] + V(3, 1, 1) * dz;
                vn[GI(i, j, nz + 1)] = T(3, 1, 2) == 0 ? 2 * V(3, 1, 2) - vn[GI(i, j, nz)] : vn[GI(i, j, nz)] + V(3, 1, 2) * dz;
                wn[GI(i, j, nz + 1)] = T(3, 1, 3) == 0 ? V(3, 1, 3) : wn[GI(i, j, nz)] + V(3, 1, 3) * dz;
            }
        }
#undef T
#undef V
}

/* ------------------------------------------------------------------ FFT path */
/* pressure_solver.f90:77-115: rhs(i,k,j) in (x,z,y) order */
void mdo_build_fft_rhs(const mdo_grid *g, const double *us, const double *vs, const double *ws,
                       double dtg, double *rhs) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int j = 1; j <= ny; ++j)
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                double dv = g->ihc ? (vs[GI(i, j + 1, k)] - vs[GI(i, j, k)]) * g->ihc[j]
                                   : (vs[GI(i, j + 1, k)] - vs[GI(i, j, k)]) / dy;
                rhs[(size_t)(i - 1) + (size_t)nx * ((size_t)(k - 1) + (size_t)nz * (size_t)(j - 1))] =
                    ((us[GI(i + 1, j, k)] - us[GI(i, j, k)]) / dx + dv + (ws[GI(i, j, k + 1)] - ws[GI(i, j, k)]) / dz) / dtg;
            }
}

/* In-place complex FFT of length n over elements x[s*q] (interleaved re,im),
 * X_m = sum_q x_q exp(sign * 2 pi i m q / n).  Radix-2 for powers of two, direct DFT otherwise. */
static void fft_line(int n, double *x, size_t s, int sign, double *tmp) {
    if ((n & (n - 1)) == 0) {
        for (int q = 0; q < n; ++q) { tmp[2 * q] = x[2 * s * q]; tmp[2 * q + 1] = x[2 * s * q + 1]; }
        for (int q = 1, r = 0; q < n; ++q) { /* bit reversal */
            int bit = n >> 1;
            for (; r & bit; bit >>= 1) r ^= bit;
            r ^= bit;
            if (q < r) {
                double a = tmp[2 * q], b = tmp[2 * q + 1];
                tmp[2 * q] = tmp[2 * r]; tmp[2 * q + 1] = tmp[2 * r + 1];
                tmp[2 * r] = a; tmp[2 * r + 1] = b;
            }
        }
        for (int len = 2; len <= n; len <<= 1) {
            int h = len >> 1;
            for (int m = 0; m < h; ++m) {
                double ang = sign * 2.0 * PI * (double)m / (double)len;
                double wr = cos(ang), wi = sin(ang);
                for (int b = 0; b < n; b += len) {
                    double *p = tmp + 2 * (b + m), *q2 = tmp + 2 * (b + m + h);
                    double tr = q2[0] * wr - q2[1] * wi, ti = q2[0] * wi + q2[1] * wr;
                    q2[0] = p[0] - tr; q2[1] = p[1] - ti;
                    p[0] = p[0] + tr; p[1] = p[1] + ti;
                }
            }
        }
        for (int q = 0; q < n; ++q) { x[2 * s * q] = tmp[2 * q]; x[2 * s * q + 1] = tmp[2 * q + 1]; }
    } else {
        for (int m = 0; m < n; ++m) {
            double sr = 0.0, si = 0.0;
            for (int q = 0; q < n; ++q) {
                double ang = sign * 2.0 * PI * (double)((long)m * q % n) / (double)n;
                double wr = cos(ang), wi = sin(ang);
                sr += x[2 * s * q] * wr - x[2 * s * q + 1] * wi;
                si += x[2 * s * q] * wi + x[2 * s * q + 1] * wr;
            }
            tmp[2 * m] = sr; tmp[2 * m + 1] = si;
        }
        for (int q = 0; q < n; ++q) { x[2 * s * q] = tmp[2 * q]; x[2 * s * q + 1] = tmp[2 * q + 1]; }
    }
}

/* pressure_backend_fftw.f90:22-49,68-72: batched 2-D r2c over (z,x), x fastest, one plane per j,
 * unnormalised, half spectrum kx = 0..nx/2.  hat(kx,kz,j) interleaved complex. */
void mdo_fft_r2c_xz(const mdo_grid *g, const double *rhs, double *hat) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = nx / 2 + 1;
#pragma omp parallel
    {
        double *line = (double *)malloc(sizeof(double) * 2 * (nx > nz ? nx : nz));
        double *tmp = (double *)malloc(sizeof(double) * 2 * (nx > nz ? nx : nz));
#pragma omp for schedule(static)
        for (int j = 0; j < ny; ++j) {
            double *h = hat + (size_t)2 * nxh * nz * j;
            for (int k = 0; k < nz; ++k) {
                const double *r = rhs + (size_t)nx * (k + (size_t)nz * j);
                for (int i = 0; i < nx; ++i) { line[2 * i] = r[i]; line[2 * i + 1] = 0.0; }
                fft_line(nx, line, 1, -1, tmp);
                for (int m = 0; m < nxh; ++m) {
                    h[2 * ((size_t)m + (size_t)nxh * k)] = line[2 * m];
                    h[2 * ((size_t)m + (size_t)nxh * k) + 1] = line[2 * m + 1];
                }
            }
            for (int m = 0; m < nxh; ++m) fft_line(nz, h + 2 * m, (size_t)nxh, -1, tmp);
        }
        free(line);
        free(tmp);
    }
}

/* pressure_backend_fftw.f90:46-48,74-78: 2-D c2r, unnormalised.  The x-direction c2r uses the
 * real parts of the kx=0 and kx=nx/2 coefficients (a Hermitian half spectrum carries no more). */
void mdo_fft_c2r_xz(const mdo_grid *g, const double *hat_in, double *rhs) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = nx / 2 + 1;
#pragma omp parallel
    {
        double *plane = (double *)malloc(sizeof(double) * 2 * nxh * nz);
        double *line = (double *)malloc(sizeof(double) * 2 * (nx > nz ? nx : nz));
        double *tmp = (double *)malloc(sizeof(double) * 2 * (nx > nz ? nx : nz));
#pragma omp for schedule(static)
        for (int j = 0; j < ny; ++j) {
            memcpy(plane, hat_in + (size_t)2 * nxh * nz * j, sizeof(double) * 2 * nxh * nz);
            for (int m = 0; m < nxh; ++m) fft_line(nz, plane + 2 * m, (size_t)nxh, +1, tmp);
            for (int k = 0; k < nz; ++k) {
                const double *h = plane + (size_t)2 * nxh * k;
                for (int m = 0; m < nx; ++m) {
                    if (m == 0 || (2 * m == nx)) { line[2 * m] = h[2 * m]; line[2 * m + 1] = 0.0; }
                    else if (m < nxh) { line[2 * m] = h[2 * m]; line[2 * m + 1] = h[2 * m + 1]; }
                    else { line[2 * m] = h[2 * (nx - m)]; line[2 * m + 1] = -h[2 * (nx - m) + 1]; }
                }
                fft_line(nx, line, 1, +1, tmp);
                double *r = rhs + (size_t)nx * (k + (size_t)nz * j);
                for (int i = 0; i < nx; ++i) r[i] = line[2 * i];
            }
        }
        free(plane);
        free(line);
        free(tmp);
    }
}

/* pressure_fft.f90:63,79 compute den_inv = one_c/denom as a COMPLEX division.  The compiled
 * reference evaluates it with the scaled C99 Annex G algorithm (divide by c*c after scaling c by
 * 2^-logb(c)), which can differ from 1/c in the last bit; this is that algorithm for a real c. */
static double recip_complex_div(double c) {
    double lb = logb(fabs(c));
    int e = 0;
    if (isfinite(lb)) { e = (int)lb; c = scalbn(c, -e); }
    double denom = c * c;
    return scalbn(c / denom, -e);
}

/* pressure_fft.f90:20-100.  Arrays cp/den_inv(nxh,ny,nz) real (the reference stores them as
 * complex with zero imaginary part; the products below are identical).  a_row(ny) carries the
 * sub-diagonal, constant 1/dy^2 on the reference grid. */
void mdo_tridiag_coefficients(const mdo_grid *g, double *cp, double *den_inv, double *a_row) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = nx / 2 + 1;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
    const double dyi2 = 1.0 / (dy * dy);
    for (int j = 1; j <= ny; ++j) a_row[j - 1] = g->ihc ? g->ihv[j] * g->ihc[j] : dyi2;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ikx = 1; ikx <= nxh; ++ikx)
        for (int ikz = 1; ikz <= nz; ++ikz) {
            int kx = ikx - 1;
            double s = sin(PI * (double)kx / (double)nx);
            double kx_ph = (4.0 / (dx * dx)) * (s * s);
            int kz = (ikz <= nz / 2 + 1) ? ikz - 1 : ikz - nz - 1;
            double sz = sin(PI * (double)kz / (double)nz);
            double kz_ph = (4.0 / (dz * dz)) * (sz * sz);
            double k_tot = kx_ph + kz_ph;
            int zero_mode = (kx == 0 && kz == 0);
#define CI(jj) ((size_t)(ikx - 1) + (size_t)nxh * ((size_t)((jj) - 1) + (size_t)ny * (size_t)(ikz - 1)))
            if (!g->ihc) {
                if (zero_mode) {
                    den_inv[CI(1)] = 1.0;
                    cp[CI(1)] = 0.0;
                } else {
                    double b = -dyi2 - k_tot;
                    den_inv[CI(1)] = recip_complex_div(b);
                    cp[CI(1)] = dyi2 * den_inv[CI(1)];
                }
                for (int jj = 2; jj <= ny; ++jj) {
                    double b, c;
                    if (jj < ny) { b = -2.0 * dyi2 - k_tot; c = dyi2; }
                    else { b = -dyi2 - k_tot; c = 0.0; }
                    double denom = b - dyi2 * cp[CI(jj - 1)];
                    den_inv[CI(jj)] = recip_complex_div(denom);
                    cp[CI(jj)] = (jj < ny) ? c * den_inv[CI(jj)] : 0.0;
                }
            } else {
                /* row j: a_j phi_{j-1} + b_j phi_j + c_j phi_{j+1}; walls drop a_1 and c_ny */
                const double *ihc = g->ihc, *ihv = g->ihv;
                if (zero_mode) {
                    den_inv[CI(1)] = 1.0;
                    cp[CI(1)] = 0.0;
                } else {
                    double c1 = ihv[2] * ihc[1];
                    double b = -c1 - k_tot;
                    den_inv[CI(1)] = recip_complex_div(b);
                    cp[CI(1)] = c1 * den_inv[CI(1)];
                }
                for (int jj = 2; jj <= ny; ++jj) {
                    double aj = ihv[jj] * ihc[jj];
                    double cj = (jj < ny) ? ihv[jj + 1] * ihc[jj] : 0.0;
                    double b = -(aj + cj) - k_tot;
                    double denom = b - aj * cp[CI(jj - 1)];
                    den_inv[CI(jj)] = recip_complex_div(denom);
                    cp[CI(jj)] = (jj < ny) ? cj * den_inv[CI(jj)] : 0.0;
                }
            }
#undef CI
        }
}

/* pressure_fft.f90:102-148: forward elimination and back substitution per (kx,kz) line.
 * hat(kx,kz,j) interleaved complex; dp(nxh,ny,nz) complex work array. */
void mdo_tridiag_solve(const mdo_grid *g, const double *cp, const double *den_inv,
                       const double *a_row, double *hat, double *dp) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = nx / 2 + 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ikx = 1; ikx <= nxh; ++ikx)
        for (int ikz = 1; ikz <= nz; ++ikz) {
#define CI(jj) ((size_t)(ikx - 1) + (size_t)nxh * ((size_t)((jj) - 1) + (size_t)ny * (size_t)(ikz - 1)))
#define HI(jj) ((size_t)(ikx - 1) + (size_t)nxh * ((size_t)(ikz - 1) + (size_t)nz * (size_t)((jj) - 1)))
            if (ikx == 1 && ikz == 1) {
                dp[2 * CI(1)] = 0.0;
                dp[2 * CI(1) + 1] = 0.0;
            } else {
                dp[2 * CI(1)] = hat[2 * HI(1)] * den_inv[CI(1)];
                dp[2 * CI(1) + 1] = hat[2 * HI(1) + 1] * den_inv[CI(1)];
            }
            for (int jj = 2; jj <= ny; ++jj) {
                double a = a_row[jj - 1];
                dp[2 * CI(jj)] = (hat[2 * HI(jj)] - a * dp[2 * CI(jj - 1)]) * den_inv[CI(jj)];
                dp[2 * CI(jj) + 1] = (hat[2 * HI(jj) + 1] - a * dp[2 * CI(jj - 1) + 1]) * den_inv[CI(jj)];
            }
            hat[2 * HI(ny)] = dp[2 * CI(ny)];
            hat[2 * HI(ny) + 1] = dp[2 * CI(ny) + 1];
            for (int jj = ny - 1; jj >= 1; --jj) {
                hat[2 * HI(jj)] = dp[2 * CI(jj)] - cp[CI(jj)] * hat[2 * HI(jj + 1)];
                hat[2 * HI(jj) + 1] = dp[2 * CI(jj) + 1] - cp[CI(jj)] * hat[2 * HI(jj + 1) + 1];
            }
#undef CI
#undef HI
        }
}

/* pressure_solver.f90:117-173 */
void mdo_fft_pressure_correction(const mdo_grid *g, const double *phi,
                                 const double *us, const double *vs, const double *ws,
                                 const double *cu, const double *cv, const double *cw, double dtg,
                                 double *un, double *vn, double *wn, double *pn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
#define RI(i, k, j) ((size_t)((i) - 1) + (size_t)nx * ((size_t)((k) - 1) + (size_t)nz * (size_t)((j) - 1)))
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                int im = i - 1, km = k - 1;
                if (im < 1) im = nx;
                if (km < 1) km = nz;
                size_t c = GI(i, j, k);
                un[c] = (us[c] - dtg * (phi[RI(i, k, j)] - phi[RI(im, k, j)]) / dx) / (1.0 + dtg * (cu ? cu[c] : 0.0));
                wn[c] = (ws[c] - dtg * (phi[RI(i, k, j)] - phi[RI(i, km, j)]) / dz) / (1.0 + dtg * (cw ? cw[c] : 0.0));
                pn[c] = pn[c] + phi[RI(i, k, j)];
                if (j >= 2) {
                    double gy = g->ihv ? dtg * (phi[RI(i, k, j)] - phi[RI(i, k, j - 1)]) * g->ihv[j]
                                       : dtg * (phi[RI(i, k, j)] - phi[RI(i, k, j - 1)]) / dy;
                    vn[c] = (vs[c] - gy) / (1.0 + dtg * (cv ? cv[c] : 0.0));
                }
            }
#undef RI
}

void mdo_projection_fft(const mdo_grid *g, double dtg,
                        const double *us, const double *vs, const double *ws,
                        const double *cu, const double *cv, const double *cw,
                        double *un, double *vn, double *wn, double *pn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz, nxh = nx / 2 + 1;
    size_t n = (size_t)nx * ny * nz, nh = (size_t)nxh * ny * nz;
    double *rhs = (double *)malloc(sizeof(double) * n);
    double *hat = (double *)malloc(sizeof(double) * 2 * nh);
    double *dp = (double *)malloc(sizeof(double) * 2 * nh);
    double *cp = (double *)malloc(sizeof(double) * nh);
    double *di = (double *)malloc(sizeof(double) * nh);
    double *ar = (double *)malloc(sizeof(double) * ny);
    mdo_tridiag_coefficients(g, cp, di, ar);
    mdo_build_fft_rhs(g, us, vs, ws, dtg, rhs);
    mdo_fft_r2c_xz(g, rhs, hat);
    mdo_tridiag_solve(g, cp, di, ar, hat, dp);
    mdo_fft_c2r_xz(g, hat, rhs);
    double scale = 1.0 / (double)(nx * nz); /* pressure_backend_fftw.f90:60 */
    for (size_t q = 0; q < n; ++q) rhs[q] = rhs[q] * scale;
    mdo_fft_pressure_correction(g, rhs, us, vs, ws, cu, cv, cw, dtg, un, vn, wn, pn);
    free(rhs); free(hat); free(dp); free(cp); free(di); free(ar);
}

/* ------------------------------------------------------------------ red-black SOLA */
/* pressure_redblack.f90:27-254, restated statement by statement. */
void mdo_sola_redblack(const mdo_grid *g, int niter, double sor, double dtg,
                       const double *cu, const double *cv, const double *cw,
                       const double *us, const double *vs, const double *ws,
                       double *un, double *vn, double *wn, double *pn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double dx = g->dx, dy = g->dy, dz = g->dz;
    const int nActive = nx / 2;
    const int *t = g->bc_type;
    const double *v = g->bc_value;
    const int per1 = g->periodic[0], per2 = g->periodic[1], per3 = g->periodic[2];
    const double idx = 1.0 / dx, idy = 1.0 / dy, idz = 1.0 / dz;
    const double idx2 = idx * idx, idy2 = idy * idy, idz2 = idz * idz;
    const double idt = 1.0 / dtg;
    size_t n = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    memcpy(un, us, n * sizeof(double));
    memcpy(vn, vs, n * sizeof(double));
    memcpy(wn, ws, n * sizeof(double));
#define T(d, s, q) t[MDO_BCI(d, s, q)]
#define V(d, s, q) v[MDO_BCI(d, s, q)]
#define MU(c, q) (1.0 / (1.0 + dtg * ((c) ? (c)[q] : 0.0)))
    for (int it = 0; it < niter; ++it)
        for (int parity = 0; parity <= 1; ++parity) {
#pragma omp parallel for collapse(2) schedule(static)
            for (int k = 1; k <= nz; ++k)
                for (int j = 1; j <= ny; ++j)
                    for (int ni = 1; ni <= nActive; ++ni) {
                        int jp = j + 1;
                        if (per2 && jp > ny) jp = 1;
                        int kp = k + 1;
                        if (per3 && kp > nz) kp = 1;
                        int iLo = 1 + ((j + k + parity) % 2);
                        int i = iLo + 2 * (ni - 1);
                        int ip = i + 1;
                        if (per1 && ip > nx) ip = 1;
                        size_t c = GI(i, j, k);
                        double mu_u_i = MU(cu, c), mu_u_ip = MU(cu, GI(ip, j, k));
                        double mu_v_j = MU(cv, c), mu_v_jp = MU(cv, GI(i, jp, k));
                        double mu_w_k = MU(cw, c), mu_w_kp = MU(cw, GI(i, j, kp));
                        double denom = (mu_u_i + mu_u_ip) * idx2 + (mu_w_k + mu_w_kp) * idz2 + (mu_v_j + mu_v_jp) * idy2;
                        if (!per1) {
                            if (i < 2) un[c] = T(1, 0, 1) == 0 ? V(1, 0, 1) : un[GI(ip, j, k)] - V(1, 0, 1) * dx;
                            if (i >= nx) un[GI(ip, j, k)] = T(1, 1, 1) == 0 ? V(1, 1, 1) : un[GI(nx, j, k)] + V(1, 1, 1) * dx;
                        }
                        if (!per2) {
                            if (j < 2) vn[c] = T(2, 0, 2) == 0 ? V(2, 0, 2) : vn[GI(i, jp, k)] - V(2, 0, 2) * dy;
                            if (j >= ny) vn[GI(i, jp, k)] = T(2, 1, 2) == 0 ? V(2, 1, 2) : vn[c] + V(2, 1, 2) * dy;
                        }
                        if (!per3) {
                            if (k < 2) wn[c] = T(3, 0, 3) == 0 ? V(3, 0, 3) : wn[GI(i, j, kp)] - V(3, 0, 3) * dz;
                            if (k >= nz) wn[GI(i, j, kp)] = T(3, 1, 3) == 0 ? V(3, 1, 3) : wn[c] + V(3, 1, 3) * dz;
                        }
                        double phi = ((un[GI(ip, j, k)] - un[c]) * idx + (vn[GI(i, jp, k)] - vn[c]) * idy
                                      + (wn[GI(i, j, kp)] - wn[c]) * idz) * (-sor / denom);
                        pn[c] = pn[c] + phi * idt;
                        un[c] = un[c] - phi * idx * mu_u_i;
                        if (!per1 && i == 1) {
                            if (T(1, 0, 0) == 0) un[c] = un[c] + V(1, 0, 0) * idx * mu_u_i;
                            else un[c] = un[c] + (phi * idx - V(1, 0, 0)) * mu_u_i;
                        }
                        un[GI(ip, j, k)] = un[GI(ip, j, k)] + phi * idx * mu_u_ip;
                        if (!per1 && i == nx) {
                            if (T(1, 1, 0) == 0) un[GI(ip, j, k)] = un[GI(ip, j, k)] - V(1, 1, 0) * idx * mu_u_ip;
                            else un[GI(ip, j, k)] = un[GI(ip, j, k)] + (V(1, 1, 0) - phi * idx) * mu_u_ip;
                        }
                        vn[c] = vn[c] - phi * idy * mu_v_j;
                        if (!per2 && j == 1) {
                            if (T(2, 0, 0) == 0) vn[c] = vn[c] + V(2, 0, 0) * idy * mu_v_j;
                            else vn[c] = vn[c] + (phi * idy - V(2, 0, 0)) * mu_v_j;
                        }
                        vn[GI(i, jp, k)] = vn[GI(i, jp, k)] + phi * idy * mu_v_jp;
                        if (!per2 && j == ny) {
                            if (T(2, 1, 0) == 0) vn[GI(i, jp, k)] = vn[GI(i, jp, k)] - V(2, 1, 0) * idy * mu_v_jp;
                            else vn[GI(i, jp, k)] = vn[GI(i, jp, k)] + (V(2, 1, 0) - phi * idy) * mu_v_jp;
                        }
                        wn[c] = wn[c] - phi * idz * mu_w_k;
                        if (!per3 && k == 1) {
                            if (T(3, 0, 0) == 0) wn[c] = wn[c] + V(3, 0, 0) * idz * mu_w_k;
                            else wn[c] = wn[c] + (phi * idz - V(3, 0, 0)) * mu_w_k;
                        }
                        wn[GI(i, j, kp)] = wn[GI(i, j, kp)] + phi * idz * mu_w_kp;
                        if (!per3 && k == nz) {
                            /* pressure_redblack.f90:229-230: no mu factor on the Dirichlet branch */
                            if (T(3, 1, 0) == 0) wn[GI(i, j, kp)] = wn[GI(i, j, kp)] - V(3, 1, 0) * idz;
                            else wn[GI(i, j, kp)] = wn[GI(i, j, kp)] + (V(3, 1, 0) - phi * idz) * mu_w_kp;
                        }
                    }
        }
#undef T
#undef V
}

/* ------------------------------------------------------------------ pressure-correction SOR */
/* The SOLA sweep above updates the six face velocities of a cell by phi*mu/h.  Summing those
 * updates, u = u* + mu * (Phi_west - Phi_east)/h with Phi the accumulated phi, so the residual
 * div(u) = div(u*) - L_mu(Phi).  The sweeps below iterate on Phi alone (x,z periodic; y walls
 * with Dirichlet v and Neumann p, or y periodic) and keep the reference's relaxation
 * denominator, wall faces included, so the iterates are the SOLA iterates. */
#define MUQ(c, q) (1.0 / (1.0 + dtg * ((c) ? (c)[q] : 0.0)))

void mdo_sor_div_rhs(const mdo_grid *g, const double *un, const double *vn, const double *wn,
                     double *rhs) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy, idz = 1.0 / g->dz;
    const int per2 = g->periodic[1];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                int ip = i == nx ? 1 : i + 1;
                int kp = k == nz ? 1 : k + 1;
                int jp = (per2 && j == ny) ? 1 : j + 1;
                size_t c = GI(i, j, k);
                double dvy = g->ihc ? (vn[GI(i, jp, k)] - vn[c]) * g->ihc[j] : (vn[GI(i, jp, k)] - vn[c]) * idy;
                rhs[II(i, j, k)] = (un[GI(ip, j, k)] - un[c]) * idx + dvy + (wn[GI(i, j, kp)] - wn[c]) * idz;
            }
}

static inline double sor_cell_residual(const mdo_grid *g, int i, int j, int k, double dtg,
                                       const double *cu, const double *cv, const double *cw,
                                       const double *rhs, const double *phi, double *denom_out) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int per2 = g->periodic[1];
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy, idz = 1.0 / g->dz;
    const double idx2 = idx * idx, idy2 = idy * idy, idz2 = idz * idz;
    int ip = i == nx ? 1 : i + 1, im = i == 1 ? nx : i - 1;
    int kp = k == nz ? 1 : k + 1, km = k == 1 ? nz : k - 1;
    int jp = j + 1, jm = j - 1;
    if (per2) { if (jp > ny) jp = 1; if (jm < 1) jm = ny; }
    size_t c = GI(i, j, k);
    double mu_u_i = MUQ(cu, c), mu_u_ip = MUQ(cu, GI(ip, j, k));
    double mu_v_j = MUQ(cv, c), mu_v_jp = MUQ(cv, GI(i, jp, k));
    double mu_w_k = MUQ(cw, c), mu_w_kp = MUQ(cw, GI(i, j, kp));
    double p0 = phi[II(i, j, k)];
    double sn = (!per2 && j == ny) ? 0.0 : mu_v_jp * (phi[II(i, jp, k)] - p0);
    double ss = (!per2 && j == 1) ? 0.0 : mu_v_j * (p0 - phi[II(i, jm, k)]);
    double lx = (mu_u_ip * (phi[II(ip, j, k)] - p0) - mu_u_i * (p0 - phi[II(im, j, k)])) * idx2;
    double lz = (mu_w_kp * (phi[II(i, j, kp)] - p0) - mu_w_k * (p0 - phi[II(i, j, km)])) * idz2;
    double ly, denom;
    if (!g->ihc) {
        ly = (sn - ss) * idy2;
        denom = (mu_u_i + mu_u_ip) * idx2 + (mu_w_k + mu_w_kp) * idz2 + (mu_v_j + mu_v_jp) * idy2;
    } else {
        int jn = j + 1; /* north face index in the metric arrays */
        double an = g->ihv[jn] * g->ihc[j], as = g->ihv[j] * g->ihc[j];
        ly = sn * an - ss * as;
        denom = (mu_u_i + mu_u_ip) * idx2 + (mu_w_k + mu_w_kp) * idz2 + (mu_v_j * as + mu_v_jp * an);
    }
    if (denom_out) *denom_out = denom;
    return rhs[II(i, j, k)] - (lx + ly + lz);
}

void mdo_sor_sweep(const mdo_grid *g, int color, double omega, double dtg,
                   const double *cu, const double *cv, const double *cw,
                   const double *rhs, double *phi) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j) {
            int i0 = 1 + ((j + k + color) % 2);
            for (int i = i0; i <= nx; i += 2) {
                double denom;
                double r = sor_cell_residual(g, i, j, k, dtg, cu, cv, cw, rhs, phi, &denom);
                phi[II(i, j, k)] = phi[II(i, j, k)] + r * (-omega / denom);
            }
        }
}

/* The relaxation of mdo_sor_sweep applied to an explicit list of cells (i, j, k triples, 1-based),
 * one after another in the order given.  Test hook: a colour's cells relaxed in any order must
 * give the sweep's result bit for bit (no cell of one colour reads another of the same colour). */
void mdo_sor_relax_cells(const mdo_grid *g, double omega, double dtg,
                         const double *cu, const double *cv, const double *cw,
                         const double *rhs, double *phi, const int *cells, int ncells) {
    const int nx = g->nx, ny = g->ny;
    for (int q = 0; q < ncells; ++q) {
        const int i = cells[3 * q], j = cells[3 * q + 1], k = cells[3 * q + 2];
        double denom;
        double r = sor_cell_residual(g, i, j, k, dtg, cu, cv, cw, rhs, phi, &denom);
        phi[II(i, j, k)] = phi[II(i, j, k)] + r * (-omega / denom);
    }
}

double mdo_sor_residual_inf(const mdo_grid *g, double dtg,
                            const double *cu, const double *cv, const double *cw,
                            const double *rhs, const double *phi) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double m = 0.0;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double r = fabs(sor_cell_residual(g, i, j, k, dtg, cu, cv, cw, rhs, phi, NULL));
                if (r > m) m = r;
            }
    return m;
}

void mdo_sor_apply(const mdo_grid *g, double dtg,
                   const double *cu, const double *cv, const double *cw,
                   const double *phi, double *un, double *vn, double *wn, double *pn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy, idz = 1.0 / g->dz;
    const double idt = 1.0 / dtg;
    const int per2 = g->periodic[1];
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                int im = i == 1 ? nx : i - 1, km = k == 1 ? nz : k - 1;
                size_t c = GI(i, j, k);
                double p0 = phi[II(i, j, k)];
                un[c] = un[c] + MUQ(cu, c) * (phi[II(im, j, k)] - p0) * idx;
                wn[c] = wn[c] + MUQ(cw, c) * (phi[II(i, j, km)] - p0) * idz;
                if (j >= 2 || per2) {
                    int jm = j == 1 ? ny : j - 1;
                    double gy = g->ihv ? (phi[II(i, jm, k)] - p0) * g->ihv[j] : (phi[II(i, jm, k)] - p0) * idy;
                    vn[c] = vn[c] + MUQ(cv, c) * gy;
                }
                pn[c] = pn[c] + p0 * idt;
            }
}

void mdo_projection_sor(const mdo_grid *g, double omega, double tol, int max_iter, int check_every,
                        double dtg, const double *cu, const double *cv, const double *cw,
                        const double *us, const double *vs, const double *ws,
                        double *un, double *vn, double *wn, double *pn,
                        double *phi, double *rhs, mdo_sor_stats *st) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    size_t n = (size_t)(nx + 2) * (ny + 2) * (nz + 2);
    memcpy(un, us, n * sizeof(double)); /* pressure_redblack.f90:64-84 */
    memcpy(vn, vs, n * sizeof(double));
    memcpy(wn, ws, n * sizeof(double));
    if (!g->periodic[1]) {
        for (int k = 1; k <= nz; ++k)
            for (int i = 1; i <= nx; ++i) {
                vn[GI(i, 1, k)] = g->bc_value[MDO_BCI(2, 0, 2)];
                vn[GI(i, ny + 1, k)] = g->bc_value[MDO_BCI(2, 1, 2)];
            }
    }
    mdo_sor_div_rhs(g, un, vn, wn, rhs);
    memset(phi, 0, sizeof(double) * (size_t)nx * ny * nz);
    int it = 0, conv = 0;
    double r = mdo_sor_residual_inf(g, dtg, cu, cv, cw, rhs, phi);
    st->div0_inf = r;
    if (check_every < 1) check_every = 1;
    while (1) {
        if (r < tol) { conv = 1; break; }
        if (it >= max_iter) break;
        for (int q = 0; q < check_every && it < max_iter; ++q, ++it) {
            mdo_sor_sweep(g, 0, omega, dtg, cu, cv, cw, rhs, phi);
            mdo_sor_sweep(g, 1, omega, dtg, cu, cv, cw, rhs, phi);
        }
        r = mdo_sor_residual_inf(g, dtg, cu, cv, cw, rhs, phi);
    }
    mdo_sor_apply(g, dtg, cu, cv, cw, phi, un, vn, wn, pn);
    st->iterations = it;
    st->converged = conv;
    st->div_inf = mdo_div_inf(g, un, vn, wn);
}

/* ------------------------------------------------------------------ norms, flow rate */
double mdo_div_inf(const mdo_grid *g, const double *un, const double *vn, const double *wn) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const double idx = 1.0 / g->dx, idy = 1.0 / g->dy, idz = 1.0 / g->dz;
    const int per2 = g->periodic[1];
    double m = 0.0;
#pragma omp parallel for collapse(2) reduction(max : m) schedule(static)
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                int ip = g->periodic[0] && i == nx ? 1 : i + 1;
                int kp = g->periodic[2] && k == nz ? 1 : k + 1;
                int jp = (per2 && j == ny) ? 1 : j + 1;
                size_t c = GI(i, j, k);
                double dvy = g->ihc ? (vn[GI(i, jp, k)] - vn[c]) * g->ihc[j] : (vn[GI(i, jp, k)] - vn[c]) * idy;
                double d = fabs((un[GI(ip, j, k)] - un[c]) * idx + dvy + (wn[GI(i, j, kp)] - wn[c]) * idz);
                if (d > m) m = d;
            }
    return m;
}

/* The two sums of the constant-flow-rate control, S1 = sum u h_j and S2 = sum mu_u h_j over interior
 * u points, in one fixed order that is part of their definition (this control has no reference
 * counterpart): 512 partial sums, partial b taking the x rows r = b, b + 512, ... (r = (j-1) +
 * ny (k-1)) split over 256 column lanes (lane t: columns t+1, t+257, ...), each lane adding its
 * values row by row; the lanes of a partial, and then the partials (lane t: partials t, t + 256),
 * are folded by a pairwise tree (s[t] += s[t + w], w = 128 ... 1).  The GPU kernels add in this
 * order too, so the control's shift is the same double on both. */
#define FR_PARTS 512
#define FR_LANES 256
static void fr_fold(double *s) {
    for (int w = FR_LANES / 2; w > 0; w >>= 1)
        for (int t = 0; t < w; ++t) s[t] += s[t + w];
}

static void fr_sums(const mdo_grid *g, const double *u, const double *cu, double dtg, double *S1, double *S2) {
    const int nx = g->nx, ny = g->ny, nrows = g->ny * g->nz;
    double p1[FR_PARTS], p2[FR_PARTS];
#pragma omp parallel for schedule(static)
    for (int b = 0; b < FR_PARTS; ++b) {
        double s1[FR_LANES], s2[FR_LANES];
        for (int t = 0; t < FR_LANES; ++t) {
            double a = 0.0, c2 = 0.0;
            for (int r = b; r < nrows; r += FR_PARTS) {
                const int j = r % ny + 1, k = r / ny + 1;
                const double h = g->ihc ? 1.0 / g->ihc[j] : g->dy;
                for (int i = t + 1; i <= nx; i += FR_LANES) {
                    const size_t q = GI(i, j, k);
                    a += u[q] * h;
                    c2 += MUQ(cu, q) * h;
                }
            }
            s1[t] = a;
            s2[t] = c2;
        }
        fr_fold(s1);
        fr_fold(s2);
        p1[b] = s1[0];
        p2[b] = s2[0];
    }
    double s1[FR_LANES], s2[FR_LANES];
    for (int t = 0; t < FR_LANES; ++t) {
        double a = 0.0, c2 = 0.0;
        for (int q = t; q < FR_PARTS; q += FR_LANES) {
            a += p1[q];
            c2 += p2[q];
        }
        s1[t] = a;
        s2[t] = c2;
    }
    fr_fold(s1);
    fr_fold(s2);
    *S1 = s1[0];
    *S2 = s2[0];
}

/* The order of the same sums inside a time step (the GPU forms them in its momentum kernel, the
 * flow-rate control's only pass over u* there): tiles of 64 x ty columns over kz planes, tile b =
 * tx + nbx (ty_ + nby tz).  ty = 4 when fewer than 2048 tiles of 64 x 8 columns cover the grid,
 * else 8; kz = 12 when its tiles fill their last round of 4096 / ty better than those of kz = 8,
 * else 8 (the GPU's momentum tiling).
 * In a tile the lane of column (i, j) adds its cells plane by plane; each row of 64 lanes (lanes
 * beyond nx or ny adding nothing) is folded by a pairwise tree (s[l] += s[l + w], w = 32 ... 1) and
 * the rows' sums are added in row order.  The tile sums are folded as the partials above (lane t:
 * tiles t, t + 256, ...; pairwise tree over the 256 lanes). */
#define FR_TX 64
static int fr_tile_rows(const mdo_grid *g, int kz) {
    return (long)((g->nx + FR_TX - 1) / FR_TX) * ((g->ny + 7) / 8) * ((g->nz + kz - 1) / kz) < 2048 ? 4 : 8;
}

static double fr_tile_fill(const mdo_grid *g, int kz) {
    const int ty = fr_tile_rows(g, kz);
    const long n = (long)((g->nx + FR_TX - 1) / FR_TX) * ((g->ny + ty - 1) / ty) * ((g->nz + kz - 1) / kz);
    const long slots = 4096 / ty, rounds = (n + slots - 1) / slots;
    return (double)n / (double)(rounds * slots);
}

static void fr_sums_step(const mdo_grid *g, const double *u, const double *cu, double dtg, double *S1, double *S2) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int kz = fr_tile_fill(g, 12) > fr_tile_fill(g, 8) ? 12 : 8;
    const int nbx = (nx + FR_TX - 1) / FR_TX, nbz = (nz + kz - 1) / kz;
    const int ty = fr_tile_rows(g, kz);
    const int nby = (ny + ty - 1) / ty, nparts = nbx * nby * nbz;
    double *p1 = (double *)malloc(sizeof(double) * 2 * (size_t)nparts), *p2 = p1 + nparts;
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nparts; ++b) {
        const int tx = b % nbx, tr = (b / nbx) % nby, tz = b / (nbx * nby);
        double a = 0.0, c2 = 0.0;
        for (int r = 0; r < ty; ++r) {
            const int j = tr * ty + r + 1;
            double s1[FR_TX], s2[FR_TX];
            for (int l = 0; l < FR_TX; ++l) {
                const int i = tx * FR_TX + l + 1;
                double x = 0.0, y = 0.0;
                if (i <= nx && j <= ny) {
                    const double h = g->ihc ? 1.0 / g->ihc[j] : g->dy;
                    for (int k = tz * kz + 1; k <= (tz + 1) * kz && k <= nz; ++k) {
                        const size_t q = GI(i, j, k);
                        x += u[q] * h;
                        y += MUQ(cu, q) * h;
                    }
                }
                s1[l] = x;
                s2[l] = y;
            }
            for (int w = FR_TX / 2; w > 0; w >>= 1)
                for (int l = 0; l < w; ++l) {
                    s1[l] += s1[l + w];
                    s2[l] += s2[l + w];
                }
            a += s1[0];
            c2 += s2[0];
        }
        p1[b] = a;
        p2[b] = c2;
    }
    double s1[FR_LANES], s2[FR_LANES];
    for (int t = 0; t < FR_LANES; ++t) {
        double a = 0.0, c2 = 0.0;
        for (int q = t; q < nparts; q += FR_LANES) {
            a += p1[q];
            c2 += p2[q];
        }
        s1[t] = a;
        s2[t] = c2;
    }
    fr_fold(s1);
    fr_fold(s2);
    free(p1);
    *S1 = s1[0];
    *S2 = s2[0];
}

double mdo_bulk_u_step(const mdo_grid *g, const double *u, double height) {
    const double dtg = 0.0;
    double s1, s2;
    fr_sums_step(g, u, NULL, dtg, &s1, &s2);
    return s1 / ((double)g->nx * (double)g->nz * height);
}

double mdo_bulk_u(const mdo_grid *g, const double *u, double height) {
    const double dtg = 0.0;
    double s1, s2;
    fr_sums(g, u, NULL, dtg, &s1, &s2);
    return s1 / ((double)g->nx * (double)g->nz * height);
}

/* Constant flow rate: a uniform x-force increment chosen so that the predictor's bulk velocity
 * equals the target; like the body force in the momentum equation it is damped by mu_u. */
static double flowrate_shift_with(const mdo_grid *g, double *us, const double *cu, double dtg, double target,
                                  double height, int step_order) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    double s1, s2;
    if (step_order)
        fr_sums_step(g, us, cu, dtg, &s1, &s2);
    else
        fr_sums(g, us, cu, dtg, &s1, &s2);
    const double norm = (double)nx * (double)nz * height;
    const double ub = s1 / norm;
    const double sm = s2 / norm;
    const double du = (target - ub) / sm;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) us[GI(i, j, k)] = us[GI(i, j, k)] + du * MUQ(cu, GI(i, j, k));
    return du;
}

double mdo_flowrate_shift(const mdo_grid *g, double *us, const double *cu, double dtg, double target,
                          double height) {
    return flowrate_shift_with(g, us, cu, dtg, target, height, 0);
}

/* ------------------------------------------------------------------ immersed boundary */
/* ibm.f90:84-99 (kind 0) and the two-sided wavy channel (kind 1). */
int mdo_is_in_body(const mdo_grid *g, const mdo_geometry *geo, const double x[3]) {
    if (geo->kind == 0) {
        double y_body = geo->amp_x * 0.5 * (1.0 + sin(2.0 * PI * (double)geo->n_wave_x * x[0] / g->lx + geo->phase_x)) + 2 * g->dy;
        return x[1] < y_body;
    }
    double h = geo->half_height - geo->A * cos(2.0 * PI * x[0] / geo->lambda);
    return fabs(x[1] - 0.5 * g->ly) > h;
}

/* ibm.f90:101-130 */
void mdo_bisection(const mdo_grid *g, const mdo_geometry *geo, const double xa_in[3], double xb[3]) {
    double xa[3] = {xa_in[0], xa_in[1], xa_in[2]}, xm[3] = {0, 0, 0};
    for (int it = 1; it <= 200; ++it) {
        for (int d = 0; d < 3; ++d) xm[d] = 0.5 * (xa[d] + xb[d]);
        double e0 = xb[0] - xa[0], e1 = xb[1] - xa[1], e2 = xb[2] - xa[2];
        if (sqrt(e0 * e0 + e1 * e1 + e2 * e2) < 1.0e-10) break;
        int la = mdo_is_in_body(g, geo, xa);
        int lm = mdo_is_in_body(g, geo, xm);
        if (lm == la) { for (int d = 0; d < 3; ++d) xa[d] = xm[d]; }
        else { for (int d = 0; d < 3; ++d) xb[d] = xm[d]; }
    }
    for (int d = 0; d < 3; ++d) xb[d] = xm[d];
}

/* ibm.f90:133-187 over the ghosted array (assumed-shape index ix = i+1). */
void mdo_set_ibm_coeff(const mdo_grid *g, const mdo_geometry *geo, int dix, int diy, int diz,
                       double *coeff) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    static const int nb[6][3] = {{-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1}};
#pragma omp parallel for collapse(2) schedule(dynamic, 4)
    for (int iz = 1; iz <= nz + 2; ++iz)
        for (int iy = 1; iy <= ny + 2; ++iy)
            for (int ix = 1; ix <= nx + 2; ++ix) {
                double xa[3];
                xa[0] = ((double)ix - (double)dix * 0.5 - 1.5) * g->dx;
                xa[1] = ((double)iy - (double)diy * 0.5 - 1.5) * g->dy;
                xa[2] = ((double)iz - (double)diz * 0.5 - 1.5) * g->dz;
                size_t c = GI(ix - 1, iy - 1, iz - 1);
                double acc = 0.0;
                if (mdo_is_in_body(g, geo, xa)) {
                    acc = 1.0e30;
                } else if (geo->second_order) {
                    for (int q = 0; q < 6; ++q) {
                        double xb[3];
                        xb[0] = ((double)(ix + nb[q][0]) - (double)dix * 0.5 - 1.5) * g->dx;
                        xb[1] = ((double)(iy + nb[q][1]) - (double)diy * 0.5 - 1.5) * g->dy;
                        xb[2] = ((double)(iz + nb[q][2]) - (double)diz * 0.5 - 1.5) * g->dz;
                        double e0 = xb[0] - xa[0], e1 = xb[1] - xa[1], e2 = xb[2] - xa[2];
                        double d0 = sqrt(e0 * e0 + e1 * e1 + e2 * e2);
                        if (mdo_is_in_body(g, geo, xb)) {
                            mdo_bisection(g, geo, xa, xb);
                            double f0 = xb[0] - xa[0], f1 = xb[1] - xa[1], f2 = xb[2] - xa[2];
                            double d = sqrt(f0 * f0 + f1 * f1 + f2 * f2);
                            acc = acc + ((d0 - d) / d) / (d0 * d0);
                        }
                    }
                }
                coeff[c] = acc / g->re;
            }
}

/* ------------------------------------------------------------------ initial condition */
uint64_t mdo_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* counter-based uniform in [0,1): key = seed<<34 | comp<<32, plus the ghosted linear index */
double mdo_uniform01(uint32_t seed, uint32_t comp, uint64_t index) {
    uint64_t key = ((uint64_t)seed << 34) + ((uint64_t)comp << 32) + index;
    return (double)(mdo_splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

/* Poiseuille u = 1.5(1 - (y/h)^2) plus amp*(2U-1) per component at interior faces; with a
 * geometry, samples inside the body are zero and h is the local half height. */
void mdo_initial_field(const mdo_grid *g, const mdo_geometry *geo, int use_geometry,
                       const double *yc, const double *yv,
                       uint32_t seed, double amp, double *us, double *vs, double *ws) {
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int ny0 = g->periodic[1] ? 1 : 2;
    const double ymid = 0.5 * g->ly;
    for (int k = 1; k <= nz; ++k)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                size_t c = GI(i, j, k);
                double xu[3] = {((double)(i + 1) - 0.5 - 1.5) * g->dx, yc[j], ((double)(k + 1) - 1.5) * g->dz};
                double xv[3] = {((double)(i + 1) - 1.5) * g->dx, yv[j], ((double)(k + 1) - 1.5) * g->dz};
                double xw[3] = {((double)(i + 1) - 1.5) * g->dx, yc[j], ((double)(k + 1) - 0.5 - 1.5) * g->dz};
                double yp = yc[j] - ymid;
                double base, h = 1.0;
                int fu = 1, fv = 1, fw = 1;
                if (use_geometry) {
                    h = geo->half_height - geo->A * cos(2.0 * PI * xu[0] / geo->lambda);
                    fu = !mdo_is_in_body(g, geo, xu);
                    fv = !mdo_is_in_body(g, geo, xv);
                    fw = !mdo_is_in_body(g, geo, xw);
                }
                double q = yp / h;
                base = 1.5 * (1.0 - q * q);
                us[c] = fu ? base + amp * (2.0 * mdo_uniform01(seed, 0, c) - 1.0) : 0.0;
                if (j >= ny0) vs[c] = fv ? amp * (2.0 * mdo_uniform01(seed, 1, c) - 1.0) : 0.0;
                ws[c] = fw ? amp * (2.0 * mdo_uniform01(seed, 2, c) - 1.0) : 0.0;
            }
}

/* ------------------------------------------------------------------ RK3 step */
/* main.f90:66-89 stage sequence. */
void mdo_rk3_step(const mdo_grid *g, const mdo_step_opts *o, double dt,
                  double *un, double *vn, double *wn, double *pn,
                  double *us, double *vs, double *ws,
                  double *oru, double *orv, double *orw,
                  const double *cu, const double *cv, const double *cw,
                  double *phi, double *rhs, mdo_sor_stats *st) {
    for (int s = 0; s < 3; ++s) {
        double dta = dt * mdo_rk_alpha[s], dtb = dt * mdo_rk_beta[s], dtg = dt * mdo_rk_gamma[s];
        mdo_momentum(g, un, vn, wn, pn, us, vs, ws, oru, orv, orw, cu, cv, cw, dta, dtb, dtg);
        if (o->flowrate) flowrate_shift_with(g, us, cu, dtg, o->ub_target, o->height, 1);
        if (o->solver == 0) {
            mdo_apply_bc_fft(g, un, vn, wn, us, vs, ws, pn);
            mdo_projection_fft(g, dtg, us, vs, ws, cu, cv, cw, un, vn, wn, pn);
            mdo_apply_bc_fft(g, un, vn, wn, us, vs, ws, pn);
            if (st) { st[s].iterations = 0; st[s].converged = 1; st[s].div_inf = mdo_div_inf(g, un, vn, wn); }
        } else if (o->solver == 1) {
            mdo_sola_redblack(g, o->niter, o->omega, dtg, cu, cv, cw, us, vs, ws, un, vn, wn, pn);
            mdo_apply_bc_redblack(g, un, vn, wn, pn);
            if (st) { st[s].iterations = o->niter; st[s].converged = 0; st[s].div_inf = mdo_div_inf(g, un, vn, wn); }
        } else {
            mdo_sor_stats tmp;
            mdo_projection_sor(g, o->omega, o->tol, o->max_iter, o->check_every, dtg, cu, cv, cw,
                               us, vs, ws, un, vn, wn, pn, phi, rhs, &tmp);
            mdo_apply_bc_redblack(g, un, vn, wn, pn);
            if (st) st[s] = tmp;
        }
    }
}
