// mdx_stencil.hip -- momentum predictor, ghost-cell kernels, reductions, flow-rate control and
// the initial field, for gfx950.  FP64 throughout; compiled with -ffp-contract=off so every
// expression rounds exactly as the reference's (step.f90, boundary.f90) does on the CPU.
#include <mutex>

#include "mdx_internal.h"

namespace mdx {

static constexpr int BX = 64;   // one wavefront along x: coalesced rows of the ghosted arrays
static constexpr int BY = 4;

// ------------------------------------------------------------------------------- momentum
// step.f90:27-207.  Updates u*, v*, w* and oldrhs of one interior cell.  Reads 10 arrays, writes 6
// (128 B/cell with IBM coefficients, 104 B without).  The stencil values come through an accessor
// L (L.u(a, b, c) = un(i+a, j+b, k+c), likewise v, w, p), so the direct and the LDS-tiled kernels
// evaluate the same expressions and round identically.
// Per-cell operands that are not stencil values: the previous stage's right-hand sides and the
// immersed-boundary coefficients of the three faces (read by the caller, so they can be prefetched).
struct MomOld {
    double ru, rv, rw, cu, cv, cw;
};

template <bool IBM>
__device__ inline MomOld mom_old(const double *__restrict__ oru, const double *__restrict__ orv,
                                 const double *__restrict__ orw, const double *__restrict__ cu,
                                 const double *__restrict__ cv, const double *__restrict__ cw, size_t o, size_t cc) {
    MomOld m{oru[o], orv[o], orw[o], 0.0, 0.0, 0.0};
    if (IBM) {
        m.cu = cu[cc];
        m.cv = cv[cc];
        m.cw = cw[cc];
    }
    return m;
}

// returns the u* it stores (0 where the cell's u is a boundary value it does not update)
template <bool STRETCH, bool IBM, class Acc>
__device__ inline double momentum_cell(const GridK &g, const Acc &L, int i, int j, int k, size_t c, size_t o,
                                     const MomOld &m, double *__restrict__ us, double *__restrict__ vs,
                                     double *__restrict__ ws, double *__restrict__ oru, double *__restrict__ orv,
                                     double *__restrict__ orw, double dta, double dtb, double dtg, double idx,
                                     double idy, double idz, double idx2, double idy2, double idz2, double ire) {
    const int nx0 = g.per[0] ? 1 : 2, ny0 = g.per[1] ? 1 : 2, nz0 = g.per[2] ? 1 : 2;
    const int jp = j + 1, jm = j - 1;
    const double *ihc = g.ihc, *ihv = g.ihv;
    double ustar = 0.0;
    if (i >= nx0) {
        const double u0 = L.u(0, 0, 0), uip = L.u(1, 0, 0), uim = L.u(-1, 0, 0);
        const double ujp = L.u(0, 1, 0), ujm = L.u(0, -1, 0);
        const double ukp = L.u(0, 0, 1), ukm = L.u(0, 0, -1);
        double uu_p = 0.25 * ((uip + u0) * (uip + u0));
        double uu_m = 0.25 * ((u0 + uim) * (u0 + uim));
        double uv_p = 0.25 * (ujp + u0) * (L.v(0, 1, 0) + L.v(-1, 1, 0));
        double uv_m = 0.25 * (u0 + ujm) * (L.v(0, 0, 0) + L.v(-1, 0, 0));
        double uw_p = 0.25 * (u0 + ukp) * (L.w(0, 0, 1) + L.w(-1, 0, 1));
        double uw_m = 0.25 * (u0 + ukm) * (L.w(0, 0, 0) + L.w(-1, 0, 0));
        double diff_ux = (uim - 2.0 * u0 + uip) * idx2;
        double diff_uy, cy;
        if (!STRETCH) {
            diff_uy = (ujm - 2.0 * u0 + ujp) * idy2;
            cy = (uv_p - uv_m) * idy;
        } else {
            diff_uy = ((ujp - u0) * ihv[jp] - (u0 - ujm) * ihv[j]) * ihc[j];
            cy = (uv_p - uv_m) * ihc[j];
        }
        double diff_uz = (ukm - 2.0 * u0 + ukp) * idz2;
        double dpx = (L.p(0, 0, 0) - L.p(-1, 0, 0)) * idx;
        double rhsu = -(uu_p - uu_m) * idx - cy - (uw_p - uw_m) * idz + g.fx +
                      ire * (diff_ux + diff_uy + diff_uz);
        double s = u0 + dta * rhsu + dtb * m.ru - dtg * dpx;
        if (IBM) s = s * (1.0 / (1.0 + dtg * m.cu));
        us[c] = s;
        oru[o] = rhsu;
        ustar = s;
    }
    if (j >= ny0) {
        const double v0 = L.v(0, 0, 0), vip = L.v(1, 0, 0), vim = L.v(-1, 0, 0);
        const double vjp = L.v(0, 1, 0), vjm = L.v(0, -1, 0);
        const double vkp = L.v(0, 0, 1), vkm = L.v(0, 0, -1);
        double vu_p = 0.25 * (v0 + vip) * (L.u(1, 0, 0) + L.u(1, -1, 0));
        double vu_m = 0.25 * (v0 + vim) * (L.u(0, 0, 0) + L.u(0, -1, 0));
        double vv_p = 0.25 * ((v0 + vjp) * (v0 + vjp));
        double vv_m = 0.25 * ((v0 + vjm) * (v0 + vjm));
        double vw_p = 0.25 * (vkp + v0) * (L.w(0, 0, 1) + L.w(0, -1, 1));
        double vw_m = 0.25 * (vkm + v0) * (L.w(0, 0, 0) + L.w(0, -1, 0));
        double diff_vx = (vim - 2.0 * v0 + vip) * idx2;
        double diff_vy, cy, dpy;
        if (!STRETCH) {
            diff_vy = (vjm - 2.0 * v0 + vjp) * idy2;
            cy = (vv_p - vv_m) * idy;
            dpy = (L.p(0, 0, 0) - L.p(0, -1, 0)) * idy;
        } else {
            diff_vy = ((vjp - v0) * ihc[j] - (v0 - vjm) * ihc[jm]) * ihv[j];
            cy = (vv_p - vv_m) * ihv[j];
            dpy = (L.p(0, 0, 0) - L.p(0, -1, 0)) * ihv[j];
        }
        double diff_vz = (vkm - 2.0 * v0 + vkp) * idz2;
        double rhsv = -(vu_p - vu_m) * idx - cy - (vw_p - vw_m) * idz + g.fy +
                      ire * (diff_vx + diff_vy + diff_vz);
        double s = v0 + dta * rhsv + dtb * m.rv - dtg * dpy;
        if (IBM) s = s * (1.0 / (1.0 + dtg * m.cv));
        vs[c] = s;
        orv[o] = rhsv;
    }
    if (k >= nz0) {
        const double w0 = L.w(0, 0, 0), wip = L.w(1, 0, 0), wim = L.w(-1, 0, 0);
        const double wjp = L.w(0, 1, 0), wjm = L.w(0, -1, 0);
        const double wkp = L.w(0, 0, 1), wkm = L.w(0, 0, -1);
        double wu_p = 0.25 * (w0 + wip) * (L.u(1, 0, 0) + L.u(1, 0, -1));
        double wu_m = 0.25 * (w0 + wim) * (L.u(0, 0, 0) + L.u(0, 0, -1));
        double ww_p = 0.25 * ((w0 + wkp) * (w0 + wkp));
        double ww_m = 0.25 * ((w0 + wkm) * (w0 + wkm));
        double wv_p = 0.25 * (w0 + wjp) * (L.v(0, 1, 0) + L.v(0, 1, -1));
        double wv_m = 0.25 * (w0 + wjm) * (L.v(0, 0, 0) + L.v(0, 0, -1));
        double diff_wx = (wim - 2.0 * w0 + wip) * idx2;
        double diff_wy, cy;
        if (!STRETCH) {
            diff_wy = (wjm - 2.0 * w0 + wjp) * idy2;
            cy = (wv_p - wv_m) * idy;
        } else {
            diff_wy = ((wjp - w0) * ihv[jp] - (w0 - wjm) * ihv[j]) * ihc[j];
            cy = (wv_p - wv_m) * ihc[j];
        }
        double diff_wz = (wkm - 2.0 * w0 + wkp) * idz2;
        double dpz = (L.p(0, 0, 0) - L.p(0, 0, -1)) * idz;
        double rhsw = -(wu_p - wu_m) * idx - cy - (ww_p - ww_m) * idz + g.fz +
                      ire * (diff_wx + diff_wy + diff_wz);
        double s = w0 + dta * rhsw + dtb * m.rw - dtg * dpz;
        if (IBM) s = s * (1.0 / (1.0 + dtg * m.cw));
        ws[c] = s;
        orw[o] = rhsw;
    }
    return ustar;
}

// global-memory operands of the stencil
struct MomGlobal {
    const double *un, *vn, *wn, *pn;
    Dims d;
    int i, j, k;
    __device__ inline double u(int a, int b, int c) const { return un[d.g(i + a, j + b, k + c)]; }
    __device__ inline double v(int a, int b, int c) const { return vn[d.g(i + a, j + b, k + c)]; }
    __device__ inline double w(int a, int b, int c) const { return wn[d.g(i + a, j + b, k + c)]; }
    __device__ inline double p(int a, int b, int c) const { return pn[d.g(i + a, j + b, k + c)]; }
};

template <bool STRETCH, bool IBM>
__global__ __launch_bounds__(BX * BY) void momentum_kernel(
    GridK g, const double *__restrict__ un, const double *__restrict__ vn,
    const double *__restrict__ wn, const double *__restrict__ pn, double *__restrict__ us,
    double *__restrict__ vs, double *__restrict__ ws, double *__restrict__ oru,
    double *__restrict__ orv, double *__restrict__ orw, const double *__restrict__ cu,
    const double *__restrict__ cv, const double *__restrict__ cw, int coef_kstride, double dta, double dtb,
    double dtg, double idx, double idy, double idz, double idx2, double idy2, double idz2,
    double ire, int /* planes per workgroup */, double * /* flow-rate partials */) {
    // (the last two arguments serve the tiled kernel: the launcher passes both kernels one list)
    const Dims d = g.d;
    int tx, ty, tz;
    if (!xcd_tile(d.nx, d.ny, d.nz, BX, BY, tx, ty, tz)) return;
    const int i = tx * BX + threadIdx.x + 1;
    const int j = ty * BY + threadIdx.y + 1;
    const int k = tz + 1;
    if (i > d.nx || j > d.ny) return;
    const MomGlobal L{un, vn, wn, pn, d, i, j, k};
    // coefficient index: z-invariant geometries keep one x-y plane (coef_kstride 0)
    const size_t o = d.in(i, j, k);
    const MomOld m = mom_old<IBM>(oru, orv, orw, cu, cv, cw, o, d.g(i, j, k * coef_kstride));
    momentum_cell<STRETCH, IBM>(g, L, i, j, k, d.g(i, j, k), o, m, us, vs, ws, oru, orv, orw, dta, dtb, dtg, idx, idy,
                                idz, idx2, idy2, idz2, ire);
}

// LDS-tiled momentum: a 64 x TY workgroup owns a 64 x TY column of cells and marches mkz planes in z
// (8 or 12, momentum_default_planes; 4, 16 and 32 measured slower).  The u, v, w planes k-1, k, k+1 (with a one-cell
// x-y halo) and the p plane k live in LDS as a rolling window, so each value is fetched from global
// memory about 1.3 (x-y halo) x (mkz+2)/mkz (z halo) times instead of once per stencil use (about
// 30 loads per cell in the direct kernel).
// TY = 8 rows per workgroup, or 4 on grids with fewer than 2048 such workgroups (measured: 54 against
// 61 us at 192 x 128 x 96, where 576 workgroups of 512 threads leave a thin last wave; 304 against
// 312 us at 256 x 192 x 256, where the smaller halo of 8 rows wins)
// FR: the workgroup also sums the flow-rate terms of its cells, u* h_j and mu_u h_j (the sums
// bulk_partial_kernel forms), into frpart[2 wg], frpart[2 wg + 1] for the tile wg = tx + nbx (ty +
// nby tz), so the flow-rate correction needs no pass of its own over u*.
static constexpr int MTX = 64, MTW = MTX + 2;

struct MomTile {
    const double *su, *sv, *sw, *sp;   // LDS planes
    int off[3];                        // plane offsets of k-1, k, k+1
    int t;                             // index of the cell in a plane
    double pkm;                        // p(i, j, k-1)
    __device__ inline double u(int a, int b, int c) const { return su[off[c + 1] + t + a + MTW * b]; }
    __device__ inline double v(int a, int b, int c) const { return sv[off[c + 1] + t + a + MTW * b]; }
    __device__ inline double w(int a, int b, int c) const { return sw[off[c + 1] + t + a + MTW * b]; }
    __device__ inline double p(int a, int b, int c) const { return c == 0 ? sp[t + a + MTW * b] : pkm; }
};

template <bool STRETCH, bool IBM, int TY, bool FR>
__global__ __launch_bounds__(MTX * TY) void momentum_tiled_kernel(
    GridK g, const double *__restrict__ un, const double *__restrict__ vn,
    const double *__restrict__ wn, const double *__restrict__ pn, double *__restrict__ us,
    double *__restrict__ vs, double *__restrict__ ws, double *__restrict__ oru,
    double *__restrict__ orv, double *__restrict__ orw, const double *__restrict__ cu,
    const double *__restrict__ cv, const double *__restrict__ cw, int coef_kstride, double dta, double dtb,
    double dtg, double idx, double idy, double idz, double idx2, double idy2, double idz2,
    double ire, int mkz, double *__restrict__ frpart) {
    constexpr int MPL = MTW * (TY + 2);   // window plane: the tile and its x-y halo
    __shared__ double su[3 * MPL], sv[3 * MPL], sw[3 * MPL], sp[MPL];
    const Dims d = g.d;
    const int nseg = (d.nz + mkz - 1) / mkz;
    int tx, ty, tz;
    if (!xcd_tile(d.nx, d.ny, nseg, MTX, TY, tx, ty, tz)) return;   // uniform per workgroup
    const int i0 = tx * MTX + 1, j0 = ty * TY + 1, k0 = tz * mkz + 1;
    const int k1 = k0 + mkz - 1 < d.nz ? k0 + mkz - 1 : d.nz;
    const int tid = threadIdx.x + MTX * threadIdx.y;
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool active = i <= d.nx && j <= d.ny;
    const int t = (threadIdx.x + 1) + MTW * (threadIdx.y + 1);
    // plane k of u, v, w into window slot `slot`; halo cells beyond the ghost layer are not read
    auto load_uvw = [&](int slot, int k) {
        for (int q = tid; q < MPL; q += MTX * TY) {
            const int gi = i0 - 1 + q % MTW, gj = j0 - 1 + q / MTW;
            if (gi <= d.nx + 1 && gj <= d.ny + 1) {
                const size_t x = d.g(gi, gj, k);
                su[slot * MPL + q] = un[x];
                sv[slot * MPL + q] = vn[x];
                sw[slot * MPL + q] = wn[x];
            }
        }
    };
    load_uvw(0, k0 - 1);
    load_uvw(1, k0);
    load_uvw(2, k0 + 1);
    constexpr int NT_ = MTX * TY, PER = (MPL + NT_ - 1) / NT_;   // window values per thread
    for (int q = tid; q < MPL; q += NT_) {
        const int gi = i0 - 1 + q % MTW, gj = j0 - 1 + q / MTW;
        if (gi <= d.nx + 1 && gj <= d.ny + 1) sp[q] = pn[d.g(gi, gj, k0)];
    }
    double pkm = active ? pn[d.g(i, j, k0 - 1)] : 0.0;
    // the per-cell operands of plane k0 (the next plane's are fetched while a plane is computed)
    MomOld m{};
    if (active) m = mom_old<IBM>(oru, orv, orw, cu, cv, cw, d.in(i, j, k0), d.g(i, j, k0 * coef_kstride));
    const double hj = (FR && active) ? (g.ihc ? 1.0 / g.ihc[j] : g.dy) : 0.0;
    double fa = 0.0, fb = 0.0;
    __syncthreads();
    for (int k = k0; k <= k1; ++k) {
        // the next planes (u, v, w at k+2, p at k+1, the per-cell operands at k+1) are fetched into
        // registers before plane k is computed, so their latency overlaps the arithmetic; the window
        // values enter LDS after it
        double ru[PER], rv[PER], rw[PER], rp[PER];
        MomOld mn{};
        if (active && k < k1) mn = mom_old<IBM>(oru, orv, orw, cu, cv, cw, d.in(i, j, k + 1), d.g(i, j, (k + 1) * coef_kstride));
        if (k < k1) {
#pragma unroll
            for (int c = 0; c < PER; ++c) {
                const int q = tid + c * NT_;
                const int gi = i0 - 1 + q % MTW, gj = j0 - 1 + q / MTW;
                if (q < MPL && gi <= d.nx + 1 && gj <= d.ny + 1) {
                    const size_t x2 = d.g(gi, gj, k + 2), x1 = d.g(gi, gj, k + 1);
                    ru[c] = un[x2];
                    rv[c] = vn[x2];
                    rw[c] = wn[x2];
                    rp[c] = pn[x1];
                }
            }
        }
        const int s0 = (k - k0) % 3;   // slot of plane k-1
        double pnext = 0.0;
        if (active) {
            const MomTile L{su, sv, sw, sp, {s0 * MPL, ((s0 + 1) % 3) * MPL, ((s0 + 2) % 3) * MPL}, t, pkm};
            const double uw = momentum_cell<STRETCH, IBM>(g, L, i, j, k, d.g(i, j, k), d.in(i, j, k), m, us, vs, ws,
                                                          oru, orv, orw, dta, dtb, dtg, idx, idy, idz, idx2, idy2,
                                                          idz2, ire);
            if (FR) {   // x is periodic (the launcher checks), so uw is the u* of every cell
                fa += uw * hj;
                fb += (IBM ? 1.0 / (1.0 + dtg * m.cu) : 1.0) * hj;
            }
            pnext = sp[t];
        }
        __syncthreads();
        if (k < k1) {   // plane k-1 leaves the window, plane k+2 enters; p plane k+1 replaces plane k
#pragma unroll
            for (int c = 0; c < PER; ++c) {
                const int q = tid + c * NT_;
                const int gi = i0 - 1 + q % MTW, gj = j0 - 1 + q / MTW;
                if (q < MPL && gi <= d.nx + 1 && gj <= d.ny + 1) {
                    su[s0 * MPL + q] = ru[c];
                    sv[s0 * MPL + q] = rv[c];
                    sw[s0 * MPL + q] = rw[c];
                    sp[q] = rp[c];
                }
            }
        }
        pkm = pnext;
        m = mn;
        __syncthreads();
    }
    if (FR) {
        // a fixed order: a butterfly within each wave (every lane ends with the same sums), then
        // the waves in index order; the window's LDS is free after the march's last barrier
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            fa += __shfl_xor(fa, o);
            fb += __shfl_xor(fb, o);
        }
        constexpr int NW = MTX * TY / 64;
        if ((tid & 63) == 0) {
            su[tid >> 6] = fa;
            su[NW + (tid >> 6)] = fb;
        }
        __syncthreads();
        if (tid == 0) {
            double a = 0.0, b = 0.0;
            for (int w = 0; w < NW; ++w) {
                a += su[w];
                b += su[NW + w];
            }
            const size_t nbx = (d.nx + MTX - 1) / MTX, nby = (d.ny + TY - 1) / TY;
            const size_t wg = (size_t)tx + nbx * ((size_t)ty + nby * (size_t)tz);
            frpart[2 * wg] = a;
            frpart[2 * wg + 1] = b;
        }
    }
}

// rows per tiled workgroup (TY above) when each marches kz planes
static int momentum_rows_for(const Dims &d, int kz) {
    const long wg8 = (long)((d.nx + MTX - 1) / MTX) * ((d.ny + 7) / 8) * ((d.nz + kz - 1) / kz);
    return wg8 < 2048 ? 4 : 8;
}

static int momentum_rows(const mdx_handle *h) { return momentum_rows_for(h->d, h->mom_kz); }

// The planes a workgroup marches unless configured: 12 where its workgroups fill their last round
// on the device better than those of 8 planes, else 8.  The device holds 4096 / TY of them at once
// (256 CUs, 4 waves per SIMD at up to 128 VGPRs).  At 192 x 128 x 96: 768 workgroups, one round
// (75% of the slots), against 1152 in two (56%): 220-224 against 231-240 us per step of momentum;
// at 256 x 192 x 256, 8 planes fill every round.  The flow-rate sums follow this tiling.
int momentum_default_planes(const Dims &d) {
    auto fill = [&](int kz) {
        const int ty = momentum_rows_for(d, kz);
        const long n = (long)((d.nx + MTX - 1) / MTX) * ((d.ny + ty - 1) / ty) * ((d.nz + kz - 1) / kz);
        const long slots = 4096 / ty, rounds = (n + slots - 1) / slots;
        return (double)n / (double)(rounds * slots);
    };
    return fill(12) > fill(8) ? 12 : 8;
}

int momentum_tiles(const mdx_handle *h) {
    const int ty = momentum_rows(h);
    return ((h->d.nx + MTX - 1) / MTX) * ((h->d.ny + ty - 1) / ty) * ((h->d.nz + h->mom_kz - 1) / h->mom_kz);
}

hipError_t launch_momentum(mdx_handle *h, double dta, double dtb, double dtg, double *frpart) {
    GridK g = grid_k(h);
    const double dx = h->dx, dy = h->dy, dz = h->dz;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const bool tiled = h->tiled_momentum;
    const int mkz = h->mom_kz;
    const int nseg = (h->d.nz + mkz - 1) / mkz;
    const int ty = momentum_rows(h);
    dim3 blk = tiled ? dim3(MTX, ty, 1) : dim3(BX, BY, 1);
    dim3 grd(tiled ? xcd_grid(h->d.nx, h->d.ny, nseg, MTX, ty) : xcd_grid(h->d.nx, h->d.ny, h->d.nz, BX, BY));
    // the partials come from the tiled kernel, over a periodic x (as constant flow rate requires)
    if (frpart && (!tiled || !h->cfg.periodic[0])) return hipErrorInvalidValue;
#define MDX_MOM_(S, I, FR)                                                                                     \
    hipLaunchKernelGGL((tiled ? (ty == 4 ? momentum_tiled_kernel<S, I, 4, FR> : momentum_tiled_kernel<S, I, 8, FR>) \
                              : momentum_kernel<S, I>),                                                        \
                       grd, blk, 0, h->stream,                                                               \
                       g, h->un, h->vn, h->wn,                                                   \
                       h->pn, h->us, h->vs, h->ws, h->oru, h->orv, h->orw, h->cu, h->cv, h->cw,  \
                       h->coef_plane ? 0 : 1, dta, dtb, dtg, 1.0 / dx, 1.0 / dy, 1.0 / dz, 1.0 / dx2, 1.0 / dy2,      \
                       1.0 / dz2, 1.0 / h->cfg.re, mkz, frpart)
#define MDX_MOM(S, I)                \
    do {                             \
        if (frpart)                  \
            MDX_MOM_(S, I, true);    \
        else                         \
            MDX_MOM_(S, I, false);   \
    } while (0)
    if (h->stretched) {
        if (h->ibm_on) MDX_MOM(true, true); else MDX_MOM(true, false);
    } else {
        if (h->ibm_on) MDX_MOM(false, true); else MDX_MOM(false, false);
    }
#undef MDX_MOM
#undef MDX_MOM_
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- apply_bc (FFT build)
// boundary.f90:23-170: y no-slip walls (edges included), then x- and z-periodic copies.  `fields`
// selects the arrays: BC_N (un, vn, wn, pn), BC_S (us, vs, ws) or both (the reference's apply_bc).
// Every ghost value is a function of its own array's interior, so refreshing only the arrays
// written since the last call leaves every ghost as the full call would.
__global__ void bc_fft_y(Dims d, double *un, double *vn, double *wn, double *us, double *vs, double *ws,
                         int fields) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int k = blockIdx.y;
    if (i > d.nx + 1 || k > d.nz + 1) return;
    const int ny = d.ny;
    if (fields & BC_N) {
        un[d.g(i, ny + 1, k)] = -un[d.g(i, ny, k)];
        vn[d.g(i, ny + 1, k)] = 0.0;
        wn[d.g(i, ny + 1, k)] = -wn[d.g(i, ny, k)];
        un[d.g(i, 0, k)] = -un[d.g(i, 1, k)];
        vn[d.g(i, 1, k)] = 0.0;
        wn[d.g(i, 0, k)] = -wn[d.g(i, 1, k)];
    }
    if (fields & BC_S) {
        us[d.g(i, ny + 1, k)] = -us[d.g(i, ny, k)];
        vs[d.g(i, ny + 1, k)] = 0.0;
        ws[d.g(i, ny + 1, k)] = -ws[d.g(i, ny, k)];
        us[d.g(i, 0, k)] = -us[d.g(i, 1, k)];
        vs[d.g(i, 1, k)] = 0.0;
        ws[d.g(i, 0, k)] = -ws[d.g(i, 1, k)];
    }
}

// x-periodic ghost columns: one thread per (row (j,k), field, side) copy.  The copies of a row are
// a stride of nx+2 apart, so each touches its own cache line; spreading them over threads keeps
// many of these latency-bound accesses in flight.
__global__ void bc_fft_x(Dims d, double *un, double *vn, double *wn, double *us, double *vs, double *ws,
                         double *pn, int fields) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long nrow = (long)(d.ny + 2) * (d.nz + 2);
    if (q >= 14 * nrow) return;
    const int task = (int)(q % 14), side = task & 1, field = task >> 1;
    const long row = q / 14;
    const int j = (int)(row % (d.ny + 2)), k = (int)(row / (d.ny + 2));
    double *f = field == 0 ? un : field == 1 ? wn : field == 2 ? us : field == 3 ? ws
              : field == 4 ? vn : field == 5 ? vs : pn;
    const bool s_field = field == 2 || field == 3 || field == 5;
    if (!(fields & (s_field ? BC_S : BC_N))) return;
    if (field >= 4 && j < 1) return;                     // v: rows 1..ny+1
    if (field == 6 && j > d.ny) return;                  // p: rows 1..ny
    const int nx = d.nx;
    if (side == 0) f[d.g(0, j, k)] = f[d.g(nx, j, k)];
    else f[d.g(nx + 1, j, k)] = f[d.g(1, j, k)];
}

__global__ void bc_fft_z(Dims d, double *un, double *vn, double *wn, double *us, double *vs, double *ws,
                         double *pn, int fields) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y;
    if (i > d.nx + 1 || j > d.ny + 1) return;
    const int nz = d.nz;
    if (fields & BC_N) {
        un[d.g(i, j, 0)] = un[d.g(i, j, nz)];
        wn[d.g(i, j, 0)] = wn[d.g(i, j, nz)];
        un[d.g(i, j, nz + 1)] = un[d.g(i, j, 1)];
        wn[d.g(i, j, nz + 1)] = wn[d.g(i, j, 1)];
    }
    if (fields & BC_S) {
        us[d.g(i, j, 0)] = us[d.g(i, j, nz)];
        ws[d.g(i, j, 0)] = ws[d.g(i, j, nz)];
        us[d.g(i, j, nz + 1)] = us[d.g(i, j, 1)];
        ws[d.g(i, j, nz + 1)] = ws[d.g(i, j, 1)];
    }
    if (j >= 1) {
        if (fields & BC_N) {
            vn[d.g(i, j, 0)] = vn[d.g(i, j, nz)];
            vn[d.g(i, j, nz + 1)] = vn[d.g(i, j, 1)];
        }
        if (fields & BC_S) {
            vs[d.g(i, j, 0)] = vs[d.g(i, j, nz)];
            vs[d.g(i, j, nz + 1)] = vs[d.g(i, j, 1)];
        }
    }
    if ((fields & BC_N) && j >= 1 && j <= d.ny) {
        pn[d.g(i, j, 0)] = pn[d.g(i, j, nz)];
        pn[d.g(i, j, nz + 1)] = pn[d.g(i, j, 1)];
    }
}

hipError_t launch_apply_bc_fft(mdx_handle *h, int fields) {
    const Dims d = h->d;
    hipLaunchKernelGGL(bc_fft_y, dim3((d.nx + 2 + 63) / 64, d.nz + 2), dim3(64), 0, h->stream, d,
                       h->un, h->vn, h->wn, h->us, h->vs, h->ws, fields);
    const long nx_tasks = 14L * (d.ny + 2) * (d.nz + 2);
    hipLaunchKernelGGL(bc_fft_x, dim3((unsigned)((nx_tasks + 255) / 256)), dim3(256), 0, h->stream, d,
                       h->un, h->vn, h->wn, h->us, h->vs, h->ws, h->pn, fields);
    hipLaunchKernelGGL(bc_fft_z, dim3((d.nx + 2 + 63) / 64, d.ny + 2), dim3(64), 0, h->stream, d,
                       h->un, h->vn, h->wn, h->us, h->vs, h->ws, h->pn, fields);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- apply_bc_redblack
// boundary.f90:173-368.  The reference runs the six faces in one parallel loop; the faces are
// applied here in its sequential order (y, then x, then z), which is also the only order in which
// the faces' shared reads (e.g. vn(nx,1,k) after the bottom wall) are well defined.
__global__ void bc_rb_y(GridK g, BcK b, double *un, double *vn, double *wn, double *pn) {
    const Dims d = g.d;
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int k = blockIdx.y + 1;
    if (i > d.nx || k > d.nz) return;
    const int ny = d.ny;
    const double dy = g.dy;
#define T(dd, s, q) b.type[bci(dd, s, q)]
#define V(dd, s, q) b.value[bci(dd, s, q)]
    if (g.per[1]) {
        un[d.g(i, 0, k)] = un[d.g(i, ny, k)];
        vn[d.g(i, 0, k)] = vn[d.g(i, ny, k)];
        wn[d.g(i, 0, k)] = wn[d.g(i, ny, k)];
        pn[d.g(i, 0, k)] = pn[d.g(i, ny, k)];
        un[d.g(i, ny + 1, k)] = un[d.g(i, 1, k)];
        vn[d.g(i, ny + 1, k)] = vn[d.g(i, 1, k)];
        wn[d.g(i, ny + 1, k)] = wn[d.g(i, 1, k)];
        pn[d.g(i, ny + 1, k)] = pn[d.g(i, 1, k)];
    } else {
        un[d.g(i, 0, k)] = T(2, 0, 1) == 0 ? 2 * V(2, 0, 1) - un[d.g(i, 1, k)] : un[d.g(i, 1, k)] - V(2, 0, 1) * dy;
        vn[d.g(i, 1, k)] = T(2, 0, 2) == 0 ? V(2, 0, 2) : vn[d.g(i, 2, k)] - V(2, 0, 2) * dy;
        wn[d.g(i, 0, k)] = T(2, 0, 3) == 0 ? 2 * V(2, 0, 3) - wn[d.g(i, 1, k)] : wn[d.g(i, 1, k)] - V(2, 0, 3) * dy;
        un[d.g(i, ny + 1, k)] = T(2, 1, 1) == 0 ? 2 * V(2, 1, 1) - un[d.g(i, ny, k)] : un[d.g(i, ny, k)] + V(2, 1, 1) * dy;
        vn[d.g(i, ny + 1, k)] = T(2, 1, 2) == 0 ? V(2, 1, 2) : vn[d.g(i, ny, k)] + V(2, 1, 2) * dy;
        wn[d.g(i, ny + 1, k)] = T(2, 1, 3) == 0 ? 2 * V(2, 1, 3) - wn[d.g(i, ny, k)] : wn[d.g(i, ny, k)] + V(2, 1, 3) * dy;
    }
}

__global__ void bc_rb_x(GridK g, BcK b, double *un, double *vn, double *wn, double *pn) {
    const Dims d = g.d;
    int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int k = blockIdx.y + 1;
    if (j > d.ny || k > d.nz) return;
    const int nx = d.nx;
    const double dx = g.dx;
    if (g.per[0]) {
        un[d.g(0, j, k)] = un[d.g(nx, j, k)];
        vn[d.g(0, j, k)] = vn[d.g(nx, j, k)];
        wn[d.g(0, j, k)] = wn[d.g(nx, j, k)];
        pn[d.g(0, j, k)] = pn[d.g(nx, j, k)];
        un[d.g(nx + 1, j, k)] = un[d.g(1, j, k)];
        vn[d.g(nx + 1, j, k)] = vn[d.g(1, j, k)];
        wn[d.g(nx + 1, j, k)] = wn[d.g(1, j, k)];
        pn[d.g(nx + 1, j, k)] = pn[d.g(1, j, k)];
    } else {
        un[d.g(1, j, k)] = T(1, 0, 1) == 0 ? V(1, 0, 1) : un[d.g(2, j, k)] - V(1, 0, 1) * dx;
        vn[d.g(0, j, k)] = T(1, 0, 2) == 0 ? 2 * V(1, 0, 2) - vn[d.g(1, j, k)] : vn[d.g(1, j, k)] - V(1, 0, 2) * dx;
        wn[d.g(0, j, k)] = T(1, 0, 3) == 0 ? 2 * V(1, 0, 3) - wn[d.g(1, j, k)] : wn[d.g(1, j, k)] - V(1, 0, 3) * dx;
        un[d.g(nx + 1, j, k)] = T(1, 1, 1) == 0 ? V(1, 1, 1) : un[d.g(nx, j, k)] + V(1, 1, 1) * dx;
        vn[d.g(nx + 1, j, k)] = T(1, 1, 2) == 0 ? 2 * V(1, 1, 2) - vn[d.g(nx, j, k)] : vn[d.g(nx, j, k)] + V(1, 1, 2) * dx;
        wn[d.g(nx + 1, j, k)] = T(1, 1, 3) == 0 ? 2 * V(1, 1, 3) - wn[d.g(nx, j, k)] : wn[d.g(nx, j, k)] + V(1, 1, 3) * dx;
    }
}

__global__ void bc_rb_z(GridK g, BcK b, double *un, double *vn, double *wn, double *pn) {
    const Dims d = g.d;
    int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    int j = blockIdx.y + 1;
    if (i > d.nx || j > d.ny) return;
    const int nz = d.nz;
    const double dz = g.dz;
    if (g.per[2]) {
        un[d.g(i, j, 0)] = un[d.g(i, j, nz)];
        vn[d.g(i, j, 0)] = vn[d.g(i, j, nz)];
        wn[d.g(i, j, 0)] = wn[d.g(i, j, nz)];
        pn[d.g(i, j, 0)] = pn[d.g(i, j, nz)];
        un[d.g(i, j, nz + 1)] = un[d.g(i, j, 1)];
        vn[d.g(i, j, nz + 1)] = vn[d.g(i, j, 1)];
        wn[d.g(i, j, nz + 1)] = wn[d.g(i, j, 1)];
        pn[d.g(i, j, nz + 1)] = pn[d.g(i, j, 1)];
    } else {
        un[d.g(i, j, 0)] = T(3, 0, 1) == 0 ? 2 * V(3, 0, 1) - un[d.g(i, j, 1)] : un[d.g(i, j, 1)] - V(3, 0, 1) * dz;
        vn[d.g(i, j, 0)] = T(3, 0, 2) == 0 ? 2 * V(3, 0, 2) - vn[d.g(i, j, 1)] : vn[d.g(i, j, 1)] - V(3, 0, 2) * dz;
        wn[d.g(i, j, 1)] = T(3, 0, 3) == 0 ? V(3, 0, 3) : wn[d.g(i, j, 2)] - V(3, 0, 3) * dz;
        un[d.g(i, j, nz + 1)] = T(3, 1, 1) == 0 ? 2 * V(3, 1, 1) - un[d.g(i, j, nz)] : un[d.g(i, j, nz)] + V(3, 1, 1) * dz;
        vn[d.g(i, j, nz + 1)] = T(3, 1, 2) == 0 ? 2 * V(3, 1, 2) - vn[d.g(i, j, nz)] : vn[d.g(i, j, nz)] + V(3, 1, 2) * dz;
        wn[d.g(i, j, nz + 1)] = T(3, 1, 3) == 0 ? V(3, 1, 3) : wn[d.g(i, j, nz)] + V(3, 1, 3) * dz;
    }
#undef T
#undef V
}

hipError_t launch_apply_bc_redblack(mdx_handle *h) {
    GridK g = grid_k(h);
    BcK b = bc_k(h);
    const Dims d = h->d;
    hipLaunchKernelGGL(bc_rb_y, dim3((d.nx + 63) / 64, d.nz), dim3(64), 0, h->stream, g, b, h->un,
                       h->vn, h->wn, h->pn);
    hipLaunchKernelGGL(bc_rb_x, dim3((d.ny + 63) / 64, d.nz), dim3(64), 0, h->stream, g, b, h->un,
                       h->vn, h->wn, h->pn);
    hipLaunchKernelGGL(bc_rb_z, dim3((d.nx + 63) / 64, d.ny), dim3(64), 0, h->stream, g, b, h->un,
                       h->vn, h->wn, h->pn);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- reductions
// Two-pass maximum: every workgroup writes its maximum to max_part (block_max_part), one
// workgroup folds them.  (A single-address atomicMax per workgroup serialises ~10^4 atomics.)
__device__ inline double nanmax(double a, double b) { return (a != a || a > b) ? a : b; }

__global__ __launch_bounds__(1024) void max_final_kernel(const double *__restrict__ part, unsigned n,
                                                         unsigned long long *out, double *host_slot,
                                                         unsigned long long seq) {
    __shared__ double w[16];
    double m = 0.0;
    // MF_BATCH partials per thread are loaded before any is folded (one memory latency per batch
    // instead of one per partial); the maximum does not depend on the order
    constexpr unsigned MF_BATCH = 16;
    for (unsigned q0 = threadIdx.x; q0 < n; q0 += 1024 * MF_BATCH) {
        double v[MF_BATCH];
#pragma unroll
        for (unsigned u = 0; u < MF_BATCH; ++u) {
            const unsigned q = q0 + u * 1024;
            v[u] = q < n ? part[q] : 0.0;
        }
#pragma unroll
        for (unsigned u = 0; u < MF_BATCH; ++u) m = nanmax(m, v[u]);
    }
    for (int off = 32; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off, 64));
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 16; ++q) m = nanmax(m, w[q]);
        m = nanmax(w[0], m);
        out[0] = (unsigned long long)__double_as_longlong(m);
        if (host_slot) {   // the result and the error word straight into pinned host memory,
            host_slot[0] = m;  // then the sequence number the host polls for
            host_slot[1] = __longlong_as_double((long long)out[1]);
            __threadfence_system();
            reinterpret_cast<volatile unsigned long long *>(host_slot)[2] = seq;
        }
    }
}

hipError_t launch_max_final(mdx_handle *h, unsigned nparts, double *host_slot, unsigned long long seq) {
    if (nparts > MAX_PARTS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(1024), 0, h->stream, h->max_part, nparts, h->red_max,
                       host_slot, seq);
    return hipGetLastError();
}

// kept for callers that still clear the result word first (a no-op for the two-pass reductions)
hipError_t launch_reset_max(mdx_handle *h) {
    (void)h;
    return hipSuccess;
}

// ||div u||_inf over interior cells (periodic wrap as in the red-black solver)
template <bool STRETCH>
__global__ __launch_bounds__(BX * BY) void div_max_kernel(GridK g, const double *__restrict__ un,
                                                          const double *__restrict__ vn,
                                                          const double *__restrict__ wn, double idx,
                                                          double idy, double idz, double *part) {
    const Dims d = g.d;
    int tx = 0, ty = 0, tz = 0;
    const bool ok = xcd_tile(d.nx, d.ny, d.nz, BX, BY, tx, ty, tz);   // padding groups still report 0
    const int i = tx * BX + threadIdx.x + 1;
    const int j = ty * BY + threadIdx.y + 1;
    const int k = tz + 1;
    double m = 0.0;
    if (ok && i <= d.nx && j <= d.ny) {
        int ip = (g.per[0] && i == d.nx) ? 1 : i + 1;
        int kp = (g.per[2] && k == d.nz) ? 1 : k + 1;
        int jp = (g.per[1] && j == d.ny) ? 1 : j + 1;
        size_t c = d.g(i, j, k);
        double dvy = STRETCH ? (vn[d.g(i, jp, k)] - vn[c]) * g.ihc[j] : (vn[d.g(i, jp, k)] - vn[c]) * idy;
        m = fabs((un[d.g(ip, j, k)] - un[c]) * idx + dvy + (wn[d.g(i, j, kp)] - wn[c]) * idz);
    }
    block_max_part(m, part);
}

hipError_t launch_div_max(mdx_handle *h, const double *u, const double *v, const double *w) {
    GridK g = grid_k(h);
    const unsigned nb = xcd_grid(h->d.nx, h->d.ny, h->d.nz, BX, BY);
    dim3 blk(BX, BY, 1);
    if (h->stretched)
        hipLaunchKernelGGL(div_max_kernel<true>, dim3(nb), blk, 0, h->stream, g, u, v, w, 1.0 / h->dx,
                           1.0 / h->dy, 1.0 / h->dz, h->max_part);
    else
        hipLaunchKernelGGL(div_max_kernel<false>, dim3(nb), blk, 0, h->stream, g, u, v, w, 1.0 / h->dx,
                           1.0 / h->dy, 1.0 / h->dz, h->max_part);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_max_final(h, nb);
}

// step.f90:210-258: max over every ghosted entry of |u|/dx, |v|/dy, |w|/dz
__global__ void cfl_kernel(GridK g, const double *__restrict__ un, const double *__restrict__ vn,
                           const double *__restrict__ wn, size_t n, double *part) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    for (; q < n; q += (size_t)gridDim.x * blockDim.x) {
        size_t j = (q / (size_t)g.d.sx) % (size_t)g.d.sy;
        double a = fabs(un[q] / g.dx);
        double b = g.ihv ? fabs(vn[q] * g.ihv[j]) : fabs(vn[q] / g.dy);
        double c = fabs(wn[q] / g.dz);
        m = nanmax(m, nanmax(a, nanmax(b, c)));
    }
    block_max_part(m, part);
}

hipError_t launch_cfl_max(mdx_handle *h) {
    GridK g = grid_k(h);
    hipLaunchKernelGGL(cfl_kernel, dim3(1024), dim3(256), 0, h->stream, g, h->un, h->vn, h->wn, h->ng,
                       h->max_part);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_max_final(h, 1024);
}

// io.f90:184-222 center_vel: the cell-centred velocity of each interior cell, the mean of its two
// faces in each direction (the ghost layer supplies the far face of the last cell).  Output in the
// interior layout uc(1:nx,1:ny,1:nz).  56 B per cell: six face reads (three of them the
// neighbour's, shared along the row or with the next row/plane) and three writes.
__global__ __launch_bounds__(BX * BY) void center_vel_kernel(Dims d, const double *__restrict__ un,
                                                             const double *__restrict__ vn,
                                                             const double *__restrict__ wn, double *__restrict__ uc,
                                                             double *__restrict__ vc, double *__restrict__ wc) {
    int tx = 0, ty = 0, tz = 0;
    if (!xcd_tile(d.nx, d.ny, d.nz, BX, BY, tx, ty, tz)) return;
    const int i = tx * BX + threadIdx.x + 1;
    const int j = ty * BY + threadIdx.y + 1;
    const int k = tz + 1;
    if (i > d.nx || j > d.ny) return;
    const size_t c = d.g(i, j, k), o = d.in(i, j, k);
    uc[o] = 0.5 * (un[c] + un[c + 1]);
    vc[o] = 0.5 * (vn[c] + vn[d.g(i, j + 1, k)]);
    wc[o] = 0.5 * (wn[c] + wn[d.g(i, j, k + 1)]);
}

hipError_t launch_center_vel(mdx_handle *h, double *uc, double *vc, double *wc) {
    const unsigned nb = xcd_grid(h->d.nx, h->d.ny, h->d.nz, BX, BY);
    hipLaunchKernelGGL(center_vel_kernel, dim3(nb), dim3(BX, BY, 1), 0, h->stream, h->d, h->un, h->vn, h->wn, uc,
                       vc, wc);
    return hipGetLastError();
}

// Deterministic two-pass sums over interior u points: S1 = sum u*h_j, S2 = sum mu_u*h_j.
static constexpr int SUM_BLOCKS = 512;
static constexpr int SUM_THREADS = 256;

__global__ __launch_bounds__(SUM_THREADS) void bulk_partial_kernel(GridK g, const double *__restrict__ u,
                                                                   const double *__restrict__ cu, int cplane,
                                                                   double dtg, double *part) {
    const Dims d = g.d;
    __shared__ double s1[SUM_THREADS], s2[SUM_THREADS];
    double a = 0.0, b = 0.0;
    // workgroup b sums the x rows r = b, b + SUM_BLOCKS, ... (r = (j-1) + ny (k-1)), threads along x
    const int nrows = d.ny * d.nz;
    // mu_u of a cell from its coefficient; a z-invariant coefficient array is read at k = 1 (the
    // same value, with 1/nz of the traffic)
    auto term_b = [&](size_t c, double hj) { return (cu ? 1.0 / (1.0 + dtg * cu[c]) : 1.0) * hj; };
    auto cidx = [&](int i, int j, int k) { return d.g(0, j, cplane ? 1 : k) + (size_t)i; };
    if (d.nx <= SUM_THREADS) {
        // one value per row and thread: SUM_ROWS rows' loads are issued together, then added in
        // row order (the same sums as the loop below)
        constexpr int SUM_ROWS = 8;
        const int i = threadIdx.x + 1;
        if (i <= d.nx) {
            int r = blockIdx.x;
            for (; r + (SUM_ROWS - 1) * SUM_BLOCKS < nrows; r += SUM_ROWS * SUM_BLOCKS) {
                double uv[SUM_ROWS], bv[SUM_ROWS];
#pragma unroll
                for (int q = 0; q < SUM_ROWS; ++q) {
                    const int rr = r + q * SUM_BLOCKS, j = rr % d.ny + 1, k = rr / d.ny + 1;
                    const double hj = g.ihc ? 1.0 / g.ihc[j] : g.dy;
                    const size_t c = d.g(0, j, k) + (size_t)i;
                    uv[q] = u[c] * hj;
                    bv[q] = term_b(cidx(i, j, k), hj);
                }
#pragma unroll
                for (int q = 0; q < SUM_ROWS; ++q) {
                    a += uv[q];
                    b += bv[q];
                }
            }
            for (; r < nrows; r += SUM_BLOCKS) {
                const int j = r % d.ny + 1, k = r / d.ny + 1;
                const double hj = g.ihc ? 1.0 / g.ihc[j] : g.dy;
                const size_t c = d.g(0, j, k) + (size_t)i;
                a += u[c] * hj;
                b += term_b(cidx(i, j, k), hj);
            }
        }
    } else {
        for (int r = blockIdx.x; r < nrows; r += SUM_BLOCKS) {
            const int j = r % d.ny + 1, k = r / d.ny + 1;
            const double hj = g.ihc ? 1.0 / g.ihc[j] : g.dy;
            const size_t row = d.g(0, j, k);
            for (int i = threadIdx.x + 1; i <= d.nx; i += SUM_THREADS) {
                const size_t c = row + (size_t)i;
                a += u[c] * hj;
                b += term_b(cidx(i, j, k), hj);
            }
        }
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = SUM_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

// out[0] = bulk velocity, out[1] = mean mu_u, out[2] = du = (target - ub)/mean_mu
__global__ __launch_bounds__(SUM_THREADS) void bulk_final_kernel(const double *part, double norm,
                                                                 double target, double *out) {
    __shared__ double s1[SUM_THREADS], s2[SUM_THREADS];
    double a = 0.0, b = 0.0;
    for (int q = threadIdx.x; q < SUM_BLOCKS; q += SUM_THREADS) {
        a += part[2 * q];
        b += part[2 * q + 1];
    }
    s1[threadIdx.x] = a;
    s2[threadIdx.x] = b;
    __syncthreads();
    for (int w = SUM_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s1[threadIdx.x] += s1[threadIdx.x + w];
            s2[threadIdx.x] += s2[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double ub = s1[0] / norm;
        double sm = s2[0] / norm;
        out[0] = ub;
        out[1] = sm;
        out[2] = (target - ub) / sm;
    }
}

hipError_t launch_bulk_sum(mdx_handle *h, const double *u, double dtg, double *out_dev) {
    GridK g = grid_k(h);
    hipLaunchKernelGGL(bulk_partial_kernel, dim3(SUM_BLOCKS), dim3(SUM_THREADS), 0, h->stream, g, u,
                       h->ibm_on ? h->cu : nullptr, h->coef_plane ? 1 : 0, dtg, h->red_part);
    double norm = (double)h->d.nx * (double)h->d.nz * h->cfg.flow_height;
    hipLaunchKernelGGL(bulk_final_kernel, dim3(1), dim3(SUM_THREADS), 0, h->stream, h->red_part, norm,
                       h->cfg.ub_target, out_dev);
    return hipGetLastError();
}

// Constant flow rate, second half: every workgroup folds the nparts partial sums (those of
// bulk_partial_kernel, folded as bulk_final_kernel does, or those of the momentum tiles) in the
// same order (so every workgroup holds the same du; workgroup 0 also stores ub, mean mu_u and du
// to out), builds its share of the stage's face-coefficient tables when mb.n > 0 (grid-stride
// over all workgroups, padding included), then u* += du * mu_u on its cells: a 64 x 4 tile of
// columns over FS_KZ planes, mu_u formed once per column where the coefficients are z-invariant
// (cplane), else per cell: u* + du * (1 / (1 + dt_gamma c_u)) as the oracle's shift forms it.
static constexpr int FS_KZ = 16;
__global__ __launch_bounds__(SUM_THREADS) void flow_shift_fold_kernel(GridK g, double *us, const double *cu,
                                                                      int cplane, double dtg,
                                                                      const double *__restrict__ part, int nparts,
                                                                      double norm, double target, double *out,
                                                                      MuBuild mb) {
    static_assert(SUM_THREADS == 256, "the fold's lanes are the 64 x 4 threads");
    __shared__ double s1[SUM_THREADS], s2[SUM_THREADS];
    const int t = threadIdx.x + 64 * threadIdx.y;
    double a = 0.0, b = 0.0;
    for (int q = t; q < nparts; q += SUM_THREADS) {
        a += part[2 * q];
        b += part[2 * q + 1];
    }
    s1[t] = a;
    s2[t] = b;
    __syncthreads();
    for (int w = SUM_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) {
            s1[t] += s1[t + w];
            s2[t] += s2[t + w];
        }
        __syncthreads();
    }
    const double ub = s1[0] / norm;
    const double sm = s2[0] / norm;
    const double du = (target - ub) / sm;
    if (blockIdx.x == 0 && t == 0) {
        out[0] = ub;
        out[1] = sm;
        out[2] = du;
    }
    for (size_t q = (size_t)blockIdx.x * SUM_THREADS + t; q < mb.n; q += (size_t)gridDim.x * SUM_THREADS)
        mu_entry(mb, q);
    const Dims d = g.d;
    int tx_, ty_, tz_;
    if (!xcd_tile(d.nx, d.ny, (d.nz + FS_KZ - 1) / FS_KZ, 64, 4, tx_, ty_, tz_)) return;
    const int i = tx_ * 64 + threadIdx.x + 1, j = ty_ * 4 + threadIdx.y + 1, k0 = tz_ * FS_KZ + 1;
    if (i > d.nx || j > d.ny) return;
    const double mc = (cu && cplane) ? 1.0 / (1.0 + dtg * cu[d.g(i, j, 1)]) : 1.0;
    double uo[FS_KZ], mk[FS_KZ];
#pragma unroll
    for (int u = 0; u < FS_KZ; ++u) {
        const int k = k0 + u;
        uo[u] = 0.0;
        mk[u] = mc;
        if (k <= d.nz) {
            const size_t c = d.g(i, j, k);
            uo[u] = us[c];
            if (cu && !cplane) mk[u] = 1.0 / (1.0 + dtg * cu[c]);
        }
    }
#pragma unroll
    for (int u = 0; u < FS_KZ; ++u)
        if (k0 + u <= d.nz) us[d.g(i, j, k0 + u)] = uo[u] + du * mk[u];
}

hipError_t launch_flow_fold(mdx_handle *h, double dtg, const double *part, int nparts, double *out_dev,
                            const MuBuild &mb) {
    GridK g = grid_k(h);
    const double *cu = h->ibm_on ? h->cu : nullptr;
    const int cplane = h->coef_plane ? 1 : 0;
    const double norm = (double)h->d.nx * (double)h->d.nz * h->cfg.flow_height;
    const unsigned grd = xcd_grid(h->d.nx, h->d.ny, (h->d.nz + FS_KZ - 1) / FS_KZ, 64, 4);
    hipLaunchKernelGGL(flow_shift_fold_kernel, dim3(grd), dim3(64, 4), 0, h->stream, g, h->us, cu, cplane, dtg, part,
                       nparts, norm, h->cfg.ub_target, out_dev, mb);
    return hipGetLastError();
}

// the flow-rate correction of u* in two launches: the partial sums, then the fold and the shift
hipError_t launch_flowrate(mdx_handle *h, double dtg, double *out_dev) {
    GridK g = grid_k(h);
    hipLaunchKernelGGL(bulk_partial_kernel, dim3(SUM_BLOCKS), dim3(SUM_THREADS), 0, h->stream, g, h->us,
                       h->ibm_on ? h->cu : nullptr, h->coef_plane ? 1 : 0, dtg, h->red_part);
    const hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_flow_fold(h, dtg, h->red_part, SUM_BLOCKS, out_dev);
}

// ------------------------------------------------------------------------------- initial field
__device__ inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ inline double uniform01(uint32_t seed, uint32_t comp, uint64_t index) {
    uint64_t key = ((uint64_t)seed << 34) + ((uint64_t)comp << 32) + index;
    return (double)(splitmix64(key) >> 11) * (1.0 / 9007199254740992.0);
}

// Poiseuille u = 1.5(1-(y/h)^2) + amp(2U-1), v,w = amp(2U-1) on interior faces; with a geometry
// (use_geo), samples inside the body (coefficient == solid) are zero and h = hx_u[i].
__global__ __launch_bounds__(BX * BY) void initial_field_kernel(
    GridK g, const double *__restrict__ yc, const double *__restrict__ hx_u, double ymid, uint32_t seed,
    double amp, int use_geo, const double *cu, const double *cv, const double *cw, double solid,
    double *us, double *vs, double *ws) {
    const Dims d = g.d;
    const int i = blockIdx.x * BX + threadIdx.x + 1;
    const int j = blockIdx.y * BY + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > d.nx || j > d.ny) return;
    const int ny0 = g.per[1] ? 1 : 2;
    size_t c = d.g(i, j, k);
    double h = use_geo ? hx_u[i] : 1.0;
    bool fu = !use_geo || cu[c] != solid;
    bool fv = !use_geo || cv[c] != solid;
    bool fw = !use_geo || cw[c] != solid;
    double q = (yc[j] - ymid) / h;
    double base = 1.5 * (1.0 - q * q);
    us[c] = fu ? base + amp * (2.0 * uniform01(seed, 0, c) - 1.0) : 0.0;
    if (j >= ny0) vs[c] = fv ? amp * (2.0 * uniform01(seed, 1, c) - 1.0) : 0.0;
    ws[c] = fw ? amp * (2.0 * uniform01(seed, 2, c) - 1.0) : 0.0;
}

hipError_t launch_initial_field(mdx_handle *h, uint32_t seed, double amp, int use_geo,
                                const double *hx_u_dev, double ymid) {
    GridK g = grid_k(h);
    dim3 blk(BX, BY, 1), grd((h->d.nx + BX - 1) / BX, (h->d.ny + BY - 1) / BY, h->d.nz);
    hipLaunchKernelGGL(initial_field_kernel, grd, blk, 0, h->stream, g, h->d_yc, hx_u_dev, ymid, seed, amp,
                       use_geo, h->cu, h->cv, h->cw, 1.0e30 / h->cfg.re, h->us, h->vs, h->ws);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- no-allocation norm
// mdx_div_norm_noalloc keeps the per-workgroup partial maxima of div_max_kernel in this file-scope
// device array instead of a handle's buffer: one copy per GPU, held until the process ends.  Host
// code never passes the array's name to a kernel (in host code it is only a stand-in symbol); it
// passes the device address hipGetSymbolAddress returned for the current device.
static __device__ __align__(256) double scratch[MAX_PARTS];

static constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[NOALLOC_MAX_DEVICES] = {};

// the device address of `scratch` on device dev (looked up once per device, with dev current)
static hipError_t scratch_address(int dev, double **out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_dev[dev];
    return hipSuccess;
}

}  // namespace mdx

#define NOALLOC_CHECK(cond, code, msg) \
    do {                               \
        if (!(cond)) {                 \
            mdx_set_error(msg);        \
            return code;               \
        }                              \
    } while (0)

extern "C" int mdx_div_norm_noalloc(int32_t nx, int32_t ny, int32_t nz, double lx, double ly, double lz,
                                    const int32_t *periodic, const double *un, const double *vn, const double *wn,
                                    double *div_inf, void *stream) {
    using namespace mdx;
    NOALLOC_CHECK(nx > 0 && ny > 0 && nz > 0 && lx > 0.0 && ly > 0.0 && lz > 0.0, MDX_ERR_ARG,
              "mdx_div_norm_noalloc: grid sizes and lengths must be positive");
    NOALLOC_CHECK(periodic && un && vn && wn && div_inf, MDX_ERR_ARG, "mdx_div_norm_noalloc: null argument");
    const unsigned nb = xcd_grid(nx, ny, nz, BX, BY);
    NOALLOC_CHECK((double)nx * ny * nz < 2.0e9 && nb <= MAX_PARTS, MDX_ERR_ARG,
              "mdx_div_norm_noalloc: grid too large for the static partial-max buffer (" + std::to_string(MAX_PARTS) +
                  " workgroups of 64 x 4 cells)");
    int dev = 0;
    MDX_HIP(hipGetDevice(&dev));
    NOALLOC_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, MDX_ERR_UNSUPPORTED, "mdx_div_norm_noalloc: device index too large");
    double *part = nullptr;
    MDX_HIP(scratch_address(dev, &part));
    GridK g{};
    g.d = Dims{nx, ny, nz, nx + 2, ny + 2};
    g.dx = lx / (double)nx;   // the spacings mdx_create derives (init.f90:48-55)
    g.dy = ly / (double)ny;
    g.dz = lz / (double)nz;
    for (int q = 0; q < 3; ++q) g.per[q] = periodic[q] ? 1 : 0;
    g.ihc = g.ihv = nullptr;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(div_max_kernel<false>, dim3(nb), dim3(BX, BY, 1), 0, st, g, un, vn, wn, 1.0 / g.dx, 1.0 / g.dy,
                       1.0 / g.dz, part);
    MDX_HIP(hipGetLastError());
    // the fold writes the maximum's bits as one 64-bit word: the double itself
    hipLaunchKernelGGL(max_final_kernel, dim3(1), dim3(1024), 0, st, part, nb,
                       reinterpret_cast<unsigned long long *>(div_inf), nullptr, 0ull);
    MDX_HIP(hipGetLastError());
    return MDX_OK;
}
