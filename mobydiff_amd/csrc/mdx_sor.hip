// mdx_sor.hip -- red-black pressure solvers for gfx950.
//
//  * sola_sweep: the reference's SOLA pressure-velocity sweep (pressure_redblack.f90:86-250),
//    statement for statement, updating u, v, w, p in place; used for the fixed-nIter mode.
//  * sor_*: the same iteration in pressure-correction form.  Each sweep updates the accumulated
//    correction Phi only; u = u* + mu (Phi_west - Phi_east)/h is formed once at the end.  The
//    residual of a cell is div(u*) - L_mu(Phi), the relaxation denominator is the reference's
//    (wall faces included), so the iterates equal the SOLA iterates.  FP64 for the plain solve;
//    FP32 sweeps on the defect equation inside an FP64 defect-correction loop for the mixed solve.
#include "mdx_sor_core.h"

namespace mdx {

static constexpr int SX = 32;   // active cells of one colour along x per block row
static constexpr int SY = 8;

// ------------------------------------------------------------------------------- SOLA (reference)
__global__ __launch_bounds__(SX * SY) void sola_sweep_kernel(
    GridK g, BcK b, int parity, double sor, double dtg, double idx, double idy, double idz,
    double idx2, double idy2, double idz2, double idt, const double *__restrict__ cu,
    const double *__restrict__ cv, const double *__restrict__ cw, double *un, double *vn, double *wn,
    double *pn) {
    const Dims d = g.d;
    int tx_, ty_, tz_;
    if (!xcd_tile(d.nx / 2, d.ny, d.nz, SX, SY, tx_, ty_, tz_)) return;
    const int ni = tx_ * SX + threadIdx.x + 1;
    const int j = ty_ * SY + threadIdx.y + 1;
    const int k = tz_ + 1;
    if (ni > d.nx / 2 || j > d.ny) return;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const bool per1 = g.per[0], per2 = g.per[1], per3 = g.per[2];
    const double dx = g.dx, dy = g.dy, dz = g.dz;
    int jp = j + 1;
    if (per2 && jp > ny) jp = 1;
    int kp = k + 1;
    if (per3 && kp > nz) kp = 1;
    const int iLo = 1 + ((j + k + parity) % 2);
    const int i = iLo + 2 * (ni - 1);
    int ip = i + 1;
    if (per1 && ip > nx) ip = 1;
    const size_t c = d.g(i, j, k), cip = d.g(ip, j, k), cjp = d.g(i, jp, k), ckp = d.g(i, j, kp);
#define MU(a, q) (1.0 / (1.0 + dtg * ((a) ? (a)[q] : 0.0)))
#define T(dd, s, q) b.type[bci(dd, s, q)]
#define V(dd, s, q) b.value[bci(dd, s, q)]
    const double mu_u_i = MU(cu, c), mu_u_ip = MU(cu, cip);
    const double mu_v_j = MU(cv, c), mu_v_jp = MU(cv, cjp);
    const double mu_w_k = MU(cw, c), mu_w_kp = MU(cw, ckp);
    const double denom = (mu_u_i + mu_u_ip) * idx2 + (mu_w_k + mu_w_kp) * idz2 + (mu_v_j + mu_v_jp) * idy2;
    if (!per1) {
        if (i < 2) un[c] = T(1, 0, 1) == 0 ? V(1, 0, 1) : un[cip] - V(1, 0, 1) * dx;
        if (i >= nx) un[cip] = T(1, 1, 1) == 0 ? V(1, 1, 1) : un[d.g(nx, j, k)] + V(1, 1, 1) * dx;
    }
    if (!per2) {
        if (j < 2) vn[c] = T(2, 0, 2) == 0 ? V(2, 0, 2) : vn[cjp] - V(2, 0, 2) * dy;
        if (j >= ny) vn[cjp] = T(2, 1, 2) == 0 ? V(2, 1, 2) : vn[c] + V(2, 1, 2) * dy;
    }
    if (!per3) {
        if (k < 2) wn[c] = T(3, 0, 3) == 0 ? V(3, 0, 3) : wn[ckp] - V(3, 0, 3) * dz;
        if (k >= nz) wn[ckp] = T(3, 1, 3) == 0 ? V(3, 1, 3) : wn[c] + V(3, 1, 3) * dz;
    }
    const double phi = ((un[cip] - un[c]) * idx + (vn[cjp] - vn[c]) * idy + (wn[ckp] - wn[c]) * idz) *
                       (-sor / denom);
    pn[c] = pn[c] + phi * idt;
    un[c] = un[c] - phi * idx * mu_u_i;
    if (!per1 && i == 1) {
        if (T(1, 0, 0) == 0) un[c] = un[c] + V(1, 0, 0) * idx * mu_u_i;
        else un[c] = un[c] + (phi * idx - V(1, 0, 0)) * mu_u_i;
    }
    un[cip] = un[cip] + phi * idx * mu_u_ip;
    if (!per1 && i == nx) {
        if (T(1, 1, 0) == 0) un[cip] = un[cip] - V(1, 1, 0) * idx * mu_u_ip;
        else un[cip] = un[cip] + (V(1, 1, 0) - phi * idx) * mu_u_ip;
    }
    vn[c] = vn[c] - phi * idy * mu_v_j;
    if (!per2 && j == 1) {
        if (T(2, 0, 0) == 0) vn[c] = vn[c] + V(2, 0, 0) * idy * mu_v_j;
        else vn[c] = vn[c] + (phi * idy - V(2, 0, 0)) * mu_v_j;
    }
    vn[cjp] = vn[cjp] + phi * idy * mu_v_jp;
    if (!per2 && j == ny) {
        if (T(2, 1, 0) == 0) vn[cjp] = vn[cjp] - V(2, 1, 0) * idy * mu_v_jp;
        else vn[cjp] = vn[cjp] + (V(2, 1, 0) - phi * idy) * mu_v_jp;
    }
    wn[c] = wn[c] - phi * idz * mu_w_k;
    if (!per3 && k == 1) {
        if (T(3, 0, 0) == 0) wn[c] = wn[c] + V(3, 0, 0) * idz * mu_w_k;
        else wn[c] = wn[c] + (phi * idz - V(3, 0, 0)) * mu_w_k;
    }
    wn[ckp] = wn[ckp] + phi * idz * mu_w_kp;
    if (!per3 && k == nz) {
        // pressure_redblack.f90:229-230 carries no mu factor on this Dirichlet branch
        if (T(3, 1, 0) == 0) wn[ckp] = wn[ckp] - V(3, 1, 0) * idz;
        else wn[ckp] = wn[ckp] + (V(3, 1, 0) - phi * idz) * mu_w_kp;
    }
#undef MU
#undef T
#undef V
}

hipError_t launch_sola_sweep(mdx_handle *h, int parity, double sor, double dtg) {
    GridK g = grid_k(h);
    BcK b = bc_k(h);
    const double idx = 1.0 / h->dx, idy = 1.0 / h->dy, idz = 1.0 / h->dz;
    dim3 blk(SX, SY), grd(xcd_grid(h->d.nx / 2, h->d.ny, h->d.nz, SX, SY));
    const double *cu = h->ibm_on ? h->cu : nullptr, *cv = h->ibm_on ? h->cv : nullptr,
                 *cw = h->ibm_on ? h->cw : nullptr;
    hipLaunchKernelGGL(sola_sweep_kernel, grd, blk, 0, h->stream, g, b, parity, sor, dtg, idx, idy, idz,
                       idx * idx, idy * idy, idz * idz, 1.0 / dtg, cu, cv, cw, h->un, h->vn, h->wn, h->pn);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- face coefficients
__global__ void build_mu_kernel(MuBuild b) {
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < b.n) mu_entry(b, q);
}

MuBuild mu_build(const mdx_handle *h, double dtg) {
    MuBuild b;
    if (!h->ibm_on) return b;
    b.n = h->coef_plane ? (size_t)h->d.sx * h->d.sy : h->ng;
    b.dtg = dtg;
    b.cu = h->cu;
    b.cv = h->cv;
    b.cw = h->cw;
    b.mu = h->mu64[0];
    b.mv = h->mu64[1];
    b.mw = h->mu64[2];
    b.fu = h->mu32[0];
    b.fv = h->mu32[1];
    b.fw = h->mu32[2];
    return b;
}

hipError_t launch_build_mu(mdx_handle *h, double dtg) {
    const MuBuild b = mu_build(h, dtg);
    if (!b.n) return hipSuccess;
    hipLaunchKernelGGL(build_mu_kernel, dim3((unsigned)((b.n + 255) / 256)), dim3(256), 0, h->stream, b);
    return hipGetLastError();
}

template <typename T>
static MuK<T> mu_k(const mdx_handle *h, T *const *tabs) {
    MuK<T> m;
    m.u = h->ibm_on ? tabs[0] : nullptr;
    m.v = h->ibm_on ? tabs[1] : nullptr;
    m.w = h->ibm_on ? tabs[2] : nullptr;
    m.plane = h->coef_plane ? 1 : 0;
    m.sx = h->d.sx;
    m.sy = h->d.sy;
    return m;
}

// ------------------------------------------------------------------------------- residual stencil
// cell_residual (mdx_sor_core.h): r = rhs - L_mu(Phi) of one cell

template <typename T>
struct PlainPhi {
    const T *p;
    Dims d;
    __device__ inline T operator()(int i, int j, int k) const { return p[d.in(i, j, k)]; }
};

// Phi64 + e32, the corrected iterate of the defect-correction loop
// Phi_in + e of one cell; without e (the first check of a solve, e = 0) Phi_in + 0.0, the same
// value (signed zeros included)
template <bool E>
struct SumPhi {
    const double *p;
    const float *e;
    Dims d;
    __device__ inline double operator()(int i, int j, int k) const {
        size_t q = d.in(i, j, k);
        return p[q] + (E ? (double)e[q] : 0.0);
    }
};

// ------------------------------------------------------------------------------- z-marched residuals
// The FP64 residual kernels give a thread the column (i, j) of a 64 x 4 tile over RZ planes: the
// loads of its planes do not depend on the arithmetic and are issued together, Phi at k - 1 and
// k + 1 of consecutive cells are the same values (carried in registers), and a z-invariant
// coefficient table is read once per column.  Every cell's operands are the ones cell_residual
// forms, so each residual is that function's value.
static constexpr int RZ = 4;

__host__ inline unsigned rz_grid(const Dims &d) { return xcd_grid(d.nx, d.ny, (d.nz + RZ - 1) / RZ, 64, 4); }

// the column of this thread: (i, j) and its first plane k0; false for padding threads
__device__ inline bool rz_column(const Dims &d, int &i, int &j, int &k0) {
    int tx_ = 0, ty_ = 0, tz_ = 0;
    if (!xcd_tile(d.nx, d.ny, (d.nz + RZ - 1) / RZ, 64, 4, tx_, ty_, tz_)) return false;
    i = tx_ * 64 + threadIdx.x + 1;
    j = ty_ * 4 + threadIdx.y + 1;
    k0 = tz_ * RZ + 1;
    return i <= d.nx && j <= d.ny;
}

// ------------------------------------------------------------------------------- prepare / apply
// After u = u* (pressure_redblack.f90:64-84): Dirichlet v on the walls, rhs = div(u), Phi = 0.
// With part != nullptr the block maxima of |rhs| go to part: the residual max-norm of the start
// Phi = 0, whose residual is the right-hand side itself.
template <bool STRETCH>
__global__ __launch_bounds__(256) void sor_prepare_kernel(GridK g, double vbot, double vtop,
                                                          const double *__restrict__ us,
                                                          const double *__restrict__ vs,
                                                          const double *__restrict__ ws, double *rhs,
                                                          double *phi, double idx, double idy, double idz,
                                                          double *part) {
    const Dims d = g.d;
    int tx_, ty_, tz_;
    const bool tile = xcd_tile(d.nx, d.ny, d.nz, 64, 4, tx_, ty_, tz_);
    if (!tile && !part) return;
    const int i = tx_ * 64 + threadIdx.x + 1;
    const int j = ty_ * 4 + threadIdx.y + 1;
    const int k = tz_ + 1;
    double m = 0.0;
    if (tile && i <= d.nx && j <= d.ny) {
        const bool per2 = g.per[1];
        const int ip = i == d.nx ? 1 : i + 1, kp = k == d.nz ? 1 : k + 1;
        const int jp = (per2 && j == d.ny) ? 1 : j + 1;
        const size_t c = d.g(i, j, k);
        // v of u* with the wall values in place (the projection writes them into vn)
        const double v0 = (!per2 && j == 1) ? vbot : vs[c];
        const double v1 = (!per2 && j == d.ny) ? vtop : vs[d.g(i, jp, k)];
        const double dvy = STRETCH ? (v1 - v0) * g.ihc[j] : (v1 - v0) * idy;
        const double r = (us[d.g(ip, j, k)] - us[c]) * idx + dvy + (ws[d.g(i, j, kp)] - ws[c]) * idz;
        rhs[d.in(i, j, k)] = r;
        phi[d.in(i, j, k)] = 0.0;
        m = fabs(r);
    }
    if (part) block_max_part(m, part);
}

// The start of a mixed-precision solve in one pass (what sor_prepare, the previous-Phi copy and the
// first defect check compute, with the same arithmetic): rhs = div u* (wall values of v), Phi =
// fac * Phi_prev + 0.0 (or 0.0), r32 = rhs - L_mu(Phi) with every Phi operand formed the same way
// from Phi_prev, and the block maxima of |r|.
template <bool STRETCH, bool IBM>
__global__ __launch_bounds__(256) void sor_start_kernel(GridK g, MuK<double> mu, SorConst<double> s, double vbot,
                                                        double vtop, const double *__restrict__ us,
                                                        const double *__restrict__ vs,
                                                        const double *__restrict__ ws,
                                                        const double *__restrict__ prev, double fac, double *rhs,
                                                        double *phi, float *r32, double *out, double idx,
                                                        double idy, double idz) {
    const Dims d = g.d;
    int tx_ = 0, ty_ = 0, tz_ = 0;
    const bool ok = xcd_tile(d.nx, d.ny, d.nz, 64, 4, tx_, ty_, tz_);   // padding groups report 0
    const int i = tx_ * 64 + threadIdx.x + 1;
    const int j = ty_ * 4 + threadIdx.y + 1;
    const int k = tz_ + 1;
    double m = 0.0;
    if (ok && i <= d.nx && j <= d.ny) {
        const bool per2 = g.per[1];
        const int ip = i == d.nx ? 1 : i + 1, kp = k == d.nz ? 1 : k + 1;
        const int jp = (per2 && j == d.ny) ? 1 : j + 1;
        const size_t c = d.g(i, j, k), q = d.in(i, j, k);
        const double v0 = (!per2 && j == 1) ? vbot : vs[c];
        const double v1 = (!per2 && j == d.ny) ? vtop : vs[d.g(i, jp, k)];
        const double dvy = STRETCH ? (v1 - v0) * g.ihc[j] : (v1 - v0) * idy;
        const double rc = (us[d.g(ip, j, k)] - us[c]) * idx + dvy + (ws[d.g(i, j, kp)] - ws[c]) * idz;
        struct StartPhi {
            const double *p;
            double f;
            Dims d;
            __device__ double operator()(int a, int b, int e) const { return (p ? f * p[d.in(a, b, e)] : 0.0) + 0.0; }
        } P{prev, fac, d};
        double denom;
        const double r = cell_residual<double, STRETCH, IBM>(g, mu, s, P, rc, i, j, k, denom);
        rhs[q] = rc;
        phi[q] = P(i, j, k);
        r32[q] = (float)r;
        m = fabs(r);
    }
    block_max_part(m, out);
}

hipError_t launch_sor_start(mdx_handle *h, const double *prev, double fac, double *host_slot,
                            unsigned long long seq) {
    GridK g = grid_k(h);
    MuK<double> mu = mu_k<double>(h, h->mu64);
    SorConst<double> s{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    dim3 blk(64, 4), grd(xcd_grid(h->d.nx, h->d.ny, h->d.nz, 64, 4));
    const double vb = h->cfg.bc_value[bci(2, 0, 2)], vt = h->cfg.bc_value[bci(2, 1, 2)];
#define MDX_ST(S, I)                                                                                     \
    hipLaunchKernelGGL((sor_start_kernel<S, I>), grd, blk, 0, h->stream, g, mu, s, vb, vt, h->us, h->vs, h->ws, \
                       prev, fac, h->rhs, h->phi, h->r32, h->max_part, 1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz)
    if (h->stretched) { if (h->ibm_on) MDX_ST(true, true); else MDX_ST(true, false); }
    else { if (h->ibm_on) MDX_ST(false, true); else MDX_ST(false, false); }
#undef MDX_ST
    hipError_t err = hipGetLastError();
    return err != hipSuccess ? err : launch_max_final(h, grd.x, host_slot, seq);
}

// un, vn, wn on the ghost shell: the values of u*, with the top wall value of v (walled y) on the
// interior of the j = ny + 1 plane; the boundary conditions that follow the projection overwrite
// the faces, and edges and corners keep the u* values.  Thread `first` of `stride` threads.
// bc_faces: the projection's tiles write the faces the boundary conditions set (the interiors of
// the x and z faces, u and w on the interiors of the y faces), and they are left to them.
__device__ inline void proj_ghost(const Dims &d, int walled_y, double vtop, bool bc_faces, const double *__restrict__ us,
                                  const double *__restrict__ vs, const double *__restrict__ ws, double *un,
                                  double *vn, double *wn, long long first, long long stride) {
    const long long sz = d.nz + 2;
    const long long pa = (long long)d.sx * d.sy, pb = (long long)d.sx * d.nz, pc = (long long)d.ny * d.nz;
    const long long n = 2 * (pa + pb + pc);
    for (long long t = first; t < n; t += stride) {
        int i, j, k;
        long long q = t;
        if (q < 2 * pa) {            // the k = 0 and k = nz + 1 planes
            k = q >= pa ? (int)(sz - 1) : 0;
            q %= pa;
            i = (int)(q % d.sx);
            j = (int)(q / d.sx);
        } else if ((q -= 2 * pa) < 2 * pb) {   // the j = 0 and j = ny + 1 rows of the inner planes
            j = q >= pb ? d.sy - 1 : 0;
            q %= pb;
            i = (int)(q % d.sx);
            k = (int)(q / d.sx) + 1;
        } else {                     // the i = 0 and i = nx + 1 cells of the inner rows
            q -= 2 * pb;
            i = q >= pc ? d.sx - 1 : 0;
            q %= pc;
            j = (int)(q % d.ny) + 1;
            k = (int)(q / d.ny) + 1;
        }
        const size_t c = d.g(i, j, k);
        const bool in_i = i >= 1 && i <= d.nx, in_j = j >= 1 && j <= d.ny, in_k = k >= 1 && k <= d.nz;
        if (bc_faces && in_j && (in_k || in_i)) continue;   // an x or z face's interior
        if (!(bc_faces && in_i && in_k)) {                   // not a y face's interior
            un[c] = us[c];
            wn[c] = ws[c];
        }
        const bool top = walled_y && j == d.ny + 1 && in_i && in_k;
        vn[c] = top ? vtop : vs[c];
    }
}

// rhs and Phi = 0; with rhs_max_slot != nullptr also max |rhs| (the residual of Phi = 0) into that
// pinned host slot (value, error word, then seq), read after the stream's next synchronisation
hipError_t launch_sor_prepare(mdx_handle *h, double *rhs_max_slot, unsigned long long seq) {
    GridK g = grid_k(h);
    dim3 blk(64, 4), grd(xcd_grid(h->d.nx, h->d.ny, h->d.nz, 64, 4));
    const double vb = h->cfg.bc_value[bci(2, 0, 2)], vt = h->cfg.bc_value[bci(2, 1, 2)];
    double *part = rhs_max_slot ? h->max_part : nullptr;
    if (h->stretched)
        hipLaunchKernelGGL(sor_prepare_kernel<true>, grd, blk, 0, h->stream, g, vb, vt, h->us, h->vs, h->ws,
                           h->rhs, h->phi, 1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz, part);
    else
        hipLaunchKernelGGL(sor_prepare_kernel<false>, grd, blk, 0, h->stream, g, vb, vt, h->us, h->vs, h->ws,
                           h->rhs, h->phi, 1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !rhs_max_slot) return e;
    return launch_max_final(h, grd.x, rhs_max_slot, seq);
}

static constexpr unsigned PROJ_GHOST_BLOCKS = 256;

// un = u* - mu grad Phi (likewise v, w) and pn += Phi / dt on the interior, reading u* directly;
// on a walled y the bottom v face keeps its wall value.  A thread takes the column (i, j) of a
// 64 x 4 tile over RZ planes (the residual kernels' layout): its loads are issued before its
// stores, Phi at k - 1 is the previous cell's Phi, and a z-invariant coefficient table is read once.
// The `tiles` workgroups of the tiles are followed by PROJ_GHOST_BLOCKS that fill the ghost shell
// (proj_ghost: cells no tile writes, read from cells no tile reads).  hist, when given, receives
// a copy of Phi.  BC (x and z periodic, Dirichlet u, v, w on the y walls): the tiles also write the
// ghost values the boundary conditions then set, each from the one interior value it depends on
// (bc_rb_y, bc_rb_x, bc_rb_z in that order give the same values): the periodic copies of u, v, w,
// p on the x and z faces and u, w on the y faces (2 V - the wall-adjacent value; ub2 = 2 V of u at
// the bottom wall, ut2 at the top, wb2 and wt2 of w), so no fill follows the projection.
template <bool STRETCH, bool IBM, bool BC>
__global__ __launch_bounds__(256) void sor_apply_kernel(GridK g, MuK<double> mu, const double *__restrict__ phi,
                                                        const double *__restrict__ us,
                                                        const double *__restrict__ vs,
                                                        const double *__restrict__ ws, double vbot,
                                                        double *__restrict__ un, double *__restrict__ vn,
                                                        double *__restrict__ wn, double *__restrict__ pn, double idx,
                                                        double idy, double idz, double idt, unsigned tiles,
                                                        double vtop, double *__restrict__ hist, double ub2,
                                                        double ut2, double wb2, double wt2) {
    const Dims d = g.d;
    if (blockIdx.x >= tiles) {   // the workgroups past the tiles: un, vn, wn on the ghost shell
        proj_ghost(d, g.per[1] ? 0 : 1, vtop, BC, us, vs, ws, un, vn, wn,
                   (long long)(blockIdx.x - tiles) * 256 + threadIdx.x + 64 * threadIdx.y,
                   (long long)(gridDim.x - tiles) * 256);
        return;
    }
    int i = 0, j = 0, k0 = 0;
    if (!rz_column(d, i, j, k0)) return;
    const bool per2 = g.per[1];
    const bool vwall = !(j >= 2 || per2);
    const int im = i == 1 ? d.nx : i - 1, jm = j == 1 ? d.ny : j - 1;
    const int n = d.nz - k0 + 1 < RZ ? d.nz - k0 + 1 : RZ;
    double pz[RZ + 1], pxm[RZ], pym[RZ], uo[RZ], vo[RZ], wo[RZ], po[RZ];
    pz[0] = phi[d.in(i, j, k0 == 1 ? d.nz : k0 - 1)];
#pragma unroll
    for (int u = 0; u < RZ; ++u) {
        if (u < n) {
            const int k = k0 + u;
            const size_t c = d.g(i, j, k);
            pz[u + 1] = phi[d.in(i, j, k)];
            pxm[u] = phi[d.in(im, j, k)];
            pym[u] = vwall ? 0.0 : phi[d.in(i, jm, k)];
            uo[u] = us[c];
            vo[u] = vwall ? 0.0 : vs[c];
            wo[u] = ws[c];
            po[u] = pn[c];
        }
    }
    double mu_u = 1.0, mu_v = 1.0, mu_w = 1.0;
    if (IBM && mu.plane) {
        mu_u = mu.u[mu.at(i, j, 1)];
        mu_v = mu.v[mu.at(i, j, 1)];
        mu_w = mu.w[mu.at(i, j, 1)];
    }
    const double ihv = STRETCH ? g.ihv[j] : 0.0;
#pragma unroll
    for (int u = 0; u < RZ; ++u) {
        if (u < n) {
            const int k = k0 + u;
            const size_t c = d.g(i, j, k);
            if (IBM && !mu.plane) {
                mu_u = mu.u[mu.at(i, j, k)];
                mu_v = mu.v[mu.at(i, j, k)];
                mu_w = mu.w[mu.at(i, j, k)];
            }
            const double p0 = pz[u + 1];
            if (hist) hist[d.in(i, j, k)] = p0;
            const double uv = uo[u] + mu_u * (pxm[u] - p0) * idx;
            const double wv = wo[u] + mu_w * (pz[u] - p0) * idz;
            double vv = vbot;
            if (!vwall) {
                const double gy = STRETCH ? (pym[u] - p0) * ihv : (pym[u] - p0) * idy;
                vv = vo[u] + mu_v * gy;
            }
            const double pv = po[u] + p0 * idt;
            un[c] = uv;
            wn[c] = wv;
            vn[c] = vv;
            pn[c] = pv;
            if (BC) {
                auto put = [&](size_t q) {
                    un[q] = uv;
                    vn[q] = vv;
                    wn[q] = wv;
                    pn[q] = pv;
                };
                if (i == d.nx) put(d.g(0, j, k));
                if (i == 1) put(d.g(d.nx + 1, j, k));
                if (k == d.nz) put(d.g(i, j, 0));
                if (k == 1) put(d.g(i, j, d.nz + 1));
                if (j == 1) {
                    const size_t q = d.g(i, 0, k);
                    un[q] = ub2 - uv;
                    wn[q] = wb2 - wv;
                }
                if (j == d.ny) {
                    const size_t q = d.g(i, d.ny + 1, k);
                    un[q] = ut2 - uv;
                    wn[q] = wt2 - wv;
                }
            }
        }
    }
}

// the boundary conditions the projection kernel writes itself inside a step: x and z periodic,
// Dirichlet u, v, w on both y walls (the channel's), unless MOBYDIFF_FUSED_BC=0
bool apply_writes_bc(const mdx_handle *h) {
    const mdx_config &c = h->cfg;
    if (!h->fused_bc || !c.periodic[0] || !c.periodic[2] || c.periodic[1]) return false;
    for (int s = 0; s < 2; ++s)
        for (int q = 1; q <= 3; ++q)
            if (c.bc_type[bci(2, s, q)] != 0) return false;
    return true;
}

hipError_t launch_sor_apply(mdx_handle *h, const double *phi, double dtg, double *hist) {
    GridK g = grid_k(h);
    MuK<double> mu = mu_k<double>(h, h->mu64);
    const unsigned tiles = rz_grid(h->d);
    dim3 blk(64, 4), grd(tiles + PROJ_GHOST_BLOCKS);
    const double a = 1.0 / h->dx, b = 1.0 / h->dy, c = 1.0 / h->dz, t = 1.0 / dtg;
    const double vb = h->cfg.bc_value[bci(2, 0, 2)], vt = h->cfg.bc_value[bci(2, 1, 2)];
    const bool bc = h->step_bc && apply_writes_bc(h);
    const double *bv = h->cfg.bc_value;
    const double ub2 = 2 * bv[bci(2, 0, 1)], ut2 = 2 * bv[bci(2, 1, 1)];
    const double wb2 = 2 * bv[bci(2, 0, 3)], wt2 = 2 * bv[bci(2, 1, 3)];
#define MDX_APPLY_(S, I, B)                                                                                    \
    hipLaunchKernelGGL((sor_apply_kernel<S, I, B>), grd, blk, 0, h->stream, g, mu, phi, h->us, h->vs, h->ws, vb, \
                       h->un, h->vn, h->wn, h->pn, a, b, c, t, tiles, vt, hist, ub2, ut2, wb2, wt2)
#define MDX_APPLY(S, I) do { if (bc) MDX_APPLY_(S, I, true); else MDX_APPLY_(S, I, false); } while (0)
    if (h->stretched) { if (h->ibm_on) MDX_APPLY(true, true); else MDX_APPLY(true, false); }
    else { if (h->ibm_on) MDX_APPLY(false, true); else MDX_APPLY(false, false); }
#undef MDX_APPLY
#undef MDX_APPLY_
    const hipError_t e = hipGetLastError();
    h->apply_wrote_bc = e == hipSuccess && bc;
    return e;
}

// ------------------------------------------------------------------------------- z-marching sweeps
// One thread owns the x-pair of cells (i0, i0+1), i0 = 2p+1, in row j and marches KZ planes in z.
// In every plane exactly one cell of the pair has the sweep's colour, and it alternates with k, so
// the pair is loaded once per plane as one 2-element vector and kept in registers for the z
// neighbours of the next plane; the x neighbour inside the pair comes from the same vector.  Per
// updated cell this leaves one vector load, three scalar loads (outer x neighbour, y neighbours),
// rhs and one vector store of the pair (its other cell is rewritten with its own, unchanged bits:
// cells of the other colour are not written by this sweep).  z-invariant mu tables are loaded
// once per thread.  The arithmetic is sor_residual_core with the operands of cell_residual, so the
// iterates are bit-identical to the one-cell-per-thread sweep.
template <typename T> struct Pair;
template <> struct Pair<float> { using V = float2; };
template <> struct Pair<double> { using V = double2; };

static constexpr int ZX = 32;   // pairs per block row
static constexpr int ZY = 8;    // rows per block

// MU: 0 no immersed boundary (mu = 1), 1 z-invariant mu planes, 2 mu varying in z
template <typename T, bool STRETCH, int MU, int KZ>
__global__ __launch_bounds__(ZX * ZY) void sor_sweep_zm_kernel(GridK g, MuK<T> mu, SorConst<T> s, int color,
                                                               T omega, const T *__restrict__ rhs, T *phi) {
    using V = typename Pair<T>::V;
    const Dims d = g.d;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int np = nx / 2, nseg = (nz + KZ - 1) / KZ;
    int tx_, ty_, tz_;
    if (!xcd_tile(np, ny, nseg, ZX, ZY, tx_, ty_, tz_)) return;
    const int p = tx_ * ZX + threadIdx.x;
    const int j = ty_ * ZY + threadIdx.y + 1;
    if (p >= np || j > ny) return;
    const int k0 = tz_ * KZ + 1;
    const int k1 = k0 + KZ - 1 < nz ? k0 + KZ - 1 : nz;
    const int i0 = 2 * p + 1;
    const int iw = i0 == 1 ? nx : i0 - 1;          // west neighbour of i0
    const int ie = i0 + 1 == nx ? 1 : i0 + 2;      // east neighbour of i0+1
    const bool per2 = g.per[1];
    int jp = j + 1, jm = j - 1;
    if (per2) {
        if (jp > ny) jp = 1;
        if (jm < 1) jm = ny;
    }
    const bool wn_ = !per2 && j == ny, ws_ = !per2 && j == 1;
    const int jn = wn_ ? j : jp, js = ws_ ? j : jm;
    T an = 0, as = 0;
    if (STRETCH) {
        an = (T)(g.ihv[j + 1] * g.ihc[j]);
        as = (T)(g.ihv[j] * g.ihc[j]);
    }
    // mu of the two cells: u faces at i0, i0+1, ie; v faces at j, jp; w faces at k, kp
    T mu_u0 = 1, mu_u1 = 1, mu_u2 = 1, mv0 = 1, mv1 = 1, mvp0 = 1, mvp1 = 1;
    T mw0 = 1, mw1 = 1, mwp0 = 1, mwp1 = 1;
    if (MU == 1) {
        mu_u0 = mu.u[mu.at(i0, j, 0)];
        mu_u1 = mu.u[mu.at(i0 + 1, j, 0)];
        mu_u2 = mu.u[mu.at(ie, j, 0)];
        mv0 = mu.v[mu.at(i0, j, 0)];
        mv1 = mu.v[mu.at(i0 + 1, j, 0)];
        mvp0 = mu.v[mu.at(i0, jp, 0)];
        mvp1 = mu.v[mu.at(i0 + 1, jp, 0)];
        mw0 = mwp0 = mu.w[mu.at(i0, j, 0)];
        mw1 = mwp1 = mu.w[mu.at(i0 + 1, j, 0)];
    }
    const size_t rs = (size_t)nx, ps = (size_t)nx * ny;
    const size_t rowj = (size_t)(i0 - 1) + rs * (size_t)(j - 1);   // pair offset within a plane
    const size_t rown = rs * (size_t)(jn - 1);
    const size_t rows = rs * (size_t)(js - 1);
    const V *P2 = reinterpret_cast<const V *>(phi);
    V *W2 = reinterpret_cast<V *>(phi);
    const int nk = k1 - k0 + 1;
    // Every value the segment reads is fetched before the first store: loads issued after a store
    // through the same pointer cannot be reordered above it, which would serialise the planes.
    // None of these values is written by this sweep (other-colour cells, or this thread's own
    // pairs before their update), so the prefetch reads exactly what the plane-by-plane order reads.
    V pr[KZ + 2];
    T pout[KZ], pyp[KZ], pym[KZ], rcv[KZ];
    T nmu_u0[KZ], nmu_u1[KZ], nmu_u2[KZ], nmv0[KZ], nmv1[KZ], nmvp0[KZ], nmvp1[KZ];
    T nmw0[KZ], nmw1[KZ], nmwp0[KZ], nmwp1[KZ];
#pragma unroll
    for (int t = 0; t < KZ + 2; ++t) {
        if (t <= nk + 1) {
            int kk = k0 - 1 + t;
            if (kk < 1) kk = nz;
            if (kk > nz) kk = 1;
            pr[t] = P2[(rowj + ps * (size_t)(kk - 1)) / 2];
        }
    }
#pragma unroll
    for (int t = 0; t < KZ; ++t) {
        if (t < nk) {
            const int k = k0 + t;
            const int c = (j + k + color) & 1;
            const int i = i0 + c;
            const size_t pk = ps * (size_t)(k - 1);
            pout[t] = phi[(size_t)((c ? ie : iw) - 1) + rs * (size_t)(j - 1) + pk];
            pyp[t] = phi[(size_t)(i - 1) + rown + pk];
            pym[t] = phi[(size_t)(i - 1) + rows + pk];
            rcv[t] = rhs[rowj + (size_t)c + pk];
            if (MU == 2) {
                const int kp = k == nz ? 1 : k + 1;
                nmu_u0[t] = mu.u[mu.at(i0, j, k)];
                nmu_u1[t] = mu.u[mu.at(i0 + 1, j, k)];
                nmu_u2[t] = mu.u[mu.at(ie, j, k)];
                nmv0[t] = mu.v[mu.at(i0, j, k)];
                nmv1[t] = mu.v[mu.at(i0 + 1, j, k)];
                nmvp0[t] = mu.v[mu.at(i0, jp, k)];
                nmvp1[t] = mu.v[mu.at(i0 + 1, jp, k)];
                nmw0[t] = mu.w[mu.at(i0, j, k)];
                nmw1[t] = mu.w[mu.at(i0 + 1, j, k)];
                nmwp0[t] = mu.w[mu.at(i0, j, kp)];
                nmwp1[t] = mu.w[mu.at(i0 + 1, j, kp)];
            }
        }
    }
#pragma unroll
    for (int t = 0; t < KZ; ++t) {
        if (t < nk) {
            const int k = k0 + t;
            const int c = (j + k + color) & 1;
            if (MU == 2) {
                mu_u0 = nmu_u0[t];
                mu_u1 = nmu_u1[t];
                mu_u2 = nmu_u2[t];
                mv0 = nmv0[t];
                mv1 = nmv1[t];
                mvp0 = nmvp0[t];
                mvp1 = nmvp1[t];
                mw0 = nmw0[t];
                mw1 = nmw1[t];
                mwp0 = nmwp0[t];
                mwp1 = nmwp1[t];
            }
            V cur = pr[t + 1];
            const V prv = pr[t], nxt = pr[t + 2];
            const T p0 = c ? cur.y : cur.x;
            const T pin = c ? cur.x : cur.y;
            const T pxp = c ? pout[t] : pin, pxm = c ? pin : pout[t];
            const T pzp = c ? nxt.y : nxt.x, pzm = c ? prv.y : prv.x;
            const T mui = c ? mu_u1 : mu_u0, muip = c ? mu_u2 : mu_u1;
            const T mvj = c ? mv1 : mv0, mvjp = c ? mvp1 : mvp0;
            const T mwk = c ? mw1 : mw0, mwkp = c ? mwp1 : mwp0;
            T denom;
            const T r = sor_residual_core<T, STRETCH>(p0, pxp, pxm, pyp[t], pym[t], pzp, pzm, mui, muip, mvj, mvjp,
                                                      mwk, mwkp, wn_, ws_, rcv[t], s, an, as, denom);
            const T nv = p0 + r * (-omega / denom);
            if (c) cur.y = nv;
            else cur.x = nv;
            W2[(rowj + ps * (size_t)(k - 1)) / 2] = cur;
        }
    }
}

template <typename T>
static hipError_t launch_sweep_zm(mdx_handle *h, int color, T omega, const MuK<T> &mu, const SorConst<T> &s,
                                  const T *rhs, T *phi) {
    GridK g = grid_k(h);
    const int kz = h->sweep_kz;
    dim3 blk(ZX, ZY), grd(xcd_grid(h->d.nx / 2, h->d.ny, (h->d.nz + kz - 1) / kz, ZX, ZY));
#define MDX_ZM(S, I, K) \
    hipLaunchKernelGGL((sor_sweep_zm_kernel<T, S, I, K>), grd, blk, 0, h->stream, g, mu, s, color, omega, rhs, phi)
    const int mumode = h->ibm_on ? (h->coef_plane ? 1 : 2) : 0;
#define MDX_ZMK(S, I)                      \
    switch (kz) {                          \
    case 1: MDX_ZM(S, I, 1); break;        \
    case 2: MDX_ZM(S, I, 2); break;        \
    case 8: MDX_ZM(S, I, 8); break;        \
    default: MDX_ZM(S, I, 4); break;       \
    }
    if (h->stretched) {
        if (mumode == 1) { MDX_ZMK(true, 1) } else if (mumode == 2) { MDX_ZMK(true, 2) } else { MDX_ZMK(true, 0) }
    } else {
        if (mumode == 1) { MDX_ZMK(false, 1) } else if (mumode == 2) { MDX_ZMK(false, 2) } else { MDX_ZMK(false, 0) }
    }
#undef MDX_ZMK
#undef MDX_ZM
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- fused iteration
// One full red-black iteration (colour 0, then colour 1) per launch, reading the old iterate from
// src and writing the new one to dst (no workgroup reads what another writes in the same launch).
// A workgroup owns FPX x-pairs x FPY rows x KZF planes and is (FPX+2) x (FPY+2) threads: thread
// (tx, ty) owns the x-pair p0-1+tx in row j0-1+ty; the outer ring of threads only computes
// colour-0 values, for the x/y neighbours of the tile's colour-1 cells.  Every thread first loads
// all old values its segment needs (as the one-colour sweep does, so the loads overlap), then
// computes the colour-0 cells of its column, planes k0-1 .. k1+1, publishes planes k0 .. k1 in LDS,
// and after one barrier updates the colour-1 cells of planes k0 .. k1: x and y neighbours from LDS,
// z neighbours from its own colour-0 values.  Colour-0 values of the ring and of planes k0-1, k1+1
// are recomputed by every workgroup that needs them, from the same inputs and with the same
// arithmetic (sor_residual_core with the operands of cell_residual), so the iterates are
// bit-identical to two one-colour sweeps.
static constexpr int FPX = 16, FPY = 8;
static constexpr int FWX = FPX + 2, FWY = FPY + 2;

template <typename T, bool STRETCH, int MU, int KZF>
__global__ __launch_bounds__(FWX * FWY) void sor_iter_fused_kernel(GridK g, MuK<T> mu, SorConst<T> s, T omega,
                                                                   const T *__restrict__ rhs,
                                                                   const T *__restrict__ src, T *__restrict__ dst) {
    using V = typename Pair<T>::V;
    __shared__ T red[KZF][FWY][FWX];
    const Dims d = g.d;
    const int nx = d.nx, ny = d.ny, nz = d.nz, np = nx / 2;
    int tx_, ty_, tz_;
    if (!xcd_tile(np, ny, (nz + KZF - 1) / KZF, FPX, FPY, tx_, ty_, tz_)) return;   // uniform per workgroup
    const int tx = threadIdx.x, ty = threadIdx.y;
    // pair index and row of this thread (x periodic: any number of periods; y periodic or walled)
    const int pr = tx_ * FPX - 1 + tx;
    const int p = ((pr % np) + np) % np;
    int j = ty_ * FPY + ty;   // = j0 - 1 + ty with j0 = ty_ * FPY + 1
    const bool per2 = g.per[1];
    bool valid = true;
    if (per2) j = ((j - 1) % ny + ny) % ny + 1;
    else valid = j >= 1 && j <= ny;
    const bool inner = tx >= 1 && tx <= FPX && ty >= 1 && ty <= FPY && pr < np && ty_ * FPY + ty <= ny;
    const int k0 = tz_ * KZF + 1;
    const int k1 = k0 + KZF - 1 < nz ? k0 + KZF - 1 : nz;
    const int nk = k1 - k0 + 1;
    const int i0 = 2 * p + 1;
    const int iw = i0 == 1 ? nx : i0 - 1, ie = i0 + 1 == nx ? 1 : i0 + 2;
    const int jj = valid ? j : 1;   // a safe row for address arithmetic of invalid (never used) rows
    int jp = jj + 1, jm = jj - 1;
    if (per2) {
        if (jp > ny) jp = 1;
        if (jm < 1) jm = ny;
    }
    const bool wn_ = !per2 && jj == ny, ws_ = !per2 && jj == 1;
    const int jn = wn_ ? jj : jp, js = ws_ ? jj : jm;
    T an = 0, as = 0;
    if (STRETCH) {
        an = (T)(g.ihv[jj + 1] * g.ihc[jj]);
        as = (T)(g.ihv[jj] * g.ihc[jj]);
    }
    T mu_u0 = 1, mu_u1 = 1, mu_u2 = 1, mv0 = 1, mv1 = 1, mvp0 = 1, mvp1 = 1, mw0 = 1, mw1 = 1, mwp0 = 1, mwp1 = 1;
    if (MU == 1) {
        mu_u0 = mu.u[mu.at(i0, jj, 0)];
        mu_u1 = mu.u[mu.at(i0 + 1, jj, 0)];
        mu_u2 = mu.u[mu.at(ie, jj, 0)];
        mv0 = mu.v[mu.at(i0, jj, 0)];
        mv1 = mu.v[mu.at(i0 + 1, jj, 0)];
        mvp0 = mu.v[mu.at(i0, jp, 0)];
        mvp1 = mu.v[mu.at(i0 + 1, jp, 0)];
        mw0 = mwp0 = mu.w[mu.at(i0, jj, 0)];
        mw1 = mwp1 = mu.w[mu.at(i0 + 1, jj, 0)];
    }
    const size_t rs = (size_t)nx, ps = (size_t)nx * ny;
    const size_t rowj = (size_t)(i0 - 1) + rs * (size_t)(jj - 1);
    const size_t rown = rs * (size_t)(jn - 1), rows = rs * (size_t)(js - 1);
    const V *S2 = reinterpret_cast<const V *>(src);
    V *D2 = reinterpret_cast<V *>(dst);
    auto kw = [&](int k) { return ((k - 1) % nz + nz) % nz + 1; };   // periodic z, any number of periods
    // the update of cell (i0 + c, j, kk) from its 7 values and the face coefficients
    auto relax = [&](int kk, int c, T p0, T pin, T pout, T pyp, T pym, T pzp, T pzm, T rc) -> T {
        if (MU == 2) {
            const int kp = kk == nz ? 1 : kk + 1;
            mu_u0 = mu.u[mu.at(i0, jj, kk)];
            mu_u1 = mu.u[mu.at(i0 + 1, jj, kk)];
            mu_u2 = mu.u[mu.at(ie, jj, kk)];
            mv0 = mu.v[mu.at(i0, jj, kk)];
            mv1 = mu.v[mu.at(i0 + 1, jj, kk)];
            mvp0 = mu.v[mu.at(i0, jp, kk)];
            mvp1 = mu.v[mu.at(i0 + 1, jp, kk)];
            mw0 = mu.w[mu.at(i0, jj, kk)];
            mw1 = mu.w[mu.at(i0 + 1, jj, kk)];
            mwp0 = mu.w[mu.at(i0, jj, kp)];
            mwp1 = mu.w[mu.at(i0 + 1, jj, kp)];
        }
        const T pxp = c ? pout : pin, pxm = c ? pin : pout;
        const T mui = c ? mu_u1 : mu_u0, muip = c ? mu_u2 : mu_u1;
        const T mvj = c ? mv1 : mv0, mvjp = c ? mvp1 : mvp0;
        const T mwk = c ? mw1 : mw0, mwkp = c ? mwp1 : mwp0;
        T denom;
        const T r = sor_residual_core<T, STRETCH>(p0, pxp, pxm, pyp, pym, pzp, pzm, mui, muip, mvj, mvjp, mwk, mwkp,
                                                  wn_, ws_, rc, s, an, as, denom);
        return p0 + r * (-omega / denom);
    };
    // Colour-0 planes t = 0 .. nk+1 are k = k0-1+t; pair planes q = 0 .. nk+3 are k = k0-2+q.
    // The ring threads need colour 0 of planes k0 .. k1 only (t = 1 .. nk, pairs q = 1 .. nk+2).
    V pr2[KZF + 4];
    T pout[KZF + 2], pyp[KZF + 2], pym[KZF + 2], rc0[KZF + 2], rc1[KZF];
#pragma unroll
    for (int q = 0; q < KZF + 4; ++q) {
        const bool need = valid && q <= nk + 3 && (inner || (q >= 1 && q <= nk + 2));
        if (need) pr2[q] = S2[(rowj + ps * (size_t)(kw(k0 - 2 + q) - 1)) / 2];
    }
#pragma unroll
    for (int t = 0; t < KZF + 2; ++t) {
        const bool need = valid && t <= nk + 1 && (inner || (t >= 1 && t <= nk));
        if (need) {
            const int kk = kw(k0 - 1 + t);
            const int c0 = (jj + kk) & 1, ic = i0 + c0;
            const size_t pk = ps * (size_t)(kk - 1);
            pout[t] = src[(size_t)((c0 ? ie : iw) - 1) + rs * (size_t)(jj - 1) + pk];
            pyp[t] = src[(size_t)(ic - 1) + rown + pk];
            pym[t] = src[(size_t)(ic - 1) + rows + pk];
            rc0[t] = rhs[rowj + (size_t)c0 + pk];
        }
    }
#pragma unroll
    for (int u = 0; u < KZF; ++u) {
        if (inner && u < nk) {
            const int kk = k0 + u;   // k0 .. k1 lie in 1 .. nz
            rc1[u] = rhs[rowj + (size_t)((jj + kk + 1) & 1) + ps * (size_t)(kk - 1)];
        }
    }
    // colour 0 of planes k0-1 .. k1+1 (own column): every neighbour is colour 1, so old values
    T r0[KZF + 2];
#pragma unroll
    for (int t = 0; t < KZF + 2; ++t) {
        r0[t] = 0;
        const bool need = valid && t <= nk + 1 && (inner || (t >= 1 && t <= nk));
        if (need) {
            const int kk = kw(k0 - 1 + t);
            const int c0 = (jj + kk) & 1;
            const V cur = pr2[t + 1], nxt = pr2[t + 2], prv = pr2[t];
            r0[t] = relax(kk, c0, c0 ? cur.y : cur.x, c0 ? cur.x : cur.y, pout[t], pyp[t], pym[t], c0 ? nxt.y : nxt.x,
                          c0 ? prv.y : prv.x, rc0[t]);
        }
    }
#pragma unroll
    for (int u = 0; u < KZF; ++u) red[u][ty][tx] = r0[u + 1];
    __syncthreads();
    // colour 1 of planes k0 .. k1 (tile threads)
    if (!inner) return;
#pragma unroll
    for (int u = 0; u < KZF; ++u) {
        if (u < nk) {
            const int kk = k0 + u;
            const int c1 = (jj + kk + 1) & 1;
            const T pin = r0[u + 1];   // own pair's colour-0 cell
            const T po = c1 ? red[u][ty][tx + 1] : red[u][ty][tx - 1];
            const T yp = wn_ ? T(0) : red[u][ty + 1][tx], ym = ws_ ? T(0) : red[u][ty - 1][tx];
            const V cur = pr2[u + 2];
            const T bnew = relax(kk, c1, c1 ? cur.y : cur.x, pin, po, yp, ym, r0[u + 2], r0[u], rc1[u]);
            V out;
            if (c1) { out.x = pin; out.y = bnew; } else { out.x = bnew; out.y = pin; }
            D2[(rowj + ps * (size_t)(kk - 1)) / 2] = out;
        }
    }
}

template <typename T>
static hipError_t launch_iter_fused(mdx_handle *h, T omega, const MuK<T> &mu, const SorConst<T> &s, const T *rhs,
                                    const T *src, T *dst) {
    GridK g = grid_k(h);
    const int kzf = h->sweep_kz == 2 || h->sweep_kz == 8 ? h->sweep_kz : 4;
    dim3 blk(FWX, FWY), grd(xcd_grid(h->d.nx / 2, h->d.ny, (h->d.nz + kzf - 1) / kzf, FPX, FPY));
    const int mumode = h->ibm_on ? (h->coef_plane ? 1 : 2) : 0;
#define MDX_FU(S, M, K) \
    hipLaunchKernelGGL((sor_iter_fused_kernel<T, S, M, K>), grd, blk, 0, h->stream, g, mu, s, omega, rhs, src, dst)
#define MDX_FUK(S, M)                   \
    switch (kzf) {                      \
    case 2: MDX_FU(S, M, 2); break;     \
    case 8: MDX_FU(S, M, 8); break;     \
    default: MDX_FU(S, M, 4); break;    \
    }
    if (h->stretched) {
        if (mumode == 1) { MDX_FUK(true, 1) } else if (mumode == 2) { MDX_FUK(true, 2) } else { MDX_FUK(true, 0) }
    } else {
        if (mumode == 1) { MDX_FUK(false, 1) } else if (mumode == 2) { MDX_FUK(false, 2) } else { MDX_FUK(false, 0) }
    }
#undef MDX_FUK
#undef MDX_FU
    return hipGetLastError();
}

template <typename T>
__global__ void copy_kernel(T *__restrict__ dst, const T *__restrict__ src, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q < n; q += (size_t)gridDim.x * blockDim.x) dst[q] = src[q];
}

// ------------------------------------------------------------------------------- FP64 sweeps
template <bool STRETCH, bool IBM>
__global__ __launch_bounds__(SX * SY) void sor_sweep64_kernel(GridK g, MuK<double> mu, SorConst<double> s,
                                                              int color, double omega,
                                                              const double *__restrict__ rhs, double *phi) {
    const Dims d = g.d;
    int tx_, ty_, tz_;
    if (!xcd_tile(d.nx / 2, d.ny, d.nz, SX, SY, tx_, ty_, tz_)) return;
    const int ni = tx_ * SX + threadIdx.x + 1;
    const int j = ty_ * SY + threadIdx.y + 1;
    const int k = tz_ + 1;
    if (ni > d.nx / 2 || j > d.ny) return;
    const int i = 1 + ((j + k + color) & 1) + 2 * (ni - 1);
    PlainPhi<double> P{phi, d};
    double denom;
    const double r = cell_residual<double, STRETCH, IBM>(g, mu, s, P, rhs[d.in(i, j, k)], i, j, k, denom);
    const size_t q = d.in(i, j, k);
    phi[q] = phi[q] + r * (-omega / denom);
}

hipError_t launch_sor_sweep64(mdx_handle *h, int color, double omega) {
    GridK g = grid_k(h);
    MuK<double> mu = mu_k<double>(h, h->mu64);
    SorConst<double> s{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    if (h->sweep_kz > 0) return launch_sweep_zm<double>(h, color, omega, mu, s, h->rhs, h->phi);
    dim3 blk(SX, SY), grd(xcd_grid(h->d.nx / 2, h->d.ny, h->d.nz, SX, SY));
#define MDX_SW(S, I) \
    hipLaunchKernelGGL((sor_sweep64_kernel<S, I>), grd, blk, 0, h->stream, g, mu, s, color, omega, h->rhs, h->phi)
    if (h->stretched) { if (h->ibm_on) MDX_SW(true, true); else MDX_SW(true, false); }
    else { if (h->ibm_on) MDX_SW(false, true); else MDX_SW(false, false); }
#undef MDX_SW
    return hipGetLastError();
}

// the right-hand side of this column's cells (0 past the last plane)
__device__ inline void rhs_column(const Dims &d, const double *__restrict__ rhs, int i, int j, int k0,
                                  double (&rv)[RZ]) {
#pragma unroll
    for (int u = 0; u < RZ; ++u) rv[u] = k0 + u <= d.nz ? rhs[d.in(i, j, k0 + u)] : 0.0;
}

// residuals r[u] = rv[u] - L_mu(Phi) of cells (i, j, k0 + u), u < n = min(RZ, nz - k0 + 1)
// (rv: the right-hand side of those cells), and Phi of those cells in p[u]
template <bool STRETCH, bool IBM, typename PhiAt>
__device__ inline int residual_column(const GridK &g, const MuK<double> &mu, const SorConst<double> &s, PhiAt P,
                                      const double (&rv)[RZ], int i, int j, int k0, double (&r)[RZ],
                                      double (&p)[RZ]) {
    const Dims d = g.d;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const bool per2 = g.per[1];
    const int ip = i == nx ? 1 : i + 1, im = i == 1 ? nx : i - 1;
    int jp = j + 1, jm = j - 1;
    if (per2) {
        if (jp > ny) jp = 1;
        if (jm < 1) jm = ny;
    }
    const bool wn_ = !per2 && j == ny, ws_ = !per2 && j == 1;
    const int jn = wn_ ? j : jp, js = ws_ ? j : jm;
    double an = 0, as = 0;
    if (STRETCH) {
        an = g.ihv[j + 1] * g.ihc[j];
        as = g.ihv[j] * g.ihc[j];
    }
    const int n = nz - k0 + 1 < RZ ? nz - k0 + 1 : RZ;
    double pz[RZ + 2];   // Phi at k0 - 1 .. k0 + n (periodic in z), of this column
    pz[0] = P(i, j, k0 == 1 ? nz : k0 - 1);
#pragma unroll
    for (int u = 0; u <= RZ; ++u) {
        const int k = k0 + u;
        pz[u + 1] = u <= n ? P(i, j, k > nz ? k - nz : k) : 0.0;
    }
    double mui = 1, muip = 1, mvj = 1, mvjp = 1, mwc = 1;
    if (IBM && mu.plane) {   // z-invariant table: the column's six face coefficients (w's twice)
        mui = mu.u[mu.at(i, j, 1)];
        muip = mu.u[mu.at(ip, j, 1)];
        mvj = mu.v[mu.at(i, j, 1)];
        mvjp = mu.v[mu.at(i, jp, 1)];
        mwc = mu.w[mu.at(i, j, 1)];
    }
#pragma unroll
    for (int u = 0; u < RZ; ++u) {
        r[u] = 0.0;
        p[u] = pz[u + 1];
        if (u < n) {
            const int k = k0 + u, kp = k == nz ? 1 : k + 1;
            double mwk = mwc, mwkp = mwc;
            if (IBM && !mu.plane) {
                mui = mu.u[mu.at(i, j, k)];
                muip = mu.u[mu.at(ip, j, k)];
                mvj = mu.v[mu.at(i, j, k)];
                mvjp = mu.v[mu.at(i, jp, k)];
                mwk = mu.w[mu.at(i, j, k)];
                mwkp = mu.w[mu.at(i, j, kp)];
            }
            double denom;
            r[u] = sor_residual_core<double, STRETCH>(pz[u + 1], P(ip, j, k), P(im, j, k), P(i, jn, k), P(i, js, k),
                                                     pz[u + 2], pz[u], mui, muip, mvj, mvjp, mwk, mwkp, wn_, ws_,
                                                     rv[u], s, an, as, denom);
        }
    }
    return n;
}

template <bool STRETCH, bool IBM>
__global__ __launch_bounds__(256) void sor_residual64_kernel(GridK g, MuK<double> mu, SorConst<double> s,
                                                             const double *__restrict__ rhs,
                                                             const double *__restrict__ phi, double *out,
                                                             double *__restrict__ field) {
    const Dims d = g.d;
    int i = 0, j = 0, k0 = 0;
    double m = 0.0;
    if (rz_column(d, i, j, k0)) {   // padding threads report 0
        double r[RZ], p[RZ];
        double rv[RZ];
        rhs_column(d, rhs, i, j, k0, rv);
        const int n = residual_column<STRETCH, IBM>(g, mu, s, PlainPhi<double>{phi, d}, rv, i, j, k0, r, p);
#pragma unroll
        for (int u = 0; u < RZ; ++u)
            if (u < n) {
                if (field) field[d.in(i, j, k0 + u)] = r[u];
                m = nanmax_any(m, fabs(r[u]));
            }
    }
    block_max_part(m, out);
}

// max |rhs - L_mu(phi)| into red_max (and host_slot); with field != nullptr the residual itself is
// stored too
hipError_t launch_residual_field64(mdx_handle *h, const double *phi, double *field, double *host_slot) {
    GridK g = grid_k(h);
    MuK<double> mu = mu_k<double>(h, h->mu64);
    SorConst<double> s{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    dim3 blk(64, 4), grd(rz_grid(h->d));
#define MDX_RS(S, I) \
    hipLaunchKernelGGL((sor_residual64_kernel<S, I>), grd, blk, 0, h->stream, g, mu, s, h->rhs, phi, h->max_part, field)
    if (h->stretched) { if (h->ibm_on) MDX_RS(true, true); else MDX_RS(true, false); }
    else { if (h->ibm_on) MDX_RS(false, true); else MDX_RS(false, false); }
#undef MDX_RS
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_max_final(h, grd.x, host_slot, 0);
}

hipError_t launch_sor_residual64(mdx_handle *h, const double *phi) {
    return launch_residual_field64(h, phi, nullptr, nullptr);
}

// fac * Phi of one cell: the value the previous-Phi copy (scale_copy_kernel) stores
struct ScaledPhi {
    const double *p;
    double fac;
    Dims d;
    __device__ inline double operator()(int i, int j, int k) const { return fac * p[d.in(i, j, k)]; }
};

// The learned start's input pass in one sweep: Phi_0 = fac * prev into phi0 and its residual
// r0 = rhs - L_mu Phi_0 into res (the values the copy and the residual kernel store), and per
// workgroup the maximum |r0| (part[b]) and the sums of r0^2 (part[nb + b]) and Phi_0^2
// (part[2 nb + b]): the network's input statistics.
template <bool STRETCH, bool IBM>
__global__ __launch_bounds__(256) void phi0_residual_kernel(GridK g, MuK<double> mu, SorConst<double> s,
                                                            const double *__restrict__ rhs,
                                                            const double *__restrict__ prev, double fac,
                                                            double *__restrict__ phi0, double *__restrict__ res,
                                                            double *part) {
    __shared__ double red[3][4];
    const Dims d = g.d;
    int i = 0, j = 0, k0 = 0;
    double m = 0.0, sr = 0.0, sp = 0.0;
    if (rz_column(d, i, j, k0)) {   // padding threads report 0
        double r[RZ], p[RZ];
        double rv[RZ];
        rhs_column(d, rhs, i, j, k0, rv);
        const int n = residual_column<STRETCH, IBM>(g, mu, s, ScaledPhi{prev, fac, d}, rv, i, j, k0, r, p);
#pragma unroll
        for (int u = 0; u < RZ; ++u)
            if (u < n) {
                const size_t q = d.in(i, j, k0 + u);
                phi0[q] = p[u];
                res[q] = r[u];
                m = nanmax_any(m, fabs(r[u]));
                sr += r[u] * r[u];
                sp += p[u] * p[u];
            }
    }
    for (int off = 32; off > 0; off >>= 1) {
        m = nanmax_any(m, __shfl_xor(m, off, 64));
        sr += __shfl_xor(sr, off, 64);
        sp += __shfl_xor(sp, off, 64);
    }
    const int t = threadIdx.x + 64 * threadIdx.y;
    if ((t & 63) == 0) {
        red[0][t >> 6] = m;
        red[1][t >> 6] = sr;
        red[2][t >> 6] = sp;
    }
    __syncthreads();
    if (t == 0) {
        const unsigned nb = gridDim.x, b = blockIdx.x;
        part[b] = nanmax_any(nanmax_any(red[0][0], red[0][1]), nanmax_any(red[0][2], red[0][3]));
        part[nb + b] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
        part[2 * nb + b] = (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]);
    }
}

// the two sums of phi0_residual_kernel's nb workgroups (1024 threads, the loads of a thread
// issued together): stat[0] = rms(r0), stat[1] = rms(Phi_0) over n cells (1 where a field is zero)
__global__ __launch_bounds__(1024) void phi0_stats_kernel(const double *part, unsigned nb, double n, double *stat) {
    __shared__ double sa[16], sb[16];
    constexpr unsigned B = 8;
    double a = 0.0, b = 0.0;
    for (unsigned q0 = threadIdx.x; q0 < nb; q0 += 1024 * B) {
        double va[B], vb[B];
#pragma unroll
        for (unsigned u = 0; u < B; ++u) {
            const unsigned q = q0 + 1024 * u;
            va[u] = q < nb ? part[nb + q] : 0.0;
            vb[u] = q < nb ? part[2 * nb + q] : 0.0;
        }
#pragma unroll
        for (unsigned u = 0; u < B; ++u) {
            a += va[u];
            b += vb[u];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sa[threadIdx.x >> 6] = a;
        sb[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = 0.0;
        b = 0.0;
        for (int w = 0; w < 16; ++w) {
            a += sa[w];
            b += sb[w];
        }
        const double ra = sqrt(a / n), rb = sqrt(b / n);
        stat[0] = ra > 1e-300 ? ra : 1.0;
        stat[1] = rb > 1e-300 ? rb : 1.0;
    }
}

hipError_t launch_phi0_residual(mdx_handle *h, const double *prev, double fac, double *phi0, double *res,
                                double *host_slot, double *stat) {
    GridK g = grid_k(h);
    MuK<double> mu = mu_k<double>(h, h->mu64);
    SorConst<double> s{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    dim3 blk(64, 4), grd(rz_grid(h->d));
    if (3 * (size_t)grd.x > MAX_PARTS) return hipErrorInvalidValue;
#define MDX_P0(S, I)                                                                                              \
    hipLaunchKernelGGL((phi0_residual_kernel<S, I>), grd, blk, 0, h->stream, g, mu, s, h->rhs, prev, fac, phi0, res, \
                       h->max_part)
    if (h->stretched) { if (h->ibm_on) MDX_P0(true, true); else MDX_P0(true, false); }
    else { if (h->ibm_on) MDX_P0(false, true); else MDX_P0(false, false); }
#undef MDX_P0
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(phi0_stats_kernel, dim3(1), dim3(1024), 0, h->stream, h->max_part, grd.x,
                       (double)h->d.nx * h->d.ny * h->d.nz, stat);
    e = hipGetLastError();
    return e != hipSuccess ? e : launch_max_final(h, grd.x, host_slot, 0);
}

// The learned start's whole input pass in one sweep, for a network of stride S (2 or 4): rhs =
// div u* (sor_prepare's arithmetic, wall values of v), Phi_0 = fac * prev, r0 = rhs - L_mu Phi_0
// (cell_residual's operands), and the network's input before its scaling: the sums of r0 and of
// Phi_0 over each voxel's S^3 cells into vsum[2 v] and vsum[2 v + 1] (voxel v = vx + W (vy + H vz)).
// Per workgroup: part[b] = max |rhs|, part[nb + b] = max |r0|, part[2 nb + b] = sum of r0^2,
// part[3 nb + b] = sum of Phi_0^2.  phi0 and res, when given, receive Phi_0 and r0 (the start that
// falls back to Phi_0 copies them).  A workgroup's 64 x 4 x RZ cells are whole voxels (S divides
// 4 and the grid's extents); a voxel's sum adds its cells plane by plane in a column, then the S
// columns of a row (lane pairs), then the S rows, in a fixed order.
template <bool STRETCH, bool IBM, int S>
__global__ __launch_bounds__(256) void learned_input_kernel(GridK g, MuK<double> mu, SorConst<double> s,
                                                            double vbot, double vtop, const double *__restrict__ us,
                                                            const double *__restrict__ vs,
                                                            const double *__restrict__ ws, double idx, double idy,
                                                            double idz, const double *__restrict__ prev, double fac,
                                                            double *__restrict__ rhs, double *__restrict__ phi0,
                                                            double *__restrict__ res, double *__restrict__ vsum,
                                                            int W, int H, double *part) {
    static_assert(S == 2 || S == 4, "voxels of 2 or 4 cells");
    constexpr int ZG = RZ / S;   // voxel planes per column
    __shared__ double red[4][4];
    __shared__ double vr[4][64], vp[4][64];
    const Dims d = g.d;
    int i = 0, j = 0, k0 = 0;
    const bool ok = rz_column(d, i, j, k0);
    double r[RZ] = {}, p[RZ] = {};
    int n = 0;
    double mrhs = 0.0, mr = 0.0, sr = 0.0, sp = 0.0;
    if (ok) {
        const bool per2 = g.per[1];
        const int ip = i == d.nx ? 1 : i + 1;
        const int jp = (per2 && j == d.ny) ? 1 : j + 1;
        double rv[RZ];
#pragma unroll
        for (int u = 0; u < RZ; ++u) {
            rv[u] = 0.0;
            const int k = k0 + u;
            if (k <= d.nz) {
                const int kp = k == d.nz ? 1 : k + 1;
                const size_t c = d.g(i, j, k);
                const double v0 = (!per2 && j == 1) ? vbot : vs[c];
                const double v1 = (!per2 && j == d.ny) ? vtop : vs[d.g(i, jp, k)];
                const double dvy = STRETCH ? (v1 - v0) * g.ihc[j] : (v1 - v0) * idy;
                rv[u] = (us[d.g(ip, j, k)] - us[c]) * idx + dvy + (ws[d.g(i, j, kp)] - ws[c]) * idz;
            }
        }
        n = residual_column<STRETCH, IBM>(g, mu, s, ScaledPhi{prev, fac, d}, rv, i, j, k0, r, p);
#pragma unroll
        for (int u = 0; u < RZ; ++u)
            if (u >= n) p[u] = 0.0;   // past the last plane (residual_column leaves the wrapped neighbour)
#pragma unroll
        for (int u = 0; u < RZ; ++u)
            if (u < n) {
                const size_t q = d.in(i, j, k0 + u);
                rhs[q] = rv[u];
                if (phi0) phi0[q] = p[u];
                if (res) res[q] = r[u];
                mrhs = nanmax_any(mrhs, fabs(rv[u]));
                mr = nanmax_any(mr, fabs(r[u]));
                sr += r[u] * r[u];
                sp += p[u] * p[u];
            }
    }
    // ---- voxel sums (threads past the grid hold zeros)
    const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
    for (int gz = 0; gz < ZG; ++gz) {
        double cr = 0.0, cp = 0.0;
#pragma unroll
        for (int u = 0; u < S; ++u) {
            cr += r[gz * S + u];
            cp += p[gz * S + u];
        }
#pragma unroll
        for (int off = 1; off < S; off <<= 1) {
            cr += __shfl_xor(cr, off, 64);
            cp += __shfl_xor(cp, off, 64);
        }
        vr[ty][tx] = cr;
        vp[ty][tx] = cp;
        __syncthreads();
        if (ok && ty % S == 0 && tx % S == 0) {
            double a = 0.0, b = 0.0;
#pragma unroll
            for (int q = 0; q < S; ++q) {
                a += vr[ty + q][tx];
                b += vp[ty + q][tx];
            }
            const int vx = (i - 1) / S, vy = (j - 1) / S, vz = (k0 - 1) / S + gz;
            if (k0 + gz * S <= d.nz) {
                const size_t v = (size_t)vx + (size_t)W * ((size_t)vy + (size_t)H * vz);
                vsum[2 * v] = a;
                vsum[2 * v + 1] = b;
            }
        }
        __syncthreads();
    }
    // ---- workgroup partials
    for (int off = 32; off > 0; off >>= 1) {
        mrhs = nanmax_any(mrhs, __shfl_xor(mrhs, off, 64));
        mr = nanmax_any(mr, __shfl_xor(mr, off, 64));
        sr += __shfl_xor(sr, off, 64);
        sp += __shfl_xor(sp, off, 64);
    }
    const int t = tx + 64 * ty;
    if ((t & 63) == 0) {
        red[0][t >> 6] = mrhs;
        red[1][t >> 6] = mr;
        red[2][t >> 6] = sr;
        red[3][t >> 6] = sp;
    }
    __syncthreads();
    if (t == 0) {
        const unsigned nb = gridDim.x, b = blockIdx.x;
        part[b] = nanmax_any(nanmax_any(red[0][0], red[0][1]), nanmax_any(red[0][2], red[0][3]));
        part[nb + b] = nanmax_any(nanmax_any(red[1][0], red[1][1]), nanmax_any(red[1][2], red[1][3]));
        part[2 * nb + b] = (red[2][0] + red[2][1]) + (red[2][2] + red[2][3]);
        part[3 * nb + b] = (red[3][0] + red[3][1]) + (red[3][2] + red[3][3]);
    }
}

// whether learned_input_kernel serves this grid at stride s: voxels of whole workgroup tiles
bool learned_input_fits(const mdx_handle *h, int stride) {
    const Dims &d = h->d;
    return (stride == 2 || stride == 4) && RZ % stride == 0 && 4 % stride == 0 && d.nx % stride == 0 &&
           d.ny % stride == 0 && d.nz % stride == 0 && 4 * (size_t)rz_grid(d) <= MAX_PARTS;
}

hipError_t launch_learned_input(mdx_handle *h, const double *prev, double fac, double *phi0, double *res,
                                double *vsum, int stride, unsigned *nparts) {
    if (!learned_input_fits(h, stride)) return hipErrorInvalidValue;
    GridK g = grid_k(h);
    MuK<double> mu = mu_k<double>(h, h->mu64);
    SorConst<double> s{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    const unsigned nb = rz_grid(h->d);
    *nparts = nb;
    dim3 blk(64, 4), grd(nb);
    const double vb = h->cfg.bc_value[bci(2, 0, 2)], vt = h->cfg.bc_value[bci(2, 1, 2)];
    const int W = h->d.nx / stride, H = h->d.ny / stride;
#define MDX_LI(S, I, Q)                                                                                          \
    hipLaunchKernelGGL((learned_input_kernel<S, I, Q>), grd, blk, 0, h->stream, g, mu, s, vb, vt, h->us, h->vs, \
                       h->ws, 1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz, prev, fac, h->rhs, phi0, res, vsum, W, H,   \
                       h->max_part)
#define MDX_LIQ(S, I) do { if (stride == 4) MDX_LI(S, I, 4); else MDX_LI(S, I, 2); } while (0)
    if (h->stretched) { if (h->ibm_on) MDX_LIQ(true, true); else MDX_LIQ(true, false); }
    else { if (h->ibm_on) MDX_LIQ(false, true); else MDX_LIQ(false, false); }
#undef MDX_LIQ
#undef MDX_LI
    return hipGetLastError();
}

// The start of a mixed solve whose first check ran on the learned start (phi, its defect r32, and
// the nb workgroup maxima of that check in part): every workgroup folds the maxima into rg (the
// same value in each), and the start is kept when rg < margin * other (other_slot[0], the fallback
// start's residual maximum; a NaN keeps nothing), else replaced by the fallback: Phi = 0 with the
// defect rhs, or (PREV) Phi_0 = phi0 with the defect res0, each the value the first check forms for
// that start (r32 = (float) of the FP64 defect).  Workgroup 0 stores rg with the error word to
// rg_slot, and posts the chosen maximum as the check: to red_max[0], which a batch queued ahead
// reads, and with the error word and seq to host_slot.  The host repeats the comparison on the
// same pinned values to learn which start ran.
template <bool PREV>
__global__ __launch_bounds__(256) void select_start_kernel(size_t n, double margin, const double *__restrict__ part,
                                                           unsigned nb, double *rg_slot, const double *other_slot,
                                                           double *phi, float *r32,
                                                           const double *__restrict__ fb_defect,
                                                           const double *__restrict__ phi0,
                                                           unsigned long long *red_max, double *host_slot,
                                                           unsigned long long seq) {
    __shared__ double wred[4];
    const int t = threadIdx.x;
    constexpr unsigned B = 8;
    double m = 0.0;
    for (unsigned q0 = (unsigned)t; q0 < nb; q0 += 256 * B) {
        double v[B];
#pragma unroll
        for (unsigned u = 0; u < B; ++u) {
            const unsigned q = q0 + 256 * u;
            v[u] = q < nb ? part[q] : 0.0;
        }
#pragma unroll
        for (unsigned u = 0; u < B; ++u) m = nanmax_any(m, v[u]);
    }
    for (int off = 32; off > 0; off >>= 1) m = nanmax_any(m, __shfl_xor(m, off, 64));
    if ((t & 63) == 0) wred[t >> 6] = m;
    __syncthreads();
    const double rg = nanmax_any(nanmax_any(wred[0], wred[1]), nanmax_any(wred[2], wred[3]));
    const double other = other_slot[0];
    const bool take = rg < margin * other;
    if (!take) {
        for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
            phi[q] = PREV ? phi0[q] : 0.0;
            r32[q] = (float)fb_defect[q];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        const double v = take ? rg : other;
        const double ew = __longlong_as_double((long long)red_max[1]);
        rg_slot[0] = rg;
        rg_slot[1] = ew;
        red_max[0] = (unsigned long long)__double_as_longlong(v);
        host_slot[0] = v;
        host_slot[1] = ew;
        __threadfence_system();
        reinterpret_cast<volatile unsigned long long *>(host_slot)[2] = seq;
    }
}

hipError_t launch_select_start(mdx_handle *h, double margin, unsigned nb, double *rg_slot, const double *other_slot,
                               const double *phi0, const double *res0, double *host_slot, unsigned long long seq) {
    const size_t n = h->ni;
    // 256 workgroups: each folds the nb maxima (from the L2) and rewrites n / 256 cells on a fallback
    const unsigned blocks = 256;
    if (nb > MAX_PARTS) return hipErrorInvalidValue;
    if (phi0)
        hipLaunchKernelGGL(select_start_kernel<true>, dim3(blocks), dim3(256), 0, h->stream, n, margin, h->max_part,
                           nb, rg_slot, other_slot, h->phi, h->r32, res0, phi0, h->red_max, host_slot, seq);
    else
        hipLaunchKernelGGL(select_start_kernel<false>, dim3(blocks), dim3(256), 0, h->stream, n, margin, h->max_part,
                           nb, rg_slot, other_slot, h->phi, h->r32, h->rhs, nullptr, h->red_max, host_slot, seq);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- mixed precision
// One FP64 defect-correction step: Phi_out = Phi_in + e; r = rhs - L_mu(Phi_out) in FP64;
// r32 = (float) r for the next FP32 inner solve; max |r| for the stopping test.  Without e (E
// false: e = 0, the first check of a solve) Phi_out may be Phi_in itself: every read adds 0.0,
// so a neighbour read before or after its own cell's rewrite gives the same value.  STORE false:
// Phi_out is Phi_in and E is false, and Phi is left as it is (Phi_in + 0.0 differs from Phi_in only
// in the sign of a zero, which no sum with another operand keeps).
template <bool STRETCH, bool IBM, bool E, bool STORE>
__global__ __launch_bounds__(256) void defect64_kernel(GridK g, MuK<double> mu, SorConst<double> s,
                                                       const double *__restrict__ rhs, const double *phi_in,
                                                       const float *__restrict__ e, double *phi_out,
                                                       float *r32, double *out) {
    const Dims d = g.d;
    int i = 0, j = 0, k0 = 0;
    double m = 0.0;
    if (rz_column(d, i, j, k0)) {   // padding threads report 0
        double r[RZ], p[RZ];
        double rv[RZ];
        rhs_column(d, rhs, i, j, k0, rv);
        const int n = residual_column<STRETCH, IBM>(g, mu, s, SumPhi<E>{phi_in, e, d}, rv, i, j, k0, r, p);
#pragma unroll
        for (int u = 0; u < RZ; ++u)
            if (u < n) {
                const size_t q = d.in(i, j, k0 + u);
                if (STORE) phi_out[q] = p[u];
                r32[q] = (float)r[u];
                m = nanmax_any(m, fabs(r[u]));
            }
    }
    block_max_part(m, out);
}

hipError_t launch_defect64(mdx_handle *h, const double *phi_in, const float *e, double *phi_out, float *r32,
                           double *host_slot, unsigned long long seq, unsigned *parts_only) {
    GridK g = grid_k(h);
    MuK<double> mu = mu_k<double>(h, h->mu64);
    SorConst<double> s{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    dim3 blk(64, 4), grd(rz_grid(h->d));
#define MDX_DF(S, I, E, W)                                                                                       \
    hipLaunchKernelGGL((defect64_kernel<S, I, E, W>), grd, blk, 0, h->stream, g, mu, s, h->rhs, phi_in, e, phi_out, \
                       r32, h->max_part)
#define MDX_DFE(S, I)                                                    \
    do {                                                                 \
        if (e) MDX_DF(S, I, true, true);                                 \
        else if (phi_out == phi_in) MDX_DF(S, I, false, false);          \
        else MDX_DF(S, I, false, true);                                  \
    } while (0)
    if (h->stretched) { if (h->ibm_on) MDX_DFE(true, true); else MDX_DFE(true, false); }
    else { if (h->ibm_on) MDX_DFE(false, true); else MDX_DFE(false, false); }
#undef MDX_DFE
#undef MDX_DF
    hipError_t err = hipGetLastError();
    if (parts_only) {   // the caller folds the grd.x workgroup maxima itself
        *parts_only = grd.x;
        return err;
    }
    return err != hipSuccess ? err : launch_max_final(h, grd.x, host_slot, seq);
}

template <bool STRETCH, bool IBM>
__global__ __launch_bounds__(SX * SY) void sor_sweep32_kernel(GridK g, MuK<float> mu, SorConst<float> s,
                                                              int color, float omega,
                                                              const float *__restrict__ rhs, float *e) {
    const Dims d = g.d;
    int tx_, ty_, tz_;
    if (!xcd_tile(d.nx / 2, d.ny, d.nz, SX, SY, tx_, ty_, tz_)) return;
    const int ni = tx_ * SX + threadIdx.x + 1;
    const int j = ty_ * SY + threadIdx.y + 1;
    const int k = tz_ + 1;
    if (ni > d.nx / 2 || j > d.ny) return;
    const int i = 1 + ((j + k + color) & 1) + 2 * (ni - 1);
    PlainPhi<float> P{e, d};
    float denom;
    const float r = cell_residual<float, STRETCH, IBM>(g, mu, s, P, rhs[d.in(i, j, k)], i, j, k, denom);
    const size_t q = d.in(i, j, k);
    e[q] = e[q] + r * (-omega / denom);
}

hipError_t launch_sor_sweep32(mdx_handle *h, int color, float omega) {
    GridK g = grid_k(h);
    MuK<float> mu = mu_k<float>(h, h->mu32);
    SorConst<float> s{(float)(1.0 / (h->dx * h->dx)), (float)(1.0 / (h->dy * h->dy)), (float)(1.0 / (h->dz * h->dz))};
    if (h->sweep_kz > 0) return launch_sweep_zm<float>(h, color, omega, mu, s, h->r32, h->e32);
    dim3 blk(SX, SY), grd(xcd_grid(h->d.nx / 2, h->d.ny, h->d.nz, SX, SY));
#define MDX_SW(S, I) \
    hipLaunchKernelGGL((sor_sweep32_kernel<S, I>), grd, blk, 0, h->stream, g, mu, s, color, omega, h->r32, h->e32)
    if (h->stretched) { if (h->ibm_on) MDX_SW(true, true); else MDX_SW(true, false); }
    else { if (h->ibm_on) MDX_SW(false, true); else MDX_SW(false, false); }
#undef MDX_SW
    return hipGetLastError();
}

// m full red+black iterations, one launch each (sor_iter_fused_kernel), alternating between the
// iterate and a second buffer; an odd m ends with a copy back, so the result is in the iterate
hipError_t launch_sor_iters64(mdx_handle *h, int m, double omega) {
    MuK<double> mu = mu_k<double>(h, h->mu64);
    SorConst<double> s{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    hipError_t e = hipSuccess;
    for (int q = 0; q < m && e == hipSuccess; ++q)
        e = (q & 1) ? launch_iter_fused<double>(h, omega, mu, s, h->rhs, h->phi2, h->phi)
                    : launch_iter_fused<double>(h, omega, mu, s, h->rhs, h->phi, h->phi2);
    if (e == hipSuccess && (m & 1)) {
        hipLaunchKernelGGL(copy_kernel<double>, dim3(2048), dim3(256), 0, h->stream, h->phi, h->phi2, h->ni);
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_sor_iters32(mdx_handle *h, int m, float omega) {
    MuK<float> mu = mu_k<float>(h, h->mu32);
    SorConst<float> s{(float)(1.0 / (h->dx * h->dx)), (float)(1.0 / (h->dy * h->dy)), (float)(1.0 / (h->dz * h->dz))};
    hipError_t e = hipSuccess;
    for (int q = 0; q < m && e == hipSuccess; ++q)
        e = (q & 1) ? launch_iter_fused<float>(h, omega, mu, s, h->r32, h->e32b, h->e32)
                    : launch_iter_fused<float>(h, omega, mu, s, h->r32, h->e32, h->e32b);
    if (e == hipSuccess && (m & 1)) {
        hipLaunchKernelGGL(copy_kernel<float>, dim3(2048), dim3(256), 0, h->stream, h->e32, h->e32b, h->ni);
        e = hipGetLastError();
    }
    return e;
}

template <typename T>
__global__ void zero_kernel(T *p, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q < n; q += (size_t)gridDim.x * blockDim.x) p[q] = T(0);
}

__global__ void scale_copy_kernel(double *__restrict__ dst, const double *__restrict__ src, double fac, size_t n) {
    size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; q < n; q += (size_t)gridDim.x * blockDim.x) dst[q] = fac * src[q];
}

// dst = fac * src over n values
hipError_t launch_scale_copy64(mdx_handle *h, double *dst, const double *src, double fac, size_t n) {
    hipLaunchKernelGGL(scale_copy_kernel, dim3(2048), dim3(256), 0, h->stream, dst, src, fac, n);
    return hipGetLastError();
}

hipError_t launch_zero32(mdx_handle *h, float *p, size_t n) {
    hipLaunchKernelGGL(zero_kernel<float>, dim3(2048), dim3(256), 0, h->stream, p, n);
    return hipGetLastError();
}
hipError_t launch_zero64(mdx_handle *h, double *p, size_t n) {
    hipLaunchKernelGGL(zero_kernel<double>, dim3(2048), dim3(256), 0, h->stream, p, n);
    return hipGetLastError();
}

}  // namespace mdx
