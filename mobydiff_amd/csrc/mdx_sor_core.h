// mdx_sor_core.h -- the per-cell arithmetic of the red-black SOR in pressure-correction form,
// shared by the streaming sweeps (mdx_sor.hip) and the LDS-resident sweeps (mdx_sor_lds.hip) so
// both evaluate exactly the same FP expressions.
#pragma once

#include "mdx_internal.h"

namespace mdx {

template <typename T>
struct SorConst {
    T idx2, idy2, idz2;
};

// residual r = rhs - L_mu(Phi) of one cell from its 7 Phi values and 6 face coefficients, and the
// reference's relaxation denominator (wall faces included, pressure_redblack.f90:124-126).
// wall_n / wall_s: the north / south face is a no-flux wall (its Phi neighbour is not read).
template <typename T, bool STRETCH>
__device__ inline T sor_residual_core(T p0, T pxp, T pxm, T pyp, T pym, T pzp, T pzm, T mui, T muip, T mvj,
                                      T mvjp, T mwk, T mwkp, bool wall_n, bool wall_s, T rhs,
                                      const SorConst<T> &s, T an, T as, T &denom) {
    const T sn = wall_n ? T(0) : mvjp * (pyp - p0);
    const T ss = wall_s ? T(0) : mvj * (p0 - pym);
    const T lx = (muip * (pxp - p0) - mui * (p0 - pxm)) * s.idx2;
    const T lz = (mwkp * (pzp - p0) - mwk * (p0 - pzm)) * s.idz2;
    T ly;
    if (!STRETCH) {
        ly = (sn - ss) * s.idy2;
        denom = (mui + muip) * s.idx2 + (mwk + mwkp) * s.idz2 + (mvj + mvjp) * s.idy2;
    } else {
        ly = sn * an - ss * as;
        denom = (mui + muip) * s.idx2 + (mwk + mwkp) * s.idz2 + (mvj * as + mvjp * an);
    }
    return rhs - (lx + ly + lz);
}

// residual r = rhs - L_mu(Phi) of cell (i,j,k) and the reference's relaxation denominator
template <typename T, bool STRETCH, bool IBM, typename PhiAt>
__device__ inline T cell_residual(const GridK &g, const MuK<T> &mu, const SorConst<T> &s, PhiAt P,
                                  T rhs_c, int i, int j, int k, T &denom) {
    const Dims d = g.d;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const bool per2 = g.per[1];
    const int ip = i == nx ? 1 : i + 1, im = i == 1 ? nx : i - 1;
    const int kp = k == nz ? 1 : k + 1, km = k == 1 ? nz : k - 1;
    int jp = j + 1, jm = j - 1;
    if (per2) {
        if (jp > ny) jp = 1;
        if (jm < 1) jm = ny;
    }
    T mui = 1, muip = 1, mvj = 1, mvjp = 1, mwk = 1, mwkp = 1;
    if (IBM) {
        mui = mu.u[mu.at(i, j, k)];
        muip = mu.u[mu.at(ip, j, k)];
        mvj = mu.v[mu.at(i, j, k)];
        mvjp = mu.v[mu.at(i, jp, k)];
        mwk = mu.w[mu.at(i, j, k)];
        mwkp = mu.w[mu.at(i, j, kp)];
    }
    const bool wn_ = !per2 && j == ny, ws_ = !per2 && j == 1;
    const T p0 = P(i, j, k);
    const T pyp = P(i, wn_ ? j : jp, k), pym = P(i, ws_ ? j : jm, k);
    T an = 0, as = 0;
    if (STRETCH) {
        an = (T)(g.ihv[j + 1] * g.ihc[j]);
        as = (T)(g.ihv[j] * g.ihc[j]);
    }
    return sor_residual_core<T, STRETCH>(p0, P(ip, j, k), P(im, j, k), pyp, pym, P(i, j, kp), P(i, j, km), mui,
                                         muip, mvj, mvjp, mwk, mwkp, wn_, ws_, rhs_c, s, an, as, denom);
}

}  // namespace mdx
