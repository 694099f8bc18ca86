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

}  // namespace mdx
