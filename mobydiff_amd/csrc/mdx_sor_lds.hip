// mdx_sor_lds.hip -- LDS-resident red-black SOR for gfx950 (multi-sweep temporal blocking).
//
// The streaming sweeps (mdx_sor.hip) read and write the whole grid once per colour; on the wavy
// channel (2.4 M cells) that is a few microseconds of HBM/MALL traffic but each launch also pays
// dispatch latency, so an iteration costs ~35 us.  Here every workgroup owns one subdomain for a
// whole batch of iterations: its Phi tile (with a one-cell halo) lives in LDS, its right-hand side
// in registers, and its face coefficients in an LDS table (the IBM geometry is z-invariant, so
// one x-y plane of mu suffices).  After each half-sweep a workgroup publishes its six boundary
// layers to global memory, raises a per-workgroup counter, waits for its (at most six)
// neighbours' counters and pulls their layers into its halo.  All workgroups must be co-resident:
// the launch is cooperative (the runtime refuses a grid that does not fit) and every wait is
// bounded (a spin limit sets an error flag and the kernel runs to completion).
//
// The arithmetic per cell is sor_residual_core(), the same expressions the streaming sweeps use,
// and the relaxation factor -omega/denom is formed from the same values, so the iterates are
// bit-identical to the streaming sweeps (and to the oracle).
#include "mdx_sor_core.h"

namespace mdx {

static constexpr int RB_THREADS = 512;

template <typename T>
struct RbArgs {
    Dims d;
    int px, py, pz;          // workgroups along x, y, z
    int bx, by, bz;          // subdomain size (bx even)
    int per_y;
    SorConst<T> s;
    T omega;
    const T *mu_u, *mu_v, *mu_w;   // z-invariant plane tables (IBM) or nullptr
    T *x;                    // solution (interior layout), updated in place
    const T *rhs;
    int zero_init;           // x starts at zero (the defect equation)
    int iters;               // full red+black iterations
    T *faces;                // [wg][face][buffer][fmax]
    int fmax;
    unsigned long long *flags;   // per-workgroup progress counters
    unsigned long long epoch;    // counters of this launch are epoch*2^32 + half-sweeps done
    int *err;
};

template <typename T, bool IBM, int MAXP>
__global__ __launch_bounds__(RB_THREADS) void rb_lds_kernel(RbArgs<T> a) {
    extern __shared__ unsigned char smem_raw[];
    const Dims d = a.d;
    const int bx = a.bx, by = a.by, bz = a.bz;
    const int tx = bx + 2, ty = by + 2;
    T *X = reinterpret_cast<T *>(smem_raw);
    T *C = X + (size_t)tx * ty * (bz + 2);          // 6 values per plane cell when IBM
    const int nwg = a.px * a.py * a.pz;
    // XCD-aware placement: workgroups are dispatched round-robin over the 8 XCDs, so consecutive
    // subdomain ids (neighbours along z first) are given to workgroups on the same XCD.
    const int per_xcd = nwg / 8;
    const int b = blockIdx.x;
    const int w = (nwg % 8 == 0) ? (b % 8) * per_xcd + b / 8 : b;
    const int wx = w % a.px, wy = (w / a.px) % a.py, wz = w / (a.px * a.py);
    const int ox = wx * bx, oy = wy * by, oz = wz * bz;
    const int t = threadIdx.x;
    auto L = [&](int i, int j, int k) { return i + tx * (j + ty * k); };

    // neighbours: 0 x-, 1 x+, 2 y-, 3 y+, 4 z-, 5 z+ (periodic in x and z; y periodic or walls)
    auto wid = [&](int ix, int iy, int iz) { return ix + a.px * (iy + a.py * iz); };
    int nb[6];
    nb[0] = wid((wx + a.px - 1) % a.px, wy, wz);
    nb[1] = wid((wx + 1) % a.px, wy, wz);
    nb[2] = (wy > 0 || a.per_y) ? wid(wx, (wy + a.py - 1) % a.py, wz) : -1;
    nb[3] = (wy < a.py - 1 || a.per_y) ? wid(wx, (wy + 1) % a.py, wz) : -1;
    nb[4] = wid(wx, wy, (wz + a.pz - 1) % a.pz);
    nb[5] = wid(wx, wy, (wz + 1) % a.pz);

    // ---- load the tile (interior and halo) from the global solution
    const int tiles = tx * ty * (bz + 2);
    for (int q = t; q < tiles; q += RB_THREADS) {
        const int i = q % tx, j = (q / tx) % ty, k = q / (tx * ty);
        T v = T(0);
        if (!a.zero_init) {
            int gi = ox + i, gj = oy + j, gk = oz + k;
            gi = gi < 1 ? d.nx : (gi > d.nx ? 1 : gi);
            gk = gk < 1 ? d.nz : (gk > d.nz ? 1 : gk);
            if (a.per_y) gj = gj < 1 ? d.ny : (gj > d.ny ? 1 : gj);
            if (gj >= 1 && gj <= d.ny) v = a.x[d.in(gi, gj, gk)];
        }
        X[q] = v;
    }
    // ---- face coefficient table (one x-y plane): mui, muip, mvj, mvjp, mw, -omega/denom
    T qconst = T(0);
    if (IBM) {
        const int sx = d.sx;
        for (int q = t; q < bx * by; q += RB_THREADS) {
            const int i = q % bx + 1, j = q / bx + 1;
            const int gi = ox + i, gj = oy + j;
            const int gip = gi == d.nx ? 1 : gi + 1;
            const int gjp = (a.per_y && gj == d.ny) ? 1 : gj + 1;
            const T mui = a.mu_u[(size_t)gi + (size_t)sx * gj], muip = a.mu_u[(size_t)gip + (size_t)sx * gj];
            const T mvj = a.mu_v[(size_t)gi + (size_t)sx * gj], mvjp = a.mu_v[(size_t)gi + (size_t)sx * gjp];
            const T mw = a.mu_w[(size_t)gi + (size_t)sx * gj];
            const T denom = (mui + muip) * a.s.idx2 + (mw + mw) * a.s.idz2 + (mvj + mvjp) * a.s.idy2;
            T *c = C + 6 * q;
            c[0] = mui; c[1] = muip; c[2] = mvj; c[3] = mvjp; c[4] = mw; c[5] = -a.omega / denom;
        }
    } else {
        const T one = T(1);
        const T denom = (one + one) * a.s.idx2 + (one + one) * a.s.idz2 + (one + one) * a.s.idy2;
        qconst = -a.omega / denom;
    }
    // the right-hand side is read from global memory at each update: a subdomain's share
    // (bx*by*bz values) stays resident in its XCD's L2 for the whole batch
    const int hbx = bx / 2, npairs = hbx * by * bz;
    const T *__restrict__ rhs_g = a.rhs;
    int my_nb = -1;                                  // the neighbour this thread waits for (t < 6)
    if (t < 6) my_nb = nb[0] * (t == 0) + nb[1] * (t == 1) + nb[2] * (t == 2) + nb[3] * (t == 3) +
                       nb[4] * (t == 4) + nb[5] * (t == 5);
    __syncthreads();

    const int fmax = a.fmax;
    const unsigned long long base = a.epoch << 32;
    const int nhalf = 2 * a.iters;
    for (int s = 0; s < nhalf; ++s) {
        const int color = s & 1;
        // ---- relax the cells of this colour: (gi + gj + gk + color) odd
#pragma unroll
        for (int m = 0; m < MAXP; ++m) {
            const int p = t + m * RB_THREADS;
            if (p < npairs) {
                const int i0 = 2 * (p % hbx) + 1, j = (p / hbx) % by + 1, k = p / (hbx * by) + 1;
                const int gj = oy + j;
                const int odd = (ox + i0 + gj + oz + k + color) & 1;
                const int i = odd ? i0 : i0 + 1;
                const T rhs = rhs_g[d.in(ox + i, gj, oz + k)];
                const bool wall_n = !a.per_y && gj == d.ny, wall_s = !a.per_y && gj == 1;
                const int c = L(i, j, k);
                T mui = 1, muip = 1, mvj = 1, mvjp = 1, mw = 1, q = qconst;
                if (IBM) {
                    const T *cc = C + 6 * ((i - 1) + bx * (j - 1));
                    mui = cc[0]; muip = cc[1]; mvj = cc[2]; mvjp = cc[3]; mw = cc[4]; q = cc[5];
                }
                T denom;
                const T p0 = X[c];
                const T r = sor_residual_core<T, false>(p0, X[c + 1], X[c - 1], X[c + tx], X[c - tx],
                                                        X[c + tx * ty], X[c - tx * ty], mui, muip, mvj, mvjp,
                                                        mw, mw, wall_n, wall_s, rhs, a.s, T(0), T(0), denom);
                X[c] = p0 + r * q;
            }
        }
        __syncthreads();
        if (s == nhalf - 1) break;

        // ---- publish the six boundary layers of this subdomain.  Face values and counters are
        // agent-scope atomic stores and loads: they are coherent across XCDs by themselves, so the
        // hand-off needs no L2 write-back or invalidate, only the completion of this workgroup's
        // stores (s_waitcnt vmcnt(0) in every wave, then the barrier) before the counter is raised.
        const int buf = s & 1;
        T *myf = a.faces + (size_t)w * 12 * fmax;
        for (int q = t; q < 2 * by * bz; q += RB_THREADS) {          // x faces (j,k)
            const int side = q / (by * bz), r = q % (by * bz), j = r % by + 1, k = r / by + 1;
            __hip_atomic_store(&myf[(size_t)((side) * 2 + buf) * fmax + r], X[L(side ? bx : 1, j, k)],
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int q = t; q < 2 * bx * bz; q += RB_THREADS) {          // y faces (i,k)
            const int side = q / (bx * bz), r = q % (bx * bz), i = r % bx + 1, k = r / bx + 1;
            __hip_atomic_store(&myf[(size_t)((2 + side) * 2 + buf) * fmax + r], X[L(i, side ? by : 1, k)],
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int q = t; q < 2 * bx * by; q += RB_THREADS) {          // z faces (i,j)
            const int side = q / (bx * by), r = q % (bx * by), i = r % bx + 1, j = r / bx + 1;
            __hip_atomic_store(&myf[(size_t)((4 + side) * 2 + buf) * fmax + r], X[L(i, j, side ? bz : 1)],
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const unsigned long long target = base + (unsigned long long)(s + 1);
        if (t == 0) __hip_atomic_store(&a.flags[w], target, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (t < 6 && my_nb >= 0) {
            long spins = 0;
            while (__hip_atomic_load(&a.flags[my_nb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++spins > (1l << 24)) {
                    __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        // ---- pull the neighbours' layers into the halo
        for (int q = t; q < 2 * by * bz; q += RB_THREADS) {
            const int side = q / (by * bz), r = q % (by * bz), j = r % by + 1, k = r / by + 1;
            const int n = nb[side];                      // side 0: x- neighbour gives its x+ face
            T *nf = a.faces + (size_t)n * 12 * fmax + (size_t)((1 - side) * 2 + buf) * fmax;
            X[L(side ? bx + 1 : 0, j, k)] = __hip_atomic_load(&nf[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        for (int q = t; q < 2 * bx * bz; q += RB_THREADS) {
            const int side = q / (bx * bz), r = q % (bx * bz), i = r % bx + 1, k = r / bx + 1;
            const int n = nb[2 + side];
            if (n >= 0) {
                T *nf = a.faces + (size_t)n * 12 * fmax + (size_t)((2 + 1 - side) * 2 + buf) * fmax;
                X[L(i, side ? by + 1 : 0, k)] = __hip_atomic_load(&nf[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        for (int q = t; q < 2 * bx * by; q += RB_THREADS) {
            const int side = q / (bx * by), r = q % (bx * by), i = r % bx + 1, j = r / bx + 1;
            const int n = nb[4 + side];
            T *nf = a.faces + (size_t)n * 12 * fmax + (size_t)((4 + 1 - side) * 2 + buf) * fmax;
            X[L(i, j, side ? bz + 1 : 0)] = __hip_atomic_load(&nf[r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __syncthreads();
    }
    // ---- write the interior back
    for (int q = t; q < bx * by * bz; q += RB_THREADS) {
        const int i = q % bx + 1, j = (q / bx) % by + 1, k = q / (bx * by) + 1;
        a.x[d.in(ox + i, oy + j, oz + k)] = X[L(i, j, k)];
    }
}

// ------------------------------------------------------------------------------- host side
static size_t rb_lds_bytes(int bx, int by, int bz, size_t tsize, bool ibm) {
    return ((size_t)(bx + 2) * (by + 2) * (bz + 2) + (ibm ? 6 * (size_t)bx * by : 0)) * tsize;
}

template <typename T, bool IBM, int MAXP>
static const void *rb_fn() {
    return reinterpret_cast<const void *>(&rb_lds_kernel<T, IBM, MAXP>);
}

template <typename T, bool IBM>
static const void *rb_fn_for(int maxp) {
    if (maxp <= 4) return rb_fn<T, IBM, 4>();
    if (maxp <= 8) return rb_fn<T, IBM, 8>();
    return rb_fn<T, IBM, 16>();
}

static int rb_maxp_class(int need) { return need <= 4 ? 4 : (need <= 8 ? 8 : (need <= 16 ? 16 : 0)); }

// workgroups of this kernel variant that fit on one CU (registers, LDS, waves), from the runtime
static int rb_occupancy(bool fp32, bool ibm, int maxp, size_t lds) {
    const void *fn = fp32 ? (ibm ? rb_fn_for<float, true>(maxp) : rb_fn_for<float, false>(maxp))
                          : (ibm ? rb_fn_for<double, true>(maxp) : rb_fn_for<double, false>(maxp));
    if (lds > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, RB_THREADS, lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

// Choose the decomposition: every workgroup co-resident (runtime occupancy), the fewest cells
// per CU, then the smallest halo surface.  `why` receives the reason when nothing fits.
bool rb_plan(const mdx_handle *h, bool fp32, RbPlanOut &out, std::string &why) {
    const int nx = h->d.nx, ny = h->d.ny, nz = h->d.nz;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) { why = "no device properties"; return false; }
    if (!prop.cooperativeLaunch) { why = "device reports no cooperative launch"; return false; }
    const int cus = prop.multiProcessorCount;
    const size_t lds_max = prop.maxSharedMemoryPerMultiProcessor > prop.sharedMemPerBlock
                               ? prop.maxSharedMemoryPerMultiProcessor : prop.sharedMemPerBlock;
    const size_t ts = fp32 ? 4 : 8;
    bool found = false;
    double best_cells = 1e300, best_surf = 1e300;
    why = "no decomposition fits (cus=" + std::to_string(cus) + ", lds/block=" + std::to_string(prop.sharedMemPerBlock) +
          ", lds/cu=" + std::to_string(prop.maxSharedMemoryPerMultiProcessor) + ")";
    for (int px = 1; px <= nx; ++px) {
        if (nx % px || (nx / px) % 2) continue;
        for (int py = 1; py <= ny; ++py) {
            if (ny % py) continue;
            for (int pz = 1; pz <= nz; ++pz) {
                if (nz % pz) continue;
                const int bx = nx / px, by = ny / py, bz = nz / pz;
                const int nwg = px * py * pz;
                const size_t lds = rb_lds_bytes(bx, by, bz, ts, h->ibm_on);
                if (lds > lds_max) continue;
                const int maxp = rb_maxp_class((bx / 2 * by * bz + RB_THREADS - 1) / RB_THREADS);
                if (!maxp) continue;
                if (nwg > 4 * cus) continue;
                const int per_cu = rb_occupancy(fp32, h->ibm_on, maxp, lds);
                if (per_cu < 1 || nwg > per_cu * cus) continue;
                const double rounds = std::ceil((double)nwg / cus);
                const double cells = rounds * bx * by * bz;
                const double surf = 2.0 * ((double)by * bz + (double)bx * bz + (double)bx * by);
                if (cells < best_cells - 0.5 || (cells < best_cells + 0.5 && surf < best_surf)) {
                    best_cells = cells;
                    best_surf = surf;
                    out = RbPlanOut{px, py, pz, bx, by, bz, maxp, lds};
                    found = true;
                }
            }
        }
    }
    return found;
}

template <typename T>
static hipError_t rb_launch_t(mdx_handle *h, const RbPlanOut &pl, T *x, const T *rhs, int zero_init, int iters,
                              T *const *mu) {
    RbArgs<T> a;
    a.d = h->d;
    a.px = pl.px; a.py = pl.py; a.pz = pl.pz;
    a.bx = pl.bx; a.by = pl.by; a.bz = pl.bz;
    a.per_y = h->cfg.periodic[1] ? 1 : 0;
    const double ix2 = 1.0 / h->dx * (1.0 / h->dx), iy2 = 1.0 / h->dy * (1.0 / h->dy), iz2 = 1.0 / h->dz * (1.0 / h->dz);
    if (sizeof(T) == 4)
        a.s = SorConst<T>{(T)(1.0 / (h->dx * h->dx)), (T)(1.0 / (h->dy * h->dy)), (T)(1.0 / (h->dz * h->dz))};
    else
        a.s = SorConst<T>{(T)ix2, (T)iy2, (T)iz2};
    a.omega = (T)h->cfg.omega;
    a.mu_u = h->ibm_on ? mu[0] : nullptr;
    a.mu_v = h->ibm_on ? mu[1] : nullptr;
    a.mu_w = h->ibm_on ? mu[2] : nullptr;
    a.x = x;
    a.rhs = rhs;
    a.zero_init = zero_init;
    a.iters = iters;
    a.faces = reinterpret_cast<T *>(h->rb_faces);
    a.fmax = std::max(pl.by * pl.bz, std::max(pl.bx * pl.bz, pl.bx * pl.by));
    a.flags = h->rb_flags;
    a.epoch = ++h->rb_epoch;
    a.err = h->rb_err;
    const void *fn = h->ibm_on ? rb_fn_for<T, true>(pl.maxp) : rb_fn_for<T, false>(pl.maxp);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds);
    if (e != hipSuccess) return e;
    void *args[] = {&a};
    return hipLaunchCooperativeKernel(fn, dim3(pl.px * pl.py * pl.pz), dim3(RB_THREADS), args,
                                      (unsigned)pl.lds, h->stream);
}

hipError_t launch_rb_lds(mdx_handle *h, bool fp32, int iters, int zero_init) {
    const RbPlanOut &pl = fp32 ? h->rb_plan32 : h->rb_plan64;
    if (fp32) return rb_launch_t<float>(h, pl, h->e32, h->r32, zero_init, iters, h->mu32);
    return rb_launch_t<double>(h, pl, h->phi, h->rhs, zero_init, iters, h->mu64);
}

}  // namespace mdx
