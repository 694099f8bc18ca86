// mdx_sor_lds.hip -- LDS-resident red-black SOR for gfx950 with multi-sweep temporal blocking.
//
// The streaming sweeps (mdx_sor.hip) read and write the whole grid once per colour; on the wavy
// channel (2.4 M cells) that is a few microseconds of HBM/MALL traffic, but every launch also pays
// dispatch latency.  Here every workgroup owns one subdomain for a whole batch of iterations: its
// Phi tile lives in LDS with a halo two cells deep, and its right-hand side and face coefficients
// (the IBM geometry is z-invariant, so one x-y plane of mu suffices) in registers of the threads
// that own its z-columns (the pair form keeps the coefficients in an LDS table).
//
// Temporal blocking: an iteration relaxes colour 0 on the subdomain AND on the one-cell ring of
// face cells around it (recomputing what the neighbours compute for those cells, from the same old
// values, so the results are identical), then colour 1 on the subdomain, whose neighbours are then
// all up to date.  One halo exchange per iteration therefore suffices instead of one per colour.
// The ring cells read the second halo layer and the halo edge lines, so a workgroup exchanges with
// up to 18 neighbours: 6 faces two layers deep and 12 edges one cell thick.  Only the colour-1
// cells of those regions travel: the ring's colour-0 cells are the only colour-0 halo cells any
// relaxation reads, and this workgroup forms them itself (the recomputation above), so half of the
// halo is exchanged.
//
// Exchange: every published word carries the number of the iteration it belongs to (a 32-bit
// payload and a 32-bit tag in one 64-bit agent-scope atomic store), and a consumer polls the face
// data itself until every word it needs shows the expected tag: one store and one load per
// hand-off, with no acknowledgement wait and no separate counter.  Faces are double-buffered by
// iteration parity; a producer cannot overwrite a buffer before its consumers have read it,
// because it first needs their next layers.  All workgroups must be co-resident: the plan never
// holds more workgroups than the CUs take at once (runtime occupancy), and every wait is bounded
// (a spin limit sets an error flag and the kernel runs to completion).
//
// The arithmetic per cell is sor_residual_core(), the same expressions the streaming sweeps use,
// and the relaxation factor -omega/denom is formed from the same values, so the iterates are
// bit-identical to the streaming sweeps (and to the oracle).
#include "mdx_sor_core.h"

namespace mdx {

static constexpr int RB_THREADS = 512;
static constexpr int RB_HE = 6;      // exchanged values per thread (a subdomain's colour-1 halo <= RB_HE * RB_THREADS)
static constexpr int RB_MAXR = 3;    // colour-0 ring cells per thread
static constexpr int RB_G = 3;       // cells relaxed per group (loads issued together)
static constexpr int RB_H = 2;       // halo depth
// column forms: cells per thread and colour, for a whole column (bz <= 2 * RB_ZC1 - 2) and for a half
// column (bz <= 4 * RB_ZC2 - 4)
static constexpr int RB_ZC1 = 13, RB_ZC2 = 7;
// plan codes (in place of the pairs per thread) of the column forms
static constexpr int RB_COLS = -1, RB_COLS2 = -2;
static constexpr int RB_XSTRIDE = RB_HE * RB_THREADS;   // exchange entries per workgroup and buffer
// Fixed LDS geometry: a tile row holds RB_TXP values and a plane RB_TYP rows, whatever the
// subdomain (bx + 4 <= RB_TXP, by + 4 <= RB_TYP), so every stencil and coefficient-plane offset is
// an immediate of the LDS instruction; the coefficient table uses the same plane shape.  The box
// is the 24 x 16 tile of the 192 x 128 x 96 channel on 256 workgroups with its halo; its 28-value
// rows also spread a wave's cells (one colour: every other value) well over the 32 banks (about
// 1.7-way on average, against 2 for rows of 32 and 3 for rows of 33).
static constexpr int RB_TXP = 28, RB_TYP = 20, RB_PLANE = RB_TXP * RB_TYP;

template <typename T>
struct RbArgs {
    Dims d;
    int px, py, pz;          // workgroups along x, y, z
    int bx, by, bz;          // subdomain size (all even)
    int per_y;
    SorConst<T> s;
    T omega;
    const T *mu_u, *mu_v, *mu_w;   // z-invariant plane tables (IBM) or nullptr
    T *x;                    // solution (interior layout), updated in place
    const T *rhs;
    int zero_init;           // x starts at zero (the defect equation)
    int iters;               // full red+black iterations
    int md;                  // leading pair entries of every thread that lie in the deep box
    unsigned long long *faces;   // [wg][buffer][exchange entry][words per value], tagged words
    int total;               // exchange entries in use per workgroup (faces two deep, then edges)
    const int *xl;           // exchange lists of the plan: [wg][entry] publish words, then pull words
    unsigned tag0;           // iteration s of this launch publishes tag tag0 + s + 1 (never 0)
    int *err;
    int spin_limit;          // polls of a neighbour wait before it gives up and sets *err
    // Defect epilogue (the FP32 batches of the mixed solve): after the last iteration and one more
    // halo exchange, the corrected iterate phi_out = phi_in + e, the new defect
    // r_out = (float)(rhs64 - L_mu phi_out) and the workgroup's max |defect| (max_part[wg]) are
    // formed here, with the arithmetic of the separate defect kernel (cell_residual over
    // phi_in + (double)e), instead of writing e back for that kernel.
    int defect;
    GridK g;
    MuK<double> mu64;
    SorConst<double> s64;
    const double *phi_in, *rhs64;
    double *phi_out, *max_part;
    float *r_out;
    // workgroup 0 folds the workgroups' maxima into red_max[0] and the host slot, with the error
    // word red_max[1], then writes seq to host_slot[2]
    unsigned long long *red_max;
    double *host_slot;
    unsigned long long seq;
    // skip: return at once when the previous check, red_max[0], is not >= skip_tol (converged, or
    // not a number): a batch queued ahead of the host's read of that check then costs one launch
    int skip;
    double skip_tol;
    // half-column form: overlap the interior columns' halo-free cells with the exchange
    int ovl;
};

// a value as tagged 64-bit words: the tag in the high half, 32 payload bits in the low half
template <typename T> struct Tagged;
template <> struct Tagged<float> {
    static constexpr int NW = 1;
    __device__ static void pack(float v, unsigned tag, unsigned long long *o) {
        o[0] = ((unsigned long long)tag << 32) | __float_as_uint(v);
    }
    __device__ static float unpack(const unsigned long long *o) { return __uint_as_float((unsigned)o[0]); }
};
template <> struct Tagged<double> {
    static constexpr int NW = 2;
    __device__ static void pack(double v, unsigned tag, unsigned long long *o) {
        const unsigned long long b = (unsigned long long)__double_as_longlong(v);
        o[0] = ((unsigned long long)tag << 32) | (b & 0xffffffffull);
        o[1] = ((unsigned long long)tag << 32) | (b >> 32);
    }
    __device__ static double unpack(const unsigned long long *o) {
        return __longlong_as_double((long long)((o[1] << 32) | (o[0] & 0xffffffffull)));
    }
};

// Exchange direction q: 0-5 faces (x-, x+, y-, y+, z-, z+), two layers deep; 6-17 edges, one cell.
__host__ __device__ inline void rb_dir(int q, int &dx, int &dy, int &dz) {
    dx = dy = dz = 0;
    if (q < 6) {
        const int sgn = (q & 1) ? 1 : -1;
        if (q < 2) dx = sgn; else if (q < 4) dy = sgn; else dz = sgn;
        return;
    }
    const int e = q - 6, s1 = (e & 1) ? 1 : -1, s2 = (e & 2) ? 1 : -1;
    if (e < 4) { dx = s1; dy = s2; }          // x-y edges (lines along z)
    else if (e < 8) { dx = s1; dz = s2; }     // x-z edges (lines along y)
    else { dy = s1; dz = s2; }                // y-z edges (lines along x)
}

// extents of direction q's region: the subdomain size along axes where the direction is 0, else
// the depth (2 for faces, 1 for edges)
__host__ __device__ inline void rb_extent(int q, int bx, int by, int bz, int &ex, int &ey, int &ez) {
    int dx, dy, dz;
    rb_dir(q, dx, dy, dz);
    const int h = q < 6 ? RB_H : 1;
    ex = dx ? h : bx;
    ey = dy ? h : by;
    ez = dz ? h : bz;
}

// exchanged values per workgroup: the colour-1 half of the 18 regions (every region holds an even
// number of cells, as every subdomain extent is even)
__host__ __device__ inline int rb_xchg_total(int bx, int by, int bz) {
    int t = 0;
    for (int q = 0; q < 18; ++q) {
        int ex, ey, ez;
        rb_extent(q, bx, by, bz, ex, ey, ez);
        t += ex * ey * ez / 2;
    }
    return t;
}

// colour-0 ring cells: on each face ring, half of by*bz (x faces), bx*bz (y) or bx*by (z) cells
__host__ __device__ inline int rb_ring_total(int bx, int by, int bz) { return by * bz + bx * bz + bx * by; }

// COLS: column form.  A thread owns one column (i, j) of the subdomain or of the face ring and
// marches it in z, relaxing every other cell; the face coefficients of the column stay in
// registers (z-invariant tables) and the z neighbour between two cells of one colour is loaded
// once, so a cell costs six LDS loads and one store instead of thirteen loads and a store.
// Otherwise (pair form) a thread owns MAXP x-pairs and up to RB_MAXR ring cells anywhere in the
// tile and reads their coefficients from an LDS table; the pair form also relaxes the deep box's
// colour 0 while the halo is in flight.
// FORM: 0 pair form, 1 column form, 2 column form with two threads per column (the lower and the
// upper half of its z range) in workgroups of twice the threads.
template <typename T, bool IBM, int MAXP, int FORM>
__global__ __launch_bounds__(FORM == 2 ? 2 * RB_THREADS : RB_THREADS) void rb_lds_kernel(RbArgs<T> a) {
    // threads per workgroup, exchange entries per thread (the face buffer layout is the same for
    // every form) and, in the column forms, cells per thread and colour
    constexpr bool COLS = FORM > 0;
    constexpr int NT = FORM == 2 ? 2 * RB_THREADS : RB_THREADS, HE = RB_XSTRIDE / NT;
    constexpr int ZC = FORM == 2 ? RB_ZC2 : RB_ZC1;
    if (a.skip) {
        // every workgroup reads the same word: the previous kernel's result, which this launch's
        // workgroup 0 overwrites only after all workgroups have published their epilogue maxima
        const unsigned long long rb = __hip_atomic_load(&a.red_max[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (!(__longlong_as_double((long long)rb) >= a.skip_tol)) return;
    }
    extern __shared__ unsigned char smem_raw[];
    using TW = Tagged<T>;
    constexpr int NW = TW::NW;
    const Dims d = a.d;
    const int bx = a.bx, by = a.by, bz = a.bz;
    const int tx = bx + 2 * RB_H, ty = by + 2 * RB_H, tz = bz + 2 * RB_H;   // tile extent in use
    constexpr int cw = RB_TXP, cpl = RB_PLANE;   // coefficient table: local i, j in [0, b+1]
    T *X = reinterpret_cast<T *>(smem_raw);
    // unused slots of the per-thread cell lists point at a dummy cell past the tile, whose stencil
    // reads and writes stay inside a padding of 2 planes + 3 values: the relax loops need no branches
    const int dummy = RB_PLANE * tz + RB_PLANE + 1;
    T *C = X + (size_t)RB_PLANE * tz + 2 * RB_PLANE + 3;   // IBM: 6 planes of cpl values
                                                          // (planes: threads read at pair stride)
    const int nwg = a.px * a.py * a.pz;
    // XCD-aware placement: workgroups are dispatched round-robin over the 8 XCDs, so consecutive
    // subdomain ids (neighbours along x first) are given to workgroups on the same XCD.
    const int per_xcd = nwg / 8;
    const int b = blockIdx.x;
    const int w = (nwg % 8 == 0) ? (b % 8) * per_xcd + b / 8 : b;
    const int wx = w % a.px, wy = (w / a.px) % a.py, wz = w / (a.px * a.py);
    const int ox = wx * bx, oy = wy * by, oz = wz * bz;
    const int t = threadIdx.x;
    const bool per_y = a.per_y != 0;
    // LDS index of local cell (i, j, k); the subdomain is 1..b, the halo -1..0 and b+1..b+2
    auto L = [&](int i, int j, int k) { return (i + 1) + RB_TXP * (j + 1) + RB_PLANE * (k + 1); };
    // periodic image in 1..n of a coordinate within one period of the domain (-n < g <= 2n)
    auto wrap = [](int g, int n) { return g < 1 ? g + n : (g > n ? g - n : g); };
    // the workgroup in direction (dx, dy, dz), or -1 beyond a wall
    auto neighbour = [&](int dx, int dy, int dz) {
        int ny_ = wy + dy;
        if (ny_ < 0 || ny_ >= a.py) {
            if (!per_y) return -1;
            ny_ = (ny_ + a.py) % a.py;
        }
        return (wx + dx + a.px) % a.px + a.px * (ny_ + a.py * ((wz + dz + a.pz) % a.pz));
    };
    auto wall_bits = [&](int gj) { return (!per_y && gj == d.ny ? 1 : 0) | (!per_y && gj == 1 ? 2 : 0); };

    // ---- load the tile (subdomain and halo) from the global solution, in LDS order over the fixed
    // row and plane (constant divisors; values past the tile extent in a row or plane are zero)
    for (int q = t; q < RB_PLANE * tz; q += NT) {
        const int xi = q % RB_TXP, yi = (q / RB_TXP) % RB_TYP;
        const int li = xi - 1, lj = yi - 1, lk = q / RB_PLANE - 1;
        T v = T(0);
        if (!a.zero_init && xi < tx && yi < ty) {
            const int gi = wrap(ox + li, d.nx), gk = wrap(oz + lk, d.nz);
            int gj = oy + lj;
            if (per_y) gj = wrap(gj, d.ny);
            if (gj >= 1 && gj <= d.ny) v = a.x[d.in(gi, gj, gk)];
        }
        X[q] = v;
    }
    // ---- face coefficient table (one x-y plane, ring included): mui, muip, mvj, mvjp, mw, -omega/denom
    T qconst = T(0);
    if (IBM && !COLS) {
        const int sx = d.sx;
        for (int q0 = t; q0 < (bx + 2) * (by + 2); q0 += NT) {
            const int li = q0 % (bx + 2), lj = q0 / (bx + 2), q = li + cw * lj;
            const int gi = wrap(ox + li, d.nx);
            int gj = oy + lj;
            if (per_y) gj = wrap(gj, d.ny);
            T mui = 1, muip = 1, mvj = 1, mvjp = 1, mw = 1, qq = 0;
            if (gj >= 1 && gj <= d.ny) {   // rows beyond a wall are never relaxed
                const int gip = gi == d.nx ? 1 : gi + 1;
                const int gjp = (per_y && gj == d.ny) ? 1 : gj + 1;
                mui = a.mu_u[(size_t)gi + (size_t)sx * gj];
                muip = a.mu_u[(size_t)gip + (size_t)sx * gj];
                mvj = a.mu_v[(size_t)gi + (size_t)sx * gj];
                mvjp = a.mu_v[(size_t)gi + (size_t)sx * gjp];
                mw = a.mu_w[(size_t)gi + (size_t)sx * gj];
                const T denom = (mui + muip) * a.s.idx2 + (mw + mw) * a.s.idz2 + (mvj + mvjp) * a.s.idy2;
                qq = -a.omega / denom;
            }
            C[q] = mui; C[q + cpl] = muip; C[q + 2 * cpl] = mvj; C[q + 3 * cpl] = mvjp; C[q + 4 * cpl] = mw;
            C[q + 5 * cpl] = qq;
        }
    } else {
        const T one = T(1);
        const T denom = (one + one) * a.s.idx2 + (one + one) * a.s.idz2 + (one + one) * a.s.idy2;
        qconst = -a.omega / denom;
    }

    // ---- subdomain pairs of this thread, packed in one word: LDS index of the colour-0 cell
    // (bits 0-15), its coefficient-table index (bits 16-27), wall flags (bits 28-29) and whether
    // the pair's colour-1 cell is at -1 (bit 30)
    // Pair order: first the pairs of the deep box (x-pairs 1..hbx-2, rows 2..by-1, planes 2..bz-1),
    // whose colour-0 cells read no halo, then the shell around it.  Entries m < a.md of every thread
    // are deep (a.md: whole relax groups), so their colour 0 can run while the halo is in flight.
    const int hbx = bx / 2, npairs = hbx * by * bz;
    const int dpx = hbx > 2 ? hbx - 2 : 0, dpy = by > 2 ? by - 2 : 0, dpz = bz > 2 ? bz - 2 : 0;
    const int ndeep = dpx * dpy * dpz;
    auto pair_of = [&](int p, int &pi, int &j, int &k) {
        if (p < ndeep) {
            pi = 1 + p % dpx;
            j = 2 + (p / dpx) % dpy;
            k = 2 + p / (dpx * dpy);
            return;
        }
        int q = p - ndeep;
        if (ndeep == 0) {   // no deep box: natural order
            pi = q % hbx;
            j = (q / hbx) % by + 1;
            k = q / (hbx * by) + 1;
            return;
        }
        const int s1 = 2 * hbx * by;   // planes 1 and bz
        if (q < s1) {
            k = q < hbx * by ? 1 : bz;
            const int r = q % (hbx * by);
            pi = r % hbx;
            j = r / hbx + 1;
            return;
        }
        q -= s1;
        const int rr = 2 * hbx + 2 * (by - 2);   // shell of a middle plane: rows 1, by and the end pairs
        k = q / rr + 2;
        int r = q % rr;
        if (r < 2 * hbx) {
            j = r < hbx ? 1 : by;
            pi = r % hbx;
        } else {
            r -= 2 * hbx;
            j = r / 2 + 2;
            pi = (r & 1) ? hbx - 1 : 0;
        }
    };
    auto pack = [](int c, int ti, int walls, int neg) { return c | (ti << 16) | (walls << 28) | (neg << 30); };
    constexpr int NP = COLS ? 1 : MAXP, NR = COLS ? 1 : RB_MAXR;
    int pw[NP];
    T rh[NP][2];
#pragma unroll
    for (int m = 0; m < NP; ++m) {
        const int p = t + m * NT;
        pw[m] = pack(dummy, 0, 0, 0);
        rh[m][0] = rh[m][1] = T(0);
        if (!COLS && p < npairs) {
            int pi, j, k;
            pair_of(p, pi, j, k);
            const int i0 = 2 * pi + 1;
            const int gj = oy + j;
            // colour c relaxes the cell with (gi + gj + gk + c) odd
            const int off0 = ((ox + i0 + gj + oz + k) & 1) ? 0 : 1;
            pw[m] = pack(L(i0 + off0, j, k), (i0 + off0) + cw * j, wall_bits(gj), off0);
            rh[m][0] = a.rhs[d.in(ox + i0 + off0, gj, oz + k)];
            rh[m][1] = a.rhs[d.in(ox + i0 + 1 - off0, gj, oz + k)];
        }
    }
    // ---- colour-0 cells of the face ring owned by neighbours, recomputed here.  Face f has fixed
    // coordinate 0 or b+1 along its axis A and spans 1..bU x 1..bV (U = y for x faces, else x);
    // with bU even each U-line holds bU/2 colour-0 cells.
    int rw[NR];   // packed as the pair words
    T rr[NR];
#pragma unroll
    for (int m = 0; m < NR; ++m) {
        rw[m] = pack(dummy, 0, 0, 0);
        rr[m] = T(0);
    }
    if (!COLS) {
        const int nr = rb_ring_total(bx, by, bz);
#pragma unroll
        for (int m = 0; m < NR; ++m) {
            int e = t + m * NT;
            rw[m] = pack(dummy, 0, 0, 0);
            rr[m] = T(0);
            if (e < nr) {
                int f;   // 0 x faces, 1 y faces, 2 z faces; two faces per axis
                if (e < by * bz) f = 0;
                else if ((e -= by * bz) < bx * bz) f = 1;
                else { e -= bx * bz; f = 2; }
                const int bU = f == 0 ? by : bx, bV = f == 2 ? by : bz;
                const int half = bU / 2 * bV, side = e / half;
                const int r = e - side * half, v = r / (bU / 2) + 1, uq = r % (bU / 2);
                const int dxn = f == 0 ? (side ? 1 : -1) : 0, dyn = f == 1 ? (side ? 1 : -1) : 0,
                          dzn = f == 2 ? (side ? 1 : -1) : 0;
                if (neighbour(dxn, dyn, dzn) >= 0) {
                    // local coordinates with u = 2 uq + 1 + s (s picks the colour-0 cell)
                    int li, lj, lk;
                    const int cA = side ? (f == 0 ? bx + 1 : (f == 1 ? by + 1 : bz + 1)) : 0;
                    if (f == 0) { li = cA; lj = 2 * uq + 1; lk = v; }
                    else if (f == 1) { li = 2 * uq + 1; lj = cA; lk = v; }
                    else { li = 2 * uq + 1; lj = v; lk = cA; }
                    const int s = ((ox + li + oy + lj + oz + lk) & 1) ? 0 : 1;
                    if (f == 0) lj += s; else li += s;
                    const int gi = wrap(ox + li, d.nx), gk = wrap(oz + lk, d.nz);
                    int gj = oy + lj;
                    if (per_y) gj = wrap(gj, d.ny);
                    rw[m] = pack(L(li, lj, lk), li + cw * lj, wall_bits(gj), 0);
                    rr[m] = a.rhs[d.in(gi, gj, gk)];
                }
            }
        }
    }
    // ---- column form: the column of this thread.  The subdomain's columns march z from 0 to bz+1
    // in colour 0 (the two end cells are the z faces of the ring) and from 1 to bz in colour 1;
    // after them come the columns of the ring's x faces (2*by) and y faces (2*bx), colour 0 only, z
    // from 1 to bz, none beyond a wall.  FORM 2 orders the subdomain's columns interior first:
    // those at least 3 cells from every x and y face (i in 3..bx-2, j in 3..by-2), padded to whole
    // wavefronts where the threads allow, then the frame of the two outer rings.
    // The cells of one colour in a thread's range are its end cell, next to the z face the range
    // ends at (the lowest; the highest in FORM 2's upper halves), and the rest, marched upward.  In
    // a wavefront of interior columns only (csplit), the rest of each colour reads no halo value
    // and sends none (colour 0 is at least 2 cells from every face, colour 1 at least 3), so it
    // relaxes while the halo exchange is in flight: colour 1's after the publish, the next
    // iteration's colour 0 after that.  Cells past a range's end, and every cell of a thread
    // without a column, go to the dummy cell.
    int cb0 = dummy, cr0 = dummy, nr0 = 0;   // colour 0: end cell, first of the rest, their count
    int cb1 = dummy, cr1 = dummy, nr1 = 0;   // colour 1
    bool csplit = false;
    bool cwn = false, cws = false;
    T cmui = 1, cmuip = 1, cmvj = 1, cmvjp = 1, cmw = 1, cq = qconst;
    T rb0 = T(0), rb1 = T(0);
    T r0[COLS ? ZC - 1 : 1], r1[COLS ? ZC - 1 : 1];
#pragma unroll
    for (int n = 0; n < (COLS ? ZC - 1 : 1); ++n) r0[n] = r1[n] = T(0);
    if (COLS) {
        // FORM 2: threads tc and tc + RB_THREADS share column tc, the lower and the upper half of its
        // z range
        const int tc = t % RB_THREADS, half = t / RB_THREADS;
        const int nin = bx * by, nxf = 2 * by, nyf = 2 * bx;
        int nI = 0, nIp = 0;   // interior columns, and the threads they take with their padding
        if (FORM == 2 && bx > 4 && by > 4) {
            nI = (bx - 4) * (by - 4);
            nIp = (nI + 63) / 64 * 64;
            if (nIp + (nin - nI) + nxf + nyf > RB_THREADS) nIp = nI;
        }
        const int nsub = nIp + nin - nI;   // threads of the subdomain's columns, padding included
        int cli = 0, clj = 0, kz0 = 0, kz1 = -1, kc0 = 1, kc1 = bz;   // colour-0 and colour-1 z ranges
        bool inner = false;
        if (tc < nsub) {
            if (tc < nI) {
                cli = 3 + tc % (bx - 4); clj = 3 + tc / (bx - 4); inner = true;
            } else if (tc >= nIp) {
                const int f = tc - nIp;
                if (nI == 0) {   // natural order
                    cli = f % bx + 1; clj = f / bx + 1;
                } else if (f < 4 * bx) {   // rows 1, 2, by-1, by
                    const int r = f / bx;
                    cli = f % bx + 1; clj = r < 2 ? r + 1 : by - 3 + r;
                } else {   // columns 1, 2, bx-1, bx of rows 3 .. by-2
                    const int g2 = f - 4 * bx, c = g2 % 4;
                    clj = 3 + g2 / 4; cli = c < 2 ? c + 1 : bx - 3 + c;
                }
                inner = true;
            }
            if (inner) { kz0 = 0; kz1 = bz + 1; }
        } else if (tc < nsub + nxf) {
            const int e = tc - nsub;
            cli = e < by ? 0 : bx + 1; clj = e % by + 1; kz0 = 1; kz1 = bz;
        } else if (tc < nsub + nxf + nyf) {
            const int e = tc - nsub - nxf;
            clj = e < bx ? 0 : by + 1; cli = e % bx + 1;
            if (neighbour(0, clj == 0 ? -1 : 1, 0) >= 0) { kz0 = 1; kz1 = bz; }
        }
        // the wavefront's threads tc are 64 consecutive ones: all interior (or padding)
        csplit = FORM == 2 && a.ovl && ((tc >> 6) + 1) * 64 <= nIp;
        const bool top = FORM == 2 && half == 1;   // the end cell is the highest of the range
        if (FORM == 2) {
            const int h0 = (kz1 - kz0 + 2) / 2, h1 = bz / 2;   // positions in the lower halves
            if (half == 0) { kz1 = kz0 + h0 - 1; kc1 = h1; }
            else { kz0 += h0; kc0 = h1 + 1; }
        }
        if (kz1 >= kz0) {
            const int gi = wrap(ox + cli, d.nx);
            int gj = oy + clj;
            if (per_y) gj = wrap(gj, d.ny);
            const int par = (ox + cli + oy + clj + oz) & 1;   // colour c relaxes (par + lk + c) odd
            // colour 0: cells k0, k0 + 2, ..., n0 of them
            const int k0 = kz0 + (((par + kz0) & 1) ? 0 : 1);
            const int n0 = k0 <= kz1 ? (kz1 - k0) / 2 + 1 : 0;
            if (n0 > 0) {
                const int kb = top ? k0 + 2 * (n0 - 1) : k0, kr = top ? k0 : k0 + 2;
                cb0 = L(cli, clj, kb);
                rb0 = a.rhs[d.in(gi, gj, wrap(oz + kb, d.nz))];
                nr0 = n0 - 1;
                if (nr0 > 0) cr0 = L(cli, clj, kr);
#pragma unroll
                for (int n = 0; n < ZC - 1; ++n)
                    r0[n] = n < nr0 ? a.rhs[d.in(gi, gj, wrap(oz + kr + 2 * n, d.nz))] : T(0);
            }
            if (inner) {
                const int k1 = kc0 + (((par + kc0) & 1) ? 1 : 0);   // colour 1: (par + lk) even
                const int n1 = k1 <= kc1 ? (kc1 - k1) / 2 + 1 : 0;
                if (n1 > 0) {
                    const int kb = top ? k1 + 2 * (n1 - 1) : k1, kr = top ? k1 : k1 + 2;
                    cb1 = L(cli, clj, kb);
                    rb1 = a.rhs[d.in(gi, gj, oz + kb)];
                    nr1 = n1 - 1;
                    if (nr1 > 0) cr1 = L(cli, clj, kr);
#pragma unroll
                    for (int n = 0; n < ZC - 1; ++n) r1[n] = n < nr1 ? a.rhs[d.in(gi, gj, oz + kr + 2 * n)] : T(0);
                }
            }
            const int wb = wall_bits(gj);
            cwn = (wb & 1) != 0;
            cws = (wb & 2) != 0;
            if (IBM) {   // the face coefficients of the column, as the streaming sweeps form them
                const int sx = d.sx, gip = gi == d.nx ? 1 : gi + 1, gjp = (per_y && gj == d.ny) ? 1 : gj + 1;
                cmui = a.mu_u[(size_t)gi + (size_t)sx * gj];
                cmuip = a.mu_u[(size_t)gip + (size_t)sx * gj];
                cmvj = a.mu_v[(size_t)gi + (size_t)sx * gj];
                cmvjp = a.mu_v[(size_t)gi + (size_t)sx * gjp];
                cmw = a.mu_w[(size_t)gi + (size_t)sx * gj];
                const T denom = (cmui + cmuip) * a.s.idx2 + (cmw + cmw) * a.s.idz2 + (cmvj + cmvjp) * a.s.idy2;
                cq = -a.omega / denom;
            }
        }
    }
    // ---- exchange lists (built once per plan by rb_xlist_kernel): entry e = t + jj * NT
    __shared__ int nbr_from[19];   // the workgroup that sends direction q's region to this one
    if (t < 18) {
        int dx, dy, dz;
        rb_dir(t, dx, dy, dz);
        nbr_from[t] = neighbour(-dx, -dy, -dz);
    }
    if (t == 18) nbr_from[18] = w;
    int xpub[HE], xdst[HE];
#pragma unroll
    for (int jj = 0; jj < HE; ++jj) {
        const size_t e = (size_t)t + (size_t)jj * NT;
        xpub[jj] = a.xl[(size_t)w * RB_XSTRIDE + e];
        xdst[jj] = a.xl[((size_t)nwg + w) * RB_XSTRIDE + e];
    }
    bool failed = false;
    __syncthreads();

    // relax the cell of packed word pw shifted by dlt (0: the colour-0 cell, +-1: its pair partner)
    auto relax = [&](int pwd, int dlt, T rhs) -> T {
        const int c = (pwd & 0xffff) + dlt, ti = ((pwd >> 16) & 0xfff) + dlt;
        T mui = 1, muip = 1, mvj = 1, mvjp = 1, mw = 1, q = qconst;
        if (IBM) {
            mui = C[ti]; muip = C[ti + cpl]; mvj = C[ti + 2 * cpl]; mvjp = C[ti + 3 * cpl]; mw = C[ti + 4 * cpl];
            q = C[ti + 5 * cpl];
        }
        T denom;
        const T p0 = X[c];
        const T r = sor_residual_core<T, false>(p0, X[c + 1], X[c - 1], X[c + RB_TXP], X[c - RB_TXP],
                                                X[c + RB_PLANE], X[c - RB_PLANE], mui, muip, mvj, mvjp, mw, mw,
                                                (pwd >> 28) & 1,
                                                (pwd >> 29) & 1, rhs, a.s, T(0), T(0), denom);
        return p0 + r * q;
    };
    auto partner = [](int pwd) { return ((pwd >> 30) & 1) ? -1 : 1; };
    // column form: the cell at LDS index c with its z neighbours already in registers
    auto relax_col = [&](int c, T zlo, T zhi, T rhs) -> T {
        T denom;
        const T p0 = X[c];
        const T r = sor_residual_core<T, false>(p0, X[c + 1], X[c - 1], X[c + RB_TXP], X[c - RB_TXP], zhi, zlo, cmui,
                                                cmuip, cmvj, cmvjp, cmw, cmw, cwn, cws, rhs, a.s, T(0), T(0), denom);
        return p0 + r * cq;
    };
    // the rest of one colour along this thread's column: n cells upward from LDS index c0, two
    // planes apart; the upper z neighbour of a cell is the lower one of the next
    auto march = [&](int c0, int n, const T *rhs) {
        int c = c0;
        T zlo = X[c - RB_PLANE];
#pragma unroll
        for (int q = 0; q < ZC - 1; ++q) {
            const int cc = q < n ? c : dummy;
            const T zhi = X[cc + RB_PLANE];
            const T v = relax_col(cc, zlo, zhi, rhs[q]);
            X[cc] = v;
            zlo = zhi;
            c += 2 * RB_PLANE;
            if ((q % RB_G) == RB_G - 1) __builtin_amdgcn_sched_barrier(0);
        }
    };
    // the end cell of one colour
    auto relax_end = [&](int c, T rhs) {
        const T v = relax_col(c, X[c - RB_PLANE], X[c + RB_PLANE], rhs);
        X[c] = v;
    };

    unsigned long long *F = a.faces;
    // publish the colour-1 layer cells, tagged with iteration s, into buffer s & 1
    auto publish = [&](int s) {
        const unsigned tag = a.tag0 + (unsigned)s + 1u;
        const size_t mybase = ((size_t)(w * 2 + (s & 1)) * RB_XSTRIDE) * NW;
        // every layer value first (one LDS wait), then the stores
        T val[HE];
#pragma unroll
        for (int jj = 0; jj < HE; ++jj) val[jj] = X[xpub[jj] & 0xfffff];
#pragma unroll
        for (int jj = 0; jj < HE; ++jj) asm volatile("" : "+v"(val[jj]));
#pragma unroll
        for (int jj = 0; jj < HE; ++jj) {
            unsigned long long wv[NW];
            TW::pack(val[jj], tag, wv);
            unsigned long long *dst = F + mybase + (size_t)(t + jj * NT) * NW;
#pragma unroll
            for (int q = 0; q < NW; ++q)
                __hip_atomic_store(dst + q, wv[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    };
    for (int s = 0; s < a.iters; ++s) {
        // The packed words are loop-invariant; left visible, the compiler would keep every address
        // derived from them (13 per cell and colour) live across the loop and run out of registers.
        // An empty asm that claims to redefine them makes the few address adds happen per use.
#pragma unroll
        for (int m = 0; m < NP; ++m) asm volatile("" : "+v"(pw[m]));
#pragma unroll
        for (int m = 0; m < NR; ++m) asm volatile("" : "+v"(rw[m]));
#pragma unroll
        for (int jj = 0; jj < HE; ++jj) asm volatile("" : "+v"(xpub[jj]), "+v"(xdst[jj]));
        // ---- colour 0 on the subdomain and the face ring.  Cells of one colour never neighbour
        // each other, so each is stored as soon as it is formed; groups of RB_G cells with a
        // scheduling barrier between them bound the loads in flight (registers).  After the first
        // iteration the deep groups (g < a.md) were relaxed while the halo was in flight.
        if (COLS) {
            // column data: loop-invariant, kept from being hoisted as per-cell addresses (as above)
            asm volatile("" : "+v"(cb0), "+v"(cr0), "+v"(cb1), "+v"(cr1));
            // colour 0: the end cells, and the rest where the last exchange's wait has not relaxed it
            relax_end(cb0, rb0);
            if (!csplit || s == 0) march(cr0, nr0, r0);
            __syncthreads();
            // colour 1: the end cells, and the rest where it is not left for the exchange's flight
            relax_end(cb1, rb1);
            if (!csplit) march(cr1, nr1, r1);
            __syncthreads();
        }
#pragma unroll
        for (int g = 0; g < NP; g += RB_G) {
            if (COLS) break;
            if (s > 0 && g < a.md) continue;
            T nv[RB_G];
#pragma unroll
            for (int u = 0; u < RB_G; ++u) nv[u] = relax(pw[g + u], 0, rh[g + u][0]);
#pragma unroll
            for (int u = 0; u < RB_G; ++u) X[pw[g + u] & 0xffff] = nv[u];
            __builtin_amdgcn_sched_barrier(0);
        }
        if (!COLS) {
            T nv[NR];
#pragma unroll
            for (int m = 0; m < NR; ++m) nv[m] = relax(rw[m], 0, rr[m]);
#pragma unroll
            for (int m = 0; m < NR; ++m) X[rw[m] & 0xffff] = nv[m];
            __syncthreads();
        }
        // ---- colour 1 on the subdomain: its neighbours are subdomain or ring cells, all updated
#pragma unroll
        for (int g = 0; g < NP; g += RB_G) {
            if (COLS) break;
            T nv[RB_G];
#pragma unroll
            for (int u = 0; u < RB_G; ++u) nv[u] = relax(pw[g + u], partner(pw[g + u]), rh[g + u][1]);
#pragma unroll
            for (int u = 0; u < RB_G; ++u) X[(pw[g + u] & 0xffff) + partner(pw[g + u])] = nv[u];
            __builtin_amdgcn_sched_barrier(0);
        }
        if (!COLS) __syncthreads();
        // the last iteration needs no exchange, unless the defect epilogue reads the final halo
        if (s == a.iters - 1 && !a.defect) {
            if (COLS && a.ovl) {   // the rest of colour 1 in the split wavefronts, before the write-back
                if (csplit) march(cr1, nr1, r1);
                __syncthreads();
            }
            break;
        }

        // ---- publish this subdomain's colour-1 layer cells
        const int buf = s & 1;
        const unsigned tag = a.tag0 + (unsigned)s + 1u;
        publish(s);
        // ---- column form, split wavefronts: the rest of colour 1, then (after every colour-1 cell
        // it reads is final) the next iteration's rest of colour 0, while the layers travel.  The
        // publish reads only colour-1 layer cells, which neither writes, and the halo the pull
        // writes is read by neither.
        if (COLS && a.ovl) {
            if (csplit) march(cr1, nr1, r1);
            if (s < a.iters - 1) {
                __syncthreads();
                if (csplit) march(cr0, nr0, r0);
            }
        }
        // ---- the next iteration's colour 0 on the deep groups, while the neighbours' layers travel:
        // those cells read only subdomain cells of colour 1, final for this iteration; the barrier
        // lets the publish read its layers first
        if (!COLS && s < a.iters - 1) {
            __syncthreads();
#pragma unroll
            for (int g = 0; g < NP; g += RB_G) {
                if (g >= a.md) continue;
                T nv[RB_G];
#pragma unroll
                for (int u = 0; u < RB_G; ++u) nv[u] = relax(pw[g + u], 0, rh[g + u][0]);
#pragma unroll
                for (int u = 0; u < RB_G; ++u) X[pw[g + u] & 0xffff] = nv[u];
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        // ---- pull the neighbours' layers: poll the tagged words until all carry this tag
        // source word of entry jj: the sender's buffer `buf`, entry e, in words
        auto src_of = [&](int jj) {
            const int sender = nbr_from[(xdst[jj] >> 20) & 31];
            return ((size_t)(sender * 2 + buf) * RB_XSTRIDE + (size_t)(t + jj * NT)) * NW;
        };
        unsigned long long got[HE][NW];
#pragma unroll
        for (int jj = 0; jj < HE; ++jj)
#pragma unroll
            for (int q = 0; q < NW; ++q)
                got[jj][q] = __hip_atomic_load(F + src_of(jj) + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int spins = 0; !failed; ++spins) {
            bool ready = true;
#pragma unroll
            for (int jj = 0; jj < HE; ++jj)
#pragma unroll
                for (int q = 0; q < NW; ++q)
                    ready = ready && (!((xdst[jj] >> 25) & 1) || (unsigned)(got[jj][q] >> 32) == tag);
            if (ready) break;
            if (spins >= a.spin_limit) {   // a neighbour never published: flag it and run to the end
                __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                failed = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
#pragma unroll
            for (int jj = 0; jj < HE; ++jj)
#pragma unroll
                for (int q = 0; q < NW; ++q)
                    if (((xdst[jj] >> 25) & 1) && (unsigned)(got[jj][q] >> 32) != tag)
                        got[jj][q] = __hip_atomic_load(F + src_of(jj) + q, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
        }
#pragma unroll
        for (int jj = 0; jj < HE; ++jj)
            X[xdst[jj] & 0xfffff] = TW::unpack(got[jj]);
        __syncthreads();
    }
    if (!a.defect) {
        // ---- write the subdomain back
        for (int q = t; q < bx * by * bz; q += NT) {
            const int i = q % bx + 1, j = (q / bx) % by + 1, k = q / (bx * by) + 1;
            a.x[d.in(ox + i, oy + j, oz + k)] = X[L(i, j, k)];
        }
        return;
    }
    // ---- defect epilogue: the halo holds the neighbours' final e (exchanged above).  A thread
    // owns a column (i, j) of the subdomain, or half of its planes where the workgroup has the
    // threads (EP = 2), and marches k, keeping Phi = phi_in + e of planes k-1, k, k+1 of its column;
    // the face coefficients of a column do not change with k (z-invariant tables).  The pointers
    // are declared unaliased so the loads of the next plane may issue before this plane's stores.
    // The operands handed to sor_residual_core are those cell_residual forms, so the defect equals
    // the separate kernel's.
    double mx = 0.0;
    const int ncol = bx * by;
    const int EP = NT >= 2 * ncol ? 2 : 1;
    if (t < EP * ncol) {
        const int col = t % ncol, part = t / ncol;
        const int li = col % bx + 1, lj = col / bx + 1;
        const int hz = (bz + EP - 1) / EP;
        const int lk0 = 1 + part * hz, lk1 = lk0 + hz - 1 < bz ? lk0 + hz - 1 : bz;
        const int i = ox + li, j = oy + lj, nx = d.nx, ny = d.ny, nz = d.nz;
        const bool per2 = per_y;
        const int ip = i == nx ? 1 : i + 1, im = i == 1 ? nx : i - 1;
        int jp = j + 1, jm = j - 1;
        if (per2) {
            if (jp > ny) jp = 1;
            if (jm < 1) jm = ny;
        }
        const bool wn_ = !per2 && j == ny, ws_ = !per2 && j == 1;
        const int jn = wn_ ? j : jp, js = ws_ ? j : jm;
        const int ljn = wn_ ? lj : lj + 1, ljs = ws_ ? lj : lj - 1;   // LDS rows of the y neighbours
        double mui = 1, muip = 1, mvj = 1, mvjp = 1, mwk = 1, mwkp = 1;
        if (IBM) {
            mui = a.mu64.u[a.mu64.at(i, j, 1)];
            muip = a.mu64.u[a.mu64.at(ip, j, 1)];
            mvj = a.mu64.v[a.mu64.at(i, j, 1)];
            mvjp = a.mu64.v[a.mu64.at(i, jp, 1)];
            mwk = a.mu64.w[a.mu64.at(i, j, 1)];
            mwkp = mwk;   // z-invariant table
        }
        const double *__restrict__ pin = a.phi_in;
        const double *__restrict__ r64 = a.rhs64;
        double *__restrict__ pout = a.phi_out;
        float *__restrict__ rout = a.r_out;
        auto Pc = [&](int gi, int gj, int gk, int xi, int xj, int xk) {   // phi_in + e of one cell
            return pin[d.in(gi, gj, gk)] + (double)X[L(xi, xj, xk)];
        };
        auto kw = [&](int k) { return k < 1 ? k + nz : (k > nz ? k - nz : k); };
        // z neighbours of the first plane: below the subdomain the lower neighbour wraps periodically
        double pm = Pc(i, j, kw(oz + lk0 - 1), li, lj, lk0 - 1), pc = Pc(i, j, oz + lk0, li, lj, lk0);
#pragma unroll 2
        for (int lk = lk0; lk <= lk1; ++lk) {
            const int k = oz + lk;
            const double pn = Pc(i, j, kw(k + 1), li, lj, lk + 1);
            const double pxp = Pc(ip, j, k, li + 1, lj, lk), pxm = Pc(im, j, k, li - 1, lj, lk);
            const double pyp = Pc(i, jn, k, li, ljn, lk), pym = Pc(i, js, k, li, ljs, lk);
            const size_t c = d.in(i, j, k);
            double denom;
            const double r = sor_residual_core<double, false>(pc, pxp, pxm, pyp, pym, pn, pm, mui, muip, mvj, mvjp,
                                                              mwk, mwkp, wn_, ws_, r64[c], a.s64, 0.0, 0.0,
                                                              denom);
            pout[c] = pc;
            rout[c] = (float)r;
            mx = nanmax_any(mx, fabs(r));
            pm = pc;
            pc = pn;
        }
    }
    // ---- the maximum over workgroups without fences: every workgroup publishes its partial as two
    // tagged words (tag0 + iters + 1, unique to this launch) after the face buffers; workgroup 0
    // polls them all (bounded) and writes the result to red_max[0] and the host slot
    __shared__ double wred[NT / 64];
    for (int off = 32; off > 0; off >>= 1) mx = nanmax_any(mx, __shfl_xor(mx, off, 64));
    if ((t & 63) == 0) wred[t >> 6] = mx;
    __syncthreads();
    const unsigned mtag = a.tag0 + (unsigned)a.iters + 1u;
    unsigned long long *MP = a.faces + (size_t)nwg * 2 * RB_XSTRIDE * NW;   // 2 words per workgroup
    if (t == 0) {
        double m = wred[0];
        for (int q = 1; q < NT / 64; ++q) m = nanmax_any(m, wred[q]);
        unsigned long long wv[2];
        Tagged<double>::pack(m, mtag, wv);
        __hip_atomic_store(MP + 2 * w, wv[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(MP + 2 * w + 1, wv[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (blockIdx.x != 0) return;
    double m = 0.0;
    for (int q = t; q < nwg; q += NT) {
        unsigned long long wv[2];
        for (int spins = 0;; ++spins) {
            wv[0] = __hip_atomic_load(MP + 2 * q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            wv[1] = __hip_atomic_load(MP + 2 * q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((unsigned)(wv[0] >> 32) == mtag && (unsigned)(wv[1] >> 32) == mtag) break;
            if (failed || spins >= a.spin_limit) {
                __hip_atomic_store(a.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                failed = true;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        m = nanmax_any(m, Tagged<double>::unpack(wv));
    }
    __syncthreads();   // wred reused
    for (int off = 32; off > 0; off >>= 1) m = nanmax_any(m, __shfl_xor(m, off, 64));
    if ((t & 63) == 0) wred[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        for (int q = 1; q < NT / 64; ++q) m = nanmax_any(m, wred[q]);
        m = nanmax_any(wred[0], m);
        const unsigned long long errw = __hip_atomic_load(&a.red_max[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        a.red_max[0] = (unsigned long long)__double_as_longlong(m);
        a.host_slot[0] = m;
        a.host_slot[1] = __longlong_as_double((long long)errw);
        __threadfence_system();
        reinterpret_cast<volatile unsigned long long *>(a.host_slot)[2] = a.seq;
    }
}

// Exchange lists of workgroup w (subdomain id), entry e of the colour-1 cells of the 18 regions
// (faces two layers deep, then edges one cell thick; entry e = t + jj * threads of thread t).
// Publish word: the LDS cell this workgroup sends for its neighbour in direction D (bits 0-19).
// Pull word: the halo cell where the region that the neighbour in direction -D sent lands (bits
// 0-19), D (bits 20-24) and whether to wait for it (bit 25).  Colour 1 has an even global
// coordinate sum; cells 2k and 2k+1 of a region's natural order are neighbours along its first
// extent above one, of opposite colours, and entry k of the region is the one of the two of
// colour 1, on both sides alike.
// Unused entries publish the dummy cell; an entry with no sender (beyond a wall, or past the end
// of the list) does not wait, reads this workgroup's own slot (direction 18) and lands in the
// dummy cell.  Built once per plan; the batch kernel only loads it.
struct RbGeom {
    int px, py, pz, bx, by, bz, per_y, total, dummy;
};

__global__ __launch_bounds__(256) void rb_xlist_kernel(RbGeom g, int *xl) {
    const int w = blockIdx.y, nwg = g.px * g.py * g.pz;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= RB_XSTRIDE) return;
    const int bx = g.bx, by = g.by, bz = g.bz;
    const int wx = w % g.px, wy = (w / g.px) % g.py, wz = w / (g.px * g.py);
    const int ox = wx * bx, oy = wy * by, oz = wz * bz;
    auto L = [](int i, int j, int k) { return (i + 1) + RB_TXP * (j + 1) + RB_PLANE * (k + 1); };
    auto has_sender = [&](int dx, int dy, int dz) {   // the neighbour in direction -D exists
        (void)dx;
        (void)dz;
        const int ny_ = wy - dy;
        return g.per_y || (ny_ >= 0 && ny_ < g.py);
    };
    int xpub = g.dummy, xdst = g.dummy | (18 << 20);
    if (e < g.total) {
        int q = 0, off = 0, ex = 0, ey = 0, ez = 0;
        for (; q < 18; ++q) {
            rb_extent(q, bx, by, bz, ex, ey, ez);
            if (e < off + ex * ey * ez / 2) break;
            off += ex * ey * ez / 2;
        }
        int dx, dy, dz;
        rb_dir(q, dx, dy, dz);
        const int h = q < 6 ? RB_H : 1;
        const int col = 1, k2 = 2 * (e - off);
        auto cell = [&](int r, auto side, int &cx, int &cy, int &cz) {
            cx = side(dx, r % ex, bx, h);
            cy = side(dy, (r / ex) % ey, by, h);
            cz = side(dz, r / (ex * ey), bz, h);
        };
        // sender side: the layer next to its face D; receiver side: the halo beyond face -D
        auto snd = [](int dd, int rr_, int bb, int hh) { return dd > 0 ? bb - hh + 1 + rr_ : 1 + rr_; };
        auto rcv = [](int dd, int rr_, int bb, int hh) {
            return dd < 0 ? bb + 1 + rr_ : (dd > 0 ? 1 - hh + rr_ : 1 + rr_);
        };
        auto colour = [&](int cx, int cy, int cz) { return ((ox + cx + oy + cy + oz + cz) & 1) ? 0 : 1; };
        int cx, cy, cz;
        cell(k2, snd, cx, cy, cz);
        if (colour(cx, cy, cz) != col) cell(k2 + 1, snd, cx, cy, cz);
        xpub = L(cx, cy, cz);
        if (has_sender(dx, dy, dz)) {
            cell(k2, rcv, cx, cy, cz);
            if (colour(cx, cy, cz) != col) cell(k2 + 1, rcv, cx, cy, cz);
            xdst = L(cx, cy, cz) | (q << 20) | (1 << 25);
        }
    }
    xl[(size_t)w * RB_XSTRIDE + e] = xpub;
    xl[((size_t)nwg + w) * RB_XSTRIDE + e] = xdst;
}

// ------------------------------------------------------------------------------- host side
static size_t rb_lds_bytes(int bx, int by, int bz, size_t tsize, bool ibm) {
    (void)bx;
    (void)by;
    const size_t tz = bz + 2 * RB_H;
    return ((size_t)RB_PLANE * tz + 2 * RB_PLANE + 3 + (ibm ? 6 * (size_t)RB_PLANE : 0)) * tsize;
}

template <typename T, bool IBM, int MAXP, int FORM = 0>
static const void *rb_fn() {
    return reinterpret_cast<const void *>(&rb_lds_kernel<T, IBM, MAXP, FORM>);
}

// pairs per thread: multiples of the group size RB_G; RB_COLS: the column form
template <typename T, bool IBM>
static const void *rb_fn_for(int maxp) {
    switch (maxp) {
    case RB_COLS: return rb_fn<T, IBM, RB_G, 1>();
    case RB_COLS2: return rb_fn<T, IBM, RB_G, 2>();
    case 3: return rb_fn<T, IBM, 3>();
    case 6: return rb_fn<T, IBM, 6>();
    case 9: return rb_fn<T, IBM, 9>();
    case 12: return rb_fn<T, IBM, 12>();
    default: return rb_fn<T, IBM, 15>();
    }
}

// threads per workgroup of a plan's kernel form
static int rb_threads(int maxp) { return maxp == RB_COLS2 ? 2 * RB_THREADS : RB_THREADS; }

static int rb_maxp_class(int need) {
    const int c = (need + RB_G - 1) / RB_G * RB_G;
    return c < RB_G ? RB_G : (c <= 15 ? c : 0);
}

// workgroups of this kernel variant that fit on one CU (registers, LDS, waves), from the runtime
static int rb_occupancy(bool fp32, bool ibm, int maxp, size_t lds) {
    const void *fn = fp32 ? (ibm ? rb_fn_for<float, true>(maxp) : rb_fn_for<float, false>(maxp))
                          : (ibm ? rb_fn_for<double, true>(maxp) : rb_fn_for<double, false>(maxp));
    if (lds > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, rb_threads(maxp), lds) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return nb;
}

// Choose the decomposition: every workgroup co-resident (runtime occupancy), the fewest cells
// per CU, then the smallest exchange.  `why` receives the reason when nothing fits.
bool rb_plan(const mdx_handle *h, bool fp32, RbPlanOut &out, std::string &why) {
    const int nx = h->d.nx, ny = h->d.ny, nz = h->d.nz;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) { why = "no device properties"; return false; }
    if (h->rb_coop && !prop.cooperativeLaunch) { why = "device reports no cooperative launch"; return false; }
    const int cus = prop.multiProcessorCount;
    const size_t lds_max = prop.maxSharedMemoryPerMultiProcessor > prop.sharedMemPerBlock
                               ? prop.maxSharedMemoryPerMultiProcessor : prop.sharedMemPerBlock;
    const size_t ts = fp32 ? 4 : 8;
    bool found = false;
    double best_cells = 1e300, best_x = 1e300;
    why = "no decomposition fits (cus=" + std::to_string(cus) + ", lds/block=" + std::to_string(prop.sharedMemPerBlock) +
          ", lds/cu=" + std::to_string(prop.maxSharedMemoryPerMultiProcessor) + ")";
    for (int px = 1; px <= nx; ++px) {
        if (nx % px || (nx / px) % 2) continue;
        for (int py = 1; py <= ny; ++py) {
            if (ny % py || (ny / py) % 2) continue;
            for (int pz = 1; pz <= nz; ++pz) {
                if (nz % pz || (nz / pz) % 2) continue;
                const int bx = nx / px, by = ny / py, bz = nz / pz;
                const int nwg = px * py * pz;
                // the fixed LDS row and plane; packed per-cell words: LDS index (dummy cell and its
                // padding included) < 2^16, coefficient-table index < 2^12
                if (bx + 2 * RB_H > RB_TXP || by + 2 * RB_H > RB_TYP) continue;
                if ((size_t)RB_PLANE * (bz + 2 * RB_H) + 2 * RB_PLANE + 3 >= 65536 || RB_PLANE >= 4096) continue;
                // the column form where a thread per column (subdomain and face ring) fits
                const bool colfit = bx * by + 2 * (bx + by) <= RB_THREADS && bx >= 4 && by >= 4 && bz >= 4;
                int maxp = rb_maxp_class((bx / 2 * by * bz + RB_THREADS - 1) / RB_THREADS);
                if (colfit && h->rb_form == 0 && bz <= 4 * RB_ZC2 - 4) maxp = RB_COLS2;
                if (colfit && h->rb_form == 2 && bz <= 2 * RB_ZC1 - 2) maxp = RB_COLS;
                if (!maxp) continue;
                // the column forms keep the face coefficients in registers: no LDS table
                const size_t lds = rb_lds_bytes(bx, by, bz, ts, h->ibm_on && maxp > 0);
                if (lds > lds_max) continue;
                if (rb_ring_total(bx, by, bz) > RB_MAXR * RB_THREADS) continue;
                const int total = rb_xchg_total(bx, by, bz);
                if (total > RB_HE * RB_THREADS) continue;
                if (nwg > 4 * cus) continue;
                const int per_cu = rb_occupancy(fp32, h->ibm_on, maxp, lds);
                if (per_cu < 1 || nwg > per_cu * cus) continue;
                const double rounds = std::ceil((double)nwg / cus);
                const double cells = rounds * bx * by * bz;
                if (cells < best_cells - 0.5 || (cells < best_cells + 0.5 && total < best_x)) {
                    best_cells = cells;
                    best_x = total;
                    out = RbPlanOut{px, py, pz, bx, by, bz, maxp, lds, total};
                    found = true;
                }
            }
        }
    }
    return found;
}

// the exchange lists of plan p into xl (2 * workgroups * RB_XSTRIDE ints), on the handle's stream
hipError_t rb_build_xlists(mdx_handle *h, const RbPlanOut &p, int *xl) {
    const int nwg = p.px * p.py * p.pz;
    const RbGeom g{p.px, p.py, p.pz, p.bx, p.by, p.bz, h->cfg.periodic[1] ? 1 : 0, p.total,
                   RB_PLANE * (p.bz + 2 * RB_H) + RB_PLANE + 1};
    hipLaunchKernelGGL(rb_xlist_kernel, dim3((RB_XSTRIDE + 255) / 256, nwg), dim3(256), 0, h->stream, g, xl);
    return hipGetLastError();
}

size_t rb_xlist_bytes(const RbPlanOut &p) { return 2 * (size_t)p.px * p.py * p.pz * RB_XSTRIDE * sizeof(int); }

size_t rb_faces_bytes(const RbPlanOut &p, bool fp32) {
    const size_t nwg = (size_t)p.px * p.py * p.pz;
    return nwg * 2 * (size_t)RB_XSTRIDE * 8 * (fp32 ? 1 : 2) + nwg * 16;   // faces, then the maxima
}

template <typename T>
static hipError_t rb_launch_t(mdx_handle *h, const RbPlanOut &pl, T *x, const T *rhs, int zero_init, int iters,
                              T *const *mu, const RbDefect *df) {
    RbArgs<T> a;
    a.defect = df != nullptr;
    a.g = grid_k(h);
    a.mu64.u = h->ibm_on ? h->mu64[0] : nullptr;
    a.mu64.v = h->ibm_on ? h->mu64[1] : nullptr;
    a.mu64.w = h->ibm_on ? h->mu64[2] : nullptr;
    a.mu64.plane = h->coef_plane ? 1 : 0;
    a.mu64.sx = h->d.sx;
    a.mu64.sy = h->d.sy;
    a.s64 = SorConst<double>{1.0 / h->dx * (1.0 / h->dx), 1.0 / h->dy * (1.0 / h->dy), 1.0 / h->dz * (1.0 / h->dz)};
    a.phi_in = df ? df->phi_in : nullptr;
    a.rhs64 = h->rhs;
    a.phi_out = df ? df->phi_out : nullptr;
    a.max_part = h->max_part;
    a.r_out = h->r32;
    a.red_max = h->red_max;
    a.host_slot = df ? df->host_slot : nullptr;
    a.seq = df ? df->seq : 0;
    a.skip = df ? df->skip : 0;
    a.skip_tol = df ? df->skip_tol : 0.0;
    a.ovl = h->rb_overlap ? 1 : 0;
    a.d = h->d;
    a.px = pl.px; a.py = pl.py; a.pz = pl.pz;
    a.bx = pl.bx; a.by = pl.by; a.bz = pl.bz;
    a.per_y = h->cfg.periodic[1] ? 1 : 0;
    const double ix2 = 1.0 / h->dx * (1.0 / h->dx), iy2 = 1.0 / h->dy * (1.0 / h->dy), iz2 = 1.0 / h->dz * (1.0 / h->dz);
    if (sizeof(T) == 4)
        a.s = SorConst<T>{(T)(1.0 / (h->dx * h->dx)), (T)(1.0 / (h->dy * h->dy)), (T)(1.0 / (h->dz * h->dz))};
    else
        a.s = SorConst<T>{(T)ix2, (T)iy2, (T)iz2};
    a.omega = (T)(sizeof(T) == 4 ? h->inner_omega : h->cfg.omega);
    a.mu_u = h->ibm_on ? mu[0] : nullptr;
    a.mu_v = h->ibm_on ? mu[1] : nullptr;
    a.mu_w = h->ibm_on ? mu[2] : nullptr;
    a.x = x;
    a.rhs = rhs;
    a.zero_init = zero_init;
    a.iters = iters;
    {
        const int hbx = pl.bx / 2, dpx = hbx > 2 ? hbx - 2 : 0, dpy = pl.by > 2 ? pl.by - 2 : 0,
                  dpz = pl.bz > 2 ? pl.bz - 2 : 0;
        a.md = (dpx * dpy * dpz / RB_THREADS) / RB_G * RB_G;   // whole relax groups only
    }
    a.faces = reinterpret_cast<unsigned long long *>(h->rb_faces);
    a.total = pl.total;
    a.xl = sizeof(T) == 4 ? h->rb_xl32 : h->rb_xl64;
    // tags of this launch: rb_tag + 1 .. rb_tag + iters + 1 (the last for the maxima); before the
    // 32-bit tag would wrap to 0 the face buffer is cleared and the count restarts
    if (h->rb_tag + (unsigned long long)iters + 2ull > 0xffffffffull) {
        hipError_t e = hipMemsetAsync(h->rb_faces, 0, h->rb_faces_bytes, h->stream);
        if (e != hipSuccess) return e;
        h->rb_tag = 0;
    }
    a.tag0 = (unsigned)h->rb_tag;
    h->rb_tag += (unsigned long long)iters + 1ull;
    a.err = h->rb_err;
    a.spin_limit = h->rb_spin_limit;
    const void *fn = h->ibm_on ? rb_fn_for<T, true>(pl.maxp) : rb_fn_for<T, false>(pl.maxp);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)pl.lds);
    if (e != hipSuccess) return e;
    void *args[] = {&a};
    // The plan holds at most as many workgroups as the CUs take at once (runtime occupancy), and a
    // dispatch waits for nothing but earlier work on this stream, so every workgroup becomes
    // resident with an ordinary launch; the cooperative launch gives the same guarantee formally
    // at about 30 us of extra dispatch time per launch.  Every neighbour wait is bounded and
    // raises the error word, so a workgroup that never arrived ends the batch with an error.
    if (h->rb_coop)
        return hipLaunchCooperativeKernel(fn, dim3(pl.px * pl.py * pl.pz), dim3(rb_threads(pl.maxp)), args,
                                          (unsigned)pl.lds, h->stream);
    return hipLaunchKernel(fn, dim3(pl.px * pl.py * pl.pz), dim3(rb_threads(pl.maxp)), args, (size_t)pl.lds,
                           h->stream);
}

hipError_t launch_rb_lds(mdx_handle *h, bool fp32, int iters, int zero_init, const RbDefect *df) {
    const RbPlanOut &pl = fp32 ? h->rb_plan32 : h->rb_plan64;
    if (fp32) return rb_launch_t<float>(h, pl, h->e32, h->r32, zero_init, iters, h->mu32, df);
    if (df) return hipErrorInvalidValue;   // the defect epilogue serves the FP32 batches only
    return rb_launch_t<double>(h, pl, h->phi, h->rhs, zero_init, iters, h->mu64, nullptr);
}

}  // namespace mdx

// ------------------------------------------------------------------------------- LDS read rate
// The chip's sustained LDS read bandwidth, the ceiling of the LDS-resident sweeps: every thread of
// 8 workgroups per CU reads 16-byte vectors of a 64 KB LDS array in a rotating, bank-conflict-free
// pattern (index masked into the array) and folds them, so nothing is optimised away.
namespace mdx {
static constexpr int LDSR_VEC = 4096, LDSR_ITERS = 4096;
__global__ __launch_bounds__(256) void lds_read_rate_kernel(float *out) {
    __shared__ float4 buf[LDSR_VEC];
    for (int q = threadIdx.x; q < LDSR_VEC; q += 256) buf[q] = make_float4((float)q, (float)q + 1.f, 0.f, 1.f);
    __syncthreads();
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    int idx = threadIdx.x;
#pragma unroll 8
    for (int i = 0; i < LDSR_ITERS; ++i) {
        const float4 v = buf[idx & (LDSR_VEC - 1)];
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
        idx += 256;
    }
    out[blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}
}  // namespace mdx

extern "C" int mdx_lds_read_rate(double *bytes_per_s, double *bytes_per_clock_per_cu) {
    using namespace mdx;
    if (!bytes_per_s) {
        mdx_set_error("mdx_lds_read_rate: null argument");
        return MDX_ERR_ARG;
    }
    int dev = 0;
    MDX_HIP(hipGetDevice(&dev));
    hipDeviceProp_t p;
    MDX_HIP(hipGetDeviceProperties(&p, dev));
    const int blocks = p.multiProcessorCount * 8;
    float *out = nullptr;
    MDX_HIP(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(lds_read_rate_kernel, dim3(blocks), dim3(256), 0, nullptr, out);   // warm-up
        e = hipEventRecord(e0, nullptr);
    }
    for (int r = 0; r < 10 && e == hipSuccess; ++r) {
        hipLaunchKernelGGL(lds_read_rate_kernel, dim3(blocks), dim3(256), 0, nullptr, out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(out);
    MDX_HIP(e);
    const double bytes = 10.0 * blocks * 256.0 * LDSR_ITERS * 16.0;
    *bytes_per_s = bytes / (ms * 1e-3);
    if (bytes_per_clock_per_cu) *bytes_per_clock_per_cu = *bytes_per_s / (p.clockRate * 1e3) / p.multiProcessorCount;
    return MDX_OK;
}
