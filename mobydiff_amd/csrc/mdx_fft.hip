// mdx_fft.hip -- FFT(x,z) + Thomas(y) pressure solve for gfx950 (FP64).
//
// Replaces build_fft_rhs + the FFTW/cuFFT 2-D r2c/c2r + solve_tridiagonal_y + scaling +
// apply_fft_pressure_correction (pressure_solver.f90:77-173, pressure_fft.f90:102-148,
// pressure_backend_fftw.f90:51-93) with five kernels:
//   1. divergence of u* fused with the real x-FFT of each (j,k) row          rhs -> hat
//   2. z-FFT of each (kx, j) line, TX lines per workgroup                    hat -> hat
//   3. Thomas forward elimination / back substitution per (kx,kz) line       hat -> hat
//   4. inverse z-FFT                                                         hat -> hat
//   5. inverse real x-FFT with the 1/(nx nz) scaling per row                 hat -> phi
//   6. projection u = (u* - dtg grad phi)/(1 + dtg coef), p += phi
// Transforms are radix-4 Stockham (self-sorting) in LDS with a precomputed FP64 twiddle table;
// hat keeps the reference's half-spectrum layout plane_hat(nxh, nz, ny), kx fastest.
#include "mdx_internal.h"

namespace mdx {

// complex element types of the transforms: FP64 for the plain solve, FP32 for the defect solves
template <typename C> struct Cx;
template <> struct Cx<double2> {
    using R = double;
    static __device__ inline double2 mk(double a, double b) { return make_double2(a, b); }
};
template <> struct Cx<float2> {
    using R = float;
    static __device__ inline float2 mk(float a, float b) { return make_float2(a, b); }
};

// Stockham (self-sorting) transform of LINES lines of length n = 2^p held in LDS (a, b ping-pong),
// line q at q * ls: one radix-2 pass when p is odd, then radix-4 passes.  A pass combining groups
// of ns points reads r inputs n/r apart and writes them r-interleaved with twiddles
// w^(q kk n / (r ns)), kk = position in the group.  Returns the buffer holding the result.
// tw[m] = exp(-2 pi i m / n), m < n/2, staged in LDS by the caller (stage_twiddles); larger
// exponents use w^(m + n/2) = -w^m.
template <typename C>
__device__ inline C *stockham(C *a, C *b, int n, int lines, int ls, const C *tw, bool inverse, int tid,
                              int nthr) {
    using R = typename Cx<C>::R;
    const int half = n >> 1, quarter = n >> 2;
    int ns = 1;
    if (__builtin_ctz(n) & 1) {   // radix-2 pass
        for (int t = tid; t < half * lines; t += nthr) {
            const int line = t / half, jj = t - line * half;
            const C x0 = a[line * ls + jj], x1 = a[line * ls + jj + half];
            b[line * ls + 2 * jj] = Cx<C>::mk(x0.x + x1.x, x0.y + x1.y);
            b[line * ls + 2 * jj + 1] = Cx<C>::mk(x0.x - x1.x, x0.y - x1.y);
        }
        __syncthreads();
        C *t2 = a;
        a = b;
        b = t2;
        ns = 2;
    }
    auto twid = [&](int m) {   // w^m for 0 <= m < n
        C w = m < half ? tw[m] : Cx<C>::mk(-tw[m - half].x, -tw[m - half].y);
        if (inverse) w.y = -w.y;
        return w;
    };
    for (; ns < n; ns <<= 2) {
        const int stride = n / (4 * ns);
        for (int t = tid; t < quarter * lines; t += nthr) {
            const int line = t / quarter, jj = t - line * quarter;
            const int kk = jj & (ns - 1);
            const C *src = a + line * ls + jj;
            const C x0 = src[0], x1 = src[quarter], x2 = src[2 * quarter], x3 = src[3 * quarter];
            const C w1 = twid(kk * stride), w2 = twid(2 * kk * stride), w3 = twid(3 * kk * stride);
            const R y1r = x1.x * w1.x - x1.y * w1.y, y1i = x1.x * w1.y + x1.y * w1.x;
            const R y2r = x2.x * w2.x - x2.y * w2.y, y2i = x2.x * w2.y + x2.y * w2.x;
            const R y3r = x3.x * w3.x - x3.y * w3.y, y3i = x3.x * w3.y + x3.y * w3.x;
            // 4-point DFT; the forward transform multiplies by (-i)^(pq), the inverse by i^(pq)
            const R s02r = x0.x + y2r, s02i = x0.y + y2i, d02r = x0.x - y2r, d02i = x0.y - y2i;
            const R s13r = y1r + y3r, s13i = y1i + y3i, d13r = y1r - y3r, d13i = y1i - y3i;
            const R jr = inverse ? -d13i : d13i, ji = inverse ? d13r : -d13r;   // -i (or +i) * d13
            C *dst = b + line * ls + ((jj - kk) << 2) + kk;
            dst[0] = Cx<C>::mk(s02r + s13r, s02i + s13i);
            dst[ns] = Cx<C>::mk(d02r + jr, d02i + ji);
            dst[2 * ns] = Cx<C>::mk(s02r - s13r, s02i - s13i);
            dst[3 * ns] = Cx<C>::mk(d02r - jr, d02i - ji);
        }
        __syncthreads();
        C *t2 = a;
        a = b;
        b = t2;
    }
    return a;
}

static constexpr int FFT_THREADS = 256;

// the n/2 twiddles into LDS (read by every butterfly of every pass)
template <typename C>
__device__ inline void stage_twiddles(C *dst, const C *__restrict__ tw, int n) {
    for (int t = threadIdx.x; t < n / 2; t += FFT_THREADS) dst[t] = tw[t];
}

// 1. rhs(i,k,j) = [(u*(i+1)-u*(i))/dx + (v*(j+1)-v*(j))/dy + (w*(k+1)-w*(k))/dz]/dtg, then the
//    r2c x-transform; rows r = (k-1) + nz (j-1) are contiguous in hat.
template <bool STRETCH>
__global__ __launch_bounds__(FFT_THREADS) void rhs_xfft_kernel(GridK g, const double *__restrict__ us,
                                                              const double *__restrict__ vs,
                                                              const double *__restrict__ ws, double dtg,
                                                              const double2 *__restrict__ tw, int lines,
                                                              double2 *hat) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double2 *lds = reinterpret_cast<double2 *>(lds_raw);
    const Dims d = g.d;
    const int n = d.nx, nxh = n / 2 + 1;
    const int nrows = d.ny * d.nz;
    double2 *a = lds, *b = lds + (size_t)lines * n, *twl = b + (size_t)lines * n;
    stage_twiddles(twl, tw, n);
    for (int t = threadIdx.x; t < lines * n; t += FFT_THREADS) {
        const int line = t / n, i = t - line * n + 1;
        const int r = blockIdx.x * lines + line;
        double v = 0.0;
        if (r < nrows) {
            const int k = r % d.nz + 1, j = r / d.nz + 1;
            const double dv = STRETCH ? (vs[d.g(i, j + 1, k)] - vs[d.g(i, j, k)]) * g.ihc[j]
                                      : (vs[d.g(i, j + 1, k)] - vs[d.g(i, j, k)]) / g.dy;
            v = ((us[d.g(i + 1, j, k)] - us[d.g(i, j, k)]) / g.dx + dv +
                 (ws[d.g(i, j, k + 1)] - ws[d.g(i, j, k)]) / g.dz) / dtg;
        }
        a[t] = make_double2(v, 0.0);
    }
    __syncthreads();
    double2 *res = stockham(a, b, n, lines, n, twl, false, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < lines * nxh; t += FFT_THREADS) {
        const int line = t / nxh, m = t - line * nxh;
        const int r = blockIdx.x * lines + line;
        if (r < nrows) hat[(size_t)nxh * r + m] = res[line * n + m];
    }
}

// 2./4. z-transform of TX consecutive kx columns of one j plane.
template <typename C>
__global__ __launch_bounds__(FFT_THREADS) void zfft_kernel(Dims d, C *hat, const C *__restrict__ tw, int tx,
                                                           int inverse) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *lds = reinterpret_cast<C *>(lds_raw);
    const int nz = d.nz, nxh = d.nx / 2 + 1;
    const int tiles = (nxh + tx - 1) / tx;
    const int j = blockIdx.x / tiles, kx0 = (blockIdx.x % tiles) * tx;
    // columns nz + 1 apart in LDS: the transposing loads and stores (tx columns of one kz row)
    // fall in different banks
    const int ls = nz + 1;
    C *a = lds, *b = lds + (size_t)tx * ls, *twl = b + (size_t)tx * ls;
    C *plane = hat + (size_t)nxh * nz * j;
    stage_twiddles(twl, tw, nz);
    for (int t = threadIdx.x; t < tx * nz; t += FFT_THREADS) {
        const int c = t % tx, kz = t / tx;
        const int kx = kx0 + c;
        a[c * ls + kz] = kx < nxh ? plane[(size_t)kx + (size_t)nxh * kz] : Cx<C>::mk(0, 0);
    }
    __syncthreads();
    C *res = stockham(a, b, nz, tx, ls, twl, inverse != 0, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < tx * nz; t += FFT_THREADS) {
        const int c = t % tx, kz = t / tx;
        const int kx = kx0 + c;
        if (kx < nxh) plane[(size_t)kx + (size_t)nxh * kz] = res[c * ls + kz];
    }
}

// 3. pressure_fft.f90:102-148 with precomputed cp/den_inv (pressure_fft.f90:20-100); one thread
//    per (kx,kz) line, consecutive kx in consecutive lanes (coalesced), in place.  The recurrences
//    run in blocks of TB planes whose loads are all issued before the block's stores (a load after
//    a store through the same pointer cannot be hoisted above it), so TB loads are in flight at a
//    time instead of one; the arithmetic and its order are those of the plane-by-plane loops.
static constexpr int TB = 8;

__global__ __launch_bounds__(64) void tridiag_kernel(Dims d, double2 *hat, const double *__restrict__ cp,
                                                     const double *__restrict__ den_inv,
                                                     const double *__restrict__ arow) {
    const int nxh = d.nx / 2 + 1, ny = d.ny, nz = d.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxh * nz) return;
    const int kx = t % nxh, kz = t / nxh;
    const size_t hs = (size_t)nxh * nz;                     // stride between j planes in hat
    const size_t cs = (size_t)nxh;                          // stride between j in cp/den_inv
    const size_t h0 = (size_t)kx + (size_t)nxh * kz;
    const size_t c0 = (size_t)kx + (size_t)nxh * (size_t)ny * kz;
    double2 dp;
    if (kx == 0 && kz == 0) {
        dp = make_double2(0.0, 0.0);
    } else {
        const double2 r = hat[h0];
        const double di = den_inv[c0];
        dp = make_double2(r.x * di, r.y * di);
    }
    hat[h0] = dp;
    // forward elimination, planes 2..ny
    for (int jb = 2; jb <= ny; jb += TB) {
        double2 rr[TB];
        double aa[TB], dd[TB];
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb + q;
            if (jj <= ny) {
                rr[q] = hat[h0 + hs * (jj - 1)];
                aa[q] = arow[jj - 1];
                dd[q] = den_inv[c0 + cs * (jj - 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb + q;
            if (jj <= ny) {
                dp = make_double2((rr[q].x - aa[q] * dp.x) * dd[q], (rr[q].y - aa[q] * dp.y) * dd[q]);
                hat[h0 + hs * (jj - 1)] = dp;
            }
        }
    }
    // back substitution, planes ny-1..1 (hat(ny) = dp(ny))
    double2 nxt = dp;
    for (int jb = ny - 1; jb >= 1; jb -= TB) {
        double2 qq[TB];
        double cc[TB];
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb - q;
            if (jj >= 1) {
                qq[q] = hat[h0 + hs * (jj - 1)];
                cc[q] = cp[c0 + cs * (jj - 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb - q;
            if (jj >= 1) {
                nxt = make_double2(qq[q].x - cc[q] * nxt.x, qq[q].y - cc[q] * nxt.y);
                hat[h0 + hs * (jj - 1)] = nxt;
            }
        }
    }
}

// 3'. The same tridiagonal solves with one wavefront per (kx,kz) line, every value of the line in
//     registers: the spectrum is read once and written once (16 B per cell), against twice each
//     plus the two FP64 tables for the thread-per-line recurrence.  Lane l holds rows
//     j = l, l + 64, ..., l + 64 (M-1) (0-based) with a_j (sub-diagonal), b_j = bd_j - k
//     (k = lamx[kx] + lamz[kz], the modified wavenumbers of pressure_fft.f90:20-100), c_j and the
//     complex right-hand side; rows past ny, and neighbours past the ends, are identity rows.
//       * Parallel cyclic reduction, six steps (coupling distance 1, 2, ..., 32, the neighbours'
//         coefficients by cross-lane shuffles), leaves every row coupled only to the rows 64 away:
//         each lane then holds an independent M-row system, solved by Thomas.
//       * One step of iterative refinement in the same precision: the residual d - A x with the
//         original rows (neighbours x_(j-1), x_(j+1) by shuffles) goes through the same reduction,
//         whose multipliers and Thomas factors were kept in registers, and the correction is added.
//     Reduction alone left residuals 100-1000x larger than the serial recurrence's on the
//     tanh-stretched grid (weakly dominant low-wavenumber rows); with the refinement step they are
//     within 2x of it.  Every lane works through the whole solve (no serial chain across lanes:
//     a line's serial recurrence in one wave would cost 64 lanes' worth of FP64 issue per step).
//     The mean mode (kx = kz = 0) pins row 1 to zero, as the reference does (dp(1) = 0), by making
//     it an identity row.  A workgroup takes 16 consecutive lines and stages them through LDS: for
//     a fixed row the 16 lines' values are contiguous in the spectrum (256 B in FP64), so the loads
//     and stores move whole cache lines, and each wave then reads its line's rows from LDS
//     (rows of a line contiguous, padded by one value so the transposed copy is conflict-free).
template <typename T>
__device__ inline T shfl_any(T v, int src) {
    return __shfl(v, src, 64);
}
template <>
__device__ inline double2 shfl_any<double2>(double2 v, int src) {
    return make_double2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}
template <>
__device__ inline float2 shfl_any<float2>(float2 v, int src) {
    return make_float2(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}

// a workgroup of TW_WAVES waves solves TW_LINES consecutive lines, staged through LDS
static constexpr int TW_WAVES = 4, TW_LINES = 16, TW_STEPS = 6;

// the value of row i - delta (lo) or i + delta (hi) for row set q, from the shuffled copies of all
// row sets (rows across the lane wrap belong to set q - 1 / q + 1; beyond the ends: `edge`)
template <int M, typename V>
__device__ inline V pick_lo(const V (&sh)[M], int q, bool wrap, V edge) {
    return !wrap ? sh[q] : (q > 0 ? sh[q > 0 ? q - 1 : 0] : edge);
}
template <int M, typename V>
__device__ inline V pick_hi(const V (&sh)[M], int q, bool wrap, V edge) {
    return !wrap ? sh[q] : (q + 1 < M ? sh[q + 1 < M ? q + 1 : q] : edge);
}

// C: double2 (the FP64 solve; the mean mode pins row 1) or float2 (the FP32 defect solves; the
// mean mode is solved apart in FP64, so its line is zeroed, MEAN_ZERO)
template <typename C, int M, bool MEAN_ZERO>
__global__ __launch_bounds__(64 * TW_WAVES) void tridiag_wave_kernel(int nlines, int nxh, int ny, size_t plane, C *hat,
                                                                    const typename Cx<C>::R *__restrict__ lamx,
                                                                    const typename Cx<C>::R *__restrict__ lamz,
                                                                    const typename Cx<C>::R *__restrict__ arow,
                                                                    const typename Cx<C>::R *__restrict__ bdrow,
                                                                    const typename Cx<C>::R *__restrict__ crow) {
    using R = typename Cx<C>::R;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *tile = reinterpret_cast<C *>(lds_raw);   // [TW_LINES][ny + 1]: line-major, rows contiguous
    const int ldt = ny + 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int t0 = blockIdx.x * TW_LINES;
    const C zero2 = Cx<C>::mk(R(0), R(0));
    // ---- the TW_LINES lines of this workgroup into LDS: for each row, the lines' values are
    // consecutive in the spectrum (TW_LINES * sizeof(C) contiguous bytes per row)
    {
        const int tt = threadIdx.x % TW_LINES, jj = threadIdx.x / TW_LINES;
        const int t = t0 + tt;
        for (int j = jj; j < ny; j += 64 * TW_WAVES / TW_LINES)
            tile[tt * ldt + j] = t < nlines ? hat[(size_t)t + plane * j] : zero2;
    }
    __syncthreads();
    // ---- every wave solves its lines from LDS (lane l: rows l, l + 64, ...)
    for (int l = wave; l < TW_LINES; l += TW_WAVES) {
        const int t = t0 + l;
        if (t >= nlines) break;   // whole waves
        C *line = tile + (size_t)l * ldt;
        if (MEAN_ZERO && t == 0) {
            for (int j = lane; j < ny; j += 64) line[j] = zero2;
            continue;
        }
        const int kx = t % nxh, kz = t / nxh;
        const R k = lamx[kx] + lamz[kz];
        // the original rows (kept for the residual) and the right-hand side
        R a0[M], b0[M], c0[M];
        C d0[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = lane + 64 * q;
            if (j < ny && !(t == 0 && j == 0)) {
                a0[q] = arow[j];
                c0[q] = crow[j];
                b0[q] = bdrow[j] - k;
                d0[q] = line[j];
            } else {
                a0[q] = R(0);
                c0[q] = R(0);
                b0[q] = R(1);
                d0[q] = zero2;
            }
        }
        // reduction of the rows, keeping the multipliers of every step
        R a[M], b[M], c[M], k1s[TW_STEPS][M], k2s[TW_STEPS][M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            a[q] = a0[q];
            b[q] = b0[q];
            c[q] = c0[q];
        }
#pragma unroll
        for (int s = 0; s < TW_STEPS; ++s) {
            const int delta = 1 << s;
            const int lo = (lane - delta) & 63, hi = (lane + delta) & 63;
            const bool wlo = lane < delta, whi = lane + delta >= 64;
            R rb[M], la[M], lc[M], lr[M], ha[M], hc[M], hr[M];
#pragma unroll
            for (int q = 0; q < M; ++q) rb[q] = R(1) / b[q];
#pragma unroll
            for (int q = 0; q < M; ++q) {
                la[q] = shfl_any(a[q], lo);
                lc[q] = shfl_any(c[q], lo);
                lr[q] = shfl_any(rb[q], lo);
                ha[q] = shfl_any(a[q], hi);
                hc[q] = shfl_any(c[q], hi);
                hr[q] = shfl_any(rb[q], hi);
            }
#pragma unroll
            for (int q = 0; q < M; ++q) {
                const R na = pick_lo<M>(la, q, wlo, R(0)), nc = pick_lo<M>(lc, q, wlo, R(0)),
                        nr = pick_lo<M>(lr, q, wlo, R(1));
                const R pa = pick_hi<M>(ha, q, whi, R(0)), pc = pick_hi<M>(hc, q, whi, R(0)),
                        pr = pick_hi<M>(hr, q, whi, R(1));
                const R k1 = a[q] * nr, k2 = c[q] * pr;
                k1s[s][q] = k1;
                k2s[s][q] = k2;
                a[q] = -na * k1;
                b[q] = b[q] - nc * k1 - pa * k2;
                c[q] = -pc * k2;
            }
        }
        // Thomas factors of the lane's M-row system (row q couples to rows q - 1 and q + 1 only)
        R cp[M], rden[M];
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const R den = q > 0 ? b[q] - a[q] * cp[q - 1] : b[q];
            rden[q] = R(1) / den;
            cp[q] = c[q] * rden[q];
        }
        // a right-hand side through the kept reduction and Thomas factors
        auto solve = [&](C (&d)[M]) {
#pragma unroll
            for (int s = 0; s < TW_STEPS; ++s) {
                const int delta = 1 << s;
                const int lo = (lane - delta) & 63, hi = (lane + delta) & 63;
                const bool wlo = lane < delta, whi = lane + delta >= 64;
                C ld[M], hd[M];
#pragma unroll
                for (int q = 0; q < M; ++q) {
                    ld[q] = shfl_any(d[q], lo);
                    hd[q] = shfl_any(d[q], hi);
                }
#pragma unroll
                for (int q = 0; q < M; ++q) {
                    const C nd = pick_lo<M>(ld, q, wlo, zero2), pd = pick_hi<M>(hd, q, whi, zero2);
                    d[q] = Cx<C>::mk(d[q].x - nd.x * k1s[s][q] - pd.x * k2s[s][q],
                                     d[q].y - nd.y * k1s[s][q] - pd.y * k2s[s][q]);
                }
            }
#pragma unroll
            for (int q = 0; q < M; ++q) {
                if (q > 0) d[q] = Cx<C>::mk(d[q].x - a[q] * d[q - 1].x, d[q].y - a[q] * d[q - 1].y);
                d[q] = Cx<C>::mk(d[q].x * rden[q], d[q].y * rden[q]);
            }
#pragma unroll
            for (int q = M - 2; q >= 0; --q)
                d[q] = Cx<C>::mk(d[q].x - cp[q] * d[q + 1].x, d[q].y - cp[q] * d[q + 1].y);
        };
        C x[M];
#pragma unroll
        for (int q = 0; q < M; ++q) x[q] = d0[q];
        solve(x);
        // refinement: r = d - A x with the original rows, solved the same way, added
        {
            const int lo = (lane - 1) & 63, hi = (lane + 1) & 63;
            const bool wlo = lane < 1, whi = lane + 1 >= 64;
            C lx[M], hx[M], r[M];
#pragma unroll
            for (int q = 0; q < M; ++q) {
                lx[q] = shfl_any(x[q], lo);
                hx[q] = shfl_any(x[q], hi);
            }
#pragma unroll
            for (int q = 0; q < M; ++q) {
                const C xm = pick_lo<M>(lx, q, wlo, zero2), xp = pick_hi<M>(hx, q, whi, zero2);
                r[q] = Cx<C>::mk(d0[q].x - (a0[q] * xm.x + b0[q] * x[q].x + c0[q] * xp.x),
                                 d0[q].y - (a0[q] * xm.y + b0[q] * x[q].y + c0[q] * xp.y));
            }
            solve(r);
#pragma unroll
            for (int q = 0; q < M; ++q) x[q] = Cx<C>::mk(x[q].x + r[q].x, x[q].y + r[q].y);
        }
#pragma unroll
        for (int q = 0; q < M; ++q) {
            const int j = lane + 64 * q;
            if (j < ny) line[j] = x[q];
        }
    }
    __syncthreads();
    // ---- back to the spectrum, as loaded
    {
        const int tt = threadIdx.x % TW_LINES, jj = threadIdx.x / TW_LINES;
        const int t = t0 + tt;
        if (t < nlines)
            for (int j = jj; j < ny; j += 64 * TW_WAVES / TW_LINES) hat[(size_t)t + plane * j] = tile[tt * ldt + j];
    }
}

// 3''. The same solves as a partitioned Thomas recurrence, for ny >= 64.  A workgroup of 256 threads
//     stages 16 consecutive lines through LDS (as above) and gives each line 16 threads, one per
//     segment of M consecutive rows (rows past ny are identity rows).  A segment's interior rows
//     (1 .. M-2) are eliminated by the Thomas recurrence with three right-hand sides: the data, and
//     the couplings to the segment's first row F (through a_1) and last row E (through c_(M-2)), so
//     every interior value is p + q F + s E.  Substituted into the first and last rows, that leaves
//     a tridiagonal system in F_0, E_0, F_1, E_1, ... (32 unknowns per line), which one thread per
//     line solves by the Thomas recurrence; the segments then form their interior values.  The
//     serial chain per line is 2 (M - 2) + 64 steps instead of 2 ny, the work about twice the plain
//     recurrence's, and the data stay in LDS: the kernel moves the spectrum once each way.  The
//     mean mode pins row 1 as the wave solve does; MEAN_ZERO zeroes that line.
static constexpr int TP_LINES = 16, TP_SEGS = 16;
template <typename C, int M, bool MEAN_ZERO>
__global__ __launch_bounds__(TP_LINES * TP_SEGS) void tridiag_part_kernel(int nlines, int nxh, int ny, size_t plane,
                                                                        C *hat,
                                                                        const typename Cx<C>::R *__restrict__ lamx,
                                                                        const typename Cx<C>::R *__restrict__ lamz,
                                                                        const typename Cx<C>::R *__restrict__ arow,
                                                                        const typename Cx<C>::R *__restrict__ bdrow,
                                                                        const typename Cx<C>::R *__restrict__ crow) {
    using R = typename Cx<C>::R;
    constexpr int NR = 2 * TP_SEGS;   // reduced rows per line
    extern __shared__ __align__(16) unsigned char lds_raw[];
    C *tile = reinterpret_cast<C *>(lds_raw);   // [TP_LINES][ny + 1]
    __shared__ R ra[TP_LINES][NR], rb[TP_LINES][NR], rc[TP_LINES][NR];
    __shared__ C rd[TP_LINES][NR];
    const int ldt = ny + 1;
    const int tt = threadIdx.x % TP_LINES, sg = threadIdx.x / TP_LINES;
    const int t = blockIdx.x * TP_LINES + tt;
    const C zero2 = Cx<C>::mk(R(0), R(0));
    for (int j = sg; j < ny; j += TP_SEGS) tile[tt * ldt + j] = t < nlines ? hat[(size_t)t + plane * j] : zero2;
    __syncthreads();
    C *line = tile + (size_t)tt * ldt;
    R k = R(0);
    if (t < nlines) k = lamx[t % nxh] + lamz[t / nxh];
    const bool ident_line = t >= nlines || (MEAN_ZERO && t == 0);
    const int j0 = sg * M;
    // row r of the segment: (a, b, c); the data row is line[j0 + r]
    auto coef = [&](int r, R &a, R &b, R &c) {
        const int j = j0 + r;
        if (!ident_line && j < ny && !(t == 0 && j == 0)) {
            a = arow[j];
            b = bdrow[j] - k;
            c = crow[j];
        } else {
            a = R(0);
            b = R(1);
            c = R(0);
        }
    };
    auto data = [&](int r) {
        const int j = j0 + r;
        return (!ident_line && j < ny && !(t == 0 && j == 0)) ? line[j] : zero2;
    };
    // ---- interior rows 1 .. M-2: forward elimination with the three right-hand sides; the data's
    // eliminated values go back into LDS, the couplings' stay in registers
    R cp[M], qq[M], ss[M];
#pragma unroll
    for (int r = 0; r < M; ++r) cp[r] = qq[r] = ss[r] = R(0);
    R a0, b0, c0, aL, bL, cL;
    coef(0, a0, b0, c0);
    coef(M - 1, aL, bL, cL);
    const C d0 = data(0), dL = data(M - 1);
    C pprev = zero2;
#pragma unroll
    for (int r = 1; r <= M - 2; ++r) {
        R a, b, c;
        coef(r, a, b, c);
        const C d = data(r);
        const R den = r > 1 ? b - a * cp[r - 1] : b;
        const R rden = R(1) / den;
        cp[r] = r < M - 2 ? c * rden : R(0);
        const C p = r > 1 ? Cx<C>::mk((d.x - a * pprev.x) * rden, (d.y - a * pprev.y) * rden)
                          : Cx<C>::mk(d.x * rden, d.y * rden);
        qq[r] = (r == 1 ? -a : -a * qq[r - 1]) * rden;
        ss[r] = ((r == M - 2 ? -c : R(0)) - (r > 1 ? a * ss[r - 1] : R(0))) * rden;
        if (j0 + r < ny) line[j0 + r] = p;
        pprev = p;
    }
    // back substitution of the three right-hand sides
    C pnext = zero2;
#pragma unroll
    for (int r = M - 2; r >= 1; --r) {
        C p = j0 + r < ny ? line[j0 + r] : zero2;
        if (r < M - 2) {
            p = Cx<C>::mk(p.x - cp[r] * pnext.x, p.y - cp[r] * pnext.y);
            qq[r] -= cp[r] * qq[r + 1];
            ss[r] -= cp[r] * ss[r + 1];
        }
        if (j0 + r < ny) line[j0 + r] = p;
        pnext = p;
    }
    // ---- the segment's two rows of the reduced system: F (row 0) and E (row M-1)
    {
        const C p1 = M > 2 ? (j0 + 1 < ny ? line[j0 + 1] : zero2) : zero2;
        const C pl = M > 2 ? (j0 + M - 2 < ny ? line[j0 + M - 2] : zero2) : zero2;
        if (M > 2) {
            ra[tt][2 * sg] = a0;
            rb[tt][2 * sg] = b0 + c0 * qq[1];
            rc[tt][2 * sg] = c0 * ss[1];
            rd[tt][2 * sg] = Cx<C>::mk(d0.x - c0 * p1.x, d0.y - c0 * p1.y);
            ra[tt][2 * sg + 1] = aL * qq[M - 2];
            rb[tt][2 * sg + 1] = bL + aL * ss[M - 2];
            rc[tt][2 * sg + 1] = cL;
            rd[tt][2 * sg + 1] = Cx<C>::mk(dL.x - aL * pl.x, dL.y - aL * pl.y);
        } else {
            ra[tt][2 * sg] = a0;
            rb[tt][2 * sg] = b0;
            rc[tt][2 * sg] = c0;
            rd[tt][2 * sg] = d0;
            ra[tt][2 * sg + 1] = aL;
            rb[tt][2 * sg + 1] = bL;
            rc[tt][2 * sg + 1] = cL;
            rd[tt][2 * sg + 1] = dL;
        }
    }
    __syncthreads();
    // ---- the reduced system of each line, one thread per line (Thomas; solution into rd)
    if (sg == 0) {
        R cprev = R(0);
        C dprev = zero2;
        for (int r = 0; r < NR; ++r) {
            const R a = ra[tt][r], den = rb[tt][r] - a * cprev, rden = R(1) / den;
            const C d = rd[tt][r];
            cprev = rc[tt][r] * rden;
            dprev = Cx<C>::mk((d.x - a * dprev.x) * rden, (d.y - a * dprev.y) * rden);
            rc[tt][r] = cprev;
            rd[tt][r] = dprev;
        }
        C xn = rd[tt][NR - 1];
        for (int r = NR - 2; r >= 0; --r) {
            const C d = rd[tt][r];
            xn = Cx<C>::mk(d.x - rc[tt][r] * xn.x, d.y - rc[tt][r] * xn.y);
            rd[tt][r] = xn;
        }
    }
    __syncthreads();
    // ---- every row of the segment from F and E
    {
        const C F = rd[tt][2 * sg], E = rd[tt][2 * sg + 1];
        if (j0 < ny) line[j0] = F;
#pragma unroll
        for (int r = 1; r <= M - 2; ++r) {
            if (j0 + r < ny) {
                const C p = line[j0 + r];
                line[j0 + r] = Cx<C>::mk(p.x + qq[r] * F.x + ss[r] * E.x, p.y + qq[r] * F.y + ss[r] * E.y);
            }
        }
        if (j0 + M - 1 < ny) line[j0 + M - 1] = E;
    }
    __syncthreads();
    if (t < nlines)
        for (int j = sg; j < ny; j += TP_SEGS) hat[(size_t)t + plane * j] = tile[tt * ldt + j];
}

// launches the partitioned solve for 64 <= ny <= 256 (false: the caller takes another solver)
template <typename C, bool MEAN_ZERO>
static bool launch_tridiag_part(hipStream_t st, const Dims &d, C *hat, const typename Cx<C>::R *lamx,
                                const typename Cx<C>::R *lamz, const typename Cx<C>::R *a,
                                const typename Cx<C>::R *bd, const typename Cx<C>::R *c) {
    if (d.ny < 64 || d.ny > 256) return false;
    const int nxh = d.nx / 2 + 1, nlines = nxh * d.nz, m = (d.ny + TP_SEGS - 1) / TP_SEGS;
    const size_t plane = (size_t)nxh * d.nz;
    const dim3 grid((nlines + TP_LINES - 1) / TP_LINES), blk(TP_LINES * TP_SEGS);
    const size_t lds = (size_t)TP_LINES * (d.ny + 1) * sizeof(C);   // <= 66 KB
#define TP_CASE(MM)                                                                                            \
    if (m <= MM) {                                                                                             \
        static const hipError_t attr = hipFuncSetAttribute(                                                    \
            reinterpret_cast<const void *>(&tridiag_part_kernel<C, MM, MEAN_ZERO>),                            \
            hipFuncAttributeMaxDynamicSharedMemorySize, 16 * 257 * (int)sizeof(C));                            \
        if (attr != hipSuccess) return false;                                                                  \
        hipLaunchKernelGGL((tridiag_part_kernel<C, MM, MEAN_ZERO>), grid, blk, lds, st, nlines, nxh, d.ny, plane, \
                           hat, lamx, lamz, a, bd, c);                                                         \
        return true;                                                                                           \
    }
    TP_CASE(4) TP_CASE(8) TP_CASE(12) TP_CASE(16)
#undef TP_CASE
    return false;
}

// launches the one-wave-per-line solve for ny <= 256 (false: the caller runs the Thomas kernel)
template <typename C, bool MEAN_ZERO>
static bool launch_tridiag_wave(hipStream_t st, const Dims &d, C *hat, const typename Cx<C>::R *lamx,
                                const typename Cx<C>::R *lamz, const typename Cx<C>::R *a,
                                const typename Cx<C>::R *bd, const typename Cx<C>::R *c) {
    const int nxh = d.nx / 2 + 1, nlines = nxh * d.nz, m = (d.ny + 63) / 64;
    const size_t plane = (size_t)nxh * d.nz;
    const dim3 grid((nlines + TW_LINES - 1) / TW_LINES), blk(64 * TW_WAVES);
    const size_t lds = (size_t)TW_LINES * (d.ny + 1) * sizeof(C);   // <= 66 KB (ny <= 256)
#define TW_CASE(MM)                                                                                            \
    case MM: {                                                                                                 \
        static const hipError_t attr = hipFuncSetAttribute(                                                    \
            reinterpret_cast<const void *>(&tridiag_wave_kernel<C, MM, MEAN_ZERO>),                            \
            hipFuncAttributeMaxDynamicSharedMemorySize, 16 * 257 * (int)sizeof(C));                            \
        if (attr != hipSuccess) return false;                                                                  \
        hipLaunchKernelGGL((tridiag_wave_kernel<C, MM, MEAN_ZERO>), grid, blk, lds, st, nlines, nxh, d.ny, plane, \
                           hat, lamx, lamz, a, bd, c);                                                         \
        return true;                                                                                           \
    }
    switch (m) {
    TW_CASE(1) TW_CASE(2) TW_CASE(3) TW_CASE(4)
    default: return false;
    }
#undef TW_CASE
}

// 5. c2r along x: the real parts of the kx=0 and kx=nx/2 coefficients, Hermitian extension of the
//    rest, inverse transform, times 1/(nx nz); phi in rhs layout (nx, nz, ny).
__global__ __launch_bounds__(FFT_THREADS) void ixfft_kernel(Dims d, const double2 *__restrict__ hat,
                                                           const double2 *__restrict__ tw, int lines,
                                                           double scale, double *phi) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    double2 *lds = reinterpret_cast<double2 *>(lds_raw);
    const int n = d.nx, nxh = n / 2 + 1;
    const int nrows = d.ny * d.nz;
    double2 *a = lds, *b = lds + (size_t)lines * n, *twl = b + (size_t)lines * n;
    stage_twiddles(twl, tw, n);
    for (int t = threadIdx.x; t < lines * n; t += FFT_THREADS) {
        const int line = t / n, m = t - line * n;
        const int r = blockIdx.x * lines + line;
        double2 v = make_double2(0.0, 0.0);
        if (r < nrows) {
            const double2 *row = hat + (size_t)nxh * r;
            if (m == 0 || 2 * m == n) v = make_double2(row[m].x, 0.0);
            else if (m < nxh) v = row[m];
            else v = make_double2(row[n - m].x, -row[n - m].y);
        }
        a[t] = v;
    }
    __syncthreads();
    double2 *res = stockham(a, b, n, lines, n, twl, true, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < lines * n; t += FFT_THREADS) {
        const int line = t / n, i = t - line * n;
        const int r = blockIdx.x * lines + line;
        if (r < nrows) phi[(size_t)n * r + i] = res[t].x * scale;
    }
}

// workgroup shape of the k-marching fft_div_rhs_kernel: 64 x FC_Y cells, FC_KZ planes
static constexpr int FC_Y = 4, FC_KZ = 16;

// 6. pressure_solver.f90:150-169
template <bool STRETCH, bool IBM>
__global__ __launch_bounds__(256) void fft_correct_kernel(GridK g, const double *__restrict__ phi,
                                                          const double *__restrict__ us,
                                                          const double *__restrict__ vs,
                                                          const double *__restrict__ ws,
                                                          const double *__restrict__ cu,
                                                          const double *__restrict__ cv,
                                                          const double *__restrict__ cw, double dtg,
                                                          double *un, double *vn, double *wn, double *pn) {
    const Dims d = g.d;
    const int i = blockIdx.x * 64 + threadIdx.x + 1;
    const int j = blockIdx.y * 4 + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > d.nx || j > d.ny) return;
    const int nx = d.nx, nz = d.nz;
    int im = i - 1, km = k - 1;
    if (im < 1) im = nx;
    if (km < 1) km = nz;
#define RI(ii, kk, jj) ((size_t)((ii) - 1) + (size_t)nx * ((size_t)((kk) - 1) + (size_t)nz * (size_t)((jj) - 1)))
    const size_t c = d.g(i, j, k);
    const double p0 = phi[RI(i, k, j)];
    un[c] = (us[c] - dtg * (p0 - phi[RI(im, k, j)]) / g.dx) / (1.0 + dtg * (IBM ? cu[c] : 0.0));
    wn[c] = (ws[c] - dtg * (p0 - phi[RI(i, km, j)]) / g.dz) / (1.0 + dtg * (IBM ? cw[c] : 0.0));
    pn[c] = pn[c] + p0;
    if (j >= 2) {
        const double gy = STRETCH ? dtg * (p0 - phi[RI(i, k, j - 1)]) * g.ihv[j]
                                  : dtg * (p0 - phi[RI(i, k, j - 1)]) / g.dy;
        vn[c] = (vs[c] - gy) / (1.0 + dtg * (IBM ? cv[c] : 0.0));
    }
#undef RI
}

static int rows_per_block(int n) {
    int lines = (4 * FFT_THREADS) / n;   // n/4 radix-4 butterflies per line, FFT_THREADS threads
    if (lines < 1) lines = 1;
    return lines;
}

static int z_tile(int nz) {
    int tx = 2048 / nz;                  // 2 buffers x tx x (nz + 1) x 16 B, about 64 KB
    if (tx < 1) tx = 1;
    if (tx > 16) tx = 16;
    return tx;
}

// dynamic LDS of the transforms: two ping-pong buffers of `lines` lines (stride ls) and the twiddles
template <typename C>
static size_t fft_lds(int lines, int ls, int n) {
    return (2 * (size_t)lines * ls + (size_t)(n / 2)) * sizeof(C);
}

// allow a z-transform kernel more than the default 64 KB of dynamic LDS (once per kernel)
template <typename C>
static hipError_t zfft_allow(size_t bytes) {
    static size_t allowed = 65536;
    if (bytes <= allowed) return hipSuccess;
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&zfft_kernel<C>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
    if (e == hipSuccess) allowed = bytes;
    return e;
}

hipError_t launch_fft_forward(mdx_handle *h, double dtg) {
    GridK g = grid_k(h);
    const Dims d = h->d;
    const int lines = rows_per_block(d.nx);
    const int nrows = d.ny * d.nz;
    const size_t shx = fft_lds<double2>(lines, d.nx, d.nx);
    if (h->stretched)
        hipLaunchKernelGGL(rhs_xfft_kernel<true>, dim3((nrows + lines - 1) / lines), dim3(FFT_THREADS), shx,
                           h->stream, g, h->us, h->vs, h->ws, dtg, (const double2 *)h->tw_x, lines,
                           (double2 *)h->hat);
    else
        hipLaunchKernelGGL(rhs_xfft_kernel<false>, dim3((nrows + lines - 1) / lines), dim3(FFT_THREADS), shx,
                           h->stream, g, h->us, h->vs, h->ws, dtg, (const double2 *)h->tw_x, lines,
                           (double2 *)h->hat);
    const int tx = z_tile(d.nz), nxh = d.nx / 2 + 1;
    const size_t shz = fft_lds<double2>(tx, d.nz + 1, d.nz);
    hipError_t e = zfft_allow<double2>(shz);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(zfft_kernel<double2>, dim3(d.ny * ((nxh + tx - 1) / tx)), dim3(FFT_THREADS), shz, h->stream, d,
                       (double2 *)h->hat, (const double2 *)h->tw_z, tx, 0);
    return hipGetLastError();
}

hipError_t launch_tridiag(mdx_handle *h) {
    const Dims d = h->d;
    const int nlines = (d.nx / 2 + 1) * d.nz;
    if (h->wave_tridiag && h->part_tridiag &&
        launch_tridiag_part<double2, false>(h->stream, d, (double2 *)h->hat, h->lamx, h->lamz, h->tw_a, h->bdrow,
                                            h->crow))
        return hipGetLastError();
    if (h->wave_tridiag && launch_tridiag_wave<double2, false>(h->stream, d, (double2 *)h->hat, h->lamx, h->lamz, h->tw_a,
                                                     h->bdrow, h->crow))
        return hipGetLastError();
    hipLaunchKernelGGL(tridiag_kernel, dim3((nlines + 63) / 64), dim3(64), 0, h->stream, d, (double2 *)h->hat,
                       h->cp, h->den_inv, h->arow);
    return hipGetLastError();
}

hipError_t launch_fft_backward_correct(mdx_handle *h, double dtg) {
    GridK g = grid_k(h);
    const Dims d = h->d;
    const int tx = z_tile(d.nz), nxh = d.nx / 2 + 1;
    const size_t shz = fft_lds<double2>(tx, d.nz + 1, d.nz);
    hipError_t e = zfft_allow<double2>(shz);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(zfft_kernel<double2>, dim3(d.ny * ((nxh + tx - 1) / tx)), dim3(FFT_THREADS), shz, h->stream, d,
                       (double2 *)h->hat, (const double2 *)h->tw_z, tx, 1);
    const int lines = rows_per_block(d.nx);
    const int nrows = d.ny * d.nz;
    const size_t shx = fft_lds<double2>(lines, d.nx, d.nx);
    const double scale = 1.0 / (double)(d.nx * d.nz);
    hipLaunchKernelGGL(ixfft_kernel, dim3((nrows + lines - 1) / lines), dim3(FFT_THREADS), shx, h->stream, d,
                       (const double2 *)h->hat, (const double2 *)h->tw_x, lines, scale, h->fft_rhs);
    dim3 blk(64, 4), grd((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz);
#define MDX_CORR(S, I)                                                                                     \
    hipLaunchKernelGGL((fft_correct_kernel<S, I>), grd, blk, 0, h->stream, g, h->fft_rhs, h->us, h->vs, h->ws, \
                       h->cu, h->cv, h->cw, dtg, h->un, h->vn, h->wn, h->pn)
    if (h->stretched) { if (h->ibm_on) MDX_CORR(true, true); else MDX_CORR(true, false); }
    else { if (h->ibm_on) MDX_CORR(false, true); else MDX_CORR(false, false); }
#undef MDX_CORR
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- mixed precision
// FP32 transforms and Thomas solves inside an FP64 defect correction.  Each pass takes the residual
// r = rhs - L phi, formed in FP64 with the discrete operator the transforms diagonalise (3-point
// second differences in x and z, the tridiagonal rows in y).  Its mean mode (kx = kz = 0), whose
// system is singular and pinned by phi_hat(y_1) = 0 (pressure_fft.f90:92-96, 119-121), is solved in
// FP64 from the plane sums of r; the other modes go through the FP32 transforms and Thomas solves.

// rhs(i,k,j) of rhs_xfft_kernel, in FP64, layout (nx, nz, ny).  A 64 x FC_Y workgroup owns rows
// j0 .. j0+FC_Y-1 of a 64-cell x segment and marches FC_KZ planes: w*(k+1) becomes w*(k) of the
// next plane in a register and u*(i+1) comes from the lane above (lane 63 reads it), so w* and u*
// are fetched about once; v*(j+1) is the row above's v*(j) (a cache hit).
template <bool STRETCH>
__global__ __launch_bounds__(64 * FC_Y) void fft_div_rhs_kernel(GridK g, const double *__restrict__ us,
                                                               const double *__restrict__ vs,
                                                               const double *__restrict__ ws, double dtg,
                                                               double *__restrict__ rhs) {
    const Dims d = g.d;
    const int lane = threadIdx.x;
    const int i = blockIdx.x * 64 + lane + 1;
    const int j = blockIdx.y * FC_Y + threadIdx.y + 1;
    const int k0 = blockIdx.z * FC_KZ + 1, k1 = k0 + FC_KZ - 1 < d.nz ? k0 + FC_KZ - 1 : d.nz;
    const bool act = i <= d.nx && j <= d.ny;
    double wk = act ? ws[d.g(i, j, k0)] : 0.0;
    for (int k = k0; k <= k1; ++k) {
        double uc = 0.0, up = 0.0, vj = 0.0, vjp = 0.0, wkp = 0.0;
        if (act) {
            uc = us[d.g(i, j, k)];
            vj = vs[d.g(i, j, k)];
            vjp = vs[d.g(i, j + 1, k)];
            wkp = ws[d.g(i, j, k + 1)];
        }
        up = __shfl_down(uc, 1, 64);
        if (act && (lane == 63 || i == d.nx)) up = us[d.g(i + 1, j, k)];
        if (act) {
            const double dv = STRETCH ? (vjp - vj) * g.ihc[j] : (vjp - vj) / g.dy;
            rhs[(size_t)(i - 1) + (size_t)d.nx * ((size_t)(k - 1) + (size_t)d.nz * (size_t)(j - 1))] =
                ((up - uc) / g.dx + dv + (wkp - wk) / g.dz) / dtg;
        }
        wk = wkp;
    }
}

// sums[j-1] = sum over the (x, z) plane j of rhs (one workgroup per plane, fixed order)
// partial sums of the (nx, nz) planes: workgroup (plane j, chunk c) sums chunk c of PLANE_CHUNKS
// equal parts of plane j into part[j * PLANE_CHUNKS + c]; consumers fold a plane's chunks in order
static constexpr int PLANE_CHUNKS = 16;
__global__ __launch_bounds__(256) void fft_plane_sum_kernel(Dims d, const double *__restrict__ rhs, double *part) {
    __shared__ double s[256];
    const long plane = (long)d.nx * d.nz;
    const int j = blockIdx.x / PLANE_CHUNKS, c = blockIdx.x % PLANE_CHUNKS;
    const long lo = plane * c / PLANE_CHUNKS, hi = plane * (c + 1) / PLANE_CHUNKS;
    const double *p = rhs + plane * j;
    double a = 0.0;
    for (long q = lo + threadIdx.x; q < hi; q += 256) a += p[q];
    s[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__device__ inline double plane_sum(const double *part, int j) {
    double a = 0.0;
    for (int c = 0; c < PLANE_CHUNKS; ++c) a += part[j * PLANE_CHUNKS + c];
    return a;
}

// the kx = kz = 0 line of tridiag_kernel in FP64 on the plane sums; mean[j-1] = x_j / (nx nz).
// The operands are staged in LDS first, so the sequential recurrences do not wait on global memory.
static constexpr int MEAN_MAX_NY = 1024;   // 40 KB of LDS
__global__ __launch_bounds__(256) void fft_mean_mode_kernel(Dims d, const double *__restrict__ part,
                                                            const double *__restrict__ cp,
                                                            const double *__restrict__ den_inv,
                                                            const double *__restrict__ arow, double scale,
                                                            double *mean) {
    __shared__ double ss[MEAN_MAX_NY], sa[MEAN_MAX_NY], sd[MEAN_MAX_NY], sc[MEAN_MAX_NY], sm[MEAN_MAX_NY];
    const int ny = d.ny;
    const size_t cs = (size_t)(d.nx / 2 + 1);   // stride between j of line (0, 0) in cp / den_inv
    for (int q = threadIdx.x; q < ny; q += 256) {
        ss[q] = plane_sum(part, q);
        sa[q] = arow[q];
        sd[q] = den_inv[cs * q];
        sc[q] = cp[cs * q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double dp = 0.0;
        sm[0] = 0.0;
        for (int jj = 2; jj <= ny; ++jj) {
            dp = (ss[jj - 1] - sa[jj - 1] * dp) * sd[jj - 1];
            sm[jj - 1] = dp;
        }
        double nxt = dp;
        sm[ny - 1] = nxt * scale;
        for (int jj = ny - 1; jj >= 1; --jj) {
            nxt = sm[jj - 1] - sc[jj - 1] * nxt;
            sm[jj - 1] = nxt;
        }
        for (int jj = 1; jj < ny; ++jj) sm[jj - 1] *= scale;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < ny; q += 256) mean[q] = sm[q];
}

// phi += mean[j] everywhere in plane j
__global__ void fft_add_mean_kernel(Dims d, const double *__restrict__ mean, double *__restrict__ phi) {
    const long n = (long)d.nx * d.ny * d.nz;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x)
        phi[q] += mean[q / ((long)d.nx * d.nz)];
}

// phi -= (plane-1 sum of phi) / (nx nz): re-imposes the pin phi_hat(y_1) = 0 exactly (the FP32
// inverse transforms leave plane means of the order of their rounding, and a constant is invisible
// to the residual); velocities see only differences of phi and do not change
__global__ void fft_pin_kernel(Dims d, const double *__restrict__ part, double *__restrict__ phi) {
    const long n = (long)d.nx * d.ny * d.nz;
    const double c = plane_sum(part, 0) / (double)((long)d.nx * d.nz);
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) phi[q] -= c;
}

// res = rhs - L phi (FP64), and max |res| per workgroup; layout (nx, nz, ny), x and z periodic
template <bool STRETCH>
__global__ __launch_bounds__(256) void fft_residual_kernel(GridK g, const double *__restrict__ rhs,
                                                           const double *__restrict__ phi, const double *__restrict__ arow,
                                                           double *__restrict__ res, double *part) {
    const Dims d = g.d;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long n = (long)nx * ny * nz;
    const double idx2 = 1.0 / (g.dx * g.dx), idz2 = 1.0 / (g.dz * g.dz);
    double m = 0.0;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int i = (int)(q % nx);
        const long r = q / nx;
        const int k = (int)(r % nz), j = (int)(r / nz);   // 0-based
        const long row = q - i;
        const double p0 = phi[q];
        const double pxp = phi[row + (i + 1 == nx ? 0 : i + 1)], pxm = phi[row + (i == 0 ? nx - 1 : i - 1)];
        const long plane = (long)nx * nz * j;
        const long kp = k + 1 == nz ? 0 : k + 1, km = k == 0 ? nz - 1 : k - 1;
        const double pzp = phi[plane + (long)nx * kp + i], pzm = phi[plane + (long)nx * km + i];
        double ly = 0.0;
        if (j > 0) ly += arow[j] * (phi[q - (long)nx * nz] - p0);
        if (j < ny - 1) {
            const double c = STRETCH ? g.ihv[j + 2] * g.ihc[j + 1] : arow[j + 1];
            ly += c * (phi[q + (long)nx * nz] - p0);
        }
        const double lap = (pxp - 2.0 * p0 + pxm) * idx2 + (pzp - 2.0 * p0 + pzm) * idz2 + ly;
        const double rr = rhs[q] - lap;
        res[q] = rr;
        const double a = fabs(rr);
        m = nanmax_any(m, a);
    }
    block_max_part(m, part);
}

// 1. (FP32) r2c x-transform of a real FP64 field in (nx, nz, ny) layout
__global__ __launch_bounds__(FFT_THREADS) void xfft32_kernel(Dims d, const double *__restrict__ in,
                                                             const float2 *__restrict__ tw, int lines, float2 *hat) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float2 *lds = reinterpret_cast<float2 *>(lds_raw);
    const int n = d.nx, nxh = n / 2 + 1;
    const int nrows = d.ny * d.nz;
    float2 *a = lds, *b = lds + (size_t)lines * n, *twl = b + (size_t)lines * n;
    stage_twiddles(twl, tw, n);
    for (int t = threadIdx.x; t < lines * n; t += FFT_THREADS) {
        const int line = t / n, i = t - line * n;
        const int r = blockIdx.x * lines + line;
        a[t] = make_float2(r < nrows ? (float)in[(size_t)n * r + i] : 0.f, 0.f);
    }
    __syncthreads();
    float2 *res = stockham(a, b, n, lines, n, twl, false, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < lines * nxh; t += FFT_THREADS) {
        const int line = t / nxh, m = t - line * nxh;
        const int r = blockIdx.x * lines + line;
        if (r < nrows) hat[(size_t)nxh * r + m] = res[line * n + m];
    }
}

// 3. (FP32) Thomas per (kx, kz) line with FP32 tables; the mean mode line is set to zero (it is
//    solved in FP64 separately)
__global__ __launch_bounds__(64) void tridiag32_kernel(Dims d, float2 *hat, const float *__restrict__ cp,
                                                       const float *__restrict__ den_inv,
                                                       const float *__restrict__ arow) {
    const int nxh = d.nx / 2 + 1, ny = d.ny, nz = d.nz;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nxh * nz) return;
    const int kx = t % nxh, kz = t / nxh;
    const size_t hs = (size_t)nxh * nz, cs = (size_t)nxh;
    const size_t h0 = (size_t)kx + (size_t)nxh * kz;
    const size_t c0 = (size_t)kx + (size_t)nxh * (size_t)ny * kz;
    if (kx == 0 && kz == 0) {
        for (int jj = 1; jj <= ny; ++jj) hat[h0 + hs * (jj - 1)] = make_float2(0.f, 0.f);
        return;
    }
    float2 dp;
    {
        const float2 r = hat[h0];
        const float di = den_inv[c0];
        dp = make_float2(r.x * di, r.y * di);
    }
    hat[h0] = dp;
    for (int jb = 2; jb <= ny; jb += TB) {
        float2 rr[TB];
        float aa[TB], dd[TB];
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb + q;
            if (jj <= ny) {
                rr[q] = hat[h0 + hs * (jj - 1)];
                aa[q] = arow[jj - 1];
                dd[q] = den_inv[c0 + cs * (jj - 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb + q;
            if (jj <= ny) {
                dp = make_float2((rr[q].x - aa[q] * dp.x) * dd[q], (rr[q].y - aa[q] * dp.y) * dd[q]);
                hat[h0 + hs * (jj - 1)] = dp;
            }
        }
    }
    float2 nxt = dp;
    for (int jb = ny - 1; jb >= 1; jb -= TB) {
        float2 qq[TB];
        float cc[TB];
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb - q;
            if (jj >= 1) {
                qq[q] = hat[h0 + hs * (jj - 1)];
                cc[q] = cp[c0 + cs * (jj - 1)];
            }
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int jj = jb - q;
            if (jj >= 1) {
                nxt = make_float2(qq[q].x - cc[q] * nxt.x, qq[q].y - cc[q] * nxt.y);
                hat[h0 + hs * (jj - 1)] = nxt;
            }
        }
    }
}

// 5. (FP32) c2r x-transform, scaled, added to the FP64 phi
__global__ __launch_bounds__(FFT_THREADS) void ixfft32_add_kernel(Dims d, const float2 *__restrict__ hat,
                                                                  const float2 *__restrict__ tw, int lines, float scale,
                                                                  double *phi) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    float2 *lds = reinterpret_cast<float2 *>(lds_raw);
    const int n = d.nx, nxh = n / 2 + 1;
    const int nrows = d.ny * d.nz;
    float2 *a = lds, *b = lds + (size_t)lines * n, *twl = b + (size_t)lines * n;
    stage_twiddles(twl, tw, n);
    for (int t = threadIdx.x; t < lines * n; t += FFT_THREADS) {
        const int line = t / n, m = t - line * n;
        const int r = blockIdx.x * lines + line;
        float2 v = make_float2(0.f, 0.f);
        if (r < nrows) {
            const float2 *row = hat + (size_t)nxh * r;
            if (m == 0 || 2 * m == n) v = make_float2(row[m].x, 0.f);
            else if (m < nxh) v = row[m];
            else v = make_float2(row[n - m].x, -row[n - m].y);
        }
        a[t] = v;
    }
    __syncthreads();
    float2 *res = stockham(a, b, n, lines, n, twl, true, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < lines * n; t += FFT_THREADS) {
        const int line = t / n, i = t - line * n;
        const int r = blockIdx.x * lines + line;
        if (r < nrows) phi[(size_t)n * r + i] += (double)(res[t].x * scale);
    }
}

// the FP64 right-hand side -> h->fft_rhs, phi = 0
hipError_t launch_fft_mixed_setup(mdx_handle *h, double dtg) {
    GridK g = grid_k(h);
    const Dims d = h->d;
    dim3 blk(64, FC_Y), grd((d.nx + 63) / 64, (d.ny + FC_Y - 1) / FC_Y, (d.nz + FC_KZ - 1) / FC_KZ);
    if (h->stretched)
        hipLaunchKernelGGL(fft_div_rhs_kernel<true>, grd, blk, 0, h->stream, g, h->us, h->vs, h->ws, dtg, h->fft_rhs);
    else
        hipLaunchKernelGGL(fft_div_rhs_kernel<false>, grd, blk, 0, h->stream, g, h->us, h->vs, h->ws, dtg, h->fft_rhs);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_zero64(h, h->fft_phi, h->ni);
}

// res = rhs - L phi into h->fft_res; its max |.| into the max reduction
hipError_t launch_fft_residual(mdx_handle *h) {
    GridK g = grid_k(h);
    const unsigned blocks = 2048;
    if (h->stretched)
        hipLaunchKernelGGL(fft_residual_kernel<true>, dim3(blocks), dim3(256), 0, h->stream, g, h->fft_rhs, h->fft_phi,
                           h->arow, h->fft_res, h->max_part);
    else
        hipLaunchKernelGGL(fft_residual_kernel<false>, dim3(blocks), dim3(256), 0, h->stream, g, h->fft_rhs,
                           h->fft_phi, h->arow, h->fft_res, h->max_part);
    hipError_t e = hipGetLastError();
    return e != hipSuccess ? e : launch_max_final(h, blocks);
}

// one defect-correction pass on res = rhs - L phi: the mean mode in FP64 (plane sums, pinned
// Thomas), the other modes by FP32 transforms and Thomas; both corrections added to h->fft_phi
hipError_t launch_fft32_solve_add(mdx_handle *h) {
    const Dims d = h->d;
    if (d.ny > MEAN_MAX_NY) return hipErrorInvalidValue;
    hipLaunchKernelGGL(fft_plane_sum_kernel, dim3(d.ny * PLANE_CHUNKS), dim3(256), 0, h->stream, d, h->fft_res,
                       h->fft_mean);
    double *mean = h->fft_mean + (size_t)d.ny * PLANE_CHUNKS;
    hipLaunchKernelGGL(fft_mean_mode_kernel, dim3(1), dim3(256), 0, h->stream, d, h->fft_mean, h->cp, h->den_inv,
                       h->arow, 1.0 / (double)(d.nx * d.nz), mean);
    hipLaunchKernelGGL(fft_add_mean_kernel, dim3(4096), dim3(256), 0, h->stream, d, mean, h->fft_phi);
    const int lines = rows_per_block(d.nx), nrows = d.ny * d.nz, nxh = d.nx / 2 + 1;
    const size_t shx = fft_lds<float2>(lines, d.nx, d.nx);
    hipLaunchKernelGGL(xfft32_kernel, dim3((nrows + lines - 1) / lines), dim3(FFT_THREADS), shx, h->stream, d,
                       h->fft_res, (const float2 *)h->tw_x32, lines, (float2 *)h->hat32);
    const int tx = z_tile(d.nz);
    const size_t shz = fft_lds<float2>(tx, d.nz + 1, d.nz);
    hipError_t ez = zfft_allow<float2>(shz);
    if (ez != hipSuccess) return ez;
    hipLaunchKernelGGL(zfft_kernel<float2>, dim3(d.ny * ((nxh + tx - 1) / tx)), dim3(FFT_THREADS), shz, h->stream, d,
                       (float2 *)h->hat32, (const float2 *)h->tw_z32, tx, 0);
    const int nlines = nxh * d.nz;
    if (!(h->wave_tridiag && h->part_tridiag &&
          launch_tridiag_part<float2, true>(h->stream, d, (float2 *)h->hat32, h->lamx32, h->lamz32, h->tw_a32,
                                            h->bdrow32, h->crow32)) &&
        !(h->wave_tridiag && launch_tridiag_wave<float2, true>(h->stream, d, (float2 *)h->hat32, h->lamx32, h->lamz32,
                                                     h->tw_a32, h->bdrow32, h->crow32)))
        hipLaunchKernelGGL(tridiag32_kernel, dim3((nlines + 63) / 64), dim3(64), 0, h->stream, d, (float2 *)h->hat32,
                           h->cp32, h->den32, h->arow32);
    hipLaunchKernelGGL(zfft_kernel<float2>, dim3(d.ny * ((nxh + tx - 1) / tx)), dim3(FFT_THREADS), shz, h->stream, d,
                       (float2 *)h->hat32, (const float2 *)h->tw_z32, tx, 1);
    hipLaunchKernelGGL(ixfft32_add_kernel, dim3((nrows + lines - 1) / lines), dim3(FFT_THREADS), shx, h->stream, d,
                       (const float2 *)h->hat32, (const float2 *)h->tw_x32, lines, 1.0f / (float)(d.nx * d.nz),
                       h->fft_phi);
    return hipGetLastError();
}

hipError_t launch_fft_pin(mdx_handle *h) {
    const Dims d = h->d;
    hipLaunchKernelGGL(fft_plane_sum_kernel, dim3(PLANE_CHUNKS), dim3(256), 0, h->stream, d, h->fft_phi, h->fft_mean);
    hipLaunchKernelGGL(fft_pin_kernel, dim3(4096), dim3(256), 0, h->stream, d, h->fft_mean, h->fft_phi);
    return hipGetLastError();
}

// the projection with a given FP64 phi in (nx, nz, ny) layout (step 6)
hipError_t launch_fft_correct(mdx_handle *h, const double *phi, double dtg) {
    GridK g = grid_k(h);
    const Dims d = h->d;
    dim3 blk(64, 4), grd((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz);
#define MDX_CORR(S, I)                                                                                     \
    hipLaunchKernelGGL((fft_correct_kernel<S, I>), grd, blk, 0, h->stream, g, phi, h->us, h->vs, h->ws, h->cu, \
                       h->cv, h->cw, dtg, h->un, h->vn, h->wn, h->pn)
    if (h->stretched) { if (h->ibm_on) MDX_CORR(true, true); else MDX_CORR(true, false); }
    else { if (h->ibm_on) MDX_CORR(false, true); else MDX_CORR(false, false); }
#undef MDX_CORR
    return hipGetLastError();
}

}  // namespace mdx
