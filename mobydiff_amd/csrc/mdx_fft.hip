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
// Transforms are radix-2 Stockham (self-sorting) in LDS with a precomputed FP64 twiddle table;
// hat keeps the reference's half-spectrum layout plane_hat(nxh, nz, ny), kx fastest.
#include "mdx_internal.h"

namespace mdx {

// One radix-2 Stockham pass sequence over LINES lines of length n held in LDS (a, b ping-pong).
// Returns the buffer holding the result.  tw[m] = exp(-2 pi i m / n), m < n/2.
__device__ inline double2 *stockham(double2 *a, double2 *b, int n, int lines, const double2 *__restrict__ tw,
                                    bool inverse, int tid, int nthr) {
    const int half = n >> 1;
    for (int ns = 1; ns < n; ns <<= 1) {
        const int stride = n / (2 * ns);
        for (int t = tid; t < half * lines; t += nthr) {
            const int line = t / half, jj = t - line * half;
            const int kk = jj & (ns - 1);
            const double2 x0 = a[line * n + jj], x1 = a[line * n + jj + half];
            double2 w = tw[kk * stride];
            if (inverse) w.y = -w.y;
            const double tr = x1.x * w.x - x1.y * w.y, ti = x1.x * w.y + x1.y * w.x;
            const int o = ((jj - kk) << 1) + kk;
            b[line * n + o] = make_double2(x0.x + tr, x0.y + ti);
            b[line * n + o + ns] = make_double2(x0.x - tr, x0.y - ti);
        }
        __syncthreads();
        double2 *t2 = a;
        a = b;
        b = t2;
    }
    return a;
}

static constexpr int FFT_THREADS = 256;

// 1. rhs(i,k,j) = [(u*(i+1)-u*(i))/dx + (v*(j+1)-v*(j))/dy + (w*(k+1)-w*(k))/dz]/dtg, then the
//    r2c x-transform; rows r = (k-1) + nz (j-1) are contiguous in hat.
template <bool STRETCH>
__global__ __launch_bounds__(FFT_THREADS) void rhs_xfft_kernel(GridK g, const double *__restrict__ us,
                                                              const double *__restrict__ vs,
                                                              const double *__restrict__ ws, double dtg,
                                                              const double2 *__restrict__ tw, int lines,
                                                              double2 *hat) {
    extern __shared__ double2 lds[];
    const Dims d = g.d;
    const int n = d.nx, nxh = n / 2 + 1;
    const int nrows = d.ny * d.nz;
    double2 *a = lds, *b = lds + (size_t)lines * n;
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
    double2 *res = stockham(a, b, n, lines, tw, false, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < lines * nxh; t += FFT_THREADS) {
        const int line = t / nxh, m = t - line * nxh;
        const int r = blockIdx.x * lines + line;
        if (r < nrows) hat[(size_t)nxh * r + m] = res[line * n + m];
    }
}

// 2./4. z-transform of TX consecutive kx columns of one j plane.
__global__ __launch_bounds__(FFT_THREADS) void zfft_kernel(Dims d, double2 *hat, const double2 *__restrict__ tw,
                                                           int tx, int inverse) {
    extern __shared__ double2 lds[];
    const int nz = d.nz, nxh = d.nx / 2 + 1;
    const int tiles = (nxh + tx - 1) / tx;
    const int j = blockIdx.x / tiles, kx0 = (blockIdx.x % tiles) * tx;
    double2 *a = lds, *b = lds + (size_t)tx * nz;
    double2 *plane = hat + (size_t)nxh * nz * j;
    for (int t = threadIdx.x; t < tx * nz; t += FFT_THREADS) {
        const int c = t % tx, kz = t / tx;
        const int kx = kx0 + c;
        a[c * nz + kz] = kx < nxh ? plane[(size_t)kx + (size_t)nxh * kz] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    double2 *res = stockham(a, b, nz, tx, tw, inverse != 0, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < tx * nz; t += FFT_THREADS) {
        const int c = t % tx, kz = t / tx;
        const int kx = kx0 + c;
        if (kx < nxh) plane[(size_t)kx + (size_t)nxh * kz] = res[c * nz + kz];
    }
}

// 3. pressure_fft.f90:102-148 with precomputed cp/den_inv (pressure_fft.f90:20-100); one thread
//    per (kx,kz) line, consecutive kx in consecutive lanes (coalesced), in place.
__global__ __launch_bounds__(256) void tridiag_kernel(Dims d, double2 *hat, const double *__restrict__ cp,
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
    for (int jj = 2; jj <= ny; ++jj) {
        const double2 r = hat[h0 + hs * (jj - 1)];
        const double a = arow[jj - 1], di = den_inv[c0 + cs * (jj - 1)];
        dp = make_double2((r.x - a * dp.x) * di, (r.y - a * dp.y) * di);
        hat[h0 + hs * (jj - 1)] = dp;
    }
    double2 nxt = dp;   // hat(ny) = dp(ny)
    for (int jj = ny - 1; jj >= 1; --jj) {
        const double2 q = hat[h0 + hs * (jj - 1)];
        const double c = cp[c0 + cs * (jj - 1)];
        nxt = make_double2(q.x - c * nxt.x, q.y - c * nxt.y);
        hat[h0 + hs * (jj - 1)] = nxt;
    }
}

// 5. c2r along x: the real parts of the kx=0 and kx=nx/2 coefficients, Hermitian extension of the
//    rest, inverse transform, times 1/(nx nz); phi in rhs layout (nx, nz, ny).
__global__ __launch_bounds__(FFT_THREADS) void ixfft_kernel(Dims d, const double2 *__restrict__ hat,
                                                           const double2 *__restrict__ tw, int lines,
                                                           double scale, double *phi) {
    extern __shared__ double2 lds[];
    const int n = d.nx, nxh = n / 2 + 1;
    const int nrows = d.ny * d.nz;
    double2 *a = lds, *b = lds + (size_t)lines * n;
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
    double2 *res = stockham(a, b, n, lines, tw, true, threadIdx.x, FFT_THREADS);
    for (int t = threadIdx.x; t < lines * n; t += FFT_THREADS) {
        const int line = t / n, i = t - line * n;
        const int r = blockIdx.x * lines + line;
        if (r < nrows) phi[(size_t)n * r + i] = res[t].x * scale;
    }
}

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
    int lines = (2 * FFT_THREADS) / n;   // n/2 butterflies per line, FFT_THREADS threads
    if (lines < 1) lines = 1;
    return lines;
}

static int z_tile(int nz) {
    int tx = 2048 / nz;                  // 2 buffers x tx x nz x 16 B <= 64 KB
    if (tx < 1) tx = 1;
    if (tx > 16) tx = 16;
    return tx;
}

hipError_t launch_fft_forward(mdx_handle *h, double dtg) {
    GridK g = grid_k(h);
    const Dims d = h->d;
    const int lines = rows_per_block(d.nx);
    const int nrows = d.ny * d.nz;
    const size_t shx = 2 * (size_t)lines * d.nx * sizeof(double2);
    if (h->stretched)
        hipLaunchKernelGGL(rhs_xfft_kernel<true>, dim3((nrows + lines - 1) / lines), dim3(FFT_THREADS), shx,
                           h->stream, g, h->us, h->vs, h->ws, dtg, (const double2 *)h->tw_x, lines,
                           (double2 *)h->hat);
    else
        hipLaunchKernelGGL(rhs_xfft_kernel<false>, dim3((nrows + lines - 1) / lines), dim3(FFT_THREADS), shx,
                           h->stream, g, h->us, h->vs, h->ws, dtg, (const double2 *)h->tw_x, lines,
                           (double2 *)h->hat);
    const int tx = z_tile(d.nz), nxh = d.nx / 2 + 1;
    const size_t shz = 2 * (size_t)tx * d.nz * sizeof(double2);
    hipLaunchKernelGGL(zfft_kernel, dim3(d.ny * ((nxh + tx - 1) / tx)), dim3(FFT_THREADS), shz, h->stream, d,
                       (double2 *)h->hat, (const double2 *)h->tw_z, tx, 0);
    return hipGetLastError();
}

hipError_t launch_tridiag(mdx_handle *h) {
    const Dims d = h->d;
    const int nlines = (d.nx / 2 + 1) * d.nz;
    hipLaunchKernelGGL(tridiag_kernel, dim3((nlines + 255) / 256), dim3(256), 0, h->stream, d, (double2 *)h->hat,
                       h->cp, h->den_inv, h->arow);
    return hipGetLastError();
}

hipError_t launch_fft_backward_correct(mdx_handle *h, double dtg) {
    GridK g = grid_k(h);
    const Dims d = h->d;
    const int tx = z_tile(d.nz), nxh = d.nx / 2 + 1;
    const size_t shz = 2 * (size_t)tx * d.nz * sizeof(double2);
    hipLaunchKernelGGL(zfft_kernel, dim3(d.ny * ((nxh + tx - 1) / tx)), dim3(FFT_THREADS), shz, h->stream, d,
                       (double2 *)h->hat, (const double2 *)h->tw_z, tx, 1);
    const int lines = rows_per_block(d.nx);
    const int nrows = d.ny * d.nz;
    const size_t shx = 2 * (size_t)lines * d.nx * sizeof(double2);
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

}  // namespace mdx
