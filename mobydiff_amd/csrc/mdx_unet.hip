// mdx_unet.hip -- the pressure-guess U-Net on gfx950: BF16 weights and activations, FP32
// accumulation on MFMA (v_mfma_f32_16x16x32_bf16), with forward, backward and Adam kernels.
//
// Network (3 levels, 16/32/64 channels, 3x3x3 convs, zero padding, ReLU):
//   e1a 16->16 (input channels 0,1 carry div(u*) and p; the rest are zero), e1b 16->16,
//   avg-pool 2, e2a 16->32, e2b 32->32, avg-pool 2, b3a 32->64, b3b 64->64,
//   nearest x2 + concat(e2b) -> d2a 96->32, d2b 32->32,
//   nearest x2 + concat(e1b) -> d1a 48->16, d1b 16->16, head 1x1 16->1 (linear, FP32 out).
// Activations are channels-last, voxel v = x + W (y + H z), element v*C + c.
//
// Convolution = implicit GEMM: out[v][n] = sum_k A[v][k] B[k][n], k = tap*CIN + cin, A gathered
// from the input (8 consecutive channels of one tap per lane: one 16-byte load), B from weights
// repacked into MFMA fragment order.  The input gradient is the same GEMM with the weights
// flipped in space and transposed in channels.  The weight gradient reduces over voxels:
// dW[k][n] = sum_v A[v][k] dY[v][n], computed per tap on 32-voxel chunks staged (transposed)
// in LDS, partial sums added atomically in FP32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mdx_internal.h"
#include "mdx_unet.h"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace mdx {
namespace unet {

// ------------------------------------------------------------------------------- conv forward
template <int CIN, int COUT, bool RELU>
__global__ __launch_bounds__(256) void conv3_fwd_kernel(const bf16 *__restrict__ in, const bf16x8 *__restrict__ wp,
                                                        const float *__restrict__ bias, bf16 *__restrict__ out,
                                                        int W, int H, int D) {
    constexpr int K = 27 * CIN, KC = (K + 31) / 32, NT = COUT / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long V = (long)W * H * D;
    const long v0 = ((long)blockIdx.x * 4 + wave) * 16;
    if (v0 >= V) return;
    const int r = lane & 15, g = lane >> 4;
    const long v = v0 + r;
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long)W * H));
    f32x4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = {};
    auto load_a = [&](int kc) {
        const int k0 = kc * 32 + g * 8;
        bf16x8 a = zero8;
        if (k0 < K) {
            const int tap = k0 / CIN, c0 = k0 - tap * CIN;
            const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
            if (xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D)
                a = *reinterpret_cast<const bf16x8 *>(in + (((long)zz * H + yy) * W + xx) * CIN + c0);
        }
        return a;
    };
    // operands are software-pipelined PD K chunks ahead in a ring of registers (a chunk holds only
    // NT MFMAs, too little work to cover one global fetch); the MFMA sequence is unchanged
    constexpr int PD = NT >= 4 ? 2 : 4;
    bf16x8 aq[PD], bq[PD][NT];
#pragma unroll
    for (int p = 0; p < PD; ++p) {
        aq[p] = p < KC ? load_a(p) : zero8;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bq[p][nt] = p < KC ? wp[((long)p * NT + nt) * 64 + lane] : zero8;
    }
    for (int kc0 = 0; kc0 < KC; kc0 += PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
            const int kc = kc0 + p;
            if (kc >= KC) break;
            const bf16x8 a = aq[p];
            bf16x8 b[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt] = bq[p][nt];
            if (kc + PD < KC) {
                aq[p] = load_a(kc + PD);
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bq[p][nt] = wp[((long)(kc + PD) * NT + nt) * 64 + lane];
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) {
        const int ch = nt * 16 + (lane & 15);
        const float b = bias ? bias[ch] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long vo = v0 + 4 * g + i;
            float val = acc[nt][i] + b;
            if (RELU) val = val > 0.f ? val : 0.f;
            out[vo * COUT + ch] = (bf16)val;
        }
    }
}

// ------------------------------------------------------------------------------- conv forward, split K
// The same GEMM for grids too small to fill the chip (the coarse levels of the strided network:
// 12 x 8 x 6 voxels are 36 M-tiles): S waves share one 16-voxel M-tile, wave w taking the K chunks
// w, w + S, w + 2S, ... (the direct kernel's chunks and weight fragments), and their partial
// accumulators are summed in LDS in wave order, then bias and ReLU as above.  A wave then waits
// for KC / S global fetches instead of KC.  Needs V % 16 == 0.
// The input of a decoder convolution may come unmaterialised (up != nullptr): channels [0, cu) are
// the nearest up-sampling of `up` (a grid of half the extents), channels [cu, cin) are `in` (the
// skip tensor, cin - cu channels): the values upcat_kernel would have stored.
__device__ __forceinline__ const bf16 *conv_in_at(const bf16 *in, const bf16 *up, int cu, int cin, int W, int H, int x,
                                                  int y, int z, int c0) {
    if (!up) return in + (((long)z * H + y) * W + x) * cin + c0;
    if (c0 < cu) return up + (((long)(z >> 1) * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1)) * cu + c0;
    return in + (((long)z * H + y) * W + x) * (cin - cu) + (c0 - cu);
}

template <int COUT> struct SplitK {
    static constexpr int NT = COUT / 16, S = NT >= 3 ? 8 : 16;
};
template <int CIN, int COUT, bool RELU>
__global__ __launch_bounds__(64 * SplitK<COUT>::S) void conv3_splitk_kernel(const bf16 *__restrict__ in,
                                                                            const bf16 *__restrict__ up, int cu,
                                                                            const bf16x8 *__restrict__ wp,
                                                                            const float *__restrict__ bias,
                                                                            bf16 *__restrict__ out, int W, int H,
                                                                            int D) {
    constexpr int K = 27 * CIN, KC = (K + 31) / 32, NT = SplitK<COUT>::NT, S = SplitK<COUT>::S;
    constexpr int NJ = (KC + S - 1) / S, PD = 2;
    __shared__ float red[S][NT * 4][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long v0 = (long)blockIdx.x * 16;
    const int r = lane & 15, g = lane >> 4;
    const long v = v0 + r;
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long)W * H));
    f32x4 acc[NT];
#pragma unroll
    for (int nt = 0; nt < NT; ++nt) acc[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = {};
    auto load_a = [&](int kc) {
        const int k0 = kc * 32 + g * 8;
        bf16x8 a = zero8;
        if (k0 < K) {
            const int tap = k0 / CIN, c0 = k0 - tap * CIN;
            const int xx = x + tap % 3 - 1, yy = y + (tap / 3) % 3 - 1, zz = z + tap / 9 - 1;
            if (xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D)
                a = *reinterpret_cast<const bf16x8 *>(conv_in_at(in, up, cu, CIN, W, H, xx, yy, zz, c0));
        }
        return a;
    };
    bf16x8 aq[PD], bq[PD][NT];
#pragma unroll
    for (int p = 0; p < PD; ++p) {
        const int kc = wave + p * S;
        aq[p] = kc < KC ? load_a(kc) : zero8;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bq[p][nt] = kc < KC ? wp[((long)kc * NT + nt) * 64 + lane] : zero8;
    }
    for (int j0 = 0; j0 < NJ; j0 += PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
            const int j = j0 + p, kc = wave + j * S;
            if (j >= NJ) break;
            const bf16x8 a = aq[p];
            bf16x8 b[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt] = bq[p][nt];
            // chunks past K are zero operands, not a branch around the MFMAs: kc depends on the
            // wave, so such a branch is an execution-masked region, and a version written that way
            // gave wrong sums
            const int kn = kc + PD * S;
            if (j + PD < NJ) {
                aq[p] = kn < KC ? load_a(kn) : zero8;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bq[p][nt] = kn < KC ? wp[((long)kn * NT + nt) * 64 + lane] : zero8;
            }
#pragma unroll
            for (int nt = 0; nt < NT; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[nt], 0, 0, 0);
        }
    }
#pragma unroll
    for (int nt = 0; nt < NT; ++nt)
#pragma unroll
        for (int i = 0; i < 4; ++i) red[wave][nt * 4 + i][lane] = acc[nt][i];
    __syncthreads();
    // output element e = voxel vo of the tile, channel ch; MFMA lane (ch & 15) + 16 (vo / 4), row vo % 4
    for (int e = threadIdx.x; e < 16 * COUT; e += 64 * S) {
        const int vo = e / COUT, ch = e - vo * COUT;
        const int q = (ch >> 4) * 4 + (vo & 3), ln = (ch & 15) + 16 * (vo >> 2);
        float sum = red[0][q][ln];
#pragma unroll
        for (int w = 1; w < S; ++w) sum += red[w][q][ln];
        float val = sum + (bias ? bias[ch] : 0.f);
        if (RELU) val = val > 0.f ? val : 0.f;
        out[(v0 + vo) * COUT + ch] = (bf16)val;
    }
}

// ------------------------------------------------------------------------------- conv forward, LDS-tiled
// Same GEMM, same weight fragments and the same K order and chunking as conv3_fwd_kernel (so the
// outputs are bit-identical), but a workgroup computes a 16 x 4 x TZ output tile and first stages
// its 18 x 6 x (TZ+2) input halo (all CIN channels) in LDS: every input voxel is fetched from global
// memory about 2.5 times instead of once per tap (27), and each wave reuses one weight fragment
// for TZ output rows.  Wave w owns output row y = y0 + w, its TZ M-tiles are z = z0 .. z0+TZ-1, and
// a tile row is 16 consecutive x voxels.  A voxel's channels sit CIN + 8 elements apart in LDS so
// the 16 lanes of a row read disjoint bank groups.  TZ = 4 below 48 channels, 2 at 48 and 64, 1 at
// 96 (measured fastest of TZ 1/2/4 per layer; the halo tile in dynamic LDS: 31 KB at 16 channels,
// 52 KB at 32, 48 KB at 48, 62 KB at 64, 67 KB at 96).  Needs W % 16 == 0, H % 4 == 0, D % TZ == 0.
static constexpr int CTX = 16, CTY = 4, CHX = CTX + 2, CHY = CTY + 2;
template <int CIN> struct ConvTile {
    static constexpr int TZ = CIN >= 96 ? 1 : CIN >= 48 ? 2 : 4, HV = CHX * CHY * (TZ + 2), CS = CIN + 8;
    static constexpr size_t BYTES = (size_t)HV * CS * sizeof(bf16);
};

// POOL: the workgroup also writes the 2x2x2 averages of its output tile to pool (W/2 x H/2 x D/2,
// channels last): after the tile's outputs are stored, their bf16 values pass through LDS and one
// thread per coarse voxel and channel adds its eight in pool_fwd_kernel's order (dz, dy, dx from
// 0.f) and scales by 0.125, so pool holds pool_fwd_kernel's values.
template <int CIN, int COUT, bool RELU, bool POOL = false>
__global__ __launch_bounds__(256) void conv3_tiled_kernel(const bf16 *__restrict__ in, const bf16 *__restrict__ up,
                                                          int cu, const bf16x8 *__restrict__ wp,
                                                          const float *__restrict__ bias, bf16 *__restrict__ out,
                                                          int W, int H, int D, bf16 *__restrict__ pool = nullptr) {
    constexpr int K = 27 * CIN, KC = (K + 31) / 32, NT = COUT / 16, U = CIN / 8;
    constexpr int TZ = ConvTile<CIN>::TZ, HV = ConvTile<CIN>::HV, CS = ConvTile<CIN>::CS;
    extern __shared__ __attribute__((aligned(16))) bf16 tile[];   // HV * CS values
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int ntx = W / CTX, nty = H / CTY;
    const int x0 = (int)(blockIdx.x % ntx) * CTX, y0 = (int)((blockIdx.x / ntx) % nty) * CTY;
    const int z0 = (int)(blockIdx.x / (ntx * nty)) * TZ;
    for (int q = threadIdx.x; q < HV * U; q += 256) {
        const int v = q / U, u = q - v * U;
        const int hx = v % CHX, hy = (v / CHX) % CHY, hz = v / (CHX * CHY);
        const int x = x0 + hx - 1, y = y0 + hy - 1, z = z0 + hz - 1;
        bf16x8 a = {};
        if (x >= 0 && x < W && y >= 0 && y < H && z >= 0 && z < D)
            a = *reinterpret_cast<const bf16x8 *>(conv_in_at(in, up, cu, CIN, W, H, x, y, z, u * 8));
        *reinterpret_cast<bf16x8 *>(&tile[v * CS + u * 8]) = a;
    }
    __syncthreads();
    const int r = lane & 15, g = lane >> 4;
    f32x4 acc[TZ][NT];
#pragma unroll
    for (int m = 0; m < TZ; ++m)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) acc[m][nt] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero8 = {};
    // weight fragments are software-pipelined PD K chunks ahead in a ring of registers: with few
    // output tiles per wave (COUT 16: one) a chunk's MFMAs cannot cover one global fetch; the
    // unrolled group also issues the LDS reads of PD chunks together.  MFMA order unchanged.
    constexpr int PD = NT >= 4 ? 2 : 4;
    bf16x8 bq[PD][NT];
#pragma unroll
    for (int p = 0; p < PD; ++p)
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) bq[p][nt] = p < KC ? wp[((long)p * NT + nt) * 64 + lane] : zero8;
    for (int kc0 = 0; kc0 < KC; kc0 += PD) {
#pragma unroll
        for (int p = 0; p < PD; ++p) {
            const int kc = kc0 + p;
            if (kc >= KC) break;
            const int k0 = kc * 32 + g * 8;
            const bool valid = k0 < K;
            const int tap = valid ? k0 / CIN : 0, c0 = k0 - tap * CIN;
            const int base = ((tap / 9) * CHY + wave + (tap / 3) % 3) * CHX + r + tap % 3;
            bf16x8 b[NT];
#pragma unroll
            for (int nt = 0; nt < NT; ++nt) b[nt] = bq[p][nt];
            if (kc + PD < KC) {
#pragma unroll
                for (int nt = 0; nt < NT; ++nt) bq[p][nt] = wp[((long)(kc + PD) * NT + nt) * 64 + lane];
            }
#pragma unroll
            for (int m = 0; m < TZ; ++m) {
                const bf16x8 a =
                    valid ? *reinterpret_cast<const bf16x8 *>(&tile[(base + m * CHX * CHY) * CS + c0]) : zero8;
#pragma unroll
                for (int nt = 0; nt < NT; ++nt)
                    acc[m][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[nt], acc[m][nt], 0, 0, 0);
            }
        }
    }
    const int y = y0 + wave;
    bf16 keep[POOL ? TZ : 1][4];
#pragma unroll
    for (int m = 0; m < TZ; ++m) {
        const long row = ((long)(z0 + m) * H + y) * W + x0;
#pragma unroll
        for (int nt = 0; nt < NT; ++nt) {
            const int ch = nt * 16 + (lane & 15);
            const float bb = bias ? bias[ch] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float val = acc[m][nt][i] + bb;
                if (RELU) val = val > 0.f ? val : 0.f;
                const bf16 bv = (bf16)val;
                out[(row + 4 * g + i) * COUT + ch] = bv;
                if (POOL) keep[POOL ? m : 0][i] = bv;
            }
        }
    }
    if (POOL) {
        static_assert(!POOL || (COUT == 16 && TZ % 2 == 0 && CTX * CTY / 4 * 16 == 256),
                      "pooling: 16 channels, whole 2x2x2 cells, one thread per coarse voxel and channel");
        static_assert(!POOL || TZ * CTY * CTX * 16 <= HV * CS, "the pooled tile fits the input tile's LDS");
        __syncthreads();   // every wave has read its last input operands
#pragma unroll
        for (int m = 0; m < TZ; ++m)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                tile[((m * CTY + wave) * CTX + 4 * g + i) * 16 + (lane & 15)] = keep[POOL ? m : 0][i];
        __syncthreads();
        const int ch = threadIdx.x & 15, cell = threadIdx.x >> 4;
        const int xc = cell % (CTX / 2), yc = cell / (CTX / 2);
        const int Wc = W / 2, Hc = H / 2;
#pragma unroll
        for (int zc = 0; zc < TZ / 2; ++zc) {
            float s = 0.f;
            for (int dz = 0; dz < 2; ++dz)
                for (int dy = 0; dy < 2; ++dy)
                    for (int dx = 0; dx < 2; ++dx)
                        s += (float)tile[(((2 * zc + dz) * CTY + 2 * yc + dy) * CTX + 2 * xc + dx) * 16 + ch];
            pool[(((long)(z0 / 2 + zc) * Hc + y0 / 2 + yc) * Wc + x0 / 2 + xc) * 16 + ch] = (bf16)(s * 0.125f);
        }
    }
}

// ------------------------------------------------------------------------------- weight gradient
// One workgroup = one tap x one slice of voxels.  Each 32-voxel chunk: dY (32 x COUT) and the
// tap-shifted input (32 x CIN) are staged transposed in LDS ([c][v], so 8 consecutive voxels of one
// channel are one 16-byte read); each wave owns a subset of the (CIN/16 x COUT/16) output tiles.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void conv3_wgrad_kernel(const bf16 *__restrict__ in, const bf16 *__restrict__ dy,
                                                          float *__restrict__ dw, int W, int H, int D, int slices) {
    constexpr int MT = CIN / 16, NT = COUT / 16, TILES = MT * NT;
    constexpr int TPW = (TILES + 3) / 4;   // tiles per wave
    __shared__ __attribute__((aligned(16))) bf16 sx[CIN * 32];
    __shared__ __attribute__((aligned(16))) bf16 sd[COUT * 32];
    const int tap = blockIdx.x % 27, slice = blockIdx.x / 27;
    const int dxo = tap % 3 - 1, dyo = (tap / 3) % 3 - 1, dzo = tap / 9 - 1;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long V = (long)W * H * D;
    const long nchunks = V / 32;
    const long c_begin = nchunks * slice / slices, c_end = nchunks * (slice + 1) / slices;
    f32x4 acc[TPW];
#pragma unroll
    for (int q = 0; q < TPW; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (long ch = c_begin; ch < c_end; ++ch) {
        const long vb = ch * 32;
        __syncthreads();
        // stage: thread t covers (voxel t/8... ) -- CIN/8 and COUT/8 vector loads per voxel
        for (int q = threadIdx.x; q < 32 * (CIN / 8); q += 256) {
            const int vv = q / (CIN / 8), c0 = (q % (CIN / 8)) * 8;
            const long v = vb + vv;
            const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long)W * H));
            const int xx = x + dxo, yy = y + dyo, zz = z + dzo;
            bf16x8 a = {};
            if (xx >= 0 && xx < W && yy >= 0 && yy < H && zz >= 0 && zz < D)
                a = *reinterpret_cast<const bf16x8 *>(in + (((long)zz * H + yy) * W + xx) * CIN + c0);
#pragma unroll
            for (int e = 0; e < 8; ++e) sx[(c0 + e) * 32 + vv] = a[e];
        }
        for (int q = threadIdx.x; q < 32 * (COUT / 8); q += 256) {
            const int vv = q / (COUT / 8), c0 = (q % (COUT / 8)) * 8;
            const bf16x8 a = *reinterpret_cast<const bf16x8 *>(dy + (vb + vv) * COUT + c0);
#pragma unroll
            for (int e = 0; e < 8; ++e) sd[(c0 + e) * 32 + vv] = a[e];
        }
        __syncthreads();
        // D[m = cin][n = cout] += sum_{v in chunk} X[v][m] dY[v][n]:  A[m][k=v], B[k=v][n]
        const int r = lane & 15, g = lane >> 4;
#pragma unroll
        for (int q = 0; q < TPW; ++q) {
            const int tile = wave * TPW + q;
            if (tile < TILES) {
                const int mt = tile / NT, nt = tile % NT;
                const bf16x8 a = *reinterpret_cast<const bf16x8 *>(&sx[(mt * 16 + r) * 32 + 8 * g]);
                const bf16x8 bb = *reinterpret_cast<const bf16x8 *>(&sd[(nt * 16 + r) * 32 + 8 * g]);
                acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, bb, acc[q], 0, 0, 0);
            }
        }
    }
    // accumulate: lane holds D[4g+i][lane%16] -> dw[(tap*CIN + m)*COUT + n]
    const int g = lane >> 4;
#pragma unroll
    for (int q = 0; q < TPW; ++q) {
        const int tile = wave * TPW + q;
        if (tile < TILES) {
            const int mt = tile / NT, nt = tile % NT;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = mt * 16 + 4 * g + i, n = nt * 16 + (lane & 15);
                atomicAdd(&dw[((long)tap * CIN + m) * COUT + n], acc[q][i]);
            }
        }
    }
}

// bias gradient: db[n] = sum_v dY[v][n].  The COUT channels of a voxel are G = COUT/8 16-byte
// vectors; thread t reads vectors q = t, t + stride, ... of the flat array (consecutive lanes,
// consecutive 16 bytes: whole cache lines per wave), and since the stride is a multiple of G a thread
// always sees the same channel group q % G.  The lanes of one group are then folded by shuffles
// (offsets G .. 32) and one atomic addition per channel and wave is left.
template <int COUT>
__global__ __launch_bounds__(256) void bias_grad_kernel(const bf16 *__restrict__ dy, float *__restrict__ db, long V) {
    constexpr int G = COUT / 8;
    float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const long nvec = V * G;
    const bf16x8 *__restrict__ d8 = reinterpret_cast<const bf16x8 *>(dy);
    for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < nvec; q += (long)gridDim.x * 256) {
        const bf16x8 a = d8[q];
#pragma unroll
        for (int e = 0; e < 8; ++e) s[e] += (float)a[e];
    }
#pragma unroll
    for (int e = 0; e < 8; ++e)
        for (int off = 32; off >= G; off >>= 1) s[e] += __shfl_xor(s[e], off, 64);
    const int lane = threadIdx.x & 63;
    if (lane < G) {
        const int g = (int)(((long)blockIdx.x * 256 + threadIdx.x) % G);
#pragma unroll
        for (int e = 0; e < 8; ++e) atomicAdd(&db[g * 8 + e], s[e]);
    }
}

// ------------------------------------------------------------------------------- elementwise
// avg-pool 2x2x2 (fine W,H,D -> coarse W/2,H/2,D/2)
__global__ void pool_fwd_kernel(const bf16 *__restrict__ in, bf16 *__restrict__ out, int C, int W, int H, int D) {
    // one thread per coarse voxel and 8 channels (one 16-byte vector)
    const int Wc = W / 2, Hc = H / 2, Dc = D / 2, G = C / 8;
    const long n = (long)Wc * Hc * Dc * G;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c0 = (int)(q % G) * 8;
        const long v = q / G;
        const int x = (int)(v % Wc), y = (int)((v / Wc) % Hc), z = (int)(v / ((long)Wc * Hc));
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx) {
                    const bf16x8 a = *reinterpret_cast<const bf16x8 *>(
                        in + (((long)(2 * z + dz) * H + 2 * y + dy) * W + 2 * x + dx) * C + c0);
#pragma unroll
                    for (int e = 0; e < 8; ++e) s[e] += (float)a[e];
                }
        bf16x8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = (bf16)(s[e] * 0.125f);
        *reinterpret_cast<bf16x8 *>(out + v * C + c0) = o;
    }
}

// nearest up-sampling of `coarse` into channels [0, Cu) of `cat` and a copy of `skip` into
// channels [Cu, Cu+Cs) (the concat that feeds the decoder conv); one thread per voxel and
// 8 channels
__global__ void upcat_kernel(const bf16 *__restrict__ coarse, const bf16 *__restrict__ skip, bf16 *__restrict__ cat,
                             int Cu, int Cs, int W, int H, int D) {
    const int C = Cu + Cs, G = C / 8;
    const long n = (long)W * H * D * G;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % G) * 8;
        const long v = q / G;
        bf16x8 a;
        if (c < Cu) {
            const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long)W * H));
            a = *reinterpret_cast<const bf16x8 *>(coarse + (((long)(z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2) * Cu + c);
        } else {
            a = *reinterpret_cast<const bf16x8 *>(skip + v * Cs + (c - Cu));
        }
        *reinterpret_cast<bf16x8 *>(cat + v * C + c) = a;
    }
}

// backward of upcat: d_coarse = sum over the 8 children of dcat[:, :Cu]; d_skip += dcat[:, Cu:]
__global__ void upcat_bwd_coarse_kernel(const bf16 *__restrict__ dcat, bf16 *__restrict__ dcoarse, int Cu, int Cs,
                                        int W, int H, int D) {
    const int C = Cu + Cs, Wc = W / 2, Hc = H / 2, Dc = D / 2;
    const long n = (long)Wc * Hc * Dc * Cu;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % Cu);
        const long v = q / Cu;
        const int x = (int)(v % Wc), y = (int)((v / Wc) % Hc), z = (int)(v / ((long)Wc * Hc));
        float s = 0.f;
        for (int dz = 0; dz < 2; ++dz)
            for (int dy = 0; dy < 2; ++dy)
                for (int dx = 0; dx < 2; ++dx)
                    s += (float)dcat[(((long)(2 * z + dz) * H + 2 * y + dy) * W + 2 * x + dx) * C + c];
        dcoarse[q] = (bf16)s;
    }
}

// d_out (gradient arriving at a ReLU layer's output) = (d_from_next + d_skip) * (y > 0)
__global__ void relu_bwd_kernel(bf16 *__restrict__ g, const bf16 *__restrict__ extra, int Ce, int coff, int C,
                                const bf16 *__restrict__ y, long n) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        float d = (float)g[q];
        if (extra) {
            const long v = q / C;
            const int c = (int)(q % C);
            d += (float)extra[v * Ce + coff + c];
        }
        g[q] = (bf16)((float)y[q] > 0.f ? d : 0.f);
    }
}

// avg-pool backward: d_fine[v] = d_coarse[parent(v)] / 8
__global__ void pool_bwd_kernel(const bf16 *__restrict__ dcoarse, bf16 *__restrict__ dfine, int C, int W, int H, int D) {
    const long n = (long)W * H * D * C;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q % C);
        const long v = q / C;
        const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((long)W * H));
        dfine[q] = (bf16)((float)dcoarse[(((long)(z / 2) * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c] * 0.125f);
    }
}

// head: out[v] = sum_c w[c] x[v][c] + b (FP32)
__global__ void head_fwd_kernel(const bf16 *__restrict__ x, const float *__restrict__ w, const float *__restrict__ b,
                                float *__restrict__ out, long V) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        float s = 0.f;
#pragma unroll
        for (int c = 0; c < 16; ++c) s += w[c] * (float)x[v * 16 + c];
        out[v] = s + b[0];
    }
}

// MSE loss against the target: dout = 2 (out - t) / (V * batch); loss partial sums
__global__ void mse_kernel(const float *__restrict__ out, const float *__restrict__ t, float *__restrict__ dout,
                           float scale, long V, float *__restrict__ loss) {
    float s = 0.f;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        const float e = out[v] - t[v];
        dout[v] = 2.f * e * scale;
        s += e * e;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(loss, s * scale);
}

// head backward: dx[v][c] = dout[v] w[c]; dw[c] += sum_v dout[v] x[v][c]; db += sum dout
__global__ __launch_bounds__(256) void head_bwd_kernel(const bf16 *__restrict__ x, const float *__restrict__ w,
                                                       const float *__restrict__ dout, bf16 *__restrict__ dx,
                                                       float *__restrict__ dw, float *__restrict__ db, long V) {
    float sw[16], sb = 0.f;
#pragma unroll
    for (int c = 0; c < 16; ++c) sw[c] = 0.f;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        const float d = dout[v];
        sb += d;
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            sw[c] += d * (float)x[v * 16 + c];
            dx[v * 16 + c] = (bf16)(d * w[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) {
        float t = sw[c];
        for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
        if ((threadIdx.x & 63) == 0) atomicAdd(&dw[c], t);
    }
    for (int off = 32; off > 0; off >>= 1) sb += __shfl_xor(sb, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(db, sb);
}

// ------------------------------------------------------------------------------- weights
// fragment repack: wp[(kc*NT+nt)*64 + lane][e] = W[k][n], k = 32 kc + 8 (lane/16) + e, n = 16 nt + lane%16;
// flip != 0 builds the input-gradient weights W'[tap'][cout][cin] = W[26-tap'][cin][cout]
__global__ void repack_kernel(const float *__restrict__ w, bf16 *__restrict__ wp, int cin, int cout, int flip) {
    const int ci = flip ? cout : cin, co = flip ? cin : cout;   // GEMM input / output channels
    const int K = 27 * ci, KC = (K + 31) / 32, NT = co / 16;
    const long n = (long)KC * NT * 64 * 8;
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const int e = (int)(q % 8);
        const int lane = (int)((q / 8) % 64);
        const long blk = q / 512;
        const int nt = (int)(blk % NT), kc = (int)(blk / NT);
        const int k = 32 * kc + 8 * (lane / 16) + e, nn = 16 * nt + lane % 16;
        float val = 0.f;
        if (k < K) {
            const int tap = k / ci, c = k % ci;
            if (!flip) val = w[((long)tap * cin + c) * cout + nn];
            else val = w[((long)(26 - tap) * cin + nn) * cout + c];
        }
        wp[q] = (bf16)val;
    }
}

// Adam (Kingma & Ba 2015, Alg. 1) on FP32 master weights; grads are averaged over the batch
__global__ void adam_kernel(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2, float gscale) {
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (long)gridDim.x * blockDim.x) {
        const float gg = g[q] * gscale;
        const float mm = b1 * m[q] + (1.f - b1) * gg;
        const float vv = b2 * v[q] + (1.f - b2) * gg * gg;
        m[q] = mm;
        v[q] = vv;
        p[q] = p[q] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    }
}

// ------------------------------------------------------------------------------- solver coupling
// network input from the solver's FP64 state (interior layout, voxel = cell):
// ch0 = r0 / s_r, the residual rhs - L_mu(Phi_0) of the previous-step start Phi_0 = fac * Phi_prev;
// ch1 = Phi_0 / s_p (0 without history).  s_r = rms(r0) = stats[0], s_p = rms(Phi_0) = stats[1].
__global__ void make_input_kernel(long V, const double *__restrict__ r0, const double *__restrict__ phi0,
                                  const double *__restrict__ stats, bf16 *__restrict__ x) {
    const double ir = 1.0 / stats[0], ip = phi0 ? 1.0 / stats[1] : 0.0;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        bf16x8 a = {}, b = {};
        a[0] = (bf16)(float)(r0[v] * ir);
        a[1] = (bf16)(float)(phi0 ? phi0[v] * ip : 0.0);
        *reinterpret_cast<bf16x8 *>(x + v * 16) = a;
        *reinterpret_cast<bf16x8 *>(x + v * 16 + 8) = b;
    }
}

// root mean square of a - fac * b over interior cells (b may be null): partial sums of squares,
// then one final pass (deterministic order)
__global__ __launch_bounds__(256) void input_stats_partial(long V, const double *__restrict__ a,
                                                           const double *__restrict__ b, double fac, double *part) {
    __shared__ double s[256];
    double acc = 0;
    for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < V; v += (long)gridDim.x * 256) {
        const double t = a[v] - (b ? fac * b[v] : 0.0);
        acc += t * t;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// one wave folds the partial sums: lane l adds partials l, l + 64, ... in order, then a shuffle
// tree (a fixed order, so the statistics repeat exactly)
__global__ __launch_bounds__(64) void input_stats_final(const double *part, int nb, double V, double *stat) {
    double a = 0;
    for (int q = threadIdx.x; q < nb; q += 64) a += part[q];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (threadIdx.x == 0) {
        const double r = sqrt(a / V);
        stat[0] = r > 1e-300 ? r : 1.0;
    }
}

// both input statistics in one pass: rms of a into stat[0], of b into stat[1], each summed in the
// order input_stats_partial / input_stats_final use for one array (the same values)
__global__ __launch_bounds__(256) void input_stats2_partial(long V, const double *__restrict__ a,
                                                            const double *__restrict__ b, double *part) {
    __shared__ double sa[256], sb[256];
    double acc_a = 0, acc_b = 0;
    for (long v = (long)blockIdx.x * 256 + threadIdx.x; v < V; v += (long)gridDim.x * 256) {
        const double ta = a[v], tb = b[v];
        acc_a += ta * ta;
        acc_b += tb * tb;
    }
    sa[threadIdx.x] = acc_a;
    sb[threadIdx.x] = acc_b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            sa[threadIdx.x] += sa[threadIdx.x + w];
            sb[threadIdx.x] += sb[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x] = sa[0];
        part[gridDim.x + blockIdx.x] = sb[0];
    }
}

// wave w folds the partials of array w as input_stats_final does
__global__ __launch_bounds__(128) void input_stats2_final(const double *part, int nb, double V, double *stat) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    double a = 0;
    for (int q = l; q < nb; q += 64) a += part[w * nb + q];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_xor(a, off, 64);
    if (l == 0) {
        const double r = sqrt(a / V);
        stat[w] = r > 1e-300 ? r : 1.0;
    }
}

// Phi guess (FP64, interior layout) = Phi_0 + out * s_r * c0 (in place when phi == phi0)
__global__ void guess_out_kernel(const float *__restrict__ out, const double *__restrict__ stats, double c0,
                                 const double *phi0, double *phi, long V) {
    const double s = stats[0] * c0;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x)
        phi[v] = (phi0 ? phi0[v] : 0.0) + (double)out[v] * s;
}

// training target = (Phi - fac * Phi_prev) / (s_r c0) in FP32
__global__ void make_target_kernel(const double *__restrict__ phi, const double *__restrict__ phi_prev, double fac,
                                   const double *__restrict__ stats, double c0, float *__restrict__ t, long V) {
    const double s = 1.0 / (stats[0] * c0);
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x)
        t[v] = (float)((phi[v] - (phi_prev ? fac * phi_prev[v] : 0.0)) * s);
}

// Strided coupling (network grid = solver grid / s).  Network voxel (x, y, z) covers the cells
// (s x + a, s y + b, s z + c), 0 <= a, b, c < s, of the solver's interior layout (fine index
// i + FW (j + FH k), 0-based).
struct FineGrid {
    int W, H, D;      // network grid
    int s;            // stride
    int FW, FH, FD;   // solver grid
};

// network input: the mean over each voxel's s^3 cells of r0 / s_r and Phi_0 / s_p.  S = s (2 or 4)
// as a constant.  A workgroup of 64 x S^2 threads takes 64 voxels of one (vy, vz) row (grid
// ceil(W/64) x H x D): thread (x, y) sums the S cells of fine row y (b = y % S, c = y / S) of voxel
// x, lanes reading consecutive memory, and the S^2 row sums of a voxel are added in row order.
template <int S>
__global__ __launch_bounds__(64 * S * S) void restrict_input_kernel(FineGrid f, const double *__restrict__ r0,
                                                                    const double *__restrict__ phi0,
                                                                    const double *__restrict__ stats,
                                                                    bf16 *__restrict__ x) {
    constexpr int s = S, R = S * S;
    __shared__ double pr[R][64], pp[R][64];
    const int vx = blockIdx.x * 64 + threadIdx.x, vy = blockIdx.y, vz = blockIdx.z, y = threadIdx.y;
    double sr = 0.0, sp = 0.0;
    if (vx < f.W) {
        const int b = y % s, c = y / s;
        const long row = (long)f.FW * ((long)(s * vy + b) + (long)f.FH * (s * vz + c)) + (long)s * vx;
        double vr[s], vp[s];
#pragma unroll
        for (int a = 0; a < s; ++a) {
            vr[a] = r0[row + a];
            vp[a] = phi0 ? phi0[row + a] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < s; ++a) {
            sr += vr[a];
            sp += vp[a];
        }
    }
    pr[y][threadIdx.x] = sr;
    pp[y][threadIdx.x] = sp;
    __syncthreads();
    if (y != 0 || vx >= f.W) return;
    sr = 0.0;
    sp = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) {
        sr += pr[q][threadIdx.x];
        sp += pp[q][threadIdx.x];
    }
    const double ir = 1.0 / stats[0], ip = phi0 ? 1.0 / stats[1] : 0.0;
    const double w = 1.0 / (double)(s * s * s);
    const long v = ((long)vz * f.H + vy) * f.W + vx;
    bf16x8 lo = {}, hi = {};
    lo[0] = (bf16)(float)(sr * w * ir);
    lo[1] = (bf16)(float)(sp * w * ip);
    *reinterpret_cast<bf16x8 *>(x + v * 16) = lo;
    *reinterpret_cast<bf16x8 *>(x + v * 16 + 8) = hi;
}

// The network input from the fused solver pass: every workgroup folds the nb partial sums of r0^2
// and Phi_0^2 in one fixed order (so every workgroup holds the same statistics), then scales the
// voxel sums of its voxels: x = (bf16)(float)(sum * w / s), w = 1 / s^3, s the rms of the field.
// Workgroup 0 also stores stats[0..1] and the two maxima into their pinned slots.
__global__ __launch_bounds__(1024) void input_from_sums_kernel(long V, const double *__restrict__ vsum,
                                                              const double *__restrict__ part, unsigned nb,
                                                              double n_cells, double w, double *stats,
                                                              bf16 *__restrict__ x, double *rz_slot,
                                                              double *rp_slot, const unsigned long long *err) {
    __shared__ double red[4][16];
    const int t = threadIdx.x;
    const bool first = blockIdx.x == 0;
    constexpr unsigned B = 8;
    double a = 0.0, b = 0.0, m0 = 0.0, m1 = 0.0;
    for (unsigned q0 = (unsigned)t; q0 < nb; q0 += 1024 * B) {
        double va[B], vb[B];
#pragma unroll
        for (unsigned u = 0; u < B; ++u) {
            const unsigned q = q0 + 1024 * u;
            va[u] = q < nb ? part[2 * nb + q] : 0.0;
            vb[u] = q < nb ? part[3 * nb + q] : 0.0;
        }
#pragma unroll
        for (unsigned u = 0; u < B; ++u) {
            a += va[u];
            b += vb[u];
        }
        if (first) {
            for (unsigned u = 0; u < B; ++u) {
                const unsigned q = q0 + 1024 * u;
                if (q < nb) {
                    m0 = nanmax_any(m0, part[q]);
                    m1 = nanmax_any(m1, part[nb + q]);
                }
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        m0 = nanmax_any(m0, __shfl_xor(m0, off, 64));
        m1 = nanmax_any(m1, __shfl_xor(m1, off, 64));
    }
    if ((t & 63) == 0) {
        red[0][t >> 6] = a;
        red[1][t >> 6] = b;
        red[2][t >> 6] = m0;
        red[3][t >> 6] = m1;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
    for (int w = 0; w < 16; ++w) {
        a += red[0][w];
        b += red[1][w];
    }
    const double ra = sqrt(a / n_cells), rb = sqrt(b / n_cells);
    const double sr = ra > 1e-300 ? ra : 1.0, sp = rb > 1e-300 ? rb : 1.0;
    const double ir = 1.0 / sr, ip = 1.0 / sp;
    for (long v = (long)blockIdx.x * 1024 + t; v < V; v += (long)gridDim.x * 1024) {
        bf16x8 lo = {}, hi = {};
        lo[0] = (bf16)(float)(vsum[2 * v] * w * ir);
        lo[1] = (bf16)(float)(vsum[2 * v + 1] * w * ip);
        *reinterpret_cast<bf16x8 *>(x + v * 16) = lo;
        *reinterpret_cast<bf16x8 *>(x + v * 16 + 8) = hi;
    }
    if (first && t == 0) {
        stats[0] = sr;
        stats[1] = sp;
        double mz = red[2][0], mp = red[3][0];
        for (int w = 1; w < 16; ++w) {
            mz = nanmax_any(mz, red[2][w]);
            mp = nanmax_any(mp, red[3][w]);
        }
        const double ew = __longlong_as_double((long long)err[0]);
        if (rz_slot) {
            rz_slot[0] = mz;
            rz_slot[1] = ew;
        }
        if (rp_slot) {
            rp_slot[0] = mp;
            rp_slot[1] = ew;
        }
        __threadfence_system();
        if (rz_slot) reinterpret_cast<volatile unsigned long long *>(rz_slot)[2] = 0ull;
        if (rp_slot) reinterpret_cast<volatile unsigned long long *>(rp_slot)[2] = 0ull;
    }
}

// coarse cell-centred coordinate of fine cell i: (i + 1/2)/s - 1/2 = i0 + t, i0 clamped to the grid
// (the outermost fine cells take the nearest voxel's value)
__device__ __forceinline__ void prolong_axis(int i, int s, int n, int &i0, int &i1, double &t) {
    const double xc = ((double)i + 0.5) / (double)s - 0.5;
    double fl = floor(xc);
    t = xc - fl;
    i0 = (int)fl;
    i1 = i0 + 1;
    if (i0 < 0) { i0 = 0; t = 0.0; }
    if (i1 > n - 1) { i1 = n - 1; }
}

// Phi guess on the solver grid = Phi_0 + (trilinear prolongation of out) * s_r * c0, Phi_0 read
// from phi0 or formed as fac * prev.  S = s (2 or
// 4) as a constant: the division by s is then by a power-of-two constant, the same value.  One
// thread per fine cell, on a grid of ceil(FW/64) x ceil(FH/4) x FD workgroups of 64 x 4.
template <int S>
__global__ __launch_bounds__(256) void prolong_guess_kernel(FineGrid f, const float *__restrict__ out,
                                                            const double *__restrict__ stats, double c0,
                                                            const double *phi0, const double *prev, double fac,
                                                            double *phi) {
    constexpr int s = S;
    const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
    if (i >= f.FW || j >= f.FH) return;
    const double sc = stats[0] * c0;
    int x0, x1, y0, y1, z0, z1;
    double tx, ty, tz;
    prolong_axis(i, s, f.W, x0, x1, tx);
    prolong_axis(j, s, f.H, y0, y1, ty);
    prolong_axis(k, s, f.D, z0, z1, tz);
    auto at = [&](int x, int y, int z) { return (double)out[(z * f.H + y) * f.W + x]; };
    const double c00 = at(x0, y0, z0) * (1.0 - tx) + at(x1, y0, z0) * tx;
    const double c10 = at(x0, y1, z0) * (1.0 - tx) + at(x1, y1, z0) * tx;
    const double c01 = at(x0, y0, z1) * (1.0 - tx) + at(x1, y0, z1) * tx;
    const double c11 = at(x0, y1, z1) * (1.0 - tx) + at(x1, y1, z1) * tx;
    const double val = (c00 * (1.0 - ty) + c10 * ty) * (1.0 - tz) + (c01 * (1.0 - ty) + c11 * ty) * tz;
    const long v = ((long)k * f.FH + j) * f.FW + i;
    const double base = phi0 ? phi0[v] : (prev ? fac * prev[v] : 0.0);   // Phi_0, or formed as the input pass does
    phi[v] = base + val * sc;
}

// training target at the network grid: the mean over each voxel's s^3 cells of
// (Phi - fac Phi_prev) / (s_r c0)
__global__ void restrict_target_kernel(FineGrid f, const double *__restrict__ phi, const double *__restrict__ phi_prev,
                                       double fac, const double *__restrict__ stats, double c0, float *__restrict__ t) {
    const double sc = 1.0 / (stats[0] * c0) / (double)(f.s * f.s * f.s);
    const long V = (long)f.W * f.H * f.D;
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        const int vx = (int)(v % f.W), vy = (int)((v / f.W) % f.H), vz = (int)(v / ((long)f.W * f.H));
        double acc = 0.0;
        for (int c = 0; c < f.s; ++c)
            for (int b = 0; b < f.s; ++b) {
                const long row = (long)f.FW * ((long)(f.s * vy + b) + (long)f.FH * (f.s * vz + c)) + (long)f.s * vx;
                for (int a = 0; a < f.s; ++a) acc += phi[row + a] - (phi_prev ? fac * phi_prev[row + a] : 0.0);
            }
        t[v] = (float)(acc * sc);
    }
}

// replay buffer: store / load the 2 used input channels and the target
__global__ void pack2_kernel(const bf16 *__restrict__ x, bf16 *__restrict__ x2, long V) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        x2[2 * v] = x[v * 16];
        x2[2 * v + 1] = x[v * 16 + 1];
    }
}
__global__ void unpack2_kernel(const bf16 *__restrict__ x2, bf16 *__restrict__ x, long V) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        bf16x8 a = {}, b = {};
        a[0] = x2[2 * v];
        a[1] = x2[2 * v + 1];
        *reinterpret_cast<bf16x8 *>(x + v * 16) = a;
        *reinterpret_cast<bf16x8 *>(x + v * 16 + 8) = b;
    }
}
__global__ void input_from_f32_kernel(const float *__restrict__ x2, bf16 *__restrict__ x, long V) {
    for (long v = (long)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (long)gridDim.x * blockDim.x) {
        bf16x8 a = {}, b = {};
        a[0] = (bf16)x2[2 * v];
        a[1] = (bf16)x2[2 * v + 1];
        *reinterpret_cast<bf16x8 *>(x + v * 16) = a;
        *reinterpret_cast<bf16x8 *>(x + v * 16 + 8) = b;
    }
}

// ------------------------------------------------------------------------------- launchers
// reductions into a few accumulators (bias and head gradients, the loss): a bounded grid, so each
// thread sums many voxels and the atomic additions stay few (one per wave and accumulator)
static constexpr unsigned RED_BLOCKS = 512;
static inline unsigned nred(long n) { return (unsigned)std::min<long>((n + 255) / 256, RED_BLOCKS); }

static inline unsigned nblk(long n, int t = 256) {
    long b = (n + t - 1) / t;
    return (unsigned)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

static float *Wt(mdx_net *n, int l) { return n->params + n->woff[l]; }
static float *Bs(mdx_net *n, int l) { return n->params + n->boff[l]; }
static float *dWt(mdx_net *n, int l) { return n->grads + n->woff[l]; }
static float *dBs(mdx_net *n, int l) { return n->grads + n->boff[l]; }

// 3x3x3 conv (implicit GEMM on MFMA): forward (relu, bias) or input gradient (plain)
// up != nullptr: the input's channels [0, cu) are the nearest up-sampling of `up` and the rest
// `in` (conv_in_at); only the tiled and split-K kernels take such an input
// pool != nullptr: also the output's 2x2x2 averages (the tiled 16 -> 16 ReLU convolution only;
// hipErrorInvalidValue otherwise, before anything runs)
static hipError_t conv3(mdx_net *n, int cin, int cout, bool relu, const bf16 *in, const bf16 *wp, const float *bias,
                        bf16 *out, int lvl, const bf16 *up = nullptr, int cu = 0, bf16 *pool = nullptr) {
    const int W = n->W[lvl], H = n->H[lvl], D = n->D[lvl];
    const unsigned blocks = (unsigned)(n->V[lvl] / 64);
    if (pool) {
        if (!(n->tiled_conv && cin == 16 && cout == 16 && relu && !up && W % CTX == 0 && H % CTY == 0 &&
              D % ConvTile<16>::TZ == 0))
            return hipErrorInvalidValue;
        const unsigned tiles = (unsigned)((W / CTX) * (H / CTY) * (D / ConvTile<16>::TZ));
        hipLaunchKernelGGL((conv3_tiled_kernel<16, 16, true, true>), dim3(tiles), dim3(256), ConvTile<16>::BYTES,
                           n->stream, in, nullptr, 0, reinterpret_cast<const bf16x8 *>(wp), bias, out, W, H, D, pool);
        return hipGetLastError();
    }
#define TILED_CASE(CI, CO, R)                                                                                      \
    if (n->tiled_conv && cin == CI && cout == CO && relu == R && W % CTX == 0 && H % CTY == 0 &&                   \
        D % ConvTile<CI>::TZ == 0) {                                                                               \
        const unsigned tiles = (unsigned)((W / CTX) * (H / CTY) * (D / ConvTile<CI>::TZ));                          \
        static const hipError_t attr = ConvTile<CI>::BYTES <= 65536 ? hipSuccess :                                 \
            hipFuncSetAttribute(reinterpret_cast<const void *>(&conv3_tiled_kernel<CI, CO, R>),                     \
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)ConvTile<CI>::BYTES);              \
        if (attr != hipSuccess) return attr;                                                                       \
        hipLaunchKernelGGL((conv3_tiled_kernel<CI, CO, R>), dim3(tiles), dim3(256), ConvTile<CI>::BYTES,           \
                           n->stream, in, up, cu, reinterpret_cast<const bf16x8 *>(wp), bias, out, W, H, D);       \
        return hipGetLastError();                                                                                  \
    }
    TILED_CASE(16, 16, true) TILED_CASE(16, 32, true) TILED_CASE(32, 32, true) TILED_CASE(32, 64, true)
    TILED_CASE(48, 16, true) TILED_CASE(64, 64, true) TILED_CASE(96, 32, true) TILED_CASE(16, 16, false) TILED_CASE(32, 16, false)
    TILED_CASE(32, 32, false) TILED_CASE(32, 96, false) TILED_CASE(16, 48, false) TILED_CASE(64, 32, false)
    TILED_CASE(64, 64, false)
#undef TILED_CASE
    // grids the tiled kernel cannot take and too small to fill the chip: split K
    const long mt = n->V[lvl] / 16;   // 16-voxel M-tiles
#define SPLIT_CASE(CI, CO, R)                                                                                      \
    if (n->split_conv && cin == CI && cout == CO && relu == R &&                                               \
        !(W % CTX == 0 && H % CTY == 0 && D % ConvTile<CI>::TZ == 0) &&                                           \
        n->V[lvl] % 16 == 0 && mt < 1024) {                                                                        \
        hipLaunchKernelGGL((conv3_splitk_kernel<CI, CO, R>), dim3((unsigned)mt), dim3(64 * SplitK<CO>::S), 0,       \
                           n->stream, in, up, cu, reinterpret_cast<const bf16x8 *>(wp), bias, out, W, H, D);       \
        return hipGetLastError();                                                                                  \
    }
    SPLIT_CASE(16, 16, true) SPLIT_CASE(16, 32, true) SPLIT_CASE(32, 32, true) SPLIT_CASE(32, 64, true)
    SPLIT_CASE(64, 64, true) SPLIT_CASE(96, 32, true) SPLIT_CASE(48, 16, true)
    SPLIT_CASE(16, 16, false) SPLIT_CASE(32, 16, false) SPLIT_CASE(32, 32, false) SPLIT_CASE(64, 32, false)
    SPLIT_CASE(64, 64, false) SPLIT_CASE(32, 96, false) SPLIT_CASE(16, 48, false)
#undef SPLIT_CASE
    if (up) return hipErrorInvalidValue;   // the direct kernel takes a materialised input only
#define CONV_CASE(CI, CO, R)                                                                                       \
    if (cin == CI && cout == CO && relu == R) {                                                                    \
        hipLaunchKernelGGL((conv3_fwd_kernel<CI, CO, R>), dim3(blocks), dim3(256), 0, n->stream, in,                \
                           reinterpret_cast<const bf16x8 *>(wp), bias, out, W, H, D);                              \
        return hipGetLastError();                                                                                  \
    }
    CONV_CASE(16, 16, true) CONV_CASE(16, 32, true) CONV_CASE(32, 32, true) CONV_CASE(32, 64, true)
    CONV_CASE(64, 64, true) CONV_CASE(96, 32, true) CONV_CASE(48, 16, true)
    CONV_CASE(16, 16, false) CONV_CASE(32, 16, false) CONV_CASE(32, 32, false) CONV_CASE(64, 32, false)
    CONV_CASE(64, 64, false) CONV_CASE(32, 96, false) CONV_CASE(16, 48, false)
#undef CONV_CASE
    return hipErrorInvalidValue;
}

static hipError_t wgrad(mdx_net *n, int l, const bf16 *in, const bf16 *dy) {
    const LayerDesc &L = LAYERS[l];
    const int lvl = L.level;
    const int W = n->W[lvl], H = n->H[lvl], D = n->D[lvl];
    const long chunks = n->V[lvl] / 32;
    int slices = (int)std::min<long>(48, std::max<long>(1, chunks / 16));
#define WG_CASE(CI, CO)                                                                                       \
    if (L.cin == CI && L.cout == CO) {                                                                       \
        hipLaunchKernelGGL((conv3_wgrad_kernel<CI, CO>), dim3(27 * slices), dim3(256), 0, n->stream, in, dy,    \
                           dWt(n, l), W, H, D, slices);                                                      \
    } else
    WG_CASE(16, 16) WG_CASE(16, 32) WG_CASE(32, 32) WG_CASE(32, 64) WG_CASE(64, 64) WG_CASE(96, 32) WG_CASE(48, 16)
    return hipErrorInvalidValue;
#undef WG_CASE
    const long V = n->V[lvl];
    if (L.cout == 16) hipLaunchKernelGGL(bias_grad_kernel<16>, dim3(nred(V * 2)), dim3(256), 0, n->stream, dy, dBs(n, l), V);
    else if (L.cout == 32) hipLaunchKernelGGL(bias_grad_kernel<32>, dim3(nred(V * 4)), dim3(256), 0, n->stream, dy, dBs(n, l), V);
    else hipLaunchKernelGGL(bias_grad_kernel<64>, dim3(nred(V * 8)), dim3(256), 0, n->stream, dy, dBs(n, l), V);
    return hipGetLastError();
}

static hipError_t relu_bwd(mdx_net *n, bf16 *g, const bf16 *extra, int Ce, int coff, int C, const bf16 *y, int lvl) {
    const long cnt = n->V[lvl] * C;
    hipLaunchKernelGGL(relu_bwd_kernel, dim3(nblk(cnt)), dim3(256), 0, n->stream, g, extra, Ce, coff, C, y, cnt);
    return hipGetLastError();
}

hipError_t repack_all(mdx_net *n) {
    for (int l = 0; l < NL; ++l) {
        const LayerDesc &L = LAYERS[l];
        if (L.ks != 3) continue;
        const long cf = (long)((27 * L.cin + 31) / 32) * (L.cout / 16) * 512;
        hipLaunchKernelGGL(repack_kernel, dim3(nblk(cf)), dim3(256), 0, n->stream, Wt(n, l), n->wp[l], L.cin, L.cout, 0);
        if (n->wd[l]) {
            const long cd = (long)((27 * L.cout + 31) / 32) * (L.cin / 16) * 512;
            hipLaunchKernelGGL(repack_kernel, dim3(nblk(cd)), dim3(256), 0, n->stream, Wt(n, l), n->wd[l], L.cin, L.cout, 1);
        }
    }
    return hipGetLastError();
}

hipError_t upload_input2(mdx_net *n, const float *x2) {
    const long V = n->V[0];
    hipError_t e = hipMemcpyAsync(n->x2_stage, x2, (size_t)V * 2 * 4, hipMemcpyHostToDevice, n->stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(input_from_f32_kernel, dim3(nblk(V)), dim3(256), 0, n->stream, n->x2_stage, n->x_in, V);
    return hipGetLastError();
}

hipError_t net_forward(mdx_net *n, bool infer) {
    hipError_t e;
    const int W0 = n->W[0], H0 = n->H[0], D0 = n->D[0], W1 = n->W[1], H1 = n->H[1], D1 = n->D[1];
#define TRY(x) if ((e = (x)) != hipSuccess) return e
    TRY(conv3(n, 16, 16, true, n->x_in, n->wp[0], Bs(n, 0), n->a[0], 0));
    // inference: the second convolution also writes the first pooling's averages
    e = infer && n->fuse_pool ? conv3(n, 16, 16, true, n->a[0], n->wp[1], Bs(n, 1), n->a[1], 0, nullptr, 0, n->p1)
                              : hipErrorInvalidValue;
    if (e == hipErrorInvalidValue) {
        TRY(conv3(n, 16, 16, true, n->a[0], n->wp[1], Bs(n, 1), n->a[1], 0));
        hipLaunchKernelGGL(pool_fwd_kernel, dim3(nblk(n->V[1] * 2)), dim3(256), 0, n->stream, n->a[1], n->p1, 16, W0,
                           H0, D0);
    } else {
        TRY(e);
    }
    TRY(conv3(n, 16, 32, true, n->p1, n->wp[2], Bs(n, 2), n->a[2], 1));
    TRY(conv3(n, 32, 32, true, n->a[2], n->wp[3], Bs(n, 3), n->a[3], 1));
    hipLaunchKernelGGL(pool_fwd_kernel, dim3(nblk(n->V[2] * 4)), dim3(256), 0, n->stream, n->a[3], n->p2, 32, W1, H1, D1);
    TRY(conv3(n, 32, 64, true, n->p2, n->wp[4], Bs(n, 4), n->a[4], 2));
    TRY(conv3(n, 64, 64, true, n->a[4], n->wp[5], Bs(n, 5), n->a[5], 2));
    // decoder inputs: inference reads the up-sampling and the skip tensor where they lie (the
    // convolution's loader forms the concat); training keeps the concat for the backward pass
    const bool fuse = infer && n->fuse_cat;
    e = fuse ? conv3(n, 96, 32, true, n->a[3], n->wp[6], Bs(n, 6), n->a[6], 1, n->a[5], 64) : hipErrorInvalidValue;
    if (e == hipErrorInvalidValue) {
        hipLaunchKernelGGL(upcat_kernel, dim3(nblk(n->V[1] * 12)), dim3(256), 0, n->stream, n->a[5], n->a[3], n->cat2,
                           64, 32, W1, H1, D1);
        e = conv3(n, 96, 32, true, n->cat2, n->wp[6], Bs(n, 6), n->a[6], 1);
    }
    TRY(e);
    TRY(conv3(n, 32, 32, true, n->a[6], n->wp[7], Bs(n, 7), n->a[7], 1));
    e = fuse ? conv3(n, 48, 16, true, n->a[1], n->wp[8], Bs(n, 8), n->a[8], 0, n->a[7], 32) : hipErrorInvalidValue;
    if (e == hipErrorInvalidValue) {
        hipLaunchKernelGGL(upcat_kernel, dim3(nblk(n->V[0] * 6)), dim3(256), 0, n->stream, n->a[7], n->a[1], n->cat1,
                           32, 16, W0, H0, D0);
        e = conv3(n, 48, 16, true, n->cat1, n->wp[8], Bs(n, 8), n->a[8], 0);
    }
    TRY(e);
    TRY(conv3(n, 16, 16, true, n->a[8], n->wp[9], Bs(n, 9), n->a[9], 0));
    hipLaunchKernelGGL(head_fwd_kernel, dim3(nblk(n->V[0])), dim3(256), 0, n->stream, n->a[9], Wt(n, 10), Bs(n, 10),
                       n->out, n->V[0]);
    return hipGetLastError();
}

// loss = mean (out - target)^2 of this sample; parameter gradients scaled by `scale` are added
hipError_t net_backward(mdx_net *n, float scale) {
    hipError_t e;
    const long V0 = n->V[0];
    const int W0 = n->W[0], H0 = n->H[0], D0 = n->D[0], W1 = n->W[1], H1 = n->H[1], D1 = n->D[1];
    TRY(hipMemsetAsync(n->loss, 0, 4, n->stream));
    hipLaunchKernelGGL(mse_kernel, dim3(nred(V0)), dim3(256), 0, n->stream, n->out, n->target, n->dout,
                       scale / (float)V0, V0, n->loss);
    hipLaunchKernelGGL(head_bwd_kernel, dim3(nred(V0)), dim3(256), 0, n->stream, n->a[9], Wt(n, 10), n->dout, n->g[9],
                       dWt(n, 10), dBs(n, 10), V0);
    // decoder level 0
    TRY(relu_bwd(n, n->g[9], nullptr, 0, 0, 16, n->a[9], 0));
    TRY(wgrad(n, 9, n->a[8], n->g[9]));
    TRY(conv3(n, 16, 16, false, n->g[9], n->wd[9], nullptr, n->g[8], 0));
    TRY(relu_bwd(n, n->g[8], nullptr, 0, 0, 16, n->a[8], 0));
    TRY(wgrad(n, 8, n->cat1, n->g[8]));
    TRY(conv3(n, 16, 48, false, n->g[8], n->wd[8], nullptr, n->gcat1, 0));
    // decoder level 1
    hipLaunchKernelGGL(upcat_bwd_coarse_kernel, dim3(nblk(n->V[1] * 32)), dim3(256), 0, n->stream, n->gcat1, n->g[7], 32,
                       16, W0, H0, D0);
    TRY(relu_bwd(n, n->g[7], nullptr, 0, 0, 32, n->a[7], 1));
    TRY(wgrad(n, 7, n->a[6], n->g[7]));
    TRY(conv3(n, 32, 32, false, n->g[7], n->wd[7], nullptr, n->g[6], 1));
    TRY(relu_bwd(n, n->g[6], nullptr, 0, 0, 32, n->a[6], 1));
    TRY(wgrad(n, 6, n->cat2, n->g[6]));
    TRY(conv3(n, 32, 96, false, n->g[6], n->wd[6], nullptr, n->gcat2, 1));
    // bottleneck level 2
    hipLaunchKernelGGL(upcat_bwd_coarse_kernel, dim3(nblk(n->V[2] * 64)), dim3(256), 0, n->stream, n->gcat2, n->g[5], 64,
                       32, W1, H1, D1);
    TRY(relu_bwd(n, n->g[5], nullptr, 0, 0, 64, n->a[5], 2));
    TRY(wgrad(n, 5, n->a[4], n->g[5]));
    TRY(conv3(n, 64, 64, false, n->g[5], n->wd[5], nullptr, n->g[4], 2));
    TRY(relu_bwd(n, n->g[4], nullptr, 0, 0, 64, n->a[4], 2));
    TRY(wgrad(n, 4, n->p2, n->g[4]));
    TRY(conv3(n, 64, 32, false, n->g[4], n->wd[4], nullptr, n->gp2, 2));
    // encoder level 1 (pool backward + skip gradient from the concat)
    hipLaunchKernelGGL(pool_bwd_kernel, dim3(nblk(n->V[1] * 32)), dim3(256), 0, n->stream, n->gp2, n->g[3], 32, W1, H1, D1);
    TRY(relu_bwd(n, n->g[3], n->gcat2, 96, 64, 32, n->a[3], 1));
    TRY(wgrad(n, 3, n->a[2], n->g[3]));
    TRY(conv3(n, 32, 32, false, n->g[3], n->wd[3], nullptr, n->g[2], 1));
    TRY(relu_bwd(n, n->g[2], nullptr, 0, 0, 32, n->a[2], 1));
    TRY(wgrad(n, 2, n->p1, n->g[2]));
    TRY(conv3(n, 32, 16, false, n->g[2], n->wd[2], nullptr, n->gp1, 1));
    // encoder level 0
    hipLaunchKernelGGL(pool_bwd_kernel, dim3(nblk(n->V[0] * 16)), dim3(256), 0, n->stream, n->gp1, n->g[1], 16, W0, H0, D0);
    TRY(relu_bwd(n, n->g[1], n->gcat1, 48, 32, 16, n->a[1], 0));
    TRY(wgrad(n, 1, n->a[0], n->g[1]));
    TRY(conv3(n, 16, 16, false, n->g[1], n->wd[1], nullptr, n->g[0], 0));
    TRY(relu_bwd(n, n->g[0], nullptr, 0, 0, 16, n->a[0], 0));
    TRY(wgrad(n, 0, n->x_in, n->g[0]));
    return hipGetLastError();
}

hipError_t net_adam(mdx_net *n, float lr) {
    n->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - powf(b1, (float)n->adam_t), c2 = 1.f - powf(b2, (float)n->adam_t);
    hipLaunchKernelGGL(adam_kernel, dim3(nblk(n->nparam)), dim3(256), 0, n->stream, n->params, n->grads, n->adam_m,
                       n->adam_v, (long)n->nparam, lr, b1, b2, 1e-8f, c1, c2, 1.0f);
    hipError_t e;
    TRY(hipMemsetAsync(n->grads, 0, n->nparam * 4, n->stream));
    return repack_all(n);
}

// ------------------------------------------------------------------------------- solver coupling
static FineGrid fine_grid(const mdx_net *n) {
    return FineGrid{n->W[0], n->H[0], n->D[0], n->stride, n->FW, n->FH, n->FD};
}

// rms over the solver grid's interior cells
static hipError_t rms_into(mdx_net *n, const double *a, const double *b, double fac, double *stat) {
    const long V = n->FV;
    hipLaunchKernelGGL(input_stats_partial, dim3(512), dim3(256), 0, n->stream, V, a, b, fac, n->part);
    hipLaunchKernelGGL(input_stats_final, dim3(1), dim3(64), 0, n->stream, n->part, 512, (double)V, stat);
    return hipGetLastError();
}

// stats_ready: stats[0..1] already hold rms(r0) and rms(Phi_0) (the solver's input pass)
// the prolongation at the network's stride (a compile-time constant for 2 and 4)
static void launch_prolong(mdx_net *n, const float *coarse, const double *stats, double c0, const double *phi0,
                           double *phi, const double *prev = nullptr, double fac = 0.0) {
    const FineGrid f = fine_grid(n);
    const dim3 grd((f.FW + 63) / 64, (f.FH + 3) / 4, f.FD), blk(64, 4);
    if (f.s == 4)
        hipLaunchKernelGGL(prolong_guess_kernel<4>, grd, blk, 0, n->stream, f, coarse, stats, c0, phi0, prev, fac, phi);
    else   // stride 2
        hipLaunchKernelGGL(prolong_guess_kernel<2>, grd, blk, 0, n->stream, f, coarse, stats, c0, phi0, prev, fac, phi);
}

hipError_t net_input_from_solver(mdx_net *n, const mdx_handle *h, const double *r0, const double *phi0,
                                 bool stats_ready) {
    (void)h;
    hipError_t e;
    const long V = n->V[0];
    if (stats_ready) {
    } else if (phi0) {   // rms of r0 and of Phi_0 into stats[0..1]
        hipLaunchKernelGGL(input_stats2_partial, dim3(512), dim3(256), 0, n->stream, n->FV, r0, phi0, n->part);
        hipLaunchKernelGGL(input_stats2_final, dim3(1), dim3(128), 0, n->stream, n->part, 512, (double)n->FV,
                           n->stats);
    } else {
        TRY(rms_into(n, r0, nullptr, 0.0, n->stats));
    }
    if (n->stride == 1)
        hipLaunchKernelGGL(make_input_kernel, dim3(nblk(V)), dim3(256), 0, n->stream, V, r0, phi0, n->stats, n->x_in);
    else if (n->stride == 4)
        hipLaunchKernelGGL(restrict_input_kernel<4>, dim3((n->W[0] + 63) / 64, n->H[0], n->D[0]), dim3(64, 16), 0,
                           n->stream, fine_grid(n), r0, phi0, n->stats, n->x_in);
    else   // stride 2 (the strides are 1, 2 and 4)
        hipLaunchKernelGGL(restrict_input_kernel<2>, dim3((n->W[0] + 63) / 64, n->H[0], n->D[0]), dim3(64, 4), 0,
                           n->stream, fine_grid(n), r0, phi0, n->stats, n->x_in);
    return hipGetLastError();
}

hipError_t net_guess_to_phi(mdx_net *n, const double *phi0, double *phi) {
    hipError_t e;
    TRY(net_forward(n, true));
    if (n->stride == 1)
        hipLaunchKernelGGL(guess_out_kernel, dim3(nblk(n->V[0])), dim3(256), 0, n->stream, n->out, n->stats, n->c0,
                           phi0, phi, n->V[0]);
    else
        launch_prolong(n, n->out, n->stats, n->c0, phi0, phi);
    return hipGetLastError();
}

hipError_t net_guess_to_phi_scaled(mdx_net *n, const double *prev, double fac, double *phi) {
    hipError_t e;
    if (n->stride == 1) return hipErrorInvalidValue;
    TRY(net_forward(n, true));
    launch_prolong(n, n->out, n->stats, n->c0, nullptr, phi, prev, fac);
    return hipGetLastError();
}

hipError_t net_input_from_sums(mdx_net *n, const double *part, unsigned nb, double *rz_slot, double *rp_slot,
                               const unsigned long long *err) {
    if (n->stride != 2 && n->stride != 4) return hipErrorInvalidValue;
    const long V = n->V[0];
    const unsigned blocks = (unsigned)((V + 1023) / 1024);
    hipLaunchKernelGGL(input_from_sums_kernel, dim3(blocks), dim3(1024), 0, n->stream, V, n->vsum, part, nb,
                       (double)n->FV, 1.0 / (double)(n->stride * n->stride * n->stride), n->stats, n->x_in,
                       rz_slot, rp_slot, err);
    return hipGetLastError();
}

// append (current input, target of the converged Phi) to the replay ring buffer
hipError_t net_capture(mdx_net *n, const double *phi, const double *phi_prev, double fac) {
    const long V = n->V[0];
    const int slot = n->next;
    hipLaunchKernelGGL(pack2_kernel, dim3(nblk(V)), dim3(256), 0, n->stream, n->x_in, n->rb_x + (size_t)slot * 2 * V, V);
    if (n->stride == 1)
        hipLaunchKernelGGL(make_target_kernel, dim3(nblk(V)), dim3(256), 0, n->stream, phi, phi_prev, fac, n->stats,
                           n->c0, n->rb_t + (size_t)slot * V, V);
    else
        hipLaunchKernelGGL(restrict_target_kernel, dim3(nblk(V)), dim3(256), 0, n->stream, fine_grid(n), phi, phi_prev,
                           fac, n->stats, n->c0, n->rb_t + (size_t)slot * V);
    n->next = (n->next + 1) % n->cap;
    if (n->count < n->cap) n->count += 1;
    return hipGetLastError();
}

// loss of the current output against the current target (no gradient accumulation)
hipError_t net_mse_only(mdx_net *n) {
    hipError_t e = hipMemsetAsync(n->loss, 0, 4, n->stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(mse_kernel, dim3(nred(n->V[0])), dim3(256), 0, n->stream, n->out, n->target, n->dout,
                       1.0f / (float)n->V[0], n->V[0], n->loss);
    return hipGetLastError();
}

// rms of the first training target's numerator, Phi - fac * Phi_prev, into stats[3]
hipError_t net_rms_to_stats3(mdx_net *n, const mdx_handle *h, const double *phi, const double *phi_prev, double fac) {
    (void)h;
    return rms_into(n, phi, phi_prev, fac, n->stats + 3);
}

// the transfer operators alone on device buffers: coarse = restriction of fine (mean over
// stride^3 cells), fine = trilinear prolongation of coarse (stats[6] holds the unit scale)
hipError_t net_restrict(mdx_net *n, const double *fine, float *coarse) {
    const double one = 1.0;
    hipError_t e = hipMemcpyAsync(n->stats + 6, &one, 8, hipMemcpyHostToDevice, n->stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(restrict_target_kernel, dim3(nblk(n->V[0])), dim3(256), 0, n->stream, fine_grid(n), fine, nullptr,
                       0.0, n->stats + 6, 1.0, coarse);
    return hipGetLastError();
}

hipError_t net_prolong(mdx_net *n, const float *coarse, double *fine) {
    const double one = 1.0;
    hipError_t e = hipMemcpyAsync(n->stats + 6, &one, 8, hipMemcpyHostToDevice, n->stream);
    if (e != hipSuccess) return e;
    launch_prolong(n, coarse, n->stats + 6, 1.0, nullptr, fine);
    return hipGetLastError();
}

hipError_t net_load_replay(mdx_net *n, int slot) {
    const long V = n->V[0];
    hipLaunchKernelGGL(unpack2_kernel, dim3(nblk(V)), dim3(256), 0, n->stream, n->rb_x + (size_t)slot * 2 * V, n->x_in, V);
    return hipMemcpyAsync(n->target, n->rb_t + (size_t)slot * V, (size_t)V * 4, hipMemcpyDeviceToDevice, n->stream);
}
#undef TRY

}  // namespace unet
}  // namespace mdx
