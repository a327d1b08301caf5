// Decode-fused GEMMs of the randomized-SVD initializer (FP32, as the
// reference's SGEMMs).  X is the centred genotype matrix of
// pack2bit.cu:46-77 / tools.pyx:292-324: X[m, i] = g - 2 f[m], 0 if missing.
// The reference materialises X chunk by chunk (C x N FP32) and calls cuBLAS
// (svd_gpu.py:86, :106-107, :138); here the 2-bit codes are decoded inside the
// GEMM and X never exists in memory.
//
//   svd_xw_kernel:   Y = X W          (M x L; W is N x L)    reduction over samples
//   svd_xty_kernel:  A += X^T Y       (N x L; Y is M x L)    reduction over SNPs
// When the grid splits the reduction, each split stores its slice of the
// result in the caller's workspace and svd_sum_kernel adds the slices in a
// fixed order (deterministic); without a workspace the splits meet in
// FP32 atomics.
//
// Both run 16x16x4 FP32 MFMA tiles with L padded to 16*LT rows:
//   xw:  Y^T (L x SNPs) += W^T (L x 4 samples) . X^T (4 samples x SNPs).  The
//        genotype operand comes straight from the EM kernel's per-lane 16-byte
//        row loads (lane (g, c): byte row 16t+c, samples s0+16g .. s0+16g+15).
//   xty: A^T (L x samples) += Y^T (L x 4 SNPs) . X (4 SNPs x samples).  A
//        64-row x 64-sample genotype tile and its 256 Y rows are staged in
//        LDS so that lane (g, c) reads field g of byte [row s][sample c] and
//        Y[4s + g][li] for the k-step of byte row s.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

__device__ __forceinline__ float centred(uint32_t code, float twof) {
    return code == 3u ? 0.0f : (float)code - twof;
}

template <int LT, int CW>
__global__ __launch_bounds__(256) void svd_xw_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                     int64_t N, const float* __restrict__ f,
                                                     const float* __restrict__ W, int L, int64_t n_per_split,
                                                     float* __restrict__ Y, float* __restrict__ part) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int64_t snp_w = ((int64_t)blockIdx.x * 4 + wave) * (16 * CW);
    const int64_t row_w = snp_w >> 2;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t n0 = (int64_t)blockIdx.y * n_per_split;
    const int64_t n1 = n0 + n_per_split < N ? n0 + n_per_split : N;

    float twof[CW];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int64_t snp = snp_w + 64 * (cb >> 2) + 4 * cl + (cb & 3);
        twof[cb] = snp < M ? 2.0f * f[snp] : 0.0f;
    }
    floatx4 acc[CW][LT];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb)
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) acc[cb][lt] = floatx4{0, 0, 0, 0};

    for (int64_t s0 = n0; s0 < n1; s0 += 64) {
        uint32_t gw[CW / 4][4];
#pragma unroll
        for (int t = 0; t < CW / 4; ++t) {
            const int64_t row = row_w + 16 * t + cl;
            if (row < Mb) {
                const uint4 v = *reinterpret_cast<const uint4*>(G + row * ldg + s0 + 16 * g);
                gw[t][0] = v.x; gw[t][1] = v.y; gw[t][2] = v.z; gw[t][3] = v.w;
                const int64_t nf = M - 4 * row;
                if (nf < 4) {
                    const uint32_t pm = ((0xFFu << (2 * nf)) & 0xFFu) * 0x01010101u;
#pragma unroll
                    for (int r = 0; r < 4; ++r) gw[t][r] |= pm;
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) gw[t][r] = 0xFFFFFFFFu;
            }
        }
        // k-step e: the 4 samples s0 + 16g + e (g = 0..3).
#pragma unroll 4
        for (int e = 0; e < 16; ++e) {
            const int64_t smp = s0 + 16 * g + e;
            float wv[LT];
#pragma unroll
            for (int lt = 0; lt < LT; ++lt) {
                const int li = 16 * lt + cl;
                wv[lt] = (smp < N && li < L) ? W[smp * L + li] : 0.0f;
            }
#pragma unroll
            for (int cb = 0; cb < CW; ++cb) {
                const uint32_t code = (gw[cb >> 2][e >> 2] >> (8 * (e & 3) + 2 * (cb & 3))) & 3u;
                const float x = centred(code, twof[cb]);
#pragma unroll
                for (int lt = 0; lt < LT; ++lt) acc[cb][lt] = Mma16<float>::run(wv[lt], x, acc[cb][lt]);
            }
        }
    }
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int64_t snp = snp_w + 64 * (cb >> 2) + 4 * cl + (cb & 3);
        if (snp >= M) continue;
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int li = 16 * lt + 4 * g + j;
                if (li < L) {
                    if (part)   // this sample split's slice, summed in order by svd_sum_kernel
                        part[((int64_t)blockIdx.y * M + snp) * L + li] = acc[cb][lt][j];
                    else
                        atomicAdd(&Y[snp * L + li], acc[cb][lt][j]);
                }
            }
    }
}

// Workgroup: 256 samples (64 per wave, as 4 sub-tiles of 16), one SNP range;
// chunks of 64 byte rows (256 SNPs) staged in LDS together with their Y rows
// (transposed, li-major, row stride YS = 260 floats so that lane (g, c)
// reading Y[4s + g][16 lt + c] hits bank 4c + g: all 64 lanes distinct) and
// their 2 f values.  The inner loop reads LDS only, and each Y value read
// feeds the 4 sub-tiles of the wave.
constexpr int XTY_SUB = 4;

template <int LT>
__global__ __launch_bounds__(256) void svd_xty_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                      int64_t N, const float* __restrict__ f,
                                                      const float* __restrict__ Y, int L, int64_t rows_per_split,
                                                      float* __restrict__ A, float* __restrict__ part) {
    constexpr int YS = 260;
    constexpr int WS = 64 * XTY_SUB;          // samples per workgroup
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int64_t s0 = (int64_t)blockIdx.x * WS;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_split;
    const int64_t r1 = r0 + rows_per_split < Mb ? r0 + rows_per_split : Mb;
    __shared__ __attribute__((aligned(16))) uint8_t tile[64][WS];
    __shared__ float sY[16 * LT][YS];
    __shared__ float s2f[256];

    floatx4 acc[XTY_SUB][LT];
#pragma unroll
    for (int q = 0; q < XTY_SUB; ++q)
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) acc[q][lt] = floatx4{0, 0, 0, 0};

    for (int64_t rb = r0; rb < r1; rb += 64) {
        const int nrows = (int)(r1 - rb < 64 ? r1 - rb : 64);
        const int64_t snp0 = 4 * rb;
        // stage 64 byte rows x WS samples, 16 bytes per thread and step
        for (int e = threadIdx.x; e < 64 * (WS / 16); e += 256) {
            const int r = e / (WS / 16), c = 16 * (e - r * (WS / 16));
            uint4 v = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
            if (r < nrows && s0 + c < ldg) v = *reinterpret_cast<const uint4*>(G + (rb + r) * ldg + s0 + c);
            *reinterpret_cast<uint4*>(&tile[r][c]) = v;
        }
        {   // the chunk's 256 SNPs: 2 f and Y rows (zero beyond M and L)
            const int64_t snp = snp0 + threadIdx.x;
            s2f[threadIdx.x] = snp < M ? 2.0f * f[snp] : 0.0f;
            for (int e = threadIdx.x; e < 256 * 16 * LT; e += 256) {
                const int sl = e / (16 * LT), li = e - sl * (16 * LT);
                const int64_t sn = snp0 + sl;
                sY[li][sl] = (sn < M && li < L) ? Y[sn * L + li] : 0.0f;
            }
        }
        __syncthreads();
        for (int s = 0; s < nrows; ++s) {
            const int sl = 4 * s + g;
            const bool live = snp0 + sl < M;
            const float twof = s2f[sl];
            float yv[LT];
#pragma unroll
            for (int lt = 0; lt < LT; ++lt) yv[lt] = sY[16 * lt + cl][sl];
#pragma unroll
            for (int q = 0; q < XTY_SUB; ++q) {
                uint32_t code = (tile[s][64 * wave + 16 * q + cl] >> (2 * g)) & 3u;
                if (!live) code = 3u;
                const float x = centred(code, twof);
#pragma unroll
                for (int lt = 0; lt < LT; ++lt) acc[q][lt] = Mma16<float>::run(yv[lt], x, acc[q][lt]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < XTY_SUB; ++q) {
        const int64_t smp = s0 + 64 * wave + 16 * q + cl;   // B-operand column of sub-tile q
        if (smp >= N) continue;
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int li = 16 * lt + 4 * g + j;
                if (li < L) {
                    if (part)   // this SNP split's slice, summed in order by svd_sum_kernel
                        part[((int64_t)blockIdx.y * N + smp) * L + li] = acc[q][lt][j];
                    else
                        atomicAdd(&A[smp * L + li], acc[q][lt][j]);
                }
            }
    }
}

// out[e] = sum over the nparts slices of part[p * n + e], in order.
__global__ void svd_sum_kernel(const float* __restrict__ part, int64_t nparts, int64_t n, float* __restrict__ out) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float s = 0.0f;
    for (int64_t p = 0; p < nparts; ++p) s += part[p * n + e];
    out[e] = s;
}

// Launch shapes: xw splits the samples, xty the byte rows, until the grid
// holds about 2048 workgroups.
struct SvdGrid {
    int64_t nbx, splits, per_split;
};

static SvdGrid xw_grid(int64_t M, int64_t N) {
    constexpr int CW = 4;
    const int64_t nbx = ceil_div(M, 64 * CW);
    int64_t s = ceil_div(2048, nbx);
    const int64_t max_s = ceil_div(N, 64);
    s = s > max_s ? max_s : (s < 1 ? 1 : s);
    if (s > 65535) s = 65535;
    const int64_t nps = ceil_div(ceil_div(N, s), 64) * 64;
    return {nbx, ceil_div(N, nps), nps};
}

static SvdGrid xty_grid(int64_t M, int64_t N) {
    const int64_t Mb = ceil_div(M, 4);
    const int64_t nbx = ceil_div(N, 64 * XTY_SUB);
    int64_t s = ceil_div(2048, nbx);
    const int64_t max_s = ceil_div(Mb, 64);
    s = s > max_s ? max_s : (s < 1 ? 1 : s);
    if (s > 65535) s = 65535;
    const int64_t rps = ceil_div(ceil_div(Mb, s), 64) * 64;
    return {nbx, ceil_div(Mb, rps), rps};
}

template <int LT>
static void launch_xw(hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f,
                      const float* W, int L, float* Y, float* work) {
    constexpr int CW = 4;
    const SvdGrid gd = xw_grid(M, N);
    float* part = (work && gd.splits > 1) ? work : nullptr;
    dim3 grid((unsigned)gd.nbx, (unsigned)gd.splits);
    svd_xw_kernel<LT, CW><<<grid, 256, 0, st>>>(G, ldg, M, N, f, W, L, gd.per_split, Y, part);
    if (part)
        svd_sum_kernel<<<(unsigned)ceil_div(M * L, 256), 256, 0, st>>>(part, gd.splits, M * (int64_t)L, Y);
}

template <int LT>
static void launch_xty(hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f,
                       const float* Y, int L, float* A, float* work) {
    const SvdGrid gd = xty_grid(M, N);
    float* part = (work && gd.splits > 1) ? work : nullptr;
    dim3 grid((unsigned)gd.nbx, (unsigned)gd.splits);
    svd_xty_kernel<LT><<<grid, 256, 0, st>>>(G, ldg, M, N, f, Y, L, gd.per_split, A, part);
    if (part)
        svd_sum_kernel<<<(unsigned)ceil_div(N * L, 256), 256, 0, st>>>(part, gd.splits, N * (int64_t)L, A);
}

}  // namespace adx

using namespace adx;

extern "C" {

int64_t adx_svd_work_floats(int which, int64_t M, int64_t N, int L) {
    if (M <= 0 || N <= 0 || L < 1) return 0;
    const SvdGrid gd = which == 0 ? xw_grid(M, N) : xty_grid(M, N);
    return gd.splits > 1 ? gd.splits * (which == 0 ? M : N) * (int64_t)L : 0;
}

int adx_svd_xw(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f, const float* W, int L,
               float* Y, float* work, void* stream) {
    if (M <= 0 || N <= 0) return 0;
    if (L < 1 || L > 64) return fail_msg(1, "adx_svd_xw: L must be in [1, 64]");
    if (ldg % 16 != 0 || ldg < ceil_div(N, 64) * 64) return fail_msg(1, "adx_svd_xw: bad ldg");
    hipStream_t st = (hipStream_t)stream;
    switch ((L + 15) / 16) {
        case 1: launch_xw<1>(st, G, ldg, M, N, f, W, L, Y, work); break;
        case 2: launch_xw<2>(st, G, ldg, M, N, f, W, L, Y, work); break;
        case 3: launch_xw<3>(st, G, ldg, M, N, f, W, L, Y, work); break;
        default: launch_xw<4>(st, G, ldg, M, N, f, W, L, Y, work); break;
    }
    ADX_LAUNCH_CHECK("adx_svd_xw");
    return 0;
}

int adx_svd_xty(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f, const float* Y, int L,
                float* A, float* work, void* stream) {
    if (M <= 0 || N <= 0) return 0;
    if (L < 1 || L > 64) return fail_msg(1, "adx_svd_xty: L must be in [1, 64]");
    if (ldg % 16 != 0 || ldg < ceil_div(N, 64) * 64) return fail_msg(1, "adx_svd_xty: bad ldg");
    hipStream_t st = (hipStream_t)stream;
    switch ((L + 15) / 16) {
        case 1: launch_xty<1>(st, G, ldg, M, N, f, Y, L, A, work); break;
        case 2: launch_xty<2>(st, G, ldg, M, N, f, Y, L, A, work); break;
        case 3: launch_xty<3>(st, G, ldg, M, N, f, Y, L, A, work); break;
        default: launch_xty<4>(st, G, ldg, M, N, f, Y, L, A, work); break;
    }
    ADX_LAUNCH_CHECK("adx_svd_xty");
    return 0;
}

}  // extern "C"
