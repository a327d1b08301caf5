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

// ---------------------------------------------------------------------------
// FP16 two-term forms (16x16x32 FP16 MFMA, FP32 accumulation).  They give the
// FP32 forms' results to FP32 rounding at a quarter of the matrix-core time:
//
//   svd_xw_h2_kernel:  Y = X W.  Per SNP the four values x(c) = c - 2f (FP32,
//     as centred(); x(3) = 0) are split into x_hi + x_lo (FP16; |x - x_hi -
//     x_lo| <= 2^-24 |x| + 2^-25) and looked up from the 2-bit codes with
//     byte permutes.  W s = W_hi + W_lo (FP16) with s a power of two that
//     puts max |W s| in [2^13, 2^14).  Y = (W_hi x_hi + W_hi x_lo + W_lo x_hi)
//     / s; the dropped W_lo x_lo is below 2^-22 of each product.
//   svd_xty_h2_kernel: A = X^T Y = Y^T Gm - (2f Y)^T Mask, Gm = g and Mask = 1
//     (both 0 if missing) exact in FP16; Y s and 2f Y s (s: max |Y s| in
//     [2^12, 2^13), so |2f Y s| < 2^14) split into two FP16 terms each, the
//     negated 2f Y terms against Mask, four MFMAs into one accumulator.
//
// The scaled, split operand is rebuilt by a prep kernel (transposed to the
// MFMA's k-contiguous layout) in the caller's workspace before each product;
// s comes from a fixed-order max reduction (svd_absmax_kernel), so the
// products are deterministic.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// 16x16x32 FP16 MFMA, FP32 accumulation (layouts as mma_bf16).
__device__ __forceinline__ floatx4 mma_h2(f16x8 a, f16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ uint32_t f16_bits(_Float16 h) { return (uint32_t)__builtin_bit_cast(uint16_t, h); }

constexpr int SVD_MAX_BLOCKS = 256;   // partial maxima of svd_absmax_kernel

// bmax[b] = max |X[e]| over the block's grid-stride share of the n entries.
__global__ __launch_bounds__(256) void svd_absmax_kernel(const float* __restrict__ X, int64_t n,
                                                         float* __restrict__ bmax) {
    __shared__ float red[256];
    float m = 0.0f;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < n; e += (int64_t)SVD_MAX_BLOCKS * 256)
        m = fmaxf(m, fabsf(X[e]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) bmax[blockIdx.x] = red[0];
}

// Power-of-two scale s with max * s in [2^(top-1), 2^top) (1 for a zero or
// non-finite max), from the partial maxima; every block reduces them itself.
__device__ float block_scale(const float* __restrict__ bmax, int top, float* red) {
    red[threadIdx.x] = threadIdx.x < SVD_MAX_BLOCKS ? bmax[threadIdx.x] : 0.0f;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    const float m = red[0];
    __syncthreads();
    if (!(m > 0.0f) || !isfinite(m)) return 1.0f;
    int ex;
    frexpf(m, &ex);   // m < 2^ex
    return ldexpf(1.0f, top - ex < 126 ? top - ex : 126);   // 1 / s stays a normal float
}

// Split operand of the products: from X (n x L FP32, row-major) write
// out[t][li][r] (FP16, LP = 16 LT rows of ld entries per term, zero past n
// and L) for 64 rows r per block, transposed through LDS:
//   NT = 2 (xw, W):  t0 = hi(X s), t1 = lo(X s)                  (top 14)
//   NT = 4 (xty, Y): t0 = hi(Y s), t1 = lo(Y s),
//                    t2 = -hi(2f Y s), t3 = -lo(2f Y s)           (top 13)
// Block 0 stores 1 / s in scal[0].
template <int NT>
__global__ __launch_bounds__(256) void svd_prep_h2_kernel(const float* __restrict__ X, const float* __restrict__ f,
                                                          int64_t n, int L, int LP, int64_t ld,
                                                          const float* __restrict__ bmax, float* __restrict__ scal,
                                                          _Float16* __restrict__ out) {
    __shared__ float red[256];
    __shared__ float tile[64][65];
    const float s = block_scale(bmax, NT == 2 ? 14 : 13, red);
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[0] = 1.0f / s;   // exact: s is a power of two
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * L; e += 256) {
        const int r = e / L, c = e - r * L;
        tile[c][r] = r0 + r < n ? X[(r0 + r) * L + c] : 0.0f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < LP * 64; e += 256) {
        const int li = e >> 6, r = e & 63;
        const int64_t row = r0 + r;
        const float x = li < L ? tile[li][r] : 0.0f;
        const int64_t o = (int64_t)li * ld + row;
        const float v = x * s;
        const _Float16 vh = (_Float16)v;
        out[o] = vh;
        out[(int64_t)LP * ld + o] = (_Float16)(v - (float)vh);
        if (NT == 4) {
            const float twof = row < n ? 2.0f * f[row] : 0.0f;
            const float w = (twof * x) * s;
            const _Float16 wh = (_Float16)w;
            out[2 * (int64_t)LP * ld + o] = -wh;
            out[3 * (int64_t)LP * ld + o] = -(_Float16)(w - (float)wh);
        }
    }
}

// x_hi / x_lo byte tables of one SNP: byte c of lo / hi = low / high byte of
// the FP16 term for code c.
struct XTab {
    uint32_t hlo, hhi, llo, lhi;
};

__device__ __forceinline__ XTab x_table(float twof) {
    XTab t{0u, 0u, 0u, 0u};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float x = (float)c - twof;
        const _Float16 xh = (_Float16)x;
        const uint32_t bh = f16_bits(xh), bl = f16_bits((_Float16)(x - (float)xh));
        t.hlo |= (bh & 0xFFu) << (8 * c);
        t.hhi |= (bh >> 8) << (8 * c);
        t.llo |= (bl & 0xFFu) << (8 * c);
        t.lhi |= (bl >> 8) << (8 * c);
    }
    return t;
}

// FP16 pairs of x_hi and x_lo for the four codes in the bytes of cwd
// (samples j..j+3): dwords (j, j+1) and (j+2, j+3) of each term.
__device__ __forceinline__ void x_quads(uint32_t cwd, const XTab& t, uint32_t& h01, uint32_t& h23, uint32_t& l01,
                                        uint32_t& l23) {
    const uint32_t cwp = cwd | 0x04040404u;
    const uint32_t s01 = __builtin_amdgcn_perm(cwp, cwd, 0x05010400u);   // [c0, c0 + 4, c1, c1 + 4]
    const uint32_t s23 = __builtin_amdgcn_perm(cwp, cwd, 0x07030602u);   // [c2, c2 + 4, c3, c3 + 4]
    h01 = __builtin_amdgcn_perm(t.hhi, t.hlo, s01);
    h23 = __builtin_amdgcn_perm(t.hhi, t.hlo, s23);
    l01 = __builtin_amdgcn_perm(t.lhi, t.llo, s01);
    l23 = __builtin_amdgcn_perm(t.lhi, t.llo, s23);
}

// Same tiling as svd_xw_kernel; the 16 genotype bytes of a lane give two
// k-steps of 32 samples (k = 8g + i <-> sample s0 + 16g + 8h + i).  Wx holds
// the split W^T, [2][16 LT][ldw] FP16, ldw >= the samples rounded up to 64.
template <int LT, int CW>
__global__ __launch_bounds__(256) void svd_xw_h2_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                        int64_t N, const float* __restrict__ f,
                                                        const _Float16* __restrict__ Wx, int64_t ldw,
                                                        const float* __restrict__ scal, int L, int64_t n_per_split,
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
    const int64_t WT = (int64_t)16 * LT * ldw;

    XTab tab[CW];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int64_t snp = snp_w + 64 * (cb >> 2) + 4 * cl + (cb & 3);
        tab[cb] = x_table(snp < M ? 2.0f * f[snp] : 0.0f);
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
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            f16x8 wh[LT], wl[LT];
#pragma unroll
            for (int lt = 0; lt < LT; ++lt) {
                const _Float16* p = Wx + (int64_t)(16 * lt + cl) * ldw + s0 + 16 * g + 8 * h;
                wh[lt] = *reinterpret_cast<const f16x8*>(p);
                wl[lt] = *reinterpret_cast<const f16x8*>(p + WT);
            }
#pragma unroll
            for (int cb = 0; cb < CW; ++cb) {
                const int t = cb >> 2, sh = 2 * (cb & 3);
                uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
                x_quads((gw[t][2 * h] >> sh) & 0x03030303u, tab[cb], h0, h1, l0, l1);
                x_quads((gw[t][2 * h + 1] >> sh) & 0x03030303u, tab[cb], h2, h3, l2, l3);
                const f16x8 xh = __builtin_bit_cast(f16x8, u32x4{h0, h1, h2, h3});
                const f16x8 xl = __builtin_bit_cast(f16x8, u32x4{l0, l1, l2, l3});
#pragma unroll
                for (int lt = 0; lt < LT; ++lt) {
                    acc[cb][lt] = mma_h2(wh[lt], xh, acc[cb][lt]);
                    acc[cb][lt] = mma_h2(wh[lt], xl, acc[cb][lt]);
                    acc[cb][lt] = mma_h2(wl[lt], xh, acc[cb][lt]);
                }
            }
        }
    }
    const float inv = scal[0];
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
                    const float v = acc[cb][lt][j] * inv;
                    if (part)
                        part[((int64_t)blockIdx.y * M + snp) * L + li] = v;
                    else
                        Y[snp * L + li] = v;
                }
            }
    }
}

// FP16 high bytes by code (low bytes 0): g = {0, 1, 2, -} and the mask
// {1, 1, 1, 0}; FP16 1.0 = 0x3C00, 2.0 = 0x4000.
constexpr uint32_t kGenoH = 0x00403C00u;
constexpr uint32_t kMaskH = 0x003C3C3Cu;

// The four 2-bit fields of byte b as bytes [c0, c1, c2, c3].
__device__ __forceinline__ uint32_t spread_codes(uint32_t b) {
    return (b | (b << 6) | (b << 12) | (b << 18)) & 0x03030303u;
}

// Gm and Mask FP16 pairs of the codes in the bytes of cw: dwords (c0, c1)
// and (c2, c3).
__device__ __forceinline__ void gm_quads(uint32_t cw, uint32_t& g01, uint32_t& g23, uint32_t& m01, uint32_t& m23) {
    // byte selectors [4, c_j, 4, c_j+1]: selector 4 picks the zero first operand
    const uint32_t s01 = __builtin_amdgcn_perm(0x04040404u, cw, 0x01040004u);
    const uint32_t s23 = __builtin_amdgcn_perm(0x04040404u, cw, 0x03040204u);
    g01 = __builtin_amdgcn_perm(0u, kGenoH, s01);
    g23 = __builtin_amdgcn_perm(0u, kGenoH, s23);
    m01 = __builtin_amdgcn_perm(0u, kMaskH, s01);
    m23 = __builtin_amdgcn_perm(0u, kMaskH, s23);
}

// Workgroup: NW waves x SUB sub-tiles of 16 samples, one range of byte rows.
// Chunks of 32 byte rows (128 SNPs, four k-steps of 32) are staged in LDS:
//   sY:  the chunk's Yx rows, [4 terms][16 LT][136] FP16 (272-byte rows: the
//        16-byte reads of 8 lanes hit distinct banks);
//   sG:  per sample 40 bytes, of which byte 8 g + 2 kk (+1) is byte row
//        8 kk + 2 g (+1) of the chunk: lane (g, c) reads the 8 bytes of its
//        sample for all four k-steps in one 8-byte read (16 lanes of a read
//        hit distinct banks with the 40-byte stride).
// k-step kk, k = 8 g + i <-> SNP 32 kk + 8 g + i of the chunk.  Yx holds
// [4][16 LT][ldy] FP16, ldy >= the SNPs rounded up to 128.
template <int LT, int NW, int SUB>
__global__ __launch_bounds__(64 * NW) void svd_xty_h2_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                             int64_t N, const _Float16* __restrict__ Yx, int64_t ldy,
                                                             const float* __restrict__ scal, int L,
                                                             int64_t rows_per_split, float* __restrict__ A,
                                                             float* __restrict__ part) {
    constexpr int WS = 16 * SUB * NW;   // samples per workgroup
    constexpr int CR = 32;              // byte rows per chunk
    constexpr int YR = 4 * CR + 8;      // FP16 per sY row
    constexpr int GS = 40;              // bytes per sample in sG
    constexpr int LP = 16 * LT;
    __shared__ __attribute__((aligned(16))) _Float16 sY[4 * LP * YR];
    __shared__ __attribute__((aligned(16))) uint8_t sG[WS * GS];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int64_t s0 = (int64_t)blockIdx.x * WS;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_split;
    const int64_t r1 = r0 + rows_per_split < Mb ? r0 + rows_per_split : Mb;
    const int64_t YT = (int64_t)LP * ldy;

    floatx4 acc[SUB][LT];
#pragma unroll
    for (int q = 0; q < SUB; ++q)
#pragma unroll
        for (int lt = 0; lt < LT; ++lt) acc[q][lt] = floatx4{0, 0, 0, 0};

    for (int64_t rb = r0; rb < r1; rb += CR) {
        // byte-row pairs (8 kk + 2 gp, +1) of 16 samples per task; rows past
        // the range and SNP fields past M read as missing
        for (int e = tid; e < WS; e += 64 * NW) {
            const int grp = e >> 4, kk = (e >> 2) & 3, gp = e & 3;
            const int64_t c = s0 + 16 * grp;
            uint32_t v[2][4];
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int64_t r = rb + 8 * kk + 2 * gp + b;
                if (r < r1 && c < ldg) {
                    const uint4 x = *reinterpret_cast<const uint4*>(G + r * ldg + c);
                    v[b][0] = x.x; v[b][1] = x.y; v[b][2] = x.z; v[b][3] = x.w;
                    const int64_t nf = M - 4 * r;
                    if (nf < 4) {
                        const uint32_t pm = ((0xFFu << (2 * nf)) & 0xFFu) * 0x01010101u;
#pragma unroll
                        for (int d = 0; d < 4; ++d) v[b][d] |= pm;
                    }
                } else {
#pragma unroll
                    for (int d = 0; d < 4; ++d) v[b][d] = 0xFFFFFFFFu;
                }
            }
            uint8_t* base = sG + (16 * grp) * GS + 8 * gp + 2 * kk;
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                const uint32_t p01 = __builtin_amdgcn_perm(v[1][d], v[0][d], 0x05010400u);   // [a0, b0, a1, b1]
                const uint32_t p23 = __builtin_amdgcn_perm(v[1][d], v[0][d], 0x07030602u);   // [a2, b2, a3, b3]
                *reinterpret_cast<uint16_t*>(base + (4 * d + 0) * GS) = (uint16_t)p01;
                *reinterpret_cast<uint16_t*>(base + (4 * d + 1) * GS) = (uint16_t)(p01 >> 16);
                *reinterpret_cast<uint16_t*>(base + (4 * d + 2) * GS) = (uint16_t)p23;
                *reinterpret_cast<uint16_t*>(base + (4 * d + 3) * GS) = (uint16_t)(p23 >> 16);
            }
        }
        // the chunk's 4 x LP Yx rows of 128 SNPs, 16 bytes per thread and step
        for (int e = tid; e < 4 * LP * 16; e += 64 * NW) {
            const int u = e & 15, row = e >> 4;
            const int term = row / LP, li = row - term * LP;
            *reinterpret_cast<uint4*>(sY + row * YR + 8 * u) =
                *reinterpret_cast<const uint4*>(Yx + term * YT + (int64_t)li * ldy + 4 * rb + 8 * u);
        }
        __syncthreads();
        uint2 gb[SUB];
#pragma unroll
        for (int q = 0; q < SUB; ++q)
            gb[q] = *reinterpret_cast<const uint2*>(sG + (wave * 16 * SUB + 16 * q + cl) * GS + 8 * g);
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            f16x8 a[4][LT];
#pragma unroll
            for (int t = 0; t < 4; ++t)
#pragma unroll
                for (int lt = 0; lt < LT; ++lt)
                    a[t][lt] = *reinterpret_cast<const f16x8*>(sY + (t * LP + 16 * lt + cl) * YR + 32 * kk + 8 * g);
#pragma unroll
            for (int q = 0; q < SUB; ++q) {
                const uint32_t w = (kk < 2 ? gb[q].x : gb[q].y) >> (16 * (kk & 1));
                uint32_t g0, g1, g2, g3, m0, m1, m2, m3;
                gm_quads(spread_codes(w & 0xFFu), g0, g1, m0, m1);
                gm_quads(spread_codes((w >> 8) & 0xFFu), g2, g3, m2, m3);
                const f16x8 bg = __builtin_bit_cast(f16x8, u32x4{g0, g1, g2, g3});
                const f16x8 bm = __builtin_bit_cast(f16x8, u32x4{m0, m1, m2, m3});
#pragma unroll
                for (int lt = 0; lt < LT; ++lt) {
                    acc[q][lt] = mma_h2(a[0][lt], bg, acc[q][lt]);
                    acc[q][lt] = mma_h2(a[1][lt], bg, acc[q][lt]);
                    acc[q][lt] = mma_h2(a[2][lt], bm, acc[q][lt]);
                    acc[q][lt] = mma_h2(a[3][lt], bm, acc[q][lt]);
                }
            }
        }
        __syncthreads();
    }
    const float inv = scal[0];
#pragma unroll
    for (int q = 0; q < SUB; ++q) {
        const int64_t smp = s0 + wave * 16 * SUB + 16 * q + cl;
        if (smp >= N) continue;
#pragma unroll
        for (int lt = 0; lt < LT; ++lt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int li = 16 * lt + 4 * g + j;
                if (li < L) {
                    const float v = acc[q][lt][j] * inv;
                    if (part)
                        part[((int64_t)blockIdx.y * N + smp) * L + li] = v;
                    else
                        A[smp * L + li] = v;
                }
            }
    }
}

// Launch shapes: xw splits the samples, xty the byte rows, until the grid
// holds about 2048 workgroups.
struct SvdGrid {
    int64_t nbx, splits, per_split;
};

static SvdGrid xw_grid(int64_t M, int64_t N, int CW = 4) {
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

// FP16 two-term forms: xw with 8 SNP column tiles per wave for L <= 32 (4
// above); xty with 16 waves of 4 sub-tiles (1024 samples) for L <= 32 (8
// waves above), its byte rows split until the grid holds about 1024
// workgroups (one or two resident per CU).
constexpr int H2_SUB = 4;
static int h2_xw_cw(int LT) { return LT <= 2 ? 8 : 4; }
static int h2_xty_nw(int LT) { return LT <= 2 ? 16 : 8; }

static SvdGrid xty_h2_grid(int64_t M, int64_t N, int NW) {
    const int64_t Mb = ceil_div(M, 4);
    const int64_t nbx = ceil_div(N, 16 * H2_SUB * NW);
    int64_t s = ceil_div(1024, nbx);
    const int64_t max_s = ceil_div(Mb, 32);
    s = s > max_s ? max_s : (s < 1 ? 1 : s);
    if (s > 65535) s = 65535;
    const int64_t rps = ceil_div(ceil_div(Mb, s), 32) * 32;
    return {nbx, ceil_div(Mb, rps), rps};
}

// Workspace of the FP16 forms: [0, 1024) partial maxima, [1024, 1280) the
// inverse scale, then the split operand (NT terms x 16 LT rows x ld FP16),
// then the split-reduction slices.
struct H2Layout {
    SvdGrid grid;
    int64_t ld, op_off, part_off, part_floats, total;
};

static H2Layout h2_layout(int which, int64_t M, int64_t N, int L) {
    const int LT = (L + 15) / 16;
    H2Layout w;
    int nt;
    int64_t out_rows;
    if (which == 0) {
        w.grid = xw_grid(M, N, h2_xw_cw(LT));
        w.ld = ceil_div(N, 64) * 64;
        nt = 2;
        out_rows = M;
    } else {
        w.grid = xty_h2_grid(M, N, h2_xty_nw(LT));
        w.ld = ceil_div(M, 128) * 128;
        nt = 4;
        out_rows = N;
    }
    w.op_off = 1280;
    w.part_off = w.op_off + ceil_div((int64_t)nt * 16 * LT * w.ld * 2, 256) * 256;
    w.part_floats = w.grid.splits > 1 ? w.grid.splits * out_rows * (int64_t)L : 0;
    w.total = w.part_off + w.part_floats * 4;
    return w;
}

template <int LT>
static void launch_xw_h2(hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f,
                         const float* W, int L, float* Y, uint8_t* ws) {
    constexpr int CW = LT <= 2 ? 8 : 4;
    const H2Layout lay = h2_layout(0, M, N, L);
    float* bmax = reinterpret_cast<float*>(ws);
    float* scal = reinterpret_cast<float*>(ws + 1024);
    _Float16* Wx = reinterpret_cast<_Float16*>(ws + lay.op_off);
    float* part = lay.part_floats ? reinterpret_cast<float*>(ws + lay.part_off) : nullptr;
    svd_absmax_kernel<<<SVD_MAX_BLOCKS, 256, 0, st>>>(W, N * (int64_t)L, bmax);
    svd_prep_h2_kernel<2><<<(unsigned)(lay.ld / 64), 256, 0, st>>>(W, nullptr, N, L, 16 * LT, lay.ld, bmax, scal, Wx);
    const SvdGrid& gd = lay.grid;
    svd_xw_h2_kernel<LT, CW><<<dim3((unsigned)gd.nbx, (unsigned)gd.splits), 256, 0, st>>>(
        G, ldg, M, N, f, Wx, lay.ld, scal, L, gd.per_split, Y, part);
    if (part)
        svd_sum_kernel<<<(unsigned)ceil_div(M * L, 256), 256, 0, st>>>(part, gd.splits, M * (int64_t)L, Y);
}

template <int LT>
static void launch_xty_h2(hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f,
                          const float* Y, int L, float* A, uint8_t* ws) {
    constexpr int NW = LT <= 2 ? 16 : 8;
    const H2Layout lay = h2_layout(1, M, N, L);
    float* bmax = reinterpret_cast<float*>(ws);
    float* scal = reinterpret_cast<float*>(ws + 1024);
    _Float16* Yx = reinterpret_cast<_Float16*>(ws + lay.op_off);
    float* part = lay.part_floats ? reinterpret_cast<float*>(ws + lay.part_off) : nullptr;
    svd_absmax_kernel<<<SVD_MAX_BLOCKS, 256, 0, st>>>(Y, M * (int64_t)L, bmax);
    svd_prep_h2_kernel<4><<<(unsigned)(lay.ld / 64), 256, 0, st>>>(Y, f, M, L, 16 * LT, lay.ld, bmax, scal, Yx);
    const SvdGrid& gd = lay.grid;
    svd_xty_h2_kernel<LT, NW, H2_SUB><<<dim3((unsigned)gd.nbx, (unsigned)gd.splits), 64 * NW, 0, st>>>(
        G, ldg, M, N, Yx, lay.ld, scal, L, gd.per_split, A, part);
    if (part)
        svd_sum_kernel<<<(unsigned)ceil_div(N * L, 256), 256, 0, st>>>(part, gd.splits, N * (int64_t)L, A);
}

}  // namespace adx

using namespace adx;

extern "C" {

int64_t adx_svd_h2_work_bytes(int which, int64_t M, int64_t N, int L) {
    if (M <= 0 || N <= 0 || L < 1 || L > 64 || (which != 0 && which != 1)) return 0;
    return h2_layout(which, M, N, L).total;
}

int adx_svd_xw_h2(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f, const float* W, int L,
                  float* Y, void* work, int64_t work_bytes, void* stream) {
    if (M <= 0 || N <= 0) return 0;
    if (L < 1 || L > 64) return fail_msg(1, "adx_svd_xw_h2: L must be in [1, 64]");
    if (ldg % 16 != 0 || ldg < ceil_div(N, 64) * 64) return fail_msg(1, "adx_svd_xw_h2: bad ldg");
    if (!work || work_bytes < h2_layout(0, M, N, L).total)
        return fail_msg(1, "adx_svd_xw_h2: workspace smaller than adx_svd_h2_work_bytes");
    if (reinterpret_cast<uintptr_t>(work) % 256 != 0) return fail_msg(1, "adx_svd_xw_h2: workspace not 256-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    uint8_t* ws = static_cast<uint8_t*>(work);
    switch ((L + 15) / 16) {
        case 1: launch_xw_h2<1>(st, G, ldg, M, N, f, W, L, Y, ws); break;
        case 2: launch_xw_h2<2>(st, G, ldg, M, N, f, W, L, Y, ws); break;
        case 3: launch_xw_h2<3>(st, G, ldg, M, N, f, W, L, Y, ws); break;
        default: launch_xw_h2<4>(st, G, ldg, M, N, f, W, L, Y, ws); break;
    }
    ADX_LAUNCH_CHECK("adx_svd_xw_h2");
    return 0;
}

int adx_svd_xty_h2(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const float* f, const float* Y, int L,
                   float* A, void* work, int64_t work_bytes, void* stream) {
    if (M <= 0 || N <= 0) return 0;
    if (L < 1 || L > 64) return fail_msg(1, "adx_svd_xty_h2: L must be in [1, 64]");
    if (ldg % 16 != 0 || ldg < ceil_div(N, 64) * 64) return fail_msg(1, "adx_svd_xty_h2: bad ldg");
    if (!work || work_bytes < h2_layout(1, M, N, L).total)
        return fail_msg(1, "adx_svd_xty_h2: workspace smaller than adx_svd_h2_work_bytes");
    if (reinterpret_cast<uintptr_t>(work) % 256 != 0) return fail_msg(1, "adx_svd_xty_h2: workspace not 256-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    uint8_t* ws = static_cast<uint8_t*>(work);
    switch ((L + 15) / 16) {
        case 1: launch_xty_h2<1>(st, G, ldg, M, N, f, Y, L, A, ws); break;
        case 2: launch_xty_h2<2>(st, G, ldg, M, N, f, Y, L, A, ws); break;
        case 3: launch_xty_h2<3>(st, G, ldg, M, N, f, Y, L, A, ws); break;
        default: launch_xty_h2<4>(st, G, ldg, M, N, f, Y, L, A, ws); break;
    }
    ADX_LAUNCH_CHECK("adx_svd_xty_h2");
    return 0;
}

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
