// BF16 form of the fused decode + EM-statistics pass (FP32 accumulation).
//
// Same tiling as em_pass_kernel (adx_em.hip): a workgroup owns 64*CW SNPs and
// a sample strip, lane (g, c) streams 16 genotype bytes of one byte row per
// 64-sample sub-tile, A and B live in MFMA accumulators for the strip and T
// is reduced per sub-tile.  The contractions run on 16x16x32 BF16 MFMA:
//
//   H^T  = Qx . Px^T with Qx = [Q_hi, Q_hi, Q_lo], Px = [P_hi, P_lo, P_hi]
//          (three-term split: |H - H_exact| <~ 3 * 2^-16 |H|); one MFMA for
//          K <= 10, two for K <= 16
//   A^T += (Qt_hi + Qt_lo) . a^T,  B^T += (Qt_hi + Qt_lo) . b^T
//          k = 32 samples (two 16-sample row blocks), a, b rounded to BF16
//   T   += [a b] . [P_hi ; 1 - P_hi]   k = 16 SNPs x {a, b}: a and b of one
//          genotype are rounded into one word (bf16 pair), the 16 x 16 word
//          tile is transposed through LDS (one 32-bit write per genotype, one
//          128-bit read per lane), and the colsum(b) term of em_batch_math
//          is the (1 - P) half.  The same words, permuted, are the B
//          operands of the A and B contractions.
//
// a = g/H and b = (2-g)/(1-H) are formed in FP32 from the FP32 accumulator,
// with one reciprocal of H(1-H).  The operand arrays Px, Qx, Qt are rebuilt
// from the FP64 P, Q by adx_em_prep_bf16 before every pass.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

// Per-byte genotype weights of four genotypes at once: w holds one 2-bit
// code per byte (bits above cleared); returns ga = g and gb = 2 - g per byte,
// both 0 where the code is 3 (missing).
__device__ __forceinline__ void code_weights(uint32_t w, uint32_t& ga, uint32_t& gb) {
    const uint32_t m = (w >> 1) & w & 0x01010101u;   // 1 in missing bytes
    ga = w ^ (m | (m << 1));                          // 3 -> 0
    gb = (0x02020202u ^ (m << 1)) - ga;               // 2 - g, missing -> 0 (no borrows)
}

// a = ga / H and b = gb / (1 - H) with H clamped to [1e-5, 1 - 1e-5]
// (em_adam_gpu.py:62-72), one reciprocal of H (1 - H).
__device__ __forceinline__ void em_terms_f32(float h, float ga, float gb, float& a, float& b) {
    const float hc = __builtin_amdgcn_fmed3f(h, 1e-5f, 1.0f - 1e-5f);
    const float u = 1.0f - hc;
    const float r = __builtin_amdgcn_rcpf(hc * u);
    a = ga * (u * r);
    b = gb * (hc * r);
}

template <int J>
__device__ __forceinline__ float byte_to_float(uint32_t w) {
    return (float)((w >> (8 * J)) & 0xFFu);
}

__device__ __forceinline__ bf16x8 load_bf16x8(const __bf16* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}

// Two FP32 values rounded to BF16 and packed into one word (lo = x, hi = y).
__device__ __forceinline__ uint32_t pack_bf16x2(float x, float y) {
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bf16x2 v = bf16x2{(__bf16)x, (__bf16)y};
    return __builtin_bit_cast(uint32_t, v);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int KH, int CW>
__global__ __launch_bounds__(256) void em_pass_bf16_kernel(
    const uint8_t* __restrict__ G, int64_t ldg, int64_t M, int64_t N, int K, const __bf16* __restrict__ Px,
    const __bf16* __restrict__ Qx, const __bf16* __restrict__ Qt, int64_t ldqt, int64_t n_per_split,
    float* __restrict__ Aacc, float* __restrict__ Bacc, float* __restrict__ Tacc, float* __restrict__ Tpart,
    float* __restrict__ ABpart) {
    static_assert(CW % 4 == 0, "CW must be a multiple of 4");
    constexpr int LX = 32 * KH;  // row length of Px / Qx
    constexpr int TS = 20;       // row stride (words) of the per-wave transpose tile
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int64_t snp_w = ((int64_t)blockIdx.x * 4 + wave) * (16 * CW);
    const int64_t row_w = snp_w >> 2;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t n0 = (int64_t)blockIdx.y * n_per_split;
    const int64_t n1 = n0 + n_per_split < N ? n0 + n_per_split : N;

    // Per wave: word [sample row r][SNP column c] = bf16 a | bf16 b << 16.
    __shared__ __attribute__((aligned(16))) uint32_t s_tile[4][16 * TS];
    // Per wave: T of the 64 samples of the current row block over the wave's
    // SNPs; each (sample, k) is written once per wave and row block, and the
    // four slices are summed in a fixed order at the flush (deterministic).
    // Rows padded to 17 floats so the 64 lanes of a write hit distinct banks.
    __shared__ float s_tsum[4][64 * 17];

    const bf16x8 zero8 = bf16x8{};
    bf16x8 pH[CW][KH];
    bf16x8 pT[CW];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int t = cb >> 2, f = cb & 3;
        const int64_t snp = snp_w + 64 * t + 4 * cl + f;
#pragma unroll
        for (int kh = 0; kh < KH; ++kh) pH[cb][kh] = snp < M ? load_bf16x8(Px + snp * LX + 32 * kh + 8 * g) : zero8;
        // T B-operand: k = 8g + i <-> (SNP column 4g + i/2, P for even i, 1 - P
        // for odd i), matching the (a, b) pairs of the tile words; column kp = cl.
        bf16x8 v;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t s2 = snp_w + 64 * t + 4 * (4 * g + (i >> 1)) + f;
            float x = 0.0f;
            if (s2 < M && cl < K) {
                const float p = (float)Px[s2 * LX + cl];
                x = (i & 1) ? 1.0f - p : p;
            }
            v[i] = (__bf16)x;
        }
        pT[cb] = v;
    }

    floatx4 accA[CW], accB[CW];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        accA[cb] = floatx4{0, 0, 0, 0};
        accB[cb] = floatx4{0, 0, 0, 0};
    }
    uint32_t* tile = s_tile[wave];
    __syncthreads();

    for (int64_t s0 = n0; s0 < n1; s0 += 64) {
        uint32_t gw[CW / 4][4];
#pragma unroll
        for (int t = 0; t < CW / 4; ++t) {
            const int64_t row = row_w + 16 * t + cl;
            if (row < Mb) {
                const uint4 v = *reinterpret_cast<const uint4*>(G + row * ldg + s0 + 16 * g);
                gw[t][0] = v.x;
                gw[t][1] = v.y;
                gw[t][2] = v.z;
                gw[t][3] = v.w;
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

        // Row-block pairs: accumulator element j of lane group g in row block
        // rw is sample s0 + 16g + 4rw + j; a pair (rw = 2tp, 2tp+1) gives lane
        // (g, cl) the 8 consecutive samples s0 + 16g + 8tp + i, i = 0..7.
#pragma unroll
        for (int tp = 0; tp < 2; ++tp) {
            const int64_t sq = s0 + 16 * g + 8 * tp;  // first of the 8 samples (Qt rows are padded)
            const bf16x8 qThi = load_bf16x8(Qt + (int64_t)cl * ldqt + sq);
            const bf16x8 qTlo = load_bf16x8(Qt + (16 + (int64_t)cl) * ldqt + sq);
            bf16x8 qH[2][KH];
#pragma unroll
            for (int rr = 0; rr < 2; ++rr) {
                const int64_t sH = s0 + 16 * (cl >> 2) + 4 * (2 * tp + rr) + (cl & 3);
#pragma unroll
                for (int kh = 0; kh < KH; ++kh)
                    qH[rr][kh] = sH < N ? load_bf16x8(Qx + sH * LX + 32 * kh + 8 * g) : zero8;
            }
            floatx4 accT[2] = {floatx4{0, 0, 0, 0}, floatx4{0, 0, 0, 0}};

#pragma unroll
            for (int cb = 0; cb < CW; ++cb) {
                const int t = cb >> 2, f = cb & 3;
                uint32_t wab[2][4];
#pragma unroll
                for (int rr = 0; rr < 2; ++rr) {
                    floatx4 h = floatx4{0, 0, 0, 0};
#pragma unroll
                    for (int kh = 0; kh < KH; ++kh) h = mma_bf16(qH[rr][kh], pH[cb][kh], h);
                    uint32_t ga, gb;
                    code_weights((gw[t][2 * tp + rr] >> (2 * f)) & 0x03030303u, ga, gb);
                    const float fa[4] = {byte_to_float<0>(ga), byte_to_float<1>(ga), byte_to_float<2>(ga),
                                         byte_to_float<3>(ga)};
                    const float fb[4] = {byte_to_float<0>(gb), byte_to_float<1>(gb), byte_to_float<2>(gb),
                                         byte_to_float<3>(gb)};
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        float a, b;
                        em_terms_f32(h[j], fa[j], fb[j], a, b);
                        wab[rr][j] = pack_bf16x2(a, b);
                        tile[(4 * g + j) * TS + cl] = wab[rr][j];
                    }
                    // The tile is written and read by lanes of this wave only;
                    // LDS operations of a wave complete in order, and the
                    // wave barrier keeps the compiler from moving them.
                    __builtin_amdgcn_wave_barrier();
                    // T (16 samples x 16 kp) += [a b] (16 x 32) . [P ; 1-P] (32 x 16)
                    const bf16x8 tv = __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(tile + cl * TS + 4 * g));
                    __builtin_amdgcn_wave_barrier();
                    accT[rr] = mma_bf16(tv, pT[cb], accT[rr]);
                }
                // B operands of A^T, B^T: k = sample 8g + i (rr = i / 4, j = i % 4).
                const u32x4 wa = u32x4{__builtin_amdgcn_perm(wab[0][1], wab[0][0], 0x05040100u),
                                       __builtin_amdgcn_perm(wab[0][3], wab[0][2], 0x05040100u),
                                       __builtin_amdgcn_perm(wab[1][1], wab[1][0], 0x05040100u),
                                       __builtin_amdgcn_perm(wab[1][3], wab[1][2], 0x05040100u)};
                const u32x4 wb = u32x4{__builtin_amdgcn_perm(wab[0][1], wab[0][0], 0x07060302u),
                                       __builtin_amdgcn_perm(wab[0][3], wab[0][2], 0x07060302u),
                                       __builtin_amdgcn_perm(wab[1][1], wab[1][0], 0x07060302u),
                                       __builtin_amdgcn_perm(wab[1][3], wab[1][2], 0x07060302u)};
                const bf16x8 aB = __builtin_bit_cast(bf16x8, wa);
                const bf16x8 bB = __builtin_bit_cast(bf16x8, wb);
                accA[cb] = mma_bf16(qThi, aB, accA[cb]);
                accA[cb] = mma_bf16(qTlo, aB, accA[cb]);
                accB[cb] = mma_bf16(qThi, bB, accB[cb]);
                accB[cb] = mma_bf16(qTlo, bB, accB[cb]);
            }
#pragma unroll
            for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (cl < K) s_tsum[wave][(16 * g + 4 * (2 * tp + rr) + j) * 17 + cl] = accT[rr][j];
        }

        __syncthreads();
        // T of these 64 samples over this workgroup's SNPs: one plain store
        // into the workgroup's slice of the partial buffer (every (slice,
        // sample) is written exactly once), summed by tpart_reduce_kernel.
        // Device-scope float atomics would be performed at the memory side
        // (one 4-byte read-modify-write each, ~N K M / (64 CW) per pass).
        for (int e = threadIdx.x; e < 64 * 16; e += 256) {
            const int kp = e & 15;
            const int64_t smp = s0 + (e >> 4);
            const int ex = (e >> 4) * 17 + kp;
            if (kp < K && smp < N) {
                const float v = (s_tsum[0][ex] + s_tsum[1][ex]) + (s_tsum[2][ex] + s_tsum[3][ex]);
                if (Tpart)
                    Tpart[((int64_t)blockIdx.x * N + smp) * K + kp] = v;
                else
                    atomicAdd(&Tacc[smp * K + kp], v);
            }
        }
        __syncthreads();
    }

#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int t = cb >> 2, f = cb & 3;
        const int64_t snp = snp_w + 64 * t + 4 * cl + f;
        if (snp < M) {
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int kp = 4 * g + jj;
                if (kp < K) {
                    // this sample split's slice of the A / B partial sums (each
                    // element written by one workgroup), summed in a fixed order
                    if (ABpart) {
                        const int64_t mk = M * (int64_t)K;
                        ABpart[(int64_t)blockIdx.y * mk + snp * K + kp] = accA[cb][jj];
                        ABpart[((int64_t)gridDim.y + blockIdx.y) * mk + snp * K + kp] = accB[cb][jj];
                    } else {
                        atomicAdd(&Aacc[snp * K + kp], accA[cb][jj]);
                        atomicAdd(&Bacc[snp * K + kp], accB[cb][jj]);
                    }
                }
            }
        }
    }
}

// Operand arrays of the BF16 pass from the FP64 master P (M x K) and Q (N x K):
//   Px[m]  = [P_hi(0..K), P_lo, P_hi, 0 ...]    (32*KH bf16)
//   Qx[i]  = [Q_hi, Q_hi, Q_lo, 0 ...]          (32*KH bf16)
//   Qt     = [hi rows 0..15 | lo rows 0..15] x ldqt, Qt[kp][i] = Q[i][kp]
// with x_hi = bf16(x), x_lo = bf16(x - x_hi); zero outside the valid range.
template <int KH>
__global__ void prep_px_kernel(const double* __restrict__ P, int64_t M, int K, __bf16* __restrict__ Px) {
    constexpr int LX = 32 * KH;
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    __bf16 row[LX];
#pragma unroll
    for (int e = 0; e < LX; ++e) row[e] = (__bf16)0.0f;
    for (int k = 0; k < K; ++k) {
        const float x = (float)P[m * K + k];
        const __bf16 hi = (__bf16)x;
        const __bf16 lo = (__bf16)(x - (float)hi);
        row[k] = hi;
        row[K + k] = lo;
        row[2 * K + k] = hi;
    }
#pragma unroll
    for (int e = 0; e < LX; ++e) Px[m * LX + e] = row[e];
}

template <int KH>
__global__ void prep_qx_kernel(const double* __restrict__ Q, int64_t N, int K, __bf16* __restrict__ Qx,
                               __bf16* __restrict__ Qt, int64_t ldqt) {
    constexpr int LX = 32 * KH;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ldqt) return;
    __bf16 row[LX];
#pragma unroll
    for (int e = 0; e < LX; ++e) row[e] = (__bf16)0.0f;
    for (int k = 0; k < 16; ++k) {
        __bf16 hi = (__bf16)0.0f, lo = (__bf16)0.0f;
        if (i < N && k < K) {
            const float x = (float)Q[i * K + k];
            hi = (__bf16)x;
            lo = (__bf16)(x - (float)hi);
            row[k] = hi;
            row[K + k] = hi;
            row[2 * K + k] = lo;
        }
        Qt[(int64_t)k * ldqt + i] = hi;
        Qt[(int64_t)(16 + k) * ldqt + i] = lo;
    }
    if (i < N)
#pragma unroll
        for (int e = 0; e < LX; ++e) Qx[i * LX + e] = row[e];
}

struct Bf16Grid {
    int cw;
    int64_t nbx, splits, n_per_split;
};

static Bf16Grid bf16_grid(int64_t M, int64_t N) {
    const EmTiles tl = em_tiles(ADX_BF16);
    Bf16Grid gd;
    gd.cw = tl.cw;
    gd.nbx = ceil_div(M, 64 * tl.cw);
    int64_t s = ceil_div(tl.target_wgs, gd.nbx);
    const int64_t max_s = ceil_div(N, 64);
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    if (s > 65535) s = 65535;
    gd.n_per_split = ceil_div(ceil_div(N, s), 64) * 64;
    gd.splits = ceil_div(N, gd.n_per_split);
    return gd;
}

// T[e] += sum over the nparts slices of Tpart[p * nk + e] (also used for the
// A and B slices, nk = M K).
__global__ void tpart_reduce_kernel(const float* __restrict__ Tpart, int64_t nparts, int64_t nk,
                                    float* __restrict__ T) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nk) return;
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
    int64_t p = 0;
    for (; p + 4 <= nparts; p += 4) {
        s0 += Tpart[(p + 0) * nk + e];
        s1 += Tpart[(p + 1) * nk + e];
        s2 += Tpart[(p + 2) * nk + e];
        s3 += Tpart[(p + 3) * nk + e];
    }
    for (; p < nparts; ++p) s0 += Tpart[p * nk + e];
    T[e] += (s0 + s1) + (s2 + s3);
}

template <int KH>
static void launch_bf16(const Bf16Grid& gd, hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N,
                        int K, const __bf16* Px, const __bf16* Qx, const __bf16* Qt, int64_t ldqt, float* A,
                        float* B, float* T, float* Tpart) {
    dim3 grid((unsigned)gd.nbx, (unsigned)gd.splits);
    float* ABpart = Tpart + gd.nbx * N * (int64_t)K;
    if (gd.cw == 8)
        em_pass_bf16_kernel<KH, 8>
            <<<grid, 256, 0, st>>>(G, ldg, M, N, K, Px, Qx, Qt, ldqt, gd.n_per_split, A, B, T, Tpart, ABpart);
    else
        em_pass_bf16_kernel<KH, 4>
            <<<grid, 256, 0, st>>>(G, ldg, M, N, K, Px, Qx, Qt, ldqt, gd.n_per_split, A, B, T, Tpart, ABpart);
    const int64_t nk = N * K, mk = M * K;
    tpart_reduce_kernel<<<(unsigned)ceil_div(nk, 256), 256, 0, st>>>(Tpart, gd.nbx, nk, T);
    tpart_reduce_kernel<<<(unsigned)ceil_div(mk, 256), 256, 0, st>>>(ABpart, gd.splits, mk, A);
    tpart_reduce_kernel<<<(unsigned)ceil_div(mk, 256), 256, 0, st>>>(ABpart + gd.splits * mk, gd.splits, mk, B);
}

// Byte offset of the partial-T slices in the operand buffer, and its size.
static int64_t bf16_base_bytes(int64_t M, int64_t N, int K) {
    const int kh = K <= 10 ? 1 : 2;
    const int64_t ld = ceil_div(N, 64) * 64;
    return ceil_div((int64_t)sizeof(__bf16) * (M * 32 * kh + N * 32 * kh + 32 * ld), 256) * 256;
}

}  // namespace adx

using namespace adx;

extern "C" {

int64_t adx_em_bf16_operand_bytes(int64_t M, int64_t N, int K, int64_t* ldqt) {
    const int64_t ld = ceil_div(N, 64) * 64;
    if (ldqt) *ldqt = ld;
    if (M <= 0 || N <= 0) return bf16_base_bytes(M > 0 ? M : 0, N > 0 ? N : 0, K);
    const Bf16Grid gd = bf16_grid(M, N);
    return bf16_base_bytes(M, N, K) + (int64_t)sizeof(float) * (gd.nbx * N * K + 2 * gd.splits * M * K);
}

int adx_em_prep_bf16(const double* P, const double* Q, int64_t M, int64_t N, int K, void* operands,
                     int64_t operand_bytes, void* stream) {
    if (K < 1 || K > 16) return fail_msg(1, "adx_em_prep_bf16: K must be in [1, 16]");
    int64_t ldqt = 0;
    if (operand_bytes < adx_em_bf16_operand_bytes(M, N, K, &ldqt))
        return fail_msg(1, "adx_em_prep_bf16: operand buffer too small");
    const int kh = K <= 10 ? 1 : 2;
    __bf16* Px = static_cast<__bf16*>(operands);
    __bf16* Qx = Px + M * 32 * kh;
    __bf16* Qt = Qx + N * 32 * kh;
    hipStream_t st = (hipStream_t)stream;
    if (kh == 1) {
        if (M > 0) prep_px_kernel<1><<<(unsigned)ceil_div(M, 256), 256, 0, st>>>(P, M, K, Px);
        prep_qx_kernel<1><<<(unsigned)ceil_div(ldqt, 256), 256, 0, st>>>(Q, N, K, Qx, Qt, ldqt);
    } else {
        if (M > 0) prep_px_kernel<2><<<(unsigned)ceil_div(M, 256), 256, 0, st>>>(P, M, K, Px);
        prep_qx_kernel<2><<<(unsigned)ceil_div(ldqt, 256), 256, 0, st>>>(Q, N, K, Qx, Qt, ldqt);
    }
    ADX_LAUNCH_CHECK("adx_em_prep_bf16");
    return 0;
}

int adx_em_pass_bf16(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K, void* operands,
                     int64_t operand_bytes, float* A, float* B, float* T, void* stream) {
    if (M <= 0 || N <= 0) return fail_msg(1, "adx_em_pass_bf16: empty genotype matrix");
    if (K < 1 || K > 16) return fail_msg(1, "adx_em_pass_bf16: K must be in [1, 16]");
    if (ldg % 16 != 0 || ldg < ceil_div(N, 64) * 64)
        return fail_msg(1, "adx_em_pass_bf16: ldg must be a multiple of 16 and >= N rounded up to 64");
    if (reinterpret_cast<uintptr_t>(G) % 16 != 0 || reinterpret_cast<uintptr_t>(operands) % 16 != 0)
        return fail_msg(1, "adx_em_pass_bf16: G and operands must be 16-byte aligned");
    int64_t ldqt = 0;
    if (operand_bytes < adx_em_bf16_operand_bytes(M, N, K, &ldqt))
        return fail_msg(1, "adx_em_pass_bf16: operand buffer too small");
    const int kh = K <= 10 ? 1 : 2;
    const __bf16* Px = static_cast<const __bf16*>(operands);
    const __bf16* Qx = Px + M * 32 * kh;
    const __bf16* Qt = Qx + N * 32 * kh;
    float* Tpart = reinterpret_cast<float*>(static_cast<uint8_t*>(operands) + bf16_base_bytes(M, N, K));
    const Bf16Grid gd = bf16_grid(M, N);
    hipStream_t st = (hipStream_t)stream;
    if (kh == 1)
        launch_bf16<1>(gd, st, G, ldg, M, N, K, Px, Qx, Qt, ldqt, A, B, T, Tpart);
    else
        launch_bf16<2>(gd, st, G, ldg, M, N, K, Px, Qx, Qt, ldqt, A, B, T, Tpart);
    ADX_LAUNCH_CHECK("adx_em_pass_bf16");
    return 0;
}

}  // extern "C"
