// Mixed-precision (BF16 / FP16 operand) form of the fused decode +
// EM-statistics pass, FP32 accumulation.
//
// Same tiling as em_pass_kernel (adx_em.hip): a workgroup owns 64*CW SNPs and
// a sample strip, lane (g, c) streams 16 genotype bytes of one byte row per
// 64-sample sub-tile, A and B live in MFMA accumulators for the strip and T
// is reduced per sub-tile.  The contractions run on 16x16x32 MFMA:
//
//   H^T  = Qx . Px^T (BF16) with Qx = 8 [Q_hi, Q_hi, Q_lo], Px = [P_hi, P_lo,
//          P_hi] (three-term split: |H - H_exact| <~ 3 * 2^-16 |H|); one
//          MFMA for K <= 10, two for K <= 16.  The factor 8 (exact) gives
//          h8 = 8 H.
//   A^T += Qt . a'^T,  B^T += Qt . b'^T   (FP16)
//          k = 32 samples (two 16-sample row blocks); Qt = 256 Q rounded to
//          FP16, a' = a / 8 and b' = b / 8 rounded to FP16 (the operands of
//          each product carry the same 11-bit rounding, as P does in T).
//   T   += [a' b'] . [P ; 1 - P]   (FP16)  k = 16 SNPs x {a', b'}: a' and b'
//          of one genotype share one word (f16 pair), the 16 x 16 word tile
//          is transposed through LDS (one 32-bit write per genotype, one
//          128-bit read per lane), and the colsum(b) term of em_batch_math
//          is the (1 - P) half.  P here is P_hi + P_lo rounded to FP16.
//
// Per genotype: hc8 = clamp(h8, 8e-5, 8 (1 - 1e-5)) (H clamped to [1e-5,
// 1 - 1e-5], em_adam_gpu.py:62-72), 1 / (8 H) = rcp(hc8), 1 / (8 (1 - H)) =
// rcp(8 - hc8) in FP32, both rounded to FP16 (at most 12500, in range) and
// multiplied, two genotypes per packed FP16 multiply, by g and 2 - g (0 for
// a missing genotype) looked up from the 2-bit codes with byte permutes.
// The slice reductions undo the power-of-two scales exactly (T x 8,
// A and B x 1/32).  The operand arrays Px, Qx, Qt are rebuilt from the FP64
// P, Q by adx_em_prep_bf16 before every pass.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x2 __attribute__((ext_vector_type(2)));

// 16x16x32 FP16 MFMA, FP32 accumulation; operand and accumulator layouts as
// mma_bf16.
__device__ __forceinline__ floatx4 mma_f16(f16x8 a, f16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

// High bytes of the FP16 genotype weights indexed by the 2-bit code (byte c
// of the word; the low bytes are 0): g = {0, 1, 2, -} and 2 - g = {2, 1, 0, -},
// 0 for the missing code 3.  FP16 1.0 = 0x3C00, 2.0 = 0x4000.
constexpr uint32_t kWeightA = 0x00403C00u;
constexpr uint32_t kWeightB = 0x00003C40u;

// FP16 pairs (a'_j, a'_{j+1}) and (b'_j, b'_{j+1}) of two genotypes with codes
// in bytes j, j+1 of cw (sel = byte selector 0x01040004 for j = 0, 0x03040204
// for j = 2) and FP32 reciprocals ra = 1 / (8 H), rb = 1 / (8 (1 - H)).
__device__ __forceinline__ void weight_pair(uint32_t cw, uint32_t sel, float ra0, float ra1, float rb0, float rb1,
                                            uint32_t& a, uint32_t& b) {
    // byte selector [4, c_j, 4, c_j+1]: selector 4 picks the zero first operand
    const uint32_t s = __builtin_amdgcn_perm(0x04040404u, cw, sel);
    const f16x2 ga = __builtin_bit_cast(f16x2, __builtin_amdgcn_perm(0u, kWeightA, s));
    const f16x2 gb = __builtin_bit_cast(f16x2, __builtin_amdgcn_perm(0u, kWeightB, s));
    a = __builtin_bit_cast(uint32_t, ga * f16x2{(_Float16)ra0, (_Float16)ra1});
    b = __builtin_bit_cast(uint32_t, gb * f16x2{(_Float16)rb0, (_Float16)rb1});
}

__device__ __forceinline__ bf16x8 load_bf16x8(const __bf16* p) {
    return *reinterpret_cast<const bf16x8*>(p);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

template <int KH, int CW, int WV>
__global__ __launch_bounds__(64 * WV) void em_pass_bf16_kernel(
    const uint8_t* __restrict__ G, int64_t ldg, int64_t M, int64_t N, int K, const __bf16* __restrict__ Px,
    const __bf16* __restrict__ Qx, const _Float16* __restrict__ Qt, int64_t ldqt, int64_t n_per_split,
    float* __restrict__ Tpart, float* __restrict__ ABpart, int kb0) {
    static_assert(CW % 4 == 0, "CW must be a multiple of 4");
    constexpr int LX = 32 * KH;  // row length of Px / Qx
    constexpr int TS = 20;       // row stride (words) of the per-wave transpose tile
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int64_t snp_w = ((int64_t)blockIdx.x * WV + wave) * (16 * CW);
    const int64_t row_w = snp_w >> 2;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t n0 = (int64_t)blockIdx.y * n_per_split;
    const int64_t n1 = n0 + n_per_split < N ? n0 + n_per_split : N;

    // Per wave: word [sample row r][SNP column c] = f16 a' | f16 b' << 16.
    __shared__ __attribute__((aligned(16))) uint32_t s_tile[WV][16 * TS];
    // Per wave: T of the 64 samples of the current row block over the wave's
    // SNPs; each (sample, k) is written once per wave and row block, and the
    // WV slices are summed in a fixed order at the flush (deterministic).
    // Rows padded to 17 floats so the 64 lanes of a write hit distinct banks.
    // NB = 2 (KH <= 2): this and the staged Q chunks below are double
    // buffered, so one workgroup barrier per sub-tile separates the writes of
    // one buffer from the reads of the other; with KH > 2 the Q chunks are
    // too large for two workgroups per CU and one buffer with two barriers
    // is kept.
    constexpr int NB = KH <= 2 ? 2 : 1;
    __shared__ float s_tsum[NB][WV][64 * 17];

    const bf16x8 zero8 = bf16x8{};
    bf16x8 pH[CW][KH];
    f16x8 pT[CW];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int t = cb >> 2, f = cb & 3;
        const int64_t snp = snp_w + 64 * t + 4 * cl + f;
#pragma unroll
        for (int kh = 0; kh < KH; ++kh) pH[cb][kh] = snp < M ? load_bf16x8(Px + snp * LX + 32 * kh + 8 * g) : zero8;
        // T B-operand: k = 8g + i <-> (SNP column 4g + i/2, P for even i, 1 - P
        // for odd i), matching the (a', b') pairs of the tile words; column
        // kp = cl.  P = P_hi + P_lo from the Px row.
        f16x8 v;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int64_t s2 = snp_w + 64 * t + 4 * (4 * g + (i >> 1)) + f;
            float x = 0.0f;
            if (s2 < M && kb0 + cl < K) {
                const float p = (float)Px[s2 * LX + kb0 + cl] + (float)Px[s2 * LX + K + kb0 + cl];
                x = (i & 1) ? 1.0f - p : p;
            }
            v[i] = (_Float16)x;
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

    // Genotype words: lane (g, cl) reads byte row row_w + 16t + cl (clamped
    // into the matrix), and gor[t] marks as missing the SNP fields past M
    // in the last byte row and whole rows past it.
    const uint8_t* grow[CW / 4];
    uint32_t gor[CW / 4];
#pragma unroll
    for (int t = 0; t < CW / 4; ++t) {
        const int64_t row = row_w + 16 * t + cl;
        const int64_t nf = M - 4 * row;
        grow[t] = G + (row < Mb ? row : Mb - 1) * ldg + 16 * g;
        gor[t] = row >= Mb ? 0xFFFFFFFFu : nf < 4 ? ((0xFFu << (2 * nf)) & 0xFFu) * 0x01010101u : 0u;
    }
    // The Q operands of a 64-sample sub-tile are staged once per workgroup by
    // its first 256 threads (waves 0-3; the others skip the staging):
    // thread t < 128 copies Qt chunk (row t >> 3, samples 8 (t & 7) .. + 7)
    // and every thread KH Qx chunks (16 bytes each) into LDS in lane-fragment
    // order,
    //   s_qt[tp 64 + lane]                   = Qt row cl, samples 16g + 8tp .. + 7
    //   s_qx[((2 tp + rr) KH + kh) 64 + lane] = Qx row 16 (cl >> 2) + 4 (2tp + rr) + (cl & 3),
    //                                          columns 32 kh + 8 g .. + 7 (zero past N)
    // so each wave's 128-bit operand reads hit consecutive addresses.
    __shared__ u32x4 s_qt[NB][2 * 64];
    __shared__ u32x4 s_qx[NB][4 * KH * 64];
    const int tid = threadIdx.x;
    // Qt rows of this k block: kb0 .. kb0 + 15
    const bool qt_stager = tid < 128;
    const _Float16* qt_src = Qt + (int64_t)(kb0 + ((tid >> 3) & 15)) * ldqt + 8 * (tid & 7);
    const int qt_dst = (tid & 1) * 64 + 16 * ((tid & 7) >> 1) + ((tid >> 3) & 15);
    int qx_row[KH], qx_col[KH], qx_dst[KH];
#pragma unroll
    for (int j = 0; j < KH; ++j) {
        const int idx = j * 256 + tid;
        const int q = idx / (4 * KH), ch = idx % (4 * KH);
        qx_row[j] = q;
        qx_col[j] = 8 * ch;
        qx_dst[j] = ((((q >> 2) & 3) * KH + (ch >> 2)) * 64) + 16 * (ch & 3) + (((q >> 4) << 2) | (q & 3));
    }
    const u32x4 zero4 = u32x4{0, 0, 0, 0};
    const bool stager = WV == 4 || wave < 4;
    u32x4 qn_t = zero4, qn_x[KH];
    uint32_t gn[CW / 4][4];
    auto load_next = [&](int64_t s) {
        if (stager) {
            if (qt_stager) qn_t = *reinterpret_cast<const u32x4*>(qt_src + s);
#pragma unroll
            for (int j = 0; j < KH; ++j) {
                const int64_t r = s + qx_row[j];
                const u32x4 v = *reinterpret_cast<const u32x4*>(Qx + (r < N ? r : N - 1) * LX + qx_col[j]);
                qn_x[j] = r < N ? v : zero4;
            }
        }
#pragma unroll
        for (int t = 0; t < CW / 4; ++t) {
            const uint4 v = *reinterpret_cast<const uint4*>(grow[t] + s);
            gn[t][0] = v.x;
            gn[t][1] = v.y;
            gn[t][2] = v.z;
            gn[t][3] = v.w;
        }
    };
    auto stage_next = [&](uint32_t (&gw)[CW / 4][4], int buf) {
        if (stager) {
            if (qt_stager) s_qt[buf][qt_dst] = qn_t;
#pragma unroll
            for (int j = 0; j < KH; ++j) s_qx[buf][qx_dst[j]] = qn_x[j];
        }
#pragma unroll
        for (int t = 0; t < CW / 4; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) gw[t][r] = gn[t][r] | gor[t];
    };
    uint32_t gw[CW / 4][4];
    load_next(n0);
    stage_next(gw, 0);
    __syncthreads();
    int pb = 0;  // buffer of the current sub-tile

    for (int64_t s0 = n0; s0 < n1; s0 += 64) {
        // The next sub-tile's genotype words and Q chunks are in flight
        // during this sub-tile's arithmetic (the last sub-tile reloads its
        // own, unused: the loads stay unconditional).
        load_next(s0 + 64 < n1 ? s0 + 64 : s0);

        // Row-block pairs: accumulator element j of lane group g in row block
        // rw is sample s0 + 16g + 4rw + j; a pair (rw = 2tp, 2tp+1) gives lane
        // (g, cl) the 8 consecutive samples s0 + 16g + 8tp + i, i = 0..7.
#pragma unroll
        for (int tp = 0; tp < 2; ++tp) {
            const f16x8 qT = __builtin_bit_cast(f16x8, s_qt[pb][tp * 64 + lane]);
            bf16x8 qH[2][KH];
#pragma unroll
            for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                for (int kh = 0; kh < KH; ++kh)
                    qH[rr][kh] = __builtin_bit_cast(bf16x8, s_qx[pb][((2 * tp + rr) * KH + kh) * 64 + lane]);
            floatx4 accT[2] = {floatx4{0, 0, 0, 0}, floatx4{0, 0, 0, 0}};

            // H of column block cb + 1 is issued before the weights of block
            // cb are formed, so the MFMA latency overlaps that arithmetic
            // (consuming H right after its MFMA stalled the wave).
            auto hprod = [&](int cb, floatx4 (&h)[2]) {
#pragma unroll
                for (int rr = 0; rr < 2; ++rr) {
                    h[rr] = floatx4{0, 0, 0, 0};
#pragma unroll
                    for (int kh = 0; kh < KH; ++kh) h[rr] = mma_bf16(qH[rr][kh], pH[cb][kh], h[rr]);
                }
            };
            floatx4 hc[2], hn[2];
            hprod(0, hc);
#pragma unroll
            for (int cb = 0; cb < CW; ++cb) {
                const int t = cb >> 2, f = cb & 3;
                if (cb + 1 < CW) hprod(cb + 1, hn);
                // keep the scheduler from sinking those MFMAs to their uses
                __builtin_amdgcn_sched_barrier(0);
                // wa[rr][p] / wb[rr][p]: f16 pairs of a' / b' for samples 2p, 2p + 1
                // of the 4 in accumulator order
                uint32_t wa[2][2], wb[2][2];
#pragma unroll
                for (int rr = 0; rr < 2; ++rr) {
                    const floatx4 h = hc[rr];
                    const uint32_t cw = (gw[t][2 * tp + rr] >> (2 * f)) & 0x03030303u;
                    float ra[4], rb[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const float hc8 = __builtin_amdgcn_fmed3f(h[j], 8.0f * 1e-5f, 8.0f * (1.0f - 1e-5f));
                        ra[j] = __builtin_amdgcn_rcpf(hc8);
                        rb[j] = __builtin_amdgcn_rcpf(8.0f - hc8);
                    }
                    weight_pair(cw, 0x01040004u, ra[0], ra[1], rb[0], rb[1], wa[rr][0], wb[rr][0]);
                    weight_pair(cw, 0x03040204u, ra[2], ra[3], rb[2], rb[3], wa[rr][1], wb[rr][1]);
#pragma unroll
                    for (int p = 0; p < 2; ++p) {
                        tile[(4 * g + 2 * p) * TS + cl] = __builtin_amdgcn_perm(wb[rr][p], wa[rr][p], 0x05040100u);
                        tile[(4 * g + 2 * p + 1) * TS + cl] = __builtin_amdgcn_perm(wb[rr][p], wa[rr][p], 0x07060302u);
                    }
                    // The tile is written and read by lanes of this wave only;
                    // LDS operations of a wave complete in order, and the
                    // wave barrier keeps the compiler from moving them.
                    __builtin_amdgcn_wave_barrier();
                    // T (16 samples x 16 kp) += [a' b'] (16 x 32) . [P ; 1-P] (32 x 16)
                    const f16x8 tv = __builtin_bit_cast(f16x8, *reinterpret_cast<const u32x4*>(tile + cl * TS + 4 * g));
                    __builtin_amdgcn_wave_barrier();
                    accT[rr] = mma_f16(tv, pT[cb], accT[rr]);
                }
                // B operands of A^T, B^T: k = sample 8g + i (rr = i / 4, j = i % 4).
                const f16x8 aB = __builtin_bit_cast(f16x8, u32x4{wa[0][0], wa[0][1], wa[1][0], wa[1][1]});
                const f16x8 bB = __builtin_bit_cast(f16x8, u32x4{wb[0][0], wb[0][1], wb[1][0], wb[1][1]});
                accA[cb] = mma_f16(qT, aB, accA[cb]);
                accB[cb] = mma_f16(qT, bB, accB[cb]);
                if (cb + 1 < CW) {
                    hc[0] = hn[0];
                    hc[1] = hn[1];
                }
            }
#pragma unroll
            for (int rr = 0; rr < 2; ++rr)
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    if (kb0 + cl < K) s_tsum[pb][wave][(16 * g + 4 * (2 * tp + rr) + j) * 17 + cl] = accT[rr][j];
        }

        if constexpr (NB == 2) {
            // the other buffers were last read before the previous barrier
            stage_next(gw, pb ^ 1);
            __syncthreads();
        } else {
            __syncthreads();
            // every wave is past its reads of the staged Q chunks
            stage_next(gw, 0);
        }
        // T of these 64 samples over this workgroup's SNPs: one plain store
        // into the workgroup's slice of the partial buffer (every (slice,
        // sample) is written exactly once), summed by tpart_reduce_kernel.
        // Device-scope float atomics would be performed at the memory side
        // (one 4-byte read-modify-write each, ~N K M / (64 CW) per pass).
        for (int e = threadIdx.x; e < 64 * 16; e += 64 * WV) {
            const int kp = e & 15;
            const int64_t smp = s0 + (e >> 4);
            const int ex = (e >> 4) * 17 + kp;
            if (kb0 + kp < K && smp < N)
                Tpart[((int64_t)blockIdx.x * N + smp) * K + kb0 + kp] = sum_wave_slices<WV>(s_tsum[pb], ex);
        }
        if constexpr (NB == 2)
            pb ^= 1;
        else
            __syncthreads();
    }

#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int t = cb >> 2, f = cb & 3;
        const int64_t snp = snp_w + 64 * t + 4 * cl + f;
        if (snp < M) {
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int kp = kb0 + 4 * g + jj;
                if (kp < K) {
                    // this sample split's slice of the A / B partial sums (each
                    // element written by one workgroup), summed in a fixed order
                    const int64_t mk = M * (int64_t)K;
                    ABpart[(int64_t)blockIdx.y * mk + snp * K + kp] = accA[cb][jj];
                    ABpart[((int64_t)gridDim.y + blockIdx.y) * mk + snp * K + kp] = accB[cb][jj];
                }
            }
        }
    }
}

// Operand arrays of the pass from the FP64 master P (M x K) and Q (N x K):
//   Px[m]  = [P_hi(0..K), P_lo, P_hi, 0 ...]    (32*KH bf16)
//   Qx[i]  = 8 [Q_hi, Q_hi, Q_lo, 0 ...]        (32*KH bf16)
//   Qt     = rows 0..16KB-1 x ldqt, Qt[kp][i] = 256 Q[i][kp] rounded to FP16
//            (KB = ceil(K / 16) k blocks)
// with x_hi = bf16(x), x_lo = bf16(x - x_hi); zero outside the valid range.
// The power-of-two factors are exact.
__global__ void prep_px_kernel(const double* __restrict__ P, int64_t M, int K, int LX, __bf16* __restrict__ Px) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    __bf16* row = Px + m * LX;
    for (int k = 0; k < K; ++k) {
        const float x = (float)P[m * K + k];
        const __bf16 hi = (__bf16)x;
        const __bf16 lo = (__bf16)(x - (float)hi);
        row[k] = hi;
        row[K + k] = lo;
        row[2 * K + k] = hi;
    }
    for (int e = 3 * K; e < LX; ++e) row[e] = (__bf16)0.0f;
}

__global__ void prep_qx_kernel(const double* __restrict__ Q, int64_t N, int K, int LX, int KB,
                               __bf16* __restrict__ Qx, _Float16* __restrict__ Qt, int64_t ldqt) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ldqt) return;
    __bf16* row = i < N ? Qx + i * LX : nullptr;
    for (int k = 0; k < 16 * KB; ++k) {
        _Float16 t = (_Float16)0.0f;
        if (i < N && k < K) {
            const float x = (float)Q[i * K + k];
            const __bf16 hi = (__bf16)(8.0f * x);
            const __bf16 lo = (__bf16)(8.0f * x - (float)hi);
            row[k] = hi;
            row[K + k] = hi;
            row[2 * K + k] = lo;
            t = (_Float16)(256.0f * x);
        }
        Qt[(int64_t)k * ldqt + i] = t;
    }
    if (row)
        for (int e = 3 * K; e < LX; ++e) row[e] = (__bf16)0.0f;
}

// 32-element chunks of the H operand rows (3K terms) and 16-wide k blocks.
static int bf16_kh(int K) { return K <= 10 ? 1 : K <= 21 ? 2 : K <= 32 ? 3 : K <= 42 ? 4 : 6; }
static int bf16_kb(int K) { return (K + 15) / 16; }

struct Bf16Grid {
    int cw, waves;
    int64_t nbx, splits, n_per_split;
};

// K > 16: one launch per 16-wide k block at CW = 4 (H over the full K in each).
// K = 11..16 (KH = 2) also runs at CW = 4: at CW = 8 that kernel needs more
// than 256 registers per lane (one wave per SIMD) and took 81.9 ms against
// 67.7 ms at N=200k, M=600k, K=12.  8-wave workgroups for K <= 16.
static Bf16Grid bf16_grid(int64_t M, int64_t N, int K) {
    const EmTiles tl = em_tiles(ADX_BF16);
    Bf16Grid gd;
    gd.cw = bf16_kh(K) >= 2 ? 4 : tl.cw;
    gd.waves = K > 16 ? 4 : tl.waves;
    gd.nbx = ceil_div(M, 16 * gd.cw * gd.waves);
    int64_t s = ceil_div(tl.target_wgs, gd.nbx);
    const int64_t max_s = ceil_div(N, 64);
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    if (s > 65535) s = 65535;
    gd.n_per_split = ceil_div(ceil_div(N, s), 64) * 64;
    gd.splits = ceil_div(N, gd.n_per_split);
    return gd;
}

// T[e] += scale * (sum over the nparts slices of Tpart[p * nk + e]) (also
// used for the A and B slices, nk = M K); scale is a power of two.
__global__ void tpart_reduce_kernel(const float* __restrict__ Tpart, int64_t nparts, int64_t nk, float scale,
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
    T[e] += scale * ((s0 + s1) + (s2 + s3));
}

template <int KH>
static void launch_bf16(const Bf16Grid& gd, hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N,
                        int K, const __bf16* Px, const __bf16* Qx, const _Float16* Qt, int64_t ldqt, float* A,
                        float* B, float* T, float* Tpart) {
    dim3 grid((unsigned)gd.nbx, (unsigned)gd.splits);
    float* ABpart = Tpart + gd.nbx * N * (int64_t)K;
    const int kbt = bf16_kb(K);
#define ADX_BF16_LAUNCH(cw, wv)                                                                               \
    em_pass_bf16_kernel<KH, cw, wv><<<grid, 64 * wv, 0, st>>>(G, ldg, M, N, K, Px, Qx, Qt, ldqt, gd.n_per_split, \
                                                              Tpart, ABpart, 16 * kb)
    for (int kb = 0; kb < kbt; ++kb) {
        bool done = false;
        if constexpr (KH <= 2) {
            if (gd.waves == 8) {
                if constexpr (KH == 1) {
                    if (gd.cw == 8) {
                        ADX_BF16_LAUNCH(8, 8);
                        done = true;
                    }
                }
                if (!done) {
                    ADX_BF16_LAUNCH(4, 8);
                    done = true;
                }
            }
        }
        if constexpr (KH == 1) {
            if (!done && gd.cw == 8) {
                ADX_BF16_LAUNCH(8, 4);
                done = true;
            }
        }
        if (!done) ADX_BF16_LAUNCH(4, 4);
    }
#undef ADX_BF16_LAUNCH
    // slices hold T / 8 (a' = a / 8) and 32 A, 32 B (Qt = 256 Q)
    const int64_t nk = N * K, mk = M * K;
    tpart_reduce_kernel<<<(unsigned)ceil_div(nk, 256), 256, 0, st>>>(Tpart, gd.nbx, nk, 8.0f, T);
    tpart_reduce_kernel<<<(unsigned)ceil_div(mk, 256), 256, 0, st>>>(ABpart, gd.splits, mk, 1.0f / 32.0f, A);
    tpart_reduce_kernel<<<(unsigned)ceil_div(mk, 256), 256, 0, st>>>(ABpart + gd.splits * mk, gd.splits, mk,
                                                                      1.0f / 32.0f, B);
}

// Byte offset of the partial-T slices in the operand buffer, and its size.
static int64_t bf16_base_bytes(int64_t M, int64_t N, int K) {
    const int kh = bf16_kh(K);
    const int64_t ld = ceil_div(N, 64) * 64;
    return ceil_div((int64_t)sizeof(__bf16) * (M * 32 * kh + N * 32 * kh + 16 * bf16_kb(K) * ld), 256) * 256;
}

}  // namespace adx

using namespace adx;

extern "C" {

int64_t adx_em_bf16_operand_bytes(int64_t M, int64_t N, int K, int64_t* ldqt) {
    const int64_t ld = ceil_div(N, 64) * 64;
    if (ldqt) *ldqt = ld;
    if (M <= 0 || N <= 0) return bf16_base_bytes(M > 0 ? M : 0, N > 0 ? N : 0, K);
    const Bf16Grid gd = bf16_grid(M, N, K);
    return bf16_base_bytes(M, N, K) + (int64_t)sizeof(float) * (gd.nbx * N * K + 2 * gd.splits * M * K);
}

int adx_em_prep_bf16(const double* P, const double* Q, int64_t M, int64_t N, int K, void* operands,
                     int64_t operand_bytes, void* stream) {
    if (K < 1 || K > 64) return fail_msg(1, "adx_em_prep_bf16: K must be in [1, 64]");
    int64_t ldqt = 0;
    if (operand_bytes < adx_em_bf16_operand_bytes(M, N, K, &ldqt))
        return fail_msg(1, "adx_em_prep_bf16: operand buffer too small");
    const int kh = bf16_kh(K);
    __bf16* Px = static_cast<__bf16*>(operands);
    __bf16* Qx = Px + M * 32 * kh;
    _Float16* Qt = reinterpret_cast<_Float16*>(Qx + N * 32 * kh);
    hipStream_t st = (hipStream_t)stream;
    if (M > 0) prep_px_kernel<<<(unsigned)ceil_div(M, 256), 256, 0, st>>>(P, M, K, 32 * kh, Px);
    prep_qx_kernel<<<(unsigned)ceil_div(ldqt, 256), 256, 0, st>>>(Q, N, K, 32 * kh, bf16_kb(K), Qx, Qt, ldqt);
    ADX_LAUNCH_CHECK("adx_em_prep_bf16");
    return 0;
}

int adx_em_pass_bf16(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K, void* operands,
                     int64_t operand_bytes, float* A, float* B, float* T, void* stream) {
    if (M <= 0 || N <= 0) return fail_msg(1, "adx_em_pass_bf16: empty genotype matrix");
    if (K < 1 || K > 64) return fail_msg(1, "adx_em_pass_bf16: K must be in [1, 64]");
    if (ldg % 16 != 0 || ldg < ceil_div(N, 64) * 64)
        return fail_msg(1, "adx_em_pass_bf16: ldg must be a multiple of 16 and >= N rounded up to 64");
    if (reinterpret_cast<uintptr_t>(G) % 16 != 0 || reinterpret_cast<uintptr_t>(operands) % 16 != 0)
        return fail_msg(1, "adx_em_pass_bf16: G and operands must be 16-byte aligned");
    int64_t ldqt = 0;
    if (operand_bytes < adx_em_bf16_operand_bytes(M, N, K, &ldqt))
        return fail_msg(1, "adx_em_pass_bf16: operand buffer too small");
    const int kh = bf16_kh(K);
    const __bf16* Px = static_cast<const __bf16*>(operands);
    const __bf16* Qx = Px + M * 32 * kh;
    const _Float16* Qt = reinterpret_cast<const _Float16*>(Qx + N * 32 * kh);
    float* Tpart = reinterpret_cast<float*>(static_cast<uint8_t*>(operands) + bf16_base_bytes(M, N, K));
    const Bf16Grid gd = bf16_grid(M, N, K);
    hipStream_t st = (hipStream_t)stream;
    switch (kh) {
        case 1: launch_bf16<1>(gd, st, G, ldg, M, N, K, Px, Qx, Qt, ldqt, A, B, T, Tpart); break;
        case 2: launch_bf16<2>(gd, st, G, ldg, M, N, K, Px, Qx, Qt, ldqt, A, B, T, Tpart); break;
        case 3: launch_bf16<3>(gd, st, G, ldg, M, N, K, Px, Qx, Qt, ldqt, A, B, T, Tpart); break;
        case 4: launch_bf16<4>(gd, st, G, ldg, M, N, K, Px, Qx, Qt, ldqt, A, B, T, Tpart); break;
        default: launch_bf16<6>(gd, st, G, ldg, M, N, K, Px, Qx, Qt, ldqt, A, B, T, Tpart); break;
    }
    ADX_LAUNCH_CHECK("adx_em_pass_bf16");
    return 0;
}

}  // extern "C"
