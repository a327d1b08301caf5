// Fused genotype-decode + EM-statistics pass, EM finalize and Adam step.
//
// Reference: adamixture/model/em_adam_gpu.py.  The reference materialises
// six C x N FP64 temporaries per SNP chunk and runs four skinny GEMMs
// (em_batch_math, :46-78).  Here one launch streams the 2-bit genotypes once
// and keeps every intermediate in registers:
//
//   tile (16 samples x 16 SNPs), formulation "H^T" (rows = samples):
//     H^T  = Q_tile . P_tile^T                     MFMA, k = K (padded to 4)
//     a, b = g/H, (2-g)/(1-H) on observed codes    VALU, in accumulator layout
//     A^T += Q_tile^T . a^T,  B^T += Q_tile^T . b^T  MFMA, k = samples; the
//            accumulator layout of H^T is exactly the B-operand layout here
//     T   += c^T . P_tile (c = a - b)              MFMA, k = SNPs; c^T goes
//            through a 1 KB per-wave LDS transpose
//     T   += colsum_SNP(b)                         VALU + 16-lane shuffles
//
// Work split: a workgroup (4 waves) owns 64*CW consecutive SNPs and one
// sample strip; each wave owns 16*CW SNPs = 4*CW byte rows.  Within a 64-
// sample sub-tile, lane (g, c) loads 16 contiguous bytes (samples
// s0+16g .. s0+16g+15) of byte row 16t+c, so every genotype byte is read
// once, coalesced, and supplies the four SNP column blocks cb = 4t..4t+3.
// A and B stay in MFMA accumulators for the whole strip; T is reduced per
// sub-tile across the 4 waves in LDS and stored into the workgroup's slice
// of a partial buffer (summed in a fixed order afterwards).
//
// K > 16 (up to the reference GPU limit of 64): the output k range of a
// launch is the 16-wide block kb0 .. kb0 + 15 (A^T, B^T, T tiles are 16
// wide); H is formed over the full K (KS up to 16 MFMA steps) in every
// launch, one launch per k block.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <mutex>
#include <string>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

// The weights g and 2 - g of a genotype (0 for a missing code) are read
// from two masked copies of its genotype word, built once per word:
// field i of ga holds g (0 for code 3), field i of gb holds 2 - g (0 for
// code 3).  Same values as converting the code and selecting 0 for a
// missing one, without a compare or a select per genotype.
struct WeightWords {
    uint32_t ga, gb;
};
__device__ __forceinline__ WeightWords weight_words(uint32_t w) {
    const uint32_t lo = w & 0x55555555u, hi = (w >> 1) & 0x55555555u;
    const uint32_t miss = lo & hi;
    return WeightWords{w & ~(miss | (miss << 1)), ((~(lo | hi) & 0x55555555u) << 1) | (lo & ~hi)};
}

template <typename T>
__device__ __forceinline__ void em_terms(T h, T ga, T gb, T& a, T& b);

// FP32: one reciprocal of h(1-h) gives both 1/h and 1/(1-h).  The clamp is
// one med3 (equal to fmin(fmax(h, lo), hi) for every non-NaN h).
template <>
__device__ __forceinline__ void em_terms<float>(float h, float ga, float gb, float& a, float& b) {
    const float hc = __builtin_amdgcn_fmed3f(h, 1e-5f, 1.0f - 1e-5f);
    const float u = 1.0f - hc;
    const float r = __builtin_amdgcn_rcpf(hc * u);
    a = ga * (u * r);
    b = gb * (hc * r);
}

// FP64: the reference's quotients g / H and (2-g) / (1-H)
// (em_adam_gpu.py:62-67) as g and 2-g times 1/H = (1-H) r and 1/(1-H) = H r
// with r = 1 / (H (1-H)), the hardware estimate refined by two
// Newton-Raphson steps r <- r + r (1 - x r) (x in [1e-5, 1/4]); a and b stay
// within a few ulp of the correctly rounded quotients, the size of a change
// of summation order in the contractions that follow.
__device__ __forceinline__ double recip_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

template <>
__device__ __forceinline__ void em_terms<double>(double h, double ga, double gb, double& a, double& b) {
    const double hc = fmin(fmax(h, 1e-5), 1.0 - 1e-5);
    const double u = 1.0 - hc;
    const double r = recip_nr(hc * u);
    a = ga * (u * r);
    b = gb * (hc * r);
}

// Row stride of the per-wave transpose tile (see em_pass_body).
constexpr int TSX = 17;

// The pass body.  LLFAST: the log-likelihood factors skip the clamp of H to
// [1e-10, 1 - 1e-10], which ll_domain_kernel has shown to be a no-op for
// this (P, Q) (LogProd::mul), saving three of its FP64 VALU instructions
// per genotype.  The LDS arrays are the kernel's.
template <typename T, int KS, int CW, int WV, bool STATS, bool LL, bool LLFAST>
__device__ __forceinline__ void em_pass_body(
    const uint8_t* __restrict__ G, int64_t ldg, int64_t M, int64_t N, int K, const T* __restrict__ P,
    int64_t ldp, const T* __restrict__ Q, int64_t ldq, int64_t n_per_split, T* __restrict__ Aacc,
    T* __restrict__ Bacc, T* __restrict__ Tacc, double* __restrict__ ll_part, T* __restrict__ Tpart,
    T* __restrict__ ABpart, int kb0, T (*s_tile)[16 * TSX], T (*s_tsum)[64 * 17], double (*s_ll)[1]) {
    using MM = Mma16<T>;
    using acc_t = typename MM::acc;
    static_assert(CW % 4 == 0, "CW must be a multiple of 4");

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int64_t snp_w = ((int64_t)blockIdx.x * WV + wave) * (16 * CW);
    const int64_t row_w = snp_w >> 2;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t n0 = (int64_t)blockIdx.y * n_per_split;
    const int64_t n1 = n0 + n_per_split < N ? n0 + n_per_split : N;


    // P operands, constant for the whole strip.
    T pH[CW][KS];
    T pT[CW][4];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        const int t = cb >> 2, f = cb & 3;
        const int64_t snp = snp_w + 64 * t + 4 * cl + f;
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int kp = 4 * ks + g;
            pH[cb][ks] = (snp < M && kp < K) ? P[snp * ldp + kp] : T(0);
        }
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
            const int64_t snp2 = snp_w + 64 * t + 4 * (4 * kk + g) + f;
            pT[cb][kk] = (STATS && snp2 < M && kb0 + cl < K) ? P[snp2 * ldp + kb0 + cl] : T(0);
        }
    }

    acc_t accA[CW], accB[CW];
#pragma unroll
    for (int cb = 0; cb < CW; ++cb) {
        accA[cb] = acc_t{0, 0, 0, 0};
        accB[cb] = acc_t{0, 0, 0, 0};
    }
    LogProd lp;
    // s_tsum holds two sets of WV slices, used by alternate sub-tiles: the
    // barrier after a sub-tile's slices are written also separates the
    // previous flush's reads of the other set from the next writes to it.
    int pb = 0;
    __syncthreads();

    for (int64_t s0 = n0; s0 < n1; s0 += 64) {
        // Genotype bytes: 16 samples of byte row 16t+cl per t.
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
                if (nf < 4) {  // SNP fields beyond M read as missing
                    const uint32_t pm = ((0xFFu << (2 * nf)) & 0xFFu) * 0x01010101u;
#pragma unroll
                    for (int r = 0; r < 4; ++r) gw[t][r] |= pm;
                }
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) gw[t][r] = 0xFFFFFFFFu;
            }
        }

        // Row blocks of 16 samples: the accumulator element j of lane group g
        // (D-row MM::row(g, j)) is sample s0+16g+4rw+j, so lane (g, cl) reads
        // bytes 4rw..4rw+3 of its 16.
#pragma unroll
        for (int rw = 0; rw < 4; ++rw) {
            T qH[KS];
            T qT[4];
            // D-row r holds the sample of lane group row_group(r), element row_elem(r).
            const int64_t sH = s0 + 16 * row_group<T>(cl) + 4 * rw + row_elem<T>(cl);
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int kp = 4 * ks + g;
                qH[ks] = (sH < N && kp < K) ? Q[sH * ldq + kp] : T(0);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int64_t sT = s0 + 16 * g + 4 * rw + j;
                qT[j] = (STATS && sT < N && kb0 + cl < K) ? Q[sT * ldq + kb0 + cl] : T(0);
            }
            acc_t accT = acc_t{0, 0, 0, 0};
            T sb[4] = {T(0), T(0), T(0), T(0)};
            CodeMasks cm[CW / 4];
            WeightWords ww[CW / 4];
#pragma unroll
            for (int t = 0; t < CW / 4; ++t) {
                if (LL) cm[t] = code_masks(gw[t][rw]);
                if (STATS) ww[t] = weight_words(gw[t][rw]);
            }

#pragma unroll
            for (int cb = 0; cb < CW; ++cb) {
                const int t = cb >> 2, f = cb & 3;
                acc_t h = acc_t{0, 0, 0, 0};
#pragma unroll
                for (int ks = 0; ks < KS; ++ks) h = MM::run(qH[ks], pH[cb][ks], h);
                // bytes j of wa, wb: g and 2 - g of sample j's genotype
                const uint32_t wa = (ww[t].ga >> (2 * f)) & 0x03030303u;
                const uint32_t wb = (ww[t].gb >> (2 * f)) & 0x03030303u;
                T a[4], b[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if (STATS)
                        em_terms<T>(h[j], (T)((wa >> (8 * j)) & 0xFFu), (T)((wb >> (8 * j)) & 0xFFu), a[j], b[j]);
                    if (LL) lp.mul<!LLFAST>((double)h[j], cm[t], 2 * f + 8 * j);
                }
                // 8 factors (two column blocks) between renormalisations: the
                // mantissa stays >= 1e-160; the rescale is exact, so the result
                // does not depend on where it happens.  CW is a multiple of 4,
                // so the last column block of every row block renormalises.
                if (LL && (cb & 1)) lp.renorm();
                if (STATS) {
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        accA[cb] = MM::run(qT[j], a[j], accA[cb]);
                        accB[cb] = MM::run(qT[j], b[j], accB[cb]);
                        s_tile[wave][cl * TSX + MM::row(g, j)] = a[j] - b[j];
                        sb[j] += b[j];
                    }
#pragma unroll
                    for (int kk = 0; kk < 4; ++kk)
                        accT = MM::run(s_tile[wave][(4 * kk + g) * TSX + cl], pT[cb][kk], accT);
                }
            }

            if (STATS) {
                // colsum(b) over the wave's SNPs: reduce across the 16 lanes of a group.
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    T v = sb[j];
                    v += __shfl_xor(v, 1);
                    v += __shfl_xor(v, 2);
                    v += __shfl_xor(v, 4);
                    v += __shfl_xor(v, 8);
                    if (kb0 + cl < K) s_tsum[pb * WV + wave][(16 * g + 4 * rw + j) * 17 + cl] = accT[j] + v;
                }
            }
        }

        if (STATS) {
            // T of these 64 samples over this workgroup's SNPs (the WV wave
            // slices summed in a fixed order): a plain store into the
            // workgroup's slice of Tpart (each (slice, sample) is written
            // once; tpart_reduce_kernel sums the slices), or, without a
            // workspace, a device-scope atomic (performed at the memory side).
            __syncthreads();
            for (int e = threadIdx.x; e < 64 * 16; e += 64 * WV) {
                const int kp = e & 15;
                const int64_t smp = s0 + (e >> 4);
                const int ex = (e >> 4) * 17 + kp;
                if (kb0 + kp < K && smp < N) {
                    const T v = sum_wave_slices<WV>(s_tsum + pb * WV, ex);
                    if (Tpart)
                        Tpart[((int64_t)blockIdx.x * N + smp) * K + kb0 + kp] = v;
                    else
                        atomicAdd(&Tacc[smp * K + kb0 + kp], v);
                }
            }
            pb ^= 1;
        }
    }

    if (STATS) {
#pragma unroll
        for (int cb = 0; cb < CW; ++cb) {
            const int t = cb >> 2, f = cb & 3;
            const int64_t snp = snp_w + 64 * t + 4 * cl + f;
            if (snp < M) {
#pragma unroll
                for (int jj = 0; jj < 4; ++jj) {
                    const int kp = kb0 + MM::row(g, jj);
                    if (kp < K) {
                        // With a workspace: this sample split's slice of the A / B
                        // partial sums (each element written by one workgroup),
                        // summed in a fixed order by tpart_reduce_kernel.
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

    if (LL) {
        double ll = lp.log_value();
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) ll += __shfl_xor(ll, off);
        if (lane == 0) s_ll[wave][0] = ll;
        __syncthreads();
        if (threadIdx.x == 0) ll_part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sum_wave_slices<WV>(s_ll, 0);
    }
}

template <typename T, int KS, int CW, int WV, bool STATS, bool LL>
__global__ __launch_bounds__(64 * WV) void em_pass_kernel(
    const uint8_t* __restrict__ G, int64_t ldg, int64_t M, int64_t N, int K, const T* __restrict__ P,
    int64_t ldp, const T* __restrict__ Q, int64_t ldq, int64_t n_per_split, T* __restrict__ Aacc,
    T* __restrict__ Bacc, T* __restrict__ Tacc, double* __restrict__ ll_part, T* __restrict__ Tpart,
    T* __restrict__ ABpart, int kb0, const int* __restrict__ ll_fast) {
    // Per-wave 16 x 16 transpose tile for the T contraction, rows padded to
    // TSX = 17 elements: the column writes (lane (g, c) -> row c) and the row
    // reads (lane (g, c) -> column c) then fall on distinct banks within each
    // 16-lane (FP64) or 32-lane (FP32, at most 2-way) LDS pass, where an
    // unpadded 16-element stride puts the 16 lanes of a column write on one
    // bank.
    __shared__ T s_tile[WV][16 * TSX];
    // Per wave: T of the 64 samples of the current row block over the wave's
    // SNPs.  Each wave writes each (sample, k) once per row block, so the
    // slices need no atomics and are summed in a fixed order at the flush.
    // Rows padded to 17 elements: the 64 lanes of a write (4 lane groups x
    // 16 columns, rows 16 apart) fall on distinct banks.  Two sets of slices
    // (alternate sub-tiles): one workgroup barrier per sub-tile.
    __shared__ T s_tsum[STATS ? 2 * WV : 1][64 * 17];
    __shared__ double s_ll[WV][1];
    // ll_fast (written by ll_domain_kernel earlier on the stream, the same
    // value for every thread): one uniform choice of the body per launch.
    // The log-likelihood-only pass has both bodies; the fused statistics +
    // log-likelihood kernel keeps the clamped one (hipcc 7.2 fails on the
    // two inlined copies of that body).
    if constexpr (LL && !STATS) {
        if (ll_fast && *ll_fast) {
            em_pass_body<T, KS, CW, WV, STATS, LL, true>(G, ldg, M, N, K, P, ldp, Q, ldq, n_per_split, Aacc, Bacc,
                                                         Tacc, ll_part, Tpart, ABpart, kb0, s_tile, s_tsum, s_ll);
            return;
        }
    }
    em_pass_body<T, KS, CW, WV, STATS, LL, false>(G, ldg, M, N, K, P, ldp, Q, ldq, n_per_split, Aacc, Bacc, Tacc,
                                                  ll_part, Tpart, ABpart, kb0, s_tile, s_tsum, s_ll);
}

// Whether clamping H = Q P^T to [1e-10, 1 - 1e-10] (the log-likelihood's
// clamp, em_adam_gpu.py:238-246) is a no-op for this (P, Q): every P entry
// in [2e-10, 1 - 2e-10], every Q entry >= 0 and every Q row sum within 1e-12
// of 1.  Then each H is a convex combination of one SNP's P entries up to a
// relative 1e-12 (plus FP64 rounding), so 1e-10 < H < 1 - 1e-10.  *flag
// must be 1 on entry (ll_flag_init_kernel); any failing row (NaN included)
// stores 0 (every writer stores the same value).  Grid-stride over rows.
__global__ void ll_flag_init_kernel(int* flag) {
    if (threadIdx.x == 0) *flag = 1;
}

__global__ void ll_domain_kernel(const double* __restrict__ P, int64_t M, int64_t ldp, const double* __restrict__ Q,
                                 int64_t N, int64_t ldq, int K, int* __restrict__ flag) {
    const double lo = 2e-10, hi = 1.0 - 2e-10;
    bool ok = true;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < M + N; r += stride) {
        if (r < M) {
            for (int k = 0; k < K; ++k) {
                const double v = P[r * ldp + k];
                ok = ok && (v >= lo && v <= hi);
            }
        } else {
            double s = 0.0;
            for (int k = 0; k < K; ++k) {
                const double v = Q[(r - M) * ldq + k];
                ok = ok && (v >= 0.0);
                s += v;
            }
            ok = ok && (fabs(s - 1.0) <= 1e-12);
        }
    }
    if (!ok) *flag = 0;
}

// Deterministic sum of n FP64 partials (fixed order) into out[0].
__global__ __launch_bounds__(256) void reduce_f64_kernel(const double* __restrict__ x, int64_t n,
                                                         double* __restrict__ out) {
    __shared__ double s[256];
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) v += x[i];
    s[threadIdx.x] = v;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}

// em_adam_gpu.py:101-107 (P side), elementwise; zeroes A and B.
template <typename TA>
__global__ void finalize_p_kernel(const double* __restrict__ P, TA* __restrict__ A, TA* __restrict__ B,
                                  int64_t n, double* __restrict__ P_EM) {
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const double a = (double)A[x], b = (double)B[x], p = P[x];
    double d = (a - b) * p + b;
    d = fmax(d, 1e-5);
    P_EM[x] = ((1.0 / d) * p) * a;
    A[x] = TA(0);
    B[x] = TA(0);
}

// em_adam_gpu.py:109-115 (Q side), one thread per sample; zeroes T.
template <typename TA>
__global__ void finalize_q_kernel(const double* __restrict__ Q, TA* __restrict__ T,
                                  const double* __restrict__ qbat, int64_t N, int K,
                                  double* __restrict__ Q_EM) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double sc = 1.0 / fmax(qbat[i], 1e-5);
    double s = 0.0;
    for (int k = 0; k < K; ++k) {
        const double v = (Q[i * K + k] * (double)T[i * K + k]) * sc;
        Q_EM[i * K + k] = v;
        s += v;
        T[i * K + k] = TA(0);
    }
    s = fmax(s, 1e-5);
    for (int k = 0; k < K; ++k) Q_EM[i * K + k] /= s;
}

// em_adam_gpu.py:11-42 + :222 (P side), elementwise.
// With sched != NULL the step's lr and bias corrections come from device
// memory (sched = [t, lr, 1 - b1^t, 1 - b2^t], advanced by sched_tick_kernel)
// so that the step can be replayed from a captured graph.
__global__ void adam_p_kernel(double* __restrict__ P, const double* __restrict__ PE, double* __restrict__ m,
                              double* __restrict__ v, int64_t M, int K, double bc1, double bc2, double lr,
                              double b1, double b2, double eps, float* __restrict__ P_lo, int64_t ld_lo,
                              const double* __restrict__ sched) {
    if (sched) {
        lr = sched[1];
        bc1 = sched[2];
        bc2 = sched[3];
    }
    const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M * K) return;
    const double d = PE[x] - P[x];
    const double mm = m[x] * b1 + d * (1.0 - b1);
    const double vv = v[x] * b2 + (1.0 - b2) * d * d;
    m[x] = mm;
    v[x] = vv;
    double p = P[x] + lr * (mm / bc1) / (sqrt(vv / bc2) + eps);
    p = fmin(fmax(p, 1e-5), 1.0 - 1e-5);
    P[x] = p;
    if (P_lo) P_lo[(x / K) * ld_lo + (x % K)] = (float)p;
}

// em_adam_gpu.py:11-42 + :223-224 (Q side), one thread per sample.
__global__ void adam_q_kernel(double* __restrict__ Q, const double* __restrict__ QE, double* __restrict__ m,
                              double* __restrict__ v, int64_t N, int K, double bc1, double bc2, double lr,
                              double b1, double b2, double eps, float* __restrict__ Q_lo, int64_t ld_lo,
                              const double* __restrict__ sched) {
    if (sched) {
        lr = sched[1];
        bc1 = sched[2];
        bc2 = sched[3];
    }
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    double s = 0.0;
    for (int k = 0; k < K; ++k) {
        const int64_t x = i * K + k;
        const double d = QE[x] - Q[x];
        const double mm = m[x] * b1 + d * (1.0 - b1);
        const double vv = v[x] * b2 + (1.0 - b2) * d * d;
        m[x] = mm;
        v[x] = vv;
        double q = Q[x] + lr * (mm / bc1) / (sqrt(vv / bc2) + eps);
        q = fmin(fmax(q, 1e-5), 1.0 - 1e-5);
        Q[x] = q;
        s += q;
    }
    for (int k = 0; k < K; ++k) {
        const double q = Q[i * K + k] / s;
        Q[i * K + k] = q;
        if (Q_lo) Q_lo[i * ld_lo + k] = (float)q;
    }
}

// t += 1 and the bias corrections of step t (EMAdamOptimizer.step increments
// t before use, em_adam_gpu.py:201-210).
__global__ void sched_tick_kernel(double* __restrict__ sched, double b1, double b2) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const double t = sched[0] + 1.0;
        sched[0] = t;
        sched[2] = 1.0 - pow(b1, t);
        sched[3] = 1.0 - pow(b2, t);
    }
}

// T[e] += sum over the nparts slices of Tpart[p * nk + e] (nk = N K).
template <typename T>
__global__ void tpart_reduce_kernel(const T* __restrict__ Tpart, int64_t nparts, int64_t nk, T* __restrict__ Tm) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nk) return;
    T s0 = 0, s1 = 0, s2 = 0, s3 = 0;
    int64_t p = 0;
    for (; p + 4 <= nparts; p += 4) {
        s0 += Tpart[(p + 0) * nk + e];
        s1 += Tpart[(p + 1) * nk + e];
        s2 += Tpart[(p + 2) * nk + e];
        s3 += Tpart[(p + 3) * nk + e];
    }
    for (; p < nparts; ++p) s0 += Tpart[p * nk + e];
    Tm[e] += (s0 + s1) + (s2 + s3);
}

// ---------------------------------------------------------------------------
// Host-side dispatch
// ---------------------------------------------------------------------------
// Defaults: the fastest shapes measured at N=100,000 x M=500,000, K=10.
static EmTiles g_tiles[3] = {{8, 4, 8192}, {4, 4, 4096}, {8, 4, 8192}};  // ADX_F32, ADX_F64, ADX_BF16

EmTiles em_tiles(int dtype) { return g_tiles[dtype >= 0 && dtype < 3 ? dtype : 0]; }

struct EmGrid {
    int cw, waves;
    int64_t nbx, splits, n_per_split;
    const int* ll_fast = nullptr;   // ll_domain_kernel's flag (log-likelihood passes)
};

// K > 16 ("wide"): the A/B/T contractions run in 16-wide k blocks, one
// launch per block (H over the full K in each, the log-likelihood in the
// first only), at CW = 4.
static bool em_wide(int K) { return K > 16; }

// Waves per workgroup of a launch: the tiles' choice for the statistics
// pass with K <= 16; 4 for the wide form and for every pass that carries
// the log-likelihood (the fused FP64 statistics + log-likelihood kernel
// needs more than 256 registers per lane, i.e. one wave per SIMD).
static int em_waves(int dtype, int K, bool stats, bool ll) {
    return (stats && !ll && !em_wide(K)) ? em_tiles(dtype).waves : 4;
}

static EmGrid em_grid(int dtype, int64_t M, int64_t N, int K, int waves, int64_t target = 0) {
    const EmTiles tl = em_tiles(dtype);
    EmGrid gd;
    gd.cw = em_wide(K) ? 4 : tl.cw;
    gd.waves = waves;
    gd.nbx = ceil_div(M, 16 * gd.cw * gd.waves);
    if (target <= 0) target = tl.target_wgs;
    int64_t s = ceil_div(target, gd.nbx);
    const int64_t max_s = ceil_div(N, 64);
    if (s > max_s) s = max_s;
    if (s < 1) s = 1;
    if (s > 65535) s = 65535;
    gd.n_per_split = ceil_div(ceil_div(N, s), 64) * 64;
    gd.splits = ceil_div(N, gd.n_per_split);
    return gd;
}

template <typename T, int KS, bool STATS, bool LL>
static void launch_em(const EmGrid& gd, hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N,
                      int K, const T* P, int64_t ldp, const T* Q, int64_t ldq, T* A, T* B, T* Tm,
                      double* llp, T* Tpart, T* ABpart, int kb0) {
    dim3 grid((unsigned)gd.nbx, (unsigned)gd.splits);
#define ADX_EM_LAUNCH(cw, wv)                                                                                   \
    em_pass_kernel<T, KS, cw, wv, STATS, LL><<<grid, 64 * wv, 0, st>>>(G, ldg, M, N, K, P, ldp, Q, ldq,        \
                                                                       gd.n_per_split, A, B, Tm, llp, Tpart, \
                                                                       ABpart, kb0, gd.ll_fast)
    // The FP64 and wide forms have a single tile width; 8-wave workgroups
    // exist for the statistics-only pass with K <= 16 (em_waves).
    if constexpr (STATS && !LL && KS <= 4) {
        if (gd.waves == 8) {
            if constexpr (sizeof(T) == 4) {
                if (gd.cw == 8) {
                    ADX_EM_LAUNCH(8, 8);
                    return;
                }
            }
            ADX_EM_LAUNCH(4, 8);
            return;
        }
    }
    if constexpr (sizeof(T) == 4 && KS <= 4) {
        if (gd.cw == 8) {
            ADX_EM_LAUNCH(8, 4);
            return;
        }
    }
    ADX_EM_LAUNCH(4, 4);
#undef ADX_EM_LAUNCH
}

// One launch per 16-wide k block (one for K <= 16); the log-likelihood
// rides on the first.
template <typename T, int KS, bool STATS, bool LL>
static void launch_em_blocks(const EmGrid& gd, hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M,
                             int64_t N, int K, const T* P, int64_t ldp, const T* Q, int64_t ldq, T* A, T* B,
                             T* Tm, double* llp, T* Tpart, T* ABpart) {
    launch_em<T, KS, STATS, LL>(gd, st, G, ldg, M, N, K, P, ldp, Q, ldq, A, B, Tm, llp, Tpart, ABpart, 0);
    if constexpr (STATS)
        for (int kb0 = 16; kb0 < K; kb0 += 16)
            launch_em<T, KS, true, false>(gd, st, G, ldg, M, N, K, P, ldp, Q, ldq, A, B, Tm, nullptr, Tpart,
                                          ABpart, kb0);
}

template <typename T, bool STATS, bool LL>
static void launch_em_k(const EmGrid& gd, hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M,
                        int64_t N, int K, const T* P, int64_t ldp, const T* Q, int64_t ldq, T* A, T* B,
                        T* Tm, double* llp, T* Tpart) {
    // Workspace after the T slices: A slices then B slices, one per sample split.
    T* ABpart = Tpart ? Tpart + gd.nbx * N * (int64_t)K : nullptr;
#define ADX_EM_KS(ks) \
    launch_em_blocks<T, ks, STATS, LL>(gd, st, G, ldg, M, N, K, P, ldp, Q, ldq, A, B, Tm, llp, Tpart, ABpart)
    const int ks = (K + 3) / 4;
    if (ks <= 1) ADX_EM_KS(1);
    else if (ks == 2) ADX_EM_KS(2);
    else if (ks == 3) ADX_EM_KS(3);
    else if (ks == 4) ADX_EM_KS(4);
    else if (ks <= 8) ADX_EM_KS(8);      // H steps past K multiply zeros
    else if (ks <= 12) ADX_EM_KS(12);
    else ADX_EM_KS(16);
#undef ADX_EM_KS
    if (STATS && Tpart) {
        const int64_t nk = N * K, mk = M * K;
        tpart_reduce_kernel<T><<<(unsigned)ceil_div(nk, 256), 256, 0, st>>>(Tpart, gd.nbx, nk, Tm);
        tpart_reduce_kernel<T><<<(unsigned)ceil_div(mk, 256), 256, 0, st>>>(ABpart, gd.splits, mk, A);
        tpart_reduce_kernel<T><<<(unsigned)ceil_div(mk, 256), 256, 0, st>>>(ABpart + gd.splits * mk, gd.splits, mk,
                                                                           B);
    }
}

// A pass over SNP bands.  Band b covers SNPs b*band .. min((b+1)*band, M) - 1
// (band a multiple of the 16*cw*waves SNPs of one workgroup column) and is one
// launch sequence over the sub-matrix of its byte rows; its T / A / B
// partial slices live in one region of the workspace reused band after band,
// and its log-likelihood partials are summed into one slot per band.  The
// band slots are summed in band order at the end, so for a given plan the
// pass is deterministic.  Workspace layout:
//   [LL partials of one band][LL band slots][domain flag][T slices | A slices | B slices]
struct EmPlan {
    int waves;             // waves per workgroup of every launch (em_waves)
    int64_t band, nbands, target;
    int64_t ll_slots;      // LL partials of the largest band (its nbx x splits)
    int64_t ll_bytes;      // both LL regions and the domain flag, 256-byte padded
    int64_t slice_bytes;   // largest band's partial slices (0 without STATS)
};

static int64_t pad256(int64_t b) { return ceil_div(b, 256) * 256; }

static EmPlan em_plan_for(int dtype, int64_t M, int64_t N, int K, bool stats, bool ll, int64_t band,
                          int64_t target) {
    EmPlan p;
    p.waves = em_waves(dtype, K, stats, ll);
    p.band = band < M ? band : M;
    p.nbands = ceil_div(M, p.band);
    p.target = target;
    const int64_t last = M - (p.nbands - 1) * p.band;
    const EmGrid g1 = em_grid(dtype, p.band, N, K, p.waves, target);
    const EmGrid g2 = em_grid(dtype, last, N, K, p.waves, target);
    const int64_t s1 = g1.nbx * g1.splits, s2 = g2.nbx * g2.splits;
    p.ll_slots = s1 > s2 ? s1 : s2;
    p.ll_bytes = pad256(8 * p.ll_slots) + pad256(8 * p.nbands) + 256;
    const int64_t esz = dtype == ADX_F64 ? 8 : 4;
    const int64_t b1 = esz * K * (g1.nbx * N + 2 * g1.splits * p.band);
    const int64_t b2 = esz * K * (g2.nbx * N + 2 * g2.splits * last);
    p.slice_bytes = stats ? pad256(b1 > b2 ? b1 : b2) : 0;
    return p;
}

static int64_t em_plan_bytes(const EmPlan& p) { return p.ll_bytes + p.slice_bytes; }

// The plan for a workspace of ws bytes: one band if it fits, else the widest
// band (then the largest workgroup target, halved down to 1: fewer sample
// splits, fewer A / B slices) that fits.
// False if none does.
// max_band > 0 caps the band (SNPs; the host-streamed pass's staging).
static bool em_plan(int dtype, int64_t M, int64_t N, int K, bool stats, bool ll, int64_t ws, EmPlan& out,
                    int64_t max_band = 0) {
    const EmTiles tl = em_tiles(dtype);
    const int64_t unit = 16 * (em_wide(K) ? 4 : tl.cw) * em_waves(dtype, K, stats, ll);
    out = em_plan_for(dtype, M, N, K, stats, ll, M, tl.target_wgs);
    if (em_plan_bytes(out) <= ws && (max_band <= 0 || M <= max_band)) return true;
    int64_t cmax = ceil_div(M, unit);
    if (max_band > 0 && max_band / unit < cmax) cmax = max_band / unit;
    if (cmax < 1) return false;
    for (int64_t target = tl.target_wgs; target >= 1; target /= 2) {
        // the slices grow with the band: the widest fitting band by bisection
        int64_t lo = 0, hi = cmax + 1;   // fits(lo) (or lo = 0), !fits(hi) (or hi = cmax + 1)
        while (hi - lo > 1) {
            const int64_t c = (lo + hi) / 2;
            if (em_plan_bytes(em_plan_for(dtype, M, N, K, stats, ll, c * unit, target)) <= ws)
                lo = c;
            else
                hi = c;
        }
        if (lo >= 1) {
            out = em_plan_for(dtype, M, N, K, stats, ll, lo * unit, target);
            if (em_plan_bytes(out) <= ws) return true;
        }
    }
    return false;
}

// The pass over one band (sub-matrix of byte rows m0 / 4 ..): G, P, A, B
// offset to the band, Q and T whole.
template <typename T, bool STATS, bool LL>
static void launch_band(const EmGrid& gd, hipStream_t st, const uint8_t* G, int64_t ldg, int64_t m0, int64_t Mb,
                        int64_t N, int K, const void* P, int64_t ldp, const void* Q, int64_t ldq, void* A, void* B,
                        void* Tm, double* llp, void* tpart) {
    const T* Pb = static_cast<const T*>(P) + m0 * ldp;
    T* Ab = STATS ? static_cast<T*>(A) + m0 * K : nullptr;
    T* Bb = STATS ? static_cast<T*>(B) + m0 * K : nullptr;
    launch_em_k<T, STATS, LL>(gd, st, G + (m0 >> 2) * ldg, ldg, Mb, N, K, Pb, ldp, static_cast<const T*>(Q), ldq,
                              Ab, Bb, STATS ? static_cast<T*>(Tm) : nullptr, llp, static_cast<T*>(tpart));
}

static void launch_any(int dtype, bool stats, bool ll, const EmGrid& gd, hipStream_t st, const uint8_t* G,
                       int64_t ldg, int64_t m0, int64_t Mb, int64_t N, int K, const void* P, int64_t ldp,
                       const void* Q, int64_t ldq, void* A, void* B, void* Tm, double* llp, void* tpart) {
    if (dtype == ADX_F32)
        launch_band<float, true, false>(gd, st, G, ldg, m0, Mb, N, K, P, ldp, Q, ldq, A, B, Tm, nullptr, tpart);
    else if (stats && ll)
        launch_band<double, true, true>(gd, st, G, ldg, m0, Mb, N, K, P, ldp, Q, ldq, A, B, Tm, llp, tpart);
    else if (stats)
        launch_band<double, true, false>(gd, st, G, ldg, m0, Mb, N, K, P, ldp, Q, ldq, A, B, Tm, nullptr, tpart);
    else
        launch_band<double, false, true>(gd, st, G, ldg, m0, Mb, N, K, P, ldp, Q, ldq, A, B, Tm, llp, nullptr);
}

// Enqueue the log-likelihood domain check of (P, Q) (FP64) into the plan's
// flag slot; returns the flag's device address for the passes that follow.
static const int* ll_domain_check(hipStream_t st, const EmPlan& plan, uint8_t* ws, const void* P, int64_t M,
                                  int64_t ldp, const void* Q, int64_t N, int64_t ldq, int K) {
    int* flag = reinterpret_cast<int*>(ws + pad256(8 * plan.ll_slots) + pad256(8 * plan.nbands));
    ll_flag_init_kernel<<<1, 64, 0, st>>>(flag);
    const int64_t blocks = ceil_div(M + N, 256);
    ll_domain_kernel<<<(unsigned)(blocks < 4096 ? blocks : 4096), 256, 0, st>>>(
        static_cast<const double*>(P), M, ldp, static_cast<const double*>(Q), N, ldq, K, flag);
    return flag;
}

// Argument checks of adx_em_pass (-1 if the call can go ahead; 0 if there is
// nothing to do; else the failure status).
static int em_pass_check(const char* name, int dtype, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K,
                         int64_t ldp, int64_t ldq, const void* A, const void* B, const void* T, int flags,
                         const double* ll_out) {
    const std::string who(name);
    if (M <= 0 || N <= 0) return fail_msg(1, who + ": empty genotype matrix");
    if (K < 1 || K > 64) return fail_msg(1, who + ": K must be in [1, 64]");
    if (ldg % 16 != 0 || ldg < ceil_div(N, 64) * 64)
        return fail_msg(1, who + ": ldg must be a multiple of 16 and >= N rounded up to 64");
    if (ldp < K || ldq < K) return fail_msg(1, who + ": ldp/ldq < K");
    if (reinterpret_cast<uintptr_t>(G) % 16 != 0) return fail_msg(1, who + ": G not 16-byte aligned");
    if (dtype != ADX_F32 && dtype != ADX_F64) return fail_msg(1, who + ": dtype must be ADX_F32 or ADX_F64");
    const bool stats = flags & ADX_EM_STATS;
    const bool ll = flags & ADX_EM_LOGLIK;
    if (!stats && !ll) return 0;
    if (ll && dtype != ADX_F64) return fail_msg(1, who + ": ADX_EM_LOGLIK needs ADX_F64");
    if (stats && (!A || !B || !T)) return fail_msg(1, who + ": A, B, T required");
    if (ll && !ll_out) return fail_msg(1, who + ": ADX_EM_LOGLIK needs ll_out");
    return -1;
}

// The pass with a caller-given workspace (see adx_em_pass).  Without a plan
// that fits, the statistics fall back to atomics if allow_atomics, else fail.
static int em_pass_run(const char* name, int dtype, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K,
                       const void* P, int64_t ldp, const void* Q, int64_t ldq, void* A, void* B, void* T, int flags,
                       double* ll_out, void* workspace, int64_t workspace_bytes, bool allow_atomics, void* stream) {
    const int bad = em_pass_check(name, dtype, G, ldg, M, N, K, ldp, ldq, A, B, T, flags, ll_out);
    if (bad >= 0) return bad;
    const std::string who(name);
    const bool stats = flags & ADX_EM_STATS;
    const bool ll = flags & ADX_EM_LOGLIK;
    hipStream_t st = (hipStream_t)stream;
    EmPlan plan;
    const bool planned = workspace && em_plan(dtype, M, N, K, stats, ll, workspace_bytes, plan);
    if (!planned) {
        if (ll || !allow_atomics)
            return fail_msg(1, who + ": ADX_EM_LOGLIK needs a workspace (adx_em_workspace_bytes, or enough "
                                     "for one SNP band)");
        // no workspace: T, A, B by device-scope atomics (order varies run to run)
        launch_any(dtype, true, false, em_grid(dtype, M, N, K, em_waves(dtype, K, true, false)), st, G, ldg, 0, M,
                   N, K, P, ldp, Q, ldq, A, B, T, nullptr, nullptr);
        ADX_LAUNCH_CHECK(name);
        return 0;
    }
    uint8_t* ws = static_cast<uint8_t*>(workspace);
    double* llp = reinterpret_cast<double*>(ws);
    double* slots = reinterpret_cast<double*>(ws + pad256(8 * plan.ll_slots));
    void* tpart = stats ? ws + plan.ll_bytes : nullptr;
    // only the log-likelihood-only kernel reads the flag (em_pass_kernel)
    const int* ll_fast = (ll && !stats) ? ll_domain_check(st, plan, ws, P, M, ldp, Q, N, ldq, K) : nullptr;
    for (int64_t b = 0; b < plan.nbands; ++b) {
        const int64_t m0 = b * plan.band;
        const int64_t Mb = M - m0 < plan.band ? M - m0 : plan.band;
        EmGrid gd = em_grid(dtype, Mb, N, K, plan.waves, plan.target);
        gd.ll_fast = ll_fast;
        launch_any(dtype, stats, ll, gd, st, G, ldg, m0, Mb, N, K, P, ldp, Q, ldq, A, B, T, llp, tpart);
        if (ll) reduce_f64_kernel<<<1, 256, 0, st>>>(llp, gd.nbx * gd.splits, slots + b);
        ADX_LAUNCH_CHECK(name);
    }
    if (ll) {
        reduce_f64_kernel<<<1, 256, 0, st>>>(slots, plan.nbands, ll_out);
        ADX_LAUNCH_CHECK(name);
    }
    return 0;
}

// Fixed workspace of adx_em_pass_noalloc: one copy of this array per GPU,
// held from the library's load onto that GPU until the process ends.  In host
// code its name is only a stand-in; the device address comes from
// hipGetSymbolAddress, once per device (with that device current), and only
// that address is handed to kernels.  Every call on a GPU shares the GPU's
// copy; nothing in it needs clearing (every partial and band slot a call
// reads was written earlier in the same call), and calls that overlap on
// one GPU race.
static __device__ __align__(256) uint8_t scratch[ADX_EM_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static int noalloc_scratch(void** out) {
    int dev = -1;
    ADX_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail_msg(1, "adx_em_pass_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return fail(e, "adx_em_pass_noalloc: hipGetSymbolAddress(scratch)");
        if (!ptr) return fail_msg(1, "adx_em_pass_noalloc: no device address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_set_em_tiles(int dtype, int cw, int waves, int64_t target_wgs) {
    if (dtype < 0 || dtype > 2) return fail_msg(1, "adx_set_em_tiles: unknown dtype tag");
    const bool ok_cw = dtype == ADX_F64 ? cw == 4 : (cw == 4 || cw == 8);
    if (!ok_cw) return fail_msg(1, "adx_set_em_tiles: cw must be 4 or 8 (4 for ADX_F64)");
    if (waves != 4 && waves != 8) return fail_msg(1, "adx_set_em_tiles: waves must be 4 or 8");
    if (target_wgs < 64 || target_wgs > 65536) return fail_msg(1, "adx_set_em_tiles: target_wgs not in [64, 65536]");
    g_tiles[dtype] = EmTiles{cw, waves, target_wgs};
    return 0;
}

int adx_get_em_tiles(int dtype, int* cw, int* waves, int64_t* target_wgs) {
    if (dtype < 0 || dtype > 2) return fail_msg(1, "adx_get_em_tiles: unknown dtype tag");
    if (cw) *cw = g_tiles[dtype].cw;
    if (waves) *waves = g_tiles[dtype].waves;
    if (target_wgs) *target_wgs = g_tiles[dtype].target_wgs;
    return 0;
}

int64_t adx_em_workspace_bytes(int dtype, int64_t M, int64_t N, int K) {
    if (M <= 0 || N <= 0) return 0;
    // K = 0: the log-likelihood partials only.  With statistics, enough for
    // one band both without and with the log-likelihood (their workgroup
    // shapes may differ).
    const int64_t tg = em_tiles(dtype).target_wgs;
    const int64_t b0 = em_plan_bytes(em_plan_for(dtype, M, N, K > 0 ? K : 1, K > 0, false, M, tg));
    const int64_t b1 = em_plan_bytes(em_plan_for(dtype, M, N, K > 0 ? K : 1, K > 0, true, M, tg));
    return b0 > b1 ? b0 : b1;
}

int adx_em_pass(int dtype, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K, const void* P,
                int64_t ldp, const void* Q, int64_t ldq, void* A, void* B, void* T, int flags,
                double* ll_out, void* workspace, int64_t workspace_bytes, void* stream) {
    return em_pass_run("adx_em_pass", dtype, G, ldg, M, N, K, P, ldp, Q, ldq, A, B, T, flags, ll_out, workspace,
                       workspace_bytes, true, stream);
}

int64_t adx_em_band_snps(int dtype, int64_t M, int64_t N, int K, int flags, int64_t workspace_bytes,
                          int64_t max_band) {
    if (M <= 0 || N <= 0 || K < 1 || K > 64 || (dtype != ADX_F32 && dtype != ADX_F64)) return -1;
    EmPlan plan;
    if (!em_plan(dtype, M, N, K, (flags & ADX_EM_STATS) != 0, (flags & ADX_EM_LOGLIK) != 0, workspace_bytes, plan,
                 max_band))
        return -1;
    return plan.band;
}

int adx_em_pass_host(int dtype, const uint8_t* G_host, int64_t ldg, int64_t M, int64_t N, int K, const void* P,
                     int64_t ldp, const void* Q, int64_t ldq, void* A, void* B, void* T, int flags, double* ll_out,
                     void* workspace, int64_t workspace_bytes, uint8_t* staging, int64_t staging_bytes, void* stream,
                     void* copy_stream) {
    const char* name = "adx_em_pass_host";
    if (!G_host || !staging) return fail_msg(1, "adx_em_pass_host: G_host and staging required");
    if (reinterpret_cast<uintptr_t>(staging) % 256 != 0) return fail_msg(1, "adx_em_pass_host: staging not 256-byte aligned");
    const int bad = em_pass_check(name, dtype, staging, ldg, M, N, K, ldp, ldq, A, B, T, flags, ll_out);
    if (bad >= 0) return bad;
    const bool stats = flags & ADX_EM_STATS;
    const bool ll = flags & ADX_EM_LOGLIK;
    // two staging buffers of whole byte rows, each a multiple of 256 bytes
    const int64_t half = (staging_bytes / 2) / 256 * 256;
    const int64_t max_band = (half / ldg) * 4;
    EmPlan plan;
    if (!workspace || !em_plan(dtype, M, N, K, stats, ll, workspace_bytes, plan, max_band))
        return fail_msg(1, "adx_em_pass_host: the workspace and staging do not fit one SNP band "
                           "(adx_em_band_snps gives the band a workspace allows)");
    hipStream_t st = (hipStream_t)stream;
    hipStream_t cp = copy_stream ? (hipStream_t)copy_stream : st;
    uint8_t* buf[2] = {staging, staging + half};
    hipEvent_t copied[2] = {nullptr, nullptr}, used[2] = {nullptr, nullptr};
    int rc = 0;
    for (int i = 0; i < 2 && rc == 0; ++i) {
        hipError_t e = hipEventCreateWithFlags(&copied[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&used[i], hipEventDisableTiming);
        if (e != hipSuccess) rc = fail(e, "adx_em_pass_host: hipEventCreateWithFlags");
    }
    // the copies start after the work already enqueued on `stream` (which may
    // still read the staging buffer or write P, Q)
    if (rc == 0 && cp != st) {
        hipEvent_t ready = nullptr;
        hipError_t e = hipEventCreateWithFlags(&ready, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(ready, st);
        if (e == hipSuccess) e = hipStreamWaitEvent(cp, ready, 0);
        if (ready) hipEventDestroy(ready);
        if (e != hipSuccess) rc = fail(e, "adx_em_pass_host: stream fork");
    }
    uint8_t* ws = static_cast<uint8_t*>(workspace);
    double* llp = reinterpret_cast<double*>(ws);
    double* slots = reinterpret_cast<double*>(ws + pad256(8 * plan.ll_slots));
    void* tpart = stats ? ws + plan.ll_bytes : nullptr;
    const int* ll_fast =
        (rc == 0 && ll && !stats) ? ll_domain_check(st, plan, ws, P, M, ldp, Q, N, ldq, K) : nullptr;
    for (int64_t b = 0; b < plan.nbands && rc == 0; ++b) {
        const int64_t m0 = b * plan.band;
        const int64_t Mb = M - m0 < plan.band ? M - m0 : plan.band;
        const int64_t rows = (Mb + 3) / 4;
        const int i = (int)(b & 1);
        // the copy into buffer i waits for the band that last read it
        hipError_t e = hipSuccess;
        if (b >= 2) e = hipStreamWaitEvent(cp, used[i], 0);
        if (e == hipSuccess)
            e = hipMemcpyAsync(buf[i], G_host + (m0 >> 2) * ldg, (size_t)(rows * ldg), hipMemcpyHostToDevice, cp);
        if (e == hipSuccess) e = hipEventRecord(copied[i], cp);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, copied[i], 0);
        if (e != hipSuccess) {
            rc = fail(e, "adx_em_pass_host: staging copy");
            break;
        }
        EmGrid gd = em_grid(dtype, Mb, N, K, plan.waves, plan.target);
        gd.ll_fast = ll_fast;
        // the band's rows sit at the start of buffer i: offset G by -m0 / 4 rows
        // inside launch_band's arithmetic, i.e. pass m0 for P, A, B only
        const void* Pb = static_cast<const uint8_t*>(P) + m0 * ldp * (dtype == ADX_F64 ? 8 : 4);
        void* Ab = stats ? static_cast<uint8_t*>(A) + m0 * K * (dtype == ADX_F64 ? 8 : 4) : nullptr;
        void* Bb = stats ? static_cast<uint8_t*>(B) + m0 * K * (dtype == ADX_F64 ? 8 : 4) : nullptr;
        launch_any(dtype, stats, ll, gd, st, buf[i], ldg, 0, Mb, N, K, Pb, ldp, Q, ldq, Ab, Bb, T, llp, tpart);
        if (ll) reduce_f64_kernel<<<1, 256, 0, st>>>(llp, gd.nbx * gd.splits, slots + b);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipEventRecord(used[i], st);
        if (e != hipSuccess) rc = fail(e, name);
    }
    if (rc == 0 && ll) {
        reduce_f64_kernel<<<1, 256, 0, st>>>(slots, plan.nbands, ll_out);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = fail(e, name);
    }
    // the copy stream's last use of the buffers is ordered before later work on st
    if (rc == 0 && cp != st) {
        hipEvent_t done = nullptr;
        hipError_t e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(done, cp);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, done, 0);
        if (done) hipEventDestroy(done);
        if (e != hipSuccess) rc = fail(e, "adx_em_pass_host: stream join");
    }
    for (int k = 0; k < 2; ++k) {
        if (copied[k]) hipEventDestroy(copied[k]);
        if (used[k]) hipEventDestroy(used[k]);
    }
    return rc;
}

int adx_em_pass_noalloc(int dtype, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K, const void* P,
                        int64_t ldp, const void* Q, int64_t ldq, void* A, void* B, void* T, int flags,
                        double* ll_out, void* stream) {
    // shape checks and the band plan first: a shape beyond the fixed scratch
    // is refused before any device call
    const int bad = em_pass_check("adx_em_pass_noalloc", dtype, G, ldg, M, N, K, ldp, ldq, A, B, T, flags, ll_out);
    if (bad >= 0) return bad;
    EmPlan plan;
    if (!em_plan(dtype, M, N, K, (flags & ADX_EM_STATS) != 0, (flags & ADX_EM_LOGLIK) != 0,
                 ADX_EM_NOALLOC_SCRATCH_BYTES, plan))
        return fail_msg(1, "adx_em_pass_noalloc: this shape needs more than the fixed scratch of "
                           "ADX_EM_NOALLOC_SCRATCH_BYTES for one SNP band (use adx_em_pass with a workspace)");
    void* ws = nullptr;
    const int rc = noalloc_scratch(&ws);
    if (rc != 0) return rc;
    return em_pass_run("adx_em_pass_noalloc", dtype, G, ldg, M, N, K, P, ldp, Q, ldq, A, B, T, flags, ll_out, ws,
                       ADX_EM_NOALLOC_SCRATCH_BYTES, false, stream);
}

int adx_em_finalize(int acc_dtype, const double* P, const double* Q, void* A, void* B, void* T,
                    const double* qbat, int64_t M, int64_t N, int K, double* P_EM, double* Q_EM, void* stream) {
    if (K < 1) return fail_msg(1, "adx_em_finalize: K < 1");
    hipStream_t st = (hipStream_t)stream;
    const int64_t n = M * K;
    if (acc_dtype == ADX_F32) {
        if (n > 0) finalize_p_kernel<float><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(P, (float*)A, (float*)B, n, P_EM);
        if (N > 0) finalize_q_kernel<float><<<(unsigned)ceil_div(N, 256), 256, 0, st>>>(Q, (float*)T, qbat, N, K, Q_EM);
    } else if (acc_dtype == ADX_F64) {
        if (n > 0) finalize_p_kernel<double><<<(unsigned)ceil_div(n, 256), 256, 0, st>>>(P, (double*)A, (double*)B, n, P_EM);
        if (N > 0) finalize_q_kernel<double><<<(unsigned)ceil_div(N, 256), 256, 0, st>>>(Q, (double*)T, qbat, N, K, Q_EM);
    } else {
        return fail_msg(1, "adx_em_finalize: acc_dtype must be ADX_F32 or ADX_F64");
    }
    ADX_LAUNCH_CHECK("adx_em_finalize");
    return 0;
}

int adx_adam_step(double* P, double* Q, const double* P_EM, const double* Q_EM, double* mP, double* vP,
                  double* mQ, double* vQ, int64_t M, int64_t N, int K, double t, double lr, double beta1,
                  double beta2, double eps, float* P_lo, float* Q_lo, int64_t ld_lo, void* stream) {
    if (K < 1) return fail_msg(1, "adx_adam_step: K < 1");
    if ((P_lo || Q_lo) && ld_lo < K) return fail_msg(1, "adx_adam_step: ld_lo < K");
    const double bc1 = 1.0 - pow(beta1, t);
    const double bc2 = 1.0 - pow(beta2, t);
    hipStream_t st = (hipStream_t)stream;
    if (M > 0)
        adam_p_kernel<<<(unsigned)ceil_div(M * K, 256), 256, 0, st>>>(P, P_EM, mP, vP, M, K, bc1, bc2, lr, beta1,
                                                                     beta2, eps, P_lo, ld_lo, nullptr);
    if (N > 0)
        adam_q_kernel<<<(unsigned)ceil_div(N, 256), 256, 0, st>>>(Q, Q_EM, mQ, vQ, N, K, bc1, bc2, lr, beta1, beta2,
                                                                  eps, Q_lo, ld_lo, nullptr);
    ADX_LAUNCH_CHECK("adx_adam_step");
    return 0;
}

int adx_adam_step_sched(double* P, double* Q, const double* P_EM, const double* Q_EM, double* mP, double* vP,
                        double* mQ, double* vQ, int64_t M, int64_t N, int K, double* sched, double beta1,
                        double beta2, double eps, float* P_lo, float* Q_lo, int64_t ld_lo, void* stream) {
    if (K < 1) return fail_msg(1, "adx_adam_step_sched: K < 1");
    if (!sched) return fail_msg(1, "adx_adam_step_sched: sched is NULL");
    if ((P_lo || Q_lo) && ld_lo < K) return fail_msg(1, "adx_adam_step_sched: ld_lo < K");
    hipStream_t st = (hipStream_t)stream;
    sched_tick_kernel<<<1, 64, 0, st>>>(sched, beta1, beta2);
    if (M > 0)
        adam_p_kernel<<<(unsigned)ceil_div(M * K, 256), 256, 0, st>>>(P, P_EM, mP, vP, M, K, 0.0, 0.0, 0.0, beta1,
                                                                     beta2, eps, P_lo, ld_lo, sched);
    if (N > 0)
        adam_q_kernel<<<(unsigned)ceil_div(N, 256), 256, 0, st>>>(Q, Q_EM, mQ, vQ, N, K, 0.0, 0.0, 0.0, beta1, beta2,
                                                                  eps, Q_lo, ld_lo, sched);
    ADX_LAUNCH_CHECK("adx_adam_step_sched");
    return 0;
}

}  // extern "C"
