// Gradient and Hessian blocks of the SQP block relaxation (BR-QN), FP64.
//
// Reference: adamixture/model/br_qn_gpu.py:67-214 (_grad_hess_q_chunk,
// _grad_hess_p_chunk, compute_grad_hess_{Q,P}_gpu) and the Cython twin
// sqp.pyx:258-328.  With H = P Q^T clamped to [1e-10, 1-1e-10] and, on
// observed genotypes, z1 = g/H, z2 = (2-g)/(1-H), w1 = z1/H, w2 = z2/(1-H):
//   Q side (per sample i):  Xtz_q = sum_j z1 p_j + z2 (1-p_j)
//                           XtX_q = sum_j w1 p_j p_j^T + w2 (1-p_j)(1-p_j)^T
//   P side (per SNP j):     Xtz_p = sum_i (z1 - z2) q_i
//                           XtX_p = sum_i (w1 + w2) q_i q_i^T
// The reference forms these with einsum over C x N x K x K temporaries.  Here
// the Q side uses (1-p)(1-p)^T = 1 1^T - 1 p^T - p 1^T + p p^T:
//   XtX_q[k][l] = sum_j (w1 + w2) p_k p_l + S_w2 - T_k - T_l,
//   Xtz_q[k]    = sum_j (z1 - z2) p_k + S_z2,
// with S_w2 = sum_j w2, T_k = sum_j w2 p_k, S_z2 = sum_j z2: one FMA per
// Hessian term and genotype instead of two (the corrections are formed once
// per problem and reduction split; the rounding of the subtractions is
// relative to S_w2, the scale of the Hessian entries).  Each thread owns one
// problem's accumulators in registers and streams the 2-bit genotypes once:
//   Q side: thread = sample (K <= 10) or two waves per 64 samples, each with
//           half the Hessian rows (K = 11..16); workgroup = 256 or 128
//           consecutive samples x a SNP range; one coalesced byte per 4
//           SNPs, P rows are wave-uniform.
//   P side: thread = SNP, workgroup = 256 SNPs; 256-sample chunks of the
//           genotype rows and of Q are staged in LDS (Q reads broadcast).
// Sample splits / SNP splits are combined with FP64 atomics.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

struct Weights {
    double z1, z2, w1, w2;
};

// 1/x for x in [1e-10, 1]: the hardware reciprocal estimate refined by
// two Newton-Raphson steps r <- r + r (1 - x r) (the error squares at each
// step), within a few ulp of the correctly rounded quotient.
__device__ __forceinline__ double recip_nr(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// Two reciprocals (1/H, 1/(1-H)) instead of the reference's four
// quotients; each weight then differs from its quotient by a few ulp.
__device__ __forceinline__ Weights bq_weights(double h, uint32_t code) {
    Weights w{0.0, 0.0, 0.0, 0.0};
    if (code == 3u) return w;
    const double hc = fmax(fmin(h, 1.0 - 1e-10), 1e-10);
    const double g = (double)code;
    const double rh = recip_nr(hc), ru = recip_nr(1.0 - hc);
    w.z1 = g * rh;
    w.z2 = (2.0 - g) * ru;
    w.w1 = w.z1 * rh;
    w.w2 = w.z2 * ru;
    return w;
}

// Q side, Hessian rows K0 <= k < K1 of the upper triangle (pairs k <= l <
// KM, row by row in hx) for sample i over the byte rows r0 .. r1 - 1; with
// OWN also the gradient sum hz, T (tw) and the sums S_z2, S_w2.  The P rows
// of 128 SNPs at a time are staged in LDS by the whole workgroup (sP: 128
// rows of KM doubles, zero past K and past the split) and read back by every
// lane alike (LDS broadcast); the genotype byte of the next row is loaded one
// row ahead.
constexpr int GHQ_CH = 32;   // byte rows (x 4 SNPs) per staged chunk

template <int KM, int K0, int K1, bool OWN>
__device__ __forceinline__ void gh_q_loop(const uint8_t* __restrict__ G, int64_t ldg, int64_t M, int K,
                                          const double* __restrict__ P, double* sP, const double (&q)[KM],
                                          int64_t i, bool live, int64_t r0, int64_t r1, double* hx, double* hz,
                                          double* tw, double& sz2, double& sw2) {
    for (int64_t c0 = r0; c0 < r1; c0 += GHQ_CH) {
        const int nr = (int)(r1 - c0 < GHQ_CH ? r1 - c0 : GHQ_CH);
        __syncthreads();   // the previous chunk's rows are read by every wave
        for (int e = threadIdx.x; e < 4 * GHQ_CH * KM; e += 256) {
            const int sl = e / KM, k = e - sl * KM;
            const int64_t m = 4 * c0 + sl;
            sP[e] = (sl < 4 * nr && m < M && k < K) ? P[m * K + k] : 0.0;
        }
        __syncthreads();
        uint32_t bn = live ? G[c0 * ldg + i] : 0xFFu;
        for (int rr = 0; rr < nr; ++rr) {
            const uint32_t byte = bn;
            if (rr + 1 < nr) bn = live ? G[(c0 + rr + 1) * ldg + i] : 0xFFu;
#pragma unroll 1
            for (int f = 0; f < 4; ++f) {
                // SNPs past M read as missing (code 3) and have zero P rows
                const uint32_t code = 4 * (c0 + rr) + f < M ? (byte >> (2 * f)) & 3u : 3u;
                const double* pr = sP + (4 * rr + f) * KM;
                double pc[KM];
#pragma unroll
                for (int k = 0; k < KM; ++k) pc[k] = pr[k];
                double h = 0.0;
#pragma unroll
                for (int k = 0; k < KM; ++k) h += q[k] * pc[k];
                const Weights w = bq_weights(h, code);
                const double ws = w.w1 + w.w2;
                if (OWN) {
                    const double zd = w.z1 - w.z2;
                    sz2 += w.z2;
                    sw2 += w.w2;
#pragma unroll
                    for (int k = 0; k < KM; ++k) {
                        hz[k] += zd * pc[k];
                        tw[k] += w.w2 * pc[k];
                    }
                }
                int t = 0;
#pragma unroll
                for (int k = K0; k < K1; ++k) {
                    const double a = ws * pc[k];
#pragma unroll
                    for (int l = k; l < KM; ++l, ++t) hx[t] += a * pc[l];
                }
            }
        }
    }
}

// Adds rows K0 .. K1 - 1 of sample i's Hessian (with the rank-one
// corrections) and, with OWN, its gradient to the outputs.
template <int KM, int K0, int K1, bool OWN>
__device__ __forceinline__ void gh_q_store(int K, int64_t i, const double* hx, const double* hz, const double* tw,
                                           double sz2, double sw2, double* __restrict__ XtX,
                                           double* __restrict__ Xtz) {
    if (OWN) {
#pragma unroll
        for (int k = 0; k < KM; ++k)
            if (k < K) atomicAdd(&Xtz[i * K + k], hz[k] + sz2);
    }
    int t = 0;
#pragma unroll
    for (int k = K0; k < K1; ++k)
#pragma unroll
        for (int l = k; l < KM; ++l, ++t)
            if (k < K && l < K) {
                const double v = ((hx[t] + sw2) - tw[k]) - tw[l];
                atomicAdd(&XtX[(i * K + k) * K + l], v);
                if (l != k) atomicAdd(&XtX[(i * K + l) * K + k], v);
            }
}

// Hessian pairs of rows K0 .. K1 - 1 of a KM x KM upper triangle.
constexpr int tri_pairs(int KM, int K0, int K1) { return K0 >= K1 ? 0 : (KM - K0) + tri_pairs(KM, K0 + 1, K1); }

// KS = 0: one thread per sample, 256 samples per workgroup.  KS > 0: two
// waves per 64 samples, 128 samples per workgroup; the even wave owns rows
// 0 .. KS - 1, the gradient, T and the sums, the odd wave rows KS .. KM - 1,
// and reads T and S_w2 of its samples from LDS after the reduction (the two
// roles have about equal work per genotype, with half the registers of one
// thread owning the whole triangle).
template <int KM, int KS>
__global__ __launch_bounds__(256) void grad_hess_q_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                          int64_t N, int K, const double* __restrict__ Q,
                                                          const double* __restrict__ P, int64_t rows_per_split,
                                                          double* __restrict__ XtX, double* __restrict__ Xtz) {
    constexpr int PA = tri_pairs(KM, 0, KS > 0 ? KS : KM);
    constexpr int PB = KS > 0 ? tri_pairs(KM, KS, KM) : 0;
    constexpr int PX = PA > PB ? PA : PB;
    __shared__ double s_corr[KS > 0 ? 128 : 1][KM + 1];
    __shared__ __attribute__((aligned(16))) double sP[4 * GHQ_CH * KM];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int role = KS > 0 ? (wave & 1) : 0;
    const int slot = KS > 0 ? (wave >> 1) * 64 + lane : 0;
    const int64_t i = KS > 0 ? (int64_t)blockIdx.x * 128 + slot : (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_split;
    const int64_t r1 = r0 + rows_per_split < Mb ? r0 + rows_per_split : Mb;
    if (r0 >= r1) return;                               // the whole workgroup
    const bool live = i < N;
    double q[KM], hx[PX], hz[KM], tw[KM];
    double sz2 = 0.0, sw2 = 0.0;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        q[k] = (live && k < K) ? Q[i * K + k] : 0.0;
        hz[k] = 0.0;
        tw[k] = 0.0;
    }
#pragma unroll
    for (int t = 0; t < PX; ++t) hx[t] = 0.0;
    if (KS == 0) {
        gh_q_loop<KM, 0, KM, true>(G, ldg, M, K, P, sP, q, i, live, r0, r1, hx, hz, tw, sz2, sw2);
        if (live) gh_q_store<KM, 0, KM, true>(K, i, hx, hz, tw, sz2, sw2, XtX, Xtz);
        return;
    }
    constexpr int KA = KS > 0 ? KS : KM;
    if (role == 0)
        gh_q_loop<KM, 0, KA, true>(G, ldg, M, K, P, sP, q, i, live, r0, r1, hx, hz, tw, sz2, sw2);
    else
        gh_q_loop<KM, KA, KM, false>(G, ldg, M, K, P, sP, q, i, live, r0, r1, hx, hz, tw, sz2, sw2);
    if (role == 0) {
#pragma unroll
        for (int k = 0; k < KM; ++k) s_corr[slot][k] = tw[k];
        s_corr[slot][KM] = sw2;
    }
    __syncthreads();
    if (!live) return;
    if (role == 0) {
        gh_q_store<KM, 0, KA, true>(K, i, hx, hz, tw, sz2, sw2, XtX, Xtz);
    } else {
#pragma unroll
        for (int k = 0; k < KM; ++k) tw[k] = s_corr[slot][k];
        gh_q_store<KM, KA, KM, false>(K, i, hx, hz, tw, 0.0, s_corr[slot][KM], XtX, Xtz);
    }
}

template <int KM>
__global__ __launch_bounds__(256) void grad_hess_p_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                          int64_t N, int K, const double* __restrict__ Q,
                                                          const double* __restrict__ P, int64_t n_per_split,
                                                          double* __restrict__ XtX, double* __restrict__ Xtz) {
    constexpr int KT = KM * (KM + 1) / 2;
    constexpr int CH = 256;   // samples per LDS chunk
    constexpr int SGS = CH + 4;  // byte-row stride: the 16 rows a wave reads fall on distinct banks
    // Workgroup = 256 SNPs = 64 byte rows; thread = SNP (byte row tid / 4,
    // field tid % 4).  Each chunk stages the 64 rows' genotype bytes and the
    // chunk's Q rows (read by every lane alike: LDS broadcasts).
    __shared__ uint8_t sg[64 * SGS];
    __shared__ __attribute__((aligned(16))) double sq[CH * KM];
    const int tid = threadIdx.x;
    const int64_t m = (int64_t)blockIdx.x * 256 + tid;
    const int64_t row0 = (int64_t)blockIdx.x * 64;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t n0 = (int64_t)blockIdx.y * n_per_split;
    const int64_t n1 = n0 + n_per_split < N ? n0 + n_per_split : N;
    const bool live = m < M;
    double p[KM], hz[KM], hx[KT];
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        p[k] = (live && k < K) ? P[m * K + k] : 0.0;
        hz[k] = 0.0;
    }
#pragma unroll
    for (int t = 0; t < KT; ++t) hx[t] = 0.0;
    const int rr = tid >> 2, sh = 2 * (tid & 3);
    for (int64_t c0 = n0; c0 < n1; c0 += CH) {
        const int nc = (int)(n1 - c0 < CH ? n1 - c0 : CH);
        __syncthreads();
        // 64 rows x CH bytes: thread -> row e / 16, 16-byte column e % 16
        for (int e = tid; e < 64 * (CH / 16); e += 256) {
            const int r = e / (CH / 16), c = 16 * (e - r * (CH / 16));
            uint4 v = make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
            if (row0 + r < Mb && c < nc) v = *reinterpret_cast<const uint4*>(G + (row0 + r) * ldg + c0 + c);
            uint32_t* d = reinterpret_cast<uint32_t*>(sg + r * SGS + c);
            d[0] = v.x;
            d[1] = v.y;
            d[2] = v.z;
            d[3] = v.w;
        }
        for (int e = tid; e < CH * KM; e += 256) {
            const int c = e / KM, k = e - c * KM;
            sq[e] = (c < nc && k < K) ? Q[(c0 + c) * K + k] : 0.0;
        }
        __syncthreads();
        if (!live) continue;
        for (int c = 0; c < nc; ++c) {
            const uint32_t code = (sg[rr * SGS + c] >> sh) & 3u;
            const double* qc = sq + c * KM;
            double qv[KM];
            double h = 0.0;
#pragma unroll
            for (int k = 0; k < KM; ++k) {
                qv[k] = qc[k];
                h += qv[k] * p[k];
            }
            const Weights w = bq_weights(h, code);
            const double zd = w.z1 - w.z2, ws = w.w1 + w.w2;
            double wq[KM];
#pragma unroll
            for (int k = 0; k < KM; ++k) {
                hz[k] += zd * qv[k];
                wq[k] = ws * qv[k];
            }
            int t = 0;
#pragma unroll
            for (int k = 0; k < KM; ++k)
#pragma unroll
                for (int l = k; l < KM; ++l, ++t) hx[t] += wq[k] * qv[l];
        }
    }
    if (!live) return;
    int t = 0;
#pragma unroll
    for (int k = 0; k < KM; ++k) {
        if (k < K) atomicAdd(&Xtz[m * K + k], hz[k]);
#pragma unroll
        for (int l = k; l < KM; ++l, ++t)
            if (k < K && l < K) {
                atomicAdd(&XtX[(m * K + k) * K + l], hx[t]);
                if (l != k) atomicAdd(&XtX[(m * K + l) * K + k], hx[t]);
            }
    }
}

// ---------------------------------------------------------------------------
// K = 17 .. 64: one problem (sample or SNP) per wavefront.  K(K+1)/2 FP64
// Hessian terms do not fit one thread's registers, so lane l of the wave owns
// the 4 x 4 blocks t = l, l + 64, l + 128 of the upper block triangle (at most
// 136 blocks at K = 64) and lanes 0 .. 15 own 4 gradient terms each.  The
// four problems of a workgroup share one LDS-staged chunk of 64 operand rows
// (P rows on the Q side, Q rows on the P side): lane j of a wave forms H and
// the weights of row j of the chunk for its own problem, then every lane
// accumulates its blocks over the 64 rows.  Each output element is written
// by exactly one lane: no atomics, fixed summation order.
constexpr int GHW_CH = 64;            // operand rows per chunk
constexpr int GHW_LD = 65;            // LDS row stride (doubles): the H dot products of a wave hit distinct banks
constexpr int GHW_TASKS = 3;          // 4 x 4 blocks per lane (ceil(136 / 64))

template <int SIDE>
__global__ __launch_bounds__(256) void grad_hess_wide_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                             int64_t N, int K, const double* __restrict__ Q,
                                                             const double* __restrict__ P, double* __restrict__ XtX,
                                                             double* __restrict__ Xtz) {
    __shared__ double s_op[GHW_CH * GHW_LD];   // chunk operand rows (zero past K and past the end)
    __shared__ double s_own[4][64];            // each wave's own row (q_i or p_j)
    __shared__ double s_w[4][4][GHW_CH];       // per wave: weights of the chunk rows
    __shared__ double s_tw[4][64];             // Q side: each wave's T_k = sum_j w2 p_k
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int64_t nprob = SIDE == 0 ? N : M;   // problems: samples (Q side) or SNPs (P side)
    const int64_t nrow = SIDE == 0 ? M : N;    // reduction rows: SNPs or samples
    const double* own_src = SIDE == 0 ? Q : P;
    const double* op_src = SIDE == 0 ? P : Q;
    const int64_t prob = (int64_t)blockIdx.x * 4 + wave;
    const bool live = prob < nprob;
    const int nb = (K + 3) >> 2;
    const int ntask = nb * (nb + 1) / 2;
    s_own[wave][lane] = (live && lane < K) ? own_src[prob * K + lane] : 0.0;
    // lane's blocks (bk, bl), bk <= bl, enumerated row by row
    int tk[GHW_TASKS], tl[GHW_TASKS];
#pragma unroll
    for (int u = 0; u < GHW_TASKS; ++u) {
        int t = lane + 64 * u, bk = 0;
        while (bk < nb && t >= nb - bk) {
            t -= nb - bk;
            ++bk;
        }
        tk[u] = bk < nb ? 4 * bk : -1;
        tl[u] = bk < nb ? 4 * (bk + t) : -1;
    }
    double hx[GHW_TASKS][16];
    double hz[4], tw[4];
    double sz2 = 0.0, sw2 = 0.0;   // Q side: S_z2, S_w2 (every lane alike)
#pragma unroll
    for (int u = 0; u < GHW_TASKS; ++u)
#pragma unroll
        for (int e = 0; e < 16; ++e) hx[u][e] = 0.0;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        hz[a] = 0.0;
        tw[a] = 0.0;
    }

    for (int64_t r0 = 0; r0 < nrow; r0 += GHW_CH) {
        __syncthreads();
        for (int e = threadIdx.x; e < GHW_CH * 64; e += 256) {
            const int r = e >> 6, k = e & 63;
            s_op[r * GHW_LD + k] = (r0 + r < nrow && k < K) ? op_src[(r0 + r) * K + k] : 0.0;
        }
        __syncthreads();
        // weights of row lane of the chunk for this wave's problem
        {
            const int64_t r = r0 + lane;
            uint32_t code = 3u;
            if (live && r < nrow) {
                const int64_t snp = SIDE == 0 ? r : prob, smp = SIDE == 0 ? prob : r;
                code = (G[(snp >> 2) * ldg + smp] >> (2 * (snp & 3))) & 3u;
            }
            double h = 0.0;
            for (int k = 0; k < K; ++k) h += s_own[wave][k] * s_op[lane * GHW_LD + k];
            const Weights w = bq_weights(h, code);
            if (SIDE == 0) {
                s_w[wave][0][lane] = w.z1 - w.z2;
                s_w[wave][1][lane] = w.z2;
                s_w[wave][2][lane] = w.w1 + w.w2;
                s_w[wave][3][lane] = w.w2;
            } else {
                s_w[wave][0][lane] = w.z1 - w.z2;
                s_w[wave][1][lane] = w.w1 + w.w2;
            }
        }
        __syncthreads();
        const int nr = (int)(nrow - r0 < GHW_CH ? nrow - r0 : GHW_CH);
        for (int j = 0; j < nr; ++j) {
            const double* x = s_op + j * GHW_LD;
            if (SIDE == 0) {
                // XtX_q += (w1 + w2) p p^T, Xtz_q += (z1 - z2) p, and the
                // sums S_z2, S_w2, T of the (1-p) terms
                const double zd = s_w[wave][0][j], z2 = s_w[wave][1][j];
                const double ws = s_w[wave][2][j], w2 = s_w[wave][3][j];
                sz2 += z2;
                sw2 += w2;
                if (lane < 16)
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const double pk = x[(4 * lane + a) & 63];
                        hz[a] += zd * pk;
                        tw[a] += w2 * pk;
                    }
#pragma unroll
                for (int u = 0; u < GHW_TASKS; ++u) {
                    if (tk[u] < 0) continue;
                    double ak[4], pl[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        ak[a] = ws * x[tk[u] + a];
                        pl[a] = x[tl[u] + a];
                    }
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) hx[u][4 * a + b] += ak[a] * pl[b];
                }
            } else {
                // XtX_p += (w1 + w2) q q^T, Xtz_p += (z1 - z2) q
                const double zd = s_w[wave][0][j], ws = s_w[wave][1][j];
                if (lane < 16)
#pragma unroll
                    for (int a = 0; a < 4; ++a) hz[a] += zd * x[(4 * lane + a) & 63];
#pragma unroll
                for (int u = 0; u < GHW_TASKS; ++u) {
                    if (tk[u] < 0) continue;
                    double wk[4], ql[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        wk[a] = ws * x[tk[u] + a];
                        ql[a] = x[tl[u] + a];
                    }
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) hx[u][4 * a + b] += wk[a] * ql[b];
                }
            }
        }
    }
    if (SIDE == 0) {   // T_k of this wave's problem for every lane's blocks
        if (lane < 16)
#pragma unroll
            for (int a = 0; a < 4; ++a) s_tw[wave][4 * lane + a] = tw[a];
        __syncthreads();
    }
    if (!live) return;
    if (lane < 16)
#pragma unroll
        for (int a = 0; a < 4; ++a)
            if (4 * lane + a < K) Xtz[prob * K + 4 * lane + a] = SIDE == 0 ? hz[a] + sz2 : hz[a];
    double* H = XtX + prob * (int64_t)K * K;
#pragma unroll
    for (int u = 0; u < GHW_TASKS; ++u) {
        if (tk[u] < 0) continue;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const int k = tk[u] + a, l = tl[u] + b;
                if (k < K && l < K) {
                    const double v = SIDE == 0 ? ((hx[u][4 * a + b] + sw2) - s_tw[wave][k]) - s_tw[wave][l]
                                               : hx[u][4 * a + b];
                    H[k * K + l] = v;
                    H[l * K + k] = v;
                }
            }
    }
    (void)ntask;
}

template <int KM>
static void launch_gh(int side, hipStream_t st, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K,
                      const double* Q, const double* P, double* XtX, double* Xtz) {
    const int64_t Mb = ceil_div(M, 4);
    if (side == 0) {
        // rows of the Hessian triangle owned by the first of two waves (0: one
        // thread per sample); chosen so both roles do about equal work.  One
        // thread per sample up to K = 10 (218 VGPRs, two waves per SIMD): at
        // N=100k, M=500k, K=10 it took 216 ms against 281 ms for the split,
        // which forms H and the weights in both waves; at K = 12 it needs
        // more than 256 VGPRs and the split is faster (902 against 965 ms
        // at N=200k, M=600k).
        constexpr int KS = KM <= 10 ? 0 : KM <= 14 ? 3 : 4;
        const int64_t nbx = ceil_div(N, KS > 0 ? 128 : 256);
        int64_t s = ceil_div(4096, nbx);
        s = s > Mb ? Mb : (s < 1 ? 1 : s);
        if (s > 65535) s = 65535;
        const int64_t rps = ceil_div(Mb, s);
        dim3 grid((unsigned)nbx, (unsigned)ceil_div(Mb, rps));
        grad_hess_q_kernel<KM, KS><<<grid, 256, 0, st>>>(G, ldg, M, N, K, Q, P, rps, XtX, Xtz);
    } else {
        const int64_t nbx = ceil_div(M, 256);
        int64_t s = ceil_div(2048, nbx);
        const int64_t max_s = ceil_div(N, 256);
        s = s > max_s ? max_s : (s < 1 ? 1 : s);
        if (s > 65535) s = 65535;
        const int64_t nps = ceil_div(ceil_div(N, s), 256) * 256;
        dim3 grid((unsigned)nbx, (unsigned)ceil_div(N, nps));
        grad_hess_p_kernel<KM><<<grid, 256, 0, st>>>(G, ldg, M, N, K, Q, P, nps, XtX, Xtz);
    }
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_grad_hess(int side, const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K, const double* Q,
                  const double* P, double* XtX, double* Xtz, void* stream) {
    if (M <= 0 || N <= 0) return 0;
    if (K < 1 || K > 64) return fail_msg(1, "adx_grad_hess: K must be in [1, 64]");
    if (side != 0 && side != 1) return fail_msg(1, "adx_grad_hess: side must be 0 (Q) or 1 (P)");
    if (ldg < N || ldg % 16 != 0) return fail_msg(1, "adx_grad_hess: ldg must be >= N and a multiple of 16");
    if (reinterpret_cast<uintptr_t>(G) % 16 != 0) return fail_msg(1, "adx_grad_hess: G not 16-byte aligned");
    hipStream_t st = (hipStream_t)stream;
    const int64_t rows = side == 0 ? N : M;
    if (K > 16) {   // every output element is stored once: no clearing needed
        const unsigned grid = (unsigned)ceil_div(rows, 4);
        if (side == 0)
            grad_hess_wide_kernel<0><<<grid, 256, 0, st>>>(G, ldg, M, N, K, Q, P, XtX, Xtz);
        else
            grad_hess_wide_kernel<1><<<grid, 256, 0, st>>>(G, ldg, M, N, K, Q, P, XtX, Xtz);
        ADX_LAUNCH_CHECK("adx_grad_hess");
        return 0;
    }
    ADX_CHECK(hipMemsetAsync(XtX, 0, sizeof(double) * (size_t)(rows * K * K), st));
    ADX_CHECK(hipMemsetAsync(Xtz, 0, sizeof(double) * (size_t)(rows * K), st));
    // register blocks padded to an even K (the K(K+1)/2 Hessian terms of
    // each thread are the bulk of the work)
    switch ((K + 1) / 2) {
        case 1: case 2: launch_gh<4>(side, st, G, ldg, M, N, K, Q, P, XtX, Xtz); break;
        case 3: launch_gh<6>(side, st, G, ldg, M, N, K, Q, P, XtX, Xtz); break;
        case 4: launch_gh<8>(side, st, G, ldg, M, N, K, Q, P, XtX, Xtz); break;
        case 5: launch_gh<10>(side, st, G, ldg, M, N, K, Q, P, XtX, Xtz); break;
        case 6: launch_gh<12>(side, st, G, ldg, M, N, K, Q, P, XtX, Xtz); break;
        case 7: launch_gh<14>(side, st, G, ldg, M, N, K, Q, P, XtX, Xtz); break;
        default: launch_gh<16>(side, st, G, ldg, M, N, K, Q, P, XtX, Xtz); break;
    }
    ADX_LAUNCH_CHECK("adx_grad_hess");
    return 0;
}

}  // extern "C"
