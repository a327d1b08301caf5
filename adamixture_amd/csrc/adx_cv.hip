// Cross-validation kernels (adamixture/src/cv.py:171-267, br_qn.pyx:92-246).
//
// A fold holds out a set of observed genotypes, refits from the global
// (P, Q) on the matrix with those entries missing, and scores the held-out
// entries by their binomial deviance.  The reference does all of it on the
// host over the unpacked matrix; here the fold's matrix is built and scored
// on the device from the packed one.
//
// Two ways to name a fold, as in the reference:
//   list:   explicit flat indices m N + i (the host's shuffled int32 array,
//           fold f = every n_folds-th entry from f; cv.py:214-231);
//   hash:   entry (m, i) is in fold f iff splitmix64(m N + i + seed) mod
//           n_folds == f (cv.py:61-81, used when M N exceeds int32).
// Kernels:
//   cv_mask_list_kernel   sets the held-out 2-bit fields of a copy of G to 3
//                         (missing) with 32-bit atomic ORs (several entries
//                         may share a word; OR with 3 is idempotent);
//   cv_mask_hash_kernel   writes the masked copy in one pass, one byte per
//                         thread (4 SNPs of one sample);
//   cv_dev_*_kernel       deviance of the held-out entries at (P_cv, Q_cv),
//                         per-block FP64 partial sums reduced in a fixed
//                         order by cv_sum_kernel (deterministic);
//   project_simplex_kernel  Q rows onto the simplex with pseudocount, the
//                         CPU optimizer's quasi-Newton map (sqp.pyx:112-146).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Fold test of the hashed path, splitmix64(idx + seed) mod n_folds == fold,
// with the modulo by the run-time n_folds done as a multiply-high by a
// precomputed magic (the division-by-invariant-integer method: q =
// (hi(x m) + ((x - hi(x m)) >> 1)) >> (s - 1), exact for every 64-bit x;
// a 64-bit % by a variable is a long instruction sequence on the GPU).
struct FoldHash {
    uint64_t seed, magic;
    uint64_t d;
    int shift, fold;
    __device__ __forceinline__ bool held(uint64_t idx) const {
        const uint64_t x = splitmix64(idx + seed);
        uint64_t r = 0;
        if (shift > 0) {
            const uint64_t hi = __umul64hi(x, magic);
            const uint64_t q = ((((x - hi) >> 1) + hi) >> (shift - 1));
            r = x - q * d;
        }
        return r == (uint64_t)fold;
    }
};

static FoldHash fold_hash(uint64_t seed, int n_folds, int fold) {
    FoldHash h;
    h.seed = seed;
    h.d = (uint64_t)n_folds;
    h.fold = fold;
    if (n_folds == 1) {
        h.shift = 0;
        h.magic = 0;
    } else {
        h.shift = 64 - __builtin_clzll(h.d - 1);
        const unsigned __int128 one = 1;
        h.magic = (uint64_t)(((one << (64 + h.shift)) / h.d) - (one << 64) + 1);
    }
    return h;
}

template <typename I>
__device__ __forceinline__ int64_t list_entry(const void* flat, int64_t e, int64_t first, int64_t stride) {
    return (int64_t) reinterpret_cast<const I*>(flat)[first + e * stride];
}

// br_qn.pyx:121-147: one held-out entry's deviance term, from the dot
// product q . p (summed over k in ascending order).
__device__ __forceinline__ double deviance_from_mu(uint32_t code, double mu) {
    const double g = (double)code;
    mu = 2.0 * mu;
    mu = mu < 1e-10 ? 1e-10 : (mu > 2.0 - 1e-10 ? 2.0 - 1e-10 : mu);
    const double rem = 2.0 - g;
    const double ta = g > 0.0 ? g * log(g / mu) : 0.0;
    const double tb = rem > 0.0 ? rem * log(rem / (2.0 - mu)) : 0.0;
    return ta + tb;
}

__device__ __forceinline__ double deviance_term(uint32_t code, const double* __restrict__ p,
                                                const double* __restrict__ q, int K) {
    double mu = 0.0;
    for (int k = 0; k < K; ++k) mu += q[k] * p[k];
    return deviance_from_mu(code, mu);
}

// Block sum of v into part[blockIdx.x] (256 threads, fixed tree).
__device__ __forceinline__ void block_sum_store(double v, double* __restrict__ part) {
    __shared__ double s[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (s[0] + s[1]) + (s[2] + s[3]);
}

template <typename I>
__global__ __launch_bounds__(256) void cv_mask_list_kernel(uint8_t* __restrict__ G, int64_t ldg, int64_t N,
                                                           const void* __restrict__ flat, int64_t first,
                                                           int64_t stride, int64_t n) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t f = list_entry<I>(flat, e, first, stride);
        const int64_t m = f / N, i = f - m * N;
        const int64_t byte = (m >> 2) * ldg + i;
        const uint32_t sh = 8u * (uint32_t)(byte & 3) + 2u * (uint32_t)(m & 3);
        atomicOr(reinterpret_cast<unsigned int*>(G + (byte & ~(int64_t)3)), 3u << sh);
    }
}

__global__ __launch_bounds__(256) void cv_mask_hash_kernel(const uint8_t* __restrict__ G, uint8_t* __restrict__ Gm,
                                                           int64_t ldg, int64_t M, int64_t N, FoldHash fh) {
    const int64_t Mb = (M + 3) >> 2;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < Mb * ldg;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / ldg, i = e - j * ldg;
        uint32_t b = G[e];
        if (i < N) {
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const int64_t m = 4 * j + f;
                const uint32_t code = (b >> (2 * f)) & 3u;
                if (m < M && code != 3u && fh.held((uint64_t)(m * N + i))) b |= 3u << (2 * f);
            }
        }
        Gm[e] = (uint8_t)b;
    }
}

// Entries [blockIdx.x * chunk, ...) of the list, threads strided inside.
template <typename I>
__global__ __launch_bounds__(256) void cv_dev_list_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t N,
                                                          const void* __restrict__ flat, int64_t first,
                                                          int64_t stride, int64_t n, int64_t chunk,
                                                          const double* __restrict__ P, const double* __restrict__ Q,
                                                          int K, double* __restrict__ part) {
    const int64_t e0 = (int64_t)blockIdx.x * chunk;
    const int64_t e1 = e0 + chunk < n ? e0 + chunk : n;
    double acc = 0.0;
    for (int64_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
        const int64_t f = list_entry<I>(flat, e, first, stride);
        const int64_t m = f / N, i = f - m * N;
        const uint32_t code = (G[(m >> 2) * ldg + i] >> (2 * (m & 3))) & 3u;
        if (code != 3u) acc += deviance_term(code, P + m * K, Q + i * K, K);
    }
    block_sum_store(acc, part);
}

// Thread = sample i (its Q row in registers when K <= KQ); workgroup = 256
// consecutive samples (blocks strided by gridDim.x) x one range of byte rows
// (blockIdx.y).  Each genotype byte is a coalesced read and the 4 P rows of
// a byte row are the same address in every lane.
constexpr int KQ = 16;

__global__ __launch_bounds__(256) void cv_dev_hash_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                          int64_t N, FoldHash fh, const double* __restrict__ P,
                                                          const double* __restrict__ Q, int K,
                                                          int64_t rows_per_split, double* __restrict__ part) {
    const int64_t Mb = (M + 3) >> 2;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_split;
    const int64_t r1 = r0 + rows_per_split < Mb ? r0 + rows_per_split : Mb;
    const int64_t nsb = (N + 255) / 256;
    // Per held-out entry, g log(g/mu) + (2-g) log((2-g)/(2-mu)) =
    //   [g log g + (2-g) log(2-g)] - log(mu^g (2-mu)^(2-g)):
    // the first part is 2 ln 2 for g = 0 or 2 and 0 for g = 1 (counted), the
    // second a log-product (factors >= 1e-20, one log per thread at the end).
    LogProd lp;
    int64_t n_hom = 0;
    for (int64_t sb = blockIdx.x; sb < nsb; sb += gridDim.x) {
        const int64_t i = sb * 256 + threadIdx.x;
        if (i >= N) continue;
        double q[KQ];
#pragma unroll
        for (int k = 0; k < KQ; ++k) q[k] = k < K ? Q[i * K + k] : 0.0;
        for (int64_t r = r0; r < r1; ++r) {
            const uint32_t b = G[r * ldg + i];
#pragma unroll
            for (int f = 0; f < 4; ++f) {
                const int64_t m = 4 * r + f;
                const uint32_t code = (b >> (2 * f)) & 3u;
                if (m < M && code != 3u && fh.held((uint64_t)(m * N + i))) {
                    double mu = 0.0;
                    if (K <= KQ) {
#pragma unroll
                        for (int k = 0; k < KQ; ++k)
                            if (k < K) mu += q[k] * P[m * K + k];
                    } else {
                        for (int k = 0; k < K; ++k) mu += Q[i * K + k] * P[m * K + k];
                    }
                    mu = 2.0 * mu;
                    mu = mu < 1e-10 ? 1e-10 : (mu > 2.0 - 1e-10 ? 2.0 - 1e-10 : mu);
                    const double x = code == 2u ? mu : 2.0 - mu, y = code == 0u ? 2.0 - mu : mu;
                    lp.mul_factor(x * y);
                    n_hom += code != 1u;
                }
            }
            lp.renorm();
        }
    }
    const double ln2_hi = 0.6931471824645996, ln2_lo = -1.904654299957768e-09;
    const double acc = 2.0 * ((double)n_hom * ln2_hi + (double)n_hom * ln2_lo) - lp.log_value();
    block_sum_store(acc, part + (int64_t)blockIdx.y * gridDim.x);
}

// out[0] = sum of part[0 .. n) in a fixed order (one workgroup).
__global__ __launch_bounds__(256) void cv_sum_kernel(const double* __restrict__ part, int n,
                                                     double* __restrict__ out) {
    double acc = 0.0;
    for (int e = threadIdx.x; e < n; e += blockDim.x) acc += part[e];
    __shared__ double s[4];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (s[0] + s[1]) + (s[2] + s[3]);
}

__global__ void project_simplex_kernel(double* __restrict__ X, int64_t n, int K, double pc) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double b[64];
    project_simplex_row(X + r * K, X + r * K, b, K, pc);
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_cv_mask_list(uint8_t* G, int64_t ldg, int64_t M, int64_t N, const void* flat, int idx64, int64_t first,
                     int64_t stride, int64_t n, void* stream) {
    if (M <= 0 || N <= 0 || ldg < N || ldg % 4 != 0) return fail_msg(1, "adx_cv_mask_list: bad shape");
    if (first < 0 || stride < 1) return fail_msg(1, "adx_cv_mask_list: bad first / stride");
    if (n <= 0) return 0;
    int64_t blocks = ceil_div(n, 256);
    if (blocks > 16384) blocks = 16384;
    hipStream_t st = (hipStream_t)stream;
    if (idx64)
        cv_mask_list_kernel<int64_t><<<(unsigned)blocks, 256, 0, st>>>(G, ldg, N, flat, first, stride, n);
    else
        cv_mask_list_kernel<int32_t><<<(unsigned)blocks, 256, 0, st>>>(G, ldg, N, flat, first, stride, n);
    ADX_LAUNCH_CHECK("adx_cv_mask_list");
    return 0;
}

int adx_cv_mask_hash(const uint8_t* G, uint8_t* Gm, int64_t ldg, int64_t M, int64_t N, uint64_t seed, int n_folds,
                     int fold, void* stream) {
    if (M <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_cv_mask_hash: bad shape");
    if (n_folds < 1 || fold < 0 || fold >= n_folds) return fail_msg(1, "adx_cv_mask_hash: bad fold");
    int64_t blocks = ceil_div(ceil_div(M, 4) * ldg, 256);
    if (blocks > 65536) blocks = 65536;
    cv_mask_hash_kernel<<<(unsigned)blocks, 256, 0, (hipStream_t)stream>>>(G, Gm, ldg, M, N,
                                                                        fold_hash(seed, n_folds, fold));
    ADX_LAUNCH_CHECK("adx_cv_mask_hash");
    return 0;
}

int adx_cv_deviance_list(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, const void* flat, int idx64,
                         int64_t first, int64_t stride, int64_t n, const double* P, const double* Q, int K,
                         double* work, double* out, void* stream) {
    if (M <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_cv_deviance_list: bad shape");
    if (K < 1 || first < 0 || stride < 1) return fail_msg(1, "adx_cv_deviance_list: bad K / first / stride");
    hipStream_t st = (hipStream_t)stream;
    const int64_t chunk = n > 0 ? ceil_div(n, ADX_CV_BLOCKS) : 1;
    const int blocks = (int)(n > 0 ? ceil_div(n, chunk) : 0);
    if (blocks > 0) {
        if (idx64)
            cv_dev_list_kernel<int64_t><<<blocks, 256, 0, st>>>(G, ldg, N, flat, first, stride, n, chunk, P, Q, K,
                                                                 work);
        else
            cv_dev_list_kernel<int32_t><<<blocks, 256, 0, st>>>(G, ldg, N, flat, first, stride, n, chunk, P, Q, K,
                                                                 work);
    }
    cv_sum_kernel<<<1, 256, 0, st>>>(work, blocks, out);
    ADX_LAUNCH_CHECK("adx_cv_deviance_list");
    return 0;
}

int adx_cv_deviance_hash(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, uint64_t seed, int n_folds, int fold,
                         const double* P, const double* Q, int K, double* work, double* out, void* stream) {
    if (M <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_cv_deviance_hash: bad shape");
    if (K < 1 || n_folds < 1 || fold < 0 || fold >= n_folds) return fail_msg(1, "adx_cv_deviance_hash: bad K / fold");
    hipStream_t st = (hipStream_t)stream;
    const int64_t Mb = ceil_div(M, 4);
    // at most ADX_CV_BLOCKS workgroups (one partial each): sample blocks x
    // byte-row splits
    const int64_t nsb = ceil_div(N, 256);
    const int64_t gx = nsb < ADX_CV_BLOCKS ? nsb : ADX_CV_BLOCKS;
    int64_t gy = ADX_CV_BLOCKS / gx;
    if (gy > Mb) gy = Mb;
    if (gy < 1) gy = 1;
    const int64_t rps = ceil_div(Mb, gy);
    gy = ceil_div(Mb, rps);
    dim3 grid((unsigned)gx, (unsigned)gy);
    cv_dev_hash_kernel<<<grid, 256, 0, st>>>(G, ldg, M, N, fold_hash(seed, n_folds, fold), P, Q, K, rps, work);
    cv_sum_kernel<<<1, 256, 0, st>>>(work, (int)(gx * gy), out);
    ADX_LAUNCH_CHECK("adx_cv_deviance_hash");
    return 0;
}

int adx_project_simplex(double* X, int64_t n, int K, double pc, void* stream) {
    if (K < 1 || K > 64) return fail_msg(1, "adx_project_simplex: need 1 <= K <= 64");
    if (n <= 0) return 0;
    project_simplex_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(X, n, K, pc);
    ADX_LAUNCH_CHECK("adx_project_simplex");
    return 0;
}

}  // extern "C"
