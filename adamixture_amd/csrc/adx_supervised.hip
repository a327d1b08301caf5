// Supervised-mode kernels (adamixture/src/supervised.py).
//
// label_freq_kernel -- init_p_supervised_packed (supervised.py:84-107):
//   P[m][k] = (sum of the observed allele counts of SNP m over the samples
//   labelled k + 1) / (2 x their observed count); 0.5 where none is observed;
//   clipped to [eps, 1 - eps].  The host hands over the labelled samples
//   grouped by population, ascending sample index inside a group
//   (order[off[k] .. off[k + 1])), so the 64 lanes of a wave read nearby bytes
//   of one byte row.  One wave per (byte row, population); the four SNPs of a
//   byte are counted in integer registers (exact) and the quotient is one
//   FP64 division of exact integers, as numpy computes it, so P is bit-exact.
// snap_labels_kernel -- _snap_q_gpu (supervised.py:308-339): the rows of
//   labelled samples are set to eps / max(K - 1, 1), with 1 - eps on the
//   labelled population; unlabelled rows are left alone.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

__global__ __launch_bounds__(256) void label_freq_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                          const int32_t* __restrict__ order,
                                                          const int64_t* __restrict__ off, double eps,
                                                          double* __restrict__ P, int K) {
    const int lane = threadIdx.x & 63;
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int k = blockIdx.y;
    if (j >= ((M + 3) >> 2)) return;
    const int64_t s0 = off[k], s1 = off[k + 1];
    uint32_t sum[4] = {0u, 0u, 0u, 0u}, obs[4] = {0u, 0u, 0u, 0u};
    for (int64_t s = s0 + lane; s < s1; s += 64) {
        const uint32_t b = G[j * ldg + order[s]];
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const uint32_t code = (b >> (2 * f)) & 3u;
            if (code != 3u) {
                sum[f] += code;
                obs[f] += 1u;
            }
        }
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            sum[f] += __shfl_xor(sum[f], o);
            obs[f] += __shfl_xor(obs[f], o);
        }
    }
    if (lane == 0) {
#pragma unroll
        for (int f = 0; f < 4; ++f) {
            const int64_t m = 4 * j + f;
            if (m < M) {
                double p = obs[f] > 0u ? (double)sum[f] / (2.0 * (double)obs[f]) : 0.5;
                p = fmin(fmax(p, eps), 1.0 - eps);
                P[m * K + k] = p;
            }
        }
    }
}

__global__ void snap_labels_kernel(double* __restrict__ Q, int64_t N, int K, const int32_t* __restrict__ y,
                                   double eps) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int pop = y[i];
    if (pop <= 0) return;
    const double fill = eps / (double)(K - 1 > 1 ? K - 1 : 1);
    for (int k = 0; k < K; ++k) Q[i * K + k] = (k == pop - 1) ? 1.0 - eps : fill;
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_label_freq(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int K, const int32_t* order,
                   const int64_t* offsets, double eps, double* P, void* stream) {
    if (M <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_label_freq: bad shape");
    if (K < 1 || K > 65535) return fail_msg(1, "adx_label_freq: need 1 <= K <= 65535");
    if (N >= (int64_t)1 << 30) return fail_msg(1, "adx_label_freq: N must be below 2^30 (32-bit counts)");
    const int64_t Mb = ceil_div(M, 4);
    label_freq_kernel<<<dim3((unsigned)ceil_div(Mb, 4), (unsigned)K), 256, 0, (hipStream_t)stream>>>(
        G, ldg, M, order, offsets, eps, P, K);
    ADX_LAUNCH_CHECK("adx_label_freq");
    return 0;
}

int adx_snap_labels(double* Q, int64_t N, int K, const int32_t* y, double eps, void* stream) {
    if (K < 1) return fail_msg(1, "adx_snap_labels: K < 1");
    if (N <= 0) return 0;
    snap_labels_kernel<<<(unsigned)ceil_div(N, 256), 256, 0, (hipStream_t)stream>>>(Q, N, K, y, eps);
    ADX_LAUNCH_CHECK("adx_snap_labels");
    return 0;
}

}  // extern "C"
