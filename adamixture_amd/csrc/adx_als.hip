// Skinny FP64 products of the ALS warm start (adamixture/model/als_gpu.py:
// 90-113): every product there has one long dimension (M SNPs or N samples)
// and all others <= 64, e.g. Q^T Q (K x K over N rows), V^T I_q (L x K over
// N rows), Z (V^T I_q) (M x L times L x K).  General GEMM libraries tile for
// large square blocks; here
//   gemm_tn:  C = alpha X^T Y (a x b), X (n x a), Y (n x b): each workgroup
//             reduces a range of rows with the a x b partial sums spread over
//             its threads (row chunks staged in LDS) and stores them in its
//             slice of the caller's workspace; one more kernel sums the
//             slices in a fixed order (deterministic; without a workspace,
//             one FP64 atomic per output and workgroup);
//   gemm_nn:  C = alpha X W + u v^T (+ beta C), X (n x a), W (a x b) small:
//             W (and v) staged in LDS, one thread per (row, output column
//             group), the row of X read once per thread group.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

constexpr int TN_ROWS = 32;   // rows per LDS chunk

// C[i][j] = 0 for i < a, j < b (a kernel, not a memset: graph-safe).
__global__ void zero_block_kernel(double* __restrict__ C, int a, int b, int64_t ldc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < a * b) C[(int64_t)(e / b) * ldc + e % b] = 0.0;
}

// C[i][j] += alpha sum_r X[r][i] Y[r][j]; thread t owns outputs t, t + 256, ...
__global__ __launch_bounds__(256) void gemm_tn_kernel(const double* __restrict__ X, int64_t ldx,
                                                       const double* __restrict__ Y, int64_t ldy, int64_t n, int a,
                                                       int b, double alpha, int64_t rows_per_block,
                                                       double* __restrict__ C, int64_t ldc,
                                                       double* __restrict__ part) {
    __shared__ double sx[TN_ROWS][64];
    __shared__ double sy[TN_ROWS][64];
    const int tid = threadIdx.x;
    const int nout = a * b;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};          // a * b <= 1024 = 4 x 256
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < n ? r0 + rows_per_block : n;
    for (int64_t c0 = r0; c0 < r1; c0 += TN_ROWS) {
        const int nr = (int)(r1 - c0 < TN_ROWS ? r1 - c0 : TN_ROWS);
        __syncthreads();
        for (int e = tid; e < TN_ROWS * 64; e += 256) {
            const int rr = e >> 6, c = e & 63;
            sx[rr][c] = (rr < nr && c < a) ? X[(c0 + rr) * ldx + c] : 0.0;
            sy[rr][c] = (rr < nr && c < b) ? Y[(c0 + rr) * ldy + c] : 0.0;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int o = tid + 256 * q;
            if (o < nout) {
                const int i = o / b, j = o - i * b;
                double s = 0.0;
                for (int rr = 0; rr < nr; ++rr) s += sx[rr][i] * sy[rr][j];
                acc[q] += s;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int o = tid + 256 * q;
        if (o < nout) {
            if (part) {
                part[(int64_t)blockIdx.x * nout + o] = alpha * acc[q];
            } else {
                const int i = o / b, j = o - i * b;
                atomicAdd(&C[i * ldc + j], alpha * acc[q]);
            }
        }
    }
}

// C[i][j] = sum over the nblk slices of part[blk * a b + i b + j]: one
// wavefront per output, lane l adds slices l, l + 64, ... in order, then a
// fixed shuffle tree (deterministic).
__global__ __launch_bounds__(64) void gemm_tn_reduce_kernel(const double* __restrict__ part, int64_t nblk, int a,
                                                            int b, double* __restrict__ C, int64_t ldc) {
    const int o = blockIdx.x;
    const int nout = a * b;
    double s = 0.0;
    for (int64_t k = threadIdx.x; k < nblk; k += 64) s += part[k * nout + o];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (threadIdx.x == 0) C[(int64_t)(o / b) * ldc + o % b] = s;
}

// C[r][j] = alpha sum_i X[r][i] W[i][j] + u[r] v[j] + beta C[r][j]
__global__ __launch_bounds__(256) void gemm_nn_kernel(const double* __restrict__ X, int64_t ldx, int64_t n, int a,
                                                       int b, const double* __restrict__ W, int64_t ldw,
                                                       double alpha, const double* __restrict__ u,
                                                       const double* __restrict__ v, double beta,
                                                       double* __restrict__ C, int64_t ldc) {
    __shared__ double sw[64][17];
    __shared__ double sv[16];
    for (int e = threadIdx.x; e < 64 * 16; e += 256) {
        const int i = e >> 4, j = e & 15;
        sw[i][j] = (i < a && j < b) ? W[i * ldw + j] : 0.0;
    }
    if (threadIdx.x < 16) sv[threadIdx.x] = (v && (int)threadIdx.x < b) ? v[threadIdx.x] : 0.0;
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    for (int i = 0; i < a; ++i) {
        const double x = X[r * ldx + i];
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] += x * sw[i][j];
    }
    const double ur = u ? u[r] : 0.0;
#pragma unroll
    for (int j = 0; j < 16; ++j)
        if (j < b) {
            double c = alpha * acc[j] + ur * sv[j];
            if (beta != 0.0) c += beta * C[r * ldc + j];
            C[r * ldc + j] = c;
        }
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_gemm_tn_f64(const double* X, int64_t ldx, const double* Y, int64_t ldy, int64_t n, int a, int b,
                    double alpha, double* C, int64_t ldc, double* work, void* stream) {
    if (a < 1 || b < 1 || a > 64 || b > 64 || a * b > 1024)
        return fail_msg(1, "adx_gemm_tn_f64: need 1 <= a, b <= 64 and a * b <= 1024");
    if (ldx < a || ldy < b || ldc < b) return fail_msg(1, "adx_gemm_tn_f64: bad stride");
    hipStream_t st = (hipStream_t)stream;
    int64_t nblk = 0, rpb = TN_ROWS;
    if (n > 0) {
        int64_t blocks = ceil_div(n, 4 * TN_ROWS);
        if (blocks > ADX_GEMM_TN_BLOCKS) blocks = ADX_GEMM_TN_BLOCKS;
        rpb = ceil_div(ceil_div(n, blocks), TN_ROWS) * TN_ROWS;
        nblk = ceil_div(n, rpb);
    }
    if (work && nblk > 0) {
        gemm_tn_kernel<<<(unsigned)nblk, 256, 0, st>>>(X, ldx, Y, ldy, n, a, b, alpha, rpb, C, ldc, work);
        gemm_tn_reduce_kernel<<<(unsigned)(a * b), 64, 0, st>>>(work, nblk, a, b, C, ldc);
    } else {
        // C is overwritten: zero it with a kernel (graph-safe), then accumulate.
        zero_block_kernel<<<(unsigned)ceil_div(a * b, 256), 256, 0, st>>>(C, a, b, ldc);
        if (nblk > 0)
            gemm_tn_kernel<<<(unsigned)nblk, 256, 0, st>>>(X, ldx, Y, ldy, n, a, b, alpha, rpb, C, ldc, nullptr);
    }
    ADX_LAUNCH_CHECK("adx_gemm_tn_f64");
    return 0;
}

int adx_gemm_nn_f64(const double* X, int64_t ldx, int64_t n, int a, int b, const double* W, int64_t ldw,
                    double alpha, const double* u, const double* v, double beta, double* C, int64_t ldc,
                    void* stream) {
    if (a < 1 || a > 64 || b < 1 || b > 16) return fail_msg(1, "adx_gemm_nn_f64: need 1 <= a <= 64, 1 <= b <= 16");
    if (ldx < a || ldw < b || ldc < b) return fail_msg(1, "adx_gemm_nn_f64: bad stride");
    if ((u == nullptr) != (v == nullptr)) return fail_msg(1, "adx_gemm_nn_f64: u and v go together");
    if (n <= 0) return 0;
    gemm_nn_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(X, ldx, n, a, b, W, ldw, alpha, u,
                                                                               v, beta, C, ldc);
    ADX_LAUNCH_CHECK("adx_gemm_nn_f64");
    return 0;
}

}  // extern "C"
