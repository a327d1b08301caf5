// Skinny FP64 products of the ALS warm start (adamixture/model/als_gpu.py:
// 90-113): every product there has one long dimension (M SNPs or N samples)
// and all others <= 64, e.g. Q^T Q (K x K over N rows), V^T I_q (L x K over
// N rows), Z (V^T I_q) (M x L times L x K).  General GEMM libraries tile for
// large square blocks; here
//   gemm_tn:  C = alpha X^T Y (a x b), X (n x a), Y (n x b): each workgroup
//             reduces a range of rows on FP64 MFMA tiles and stores its
//             a x b partial sums in its slice of the caller's workspace; one
//             more kernel sums the slices in a fixed order (deterministic;
//             without a workspace, one FP64 atomic per output and workgroup);
//   gemm_nn:  C = alpha X W + u v^T (+ beta C), X (n x a), W (a x b) small:
//             16-row tiles on FP64 MFMA with W held in registers.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

constexpr int TN_ROWS = 32;   // row granule of a gemm_tn workgroup's range

// C[i][j] = 0 for i < a, j < b (a kernel, not a memset: graph-safe).
__global__ void zero_block_kernel(double* __restrict__ C, int a, int b, int64_t ldc) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < a * b) C[(int64_t)(e / b) * ldc + e % b] = 0.0;
}

// C[i][j] (+)= alpha sum_r X[r][i] Y[r][j] on 16x16x4 FP64 MFMA tiles: the
// a x b output is TA x TB tiles of 16 x 16; lane (g, c) of a wave supplies
// X[r + g][16 ta + c] and Y[r + g][16 tb + c] for the 4-row k-step at r
// (A = X^T, B = Y, as the MFMA operand layouts), and holds the accumulator
// rows 16 ta + g + 4 j, column 16 tb + c.  The four waves of a workgroup
// take interleaved 4-row groups of its row range (together 16 consecutive
// rows per step) and are summed in a fixed order, (w0 + w2) + (w1 + w3); the
// workgroup's tiles go to its slice of part (or, without a workspace, into C
// with one FP64 atomic per element).
template <int TA, int TB>
__global__ __launch_bounds__(256) void gemm_tn_kernel(const double* __restrict__ X, int64_t ldx,
                                                       const double* __restrict__ Y, int64_t ldy, int64_t n, int a,
                                                       int b, double alpha, int64_t rows_per_block,
                                                       double* __restrict__ C, int64_t ldc,
                                                       double* __restrict__ part) {
    using MM = Mma16<double>;
    using acc_t = MM::acc;
    constexpr int NT = TA * TB;
    __shared__ double red[2][NT][256];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < n ? r0 + rows_per_block : n;
    acc_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = acc_t{0, 0, 0, 0};
    bool xok[TA], yok[TB];
#pragma unroll
    for (int ta = 0; ta < TA; ++ta) xok[ta] = 16 * ta + cl < a;
#pragma unroll
    for (int tb = 0; tb < TB; ++tb) yok[tb] = 16 * tb + cl < b;
    // two 4-row groups per iteration (rows r + g and r + 16 + g of this wave)
    for (int64_t r = r0 + 4 * wave; r < r1; r += 32) {
        double xv[2][TA], yv[2][TB];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int64_t row = r + 16 * u + g;
            const bool live = row < r1;
#pragma unroll
            for (int ta = 0; ta < TA; ++ta) xv[u][ta] = (live && xok[ta]) ? X[row * ldx + 16 * ta + cl] : 0.0;
#pragma unroll
            for (int tb = 0; tb < TB; ++tb) yv[u][tb] = (live && yok[tb]) ? Y[row * ldy + 16 * tb + cl] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int ta = 0; ta < TA; ++ta)
#pragma unroll
                for (int tb = 0; tb < TB; ++tb)
                    acc[ta * TB + tb] = MM::run(xv[u][ta], yv[u][tb], acc[ta * TB + tb]);
    }
    if (wave >= 2) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) red[wave - 2][t][64 * j + lane] = acc[t][j];
    }
    __syncthreads();
    if (wave < 2) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[t][j] += red[wave][t][64 * j + lane];
    }
    __syncthreads();
    if (wave == 1) {
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
            for (int j = 0; j < 4; ++j) red[0][t][64 * j + lane] = acc[t][j];
    }
    __syncthreads();
    if (wave != 0) return;
    const int nout = a * b;
#pragma unroll
    for (int ta = 0; ta < TA; ++ta)
#pragma unroll
        for (int tb = 0; tb < TB; ++tb)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int t = ta * TB + tb;
                const int i = 16 * ta + MM::row(g, j), jc = 16 * tb + cl;
                if (i < a && jc < b) {
                    const double v = alpha * (acc[t][j] + red[0][t][64 * j + lane]);
                    if (part)
                        part[(int64_t)blockIdx.x * nout + i * b + jc] = v;
                    else
                        atomicAdd(&C[i * ldc + jc], v);
                }
            }
}

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

// C[r][j] = alpha sum_i X[r][i] W[i][j] + u[r] v[j] + beta C[r][j] on
// 16x16x4 FP64 MFMA tiles of 16 rows: lane (g, c) supplies X[r0 + c][4 ks +
// g] (A) and holds W[4 ks + g][c] (B) for the KS k-steps (W stays in
// registers for the whole grid-stride loop); the accumulator row g + 4 j is
// row r0 + g + 4 j of C, column c, so the stores of one instruction cover
// four row segments of C.
template <int KS>
__global__ __launch_bounds__(256) void gemm_nn_kernel(const double* __restrict__ X, int64_t ldx, int64_t n, int a,
                                                       int b, const double* __restrict__ W, int64_t ldw,
                                                       double alpha, const double* __restrict__ u,
                                                       const double* __restrict__ v, double beta,
                                                       double* __restrict__ C, int64_t ldc) {
    using MM = Mma16<double>;
    using acc_t = MM::acc;
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4;
    const int cl = lane & 15;
    const bool col_ok = cl < b;
    double wv[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
        const int i = 4 * ks + g;
        wv[ks] = (i < a && col_ok) ? W[i * ldw + cl] : 0.0;
    }
    const double vc = (v && col_ok) ? v[cl] : 0.0;
    const int64_t ntiles = (n + 15) >> 4;
    const int64_t wave_id = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * 4;
    for (int64_t tile = wave_id; tile < ntiles; tile += nwaves) {
        const int64_t r0 = tile << 4;
        const int64_t rx = r0 + cl;
        double xv[KS];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const int i = 4 * ks + g;
            xv[ks] = (rx < n && i < a) ? X[rx * ldx + i] : 0.0;
        }
        acc_t acc = acc_t{0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) acc = MM::run(xv[ks], wv[ks], acc);
        if (!col_ok) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t r = r0 + MM::row(g, j);
            if (r < n) {
                double c = alpha * acc[j];
                if (u) c += u[r] * vc;
                if (beta != 0.0) c += beta * C[r * ldc + cl];
                C[r * ldc + cl] = c;
            }
        }
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
        rpb = ceil_div(ceil_div(n, blocks), TN_ROWS) * TN_ROWS;   // a multiple of the 16 rows of one step
        nblk = ceil_div(n, rpb);
    }
    const int TA = (a + 15) / 16, TB = (b + 15) / 16;   // a b <= 1024: TA TB <= 8
    auto launch = [&](double* part) {
#define ADX_TN(ta, tb)                                                                                        \
    if (TA == ta && TB == tb)                                                                                 \
        gemm_tn_kernel<ta, tb><<<(unsigned)nblk, 256, 0, st>>>(X, ldx, Y, ldy, n, a, b, alpha, rpb, C, ldc, part);
        ADX_TN(1, 1) ADX_TN(1, 2) ADX_TN(1, 3) ADX_TN(1, 4) ADX_TN(2, 1) ADX_TN(2, 2) ADX_TN(2, 3) ADX_TN(2, 4)
        ADX_TN(3, 1) ADX_TN(3, 2) ADX_TN(4, 1) ADX_TN(4, 2)
#undef ADX_TN
    };
    if (work && nblk > 0) {
        launch(work);
        gemm_tn_reduce_kernel<<<(unsigned)(a * b), 64, 0, st>>>(work, nblk, a, b, C, ldc);
    } else {
        // C is overwritten: zero it with a kernel (graph-safe), then accumulate.
        zero_block_kernel<<<(unsigned)ceil_div(a * b, 256), 256, 0, st>>>(C, a, b, ldc);
        if (nblk > 0) launch(nullptr);
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
    // 16-row tiles, 4 per workgroup, grid-stride beyond 4096 workgroups
    int64_t blocks = ceil_div(ceil_div(n, 16), 4);
    if (blocks > 4096) blocks = 4096;
    hipStream_t st = (hipStream_t)stream;
    const int ks = (a + 3) / 4;
#define ADX_NN(k) gemm_nn_kernel<k><<<(unsigned)blocks, 256, 0, st>>>(X, ldx, n, a, b, W, ldw, alpha, u, v, beta, C, ldc)
    if (ks <= 1) ADX_NN(1);
    else if (ks == 2) ADX_NN(2);
    else if (ks == 3) ADX_NN(3);
    else if (ks == 4) ADX_NN(4);
    else if (ks <= 8) ADX_NN(8);      // k-steps past a multiply zeros
    else if (ks <= 12) ADX_NN(12);
    else ADX_NN(16);
#undef ADX_NN
    ADX_LAUNCH_CHECK("adx_gemm_nn_f64");
    return 0;
}

}  // extern "C"
