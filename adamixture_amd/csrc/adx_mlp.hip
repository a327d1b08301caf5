// Kernels of the learned Q warm start: a 3-layer MLP (64 -> 256 -> 256 -> 16,
// ReLU) that maps per-sample randomized-SVD features to a Q row.  It has no
// counterpart in the reference, which starts from SVD + ALS only
// (adamixture/src/adamixture.py:185-250); SVD + ALS stays the fallback.
//
// Training (FP32):
//   gemm          C (+)= alpha op(A) op(B): 64x64 output tiles, 16-deep k
//                 panels in LDS, 4x4 outputs per thread; split-k with FP32
//                 atomics for the weight gradients (reduction over the batch).
//   bias_act      Z += b, optional ReLU;   relu_bwd   dZ = dY (Z > 0)
//   colsum        db += 1^T dY (atomic)
//   softmax_xent  z = logits + log(max(prior, 1e-5)) over the first K
//                 columns, p = softmax(z), loss = -sum t log p,
//                 dlogits = (p - t) / rows
//   adam          Adam on the flat FP32 parameter vector
// Inference (BF16): one fused kernel, 64 samples per workgroup, weights
// rounded once to BF16 in an n-major layout (pack), activations BF16 in LDS,
// 16x16x32 BF16 MFMA with FP32 accumulation, the softmax, the [1e-5, 1-1e-5]
// clamp and the row renormalisation of the reference's Q map
// (adamixture/model/als_gpu.py:23-36) in the epilogue; Q is written in FP64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

constexpr int TM = 64, TN = 64, TK = 16;
constexpr int D0 = ADX_MLP_IN, D1 = ADX_MLP_WIDTH, D2 = ADX_MLP_WIDTH, D3 = ADX_MLP_OUT;

// op(A) is m x k and op(B) is k x n; element (r, c) of op(X) is X[r * ld + c]
// without transpose and X[c * ld + r] with it.  Panel loads run along the
// contiguous index of the stored matrix so that they coalesce.
__global__ __launch_bounds__(256) void gemm_kernel(int ta, int tb, int64_t m, int64_t n, int64_t k, float alpha,
                                                   const float* __restrict__ A, int64_t lda,
                                                   const float* __restrict__ B, int64_t ldb, float beta,
                                                   float* __restrict__ C, int64_t ldc, int64_t k_per_split,
                                                   int atomic_out) {
    __shared__ float sa[TK][TM + 1];
    __shared__ float sb[TK][TN + 1];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int64_t r0 = (int64_t)blockIdx.y * TM, c0 = (int64_t)blockIdx.x * TN;
    const int64_t k0 = (int64_t)blockIdx.z * k_per_split;
    const int64_t k1 = k0 + k_per_split < k ? k0 + k_per_split : k;
    float acc[4][4] = {};
    for (int64_t kb = k0; kb < k1; kb += TK) {
        for (int e = threadIdx.x; e < TK * TM; e += 256) {
            int kk, rr;
            if (ta) { kk = e / TM; rr = e - kk * TM; } else { rr = e / TK; kk = e - rr * TK; }
            const int64_t r = r0 + rr, kx = kb + kk;
            float v = 0.0f;
            if (r < m && kx < k1) v = ta ? A[kx * lda + r] : A[r * lda + kx];
            sa[kk][rr] = v;
        }
        for (int e = threadIdx.x; e < TK * TN; e += 256) {
            int kk, cc;
            if (tb) { cc = e / TK; kk = e - cc * TK; } else { kk = e / TN; cc = e - kk * TN; }
            const int64_t c = c0 + cc, kx = kb + kk;
            float v = 0.0f;
            if (c < n && kx < k1) v = tb ? B[c * ldb + kx] : B[kx * ldb + c];
            sb[kk][cc] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TK; ++kk) {
            float a[4], b[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) a[i] = sa[kk][ty + 16 * i];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = sb[kk][tx + 16 * j];
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int64_t r = r0 + ty + 16 * i, c = c0 + tx + 16 * j;
            if (r < m && c < n) {
                float* out = C + r * ldc + c;
                if (atomic_out)
                    atomicAdd(out, alpha * acc[i][j]);
                else
                    *out = alpha * acc[i][j] + (beta != 0.0f ? beta * *out : 0.0f);
            }
        }
}

__global__ void bias_act_kernel(float* __restrict__ Z, const float* __restrict__ b, int64_t rows, int64_t cols,
                                int64_t ld, int relu) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const int64_t r = e / cols, c = e - r * cols;
    float v = Z[r * ld + c] + b[c];
    if (relu && v < 0.0f) v = 0.0f;
    Z[r * ld + c] = v;
}

__global__ void relu_bwd_kernel(float* __restrict__ dY, const float* __restrict__ Z, int64_t n) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n && Z[e] <= 0.0f) dY[e] = 0.0f;
}

__global__ void colsum_kernel(const float* __restrict__ dY, int64_t rows, int64_t cols, float* __restrict__ db) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= cols) return;
    float s = 0.0f;
    for (int64_t r = blockIdx.y; r < rows; r += gridDim.y) s += dY[r * cols + c];
    atomicAdd(&db[c], s);
}

__global__ void softmax_xent_kernel(const float* __restrict__ logits, int64_t ldz, const float* __restrict__ prior,
                                    int64_t ldp, const float* __restrict__ target, int64_t ldt,
                                    const int* __restrict__ kvec, int64_t rows, float* __restrict__ probs,
                                    float* __restrict__ dlogits, float* __restrict__ loss) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const int K = kvec[r];
    float z[D3];
    float mx = -INFINITY;
#pragma unroll
    for (int c = 0; c < D3; ++c) {
        z[c] = logits[r * ldz + c] + logf(fmaxf(prior[r * ldp + c], 1e-5f));
        if (c < K) mx = fmaxf(mx, z[c]);
    }
    float s = 0.0f;
#pragma unroll
    for (int c = 0; c < D3; ++c) s += c < K ? expf(z[c] - mx) : 0.0f;
    const float ls = logf(s);
    float l = 0.0f;
#pragma unroll
    for (int c = 0; c < D3; ++c) {
        float p = 0.0f, t = 0.0f;
        if (c < K) {
            const float lp = z[c] - mx - ls;
            p = expf(lp);
            if (target) {
                t = target[r * ldt + c];
                l -= t * lp;
            }
        }
        if (probs) probs[r * D3 + c] = p;
        if (dlogits) dlogits[r * D3 + c] = c < K ? (p - t) / (float)rows : 0.0f;
    }
    if (loss) loss[r] = l;
}

__global__ void adam_flat_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                 float* __restrict__ v, int64_t n, float lr, float b1, float b2, float eps,
                                 float bc1, float bc2) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const float gg = g[e];
    const float mm = b1 * m[e] + (1.0f - b1) * gg;
    const float vv = b2 * v[e] + (1.0f - b2) * gg * gg;
    m[e] = mm;
    v[e] = vv;
    w[e] -= lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
}

// Offsets (elements) of the flat FP32 parameter vector:
//   W1 [D0][D1] | b1 [D1] | W2 [D1][D2] | b2 [D2] | W3 [D2][D3] | b3 [D3]
constexpr int64_t OW1 = 0, OB1 = OW1 + D0 * D1, OW2 = OB1 + D1, OB2 = OW2 + D1 * D2, OW3 = OB2 + D2,
                  OB3 = OW3 + D2 * D3, NPARAM = OB3 + D3;
// Packed inference layout: BF16 W1t [D1][D0] | W2t [D2][D1] | W3t [D3][D2],
// then FP32 b1 | b2 | b3 (byte offsets).
constexpr int64_t PW1 = 0, PW2 = PW1 + 2 * D1 * D0, PW3 = PW2 + 2 * D2 * D1, PB = PW3 + 2 * D3 * D2,
                  PACKED_BYTES = PB + 4 * (D1 + D2 + D3);
static_assert(NPARAM == ADX_MLP_NPARAM, "parameter count");
static_assert(PACKED_BYTES == ADX_MLP_PACKED_BYTES, "packed size");

__global__ void pack_kernel(const float* __restrict__ w, uint8_t* __restrict__ packed) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    __bf16* w1t = (__bf16*)(packed + PW1);
    __bf16* w2t = (__bf16*)(packed + PW2);
    __bf16* w3t = (__bf16*)(packed + PW3);
    float* bias = (float*)(packed + PB);
    if (e < D0 * D1) {
        const int nn = e / D0, kk = e - nn * D0;
        w1t[e] = (__bf16)w[OW1 + kk * D1 + nn];
    } else if (e < D0 * D1 + D1 * D2) {
        const int64_t f = e - D0 * D1;
        const int nn = f / D1, kk = f - nn * D1;
        w2t[f] = (__bf16)w[OW2 + kk * D2 + nn];
    } else if (e < D0 * D1 + D1 * D2 + D2 * D3) {
        const int64_t f = e - D0 * D1 - D1 * D2;
        const int nn = f / D2, kk = f - nn * D2;
        w3t[f] = (__bf16)w[OW3 + kk * D3 + nn];
    } else if (e < D0 * D1 + D1 * D2 + D2 * D3 + D1 + D2 + D3) {
        const int64_t f = e - D0 * D1 - D1 * D2 - D2 * D3;
        bias[f] = f < D1 ? w[OB1 + f] : (f < D1 + D2 ? w[OB2 + f - D1] : w[OB3 + f - D1 - D2]);
    }
}

constexpr int XS = D0 + 8, HS = D1 + 8;  // LDS row strides (BF16 elements)

// One hidden layer on MFMA: out[64][ND] = relu(in[64][KD] Wt^T + b).
// Wave w owns output columns [w ND/4, (w+1) ND/4) for all 4 row tiles.
template <int KD, int ND, int LDI>
__device__ __forceinline__ void hidden_layer(const __bf16* __restrict__ in, const __bf16* __restrict__ Wt,
                                             const float* __restrict__ b, __bf16* __restrict__ out, int wave,
                                             int lane) {
    constexpr int CT = ND / 64;
    const int lr = lane & 15, lg = lane >> 4;
    floatx4 acc[4][CT];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 2
    for (int kb = 0; kb < KD / 32; ++kb) {
        bf16x8 a[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) a[rt] = *(const bf16x8*)(in + (16 * rt + lr) * LDI + 32 * kb + 8 * lg);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) {
            const int nn = 16 * (wave * CT + ct) + lr;
            const bf16x8 bv = *(const bf16x8*)(Wt + (int64_t)nn * KD + 32 * kb + 8 * lg);
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) acc[rt][ct] = mma_bf16(a[rt], bv, acc[rt][ct]);
        }
    }
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
        const int col = 16 * (wave * CT + ct) + lr;
        const float bb = b[col];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float v = fmaxf(acc[rt][ct][j] + bb, 0.0f);
                out[(16 * rt + 4 * lg + j) * HS + col] = (__bf16)v;
            }
    }
}

__global__ __launch_bounds__(256) void mlp_infer_kernel(const float* __restrict__ X, int64_t ldx, int64_t N, int K,
                                                        const uint8_t* __restrict__ packed,
                                                        double* __restrict__ Q, int64_t ldq) {
    __shared__ __attribute__((aligned(16))) __bf16 sx[64 * XS];
    __shared__ __attribute__((aligned(16))) __bf16 sh1[64 * HS];
    __shared__ __attribute__((aligned(16))) __bf16 sh2[64 * HS];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int64_t r0 = (int64_t)blockIdx.x * 64;
    for (int e = tid; e < 64 * D0; e += 256) {
        const int rr = e / D0, c = e - rr * D0;
        const int64_t r = r0 + rr;
        sx[rr * XS + c] = (__bf16)(r < N ? X[r * ldx + c] : 0.0f);
    }
    __syncthreads();
    const __bf16* w1t = (const __bf16*)(packed + PW1);
    const __bf16* w2t = (const __bf16*)(packed + PW2);
    const __bf16* w3t = (const __bf16*)(packed + PW3);
    const float* bias = (const float*)(packed + PB);
    hidden_layer<D0, D1, XS>(sx, w1t, bias, sh1, wave, lane);
    __syncthreads();
    hidden_layer<D1, D2, HS>(sh1, w2t, bias + D1, sh2, wave, lane);
    __syncthreads();
    // Output layer: wave w owns rows 16w..16w+15, all D3 = 16 columns.
    const int lr = lane & 15, lg = lane >> 4;
    floatx4 acc = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < D2 / 32; ++kb) {
        const bf16x8 a = *(const bf16x8*)(sh2 + (16 * wave + lr) * HS + 32 * kb + 8 * lg);
        const bf16x8 bv = *(const bf16x8*)(w3t + lr * D2 + 32 * kb + 8 * lg);
        acc = mma_bf16(a, bv, acc);
    }
    const int c = lr;
    const float b3 = bias[D1 + D2 + c];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int64_t r = r0 + 16 * wave + 4 * lg + j;
        const bool live = r < N;
        const float prior = live ? X[r * ldx + c] : 1.0f;
        float z = c < K ? acc[j] + b3 + logf(fmaxf(prior, 1e-5f)) : -INFINITY;
        float mx = z;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        float p = c < K ? expf(z - mx) : 0.0f;
        float s = p;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o);
        p = c < K ? fminf(fmaxf(p / s, 1e-5f), 1.0f - 1e-5f) : 0.0f;
        // renormalise in FP64 so that the stored row sums to 1 in FP64
        double pd = (double)p, sd = pd;
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) sd += __shfl_xor(sd, o);
        if (live && c < K) Q[r * ldq + c] = pd / sd;
    }
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_mlp_gemm_f32(int ta, int tb, int64_t m, int64_t n, int64_t k, float alpha, const float* A, int64_t lda,
                     const float* B, int64_t ldb, float beta, float* C, int64_t ldc, int splits, void* stream) {
    if (m <= 0 || n <= 0) return 0;
    if (k < 0) return fail_msg(1, "adx_mlp_gemm_f32: k < 0");
    if (splits < 1) splits = 1;
    const int64_t kps = k > 0 ? ceil_div(ceil_div(k, splits), TK) * TK : TK;
    const int64_t nz = k > 0 ? ceil_div(k, kps) : 1;
    const int atomic_out = nz > 1 ? 1 : 0;
    if (atomic_out && beta != 1.0f)
        return fail_msg(1, "adx_mlp_gemm_f32: split-k accumulates into C and needs beta = 1");
    if (ceil_div(n, TN) > 0x7fffffff || ceil_div(m, TM) > 65535 || nz > 65535)
        return fail_msg(1, "adx_mlp_gemm_f32: grid too large");
    dim3 grid((unsigned)ceil_div(n, TN), (unsigned)ceil_div(m, TM), (unsigned)nz);
    gemm_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, kps,
                                                      atomic_out);
    ADX_LAUNCH_CHECK("adx_mlp_gemm_f32");
    return 0;
}

int adx_mlp_bias_act(float* Z, const float* b, int64_t rows, int64_t cols, int64_t ld, int relu, void* stream) {
    if (rows <= 0 || cols <= 0) return 0;
    bias_act_kernel<<<(unsigned)ceil_div(rows * cols, 256), 256, 0, (hipStream_t)stream>>>(Z, b, rows, cols, ld,
                                                                                          relu);
    ADX_LAUNCH_CHECK("adx_mlp_bias_act");
    return 0;
}

int adx_mlp_relu_bwd(float* dY, const float* Z, int64_t n, void* stream) {
    if (n <= 0) return 0;
    relu_bwd_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(dY, Z, n);
    ADX_LAUNCH_CHECK("adx_mlp_relu_bwd");
    return 0;
}

int adx_mlp_colsum(const float* dY, int64_t rows, int64_t cols, float* db, void* stream) {
    if (rows <= 0 || cols <= 0) return 0;
    const int64_t gy = rows < 512 ? rows : 512;
    colsum_kernel<<<dim3((unsigned)ceil_div(cols, 64), (unsigned)gy), 64, 0, (hipStream_t)stream>>>(dY, rows, cols,
                                                                                                   db);
    ADX_LAUNCH_CHECK("adx_mlp_colsum");
    return 0;
}

int adx_mlp_softmax_xent(const float* logits, int64_t ldz, const float* prior, int64_t ldp, const float* target,
                         int64_t ldt, const int* kvec, int64_t rows, float* probs, float* dlogits, float* loss,
                         void* stream) {
    if (rows <= 0) return 0;
    if (ldz < D3 || ldp < D3 || (target && ldt < D3)) return fail_msg(1, "adx_mlp_softmax_xent: stride < 16");
    softmax_xent_kernel<<<(unsigned)ceil_div(rows, 256), 256, 0, (hipStream_t)stream>>>(
        logits, ldz, prior, ldp, target, ldt, kvec, rows, probs, dlogits, loss);
    ADX_LAUNCH_CHECK("adx_mlp_softmax_xent");
    return 0;
}

int adx_mlp_adam(float* w, const float* g, float* m, float* v, int64_t n, float lr, float beta1, float beta2,
                 float eps, int t, void* stream) {
    if (n <= 0) return 0;
    if (t < 1) return fail_msg(1, "adx_mlp_adam: t must be >= 1");
    const float bc1 = 1.0f - powf(beta1, (float)t), bc2 = 1.0f - powf(beta2, (float)t);
    adam_flat_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(w, g, m, v, n, lr, beta1, beta2,
                                                                                 eps, bc1, bc2);
    ADX_LAUNCH_CHECK("adx_mlp_adam");
    return 0;
}

int adx_mlp_pack_bf16(const float* params, void* packed, void* stream) {
    const int64_t total = D0 * D1 + D1 * D2 + D2 * D3 + D1 + D2 + D3;
    pack_kernel<<<(unsigned)ceil_div(total, 256), 256, 0, (hipStream_t)stream>>>(params, (uint8_t*)packed);
    ADX_LAUNCH_CHECK("adx_mlp_pack_bf16");
    return 0;
}

int adx_mlp_infer_bf16(const float* X, int64_t ldx, int64_t N, int K, const void* packed, double* Q, int64_t ldq,
                       void* stream) {
    if (N <= 0) return 0;
    if (K < 2 || K > D3) return fail_msg(1, "adx_mlp_infer_bf16: K must be in [2, 16]");
    if (ldx < D0 || ldq < K) return fail_msg(1, "adx_mlp_infer_bf16: bad stride");
    if (((uintptr_t)packed & 15) != 0) return fail_msg(1, "adx_mlp_infer_bf16: packed weights not 16-byte aligned");
    mlp_infer_kernel<<<(unsigned)ceil_div(N, 64), 256, 0, (hipStream_t)stream>>>(X, ldx, N, K,
                                                                                (const uint8_t*)packed, Q, ldq);
    ADX_LAUNCH_CHECK("adx_mlp_infer_bf16");
    return 0;
}

}  // extern "C"
