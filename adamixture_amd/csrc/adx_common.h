// Shared helpers for the gfx950 kernels of libadx.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace adx {

// Last error message of the calling host thread (returned by adx_last_error).
void set_error(const std::string& msg);

// Translate a HIP status into the library status.  Out-of-memory keeps the
// phrase "out of memory" in the message so Python callers can classify it.
int fail(hipError_t err, const char* where);
int fail_msg(int code, const std::string& msg);

#define ADX_CHECK(expr)                                   \
    do {                                                  \
        hipError_t _e = (expr);                           \
        if (_e != hipSuccess) return adx::fail(_e, #expr); \
    } while (0)

#define ADX_LAUNCH_CHECK(where)                            \
    do {                                                   \
        hipError_t _e = hipGetLastError();                 \
        if (_e != hipSuccess) return adx::fail(_e, where); \
    } while (0)

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Tile configuration of the EM passes for one dtype tag (ADX_F32, ADX_F64,
// ADX_BF16): 16-SNP column tiles per wave, waves per workgroup (4 or 8; a
// workgroup owns 16 * cw * waves SNPs and writes one partial T slice per
// sample for them), and the target number of workgroups (the sample range
// is split until the grid reaches it).  Process-wide, set by
// adx_set_em_tiles (the autotuner); read at every launch.
struct EmTiles {
    int cw;
    int waves;
    int64_t target_wgs;
};
EmTiles em_tiles(int dtype);

// Sum of the WV per-wave LDS slices s[0..WV-1][e] in a fixed order: adjacent
// pairs first, ((s0 + s1) + (s2 + s3)) + ((s4 + s5) + (s6 + s7)).
template <int WV, typename T, int LEN>
__device__ __forceinline__ T sum_wave_slices(const T (*s)[LEN], int e) {
    static_assert(WV == 4 || WV == 8, "4 or 8 waves per workgroup");
    const T a = (s[0][e] + s[1][e]) + (s[2][e] + s[3][e]);
    if constexpr (WV == 4) {
        return a;
    } else {
        return a + ((s[4][e] + s[5][e]) + (s[6][e] + s[7][e]));
    }
}

// One 16x16x4 MFMA step.  Operand layouts are shared by FP32 and FP64:
//   A (16x4):  lane l supplies A[l % 16][l / 16]
//   B (4x16):  lane l supplies B[l / 16][l % 16]
// The accumulator layouts differ; with g = l / 16 and j = 0..3 lane l holds
//   FP32: D[4 g + j][l % 16]        FP64: D[g + 4 j][l % 16]
// row(g, j) gives that row index (pinned by the adx_probe_mfma test).
template <typename T>
struct Mma16;

template <>
struct Mma16<float> {
    using acc = floatx4;
    __device__ static inline acc run(float a, float b, acc c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __host__ __device__ static constexpr int row(int g, int j) { return 4 * g + j; }
};

template <>
struct Mma16<double> {
    using acc = doublex4;
    __device__ static inline acc run(double a, double b, acc c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __host__ __device__ static constexpr int row(int g, int j) { return g + 4 * j; }
};

// sqp.pyx:112-146 / sqp_kernel.cu:125-166 project_q_simplex_row, serial:
// out = the Euclidean projection of x onto {v >= pc, sum v = 1}, by sorting
// x - pc in descending order (insertion sort, K <= 64) and finding the
// threshold.  b (K doubles) is scratch; out may alias x but not b.
__device__ inline void project_simplex_row(const double* x, double* out, double* b, int K, double pc) {
    int idx[64];
    for (int i = 0; i < K; ++i) {
        b[i] = x[i] - pc;
        idx[i] = i;
    }
    for (int i = 1; i < K; ++i) {
        const int key = idx[i];
        const double kv = b[key];
        int j = i - 1;
        while (j >= 0 && b[idx[j]] < kv) {
            idx[j + 1] = idx[j];
            --j;
        }
        idx[j + 1] = key;
    }
    const double tau = 1.0 - K * pc;
    double tsum = 0.0, tmax = 0.0;
    bool bget = false;
    for (int i = 0; i < K - 1; ++i) {
        tsum += b[idx[i]];
        tmax = (tsum - tau) / (i + 1);
        if (tmax >= b[idx[i + 1]]) {
            bget = true;
            break;
        }
    }
    if (!bget) tmax = (tsum + b[idx[K - 1]] - tau) / K;
    for (int i = 0; i < K; ++i) {
        const double v = b[i] - tmax;
        out[i] = (v < 0.0 ? 0.0 : v) + pc;
    }
}

// Log-likelihood of em_adam_gpu.py:238-246 (H clamped to [1e-10, 1-1e-10],
// g log H + (2-g) log(1-H) over observed genotypes) evaluated without a log
// per genotype: each lane multiplies the per-genotype factors
//   g = 0: (1-H)^2,  g = 1: H (1-H),  g = 2: H^2,  missing: 1
// into an FP64 mantissa in [1, 2) and an integer binary exponent,
// renormalised after at most 8 factors (each >= 1e-20, so no underflow), and
// takes a single log at the end.  The result differs from the sum of FP64
// logs only by the product's roundings (<= 1.1e-16 each, random-walk
// accumulated), which is below the FP64 summation error of that sum.
//
// The factor is formed without selects: with s1 = [code != 2] and
// s2 = [code is 0 or 3] it is |H' - s1| |H' - s2|, where |H - 1| is exactly
// the rounded 1 - H (round to nearest is symmetric) and |H - 0| = H, and
// H' is H with its high word cleared for a missing code: a subnormal, so
// that both differences are exactly -1 and the factor exactly 1.  Every
// factor equals the select form's f1 * f2 (or 1) bit for bit.
struct CodeMasks {
    // bit 2i of each word describes the 2-bit code in field i of a genotype word
    uint32_t obs;  // code != 3 (observed)
    uint32_t s1;   // code != 2
    uint32_t s2;   // code == 0 or code == 3
};
__device__ __forceinline__ CodeMasks code_masks(uint32_t w) {
    const uint32_t hi = w >> 1;
    return CodeMasks{~(w & hi), ~(hi & ~w), ~(w ^ hi)};
}

struct LogProd {
    double m = 1.0;
    int64_t e = 0;
    // The factor of the genotype whose code sits at bits p, p + 1 of the
    // word cm was made from.  CLAMP = false: the caller knows the clamp is a
    // no-op (H in range).
    template <bool CLAMP = true>
    __device__ __forceinline__ void mul(double h, const CodeMasks& cm, int p) {
        const double hc = CLAMP ? fmin(fmax(h, 1e-10), 1.0 - 1e-10) : h;
        const uint32_t keep = (uint32_t)((int32_t)(cm.obs << (31 - p)) >> 31);
        const double hk = __longlong_as_double(
            (long long)((uint64_t)__double_as_longlong(hc) & (((uint64_t)keep << 32) | 0xFFFFFFFFull)));
        const double s1 = (double)((cm.s1 >> p) & 1u);
        const double s2 = (double)((cm.s2 >> p) & 1u);
        m *= fabs(hk - s1) * fabs(hk - s2);
    }
    // a factor of another product of the same form (>= 1e-20)
    __device__ __forceinline__ void mul_factor(double f) { m *= f; }
    __device__ __forceinline__ void renorm() {
        const uint64_t bits = (uint64_t)__double_as_longlong(m);
        e += (int64_t)((bits >> 52) & 0x7FF) - 1023;
        m = __longlong_as_double((long long)((bits & 0x800FFFFFFFFFFFFFULL) | 0x3FF0000000000000ULL));
    }
    __device__ __forceinline__ double log_value() const {
        // e * ln2 with ln2 split so that e * ln2_hi is exact for |e| < 2^29.
        const double ln2_hi = 0.6931471824645996, ln2_lo = -1.904654299957768e-09;
        return (double)e * ln2_hi + ((double)e * ln2_lo + log(m));
    }
};

// Inverse of row(): the lane group g and element j holding row r.
template <typename T>
__host__ __device__ constexpr int row_group(int r) {
    return Mma16<T>::row(1, 0) == 4 ? r >> 2 : r & 3;
}
template <typename T>
__host__ __device__ constexpr int row_elem(int r) {
    return Mma16<T>::row(1, 0) == 4 ? r & 3 : r >> 2;
}

}  // namespace adx

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

namespace adx {

// 16x16x32 BF16 MFMA with FP32 accumulation (gfx950):
//   A (16x32): lane l supplies A[l % 16][8 (l / 16) + i], i = 0..7
//   B (32x16): lane l supplies B[8 (l / 16) + i][l % 16]
//   D (16x16): lane l holds D[4 (l / 16) + j][l % 16]   (as FP32 16x16x4)
__device__ inline floatx4 mma_bf16(bf16x8 a, bf16x8 b, floatx4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

}  // namespace adx
