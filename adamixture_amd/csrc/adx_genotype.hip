// Genotype-layout kernels of libadx: synthetic generator, chunk unpackers,
// allele frequencies, per-sample observation counts, padded repack, and the
// MFMA lane-layout probe used by the tests.
//
// Device layout ("SNP-packed", as produced by the reference reader,
// snp_reader.pyx:76-120): byte [j, i] holds SNPs 4j..4j+3 of sample i, SNP
// 4j+k at bits 2k..2k+1, codes 0/1/2 = allele count, 3 = missing.  Rows are
// ldg bytes apart; the kernels that stream whole rows (frequencies, counts,
// EM) require ldg % 16 == 0 and treat bytes beyond N as padding.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <mutex>
#include <string>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int fail(hipError_t err, const char* where) {
    std::string s = std::string(where) + ": " + hipGetErrorString(err);
    if (err == hipErrorOutOfMemory || err == hipErrorMemoryAllocation) s += " (out of memory)";
    set_error(s);
    return (int)err;
}

int fail_msg(int code, const std::string& msg) {
    set_error(msg);
    return code;
}

// ---------------------------------------------------------------------------
// Counter-based synthetic genotypes: G_ij ~ Binomial(2, h_ij) with
// h_ij = sum_k Q_ik P_jk and independent missingness.  Every genotype is a
// pure function of (seed, SNP, sample); the integer mixing and the fma chain
// are the ones of the CPU twin so blocks can be regenerated bit-for-bit.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

__global__ void synth_kernel(const double* __restrict__ P, const double* __restrict__ Q, int64_t M,
                             int64_t N, int K, uint64_t seed, uint32_t miss_thr, uint8_t* __restrict__ G,
                             int64_t ldg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y + (int64_t)blockIdx.z * gridDim.y;
    const int64_t Mb = (M + 3) >> 2;
    if (i >= ldg || j >= Mb) return;
    uint8_t out = 0xFF;
    if (i < N) {
        out = 0;
        const double* q = Q + i * K;
        for (int k = 0; k < 4; ++k) {
            const int64_t m = j * 4 + k;
            if (m >= M) break;
            const double* p = P + m * K;
            double h = 0.0;
            for (int c = 0; c < K; ++c) h = fma(q[c], p[c], h);
            h = h < 0.0 ? 0.0 : (h > 1.0 ? 1.0 : h);
            const uint64_t thr = (uint64_t)(h * 4294967296.0);
            const uint64_t key = seed ^ ((uint64_t)m * 0xD1B54A32D192ED03ULL) ^
                                 ((uint64_t)i * 0x8CB92BA72F3D8DD7ULL);
            const uint64_t r1 = mix64(key);
            const uint64_t r2 = mix64(key ^ 0x5851F42D4C957F2DULL);
            int code;
            if ((uint32_t)r2 < miss_thr) {
                code = 3;
            } else {
                code = ((uint64_t)(uint32_t)r1 < thr) + ((uint64_t)(uint32_t)(r1 >> 32) < thr);
            }
            out |= (uint8_t)(code << (2 * k));
        }
    }
    G[j * ldg + i] = out;
}

// ---------------------------------------------------------------------------
// Chunk unpackers (pack2bit.cu:15-77 semantics, any ldg).  byte_offset is the
// index of the first byte row present in G (host-streamed sub-slices).
// ---------------------------------------------------------------------------
__global__ void unpack_u8_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t N,
                                 int64_t m_start, int64_t rows, int64_t byte_offset,
                                 uint8_t* __restrict__ out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int64_t c = blockIdx.y; c < rows; c += gridDim.y) {
        const int64_t m = m_start + c;
        const uint8_t b = G[((m >> 2) - byte_offset) * ldg + n];
        out[c * N + n] = (b >> ((m & 3) << 1)) & 3;
    }
}

__global__ void unpack_center_kernel(const uint8_t* __restrict__ G, int64_t ldg, const float* __restrict__ f,
                                     int64_t N, int64_t m_start, int64_t rows, int64_t byte_offset,
                                     float* __restrict__ out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    for (int64_t c = blockIdx.y; c < rows; c += gridDim.y) {
        const int64_t m = m_start + c;
        const uint8_t b = G[((m >> 2) - byte_offset) * ldg + n];
        const uint32_t v = (b >> ((m & 3) << 1)) & 3;
        out[c * N + n] = (v == 3u) ? 0.0f : ((float)v - 2.0f * f[m]);
    }
}

// ---------------------------------------------------------------------------
// Allele frequencies (utils.py:781-829): one workgroup per byte row (4 SNPs),
// dword loads across samples, integer sums, f = s / (2 c) or 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void freq_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                   int64_t N, float* __restrict__ f) {
    const int64_t j = blockIdx.x;
    const uint32_t* row = reinterpret_cast<const uint32_t*>(G + j * ldg);
    uint32_t s[4] = {0, 0, 0, 0}, c[4] = {0, 0, 0, 0};
    const int64_t nw = (N + 3) >> 2;
    for (int64_t w = threadIdx.x; w < nw; w += blockDim.x) {
        uint32_t v = row[w];
        const int64_t nvalid = N - (w << 2);
        if (nvalid < 4) v |= 0xFFFFFFFFu << (8 * nvalid);  // bytes beyond N read as missing
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint32_t x = (v >> (2 * k)) & 0x03030303u;
            const uint32_t miss = x & (x >> 1) & 0x01010101u;   // code 3
            const uint32_t obs = 0x01010101u & ~miss;
            const uint32_t val = x & ~(miss * 3u);
            s[k] += __builtin_popcount(val & 0x01010101u) + 2u * __builtin_popcount(val & 0x02020202u);
            c[k] += __builtin_popcount(obs);
        }
    }
    __shared__ uint32_t red[8][256];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        red[k][threadIdx.x] = s[k];
        red[4 + k][threadIdx.x] = c[k];
    }
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off)
            for (int r = 0; r < 8; ++r) red[r][threadIdx.x] += red[r][threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x < 4) {
        const int64_t m = j * 4 + threadIdx.x;
        if (m < M) {
            const uint32_t ss = red[threadIdx.x][0], cc = red[4 + threadIdx.x][0];
            f[m] = cc > 0 ? (float)ss / (2.0f * (float)cc) : 0.0f;
        }
    }
}

// Per-sample 2 x (number of observed genotypes): the q_bat accumulator of
// em_adam_gpu.py:76 summed over all SNPs (a constant of G).
__global__ __launch_bounds__(256) void qbat_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M,
                                                   int64_t N, int64_t rows_per_block,
                                                   double* __restrict__ qbat) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int64_t Mb = (M + 3) >> 2;
    const int64_t r0 = (int64_t)blockIdx.y * rows_per_block;
    const int64_t r1 = r0 + rows_per_block < Mb ? r0 + rows_per_block : Mb;
    uint32_t cnt = 0;
    for (int64_t j = r0; j < r1; ++j) {
        uint32_t b = G[j * ldg + i];
        const int64_t nf = M - 4 * j;
        if (nf < 4) b |= 0xFFu << (2 * nf);   // SNPs beyond M read as missing
        const uint32_t miss = b & (b >> 1) & 0x55u;
        cnt += 4u - __builtin_popcount(miss);
    }
    if (cnt) atomicAdd(&qbat[i], 2.0 * (double)cnt);
}

__global__ void repack_kernel(const uint8_t* __restrict__ src, int64_t ld_src, int64_t Mb, int64_t N,
                              uint8_t* __restrict__ dst, int64_t ldg) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ldg) return;
    for (int64_t j = blockIdx.y; j < Mb; j += gridDim.y) dst[j * ldg + i] = i < N ? src[j * ld_src + i] : 0xFF;
}

// Lane-layout probe: one wave runs one 16x16x4 MFMA on A (16x4) and B (4x16)
// given row-major, using the operand mapping documented in adx_common.h,
// and writes D back row-major with the documented accumulator mapping.
// raw != 0 writes D[l * 4 + j] = accumulator element j of lane l instead.
template <typename T>
__global__ void probe_mma_kernel(const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ D, int raw) {
    const int l = threadIdx.x;
    using acc_t = typename Mma16<T>::acc;
    acc_t d = {0, 0, 0, 0};
    d = Mma16<T>::run(A[(l % 16) * 4 + l / 16], B[(l / 16) * 16 + l % 16], d);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (raw)
            D[l * 4 + j] = d[j];
        else
            D[Mma16<T>::row(l / 16, j) * 16 + (l % 16)] = d[j];
    }
}

// BF16 probe: A (16x32) and B (32x16) given row-major as FP32 (values exact
// in bf16); raw accumulators written to D[lane * 4 + j].
__global__ void probe_bf16_kernel(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
    const int l = threadIdx.x;
    bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = (__bf16)A[(l % 16) * 32 + 8 * (l / 16) + i];
        b[i] = (__bf16)B[(8 * (l / 16) + i) * 16 + l % 16];
    }
    floatx4 d = mma_bf16(a, b, floatx4{0, 0, 0, 0});
#pragma unroll
    for (int j = 0; j < 4; ++j) D[l * 4 + j] = d[j];
}

// FP16 probe of the 16x16x32 FP16 MFMA, same operand and output convention
// as the BF16 probe (values exact in fp16).
__global__ void probe_f16_kernel(const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ D) {
    typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
    const int l = threadIdx.x;
    f16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = (_Float16)A[(l % 16) * 32 + 8 * (l / 16) + i];
        b[i] = (_Float16)B[(8 * (l / 16) + i) * 16 + l % 16];
    }
    floatx4 d = __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, floatx4{0, 0, 0, 0}, 0, 0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) D[l * 4 + j] = d[j];
}

}  // namespace adx

using namespace adx;

extern "C" {

const char* adx_last_error(void) { return adx::g_err.c_str(); }

int adx_synth_packed(const double* P, const double* Q, int64_t M, int64_t N, int K, uint64_t seed,
                     double miss_rate, uint8_t* G, int64_t ldg, void* stream) {
    if (M <= 0 || N <= 0 || K <= 0 || ldg < N) return fail_msg(1, "adx_synth_packed: bad shape");
    if (!(miss_rate >= 0.0 && miss_rate < 1.0)) return fail_msg(1, "adx_synth_packed: miss_rate not in [0,1)");
    const int64_t Mb = (M + 3) / 4;
    const uint32_t miss_thr = (uint32_t)(miss_rate * 4294967296.0);
    const int64_t gy = Mb < 65535 ? Mb : 65535;
    dim3 grid((unsigned)ceil_div(ldg, 256), (unsigned)gy, (unsigned)ceil_div(Mb, gy));
    synth_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(P, Q, M, N, K, seed, miss_thr, G, ldg);
    ADX_LAUNCH_CHECK("adx_synth_packed");
    return 0;
}

int adx_unpack_u8(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, int64_t m_start, int64_t C,
                  int64_t byte_offset, uint8_t* out, void* stream) {
    const int64_t rows = C < M - m_start ? C : M - m_start;
    if (rows <= 0 || N <= 0) return 0;
    if (m_start < 0 || (m_start >> 2) < byte_offset) return fail_msg(1, "adx_unpack_u8: m_start before byte_offset");
    dim3 grid((unsigned)ceil_div(N, 256), (unsigned)(rows < 65535 ? rows : 65535));
    unpack_u8_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(G, ldg, N, m_start, rows, byte_offset, out);
    ADX_LAUNCH_CHECK("adx_unpack_u8");
    return 0;
}

int adx_unpack_center_f32(const uint8_t* G, int64_t ldg, const float* f, int64_t M, int64_t N,
                          int64_t m_start, int64_t C, int64_t byte_offset, float* out, void* stream) {
    const int64_t rows = C < M - m_start ? C : M - m_start;
    if (rows <= 0 || N <= 0) return 0;
    if (m_start < 0 || (m_start >> 2) < byte_offset)
        return fail_msg(1, "adx_unpack_center_f32: m_start before byte_offset");
    dim3 grid((unsigned)ceil_div(N, 256), (unsigned)(rows < 65535 ? rows : 65535));
    unpack_center_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(G, ldg, f, N, m_start, rows, byte_offset, out);
    ADX_LAUNCH_CHECK("adx_unpack_center_f32");
    return 0;
}

int adx_freq_f32(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, float* f, void* stream) {
    if (M <= 0 || N <= 0) return 0;
    if (ldg % 16 != 0 || ldg < N) return fail_msg(1, "adx_freq_f32: ldg must be a multiple of 16 and >= N");
    freq_kernel<<<(unsigned)ceil_div(M, 4), 256, 0, (hipStream_t)stream>>>(G, ldg, M, N, f);
    ADX_LAUNCH_CHECK("adx_freq_f32");
    return 0;
}

int adx_qbat(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, double* qbat, void* stream) {
    if (N <= 0) return 0;
    ADX_CHECK(hipMemsetAsync(qbat, 0, sizeof(double) * (size_t)N, (hipStream_t)stream));
    if (M <= 0) return 0;
    const int64_t Mb = (M + 3) / 4;
    const int64_t rpb = 512;
    dim3 grid((unsigned)ceil_div(N, 256), (unsigned)ceil_div(Mb, rpb));
    qbat_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(G, ldg, M, N, rpb, qbat);
    ADX_LAUNCH_CHECK("adx_qbat");
    return 0;
}

int adx_repack_padded(const uint8_t* src, int64_t ld_src, int64_t Mb, int64_t N, uint8_t* dst,
                      int64_t ldg, void* stream) {
    if (Mb <= 0 || N <= 0) return 0;
    if (ldg < N || ld_src < N) return fail_msg(1, "adx_repack_padded: bad strides");
    dim3 grid((unsigned)ceil_div(ldg, 256), (unsigned)(Mb < 65535 ? Mb : 65535));
    repack_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(src, ld_src, Mb, N, dst, ldg);
    ADX_LAUNCH_CHECK("adx_repack_padded");
    return 0;
}

int adx_probe_mfma(int kind, const void* A, const void* B, void* D, void* stream) {
    const int raw = kind >= 2;
    if (kind == 4 || kind == 5) {
        if (kind == 4)
            probe_bf16_kernel<<<1, 64, 0, (hipStream_t)stream>>>((const float*)A, (const float*)B, (float*)D);
        else
            probe_f16_kernel<<<1, 64, 0, (hipStream_t)stream>>>((const float*)A, (const float*)B, (float*)D);
        ADX_LAUNCH_CHECK("adx_probe_mfma");
        return 0;
    }
    if (kind == 0 || kind == 2)
        probe_mma_kernel<float><<<1, 64, 0, (hipStream_t)stream>>>((const float*)A, (const float*)B, (float*)D, raw);
    else if (kind == 1 || kind == 3)
        probe_mma_kernel<double><<<1, 64, 0, (hipStream_t)stream>>>((const double*)A, (const double*)B, (double*)D,
                                                                     raw);
    else
        return fail_msg(1, "adx_probe_mfma: kind must be 0..5");
    ADX_LAUNCH_CHECK("adx_probe_mfma");
    return 0;
}

}  // extern "C"
