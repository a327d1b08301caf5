// PLINK .bed -> device SNP-packed layout on the GPU.
//
// Reference: adamixture/src/utils_c/cython/snp_reader.pyx:76-120
// (read_bed_packed / pack_genotypes: BED code map 00 -> 2, 01 -> 3 missing,
// 10 -> 1, 11 -> 0, four SNPs per output byte), :215-251 (get_mean_packed)
// and :254-302 (flip_packed: 0 <-> 2 when the mean allele frequency is
// >= 0.5, padding fields of the last byte row cleared), driven by
// adamixture/src/snp_reader.py:353-443, :747-757.  The reference transposes
// on the host with OpenMP over samples (stride-N_bytes reads); here one
// thread builds one output byte (4 SNPs of one sample): for each of the 4
// BED rows, the 64 threads of a wave read 16 consecutive BED bytes
// (4 samples per byte), and the output row is written coalesced.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

__device__ __forceinline__ uint32_t bed_code(uint32_t two_bits) {
    // 00 -> 2, 01 -> 3, 10 -> 1, 11 -> 0  (table 0b00_01_11_10 read 2 bits at a time)
    return (0x1Eu >> (2 * two_bits)) & 3u;
}

// Output rows j = 0 .. ceil(rows / 4) - 1 of this block at G + (row0 + j) ldg;
// SNP fields beyond `rows` are 0, bytes i >= N are 0xFF (missing).
__global__ void bed_transpose_kernel(const uint8_t* __restrict__ bed, int64_t nb, int64_t rows, int64_t N,
                                     uint8_t* __restrict__ G, int64_t ldg, int64_t row0) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y + (int64_t)gridDim.y * blockIdx.z;
    if (i >= ldg || 4 * j >= rows) return;
    uint32_t out = 0xFFu;
    if (i < N) {
        out = 0;
        const int64_t byte = i >> 2;
        const int sh = 2 * (int)(i & 3);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t m = 4 * j + k;
            if (m < rows) out |= bed_code((bed[m * nb + byte] >> sh) & 3u) << (2 * k);
        }
    }
    G[(row0 + j) * ldg + i] = (uint8_t)out;
}

// out[0] += sum of observed allele counts, out[1] += 2 x observed genotypes
// (snp_reader.pyx:215-251), exact in 64-bit integers.
__global__ void packed_sum_kernel(const uint8_t* __restrict__ G, int64_t ldg, int64_t M, int64_t N,
                                  unsigned long long* __restrict__ out) {
    const int64_t Mb = (M + 3) >> 2;
    const int last = (int)(M - 4 * (Mb - 1));
    unsigned long long s = 0, c = 0;
    const int64_t total = Mb * N;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = e / N, i = e - j * N;
        const uint32_t b = G[j * ldg + i];
        const int nf = j == Mb - 1 ? last : 4;
        for (int k = 0; k < nf; ++k) {
            const uint32_t code = (b >> (2 * k)) & 3u;
            if (code != 3u) {
                s += code;
                c += 2;
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_xor(s, off);
        c += __shfl_xor(c, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&out[0], s);
        atomicAdd(&out[1], c);
    }
}

// 0 <-> 2 in every field, missing kept; fields of the last byte row beyond M
// cleared (snp_reader.pyx:254-302).  Bytes i >= N are left untouched.
__global__ void flip_kernel(uint8_t* __restrict__ G, int64_t ldg, int64_t M, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y + (int64_t)gridDim.y * blockIdx.z;
    const int64_t Mb = (M + 3) >> 2;
    if (i >= N || j >= Mb) return;
    const uint32_t b = G[j * ldg + i];
    const int nf = j == Mb - 1 ? (int)(M - 4 * (Mb - 1)) : 4;
    uint32_t out = 0;
    for (int k = 0; k < nf; ++k) {
        const uint32_t code = (b >> (2 * k)) & 3u;
        out |= (code == 3u ? 3u : 2u - code) << (2 * k);
    }
    G[j * ldg + i] = (uint8_t)out;
}

__global__ void zero_u64_pair_kernel(unsigned long long* p) {
    if (threadIdx.x < 2) p[threadIdx.x] = 0;
}

static dim3 row_grid(int64_t ncols, int64_t nrows) {
    const int64_t gy = nrows < 65535 ? nrows : 65535;
    return dim3((unsigned)ceil_div(ncols, 256), (unsigned)(gy > 0 ? gy : 1), (unsigned)ceil_div(nrows, gy > 0 ? gy : 1));
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_bed_to_packed(const uint8_t* bed, int64_t nb, int64_t rows, int64_t N, uint8_t* G, int64_t ldg,
                      int64_t row0, void* stream) {
    if (rows <= 0) return 0;
    if (N <= 0 || nb < ceil_div(N, 4) || ldg < N) return fail_msg(1, "adx_bed_to_packed: bad shape");
    const int64_t out_rows = ceil_div(rows, 4);
    bed_transpose_kernel<<<row_grid(ldg, out_rows), 256, 0, (hipStream_t)stream>>>(bed, nb, rows, N, G, ldg, row0);
    ADX_LAUNCH_CHECK("adx_bed_to_packed");
    return 0;
}

int adx_packed_allele_sum(const uint8_t* G, int64_t ldg, int64_t M, int64_t N, unsigned long long* out2,
                          void* stream) {
    if (M <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_packed_allele_sum: bad shape");
    hipStream_t st = (hipStream_t)stream;
    zero_u64_pair_kernel<<<1, 64, 0, st>>>(out2);
    const int64_t total = ceil_div(M, 4) * N;
    int64_t blocks = ceil_div(total, 256);
    if (blocks > 4096) blocks = 4096;
    packed_sum_kernel<<<(unsigned)blocks, 256, 0, st>>>(G, ldg, M, N, out2);
    ADX_LAUNCH_CHECK("adx_packed_allele_sum");
    return 0;
}

int adx_flip_packed(uint8_t* G, int64_t ldg, int64_t M, int64_t N, void* stream) {
    if (M <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_flip_packed: bad shape");
    flip_kernel<<<row_grid(N, ceil_div(M, 4)), 256, 0, (hipStream_t)stream>>>(G, ldg, M, N);
    ADX_LAUNCH_CHECK("adx_flip_packed");
    return 0;
}

}  // extern "C"
