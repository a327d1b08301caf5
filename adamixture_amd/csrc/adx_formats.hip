// VCF and BCF genotype decoding into the device SNP-packed layout
// (adamixture/src/utils_c/cython/snp_reader.pyx:305-463 VCF, :829-1108 BCF).
//
// The host streams the decompressed file in blocks of whole records and
// hands over, per kept record (one SNP), where it lies in the block; the
// genotype fields are decoded here and written into byte row m / 4, bit
// pair m % 4 of the packed matrix, which the host fills with 0xFF (every
// field missing) beforehand.  Writes use 32-bit atomic ANDs that clear
// the bits a code does not have, so the four SNPs sharing a byte (or
// samples sharing a word) need no ordering.
//
//   vcf_gt_kernel: one wavefront per data line.  The wave walks the line 64
//     bytes at a time; a ballot of the tab characters numbers the fields, and
//     the lane holding the tab that opens sample field f >= 9 parses that
//     sample's GT (the first FORMAT subfield) with the reference's rules:
//     '.', '-' or no digits = missing; a haploid call counts 2 a (missing if
//     a > 1); a diploid call a1 + a2 in uint8 arithmetic, missing if > 2.
//     Samples a short line does not reach stay missing.
//   bcf_gt_kernel: one thread per (record, sample) over the record's GT
//     vector (int8/16/32, n values per sample): allele = (v >> 1) - 1, any
//     negative allele = missing; a1 + a2 (or a1 for n = 1) masked to 2 bits,
//     as the reference's packing does.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

// Set the 2-bit field of SNP m, sample i (currently 3) to `code`.
__device__ __forceinline__ void store_code(uint8_t* G, int64_t ldg, int64_t m, int64_t i, uint32_t code) {
    const int64_t byte = (m >> 2) * ldg + i;
    const uint32_t sh = 8u * (uint32_t)(byte & 3) + 2u * (uint32_t)(m & 3);
    const uint32_t clear = ((~code) & 3u) << sh;
    if (clear) atomicAnd(reinterpret_cast<unsigned int*>(G + (byte & ~(int64_t)3)), ~clear);
}

// Byte q of a line whose '\n' sits at e: the line itself, then '\n', then
// NUL (the reference parses NUL-terminated lines that keep their '\n').
__device__ __forceinline__ uint32_t line_char(const uint8_t* __restrict__ t, int64_t q, int64_t e) {
    return q < e ? (uint32_t)t[q] : (q == e ? 10u : 0u);
}

// snp_reader.pyx:305-341 _parse_gt_allele: 255 = missing.
__device__ __forceinline__ uint32_t vcf_allele(const uint8_t* __restrict__ t, int64_t& q, int64_t e) {
    uint32_t c = line_char(t, q, e);
    if (c == 46u || c == 45u || c == 0u) {
        if (c != 0u) ++q;
        return 255u;
    }
    const int64_t start = q;
    uint32_t val = 0;
    while (true) {
        c = line_char(t, q, e);
        if (c < 48u || c > 57u) break;
        val = (val * 10u + (c - 48u)) & 0xFFu;
        ++q;
    }
    return q == start ? 255u : val;
}

// snp_reader.pyx:343-375 _parse_gt_field_direct.
__device__ __forceinline__ uint32_t vcf_gt(const uint8_t* __restrict__ t, int64_t q, int64_t e) {
    const uint32_t a1 = vcf_allele(t, q, e);
    if (a1 == 255u) return 3u;
    const uint32_t sep = line_char(t, q, e);
    if (sep == 58u || sep == 9u || sep == 10u || sep == 0u) return a1 <= 1u ? 2u * a1 : 3u;
    ++q;
    const uint32_t a2 = vcf_allele(t, q, e);
    if (a2 == 255u) return 3u;
    const uint32_t total = (a1 + a2) & 0xFFu;
    return total <= 2u ? total : 3u;
}

__global__ __launch_bounds__(256) void vcf_gt_kernel(const uint8_t* __restrict__ text,
                                                     const int64_t* __restrict__ lstart,
                                                     const int64_t* __restrict__ lend, int64_t n_lines, int64_t m0,
                                                     int64_t N, uint8_t* __restrict__ G, int64_t ldg) {
    const int lane = threadIdx.x & 63;
    const int64_t l = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (l >= n_lines) return;
    const int64_t s = lstart[l], e = lend[l];
    const int64_t m = m0 + l;
    const uint64_t below = lane ? (~0ull >> (64 - lane)) : 0ull;
    int64_t field = 0;                       // tabs seen before this chunk (wave-uniform)
    for (int64_t base = s; base < e && field < 9 + N; base += 64) {
        const int64_t p = base + lane;
        const bool tab = p < e && text[p] == 9u;
        const uint64_t mask = __ballot(tab);
        if (tab) {
            const int64_t f = field + __popcll(mask & below) + 1;   // field that starts at p + 1
            if (f >= 9 && f - 9 < N) store_code(G, ldg, m, f - 9, vcf_gt(text, p + 1, e));
        }
        field += __popcll(mask);
    }
}

__global__ __launch_bounds__(256) void bcf_gt_kernel(const uint8_t* __restrict__ buf,
                                                     const int64_t* __restrict__ gt_off,
                                                     const int32_t* __restrict__ gt_desc, int64_t n_rec,
                                                     int64_t m0, int64_t N, uint8_t* __restrict__ G, int64_t ldg) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_rec * N) return;
    const int64_t r = e / N, i = e - r * N;
    const int32_t desc = gt_desc[r];
    uint32_t code = 3u;
    if (desc >= 0) {
        const int nv = desc & 0xFFFF, sz = desc >> 16;
        const uint8_t* p = buf + gt_off[r] + (int64_t)i * nv * sz;
        int32_t v1, v2 = 0;
        if (sz == 1) {
            v1 = (int8_t)p[0];
            if (nv >= 2) v2 = (int8_t)p[1];
        } else if (sz == 2) {
            v1 = (int16_t)(p[0] | (p[1] << 8));
            if (nv >= 2) v2 = (int16_t)(p[2] | (p[3] << 8));
        } else {
            v1 = (int32_t)((uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24));
            if (nv >= 2)
                v2 = (int32_t)((uint32_t)p[4] | ((uint32_t)p[5] << 8) | ((uint32_t)p[6] << 16) | ((uint32_t)p[7] << 24));
        }
        const int32_t a1 = (v1 >> 1) - 1;
        if (nv >= 2) {
            const int32_t a2 = (v2 >> 1) - 1;
            code = (a1 < 0 || a2 < 0) ? 3u : ((uint32_t)(a1 + a2) & 3u);
        } else {
            code = a1 < 0 ? 3u : ((uint32_t)a1 & 3u);
        }
    }
    store_code(G, ldg, m0 + r, i, code);
}

// ---- PGEN (variable-width, storage mode 0x10) -------------------------
//
// Hardcall records decoded one wavefront per variant.  Every contribution
// is an XOR on the packed field (preset 3): the dense default of all
// samples contributes 3 ^ d, each difflist entry d_i ^ v, so the field ends
// at v whatever the order of the atomics.  Record types: 0 = 2-bit array;
// 1 = bit array choosing between two values (first byte: low = b >> 2,
// high = low + (b & 3)) plus a difflist; 4 = all 0 plus a difflist;
// 2 / 3 = the previous non-LD variant (`base`, decoded by the earlier
// launch) patched by a difflist, type 3 then inverted 0 <-> 2.  Lane 0
// walks the difflist (varint length; first sample ID of every 64-entry
// group, idb bytes each; one skip byte per group but the last; 2-bit values;
// varint gaps), the other lanes fill the default meanwhile.

__device__ __forceinline__ void xor_code(uint8_t* G, int64_t ldg, int64_t m, int64_t i, uint32_t x) {
    if (!x) return;
    const int64_t byte = (m >> 2) * ldg + i;
    const uint32_t sh = 8u * (uint32_t)(byte & 3) + 2u * (uint32_t)(m & 3);
    atomicXor(reinterpret_cast<unsigned int*>(G + (byte & ~(int64_t)3)), x << sh);
}

__device__ __forceinline__ uint32_t field_at(const uint8_t* G, int64_t ldg, int64_t m, int64_t i) {
    return (G[(m >> 2) * ldg + i] >> (2 * (m & 3))) & 3u;
}

__device__ __forceinline__ uint32_t pgen_default(const uint8_t* __restrict__ rec, int type, const uint8_t* G,
                                                 int64_t ldg, int64_t base, int64_t i) {
    if (type == 0) return (rec[i >> 2] >> (2 * (i & 3))) & 3u;
    if (type == 1) {
        const uint32_t lo = rec[0] >> 2, hi = lo + (rec[0] & 3u);
        return ((rec[1 + (i >> 3)] >> (i & 7)) & 1u) ? hi : lo;
    }
    if (type == 4) return 0u;
    const uint32_t b = field_at(G, ldg, base, i);              // types 2 and 3
    return type == 3 ? (b == 3u ? 3u : 2u - b) : b;
}

// Varint at p[q] within a record of `len` bytes; q = -1 if it runs past it.
__device__ __forceinline__ uint64_t read_varint(const uint8_t* __restrict__ p, int64_t& q, int64_t len) {
    uint64_t v = 0;
    int sh = 0;
    while (q >= 0 && q < len && sh < 63) {
        const uint32_t b = p[q++];
        v |= (uint64_t)(b & 0x7Fu) << sh;
        sh += 7;
        if (b < 0x80u) return v;
    }
    q = -1;
    return 0;
}

// The host guarantees each record's dense part fits (type 0: ceil(N/4)
// bytes, type 1: 1 + ceil(N/8) + 1); the difflist is bounds-checked here
// against the record length, and a list that runs past it stops there.
__global__ __launch_bounds__(256) void pgen_kernel(const uint8_t* __restrict__ buf, const int64_t* __restrict__ rstart,
                                                   const int64_t* __restrict__ rlen, const uint8_t* __restrict__ rtype,
                                                   const int64_t* __restrict__ rbase, const int64_t* __restrict__ list,
                                                   int64_t n, int64_t N, int idb, uint8_t* __restrict__ G,
                                                   int64_t ldg) {
    const int lane = threadIdx.x & 63;
    const int64_t w = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w >= n) return;
    const int64_t m = list[w];
    const int type = rtype[m] & 7;
    const uint8_t* rec = buf + rstart[m];
    const int64_t len = rlen[m];
    const int64_t base = rbase[m];
    for (int64_t i = lane; i < N; i += 64) xor_code(G, ldg, m, i, 3u ^ pgen_default(rec, type, G, ldg, base, i));
    if (type == 0 || lane != 0) return;
    int64_t q = type == 1 ? 1 + ((N + 7) >> 3) : 0;
    const int64_t L = (int64_t)read_varint(rec, q, len);
    if (q < 0 || L == 0) return;
    const int64_t gc = (L + 63) >> 6;
    const int64_t ids0 = q;                                    // first IDs of the groups
    const int64_t vals = ids0 + idb * gc + (gc - 1);           // 2-bit values
    q = vals + ((L + 3) >> 2);                                 // varint gaps
    if (q > len) return;
    for (int64_t g = 0; g < gc; ++g) {
        int64_t id = 0;
        for (int k = 0; k < idb; ++k) id |= (int64_t)rec[ids0 + idb * g + k] << (8 * k);
        const int64_t cnt = L - 64 * g < 64 ? L - 64 * g : 64;
        for (int64_t k = 0; k < cnt; ++k) {
            if (k) {
                id += (int64_t)read_varint(rec, q, len);
                if (q < 0) return;
            }
            if (id < 0 || id >= N) return;
            const int64_t e = 64 * g + k;
            uint32_t v = (rec[vals + (e >> 2)] >> (2 * (e & 3))) & 3u;
            uint32_t d = type == 3 ? field_at(G, ldg, base, id) : pgen_default(rec, type, G, ldg, base, id);
            if (type == 3) v = v == 3u ? 3u : 2u - v;          // patched, then inverted
            if (type == 3) d = d == 3u ? 3u : 2u - d;
            xor_code(G, ldg, m, id, d ^ v);
        }
    }
}

// out field (m', i) = in field (idx[m'], i), fields of the last byte row
// beyond M_out left 0; one thread per output byte.
__global__ void select_snps_kernel(const uint8_t* __restrict__ Gin, int64_t ldg, const int64_t* __restrict__ idx,
                                   int64_t M_out, int64_t N, uint8_t* __restrict__ Gout) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t Mb = (M_out + 3) >> 2;
    if (e >= Mb * ldg) return;
    const int64_t j = e / ldg, i = e - j * ldg;
    uint32_t out = 0xFFu;
    if (i < N) {
        out = 0;
        for (int f = 0; f < 4; ++f) {
            const int64_t mo = 4 * j + f;
            if (mo < M_out) out |= field_at(Gin, ldg, idx[mo], i) << (2 * f);
        }
    }
    Gout[e] = (uint8_t)out;
}

// Fields of byte row `row` beyond SNP M - 1 set to 0 (the reference's
// packing leaves them 0), samples i < N.
__global__ void clear_tail_fields_kernel(uint8_t* __restrict__ G, int64_t ldg, int64_t row, int nf, int64_t N) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < N) G[row * ldg + i] &= (uint8_t)((1u << (2 * nf)) - 1u);
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_vcf_gt(const uint8_t* text, const int64_t* line_start, const int64_t* line_end, int64_t n_lines, int64_t m0,
               int64_t N, uint8_t* G, int64_t ldg, void* stream) {
    if (N <= 0 || ldg < N || ldg % 4 != 0 || m0 < 0) return fail_msg(1, "adx_vcf_gt: bad shape");
    if (n_lines <= 0) return 0;
    vcf_gt_kernel<<<(unsigned)ceil_div(n_lines, 4), 256, 0, (hipStream_t)stream>>>(text, line_start, line_end,
                                                                                   n_lines, m0, N, G, ldg);
    ADX_LAUNCH_CHECK("adx_vcf_gt");
    return 0;
}

int adx_bcf_gt(const uint8_t* buf, const int64_t* gt_off, const int32_t* gt_desc, int64_t n_rec, int64_t m0,
               int64_t N, uint8_t* G, int64_t ldg, void* stream) {
    if (N <= 0 || ldg < N || ldg % 4 != 0 || m0 < 0) return fail_msg(1, "adx_bcf_gt: bad shape");
    if (n_rec <= 0) return 0;
    bcf_gt_kernel<<<(unsigned)ceil_div(n_rec * N, 256), 256, 0, (hipStream_t)stream>>>(buf, gt_off, gt_desc, n_rec,
                                                                                       m0, N, G, ldg);
    ADX_LAUNCH_CHECK("adx_bcf_gt");
    return 0;
}

int adx_pgen_decode(const uint8_t* buf, const int64_t* rstart, const int64_t* rlen, const uint8_t* rtype,
                    const int64_t* rbase, const int64_t* list, int64_t n, int64_t N, int idb, uint8_t* G, int64_t ldg,
                    void* stream) {
    if (N <= 0 || ldg < N || ldg % 4 != 0) return fail_msg(1, "adx_pgen_decode: bad shape");
    if (idb < 1 || idb > 4) return fail_msg(1, "adx_pgen_decode: sample ID width must be 1..4 bytes");
    if (n <= 0) return 0;
    pgen_kernel<<<(unsigned)ceil_div(n, 4), 256, 0, (hipStream_t)stream>>>(buf, rstart, rlen, rtype, rbase, list, n,
                                                                           N, idb, G, ldg);
    ADX_LAUNCH_CHECK("adx_pgen_decode");
    return 0;
}

int adx_select_snps(const uint8_t* Gin, int64_t ldg, const int64_t* idx, int64_t M_out, int64_t N, uint8_t* Gout,
                    void* stream) {
    if (M_out <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_select_snps: bad shape");
    const int64_t total = ceil_div(M_out, 4) * ldg;
    select_snps_kernel<<<(unsigned)ceil_div(total, 256), 256, 0, (hipStream_t)stream>>>(Gin, ldg, idx, M_out, N,
                                                                                       Gout);
    ADX_LAUNCH_CHECK("adx_select_snps");
    return 0;
}

int adx_clear_tail_fields(uint8_t* G, int64_t ldg, int64_t M, int64_t N, void* stream) {
    if (M <= 0 || N <= 0 || ldg < N) return fail_msg(1, "adx_clear_tail_fields: bad shape");
    const int nf = (int)(M & 3);
    if (nf == 0) return 0;
    clear_tail_fields_kernel<<<(unsigned)ceil_div(N, 256), 256, 0, (hipStream_t)stream>>>(G, ldg, (M - 1) >> 2, nf,
                                                                                         N);
    ADX_LAUNCH_CHECK("adx_clear_tail_fields");
    return 0;
}

}  // extern "C"
