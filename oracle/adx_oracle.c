/*
 * adx_oracle.c -- CPU restatement of ADAMIXTURE's likelihood hot path, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * adamixture_amd/csrc.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load the shared objects built from it; the product path
 * never links, calls or falls back to it.
 *
 * Every function cites the reference file:line whose behaviour it restates
 * (paths relative to the reference repository root).  Where the reference's
 * CPU path (Cython, adamixture/src/utils_c/cython/*.pyx) and its GPU path
 * (adamixture/model/*_gpu.py + src/utils_c/cuda/*.cu) compute different
 * things, both variants are given: *_gpu follows the GPU semantics that the
 * `--device gpu` drop-in must reproduce, *_cpu follows the Cython semantics
 * that the CPU fixtures were produced with.
 *
 * Build: see oracle/Makefile (liboracle.so for checking, liboracle_fast.so
 * with the reference's compile flags -O3 -ffast-math -fno-wrapv -fopenmp for
 * the CPU baseline timing).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

static inline double clip_domain(double v) {
    /* em.pyx:306-317 _clip_to_domain */
    return fmin(fmax(v, 1e-5), 1.0 - 1e-5);
}

static inline int code_at(const uint8_t *Gp, i64 ldg, i64 m, i64 i) {
    /* pack2bit.cu:33-40: SNP m of sample i lives in byte row m/4 at bits 2*(m%4) */
    return (Gp[(m >> 2) * ldg + i] >> ((m & 3) << 1)) & 3;
}

/* ------------------------------------------------------------------------ */
/* 1. Genotype layout                                                        */
/* ------------------------------------------------------------------------ */

/* snp_reader.pyx:76-120 read_bed_packed: PLINK SNP-major BED (4 samples per
 * byte, code map {00->2, 01->3, 10->1, 11->0}) -> SNP-packed Mb x N. */
void oc_bed_to_packed(const uint8_t *bed, i64 M, i64 N_bytes, i64 N, uint8_t *Gp, i64 ldg) {
    static const uint8_t lut[4] = {2, 3, 1, 0};
    i64 Mb = (M + 3) / 4;
    for (i64 j = 0; j < Mb; ++j) {
        for (i64 i = 0; i < N; ++i) {
            uint8_t out = 0;
            for (int k = 0; k < 4; ++k) {
                i64 snp = j * 4 + k;
                if (snp >= M) break;
                uint8_t b = bed[snp * N_bytes + i / 4];
                uint8_t v = lut[(b >> ((i % 4) * 2)) & 3];
                out |= (uint8_t)((v & 3) << (2 * k));
            }
            Gp[j * ldg + i] = out;
        }
    }
}

/* snp_reader.pyx:35-73 read_bed: BED -> unpacked M x N uint8. */
void oc_bed_to_unpacked(const uint8_t *bed, i64 M, i64 N_bytes, i64 N, uint8_t *G) {
    static const uint8_t lut[4] = {2, 3, 1, 0};
    for (i64 m = 0; m < M; ++m)
        for (i64 i = 0; i < N; ++i)
            G[m * N + i] = lut[(bed[m * N_bytes + i / 4] >> ((i % 4) * 2)) & 3];
}

/* snp_reader.pyx:123-156 pack_genotypes: unpacked M x N -> SNP-packed. */
void oc_pack(const uint8_t *G, i64 M, i64 N, uint8_t *Gp, i64 ldg) {
    i64 Mb = (M + 3) / 4;
    for (i64 j = 0; j < Mb; ++j)
        for (i64 i = 0; i < N; ++i) {
            uint8_t out = 0;
            for (int k = 0; k < 4; ++k) {
                i64 snp = (j << 2) | k;
                if (snp < M) out |= (uint8_t)((G[snp * N + i] & 3) << (k << 1));
            }
            Gp[j * ldg + i] = out;
        }
}

/* pack2bit.cu:15-44 unpack2bit_chunk_kernel_uint8_SNPsPacked (byte_offset 0):
 * out[c, n] = code of SNP m_start + c, for c < min(C, M - m_start). */
void oc_unpack(const uint8_t *Gp, i64 ldg, i64 M, i64 N, i64 m_start, i64 C, uint8_t *out) {
    i64 rows = C < M - m_start ? C : M - m_start;
    for (i64 c = 0; c < rows; ++c)
        for (i64 i = 0; i < N; ++i)
            out[c * N + i] = (uint8_t)code_at(Gp, ldg, m_start + c, i);
}

/* pack2bit.cu:46-77 unpack2bit_chunk_kernel_center_SNPsPacked:
 * out = 0 if missing else g - 2 f[m], FP32. */
void oc_unpack_center(const uint8_t *Gp, i64 ldg, const float *f, i64 M, i64 N, i64 m_start,
                      i64 C, float *out) {
    i64 rows = C < M - m_start ? C : M - m_start;
    for (i64 c = 0; c < rows; ++c) {
        i64 m = m_start + c;
        for (i64 i = 0; i < N; ++i) {
            int g = code_at(Gp, ldg, m, i);
            out[c * N + i] = (g == 3) ? 0.0f : ((float)g - 2.0f * f[m]);
        }
    }
}

/* snp_reader.pyx:215-251 get_mean_packed. */
double oc_mean_packed(const uint8_t *Gp, i64 ldg, i64 M, i64 N) {
    uint64_t s = 0, c = 0;
    for (i64 m = 0; m < M; ++m)
        for (i64 i = 0; i < N; ++i) {
            int g = code_at(Gp, ldg, m, i);
            if (g != 3) { s += (uint64_t)g; c += 2; }
        }
    return c == 0 ? 0.0 : (double)s / (double)c;
}

/* snp_reader.pyx:254-302 flip_packed: 0<->2 with the last byte's padding
 * fields zeroed. */
void oc_flip_packed(uint8_t *Gp, i64 ldg, i64 M, i64 N) {
    uint8_t tab[256];
    for (int b = 0; b < 256; ++b) {
        int res = 0;
        for (int k = 0; k < 4; ++k) {
            int v = (b >> (k << 1)) & 3;
            int fv = v == 0 ? 2 : v == 1 ? 1 : v == 2 ? 0 : 3;
            res |= fv << (k << 1);
        }
        tab[b] = (uint8_t)res;
    }
    i64 Mb = (M + 3) / 4;
    i64 last = M % 4 == 0 ? 4 : M % 4;
    uint8_t mask = 0;
    for (i64 k = 0; k < last; ++k) mask |= (uint8_t)(3 << (k << 1));
    for (i64 j = 0; j < Mb; ++j)
        for (i64 i = 0; i < N; ++i) {
            uint8_t v = tab[Gp[j * ldg + i]];
            Gp[j * ldg + i] = (j == Mb - 1) ? (uint8_t)(v & mask) : v;
        }
}

/* tools.pyx:123-149 alleleFrequency (CPU, FP32 sums; 0/0 when all missing). */
void oc_freq_cpu(const uint8_t *G, i64 M, i64 N, float *f) {
    for (i64 m = 0; m < M; ++m) {
        float s = 0.0f, d = 0.0f;
        for (i64 i = 0; i < N; ++i)
            if (G[m * N + i] != 3) { s += (float)G[m * N + i]; d += 2.0f; }
        f[m] = s / d;
    }
}

/* utils.py:781-829 freq_batch_math + _calculate_frequencies_gpu_once (GPU):
 * f = sum / (2 * count) where count > 0, else 0.  Integer counts are exact,
 * so the FP32 quotient is the same whatever the summation order. */
void oc_freq_gpu(const uint8_t *Gp, i64 ldg, i64 M, i64 N, float *f) {
    for (i64 m = 0; m < M; ++m) {
        i64 s = 0, c = 0;
        for (i64 i = 0; i < N; ++i) {
            int g = code_at(Gp, ldg, m, i);
            if (g != 3) { s += g; c += 1; }
        }
        f[m] = c > 0 ? (float)s / (2.0f * (float)c) : 0.0f;
    }
}

/* Per-sample count of observed genotypes times two: the q_bat accumulator of
 * em_adam_gpu.py:76 / em.pyx:259 summed over all SNPs. */
void oc_qbat(const uint8_t *Gp, i64 ldg, i64 M, i64 N, double *qbat) {
    for (i64 i = 0; i < N; ++i) {
        i64 c = 0;
        for (i64 m = 0; m < M; ++m) c += code_at(Gp, ldg, m, i) != 3;
        qbat[i] = 2.0 * (double)c;
    }
}

/* ------------------------------------------------------------------------ */
/* 2. Synthetic genotypes (the build's own generator, not in the reference)  */
/* ------------------------------------------------------------------------ */
/* Counter-based: every genotype is a pure function of (seed, SNP, sample),
 * so any block can be regenerated on the CPU and compared bit-for-bit with
 * the device generator.  h = sum_k fma(Q[i,k], P[m,k], .) in FP64, draws are
 * 32-bit integers compared with floor(h * 2^32).  The device kernel
 * (adx_synth_packed) performs the same integer and fma sequence. */
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static inline int synth_code(const double *P, const double *Q, int K, i64 m, i64 i,
                             uint64_t seed, uint32_t miss_thr) {
    double h = 0.0;
    for (int k = 0; k < K; ++k) h = fma(Q[i * K + k], P[m * K + k], h);
    if (h < 0.0) h = 0.0;
    if (h > 1.0) h = 1.0;
    uint64_t thr = (uint64_t)(h * 4294967296.0);
    uint64_t key = seed ^ ((uint64_t)m * 0xD1B54A32D192ED03ULL) ^ ((uint64_t)i * 0x8CB92BA72F3D8DD7ULL);
    uint64_t r1 = mix64(key);
    uint64_t r2 = mix64(key ^ 0x5851F42D4C957F2DULL);
    if ((uint32_t)r2 < miss_thr) return 3;
    int g = ((uint64_t)(uint32_t)r1 < thr) + ((uint64_t)(uint32_t)(r1 >> 32) < thr);
    return g;
}

void oc_synth_packed(const double *P, const double *Q, i64 M, i64 N, int K, uint64_t seed,
                     double miss_rate, uint8_t *Gp, i64 ldg, i64 row0, i64 rows) {
    uint32_t miss_thr = (uint32_t)(miss_rate * 4294967296.0);
    i64 Mb = (M + 3) / 4;
    i64 r1 = row0 + rows < Mb ? row0 + rows : Mb;
#pragma omp parallel for schedule(static)
    for (i64 j = row0; j < r1; ++j)
        for (i64 i = 0; i < N; ++i) {
            uint8_t out = 0;
            for (int k = 0; k < 4; ++k) {
                i64 m = j * 4 + k;
                if (m >= M) break;
                out |= (uint8_t)(synth_code(P, Q, K, m, i, seed, miss_thr) << (2 * k));
            }
            Gp[(j - row0) * ldg + i] = out;
        }
}

/* ------------------------------------------------------------------------ */
/* 3. Adam-EM, GPU semantics (adamixture/model/em_adam_gpu.py)               */
/* ------------------------------------------------------------------------ */

/* em_adam_gpu.py:46-78 em_batch_math accumulated over all chunks as in
 * run_em_step (:182-197): H clamped to [1e-5, 1-1e-5]; masked a = g/H,
 * b = (2-g)/(1-H); A = a Q, B = b Q, T = (a-b)^T P + colsum(b),
 * q_bat = 2 * colsum(mask).  Deterministic: A/B by SNP rows, T by sample. */
void oc_em_stats_gpu(const uint8_t *Gp, i64 ldg, const double *P, const double *Q, i64 M, i64 N,
                     int K, double *A, double *B, double *T, double *qbat) {
#pragma omp parallel for schedule(static)
    for (i64 m = 0; m < M; ++m) {
        double a_acc[64], b_acc[64];
        for (int k = 0; k < K; ++k) a_acc[k] = b_acc[k] = 0.0;
        const double *p = P + m * K;
        for (i64 i = 0; i < N; ++i) {
            int g = code_at(Gp, ldg, m, i);
            if (g == 3) continue;
            const double *q = Q + i * K;
            double h = 0.0;
            for (int k = 0; k < K; ++k) h += p[k] * q[k];
            h = fmin(fmax(h, 1e-5), 1.0 - 1e-5);
            double a = (double)g / h;
            double b = (2.0 - (double)g) / (1.0 - h);
            for (int k = 0; k < K; ++k) { a_acc[k] += a * q[k]; b_acc[k] += b * q[k]; }
        }
        for (int k = 0; k < K; ++k) { A[m * K + k] = a_acc[k]; B[m * K + k] = b_acc[k]; }
    }
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < N; ++i) {
        double t_acc[64];
        double sb = 0.0, cnt = 0.0;
        for (int k = 0; k < K; ++k) t_acc[k] = 0.0;
        const double *q = Q + i * K;
        for (i64 m = 0; m < M; ++m) {
            int g = code_at(Gp, ldg, m, i);
            if (g == 3) continue;
            const double *p = P + m * K;
            double h = 0.0;
            for (int k = 0; k < K; ++k) h += p[k] * q[k];
            h = fmin(fmax(h, 1e-5), 1.0 - 1e-5);
            double a = (double)g / h;
            double b = (2.0 - (double)g) / (1.0 - h);
            for (int k = 0; k < K; ++k) t_acc[k] += (a - b) * p[k];
            sb += b;
            cnt += 2.0;
        }
        for (int k = 0; k < K; ++k) T[i * K + k] = t_acc[k] + sb;
        if (qbat) qbat[i] = cnt;
    }
}

/* em_adam_gpu.py:82-117 em_final_update: P_EM = A P / max((A-B) P + B, 1e-5)
 * (as reciprocal, *P, *A); Q_EM = Q T / max(q_bat, 1e-5), rows divided by
 * max(row sum, 1e-5). */
void oc_em_final_gpu(const double *P, const double *Q, const double *A, const double *B,
                     const double *T, const double *qbat, i64 M, i64 N, int K, double *P_EM,
                     double *Q_EM) {
    for (i64 x = 0; x < M * K; ++x) {
        double d = (A[x] - B[x]) * P[x] + B[x];
        d = fmax(d, 1e-5);
        P_EM[x] = ((1.0 / d) * P[x]) * A[x];
    }
    for (i64 i = 0; i < N; ++i) {
        double sc = 1.0 / fmax(qbat[i], 1e-5);
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            double v = (Q[i * K + k] * T[i * K + k]) * sc;
            Q_EM[i * K + k] = v;
            s += v;
        }
        s = fmax(s, 1e-5);
        for (int k = 0; k < K; ++k) Q_EM[i * K + k] /= s;
    }
}

/* em_adam_gpu.py:11-42 adam_update applied elementwise to n values. */
static void adam_update_gpu(double *p, const double *tgt, double *m, double *v, i64 n, double t,
                            double lr, double b1, double b2, double eps) {
    double bc1 = 1.0 - pow(b1, t);
    double bc2 = 1.0 - pow(b2, t);
    for (i64 x = 0; x < n; ++x) {
        double d = tgt[x] - p[x];
        m[x] = m[x] * b1 + d * (1.0 - b1);
        v[x] = v[x] * b2 + (1.0 - b2) * d * d;
        double mh = m[x] / bc1;
        double vh = v[x] / bc2;
        p[x] += lr * mh / (sqrt(vh) + eps);
    }
}

/* em_adam_gpu.py:201-224 EMAdamOptimizer.step with the step counter already
 * incremented to t: Adam on P and Q toward the EM targets, clamp both to
 * [1e-5, 1-1e-5], renormalise Q rows. */
void oc_adam_step_gpu(double *P, double *Q, const double *P_EM, const double *Q_EM, double *mP,
                      double *vP, double *mQ, double *vQ, i64 M, i64 N, int K, double t, double lr,
                      double b1, double b2, double eps) {
    adam_update_gpu(P, P_EM, mP, vP, M * K, t, lr, b1, b2, eps);
    adam_update_gpu(Q, Q_EM, mQ, vQ, N * K, t, lr, b1, b2, eps);
    for (i64 x = 0; x < M * K; ++x) P[x] = fmin(fmax(P[x], 1e-5), 1.0 - 1e-5);
    for (i64 i = 0; i < N; ++i) {
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            double v = fmin(fmax(Q[i * K + k], 1e-5), 1.0 - 1e-5);
            Q[i * K + k] = v;
            s += v;
        }
        for (int k = 0; k < K; ++k) Q[i * K + k] /= s;
    }
}

/* em_adam_gpu.py:226-279 logl_batch_math / loglikelihood_gpu: FP64 sum over
 * observed genotypes of g log H + (2-g) log1p(-H), H clamped to
 * [1e-10, 1-1e-10].  Row sums in parallel, rows summed in order. */
double oc_loglik_gpu(const uint8_t *Gp, i64 ldg, const double *P, const double *Q, i64 M, i64 N,
                     int K) {
    double *rows = (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
#pragma omp parallel for schedule(static)
    for (i64 m = 0; m < M; ++m) {
        const double *p = P + m * K;
        double s = 0.0;
        for (i64 i = 0; i < N; ++i) {
            int g = code_at(Gp, ldg, m, i);
            if (g == 3) continue;
            const double *q = Q + i * K;
            double h = 0.0;
            for (int k = 0; k < K; ++k) h += p[k] * q[k];
            h = fmin(fmax(h, 1e-10), 1.0 - 1e-10);
            s += (double)g * log(h) + (2.0 - (double)g) * log1p(-h);
        }
        rows[m] = s;
    }
    double ll = 0.0;
    for (i64 m = 0; m < M; ++m) ll += rows[m];
    free(rows);
    return ll;
}

/* ------------------------------------------------------------------------ */
/* 4. Adam-EM, CPU semantics (em.pyx, tools.pyx)                             */
/* ------------------------------------------------------------------------ */

/* em.pyx:212-277 P_step on unpacked G (M x N).  Adds into Q_T and q_bat (the
 * caller zeroes them, as Q_step does).  With nthreads > 1 the per-thread
 * buffers are merged under a lock in completion order, like the reference;
 * nthreads == 1 is deterministic. */
void oc_P_step_cpu(const uint8_t *G, const double *P, double *P_EM, const double *Q, double *Q_T,
                   double *q_bat, int K, i64 M, i64 N, int nthreads) {
#ifdef _OPENMP
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *A = (double *)calloc((size_t)K, sizeof(double));
        double *B = (double *)calloc((size_t)K, sizeof(double));
        double *t = (double *)calloc((size_t)(N * K), sizeof(double));
        double *ql = (double *)calloc((size_t)N, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(guided)
#endif
        for (i64 row = 0; row < M; ++row) {
            const double *p = P + row * K;
            const uint8_t *g = G + row * N;
            for (i64 col = 0; col < N; ++col) {
                if (g[col] == 3) continue;
                ql[col] += 2.0;
                const double *q = Q + col * K;
                double rec = 0.0;
                for (int k = 0; k < K; ++k) rec += p[k] * q[k];
                double gf = (double)g[col];
                double a = gf / rec;
                double b = (2.0 - gf) / (1.0 - rec);
                double amb = a - b;
                double *tt = t + col * K;
                for (int k = 0; k < K; ++k) {
                    A[k] += q[k] * a;
                    B[k] += q[k] * b;
                    tt[k] += p[k] * amb + b;
                }
            }
            for (int k = 0; k < K; ++k) {
                P_EM[row * K + k] = clip_domain((A[k] * p[k]) / (p[k] * (A[k] - B[k]) + B[k]));
                A[k] = 0.0;
                B[k] = 0.0;
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (i64 n = 0; n < N; ++n) {
                q_bat[n] += ql[n];
                for (int k = 0; k < K; ++k) Q_T[n * K + k] += t[n * K + k];
            }
        }
        free(A); free(B); free(t); free(ql);
    }
}

/* em.pyx:280-303 Q_step + _updateEM_Q (:64-89): element clip, then divide
 * by the row total; zeroes T and q_bat. */
void oc_Q_step_cpu(const double *Q, double *Q_EM, double *T, double *q_bat, int K, i64 N) {
    for (i64 i = 0; i < N; ++i) {
        double a = 1.0 / q_bat[i];
        double tot = 0.0;
        for (int k = 0; k < K; ++k) {
            Q_EM[i * K + k] = clip_domain(Q[i * K + k] * T[i * K + k] * a);
            tot += Q_EM[i * K + k];
        }
        double inv = 1.0 / tot;
        for (int k = 0; k < K; ++k) { Q_EM[i * K + k] *= inv; T[i * K + k] = 0.0; }
        q_bat[i] = 0.0;
    }
}

/* em.pyx:92-145 adamUpdateP. */
void oc_adamP_cpu(double *P0, const double *P1, double *mP, double *vP, double alpha, double b1,
                  double b2, double eps, int t, i64 M, int K) {
    double b1t = pow(b1, t), b2t = pow(b2, t);
    double ms = b1t != 1.0 ? 1.0 / (1.0 - b1t) : 1.0;
    double vs = b2t != 1.0 ? 1.0 / (1.0 - b2t) : 1.0;
    for (i64 x = 0; x < M * K; ++x) {
        double d = P1[x] - P0[x];
        mP[x] = b1 * mP[x] + (1.0 - b1) * d;
        vP[x] = b2 * vP[x] + (1.0 - b2) * d * d;
        double step = alpha * (mP[x] * ms) / (sqrt(vP[x] * vs) + eps);
        P0[x] = clip_domain(P0[x] + step);
    }
}

/* em.pyx:148-209 adamUpdateQ. */
void oc_adamQ_cpu(double *Q0, const double *Q1, double *mQ, double *vQ, double alpha, double b1,
                  double b2, double eps, int t, i64 N, int K) {
    double b1t = pow(b1, t), b2t = pow(b2, t);
    double ms = b1t != 1.0 ? 1.0 / (1.0 - b1t) : 1.0;
    double vs = b2t != 1.0 ? 1.0 / (1.0 - b2t) : 1.0;
    for (i64 i = 0; i < N; ++i) {
        double s = 0.0;
        for (int k = 0; k < K; ++k) {
            i64 x = i * K + k;
            double d = Q1[x] - Q0[x];
            mQ[x] = b1 * mQ[x] + (1.0 - b1) * d;
            vQ[x] = b2 * vQ[x] + (1.0 - b2) * d * d;
            double step = alpha * (mQ[x] * ms) / (sqrt(vQ[x] * vs) + eps);
            Q0[x] = clip_domain(Q0[x] + step);
            s += Q0[x];
        }
        double inv = 1.0 / s;
        for (int k = 0; k < K; ++k) Q0[i * K + k] *= inv;
    }
}

/* tools.pyx:59-85 loglikelihood on unpacked G (no clamp).  Row sums in
 * parallel, summed in row order. */
double oc_loglik_cpu(const uint8_t *G, const double *P, const double *Q, i64 M, i64 N, int K) {
    double *rows = (double *)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
#pragma omp parallel for schedule(static)
    for (i64 m = 0; m < M; ++m) {
        double s = 0.0;
        for (i64 i = 0; i < N; ++i) {
            uint8_t g = G[m * N + i];
            if (g == 3) continue;
            double rec = 0.0;
            for (int k = 0; k < K; ++k) rec += P[m * K + k] * Q[i * K + k];
            s += (double)g * log(rec) + (2.0 - (double)g) * log1p(-rec);
        }
        rows[m] = s;
    }
    double ll = 0.0;
    for (i64 m = 0; m < M; ++m) ll += rows[m];
    free(rows);
    return ll;
}

/* tools.pyx:180-255 mapQ_d / mapP_d; tools.pyx:152-177 rmse_d. */
void oc_mapQ(double *Q, i64 N, int K) {
    for (i64 i = 0; i < N; ++i) {
        double t = 0.0;
        for (int k = 0; k < K; ++k) {
            double r = Q[i * K + k];
            double c = r < 1e-5 ? 1e-5 : (r > 1.0 - 1e-5 ? 1.0 - 1e-5 : r);
            t += c;
            Q[i * K + k] = c;
        }
        double v = 1.0 / t;
        for (int k = 0; k < K; ++k) Q[i * K + k] *= v;
    }
}

void oc_mapP(double *P, i64 M, int K) {
    for (i64 x = 0; x < M * K; ++x) {
        if (P[x] < 1e-5) P[x] = 1e-5;
        else if (P[x] > 1.0 - 1e-5) P[x] = 1.0 - 1e-5;
    }
}

/* ------------------------------------------------------------------------ */
/* 5. Bounded least squares by block principal pivoting                      */
/* ------------------------------------------------------------------------ */

/* bvls.pyx:7-72 solve_subset_ptr (= bvls_kernel.cu:11-72 solve_subset):
 * Gauss-Jordan with partial pivoting on the free subset; singular pivot
 * (< 1e-12) -> failure. */
static int solve_subset(const double *A, const double *b, double *x, const int *F, int K,
                        double *aug, int *map_idx) {
    int n = 0;
    for (int i = 0; i < K; ++i)
        if (F[i] == 1) map_idx[n++] = i;
    if (n == 0) return 1;
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) aug[i * (n + 1) + j] = A[map_idx[i] * K + map_idx[j]];
        aug[i * (n + 1) + n] = b[map_idx[i]];
    }
    for (int i = 0; i < n; ++i) {
        int mr = i;
        double mv = fabs(aug[i * (n + 1) + i]);
        for (int k = i + 1; k < n; ++k) {
            double t = fabs(aug[k * (n + 1) + i]);
            if (t > mv) { mv = t; mr = k; }
        }
        if (mv < 1e-12) return 0;
        if (mr != i)
            for (int j = i; j <= n; ++j) {
                double t = aug[i * (n + 1) + j];
                aug[i * (n + 1) + j] = aug[mr * (n + 1) + j];
                aug[mr * (n + 1) + j] = t;
            }
        double piv = aug[i * (n + 1) + i];
        for (int j = i; j <= n; ++j) aug[i * (n + 1) + j] /= piv;
        for (int k = 0; k < n; ++k)
            if (k != i) {
                double t = aug[k * (n + 1) + i];
                for (int j = i; j <= n; ++j) aug[k * (n + 1) + j] -= t * aug[i * (n + 1) + j];
            }
    }
    for (int i = 0; i < n; ++i) x[map_idx[i]] = aug[i * (n + 1) + n];
    return 1;
}

/* Shared BPP loop of bvls.pyx:176-226 and bvls_kernel.cu:115-163, entered
 * with x, F, U set from the unconstrained solution. */
static void bpp_loop(const double *A, const double *b, double *x, int K, int *F, int *U,
                     double *y, double *bs, double *aug, int *mi, double lo, double hi) {
    for (int it = 1; it < 50; ++it) {
        for (int i = 0; i < K; ++i)
            if (F[i] == 1) {
                bs[i] = b[i];
                for (int j = 0; j < K; ++j)
                    if (F[j] == 0) bs[i] -= A[i * K + j] * (U[j] == 1 ? hi : lo);
            }
        if (!solve_subset(A, bs, x, F, K, aug, mi)) break;
        for (int i = 0; i < K; ++i)
            if (F[i] == 0) x[i] = U[i] == 1 ? hi : lo;
        for (int i = 0; i < K; ++i) {
            y[i] = -b[i];
            for (int j = 0; j < K; ++j) y[i] += A[i * K + j] * x[j];
        }
        int changed = 0;
        for (int i = 0; i < K; ++i) {
            if (F[i] == 1) {
                if (x[i] < lo - 1e-8) { F[i] = 0; U[i] = 0; changed = 1; }
                else if (x[i] > hi + 1e-8) { F[i] = 0; U[i] = 1; changed = 1; }
            } else if (U[i] == 0) {
                if (y[i] < -1e-8) { F[i] = 1; U[i] = 0; changed = 1; }
            } else if (y[i] > 1e-8) {
                F[i] = 1; U[i] = 0; changed = 1;
            }
        }
        if (!changed) break;
    }
}

/* bvls.pyx:130-226 _exact_bvls_bpp_ptr via batch_bvls_bpp (:268-312): the
 * unconstrained start is solved by Gauss-Jordan on all K variables. */
void oc_bvls_cpu(const double *A, const double *Bt, double *X, i64 M, int K, double lo, double hi) {
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < M; ++r) {
        double aug[64 * 65], y[64], bs[64];
        int mi[64], F[64], U[64];
        const double *b = Bt + r * K;
        double *x = X + r * K;
        int changed = 0;
        for (int i = 0; i < K; ++i) { F[i] = 1; U[i] = 0; bs[i] = b[i]; }
        if (!solve_subset(A, bs, x, F, K, aug, mi)) continue;
        for (int i = 0; i < K; ++i) {
            if (x[i] < lo - 1e-8) { F[i] = 0; U[i] = 0; changed = 1; }
            else if (x[i] > hi + 1e-8) { F[i] = 0; U[i] = 1; changed = 1; }
        }
        if (!changed) continue;
        bpp_loop(A, b, x, K, F, U, y, bs, aug, mi, lo, hi);
    }
}

/* bvls_kernel.cu:74-164 bvls_bpp_kernel: same BPP loop entered from a given
 * unconstrained solution X0 (the host computes X0 = B inv(A)^T,
 * bvls_kernel.cu:200-201). */
void oc_bvls_gpu(const double *A, const double *Bt, const double *X0, double *X, i64 M, int K,
                 double lo, double hi) {
#pragma omp parallel for schedule(static)
    for (i64 r = 0; r < M; ++r) {
        double aug[64 * 65], y[64], bs[64];
        int mi[64], F[64], U[64];
        const double *b = Bt + r * K;
        double *x = X + r * K;
        int changed = 0;
        for (int i = 0; i < K; ++i) {
            F[i] = 1; U[i] = 0;
            x[i] = X0[r * K + i];
            if (x[i] < lo - 1e-8) { F[i] = 0; U[i] = 0; changed = 1; }
            else if (x[i] > hi + 1e-8) { F[i] = 0; U[i] = 1; changed = 1; }
        }
        if (!changed) continue;
        bpp_loop(A, b, x, K, F, U, y, bs, aug, mi, lo, hi);
    }
}

/* ------------------------------------------------------------------------ */
/* 6. SQP block updates (sqp.pyx, sqp_kernel.cu)                             */
/* ------------------------------------------------------------------------ */

/* sqp.pyx:6-29 sweep operator on an sz x sz tableau. */
static void sweep(double *a, int sz, int k, double *tmp, int inverse) {
    double piv = a[k * sz + k];
    if (piv == 0.0) return;
    double p = 1.0 / piv;
    for (int i = 0; i < sz; ++i) {
        tmp[i] = a[i * sz + k];
        a[i * sz + k] = 0.0;
        a[k * sz + i] = 0.0;
    }
    tmp[k] = inverse ? 1.0 : -1.0;
    for (int i = 0; i < sz; ++i) {
        double pv = p * tmp[i];
        for (int j = 0; j < sz; ++j) a[j * sz + i] -= pv * tmp[j];
    }
}

/* sqp.pyx:31-110 quadratic_program_local: active-set QP by sweeping, with
 * p box-constrained variables (0 <= par + delta <= 1) and c equality rows.
 * Returns the iteration count, 0 on a non-PD tableau or no convergence. */
int oc_qp_sweep(double *delta, double *tab, const double *par, int p, int c) {
    int sz = p + c + 1;
    double d[132], tmp[132];
    uint8_t swept[128];
    const double small = 1e-5, tol = 1e-8;
    for (int i = 0; i < p; ++i) delta[i] = 0.0;
    for (int i = 0; i < sz; ++i) d[i] = tab[i * sz + i];
    for (int i = 0; i < p; ++i) {
        if (d[i] <= 0.0 || tab[i * sz + i] < d[i] * tol) return 0;
        sweep(tab, sz, i, tmp, 0);
    }
    for (int i = 0; i < p; ++i) swept[i] = 1;
    for (int i = p; i < p + c; ++i) {
        if (tab[i * sz + i] >= 0.0) return 0;
        sweep(tab, sz, i, tmp, 0);
    }
    for (int it = 1; it <= 1000; ++it) {
        double a = 1.0;
        for (int i = 0; i < p; ++i)
            if (swept[i]) {
                double ui = tab[i * sz + sz - 1];
                double ai = ui > 0.0 ? 1.0 - par[i] - delta[i] : 0.0 - par[i] - delta[i];
                if (fabs(ui) > 1e-10) {
                    double t = ai / ui;
                    if (t < a) a = t;
                }
            }
        for (int i = 0; i < p; ++i)
            if (swept[i]) {
                double ui = tab[i * sz + sz - 1];
                delta[i] += a * ui;
                tab[i * sz + sz - 1] = (1.0 - a) * ui;
                tab[(sz - 1) * sz + i] = tab[i * sz + sz - 1];
            }
        int cyc = 0;
        for (int i = 0; i < p; ++i) {
            int crit = (0.0 >= par[i] + delta[i] - small) || (1.0 <= par[i] + delta[i] + small);
            if (swept[i] && fabs(tab[i * sz + i]) > 1e-10 && crit) {
                sweep(tab, sz, i, tmp, 1);
                swept[i] = 0;
                cyc = 1;
                break;
            }
        }
        if (cyc) continue;
        for (int i = 0; i < p; ++i) {
            double ui = tab[i * sz + sz - 1];
            int viol = (ui > 0.0 && 0.0 >= par[i] + delta[i] - small) ||
                       (ui < 0.0 && 1.0 <= par[i] + delta[i] + small);
            if (!swept[i] && viol) {
                sweep(tab, sz, i, tmp, 0);
                swept[i] = 1;
                cyc = 1;
                break;
            }
        }
        if (cyc) continue;
        return it;
    }
    return 0;
}

/* sqp.pyx:112-146 project_q_simplex_row (sort-based, pseudocount). */
void oc_project_simplex(double *b, int n, double pc) {
    double tau = 1.0 - n * pc, tsum = 0.0, tmax = 0.0;
    int bget = 0, idx[128];
    for (int i = 0; i < n; ++i) { b[i] -= pc; idx[i] = i; }
    for (int i = 1; i < n; ++i) {
        int key = idx[i];
        double kv = b[key];
        int j = i - 1;
        while (j >= 0 && b[idx[j]] < kv) { idx[j + 1] = idx[j]; --j; }
        idx[j + 1] = key;
    }
    for (int i = 0; i < n - 1; ++i) {
        tsum += b[idx[i]];
        tmax = (tsum - tau) / (i + 1);
        if (tmax >= b[idx[i + 1]]) { bget = 1; break; }
    }
    if (!bget) tmax = (tsum + b[idx[n - 1]] - tau) / n;
    for (int i = 0; i < n; ++i) b[i] = fmax(b[i] - tmax, 0.0) + pc;
}

/* sqp.pyx:153-202 create_tableau_simplex_local (gradient r enters with +
 * sign; sqp_kernel.cu:177-237 negates a host-negated gradient, the same). */
static void tableau_simplex(double *tab, const double *H, const double *r, const double *x,
                            const double *vkk, int K) {
    int sz = K + 2;
    double t1[128], t2[128];
    for (int i = 0; i < sz * sz; ++i) tab[i] = 0.0;
    for (int i = 0; i < K; ++i) {
        t1[i] = 0.0;
        for (int j = 0; j < K; ++j) t1[i] += H[i * K + j] * vkk[j * K + 0];
    }
    for (int i = 0; i < K; ++i) {
        t2[i] = 0.0;
        for (int j = 0; j < K; ++j) t2[i] += vkk[i * K + j] * t1[j];
    }
    double n1 = 0.0;
    for (int i = 0; i < K; ++i) n1 += fabs(t2[i]);
    double mu = 2.0 * ((n1 - 2.0 * fabs(t2[0])) / K);
    if (mu < 0.0) mu = 0.0;
    for (int i = 0; i < K; ++i)
        for (int j = 0; j < K; ++j) tab[i * sz + j] = H[i * K + j] + mu;
    for (int i = 0; i < K; ++i) {
        tab[i * sz + K] = 1.0;
        tab[K * sz + i] = 1.0;
        tab[i * sz + K + 1] = r[i];
        tab[(K + 1) * sz + i] = r[i];
    }
    double sx = 0.0;
    for (int i = 0; i < K; ++i) sx += x[i];
    tab[K * sz + K] = 0.0;
    tab[K * sz + K + 1] = 1.0 - sx;
    tab[(K + 1) * sz + K] = 1.0 - sx;
    tab[(K + 1) * sz + K + 1] = 0.0;
}

/* sqp.pyx:223-239 update_single_q over all samples. */
void oc_sqp_q(const double *XtX, const double *Xtz, const double *Q, double *Qn, const double *vkk,
              i64 N, int K) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < N; ++i) {
        double *tab = (double *)malloc(sizeof(double) * (size_t)((K + 2) * (K + 2)));
        double delta[128];
        tableau_simplex(tab, XtX + i * K * K, Xtz + i * K, Q + i * K, vkk, K);
        oc_qp_sweep(delta, tab, Q + i * K, K, 1);
        for (int k = 0; k < K; ++k) Qn[i * K + k] = Q[i * K + k] + delta[k];
        oc_project_simplex(Qn + i * K, K, 1e-5);
        free(tab);
    }
}

/* sqp.pyx:204-256 create_tableau_box_local + update_single_p + box clip. */
void oc_sqp_p(const double *XtX, const double *Xtz, const double *P, double *Pn, i64 M, int K) {
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < M; ++j) {
        int sz = K + 1;
        double *tab = (double *)malloc(sizeof(double) * (size_t)(sz * sz));
        double delta[128];
        for (int x = 0; x < sz * sz; ++x) tab[x] = 0.0;
        for (int a = 0; a < K; ++a)
            for (int b = 0; b < K; ++b) tab[a * sz + b] = XtX[j * K * K + a * K + b];
        for (int a = 0; a < K; ++a) {
            tab[a * sz + K] = Xtz[j * K + a];
            tab[K * sz + a] = Xtz[j * K + a];
        }
        tab[K * sz + K] = 0.0;
        oc_qp_sweep(delta, tab, P + j * K, K, 0);
        for (int k = 0; k < K; ++k) {
            double v = P[j * K + k] + delta[k];
            Pn[j * K + k] = fmin(fmax(v, 1e-5), 1.0 - 1e-5);
        }
        free(tab);
    }
}

/* sqp.pyx:258-291 compute_grad_hess_Q (packed G; H clamped to
 * [1e-10, 1-1e-10]; positive gradient convention). */
void oc_grad_hess_q(const uint8_t *Gp, i64 ldg, const double *Q, const double *P, double *XtX,
                    double *Xtz, i64 M, i64 N, int K) {
#pragma omp parallel for schedule(static)
    for (i64 i = 0; i < N; ++i) {
        double *H = XtX + i * K * K, *z = Xtz + i * K;
        for (int k = 0; k < K * K; ++k) H[k] = 0.0;
        for (int k = 0; k < K; ++k) z[k] = 0.0;
        for (i64 j = 0; j < M; ++j) {
            int gc = code_at(Gp, ldg, j, i);
            if (gc == 3) continue;
            double g = (double)gc, qp = 0.0;
            const double *p = P + j * K;
            for (int k = 0; k < K; ++k) qp += Q[i * K + k] * p[k];
            qp = fmax(fmin(qp, 1.0 - 1e-10), 1e-10);
            double t1z = g / qp, t2z = (2.0 - g) / (1.0 - qp);
            double t1 = t1z / qp, t2 = t2z / (1.0 - qp);
            for (int k = 0; k < K; ++k) {
                z[k] += t1z * p[k] + t2z * (1.0 - p[k]);
                for (int k2 = 0; k2 < K; ++k2)
                    H[k * K + k2] += t1 * p[k] * p[k2] + t2 * (1.0 - p[k]) * (1.0 - p[k2]);
            }
        }
    }
}

/* sqp.pyx:293-328 compute_grad_hess_P. */
void oc_grad_hess_p(const uint8_t *Gp, i64 ldg, const double *Q, const double *P, double *XtX,
                    double *Xtz, i64 M, i64 N, int K) {
#pragma omp parallel for schedule(static)
    for (i64 j = 0; j < M; ++j) {
        double *H = XtX + j * K * K, *z = Xtz + j * K;
        for (int k = 0; k < K * K; ++k) H[k] = 0.0;
        for (int k = 0; k < K; ++k) z[k] = 0.0;
        for (i64 i = 0; i < N; ++i) {
            int gc = code_at(Gp, ldg, j, i);
            if (gc == 3) continue;
            double g = (double)gc, qp = 0.0;
            const double *q = Q + i * K;
            for (int k = 0; k < K; ++k) qp += q[k] * P[j * K + k];
            qp = fmax(fmin(qp, 1.0 - 1e-10), 1e-10);
            double t1z = g / qp, t2z = (2.0 - g) / (1.0 - qp);
            double t1 = t1z / qp, t2 = t2z / (1.0 - qp);
            double zd = t1z - t2z, ts = t1 + t2;
            for (int k = 0; k < K; ++k) {
                z[k] += zd * q[k];
                for (int k2 = 0; k2 < K; ++k2) H[k * K + k2] += ts * q[k] * q[k2];
            }
        }
    }
}

int oc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
