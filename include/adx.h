/*
 * adx.h -- C ABI of libadx, the gfx950 (MI355X) kernels behind the
 * `--device gpu` likelihood path of ADAMIXTURE.
 *
 * Conventions
 *   - Plain device pointers and sizes; no torch types.  `stream` is a
 *     hipStream_t passed as void* (NULL = default stream).  All entry points
 *     are asynchronous on that stream unless stated otherwise.
 *   - Return 0 on success, non-zero on failure; adx_last_error() then holds a
 *     message for the calling thread.  HIP out-of-memory failures carry the
 *     text "out of memory" (the reference classifies OOM by message,
 *     adamixture/src/utils.py:19-43).
 *   - Genotypes use the reference's device layout (adamixture/src/utils_c/
 *     cython/snp_reader.pyx:76-120): Mb = ceil(M/4) byte rows of `ldg` bytes,
 *     byte [j, i] holds SNPs 4j..4j+3 of sample i at bits 2k..2k+1; codes
 *     0/1/2 = allele count, 3 = missing.  Entry points that stream whole
 *     rows need ldg % 16 == 0; bytes i >= N of a row are padding and must
 *     hold missing codes (0xFF, as adx_repack_padded and the readers write
 *     them): the log-likelihood pass counts any other padding code.
 *   - Matrices are row-major: P is M x K (stride ldp), Q is N x K (stride ldq).
 *   - dtype tags: ADX_F32 = 0, ADX_F64 = 1, ADX_BF16 = 2.
 *
 * Reference citations are paths relative to the reference repository root.
 */
#ifndef ADX_H
#define ADX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ADX_F32 = 0, ADX_F64 = 1, ADX_BF16 = 2 };

/* Flags of adx_em_pass. */
enum { ADX_EM_STATS = 1, ADX_EM_LOGLIK = 2 };

const char *adx_last_error(void);

/* --- genotype layout ----------------------------------------------------- */

/* Replaces torch.ops.pack2bit.unpack2bit_gpu_chunk_uint8
 * (adamixture/src/utils_c/cuda/pack2bit.cu:83-110, registered at :154):
 * out (C' x N, uint8) = codes of SNPs m_start .. m_start+C'-1, with
 * C' = min(C, M - m_start).  byte_offset = first byte row present in G. */
int adx_unpack_u8(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, int64_t m_start, int64_t C,
                  int64_t byte_offset, uint8_t *out, void *stream);

/* Replaces torch.ops.pack2bit.unpack2bit_gpu_chunk_center
 * (pack2bit.cu:112-141, registered at :153): out = g - 2 f[m], 0 if missing. */
int adx_unpack_center_f32(const uint8_t *G, int64_t ldg, const float *f, int64_t M, int64_t N,
                          int64_t m_start, int64_t C, int64_t byte_offset, float *out, void *stream);

/* Replaces freq_batch_math + _calculate_frequencies_gpu_once
 * (adamixture/src/utils.py:781-829): f[m] = sum g / (2 #observed), 0 if none. */
int adx_freq_f32(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, float *f, void *stream);

/* Replaces the q_bat accumulation of em_batch_math
 * (adamixture/model/em_adam_gpu.py:76, summed at :197): qbat[i] =
 * 2 x #observed genotypes of sample i (FP64, exact). */
int adx_qbat(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, double *qbat, void *stream);

/* GPU form of the reader's transpose (adamixture/src/utils_c/cython/
 * snp_reader.pyx:76-120): `rows` consecutive SNP-major BED rows of nb bytes
 * (4 samples per byte, codes 00 -> 2, 01 -> 3 missing, 10 -> 1, 11 -> 0)
 * become output byte rows row0 .. row0 + ceil(rows/4) - 1 of G (4 SNPs per
 * byte); fields beyond `rows` are 0 and bytes i >= N are 0xFF. */
int adx_bed_to_packed(const uint8_t *bed, int64_t nb, int64_t rows, int64_t N, uint8_t *G, int64_t ldg,
                      int64_t row0, void *stream);

/* get_mean_packed (snp_reader.pyx:215-251): out2[0] = sum of observed allele
 * counts, out2[1] = 2 x observed genotypes (exact, device uint64). */
int adx_packed_allele_sum(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, unsigned long long *out2,
                          void *stream);

/* flip_packed (snp_reader.pyx:254-302): 0 <-> 2 in place, missing kept,
 * fields of the last byte row beyond M cleared. */
int adx_flip_packed(uint8_t *G, int64_t ldg, int64_t M, int64_t N, void *stream);

/* No reference counterpart (layout): copy Mb rows of N bytes from a tensor
 * with row stride ld_src into rows of ldg bytes, padding bytes set to 0xFF. */
int adx_repack_padded(const uint8_t *src, int64_t ld_src, int64_t Mb, int64_t N, uint8_t *dst,
                      int64_t ldg, void *stream);

/* No reference counterpart (synthetic workloads): G_ij ~ Binomial(2, Q_i.P_j)
 * with missing rate miss_rate, counter-based on (seed, SNP, sample); P (M x K)
 * and Q (N x K) are FP64 device arrays.  Padding bytes are set to 0xFF. */
int adx_synth_packed(const double *P, const double *Q, int64_t M, int64_t N, int K, uint64_t seed,
                     double miss_rate, uint8_t *G, int64_t ldg, void *stream);

/* --- Adam-EM -------------------------------------------------------------- */

/* Tile shapes of the EM passes (project autotuner, no counterpart in the
 * reference): for dtype tag ADX_F32 / ADX_F64 (adx_em_pass) or ADX_BF16
 * (adx_em_pass_bf16), cw = 16-SNP column tiles per wavefront (4 or 8; 4 for
 * ADX_F64), waves = wavefronts per workgroup (4 or 8: a workgroup owns
 * 16 cw waves SNPs, and the partial-T slices shrink as that grows) and
 * target_wgs = workgroups the sample range is split to reach (64..65536).
 * waves applies to the statistics pass with K <= 16; passes that carry the
 * log-likelihood run 4-wave workgroups.  cw is 4 for K > 16, and in the
 * BF16 form for K > 10 (two BF16 k slices).  Process-wide and not synchronised between host threads
 * (set it before passes run on other threads); affects later launches and
 * adx_em_workspace_bytes (query the workspace after changing it).  Defaults:
 * F32 (8, 4, 8192), F64 (4, 4, 4096), BF16 (8, 4, 8192). */
int adx_set_em_tiles(int dtype, int cw, int waves, int64_t target_wgs);
int adx_get_em_tiles(int dtype, int *cw, int *waves, int64_t *target_wgs);

/* Workspace (bytes) for adx_em_pass to run in one SNP band for the given
 * shape, dtype and current tile shape: the log-likelihood partial sums, the
 * per-workgroup partial sums of T (ceil(M / (16 cw waves)) x N x K of the
 * pass dtype) and the per-sample-split partial sums of A and B (2 x splits x
 * M x K).  Enough for either flag set (with or without the log-likelihood).  K = 0 sizes the log-likelihood partials only.  Given less, the pass
 * runs in SNP bands (below). */
int64_t adx_em_workspace_bytes(int dtype, int64_t M, int64_t N, int K);

/* Replaces the chunk loop of EMAdamOptimizer.run_em_step with
 * em_batch_math (adamixture/model/em_adam_gpu.py:166-197, :46-78), and, with
 * ADX_EM_LOGLIK, logl_batch_math (:226-248) at the same (P, Q).
 * One pass over all genotypes; 1 <= K <= 64, the reference GPU path's limit
 * (adamixture/entry.py:299-303).  For K > 16 the A, B, T contractions run in
 * 16-wide k blocks, one launch per block (H over the full K in each, the
 * log-likelihood in the first) at a 4-tile width.
 *   ADX_EM_STATS:  A += a Q, B += b Q (M x K), T += (a-b)^T P + colsum(b)
 *                  (N x K), with H = P Q^T clamped to [1e-5, 1-1e-5],
 *                  a = g/H, b = (2-g)/(1-H) on observed genotypes.  A, B, T
 *                  are accumulated (caller zeroes them; adx_em_finalize does).
 *                  With a workspace of adx_em_workspace_bytes, T, A and B
 *                  are reduced through partial sums (plain stores, summed in
 *                  a fixed order: deterministic).  With less, the pass runs
 *                  over SNP bands (multiples of one workgroup column) whose
 *                  partial sums fit, band after band in a fixed order, with
 *                  fewer sample splits if one column needs it: deterministic
 *                  for a given workspace size, equal to the one-band result
 *                  up to summation order.  With too little for one band (or
 *                  none), T, A, B by device-scope atomics (order varies).
 *   ADX_EM_LOGLIK: *ll_out (device, FP64) = sum g log H + (2-g) log1p(-H),
 *                  H clamped to [1e-10, 1-1e-10]; FP64 dtype only; needs a
 *                  workspace of at least one band (an error otherwise).
 * dtype selects the arithmetic of P, Q, A, B, T: ADX_F64 or ADX_F32. */
int adx_em_pass(int dtype, const uint8_t *G, int64_t ldg, int64_t M, int64_t N, int K,
                const void *P, int64_t ldp, const void *Q, int64_t ldq, void *A, void *B, void *T,
                int flags, double *ll_out, void *workspace, int64_t workspace_bytes, void *stream);

/* The SNP band (in SNPs) adx_em_pass runs with for this shape and workspace
 * (M if one band; max_band > 0 caps it), or -1 if the workspace does not fit
 * one band. */
int64_t adx_em_band_snps(int dtype, int64_t M, int64_t N, int K, int flags, int64_t workspace_bytes,
                         int64_t max_band);

/* Host-streamed form of adx_em_pass for genotype matrices kept in host memory
 * because they do not fit the GPU (the reference keeps G on the host then and
 * copies each byte-row chunk per use, adamixture/src/utils.py:425-428,
 * :538-557).  G_host is the padded layout (ldg, 0xFF padding) in host memory
 * (pinned for asynchronous copies).  The pass runs the SNP bands of
 * adx_em_pass's plan for workspace_bytes, capped to what half the staging
 * buffer (device, 256-byte aligned) holds: each band's byte rows are copied
 * into one of two staging halves on copy_stream (NULL: on stream), ordered
 * by events against the band that last read that half, then processed as
 * adx_em_pass processes a resident band.  With the same workspace and a
 * staging that holds the plan's band twice, A, B, T and the log-likelihood
 * are bit-identical to adx_em_pass on the resident matrix.  Returns after
 * enqueueing; later work on `stream` is ordered after every copy. */
int adx_em_pass_host(int dtype, const uint8_t *G_host, int64_t ldg, int64_t M, int64_t N, int K, const void *P,
                     int64_t ldp, const void *Q, int64_t ldq, void *A, void *B, void *T, int flags,
                     double *ll_out, void *workspace, int64_t workspace_bytes, uint8_t *staging,
                     int64_t staging_bytes, void *stream, void *copy_stream);

/* No counterpart in the reference: adx_em_pass for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments, contract and results as
 * adx_em_pass minus the workspace: the partial sums live in a fixed
 * ADX_EM_NOALLOC_SCRATCH_BYTES array of device memory that the library holds
 * on each GPU it runs on until the process ends; the pass runs in as many SNP
 * bands as that needs (deterministic for a given shape).  Call with the GPU
 * holding G current.  Allocates nothing (no hipMalloc / hipFree, no
 * allocating library call).  A shape whose single SNP band does not fit
 * (about N K x 8 bytes plus the band's A / B partial sums) is refused with an
 * error before any device work.  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (concurrent streams or host threads) race. */
#define ADX_EM_NOALLOC_SCRATCH_BYTES (128LL << 20)
int adx_em_pass_noalloc(int dtype, const uint8_t *G, int64_t ldg, int64_t M, int64_t N, int K,
                        const void *P, int64_t ldp, const void *Q, int64_t ldq, void *A, void *B, void *T,
                        int flags, double *ll_out, void *stream);

/* Mixed-precision form of adx_em_pass (ADX_EM_STATS only, 1 <= K <= 64;
 * K > 16 in 16-wide k blocks as adx_em_pass),
 * FP32 A, B, T (dtype tag ADX_BF16).  Same contract as adx_em_pass with
 * dtype ADX_F32, but the contractions run on 16x16x32 MFMA with FP32
 * accumulation from operand arrays that adx_em_prep_bf16 builds from the
 * FP64 P and Q: H on BF16 MFMA from a three-term hi/lo split; the EM weights
 * a, b rounded to FP16 and the A, B, T contractions on FP16 MFMA (Q and P
 * rounded to FP16).  Stated accuracy against the FP64 pass: 1e-3 relative on
 * A, B, T.
 * The operand buffer also holds the per-workgroup partial sums of T
 * (ceil(M / (16 cw waves)) x N x K FP32, cw and waves from adx_set_em_tiles) and the
 * per-sample-split partial sums of A and B (2 x splits x M x K), so its size
 * depends on the current tile shape: query adx_em_bf16_operand_bytes after
 * changing it.  The slices are summed in a fixed order: the pass is
 * deterministic.  Replaces the same reference code (em_adam_gpu.py:46-78,
 * :166-197). */
int64_t adx_em_bf16_operand_bytes(int64_t M, int64_t N, int K, int64_t *ldqt);
int adx_em_prep_bf16(const double *P, const double *Q, int64_t M, int64_t N, int K, void *operands,
                     int64_t operand_bytes, void *stream);
int adx_em_pass_bf16(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, int K, void *operands,
                     int64_t operand_bytes, float *A, float *B, float *T, void *stream);

/* Replaces em_final_update (adamixture/model/em_adam_gpu.py:82-117):
 * P_EM = A P / max((A-B) P + B, 1e-5); Q_EM = Q T / max(qbat, 1e-5), rows
 * divided by max(row sum, 1e-5).  A, B, T have dtype acc_dtype (F32/F64) and
 * are zeroed on return.  P, Q, qbat, P_EM, Q_EM are FP64, contiguous. */
int adx_em_finalize(int acc_dtype, const double *P, const double *Q, void *A, void *B, void *T,
                    const double *qbat, int64_t M, int64_t N, int K, double *P_EM, double *Q_EM,
                    void *stream);

/* Replaces EMAdamOptimizer.step with adam_update (em_adam_gpu.py:201-224,
 * :11-42): Adam with bias correction at step t (already incremented) toward
 * the EM targets, then P, Q clamped to [1e-5, 1-1e-5] and Q rows renormalised.
 * If P_lo / Q_lo are non-NULL, FP32 copies (row stride ld_lo) of the updated
 * P and Q are written, zero-padded to ld_lo columns. */
int adx_adam_step(double *P, double *Q, const double *P_EM, const double *Q_EM, double *mP, double *vP,
                  double *mQ, double *vQ, int64_t M, int64_t N, int K, double t, double lr, double beta1,
                  double beta2, double eps, float *P_lo, float *Q_lo, int64_t ld_lo, void *stream);

/* adx_adam_step for replay from a captured graph (project extension for the
 * batched multi-K schedule): the step count, lr and bias corrections live in
 * device memory, sched = [t, lr, 1 - beta1^t, 1 - beta2^t] (FP64).  The call
 * first advances t by one and recomputes the corrections on the device, then
 * runs the same update as adx_adam_step.  The host changes lr by writing
 * sched[1] between replays. */
int adx_adam_step_sched(double *P, double *Q, const double *P_EM, const double *Q_EM, double *mP, double *vP,
                        double *mQ, double *vQ, int64_t M, int64_t N, int K, double *sched, double beta1,
                        double beta2, double eps, float *P_lo, float *Q_lo, int64_t ld_lo, void *stream);

/* --- ALS products (FP64) ------------------------------------------------------ */

/* The skinny products of the ALS update (adamixture/model/als_gpu.py:90-113,
 * torch matmuls there): C (a x b) = alpha X^T Y with X (n x a), Y (n x b),
 * 1 <= a, b <= 64, a b <= 1024 (C is overwritten); and C (n x b) =
 * alpha X W + u v^T + beta C with X (n x a), W (a x b), a <= 64, b <= 16,
 * u (n) and v (b) both NULL or both given.  Row-major with row strides. */
int adx_gemm_tn_f64(const double *X, int64_t ldx, const double *Y, int64_t ldy, int64_t n, int a, int b,
                    double alpha, double *C, int64_t ldc, double *work, void *stream);

/* Workspace of adx_gemm_tn_f64: ADX_GEMM_TN_BLOCKS * a * b doubles give a
 * deterministic result (per-workgroup partial sums reduced in a fixed
 * order); work = NULL reduces with FP64 atomics (order varies run to run). */
#define ADX_GEMM_TN_BLOCKS 512
int adx_gemm_nn_f64(const double *X, int64_t ldx, int64_t n, int a, int b, const double *W, int64_t ldw,
                    double alpha, const double *u, const double *v, double beta, double *C, int64_t ldc,
                    void *stream);

/* --- randomized-SVD initializer GEMMs (FP32) -------------------------------- */

/* X is the centred genotype matrix of pack2bit.cu:46-77 (x = g - 2 f[m], 0 if
 * missing), decoded inside the GEMM.  Replace the chunked unpack + SGEMM of
 * adamixture/model/svd_gpu.py:104-107 and :133-138:
 *   adx_svd_xw:  Y (M x L, row-major) = X W, W is N x L
 *   adx_svd_xty: A (N x L, row-major) = X^T Y, Y is M x L
 * Zero the output first.  1 <= L <= 64.  With a workspace of
 * adx_svd_work_floats(0 for xw / 1 for xty, M, N, L) floats the split
 * reduction is summed in a fixed order (deterministic); work = NULL adds
 * the splits with FP32 atomics. */
int adx_svd_xw(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, const float *f, const float *W, int L,
               float *Y, float *work, void *stream);
int adx_svd_xty(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, const float *f, const float *Y, int L,
                float *A, float *work, void *stream);
int64_t adx_svd_work_floats(int which, int64_t M, int64_t N, int L);

/* The same products in FP16 arithmetic on two-term operand splits (16x16x32
 * FP16 MFMA, FP32 accumulation), accurate to FP32 rounding like the FP32
 * forms: xw splits x = g - 2f per SNP and W s (s a power of two), xty
 * computes Y^T Gm - (2f Y)^T Mask on splits of Y s and 2f Y s.  They replace
 * the same reference SGEMMs as adx_svd_xw / adx_svd_xty.  The output is
 * overwritten (no zeroing needed).  work: at least adx_svd_h2_work_bytes(0
 * for xw / 1 for xty, M, N, L) bytes, 256-byte aligned, holding the split
 * operand, its scale and the split-reduction slices (deterministic).
 * 1 <= L <= 64. */
int adx_svd_xw_h2(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, const float *f, const float *W, int L,
                  float *Y, void *work, int64_t work_bytes, void *stream);
int adx_svd_xty_h2(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, const float *f, const float *Y, int L,
                   float *A, void *work, int64_t work_bytes, void *stream);
int64_t adx_svd_h2_work_bytes(int which, int64_t M, int64_t N, int L);

/* --- SQP block-relaxation gradient / Hessian (FP64, 1 <= K <= 64) ---------- */

/* Replaces compute_grad_hess_Q_gpu (side 0) and compute_grad_hess_P_gpu
 * (side 1) with their chunk math (adamixture/model/br_qn_gpu.py:67-214):
 * with H = P Q^T clamped to [1e-10, 1-1e-10], on observed genotypes
 *   side 0: Xtz (N x K)  = sum_j g/H p_j + (2-g)/(1-H) (1-p_j)
 *           XtX (N x K x K) = sum_j g/H^2 p_j p_j^T + (2-g)/(1-H)^2 (1-p_j)(1-p_j)^T
 *   side 1: Xtz (M x K)  = sum_i (g/H - (2-g)/(1-H)) q_i
 *           XtX (M x K x K) = sum_i (g/H^2 + (2-g)/(1-H)^2) q_i q_i^T
 * (positive gradient convention; the reference negates before the SQP
 * solve).  Outputs are overwritten.  G 16-byte aligned, ldg >= N and a
 * multiple of 16 (the padded device layout).  K <= 16: one thread per
 * problem, register blocks; K > 16: one wavefront per problem, each lane
 * owning 4 x 4 blocks of the Hessian (deterministic, no atomics). */
int adx_grad_hess(int side, const uint8_t *G, int64_t ldg, int64_t M, int64_t N, int K, const double *Q,
                  const double *P, double *XtX, double *Xtz, void *stream);

/* --- batched K x K solvers (FP64, 1 <= K <= 64) ---------------------------- */

/* Replaces torch.ops.sqp_kernel.sqp_solve_q_cuda (simplex = 1) and
 * sqp_solve_p_cuda (simplex = 0) (adamixture/src/utils_c/cuda/
 * sqp_kernel.cu:262-370): for each of nprob problems, the sweep-operator
 * active-set QP on the tableau built from H (nprob x K x K), r (nprob x K)
 * and the current point X (nprob x K) -- r enters negated, as the reference
 * op expects the caller's negated gradient (br_qn_gpu.py:295-300) -- then
 * Xn = X + delta projected on the simplex with pseudocount 1e-5 (Q rows, v_kk
 * of svd(ones(1, K)) sets the mu shift, br_qn_gpu.py:248-250) or clipped to
 * [1e-5, 1-1e-5] (P rows).  workspace: >= 8 bytes (problem queue). */
int adx_sqp_solve(int simplex, const double *H, const double *r, const double *X, const double *vkk,
                  double *Xn, int64_t nprob, int K, void *workspace, int64_t workspace_bytes, void *stream);

/* Replaces torch.ops.bvls_kernel.batch_bvls_bpp_cuda (adamixture/src/utils_c/
 * cuda/bvls_kernel.cu:188-218): for each row b of B (nprob x K), the bounded
 * least squares min 1/2 x'Ax - b'x, lo <= x <= hi by block principal pivoting.
 * With Ainv = inv(A) the pivoting starts from X0 = inv(A) b as the CUDA op
 * does; with Ainv = NULL it starts from the Gauss-Jordan solution as the
 * Cython batch_bvls_bpp (bvls.pyx:130-226).  workspace: >= 8 bytes; with
 * Ainv given, K <= 12 and adx_bvls_workspace_bytes(K) bytes, the inverses of
 * all 2^K principal submatrices of A are formed once and each problem runs on
 * one thread (same pivoting rules; the free-set solve is inv(A_FF) b_F,
 * equal to the Gauss-Jordan solution up to rounding). */
int64_t adx_bvls_workspace_bytes(int K);
int adx_bvls_batch(const double *A, const double *Ainv, const double *B, double *X, int64_t nprob, int K,
                   double lo, double hi, void *workspace, int64_t workspace_bytes, void *stream);

/* --- tests ----------------------------------------------------------------- */

/* No reference counterpart: runs one 16x16x4 MFMA (kind 0 = FP32, 1 = FP64)
 * on row-major A (16x4) and B (4x16) with the lane mapping the EM kernels
 * assume, writing D (16x16) row-major.  kind 2 / 3 (FP32 / FP64) write the
 * raw accumulators instead, D[lane * 4 + j]; kind 4 (5) runs one 16x16x32
 * BF16 (FP16) MFMA on FP32 A (16x32) and B (32x16) holding bf16-exact
 * (fp16-exact) values, writing raw accumulators.  Used to pin the lane
 * mappings. */
int adx_probe_mfma(int kind, const void *A, const void *B, void *D, void *stream);

/* --- learned Q warm start (project extension) ----------------------------
 * No counterpart in the reference: it initialises Q by randomized SVD + ALS
 * only (adamixture/src/adamixture.py:185-250, model/als_gpu.py:45-127), and
 * that path stays the fallback.  A 3-layer MLP, ADX_MLP_IN -> ADX_MLP_WIDTH
 * -> ADX_MLP_WIDTH -> ADX_MLP_OUT with ReLU, maps per-sample features (first
 * ADX_MLP_OUT columns: the ALS Q row, zero-padded) to logits z; the warm
 * start is softmax(z + log max(x[:16], 1e-5)) over the first K columns.
 * Flat FP32 parameters: W1 [IN][W] | b1 [W] | W2 [W][W] | b2 [W] |
 * W3 [W][OUT] | b3 [OUT]  (ADX_MLP_NPARAM floats). */
enum {
    ADX_MLP_IN = 64,
    ADX_MLP_WIDTH = 256,
    ADX_MLP_OUT = 16,
    ADX_MLP_NPARAM = 86544,
    ADX_MLP_PACKED_BYTES = 174144
};

/* Training GEMM, FP32: C = alpha op(A) op(B) + beta C, op(X) = X or X^T
 * (ta / tb), op(A) m x k, op(B) k x n, row-major strides.  splits > 1 splits
 * k over workgroups with atomic accumulation (needs beta = 1). */
int adx_mlp_gemm_f32(int ta, int tb, int64_t m, int64_t n, int64_t k, float alpha, const float *A, int64_t lda,
                     const float *B, int64_t ldb, float beta, float *C, int64_t ldc, int splits, void *stream);

/* Z[r, c] += b[c] (c < cols), then ReLU if relu != 0. */
int adx_mlp_bias_act(float *Z, const float *b, int64_t rows, int64_t cols, int64_t ld, int relu, void *stream);

/* dY[e] = 0 where Z[e] <= 0 (backward of ReLU, Z = its output). */
int adx_mlp_relu_bwd(float *dY, const float *Z, int64_t n, void *stream);

/* db[c] += sum_r dY[r, c] (atomic; db zeroed by the caller). */
int adx_mlp_colsum(const float *dY, int64_t rows, int64_t cols, float *db, void *stream);

/* Masked softmax cross-entropy over the first kvec[r] of ADX_MLP_OUT logits:
 * z = logits + log(max(prior, 1e-5)), p = softmax(z), loss[r] = -sum t log p,
 * dlogits = (p - t) / rows.  target, probs, dlogits and loss may be NULL. */
int adx_mlp_softmax_xent(const float *logits, int64_t ldz, const float *prior, int64_t ldp, const float *target,
                         int64_t ldt, const int *kvec, int64_t rows, float *probs, float *dlogits, float *loss,
                         void *stream);

/* Adam on n FP32 parameters, step t >= 1 (bias-corrected). */
int adx_mlp_adam(float *w, const float *g, float *m, float *v, int64_t n, float lr, float beta1, float beta2,
                 float eps, int t, void *stream);

/* Round the FP32 parameters to the BF16 inference layout
 * (ADX_MLP_PACKED_BYTES, 16-byte aligned). */
int adx_mlp_pack_bf16(const float *params, void *packed, void *stream);

/* Fused BF16 inference (MFMA, FP32 accumulation): Q[i, :K] (FP64) = the warm
 * start of sample i, clamped to [1e-5, 1-1e-5] and renormalised as the
 * reference maps Q (adamixture/model/als_gpu.py:23-36).  2 <= K <= 16. */
int adx_mlp_infer_bf16(const float *X, int64_t ldx, int64_t N, int K, const void *packed, double *Q, int64_t ldq,
                       void *stream);

/* --- supervised mode ------------------------------------------------------ */

/* Replaces init_p_supervised_packed (adamixture/src/supervised.py:84-107):
 * P (M x K, FP64, row stride K) = per-population allele frequencies of the
 * labelled samples, 0.5 where a population has no observed genotype at a
 * SNP, clipped to [eps, 1 - eps]; bit-exact with the reference.  order holds
 * the labelled sample indices (each in [0, N)) grouped by population,
 * population k (0-based) at order[offsets[k] .. offsets[k + 1]); ascending
 * indices inside a group read best.  order and offsets are device arrays.
 * N < 2^30. */
int adx_label_freq(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, int K, const int32_t *order,
                   const int64_t *offsets, double eps, double *P, void *stream);

/* Replaces _snap_q_gpu (supervised.py:308-339): for every sample with
 * y[i] > 0, Q[i, :] = eps / max(K - 1, 1) and Q[i, y[i] - 1] = 1 - eps;
 * rows with y[i] <= 0 are unchanged.  y is a device int32 array. */
int adx_snap_labels(double *Q, int64_t N, int K, const int32_t *y, double eps, void *stream);

/* --- VCF / BCF decoding ---------------------------------------------------- */

/* Replaces the GT parsing of read_vcf_file(_packed)
 * (adamixture/src/utils_c/cython/snp_reader.pyx:343-463, 554-827): data line
 * l of `text` (device) spans [line_start[l], line_end[l]) with text[line_end[l]]
 * == '\n'; its genotypes become SNP m0 + l of G (ceil(M/4) x ldg, ldg a
 * multiple of 4), whose fields the caller set to 3 (bytes 0xFF) beforehand.
 * GT must be the first FORMAT subfield, as the reference assumes. */
int adx_vcf_gt(const uint8_t *text, const int64_t *line_start, const int64_t *line_end, int64_t n_lines, int64_t m0,
               int64_t N, uint8_t *G, int64_t ldg, void *stream);

/* Replaces the GT decoding of read_bcf_file(_packed) (snp_reader.pyx:
 * 1003-1108): record r's GT vector starts at buf[gt_off[r]] with
 * gt_desc[r] = n_values | (bytes per value << 16), or gt_desc[r] < 0 for a
 * record without a usable GT field (all missing); it becomes SNP m0 + r of G
 * (fields preset to 3 as above).  Codes are masked to 2 bits as the
 * reference's packing masks them. */
int adx_bcf_gt(const uint8_t *buf, const int64_t *gt_off, const int32_t *gt_desc, int64_t n_rec, int64_t m0,
               int64_t N, uint8_t *G, int64_t ldg, void *stream);

/* PLINK 2 .pgen hardcalls (the reference reads them through pgenlib,
 * snp_reader.pyx:1110-1135).  buf holds the file; record m starts at
 * buf[rstart[m]], is rlen[m] bytes long and has type rtype[m] (low 3 bits:
 * 0 = 2-bit array, 1 = two values + bit array + difflist, 4 = all 0 +
 * difflist, 2 / 3 = LD-compressed against record rbase[m], 3 inverted).
 * Decodes the records list[0 .. n) into SNP rows m of G (fields preset to 3);
 * call once for the non-LD records, then once for the LD-compressed ones.
 * idb = bytes per sample ID in difflists (1 for N <= 256, ... 4).  The
 * caller checks that each record holds its dense track; difflists are
 * bounds-checked against rlen. */
int adx_pgen_decode(const uint8_t *buf, const int64_t *rstart, const int64_t *rlen, const uint8_t *rtype,
                    const int64_t *rbase, const int64_t *list, int64_t n, int64_t N, int idb, uint8_t *G, int64_t ldg,
                    void *stream);

/* Gout (ceil(M_out/4) x ldg) field (m, i) = Gin field (idx[m], i); padding
 * bytes 0xFF, tail fields 0 (the chromosome filter of snp_reader.py:657-660). */
int adx_select_snps(const uint8_t *Gin, int64_t ldg, const int64_t *idx, int64_t M_out, int64_t N, uint8_t *Gout,
                    void *stream);

/* Fields of the last byte row beyond SNP M - 1 set to 0 (pack_genotypes,
 * snp_reader.pyx:123-157, leaves them 0). */
int adx_clear_tail_fields(uint8_t *G, int64_t ldg, int64_t M, int64_t N, void *stream);

/* --- cross-validation ------------------------------------------------------ */

/* Partial-sum slots the deviance entry points need in `work` (doubles). */
#define ADX_CV_BLOCKS 4096

/* A fold given as a list of flat indices f = m N + i (int32 if idx64 == 0,
 * else int64), entries flat[first + e * stride] for e < n (the reference's
 * shuffled array and fold f = shuffled[f::n_folds], adamixture/src/cv.py:
 * 214-231).  Every f must lie in [0, M N).  G is a device copy of the packed
 * matrix (row stride ldg, a multiple of 4); the held-out 2-bit fields are
 * set to 3 (missing) in place -- replaces mask_entries_i32 / _i64
 * (br_qn.pyx:150-203) on the packed layout. */
int adx_cv_mask_list(uint8_t *G, int64_t ldg, int64_t M, int64_t N, const void *flat, int idx64, int64_t first,
                     int64_t stride, int64_t n, void *stream);

/* A fold given by hashing (cv.py:61-81, the streaming path): entry (m, i)
 * is held out iff its genotype is observed and splitmix64(m N + i + seed)
 * mod n_folds == fold.  Gm (same shape as G) receives G with those fields
 * set to 3. */
int adx_cv_mask_hash(const uint8_t *G, uint8_t *Gm, int64_t ldg, int64_t M, int64_t N, uint64_t seed, int n_folds,
                     int fold, void *stream);

/* Replace deviance_squared_sum / _i32 (br_qn.pyx:92-149, 205-246): *out
 * (device) = sum over the fold's observed entries of G of g log(g / mu) +
 * (2 - g) log((2 - g) / (2 - mu)), mu = 2 Q_i . P_m clamped to
 * [1e-10, 2 - 1e-10].  P (M x K) and Q (N x K) FP64 row-major; work holds
 * ADX_CV_BLOCKS doubles.  Fixed summation order (deterministic). */
int adx_cv_deviance_list(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, const void *flat, int idx64,
                         int64_t first, int64_t stride, int64_t n, const double *P, const double *Q, int K,
                         double *work, double *out, void *stream);
int adx_cv_deviance_hash(const uint8_t *G, int64_t ldg, int64_t M, int64_t N, uint64_t seed, int n_folds, int fold,
                         const double *P, const double *Q, int K, double *work, double *out, void *stream);

/* project_q_simplex (sqp.pyx:112-146) on n rows of X (n x K, FP64) in place:
 * each row becomes its Euclidean projection onto {x >= pc, sum x = 1}.
 * The quasi-Newton map of the CPU optimizer (br_qn.py:159-160), which the
 * cross-validation polish uses.  K <= 64. */
int adx_project_simplex(double *X, int64_t n, int K, double pc, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ADX_H */
