// Batched K x K solvers of the BR-QN and ALS paths, FP64.
//
//   sqp_kernel   adamixture/src/utils_c/cuda/sqp_kernel.cu:6-311 (and the
//                Cython twin sqp.pyx:6-256): per problem a sweep-operator
//                active-set QP on a (K+2)^2 simplex tableau (Q rows) or a
//                (K+1)^2 box tableau (P rows), then simplex projection with
//                pseudocount 1e-5 or the [1e-5, 1-1e-5] clip.
//   bvls_kernel  adamixture/src/utils_c/cuda/bvls_kernel.cu:11-164 (and
//                bvls.pyx:7-226): block principal pivoting for
//                min 1/2 x'Ax - b'x, lo <= x <= hi, A shared by all rows.
//
// MI355X mapping: one wavefront per problem with the tableau / augmented
// matrix in LDS and every O(K^2) update (sweep, elimination) spread over the
// 64 lanes; the scalar active-set decisions are evaluated redundantly by all
// lanes from LDS, so control flow stays wave-uniform.  Workgroups are
// persistent and take problems from an atomic queue, which balances the
// uneven iteration counts of different active sets.  The reference instead
// runs one thread per problem with an 83 KB private tableau (scratch-bound).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "adx_common.h"
#include "../../include/adx.h"

namespace adx {

__device__ __forceinline__ void wave_sync() { __syncthreads(); }  // workgroup == one wave

// sqp.pyx:6-29 sweep(), with the (i, j) updates spread over the lanes.
__device__ void sweep_par(double* a, int sz, int k, double* tmp, bool inverse, int lane) {
    const double piv = a[k * sz + k];
    if (piv == 0.0) return;
    const double p = 1.0 / piv;
    for (int i = lane; i < sz; i += 64) tmp[i] = a[i * sz + k];
    wave_sync();
    for (int i = lane; i < sz; i += 64) {
        a[i * sz + k] = 0.0;
        a[k * sz + i] = 0.0;
    }
    if (lane == 0) tmp[k] = inverse ? 1.0 : -1.0;
    wave_sync();
    for (int e = lane; e < sz * sz; e += 64) {
        const int j = e / sz, i = e - j * sz;
        a[e] -= (p * tmp[i]) * tmp[j];
    }
    wave_sync();
}

// sqp.pyx:31-110 quadratic_program_local.  `sw` holds the swept flags.
__device__ int qp_sweep_par(double* delta, double* tab, const double* par, int p, int c, double* d,
                            double* tmp, int* sw, int lane) {
    const int sz = p + c + 1;
    const double small = 1e-5, tol = 1e-8;
    for (int i = lane; i < p; i += 64) delta[i] = 0.0;
    for (int i = lane; i < sz; i += 64) d[i] = tab[i * sz + i];
    wave_sync();
    for (int i = 0; i < p; ++i) {
        if (d[i] <= 0.0 || tab[i * sz + i] < d[i] * tol) return 0;
        sweep_par(tab, sz, i, tmp, false, lane);
    }
    for (int i = lane; i < p; i += 64) sw[i] = 1;
    wave_sync();
    for (int i = p; i < p + c; ++i) {
        if (tab[i * sz + i] >= 0.0) return 0;
        sweep_par(tab, sz, i, tmp, false, lane);
    }
    for (int it = 1; it <= 1000; ++it) {
        double a = 1.0;
        for (int i = 0; i < p; ++i)
            if (sw[i]) {
                const double ui = tab[i * sz + sz - 1];
                const double ai = ui > 0.0 ? 1.0 - par[i] - delta[i] : 0.0 - par[i] - delta[i];
                if (fabs(ui) > 1e-10) {
                    const double t = ai / ui;
                    if (t < a) a = t;
                }
            }
        wave_sync();
        for (int i = lane; i < p; i += 64)
            if (sw[i]) {
                const double ui = tab[i * sz + sz - 1];
                delta[i] += a * ui;
                tab[i * sz + sz - 1] = (1.0 - a) * ui;
                tab[(sz - 1) * sz + i] = (1.0 - a) * ui;
            }
        wave_sync();
        int cyc = 0;
        for (int i = 0; i < p; ++i) {
            const bool crit = (0.0 >= par[i] + delta[i] - small) || (1.0 <= par[i] + delta[i] + small);
            if (sw[i] && fabs(tab[i * sz + i]) > 1e-10 && crit) {
                sweep_par(tab, sz, i, tmp, true, lane);
                if (lane == 0) sw[i] = 0;
                wave_sync();
                cyc = 1;
                break;
            }
        }
        if (cyc) continue;
        for (int i = 0; i < p; ++i) {
            const double ui = tab[i * sz + sz - 1];
            const bool viol = (ui > 0.0 && 0.0 >= par[i] + delta[i] - small) ||
                              (ui < 0.0 && 1.0 <= par[i] + delta[i] + small);
            if (!sw[i] && viol) {
                sweep_par(tab, sz, i, tmp, false, lane);
                if (lane == 0) sw[i] = 1;
                wave_sync();
                cyc = 1;
                break;
            }
        }
        if (cyc) continue;
        return it;
    }
    return 0;
}

// LDS layout of one SQP problem (doubles unless noted).
struct SqpLds {
    static __host__ __device__ int doubles(int K) {
        const int sz = K + 2;
        return sz * sz + 3 * sz + 3 * K + 2 + K;  // tab, d, tmp, t1/t2, delta, par, misc, sw(int)
    }
};

template <bool SIMPLEX>
__global__ __launch_bounds__(64) void sqp_kernel(const double* __restrict__ H, const double* __restrict__ r,
                                                 const double* __restrict__ X, const double* __restrict__ vkk,
                                                 double* __restrict__ Xn, int64_t nprob, int K,
                                                 unsigned long long* __restrict__ queue) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const int sz = K + (SIMPLEX ? 2 : 1);
    double* tab = lds;
    double* d = tab + (K + 2) * (K + 2);
    double* tmp = d + (K + 2);
    double* t1 = tmp + (K + 2);
    double* delta = t1 + (K + 2);
    double* par = delta + K;
    double* misc = par + K;                 // [0] = mu, [1] = problem index
    int* sw = reinterpret_cast<int*>(misc + 2 + K);

    for (;;) {
        if (lane == 0) misc[1] = (double)atomicAdd(queue, 1ULL);
        wave_sync();
        const int64_t pid = (int64_t)misc[1];
        wave_sync();
        if (pid >= nprob) break;
        const double* Hp = H + pid * K * K;
        const double* rp = r + pid * K;
        const double* xp = X + pid * K;

        // Tableau (sqp_kernel.cu:177-260): the caller's gradient enters negated.
        for (int e = lane; e < sz * sz; e += 64) tab[e] = 0.0;
        for (int i = lane; i < K; i += 64) par[i] = xp[i];
        double mu = 0.0;
        if (SIMPLEX) {
            for (int i = lane; i < K; i += 64) {
                double s = 0.0;
                for (int j = 0; j < K; ++j) s += Hp[i * K + j] * vkk[j * K + 0];
                t1[i] = s;
            }
            wave_sync();
            for (int i = lane; i < K; i += 64) {
                double s = 0.0;
                for (int j = 0; j < K; ++j) s += vkk[i * K + j] * t1[j];
                tmp[i] = s;
            }
            wave_sync();
            if (lane == 0) {
                double n1 = 0.0;
                for (int i = 0; i < K; ++i) n1 += fabs(tmp[i]);
                double m = 2.0 * ((n1 - 2.0 * fabs(tmp[0])) / K);
                misc[0] = m < 0.0 ? 0.0 : m;
            }
            wave_sync();
            mu = misc[0];
        }
        for (int e = lane; e < K * K; e += 64) {
            const int i = e / K, j = e - i * K;
            tab[i * sz + j] = Hp[e] + mu;
        }
        for (int i = lane; i < K; i += 64) {
            if (SIMPLEX) {
                tab[i * sz + K] = 1.0;
                tab[K * sz + i] = 1.0;
                tab[i * sz + K + 1] = -rp[i];
                tab[(K + 1) * sz + i] = -rp[i];
            } else {
                tab[i * sz + K] = -rp[i];
                tab[K * sz + i] = -rp[i];
            }
        }
        wave_sync();
        if (SIMPLEX && lane == 0) {
            double sx = 0.0;
            for (int i = 0; i < K; ++i) sx += par[i];
            tab[K * sz + K + 1] = 1.0 - sx;
            tab[(K + 1) * sz + K] = 1.0 - sx;
        }
        wave_sync();

        qp_sweep_par(delta, tab, par, K, SIMPLEX ? 1 : 0, d, tmp, sw, lane);

        double* out = Xn + pid * K;
        if (SIMPLEX) {
            // sqp_kernel.cu:125-166 project_q_simplex_row (serial, K <= 64).
            if (lane == 0) {
                for (int i = 0; i < K; ++i) out[i] = par[i] + delta[i];
                project_simplex_row(out, out, t1, K, 1e-5);
            }
        } else {
            for (int i = lane; i < K; i += 64) {
                double v = par[i] + delta[i];
                v = v < 1e-5 ? 1e-5 : v;
                v = v > 1.0 - 1e-5 ? 1.0 - 1e-5 : v;
                out[i] = v;
            }
        }
        wave_sync();
    }
}

// bvls.pyx:7-72 solve_subset with the row operations spread over the lanes.
// aug is n x (n+1) in LDS; returns false on a pivot below 1e-12.
__device__ bool solve_subset_par(const double* A, const double* b, double* x, const int* F, int K, double* aug,
                                 int* mi, double* mult, int* scal, int lane) {
    if (lane == 0) {
        int n = 0;
        for (int i = 0; i < K; ++i)
            if (F[i] == 1) mi[n++] = i;
        scal[0] = n;
    }
    wave_sync();
    const int n = scal[0];
    if (n == 0) return true;
    const int w = n + 1;
    for (int e = lane; e < n * w; e += 64) {
        const int i = e / w, j = e - i * w;
        aug[e] = j < n ? A[mi[i] * K + mi[j]] : b[mi[i]];
    }
    wave_sync();
    for (int i = 0; i < n; ++i) {
        if (lane == 0) {
            int mr = i;
            double mv = fabs(aug[i * w + i]);
            for (int k = i + 1; k < n; ++k) {
                const double t = fabs(aug[k * w + i]);
                if (t > mv) {
                    mv = t;
                    mr = k;
                }
            }
            scal[1] = mv < 1e-12 ? -1 : mr;
        }
        wave_sync();
        const int mr = scal[1];
        if (mr < 0) return false;
        if (mr != i)
            for (int j = i + lane; j <= n; j += 64) {
                const double t = aug[i * w + j];
                aug[i * w + j] = aug[mr * w + j];
                aug[mr * w + j] = t;
            }
        wave_sync();
        const double piv = aug[i * w + i];
        wave_sync();
        for (int j = i + lane; j <= n; j += 64) aug[i * w + j] /= piv;
        for (int k = lane; k < n; k += 64) mult[k] = aug[k * w + i];
        wave_sync();
        for (int e = lane; e < n * (n + 1 - i); e += 64) {
            const int k = e / (n + 1 - i), j = i + (e - k * (n + 1 - i));
            if (k != i) aug[k * w + j] -= mult[k] * aug[i * w + j];
        }
        wave_sync();
    }
    for (int i = lane; i < n; i += 64) x[mi[i]] = aug[i * w + n];
    wave_sync();
    return true;
}

struct BvlsLds {
    static __host__ __device__ int doubles(int K) { return K * (K + 1) + 6 * K + 4; }
};

// bvls_kernel.cu:74-164 (Ainv given: start from X0 = inv(A) b) or
// bvls.pyx:130-226 (Ainv NULL: start from the Gauss-Jordan solution).
__global__ __launch_bounds__(64) void bvls_kernel(const double* __restrict__ A, const double* __restrict__ Ainv,
                                                  const double* __restrict__ B, double* __restrict__ X,
                                                  int64_t nprob, int K, double lo, double hi,
                                                  unsigned long long* __restrict__ queue) {
    extern __shared__ double lds[];
    const int lane = threadIdx.x;
    const double* sA = A;                   // K x K, read through the caches (LDS stays < 64 KB at K = 64)
    double* aug = lds;
    double* x = aug + K * (K + 1);
    double* y = x + K;
    double* bs = y + K;
    double* mult = bs + K;
    double* bb = mult + K;
    double* misc = bb + K;
    int* F = reinterpret_cast<int*>(misc + 4);
    int* U = F + K;
    int* mi = U + K;
    int* scal = mi + K;

    for (;;) {
        if (lane == 0) misc[0] = (double)atomicAdd(queue, 1ULL);
        wave_sync();
        const int64_t pid = (int64_t)misc[0];
        wave_sync();
        if (pid >= nprob) break;
        for (int i = lane; i < K; i += 64) {
            bb[i] = B[pid * K + i];
            F[i] = 1;
            U[i] = 0;
        }
        wave_sync();
        bool ok = true;
        if (Ainv) {
            for (int i = lane; i < K; i += 64) {
                double s = 0.0;
                for (int j = 0; j < K; ++j) s += bb[j] * Ainv[i * K + j];
                x[i] = s;
            }
            wave_sync();
        } else {
            for (int i = lane; i < K; i += 64) {
                bs[i] = bb[i];
                x[i] = 0.0;
            }
            wave_sync();
            ok = solve_subset_par(sA, bs, x, F, K, aug, mi, mult, scal, lane);
        }
        if (ok) {
            if (lane == 0) {
                int changed = 0;
                for (int i = 0; i < K; ++i) {
                    if (x[i] < lo - 1e-8) {
                        F[i] = 0; U[i] = 0; changed = 1;
                    } else if (x[i] > hi + 1e-8) {
                        F[i] = 0; U[i] = 1; changed = 1;
                    }
                }
                scal[2] = changed;
            }
            wave_sync();
            if (scal[2]) {
                for (int it = 1; it < 50; ++it) {
                    for (int i = lane; i < K; i += 64)
                        if (F[i] == 1) {
                            double s = bb[i];
                            for (int j = 0; j < K; ++j)
                                if (F[j] == 0) s -= sA[i * K + j] * (U[j] == 1 ? hi : lo);
                            bs[i] = s;
                        }
                    wave_sync();
                    if (!solve_subset_par(sA, bs, x, F, K, aug, mi, mult, scal, lane)) break;
                    for (int i = lane; i < K; i += 64)
                        if (F[i] == 0) x[i] = U[i] == 1 ? hi : lo;
                    wave_sync();
                    for (int i = lane; i < K; i += 64) {
                        double s = -bb[i];
                        for (int j = 0; j < K; ++j) s += sA[i * K + j] * x[j];
                        y[i] = s;
                    }
                    wave_sync();
                    if (lane == 0) {
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
                        scal[2] = changed;
                    }
                    wave_sync();
                    if (!scal[2]) break;
                }
            }
            for (int i = lane; i < K; i += 64) X[pid * K + i] = x[i];
        } else {
            // bvls.pyx:162-163: a singular start leaves the output untouched (zeros).
            for (int i = lane; i < K; i += 64) X[pid * K + i] = 0.0;
        }
        wave_sync();
    }
}

// ---- shared-A BVLS for small K: one thread per problem ----------------
//
// Every problem of a batch shares A, so the inverse of each principal
// submatrix A_FF (F a free set, 2^K of them) is computed once per batch by
// bvls_subset_inv_kernel, with the pivoting of solve_subset_par (largest
// |pivot| of the column, first on ties; a pivot below 1e-12 marks F
// singular).  bvls_thread_kernel then runs the block-principal-pivoting
// loop of bvls_kernel per thread: the solve for the free variables becomes
// x_F = inv(A_FF) bs_F, the rest (start x0 = inv(A) b, the bound and
// gradient tests, the iteration limit, the break on a singular free set) is
// unchanged.  x_F then differs from the Gauss-Jordan result only by
// rounding.

constexpr int BVLS_FAST_K = 12;

// One wavefront per free set S (bit i = variable i free); table slot S holds
// inv(A_SS) (n x n, rows and columns in ascending variable order) in a
// K x K block; sing[S] = 1 if the elimination hit a pivot below 1e-12.
__global__ __launch_bounds__(64) void bvls_subset_inv_kernel(const double* __restrict__ A, int K,
                                                             double* __restrict__ table, int* __restrict__ sing) {
    __shared__ double aug[BVLS_FAST_K * 2 * BVLS_FAST_K];
    __shared__ double mult[BVLS_FAST_K];
    __shared__ int mi[BVLS_FAST_K];
    __shared__ int scal[2];
    const int lane = threadIdx.x;
    const unsigned S = blockIdx.x;
    if (lane == 0) {
        int n = 0;
        for (int i = 0; i < K; ++i)
            if ((S >> i) & 1u) mi[n++] = i;
        scal[0] = n;
    }
    __syncthreads();
    const int n = scal[0], w = 2 * n;
    for (int e = lane; e < n * w; e += 64) {
        const int i = e / w, j = e - i * w;
        aug[e] = j < n ? A[mi[i] * K + mi[j]] : (j - n == i ? 1.0 : 0.0);
    }
    __syncthreads();
    for (int i = 0; i < n; ++i) {
        if (lane == 0) {
            int mr = i;
            double mv = fabs(aug[i * w + i]);
            for (int k = i + 1; k < n; ++k) {
                const double t = fabs(aug[k * w + i]);
                if (t > mv) {
                    mv = t;
                    mr = k;
                }
            }
            scal[1] = mv < 1e-12 ? -1 : mr;
        }
        __syncthreads();
        const int mr = scal[1];
        if (mr < 0) {
            if (lane == 0) sing[S] = 1;
            return;
        }
        if (mr != i)
            for (int j = i + lane; j < w; j += 64) {
                const double t = aug[i * w + j];
                aug[i * w + j] = aug[mr * w + j];
                aug[mr * w + j] = t;
            }
        __syncthreads();
        const double piv = aug[i * w + i];
        __syncthreads();
        for (int j = i + lane; j < w; j += 64) aug[i * w + j] /= piv;
        for (int k = lane; k < n; k += 64) mult[k] = aug[k * w + i];
        __syncthreads();
        for (int e = lane; e < n * (w - i); e += 64) {
            const int k = e / (w - i), j = i + (e - k * (w - i));
            if (k != i) aug[k * w + j] -= mult[k] * aug[i * w + j];
        }
        __syncthreads();
    }
    for (int e = lane; e < n * n; e += 64) {
        const int r = e / n, c = e - r * n;
        table[(int64_t)S * K * K + r * n + c] = aug[r * w + n + c];
    }
    if (lane == 0) sing[S] = 0;
}

template <int KM>
__global__ __launch_bounds__(256) void bvls_thread_kernel(const double* __restrict__ A, const double* __restrict__ Ainv,
                                                          const double* __restrict__ B, double* __restrict__ X,
                                                          int64_t nprob, int K, double lo, double hi,
                                                          const double* __restrict__ table,
                                                          const int* __restrict__ sing) {
    __shared__ double sA[KM * KM], sAi[KM * KM];
    for (int e = threadIdx.x; e < K * K; e += blockDim.x) {
        sA[e] = A[e];
        sAi[e] = Ainv[e];
    }
    __syncthreads();
    const int64_t pid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= nprob) return;
    double b[KM], x[KM], bs[KM], y[KM];
#pragma unroll
    for (int i = 0; i < KM; ++i) b[i] = i < K ? B[pid * K + i] : 0.0;
#pragma unroll
    for (int i = 0; i < KM; ++i) {
        double s = 0.0;
        if (i < K)
            for (int j = 0; j < K; ++j) s += b[j] * sAi[i * K + j];
        x[i] = s;
    }
    unsigned F = (1u << K) - 1u, U = 0u;
    bool changed = false;
#pragma unroll
    for (int i = 0; i < KM; ++i)
        if (i < K) {
            if (x[i] < lo - 1e-8) {
                F &= ~(1u << i); changed = true;
            } else if (x[i] > hi + 1e-8) {
                F &= ~(1u << i); U |= 1u << i; changed = true;
            }
        }
    if (changed) {
        for (int it = 1; it < 50; ++it) {
#pragma unroll
            for (int i = 0; i < KM; ++i)
                if (i < K && ((F >> i) & 1u)) {
                    double s = b[i];
                    for (int j = 0; j < K; ++j)
                        if (!((F >> j) & 1u)) s -= sA[i * K + j] * (((U >> j) & 1u) ? hi : lo);
                    bs[i] = s;
                }
            if (sing[F]) break;
            const int n = __popc(F);
            const double* inv = table + (int64_t)F * K * K;
#pragma unroll
            for (int i = 0; i < KM; ++i)
                if (i < K && ((F >> i) & 1u)) {
                    const int r = __popc(F & ((1u << i) - 1u));
                    double s = 0.0;
                    int c = 0;
#pragma unroll
                    for (int j = 0; j < KM; ++j)
                        if (j < K && ((F >> j) & 1u)) s += inv[r * n + c++] * bs[j];
                    x[i] = s;
                }
#pragma unroll
            for (int i = 0; i < KM; ++i)
                if (i < K && !((F >> i) & 1u)) x[i] = ((U >> i) & 1u) ? hi : lo;
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                double s = -b[i];
                if (i < K)
                    for (int j = 0; j < K; ++j) s += sA[i * K + j] * x[j];
                y[i] = s;
            }
            changed = false;
#pragma unroll
            for (int i = 0; i < KM; ++i) {
                if (i >= K) continue;
                const unsigned bit = 1u << i;
                if (F & bit) {
                    if (x[i] < lo - 1e-8) { F &= ~bit; U &= ~bit; changed = true; }
                    else if (x[i] > hi + 1e-8) { F &= ~bit; U |= bit; changed = true; }
                } else if (!(U & bit)) {
                    if (y[i] < -1e-8) { F |= bit; U &= ~bit; changed = true; }
                } else if (y[i] > 1e-8) {
                    F |= bit; U &= ~bit; changed = true;
                }
            }
            if (!changed) break;
        }
    }
#pragma unroll
    for (int i = 0; i < KM; ++i)
        if (i < K) X[pid * K + i] = x[i];
}

__global__ void zero_u64_kernel(unsigned long long* p) { *p = 0ULL; }

static int persistent_blocks(int64_t nprob) {
    int64_t b = nprob < 4096 ? nprob : 4096;
    return (int)(b < 1 ? 1 : b);
}

}  // namespace adx

using namespace adx;

extern "C" {

int adx_sqp_solve(int simplex, const double* H, const double* r, const double* X, const double* vkk, double* Xn,
                  int64_t nprob, int K, void* workspace, int64_t workspace_bytes, void* stream) {
    if (K < 1 || K > 64) return fail_msg(1, "adx_sqp_solve: K must be in [1, 64]");
    if (nprob <= 0) return 0;
    if (!workspace || workspace_bytes < 8) return fail_msg(1, "adx_sqp_solve: needs 8 bytes of workspace");
    if (simplex && !vkk) return fail_msg(1, "adx_sqp_solve: simplex problems need v_kk");
    hipStream_t st = (hipStream_t)stream;
    auto* q = static_cast<unsigned long long*>(workspace);
    zero_u64_kernel<<<1, 1, 0, st>>>(q);
    const size_t lds = sizeof(double) * (size_t)SqpLds::doubles(K);
    if (simplex)
        sqp_kernel<true><<<persistent_blocks(nprob), 64, lds, st>>>(H, r, X, vkk, Xn, nprob, K, q);
    else
        sqp_kernel<false><<<persistent_blocks(nprob), 64, lds, st>>>(H, r, X, vkk, Xn, nprob, K, q);
    ADX_LAUNCH_CHECK("adx_sqp_solve");
    return 0;
}

int64_t adx_bvls_workspace_bytes(int K) {
    if (K < 1 || K > BVLS_FAST_K) return 8;
    const int64_t sets = (int64_t)1 << K;
    return 256 + sets * K * K * (int64_t)sizeof(double) + sets * (int64_t)sizeof(int);
}

int adx_bvls_batch(const double* A, const double* Ainv, const double* B, double* X, int64_t nprob, int K,
                   double lo, double hi, void* workspace, int64_t workspace_bytes, void* stream) {
    if (K < 1 || K > 64) return fail_msg(1, "adx_bvls_batch: K must be in [1, 64]");
    if (nprob <= 0) return 0;
    if (!workspace || workspace_bytes < 8) return fail_msg(1, "adx_bvls_batch: needs 8 bytes of workspace");
    hipStream_t st = (hipStream_t)stream;
    if (Ainv && K <= BVLS_FAST_K && workspace_bytes >= adx_bvls_workspace_bytes(K)) {
        const int64_t sets = (int64_t)1 << K;
        double* table = reinterpret_cast<double*>(static_cast<char*>(workspace) + 256);
        int* sing = reinterpret_cast<int*>(table + sets * K * K);
        bvls_subset_inv_kernel<<<(unsigned)sets, 64, 0, st>>>(A, K, table, sing);
        const unsigned blocks = (unsigned)ceil_div(nprob, 256);
        if (K <= 4)
            bvls_thread_kernel<4><<<blocks, 256, 0, st>>>(A, Ainv, B, X, nprob, K, lo, hi, table, sing);
        else if (K <= 8)
            bvls_thread_kernel<8><<<blocks, 256, 0, st>>>(A, Ainv, B, X, nprob, K, lo, hi, table, sing);
        else
            bvls_thread_kernel<12><<<blocks, 256, 0, st>>>(A, Ainv, B, X, nprob, K, lo, hi, table, sing);
        ADX_LAUNCH_CHECK("adx_bvls_batch");
        return 0;
    }
    auto* q = static_cast<unsigned long long*>(workspace);
    zero_u64_kernel<<<1, 1, 0, st>>>(q);
    const size_t lds = sizeof(double) * (size_t)BvlsLds::doubles(K) + sizeof(int) * (size_t)(3 * K + 4);
    bvls_kernel<<<persistent_blocks(nprob), 64, lds, st>>>(A, Ainv, B, X, nprob, K, lo, hi, q);
    ADX_LAUNCH_CHECK("adx_bvls_batch");
    return 0;
}

}  // extern "C"
