"""Python side of the CPU oracle: ctypes bindings to adx_oracle.c plus the
optimizer drivers restated from the reference in numpy.

TEST INFRASTRUCTURE ONLY (see the package docstring).  Each driver cites the
reference lines it follows (paths relative to the reference repository root).

Parity status of each function is listed in ``PINS`` at the bottom: what
known-answer data or mathematical property pins it, checked in
``tests/test_oracle.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_BUILD = _HERE / "_build"

_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_c_u64 = ctypes.c_uint64
_ptr = ctypes.c_void_p


def build(force: bool = False) -> None:
    """Compile the oracle shared objects with the committed Makefile."""
    if force or not (_BUILD / "liboracle.so").exists() or not (_BUILD / "liboracle_fast.so").exists():
        subprocess.run(["make", "-s"] + (["-B"] if force else []) + ["-C", str(_HERE)], check=True)


def _load(name: str) -> ctypes.CDLL:
    path = _BUILD / name
    if not path.exists():
        build()
    return ctypes.CDLL(str(path))


_LIB = None
_LIB_FAST = None

_SIGS = {
    "oc_bed_to_packed": (None, [_ptr, _c_i64, _c_i64, _c_i64, _ptr, _c_i64]),
    "oc_bed_to_unpacked": (None, [_ptr, _c_i64, _c_i64, _c_i64, _ptr]),
    "oc_pack": (None, [_ptr, _c_i64, _c_i64, _ptr, _c_i64]),
    "oc_unpack": (None, [_ptr, _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _ptr]),
    "oc_unpack_center": (None, [_ptr, _c_i64, _ptr, _c_i64, _c_i64, _c_i64, _c_i64, _ptr]),
    "oc_mean_packed": (_c_dbl, [_ptr, _c_i64, _c_i64, _c_i64]),
    "oc_flip_packed": (None, [_ptr, _c_i64, _c_i64, _c_i64]),
    "oc_freq_cpu": (None, [_ptr, _c_i64, _c_i64, _ptr]),
    "oc_freq_gpu": (None, [_ptr, _c_i64, _c_i64, _c_i64, _ptr]),
    "oc_qbat": (None, [_ptr, _c_i64, _c_i64, _c_i64, _ptr]),
    "oc_synth_packed": (None, [_ptr, _ptr, _c_i64, _c_i64, _c_int, _c_u64, _c_dbl, _ptr, _c_i64,
                               _c_i64, _c_i64]),
    "oc_em_stats_gpu": (None, [_ptr, _c_i64, _ptr, _ptr, _c_i64, _c_i64, _c_int, _ptr, _ptr, _ptr,
                               _ptr]),
    "oc_em_final_gpu": (None, [_ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _c_i64, _c_i64, _c_int, _ptr,
                               _ptr]),
    "oc_adam_step_gpu": (None, [_ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _c_i64, _c_i64,
                                _c_int, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_dbl]),
    "oc_loglik_gpu": (_c_dbl, [_ptr, _c_i64, _ptr, _ptr, _c_i64, _c_i64, _c_int]),
    "oc_P_step_cpu": (None, [_ptr, _ptr, _ptr, _ptr, _ptr, _ptr, _c_int, _c_i64, _c_i64, _c_int]),
    "oc_Q_step_cpu": (None, [_ptr, _ptr, _ptr, _ptr, _c_int, _c_i64]),
    "oc_adamP_cpu": (None, [_ptr, _ptr, _ptr, _ptr, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_int, _c_i64,
                            _c_int]),
    "oc_adamQ_cpu": (None, [_ptr, _ptr, _ptr, _ptr, _c_dbl, _c_dbl, _c_dbl, _c_dbl, _c_int, _c_i64,
                            _c_int]),
    "oc_loglik_cpu": (_c_dbl, [_ptr, _ptr, _ptr, _c_i64, _c_i64, _c_int]),
    "oc_mapQ": (None, [_ptr, _c_i64, _c_int]),
    "oc_mapP": (None, [_ptr, _c_i64, _c_int]),
    "oc_bvls_cpu": (None, [_ptr, _ptr, _ptr, _c_i64, _c_int, _c_dbl, _c_dbl]),
    "oc_bvls_gpu": (None, [_ptr, _ptr, _ptr, _ptr, _c_i64, _c_int, _c_dbl, _c_dbl]),
    "oc_qp_sweep": (_c_int, [_ptr, _ptr, _ptr, _c_int, _c_int]),
    "oc_project_simplex": (None, [_ptr, _c_int, _c_dbl]),
    "oc_sqp_q": (None, [_ptr, _ptr, _ptr, _ptr, _ptr, _c_i64, _c_int]),
    "oc_sqp_p": (None, [_ptr, _ptr, _ptr, _ptr, _c_i64, _c_int]),
    "oc_grad_hess_q": (None, [_ptr, _c_i64, _ptr, _ptr, _ptr, _ptr, _c_i64, _c_i64, _c_int]),
    "oc_grad_hess_p": (None, [_ptr, _c_i64, _ptr, _ptr, _ptr, _ptr, _c_i64, _c_i64, _c_int]),
    "oc_num_threads": (_c_int, []),
}


def _bind(lib: ctypes.CDLL) -> ctypes.CDLL:
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _bind(_load("liboracle.so"))
    return _LIB


def lib_fast() -> ctypes.CDLL:
    global _LIB_FAST
    if _LIB_FAST is None:
        _LIB_FAST = _bind(_load("liboracle_fast.so"))
    return _LIB_FAST


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# Genotype layout
# --------------------------------------------------------------------------

def bed_to_packed(bed: np.ndarray, M: int, N: int) -> np.ndarray:
    """snp_reader.pyx:76-120 read_bed_packed; ``bed`` is M x ceil(N/4)."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    out = np.zeros(((M + 3) // 4, N), dtype=np.uint8)
    lib().oc_bed_to_packed(_p(bed), M, bed.shape[1], N, _p(out), N)
    return out


def bed_to_unpacked(bed: np.ndarray, M: int, N: int) -> np.ndarray:
    """snp_reader.pyx:35-73 read_bed."""
    bed = np.ascontiguousarray(bed, dtype=np.uint8)
    out = np.zeros((M, N), dtype=np.uint8)
    lib().oc_bed_to_unpacked(_p(bed), M, bed.shape[1], N, _p(out))
    return out


def pack(G: np.ndarray) -> np.ndarray:
    """snp_reader.pyx:123-156 pack_genotypes."""
    G = np.ascontiguousarray(G, dtype=np.uint8)
    M, N = G.shape
    out = np.zeros(((M + 3) // 4, N), dtype=np.uint8)
    lib().oc_pack(_p(G), M, N, _p(out), N)
    return out


def unpack(Gp: np.ndarray, M: int, m_start: int = 0, C: int | None = None) -> np.ndarray:
    """pack2bit.cu:15-44: rows m_start .. m_start+C of the unpacked matrix."""
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    N = Gp.shape[1]
    C = M - m_start if C is None else min(C, M - m_start)
    out = np.zeros((C, N), dtype=np.uint8)
    lib().oc_unpack(_p(Gp), N, M, N, m_start, C, _p(out))
    return out


def unpack_center(Gp: np.ndarray, f: np.ndarray, M: int, m_start: int, C: int) -> np.ndarray:
    """pack2bit.cu:46-77."""
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    f = np.ascontiguousarray(f, dtype=np.float32)
    N = Gp.shape[1]
    C = min(C, M - m_start)
    out = np.zeros((C, N), dtype=np.float32)
    lib().oc_unpack_center(_p(Gp), N, _p(f), M, N, m_start, C, _p(out))
    return out


def mean_packed(Gp: np.ndarray, M: int) -> float:
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    return lib().oc_mean_packed(_p(Gp), Gp.shape[1], M, Gp.shape[1])


def flip_packed(Gp: np.ndarray, M: int) -> None:
    """snp_reader.pyx:254-302, in place."""
    assert Gp.flags["C_CONTIGUOUS"] and Gp.dtype == np.uint8
    lib().oc_flip_packed(_p(Gp), Gp.shape[1], M, Gp.shape[1])


def freq_cpu(G: np.ndarray) -> np.ndarray:
    G = np.ascontiguousarray(G, dtype=np.uint8)
    f = np.zeros(G.shape[0], dtype=np.float32)
    with np.errstate(all="ignore"):
        lib().oc_freq_cpu(_p(G), G.shape[0], G.shape[1], _p(f))
    return f


def freq_gpu(Gp: np.ndarray, M: int) -> np.ndarray:
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    f = np.zeros(M, dtype=np.float32)
    lib().oc_freq_gpu(_p(Gp), Gp.shape[1], M, Gp.shape[1], _p(f))
    return f


def qbat(Gp: np.ndarray, M: int) -> np.ndarray:
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    out = np.zeros(Gp.shape[1], dtype=np.float64)
    lib().oc_qbat(_p(Gp), Gp.shape[1], M, Gp.shape[1], _p(out))
    return out


def synth_packed(P: np.ndarray, Q: np.ndarray, seed: int, miss_rate: float,
                 row0: int = 0, rows: int | None = None) -> np.ndarray:
    """CPU twin of the device genotype generator (adx_synth_packed)."""
    P = _f64(P)
    Q = _f64(Q)
    M, K = P.shape
    N = Q.shape[0]
    Mb = (M + 3) // 4
    rows = Mb - row0 if rows is None else min(rows, Mb - row0)
    out = np.zeros((rows, N), dtype=np.uint8)
    lib().oc_synth_packed(_p(P), _p(Q), M, N, K, seed & 0xFFFFFFFFFFFFFFFF, miss_rate,
                          _p(out), N, row0, rows)
    return out


# --------------------------------------------------------------------------
# Adam-EM, GPU semantics (em_adam_gpu.py)
# --------------------------------------------------------------------------

def em_stats_gpu(Gp, P, Q, M):
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    P = _f64(P)
    Q = _f64(Q)
    K = P.shape[1]
    N = Q.shape[0]
    A = np.zeros((M, K))
    B = np.zeros((M, K))
    T = np.zeros((N, K))
    qb = np.zeros(N)
    lib().oc_em_stats_gpu(_p(Gp), Gp.shape[1], _p(P), _p(Q), M, N, K, _p(A), _p(B), _p(T), _p(qb))
    return A, B, T, qb


def em_final_gpu(P, Q, A, B, T, qb):
    P, Q, A, B, T, qb = map(_f64, (P, Q, A, B, T, qb))
    M, K = P.shape
    N = Q.shape[0]
    PE = np.zeros_like(P)
    QE = np.zeros_like(Q)
    lib().oc_em_final_gpu(_p(P), _p(Q), _p(A), _p(B), _p(T), _p(qb), M, N, K, _p(PE), _p(QE))
    return PE, QE


def loglik_gpu(Gp, P, Q, M) -> float:
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    P = _f64(P)
    Q = _f64(Q)
    return lib().oc_loglik_gpu(_p(Gp), Gp.shape[1], _p(P), _p(Q), M, Q.shape[0], P.shape[1])


class AdamEMGPU:
    """em_adam_gpu.py:121-224 EMAdamOptimizer restated on numpy buffers."""

    def __init__(self, M, N, K, lr, beta1, beta2, eps):
        self.lr, self.b1, self.b2, self.eps = lr, beta1, beta2, eps
        self.t = 0.0
        self.mP = np.zeros((M, K))
        self.vP = np.zeros((M, K))
        self.mQ = np.zeros((N, K))
        self.vQ = np.zeros((N, K))

    def run_em_step(self, Gp, P, Q, M):
        A, B, T, qb = em_stats_gpu(Gp, P, Q, M)
        return em_final_gpu(P, Q, A, B, T, qb)

    def step(self, P, Q, PE, QE):
        self.t += 1.0
        M, K = P.shape
        lib().oc_adam_step_gpu(_p(P), _p(Q), _p(_f64(PE)), _p(_f64(QE)), _p(self.mP), _p(self.vP),
                               _p(self.mQ), _p(self.vQ), M, Q.shape[0], K, self.t, self.lr,
                               self.b1, self.b2, self.eps)


def adamem_gpu(Gp, P, Q, M, *, lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, max_iter=10000,
               check=5, lr_decay=0.5, min_lr=1e-4, patience=3, tol=0.1, history=None, noise=None, labels=None):
    """em_adam_gpu.py:281-356 _optimize_parameters_gpu_once.

    Returns (P_best, Q_best).  If ``history`` is a list, (iteration, logL)
    pairs of every check are appended to it.  The third priming call of the
    reference (:322) only refills the EM target buffers, so it is omitted.
    ``noise = (seed, amp)`` multiplies every EM target by (1 + amp z),
    z ~ N(0, 1): with amp ~ 1e-16 it models the rounding differences of
    another FP64 summation order, to measure how far the trajectory moves
    under them (the chaos floor a GPU run is compared against).  With
    ``labels`` (y, one int per sample, 0 = unlabelled) it is supervised.py:
    342-462 optimize_supervised_gpu: the labelled rows of Q are snapped
    after every Adam step (:404-406).
    """
    P = _f64(P).copy()
    Q = _f64(Q).copy()
    N, K = Q.shape
    rng = None if noise is None else np.random.default_rng(noise[0])

    def perturb(PE, QE):
        if rng is None:
            return PE, QE
        return (PE * (1.0 + noise[1] * rng.standard_normal(PE.shape)),
                QE * (1.0 + noise[1] * rng.standard_normal(QE.shape)))

    opt = AdamEMGPU(M, N, K, lr, beta1, beta2, reg_adam)
    PE, QE = opt.run_em_step(Gp, P, Q, M)
    opt.step(P, Q, PE, QE)
    if labels is not None:
        snap_q(Q, labels, K)
    L_best = loglik_gpu(Gp, P, Q, M)
    if history is not None:
        history.append((0, L_best))
    P_best, Q_best = P.copy(), Q.copy()
    wait = 0
    for it in range(max_iter):
        PE, QE = perturb(*opt.run_em_step(Gp, P, Q, M))
        opt.step(P, Q, PE, QE)
        if labels is not None:
            snap_q(Q, labels, K)
        if (it + 1) % check == 0:
            L = loglik_gpu(Gp, P, Q, M)
            if history is not None:
                history.append((it + 1, L))
            if L > L_best + tol:
                L_best = L
                P_best[:] = P
                Q_best[:] = Q
                wait = 0
            else:
                wait += 1
                if wait >= patience:
                    opt.lr = max(opt.lr * lr_decay, min_lr)
                    if opt.lr <= min_lr:
                        break
                    wait = 0
    return P_best, Q_best


def projection_gpu(Gp, P, Q, M, *, lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, max_iter=10000, check=5,
                   lr_decay=0.5, min_lr=1e-4, patience=3, tol=0.1, history=None, noise=None):
    """adamixture/src/projection.py:178-321 optimize_projection_gpu: P fixed,
    Adam-EM on Q only (the P-side EM output is discarded).  The reference's
    third priming EM step (:300) does not change any state and is omitted.
    ``noise`` as in adamem_gpu (applied to the EM targets of the loop)."""
    P = _f64(P)
    Q = _f64(Q).copy()
    N, K = Q.shape
    mQ = np.zeros((N, K))
    vQ = np.zeros((N, K))
    t = 0.0
    dummy = np.zeros((1, K))
    rng = None if noise is None else np.random.default_rng(noise[0])
    primed = False

    def em_q():
        A, B, T, qb = em_stats_gpu(Gp, P, Q, M)
        QE = em_final_gpu(P, Q, A, B, T, qb)[1]
        if rng is not None and primed:
            QE = QE * (1.0 + noise[1] * rng.standard_normal(QE.shape))
        return QE

    def q_step(QE, lr_now):
        nonlocal t
        t += 1.0
        lib().oc_adam_step_gpu(_p(dummy), _p(Q), _p(dummy), _p(_f64(QE)), _p(dummy), _p(dummy), _p(mQ), _p(vQ),
                               0, N, K, t, lr_now, beta1, beta2, reg_adam)

    q_step(em_q(), lr)
    primed = True
    L_best = loglik_gpu(Gp, P, Q, M)
    if history is not None:
        history.append((0, L_best))
    Q_best = Q.copy()
    wait = 0
    for it in range(max_iter):
        q_step(em_q(), lr)
        if (it + 1) % check == 0:
            L = loglik_gpu(Gp, P, Q, M)
            if history is not None:
                history.append((it + 1, L))
            if L > L_best + tol:
                L_best = L
                Q_best[:] = Q
                wait = 0
            else:
                wait += 1
                if wait >= patience:
                    lr = max(lr * lr_decay, min_lr)
                    if lr <= min_lr:
                        break
                    wait = 0
    return Q_best


# --------------------------------------------------------------------------
# Adam-EM, CPU semantics (em_adam.py, em.pyx, tools.pyx)
# --------------------------------------------------------------------------

def loglik_cpu(G, P, Q) -> float:
    """tools.pyx:59-85 on unpacked G."""
    G = np.ascontiguousarray(G, dtype=np.uint8)
    P = _f64(P)
    Q = _f64(Q)
    with np.errstate(all="ignore"):
        return lib().oc_loglik_cpu(_p(G), _p(P), _p(Q), G.shape[0], G.shape[1], P.shape[1])


def em_step_cpu(G, P0, Q0, T, P1, Q1, qb, nthreads=1, fast=False):
    """em_adam.py:56-80 emStep core: P_step then Q_step (targets in P1/Q1)."""
    L = lib_fast() if fast else lib()
    M, K = P0.shape
    N = Q0.shape[0]
    L.oc_P_step_cpu(_p(G), _p(P0), _p(P1), _p(Q0), _p(T), _p(qb), K, M, N, nthreads)
    L.oc_Q_step_cpu(_p(Q0), _p(Q1), _p(T), _p(qb), K, N)


def adamem_cpu(G, P, Q, *, lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, max_iter=10000,
               check=5, lr_decay=0.5, min_lr=1e-4, patience=3, tol=0.1, history=None, nthreads=1, fast=False):
    """em_adam.py:82-179 optimize_parameters (CPU Adam-EM).  ``nthreads`` /
    ``fast``: the EM statistics loop's OpenMP threads and the build with the
    reference's fast-math flags (timing only; the default is the checker)."""
    G = np.ascontiguousarray(G, dtype=np.uint8)
    P = _f64(P).copy()
    Q = _f64(Q).copy()
    M, K = P.shape
    N = Q.shape[0]
    mP, vP, mQ, vQ = (np.zeros_like(P), np.zeros_like(P), np.zeros_like(Q), np.zeros_like(Q))
    P1, Q1, T, qb = np.zeros_like(P), np.zeros_like(Q), np.zeros_like(Q), np.zeros(N)
    t = 0

    def em_step():
        em_step_cpu(G, P, Q, T, P1, Q1, qb, nthreads, fast)
        P[:] = P1
        Q[:] = Q1

    def adam_step():
        nonlocal t
        em_step_cpu(G, P, Q, T, P1, Q1, qb, nthreads, fast)
        t += 1
        lib().oc_adamP_cpu(_p(P), _p(P1), _p(mP), _p(vP), lr_cur[0], beta1, beta2, reg_adam, t, M, K)
        lib().oc_adamQ_cpu(_p(Q), _p(Q1), _p(mQ), _p(vQ), lr_cur[0], beta1, beta2, reg_adam, t, N, K)

    lr_cur = [lr]
    em_step()
    adam_step()
    em_step()
    L_best = loglik_cpu(G, P, Q)
    if history is not None:
        history.append((0, L_best))
    P_best, Q_best = P.copy(), Q.copy()
    wait = 0
    for it in range(max_iter):
        adam_step()
        if (it + 1) % check == 0:
            L = loglik_cpu(G, P, Q)
            if history is not None:
                history.append((it + 1, L))
            if L > L_best + tol:
                L_best = L
                P_best[:] = P
                Q_best[:] = Q
                wait = 0
            else:
                wait += 1
                if wait >= patience:
                    lr_cur[0] = max(lr_cur[0] * lr_decay, min_lr)
                    if lr_cur[0] <= min_lr:
                        break
                    wait = 0
    return P_best, Q_best


# --------------------------------------------------------------------------
# Randomized SVD and ALS initialisation (svd.py, svd_gpu.py, als.py, als_gpu.py)
# --------------------------------------------------------------------------

def eig_svd(X: np.ndarray):
    """svd.py:13-28 eigSVD: SVD of X (m x n) via eigh of X^T X, descending."""
    D, V = np.linalg.eigh(X.T @ X)
    S = np.sqrt(D)
    U = X @ (V * (1.0 / S))
    return (np.ascontiguousarray(U[:, ::-1]), np.ascontiguousarray(S[::-1]),
            np.ascontiguousarray(V[:, ::-1]))


def rsvd(X_rows, M: int, N: int, k: int, seed: int, power: int, tol: float, chunk: int,
         proj_basis: np.ndarray | None = None):
    """svd.py:30-140 _RSVD_once on the centred genotype matrix.

    ``X_rows(start, n)`` returns the centred FP32 rows [start, start+n) (the
    decompress_block of tools.pyx:292-324).  ``proj_basis`` (M x L float32)
    replaces the seeded standard-normal draw when given.
    """
    rng = np.random.default_rng(seed)
    n_batches = int(np.ceil(M / chunk))
    alpha = 0.0
    L = max(k + 10, 20)
    accum = np.zeros((N, L), dtype=np.float32)
    if proj_basis is None:
        proj_basis = rng.standard_normal(size=(M, L), dtype=np.float32)
    else:
        proj_basis = np.array(proj_basis, dtype=np.float32, copy=True)
    for w in range(n_batches):
        s = w * chunk
        n = min(chunk, M - s)
        accum += X_rows(s, n).T @ proj_basis[s:s + n]
    orth, _, _ = eig_svd(accum)
    accum.fill(0.0)
    sv = np.zeros(L, dtype=np.float32)
    for i in range(power):
        for w in range(n_batches):
            s = w * chunk
            n = min(chunk, M - s)
            Xc = X_rows(s, n)
            proj_basis[s:s + n] = Xc @ orth
            accum += Xc.T @ proj_basis[s:s + n]
        accum -= alpha * orth
        orth, s_vals, _ = eig_svd(accum)
        accum.fill(0.0)
        if i > 0:
            s_cur = s_vals + alpha
            rel = np.abs(s_cur - sv[:len(s_cur)]) / np.maximum(s_cur, 1e-12)
            if np.max(rel[:k]) < tol:
                break
            sv[:len(s_cur)] = s_cur
        else:
            sv[:len(s_vals)] = s_vals + alpha
        if s_vals[-1] > alpha:
            alpha = 0.5 * (s_vals[-1] + alpha)
    for w in range(n_batches):
        s = w * chunk
        n = min(chunk, M - s)
        proj_basis[s:s + n] = X_rows(s, n) @ orth
    u, s_all, v = eig_svd(proj_basis)
    return (np.ascontiguousarray(u[:, :k]), np.ascontiguousarray(s_all[:k]),
            np.ascontiguousarray(orth @ v[:, :k]))


def bvls_cpu(A, B, lo=0.0, hi=1.0):
    """bvls.pyx:268-312 batch_bvls_bpp."""
    A = _f64(A)
    B = _f64(B)
    X = np.zeros_like(B)
    lib().oc_bvls_cpu(_p(A), _p(B), _p(X), B.shape[0], B.shape[1], lo, hi)
    return X


def bvls_gpu(A, B, lo=0.0, hi=1.0):
    """bvls_kernel.cu:188-218 batch_bvls_bpp_cuda: X0 = B inv(A)^T, then BPP."""
    A = _f64(A)
    B = _f64(B)
    X0 = _f64(B @ np.linalg.inv(A).T)
    X = np.zeros_like(B)
    lib().oc_bvls_gpu(_p(A), _p(B), _p(X0), _p(X), B.shape[0], B.shape[1], lo, hi)
    return X


def map_p(P):
    P = _f64(P)
    lib().oc_mapP(_p(P), P.shape[0], P.shape[1])
    return P


def map_q(Q):
    Q = _f64(Q)
    lib().oc_mapQ(_p(Q), Q.shape[0], Q.shape[1])
    return Q


def map_pq_gpu(P: np.ndarray, Q: np.ndarray):
    """br_qn_gpu.py:47-61 _mapPQ_gpu: P and Q clipped to [1e-5, 1 - 1e-5],
    then the rows of Q divided by their sums.  Returns new arrays."""
    P = np.clip(P, 1e-5, 1 - 1e-5)
    Q = np.clip(Q, 1e-5, 1 - 1e-5)
    return P, Q / Q.sum(axis=1, keepdims=True)


def als(U, S, V, f, seed, M, N, K, max_iter, tol, P_init=None, gpu_semantics=False):
    """als.py:10-88 ALS (CPU), or als_gpu.py:45-127 when ``gpu_semantics``.

    ``P_init`` (M x K) replaces the seeded uniform draw (the GPU path draws
    it from torch's device generator, als_gpu.py:69-86, which no CPU code can
    reproduce)."""
    Z = _f64(U * S)
    V = _f64(V)
    f64 = np.asarray(f, dtype=np.float64)
    if P_init is None:
        P = np.random.default_rng(seed).random(size=(M, K), dtype=np.float64)
    else:
        P = _f64(P_init).copy()
    P = map_p(P)
    solver = bvls_gpu if gpu_semantics else bvls_cpu
    I_p = P @ np.linalg.pinv(P.T @ P)
    Q = 0.5 * (V @ (Z.T @ I_p)) + (f64 @ I_p)
    Q = map_q(Q)
    Q0 = Q.copy()
    for _ in range(max_iter):
        A_q = Q.T @ Q
        I_q = Q @ np.linalg.pinv(A_q)
        P_free = 0.5 * (Z @ (V.T @ I_q)) + np.outer(f64, I_q.sum(axis=0))
        P = map_p(solver(A_q, P_free @ A_q))
        A_p = P.T @ P
        I_p = P @ np.linalg.pinv(A_p)
        Q_free = 0.5 * (V @ (Z.T @ I_p)) + (f64 @ I_p)
        Q = map_q(solver(A_p, Q_free @ A_p))
        rmse = np.sqrt(np.mean((Q - Q0) ** 2))
        if rmse < tol:
            break
        Q0[:] = Q
    return P, Q


# --------------------------------------------------------------------------
# SQP + ZAL quasi-Newton (sqp.pyx, br_qn.pyx, br_qn.py, br_qn_gpu.py)
# --------------------------------------------------------------------------

def v_kk(K: int) -> np.ndarray:
    """br_qn.py:98-99: right singular vectors of ones(1, K), transposed."""
    _, _, vt = np.linalg.svd(np.ones((1, K)), full_matrices=True)
    return np.ascontiguousarray(vt.T, dtype=np.float64)


def grad_hess_q(Gp, Q, P, M):
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    Q = _f64(Q)
    P = _f64(P)
    N, K = Q.shape
    H = np.zeros((N, K, K))
    z = np.zeros((N, K))
    lib().oc_grad_hess_q(_p(Gp), Gp.shape[1], _p(Q), _p(P), _p(H), _p(z), M, N, K)
    return H, z


def grad_hess_p(Gp, Q, P, M):
    Gp = np.ascontiguousarray(Gp, dtype=np.uint8)
    Q = _f64(Q)
    P = _f64(P)
    N, K = Q.shape
    H = np.zeros((M, K, K))
    z = np.zeros((M, K))
    lib().oc_grad_hess_p(_p(Gp), Gp.shape[1], _p(Q), _p(P), _p(H), _p(z), M, N, K)
    return H, z


def sqp_q(H, z, Q, vkk):
    H, z, Q, vkk = map(_f64, (H, z, Q, vkk))
    out = np.zeros_like(Q)
    lib().oc_sqp_q(_p(H), _p(z), _p(Q), _p(out), _p(vkk), Q.shape[0], Q.shape[1])
    return out


def sqp_p(H, z, P):
    H, z, P = map(_f64, (H, z, P))
    out = np.zeros_like(P)
    lib().oc_sqp_p(_p(H), _p(z), _p(P), _p(out), P.shape[0], P.shape[1])
    return out


def qp_sweep(tableau: np.ndarray, par: np.ndarray, p: int, c: int):
    tab = _f64(tableau).copy()
    par = _f64(par)
    delta = np.zeros(p)
    it = lib().oc_qp_sweep(_p(delta), _p(tab), _p(par), p, c)
    return delta, it


def project_simplex(b: np.ndarray, pc: float = 1e-5) -> np.ndarray:
    b = _f64(b).copy()
    lib().oc_project_simplex(_p(b), b.shape[0], pc)
    return b


def zal_extrapolate(U: np.ndarray, V: np.ndarray, x: np.ndarray, x_next: np.ndarray) -> np.ndarray:
    """br_qn.pyx:22-89 qn_extrapolate_ZAL with columns U, V (dim x q)."""
    lhs = U.T @ (U - V)
    rhs = U.T @ (x - x_next)
    alpha = np.linalg.solve(lhs, rhs)
    return x_next - V @ alpha


def brqn(Gp, P, Q, M, *, max_iter=10000, tol=0.1, Q_hist=3, patience=3, vkk=None, history=None,
         gpu_semantics=False, labels=None):
    """br_qn.py:73-205 optimize_original (CPU semantics, packed G), or with
    ``gpu_semantics`` br_qn_gpu.py:217-382, which maps the quasi-Newton point
    by clamp + row renormalisation (_mapPQ_gpu, :47-61) instead of the box
    clip + simplex projection of br_qn.py:159-160.  With ``labels`` (and
    gpu_semantics) it is supervised.py:602-781 optimize_supervised_original_gpu:
    labelled Q rows snapped after each Q solve (:689, :697) and after the
    quasi-Newton map (:741)."""
    snap = (lambda X: snap_q(X, labels, X.shape[1])) if labels is not None else (lambda X: None)
    P = _f64(P).copy()
    Q = _f64(Q).copy()
    N, K = Q.shape
    vkk = v_kk(K) if vkk is None else _f64(vkk)
    dim = (M + N) * K
    U = np.zeros((dim, Q_hist))
    V = np.zeros((dim, Q_hist))
    ll_prev = -np.inf
    ll_best = -np.inf
    wait = 0
    P_best, Q_best = P.copy(), Q.copy()
    for it in range(1, max_iter + 1):
        H, z = grad_hess_q(Gp, Q, P, M)
        Qn = sqp_q(H, z, Q, vkk)
        snap(Qn)
        H, z = grad_hess_p(Gp, Qn, P, M)
        Pn = sqp_p(H, z, P)
        H, z = grad_hess_q(Gp, Qn, Pn, M)
        Qn2 = sqp_q(H, z, Qn, vkk)
        snap(Qn2)
        H, z = grad_hess_p(Gp, Qn2, Pn, M)
        Pn2 = sqp_p(H, z, Pn)
        x = np.concatenate([P.ravel(), Q.ravel()])
        x1 = np.concatenate([Pn.ravel(), Qn.ravel()])
        x2 = np.concatenate([Pn2.ravel(), Qn2.ravel()])
        col = (it - 1) % Q_hist
        U[:, col] = x1 - x
        V[:, col] = x2 - x1
        nc = min(it, Q_hist)
        xq = zal_extrapolate(U[:, :nc], V[:, :nc], x, x1)
        Pq = np.clip(xq[:M * K].reshape(M, K), 1e-5, 1 - 1e-5)
        Qq = xq[M * K:].reshape(N, K).copy()
        if gpu_semantics:
            Pq, Qq = map_pq_gpu(Pq, Qq)
        else:
            for i in range(N):
                Qq[i] = project_simplex(Qq[i])
        snap(Qq)
        llq = loglik_gpu(Gp, Pq, Qq, M)
        if llq > ll_prev:
            P, Q, ll_new = Pq, Qq, llq
        else:
            P, Q = Pn2, Qn2
            ll_new = loglik_gpu(Gp, P, Q, M)
        best_before = ll_best
        if ll_new > ll_best:
            ll_best = ll_new
            P_best, Q_best = P.copy(), Q.copy()
        if best_before != -np.inf and ll_new <= best_before + tol:
            wait += 1
        else:
            wait = 0
        if history is not None:
            history.append((it, ll_new))
        if abs(ll_new - ll_prev) < tol or wait >= patience:
            break
        ll_prev = ll_new
    return P_best, Q_best


def projection_brqn_gpu(Gp, P, Q, M, *, max_iter=10000, tol=0.1, Q_hist=3, patience=3, history=None):
    """adamixture/src/projection.py:505-659 optimize_projection_original_gpu:
    P fixed; per iteration two SQP sweeps on Q, ZAL extrapolation over the
    flattened N K vector (:589-617), clamp + row renormalisation (:620-621),
    acceptance against the previous iteration's log-likelihood (:624-630)
    and its stopping rules (:632-656)."""
    P = _f64(P)
    Q = _f64(Q).copy()
    N, K = Q.shape
    vkk = v_kk(K)
    U = np.zeros((N * K, Q_hist))
    V = np.zeros((N * K, Q_hist))
    ll_prev = -np.inf
    ll_best = -np.inf
    wait = 0
    Q_best = Q.copy()
    for it in range(1, max_iter + 1):
        H, z = grad_hess_q(Gp, Q, P, M)
        Q1 = sqp_q(H, z, Q, vkk)
        H, z = grad_hess_q(Gp, Q1, P, M)
        Q2 = sqp_q(H, z, Q1, vkk)
        x, x1, x2 = Q.ravel(), Q1.ravel(), Q2.ravel()
        col = (it - 1) % Q_hist
        U[:, col] = x1 - x
        V[:, col] = x2 - x1
        nc = min(it, Q_hist)
        Us, Vs = U[:, :nc], V[:, :nc]
        lhs = Us.T @ (Us - Vs)
        rhs = Us.T @ (x - x1)
        try:
            alpha = np.linalg.solve(lhs, rhs)
        except np.linalg.LinAlgError:
            alpha = np.linalg.lstsq(lhs, rhs, rcond=None)[0]
        Qq = np.clip((x1 - Vs @ alpha).reshape(N, K), 1e-5, 1 - 1e-5)
        Qq /= Qq.sum(axis=1, keepdims=True)
        llq = loglik_gpu(Gp, P, Qq, M)
        if llq > ll_prev:
            Q, ll_new = Qq, llq
        else:
            Q = Q2
            ll_new = loglik_gpu(Gp, P, Q, M)
        best_before = ll_best
        if ll_new > ll_best:
            ll_best = ll_new
            Q_best = Q.copy()
        if best_before != -np.inf and ll_new <= best_before + tol:
            wait += 1
        else:
            wait = 0
        if history is not None:
            history.append((it, ll_new))
        if abs(ll_new - ll_prev) < tol or wait >= patience:
            break
        ll_prev = ll_new
    return Q_best


def labels_to_y(raw_labels):
    """supervised_entry.py:162-188: '-' -> 0, other names -> 1, 2, ... in
    order of first appearance."""
    ids = {}
    y = [0 if lbl == "-" else ids.setdefault(lbl, len(ids) + 1) for lbl in raw_labels]
    return np.array(y, dtype=np.int32), ids


def init_p_supervised(Gp, y, K, M, eps=1e-5):
    """supervised.py:84-107 init_p_supervised_packed: per-population allele
    frequencies of the labelled samples from the packed matrix."""
    Gp = np.asarray(Gp)
    y = np.asarray(y)
    P = np.full((M, K), 0.5)
    rows = np.arange(M)
    shift = (2 * (rows & 3)).astype(np.uint8)[:, None]
    for k in range(K):
        cols = np.flatnonzero(y == k + 1)
        if cols.size == 0:
            continue
        codes = (Gp[(rows >> 2)[:, None], cols[None, :]] >> shift) & np.uint8(3)
        obs = codes != 3
        n_obs = obs.sum(axis=1)
        ok = n_obs > 0
        s = np.where(obs, codes, 0).astype(np.float64).sum(axis=1)
        P[ok, k] = s[ok] / (2.0 * n_obs[ok])
    return P.clip(eps, 1.0 - eps)


def init_p_supervised_unpacked(G, y, K, M, eps=1e-5):
    """supervised.py:49-81 init_p_supervised (unpacked G, missing = 3)."""
    y = np.asarray(y)
    P = np.full((M, K), 0.5)
    for k in range(K):
        sel = y == k + 1
        if not sel.any():
            continue
        g = G[:, sel].astype(np.float64)
        miss = g == 3
        g[miss] = 0.0
        den = 2.0 * (sel.sum() - miss.sum(axis=1))
        ok = den > 0
        P[ok, k] = g.sum(axis=1)[ok] / den[ok]
    return P.clip(eps, 1.0 - eps)


def init_q_supervised(Q, y, K, eps=1e-5):
    """supervised.py:20-46: labelled rows near-one-hot, unlabelled rows
    renormalised (in place)."""
    y = np.asarray(y)
    snap_q(Q, y, K, eps)
    un = y == 0
    Q[un] /= Q[un].sum(axis=1, keepdims=True)


def snap_q(Q, y, K, eps=1e-5):
    """supervised.py:112-131 / :308-339: labelled rows set to eps / max(K-1, 1)
    with 1 - eps on the labelled population (in place)."""
    y = np.asarray(y)
    lab = np.flatnonzero(y > 0)
    Q[lab] = eps / max(K - 1, 1)
    Q[lab, y[lab] - 1] = 1.0 - eps


# --------------------------------------------------------------------------
# Cross-validation (adamixture/src/cv.py, br_qn.pyx:92-246)
# --------------------------------------------------------------------------

SPLITMIX = (0x9E3779B97F4A7C15, 0xBF58476D1CE4E5B9, 0x94D049BB133111EB)


def splitmix64(x) -> np.ndarray:
    """The first output of the splitmix64 generator seeded with x (Steele,
    Lea & Flood 2014; Vigna's reference code): z = x + 0x9E3779B97F4A7C15,
    then xor-shift 30 / multiply / xor-shift 27 / multiply / xor-shift 31.
    cv.py:61-81 _hash_fold_mask hashes flat + seed exactly this way."""
    with np.errstate(over="ignore"):
        z = np.asarray(x, dtype=np.uint64) + np.uint64(SPLITMIX[0])
        z ^= z >> np.uint64(30)
        z *= np.uint64(SPLITMIX[1])
        z ^= z >> np.uint64(27)
        z *= np.uint64(SPLITMIX[2])
        z ^= z >> np.uint64(31)
    return z


def cv_hash_folds(flat, seed: int, n_folds: int) -> np.ndarray:
    """Fold id of each flat index (m N + i) on the streaming path."""
    with np.errstate(over="ignore"):
        return (splitmix64(np.asarray(flat, dtype=np.uint64) + np.uint64(seed)) % np.uint64(n_folds)).astype(np.int64)


def cv_shuffled_entries(G: np.ndarray, seed: int) -> np.ndarray:
    """cv.py:27-42, :214-219: the observed flat indices as int32, shuffled by
    numpy's default_rng(seed); fold f is shuffled[f::n_folds]."""
    flat = np.flatnonzero(G != 3).astype(np.int32)
    np.random.default_rng(seed).shuffle(flat)
    return flat


def cv_deviance_sum(codes, flat, P, Q, N: int) -> float:
    """br_qn.pyx:92-149: sum over held-out entries of g log(g/mu) +
    (2-g) log((2-g)/(2-mu)), mu = 2 Q_i.P_m clamped to [1e-10, 2 - 1e-10],
    with 0 log 0 = 0.  (The reference sums in an OpenMP reduction.)"""
    flat = np.asarray(flat, dtype=np.int64)
    g = np.asarray(codes, dtype=np.float64)
    row, col = flat // N, flat % N
    mu = 2.0 * np.einsum("ek,ek->e", _f64(Q)[col], _f64(P)[row])
    mu = np.clip(mu, 1e-10, 2.0 - 1e-10)
    with np.errstate(divide="ignore", invalid="ignore"):
        ta = np.where(g > 0, g * np.log(g / mu), 0.0)
        tb = np.where(2.0 - g > 0, (2.0 - g) * np.log((2.0 - g) / (2.0 - mu)), 0.0)
    return float(np.sum(ta + tb))


def cv_polish(Gp, P, Q, M):
    """cv.py:146-168 + br_qn.py:50-70 polish_sqp_qn: 5 SQP + ZAL iterations
    of the CPU optimizer with early stopping disabled (tol 0, patience 6)."""
    return brqn(Gp, P, Q, M, max_iter=5, tol=0.0, Q_hist=3, patience=6, gpu_semantics=False)


def cross_validation(G: np.ndarray, P, Q, seed: int, n_folds: int, streaming: bool | None = None) -> float:
    """cv.py:171-267 run_cross_validation on an unpacked M x N matrix.
    ``streaming`` forces the hashed fold path (the reference takes it when
    M N exceeds the int32 range)."""
    M, N = G.shape
    if streaming is None:
        streaming = M * N > np.iinfo(np.int32).max
    observed = np.flatnonzero(G != 3)
    if observed.size == 0:
        raise ValueError("No non-missing genotypes available for cross-validation.")
    if streaming:
        fold_of = cv_hash_folds(observed, seed, n_folds)
    else:
        shuffled = cv_shuffled_entries(G, seed)
    total = 0.0
    for fold in range(n_folds):
        held = observed[fold_of == fold] if streaming else shuffled[fold::n_folds]
        if held.size == 0:
            continue
        Gm = G.copy()
        Gm.ravel()[held] = 3
        Pc, Qc = cv_polish(pack(Gm), P, Q, M)
        total += cv_deviance_sum(G.ravel()[held], held, Pc, Qc, N)
    return total / float(observed.size)


# --------------------------------------------------------------------------
# VCF / BCF genotype decoding (snp_reader.pyx:305-463, :829-1108)
# --------------------------------------------------------------------------

def _vcf_allele(s: bytes, q: int):
    """snp_reader.pyx:305-341: (allele or 255 for missing, next position);
    bytes past the end read as NUL."""
    c = s[q] if q < len(s) else 0
    if c in (46, 45, 0):
        return 255, (q + 1 if c != 0 else q)
    start, val = q, 0
    while q < len(s) and 48 <= s[q] <= 57:
        val = (val * 10 + s[q] - 48) & 0xFF
        q += 1
    return (255 if q == start else val), q


def vcf_gt_code(line: bytes, q: int) -> int:
    """snp_reader.pyx:343-375 on the sample field starting at line[q]; the
    line keeps its trailing '\\n' and reads as NUL beyond it."""
    a1, q = _vcf_allele(line, q)
    if a1 == 255:
        return 3
    sep = line[q] if q < len(line) else 0
    if sep in (58, 9, 10, 0):
        return a1 * 2 if a1 <= 1 else 3
    a2, q = _vcf_allele(line, q + 1)
    if a2 == 255:
        return 3
    t = (a1 + a2) & 0xFF
    return t if t <= 2 else 3


def vcf_to_unpacked(data: bytes, keep=lambda tok: True):
    """read_vcf_file_packed's parse (snp_reader.pyx:446-462, 710-827) on a
    decompressed VCF: (M x N codes, N, M); samples a short line lacks are
    missing.  ``keep(chrom_token)`` filters lines."""
    N, rows = 0, []
    for line in data.split(b"\n")[:-1] if data.endswith(b"\n") else data.split(b"\n"):
        line = line + b"\n"
        if line.startswith(b"#"):
            if line.startswith((b"#CHROM", b"#chrom")):
                N = len(line.rstrip(b"\n").split(b"\t")) - 9
            continue
        if not keep(line.split(b"\t", 1)[0].rstrip(b"\n")):
            continue
        row = np.full(N, 3, dtype=np.uint8)
        tabs = [k for k, c in enumerate(line) if c == 9]
        for i in range(N):
            q = tabs[8 + i] + 1 if 8 + i < len(tabs) else None
            if q is not None:
                row[i] = vcf_gt_code(line, q)
        rows.append(row)
    return np.array(rows, dtype=np.uint8).reshape(len(rows), N), N, len(rows)


def bcf_to_unpacked(data: bytes, keep=lambda tok: True):
    """read_bcf_file (snp_reader.pyx:859-1080) on a decompressed BCF: (M x N
    values, N, M).  Values are a1 + a2 (a1 for one value per sample), 3
    where an allele is missing; the reference's packed reader keeps the low
    2 bits of them."""
    import struct as _st
    l_text = _st.unpack_from("<I", data, 5)[0]
    header = data[9:9 + l_text].decode("utf-8", errors="ignore")
    contigs, samples, gt_idx = {}, [], -1
    for line in header.splitlines():
        if line.startswith("##contig=") and "ID=" in line:
            contigs[len(contigs)] = line.split("ID=")[1].split(",")[0].split(">")[0]
        elif line.startswith(("#CHROM", "#chrom")):
            samples = line.split("\t")[9:]
    for line in header.splitlines():
        if line.startswith("##FORMAT=") and "ID=GT" in line:
            if "IDX=" in line:
                gt_idx = int(line.split("IDX=")[1].split(",")[0].split(">")[0])
            break
    N = len(samples)

    def typed_int(pos):
        b = data[pos]
        t, n = b & 15, b >> 4
        pos += 1
        if n == 15:
            _, pos = typed_int(pos)
        fmt = {1: "<b", 2: "<h", 3: "<i"}.get(t)
        if fmt is None:
            return 0, pos
        return _st.unpack_from(fmt, data, pos)[0], pos + _st.calcsize(fmt)

    rows, off = [], 9 + l_text
    while off + 8 <= len(data):
        l_shared, l_indiv = _st.unpack_from("<II", data, off)
        base = off + 8
        chrom = contigs.get(_st.unpack_from("<i", data, base)[0], "")
        if keep(chrom.encode()):
            n_fmt = _st.unpack_from("<I", data, base + 20)[0] >> 24
            pos, gt = base + l_shared, None
            for _ in range(n_fmt):
                key, pos = typed_int(pos)
                b = data[pos]
                t, n = b & 15, b >> 4
                pos += 1
                if n == 15:
                    n, pos = typed_int(pos)
                sz = 1 if t in (1, 7) else (2 if t == 2 else 4)
                if key == gt_idx or gt_idx < 0:
                    gt = (pos, n, sz)
                    break
                pos += N * n * sz
            row = np.full(N, 3, dtype=np.int64)
            if gt is not None and gt[1] >= 1 and gt[0] + N * gt[1] * gt[2] <= len(data):
                p, n, sz = gt
                vals = np.frombuffer(data, dtype={1: "<i1", 2: "<i2", 4: "<i4"}[sz], count=N * n, offset=p)
                vals = vals.reshape(N, n).astype(np.int64)
                a1 = (vals[:, 0] >> 1) - 1
                if n >= 2:
                    a2 = (vals[:, 1] >> 1) - 1
                    row = np.where((a1 < 0) | (a2 < 0), 3, a1 + a2)
                else:
                    row = np.where(a1 < 0, 3, a1)
            rows.append(row)
        off = base + l_shared + l_indiv
    return np.array(rows, dtype=np.int64).reshape(len(rows), N), N, len(rows)


# --------------------------------------------------------------------------
# PGEN hardcall decoding (the reference reads PGEN through pgenlib >= 0.94.1,
# snp_reader.pyx:1110-1135; restated from the published PLINK 2 .pgen format)
# --------------------------------------------------------------------------

def _varint(buf: bytes, p: int):
    v, sh = 0, 0
    while True:
        b = buf[p]
        p += 1
        v |= (b & 0x7F) << sh
        sh += 7
        if b < 0x80:
            return v, p


def pgen_difflist(buf: bytes, p: int, N: int):
    """A difflist at buf[p]: varint length L; per group of 64 entries the
    first sample ID (ceil(log256 N) bytes each); L/64 - 1 skip bytes; the L
    2-bit values (low bits first); then, group by group, the varint gaps to
    the next sample ID.  Returns (ids, values, end)."""
    L, p = _varint(buf, p)
    if L == 0:
        return [], [], p
    idb = 1 if N <= 1 << 8 else (2 if N <= 1 << 16 else (3 if N <= 1 << 24 else 4))
    gc = (L + 63) // 64
    first = [int.from_bytes(buf[p + idb * g:p + idb * (g + 1)], "little") for g in range(gc)]
    p += idb * gc + (gc - 1)
    vals = [(buf[p + k // 4] >> (2 * (k % 4))) & 3 for k in range(L)]
    p += (L + 3) // 4
    ids = []
    for g in range(gc):
        cur = first[g]
        ids.append(cur)
        for _ in range(1, min(64, L - 64 * g)):
            gap, p = _varint(buf, p)
            cur += gap
            ids.append(cur)
    return ids, vals, p


def pgen_header(data: bytes):
    """Storage mode 0x10 header: (M, N, record starts, record lengths,
    record types); a variant block of 2^16 variants carries its 4- or 8-bit
    types and 1-4 byte record lengths (then optional allele counts and
    provisional-reference flags), blocks in order after the block offsets."""
    if data[:2] != b"\x6c\x1b" or data[2] != 0x10:
        raise ValueError("not a variable-width (mode 0x10) .pgen file")
    M = int.from_bytes(data[3:7], "little")
    N = int.from_bytes(data[7:11], "little")
    ctl = data[11]
    if ctl & 0x30:
        raise ValueError(".pgen with stored allele counts (multiallelic variants) is not supported")
    nblk = (M + 65535) // 65536
    offs = [int.from_bytes(data[12 + 8 * b:20 + 8 * b], "little") for b in range(nblk)]
    lb = (ctl & 3) + 1
    wide = (ctl >> 2) & 1
    pos = 12 + 8 * nblk
    types, lens = [], []
    for b in range(nblk):
        cnt = min(65536, M - 65536 * b)
        if wide:
            t = np.frombuffer(data, np.uint8, cnt, pos)
            pos += cnt
        else:
            v = np.frombuffer(data, np.uint8, (cnt + 1) // 2, pos)
            pos += (cnt + 1) // 2
            t = np.stack([v & 15, v >> 4], 1).reshape(-1)[:cnt]
        L = np.frombuffer(data, np.uint8, cnt * lb, pos).reshape(cnt, lb).astype(np.int64)
        pos += cnt * lb
        if (ctl >> 6) == 3:
            pos += (cnt + 7) // 8
        types.append(t)
        lens.append((L << (8 * np.arange(lb))).sum(axis=1))
    types, lens = np.concatenate(types), np.concatenate(lens)
    starts = np.empty(M, np.int64)
    for b in range(nblk):
        s, e = 65536 * b, min(M, 65536 * (b + 1))
        starts[s:e] = offs[b] + np.concatenate([[0], np.cumsum(lens[s:e])[:-1]])
    return M, N, starts, lens, types


def pgen_to_unpacked(data: bytes) -> tuple[np.ndarray, int, int]:
    """Hardcalls (ALT allele counts, 3 = missing) of every variant:
    record type 0 = 2-bit genotype array; 1 = two values chosen by a bit
    array (first byte: low value = byte >> 2, high = low + (byte & 3)) plus a
    difflist; 4 = all 0 plus a difflist; 2 = the previous non-LD variant's
    genotypes patched by a difflist; 3 = the same, then 0 <-> 2 inverted."""
    M, N, starts, lens, types = pgen_header(data)
    out = np.zeros((M, N), np.uint8)
    inv = np.array([2, 1, 0, 3], np.uint8)
    base = None
    for m in range(M):
        t = int(types[m]) & 7
        p = int(starts[m])
        if t == 0:
            v = np.frombuffer(data, np.uint8, (N + 3) // 4, p)
            row = np.stack([(v >> (2 * k)) & 3 for k in range(4)], 1).reshape(-1)[:N]
        elif t == 1:
            lo = data[p] >> 2
            hi = lo + (data[p] & 3)
            bits = np.unpackbits(np.frombuffer(data, np.uint8, (N + 7) // 8, p + 1), bitorder="little")[:N]
            row = np.where(bits == 1, hi, lo).astype(np.uint8)
            ids, vals, _ = pgen_difflist(data, p + 1 + (N + 7) // 8, N)
            row[ids] = vals
        elif t == 4:
            row = np.zeros(N, np.uint8)
            ids, vals, _ = pgen_difflist(data, p, N)
            row[ids] = vals
        elif t in (2, 3):
            row = base.copy()
            ids, vals, _ = pgen_difflist(data, p, N)
            row[ids] = vals
            if t == 3:
                row = inv[row]
        else:
            raise ValueError(f"variant {m}: .pgen record type {t} is not supported")
        if t not in (2, 3):
            base = row
        out[m] = row
    return out, N, M


PINS = {
    "bed_to_packed / bed_to_unpacked / unpack / pack": "reader fixtures demo_data.G(.packed).expected, exact",
    "flip_packed / mean_packed": "closed form on random tiles (0<->2, padding zeroed) and demo fixture",
    "unpack_center / freq_gpu / freq_cpu / qbat": "numpy closed form on random data",
    "synth_packed": "binomial moments (mean 2h, missing rate) on a large block",
    "em_stats_gpu / em_final_gpu / loglik_gpu": "pure-Python brute force on tiny inputs; EM ascent; "
                                                  "fixture trajectory outputs/gpu/adamem",
    "adamem_gpu": "outputs/gpu/adamem/demo_run.adamem.* from outputs/gpu/adamem/demo_run.als.*",
    "adamem_cpu / em_step_cpu / loglik_cpu": "outputs/cpu/adamem/demo_run.adamem.* from "
                                             "outputs/cpu/brqn/demo_run.als.*",
    "rsvd / eig_svd": "dense numpy SVD on small matrices; outputs/*/svd.* fixtures",
    "als / bvls_cpu": "outputs/cpu/brqn/demo_run.als.* from the svd fixtures; KKT; scipy lsq_linear",
    "bvls_gpu": "KKT + scipy lsq_linear",
    "qp_sweep / sqp_q / sqp_p / project_simplex": "KKT, brute-force active sets, sort-threshold projection",
    "grad_hess_q / grad_hess_p": "finite differences of loglik_gpu",
    "brqn / zal_extrapolate": "outputs/cpu/brqn/demo_run.brqn.* from the als fixtures",
    "projection_gpu": "first step equals the Q half of adamem_gpu's priming step; from a random start it "
                      "reaches the maximiser of the (concave in Q) log-likelihood at the fixture P, i.e. the "
                      "fixture Q of outputs/gpu/adamem, and agrees with projection_brqn_gpu",
    "init_p_supervised": "equals init_p_supervised_unpacked (the reference's other restatement) bit for bit "
                         "on the reader fixture; brute-force loop on a slice",
    "snap_q / init_q_supervised / labels_to_y": "closed form",
    "adamem_gpu(labels) / brqn(labels)": "no labels -> the unsupervised path bit for bit; all labelled -> P "
                                         "satisfies the KKT conditions of max_P logL(P, Q fixed)",
    "vcf_to_unpacked / vcf_gt_code": "the reference's reader fixture demo_data.G.expected from demo_data.vcf(.gz); "
                                      "hand-derived codes for haploid, multi-allelic, '.', '-', short lines",
    "bcf_to_unpacked": "demo_data.G.expected from demo_data.bcf(.gz); hand-built records (int8/int16 GT, "
                       "missing, vector end, haploid)",
    "pgen_to_unpacked / pgen_difflist / pgen_header": "demo_data.G.expected from demo_data.pgen (record types "
                                                      "0, 1, 2, 3, 4 all present); structural checks of the header",
    "splitmix64 / cv_hash_folds": "splitmix64's published first outputs for seeds 0, C1, 2 C1; folds partition "
                                  "the observed entries",
    "cv_deviance_sum": "sum of deviances = -logL(held-out entries) - 2 ln 2 x #heterozygous (logL pinned above); "
                       "zero for a perfect fit",
    "cv_polish / cross_validation": "brqn (pinned by the CPU fixture) for exactly 5 iterations; composition",
    "projection_brqn_gpu": "same concave-maximiser agreement; SQP steps pinned through sqp_q / grad_hess_q",
}
