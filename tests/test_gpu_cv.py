"""Cross-validation on the GPU against the oracle (cv.py:171-267,
br_qn.pyx:92-246, br_qn.py:50-70).  Fold membership, masking and the
deviance are index/arithmetic work compared entry for entry; the polish is
the SQP + ZAL path (compared as the BR-QN tests compare it); the CV index
end to end on the demo data on both fold paths, and on a synthetic matrix
whose hashed folds the oracle evaluates too."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import _lib
from adamixture_amd import cv as CV
from adamixture_amd.genotypes import PackedGenotypes, as_packed

pytestmark = pytest.mark.gpu


def _unpack(Gp: PackedGenotypes) -> np.ndarray:
    return O.unpack(Gp.to_host_packed(), Gp.M)


def test_project_simplex_matches_oracle(cuda):
    rng = np.random.default_rng(11)
    for K in (2, 7, 16, 64):
        X = rng.normal(0.2, 0.5, size=(300, K))
        Xd = torch.from_numpy(X.copy()).to(cuda)
        _lib.call("adx_project_simplex", Xd.data_ptr(), 300, K, 1e-5, _lib.stream_handle(cuda))
        want = np.stack([O.project_simplex(X[r].copy(), 1e-5) for r in range(300)])
        np.testing.assert_allclose(Xd.cpu().numpy(), want, rtol=0, atol=1e-15)


@pytest.mark.parametrize("fold", [0, 3])
def test_masks_match_oracle(cuda, fixtures, fold):
    G = fixtures["reader_G"]
    M, N = G.shape
    Gp = as_packed(fixtures["reader_G_packed"], M, cuda)
    st = _lib.stream_handle(cuda)
    print(f"case: CV masks demo fold {fold}", flush=True)
    # hashed folds
    Gm = PackedGenotypes(torch.empty_like(Gp.data), M, N)
    _lib.call("adx_cv_mask_hash", Gp.data.data_ptr(), Gm.data.data_ptr(), Gp.ldg, M, N, 42, 5, fold, st)
    obs = np.flatnonzero(G != 3)
    want = G.copy()
    want.ravel()[obs[O.cv_hash_folds(obs, 42, 5) == fold]] = 3
    np.testing.assert_array_equal(_unpack(Gm), want)
    assert np.all(Gm.data[:, N:].cpu().numpy() == 0xFF)
    # shuffled-list folds (int32 and int64 indices)
    sh = O.cv_shuffled_entries(G, 42)
    want = G.copy()
    want.ravel()[sh[fold::5]] = 3
    n = len(range(fold, sh.size, 5))
    for arr, idx64 in ((sh, 0), (sh.astype(np.int64), 1)):
        Gm.data.copy_(Gp.data)
        flat_d = torch.from_numpy(arr).to(cuda)
        _lib.call("adx_cv_mask_list", Gm.data.data_ptr(), Gm.ldg, M, N, flat_d.data_ptr(), idx64, fold, 5, n, st)
        np.testing.assert_array_equal(_unpack(Gm), want)


def test_deviance_matches_oracle(cuda, fixtures):
    G = fixtures["reader_G"]
    M, N = G.shape
    P, Q = fixtures["cpu_brqn_brqn_P"], fixtures["cpu_brqn_brqn_Q"]
    K = P.shape[1]
    Gp = as_packed(fixtures["reader_G_packed"], M, cuda)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    work = torch.empty(_lib.ADX_CV_BLOCKS, dtype=torch.float64, device=cuda)
    out = torch.empty(1, dtype=torch.float64, device=cuda)
    st = _lib.stream_handle(cuda)
    obs = np.flatnonzero(G != 3)
    for fold in range(5):
        held = obs[O.cv_hash_folds(obs, 9, 5) == fold]
        want = O.cv_deviance_sum(G.ravel()[held], held, P, Q, N)
        _lib.call("adx_cv_deviance_hash", Gp.data.data_ptr(), Gp.ldg, M, N, 9, 5, fold, Pd.data_ptr(),
                  Qd.data_ptr(), K, work.data_ptr(), out.data_ptr(), st)
        assert abs(out.item() - want) <= 1e-12 * abs(want)
        sh = O.cv_shuffled_entries(G, 9)
        held = sh[fold::5]
        want = O.cv_deviance_sum(G.ravel()[held], held, P, Q, N)
        flat_d = torch.from_numpy(sh).to(cuda)
        _lib.call("adx_cv_deviance_list", Gp.data.data_ptr(), Gp.ldg, M, N, flat_d.data_ptr(), 0, fold, 5,
                  held.size, Pd.data_ptr(), Qd.data_ptr(), K, work.data_ptr(), out.data_ptr(), st)
        assert abs(out.item() - want) <= 1e-12 * abs(want)
    # an empty fold sums to zero
    _lib.call("adx_cv_deviance_list", Gp.data.data_ptr(), Gp.ldg, M, N, flat_d.data_ptr(), 0, 0, 5, 0,
              Pd.data_ptr(), Qd.data_ptr(), K, work.data_ptr(), out.data_ptr(), st)
    assert out.item() == 0.0


def test_polish_matches_oracle(cuda, fixtures):
    G = fixtures["reader_G"]
    M, N = G.shape
    P, Q = fixtures["cpu_brqn_brqn_P"], fixtures["cpu_brqn_brqn_Q"]
    obs = np.flatnonzero(G != 3)
    Gmask = G.copy()
    Gmask.ravel()[obs[O.cv_hash_folds(obs, 42, 5) == 1]] = 3
    print("case: CV polish demo fold 1", flush=True)
    Pc, Qc = CV.polish_gpu(as_packed(O.pack(Gmask), M, cuda), torch.from_numpy(P).to(cuda),
                           torch.from_numpy(Q).to(cuda))
    Po, Qo = O.cv_polish(O.pack(Gmask), P, Q, M)
    print(f"polish dP {np.abs(Pc.cpu().numpy() - Po).max():.2e} dQ {np.abs(Qc.cpu().numpy() - Qo).max():.2e}")
    np.testing.assert_allclose(Pc.cpu().numpy(), Po, atol=1e-8)
    np.testing.assert_allclose(Qc.cpu().numpy(), Qo, atol=1e-8)


@pytest.mark.parametrize("streaming", [False, True])
def test_cv_index_demo_matches_oracle(cuda, fixtures, streaming):
    G = fixtures["reader_G"]
    M, N = G.shape
    P, Q = fixtures["cpu_brqn_brqn_P"], fixtures["cpu_brqn_brqn_Q"]
    print(f"case: CV index demo streaming={streaming}", flush=True)
    args = SimpleNamespace(cv=5, seed=42)
    got = CV.run_cross_validation(args, G, N, M, P.shape[1], P, Q, device=cuda, streaming=streaming)
    want = O.cross_validation(G, P, Q, 42, 5, streaming=streaming)
    print(f"CV index gpu {got:.12f} oracle {want:.12f}")
    assert abs(got - want) <= 1e-10 * abs(want)


def test_cv_index_synthetic_hashed(cuda):
    """N=1500, M=12001 synthetic genotypes, K=4, 3 folds, from the
    generating (P, Q) perturbed; packed device input."""
    from adamixture_amd.synth import Workload, make_genotypes
    w = Workload("cv", 1500, 12001, 4, "bn", 0.3, 0.02, fst=0.1, seed=5)
    G, P0, Q0 = make_genotypes(w, cuda)
    rng = np.random.default_rng(2)
    P = np.clip(P0 + rng.normal(0, 0.01, P0.shape), 1e-5, 1 - 1e-5)
    Q = Q0 + 0.02
    Q /= Q.sum(axis=1, keepdims=True)
    print("case: CV index synthetic N=1500 M=12001 K=4 hashed", flush=True)
    got = CV.run_cross_validation(SimpleNamespace(cv=3, seed=123), G, w.N, w.M, w.K, P, Q, device=cuda,
                                  streaming=True)
    want = O.cross_validation(O.unpack(G.to_host_packed(), w.M), P, Q, 123, 3, streaming=True)
    print(f"CV index gpu {got:.12f} oracle {want:.12f}")
    assert abs(got - want) <= 1e-9 * abs(want)


@pytest.mark.parametrize("n_folds", [1, 2, 3, 7, 10, 64, 1000])
def test_hash_masks_for_many_fold_counts(cuda, fixtures, n_folds):
    """The hashed fold test (a 64-bit modulo by the run-time fold count,
    done with a precomputed multiply-high magic) against the oracle's
    Python modulo, for fold counts of every kind (1, powers of two, odd,
    larger than the data's byte values)."""
    G = fixtures["reader_G"]
    M, N = G.shape
    Gp = as_packed(fixtures["reader_G_packed"], M, cuda)
    Gm = PackedGenotypes(torch.empty_like(Gp.data), M, N)
    obs = np.flatnonzero(G != 3)
    folds = O.cv_hash_folds(obs, 12345, n_folds)
    for fold in sorted({0, n_folds // 2, n_folds - 1}):
        print(f"case: hash mask n_folds={n_folds} fold={fold}", flush=True)
        _lib.call("adx_cv_mask_hash", Gp.data.data_ptr(), Gm.data.data_ptr(), Gp.ldg, M, N, 12345, n_folds, fold,
                  _lib.stream_handle(cuda))
        want = G.copy()
        want.ravel()[obs[folds == fold]] = 3
        np.testing.assert_array_equal(_unpack(Gm), want)


@pytest.mark.parametrize("K", [3, 16, 20])
def test_hash_deviance_register_and_pointer_paths(cuda, K):
    """Hashed deviance with the Q row held in registers (K <= 16) and read
    from memory (K > 16), on a synthetic matrix whose N is not a multiple of
    the 256-sample thread block, against the oracle."""
    from adamixture_amd.synth import Workload, make_genotypes
    w = Workload("cvdev", 333, 2050, K, "bn", 0.3, 0.05, fst=0.1, seed=K)
    G, P0, Q0 = make_genotypes(w, cuda)
    P = np.clip(P0, 1e-5, 1 - 1e-5)
    Q = Q0 / Q0.sum(axis=1, keepdims=True)
    Gu = O.unpack(G.to_host_packed(), w.M)
    obs = np.flatnonzero(Gu != 3)
    work = torch.empty(_lib.ADX_CV_BLOCKS, dtype=torch.float64, device=cuda)
    out = torch.empty(1, dtype=torch.float64, device=cuda)
    Pd, Qd = torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)
    for fold in (0, 2):
        print(f"case: hash deviance K={K} fold={fold}", flush=True)
        held = obs[O.cv_hash_folds(obs, 77, 3) == fold]
        want = O.cv_deviance_sum(Gu.ravel()[held], held, P, Q, w.N)
        _lib.call("adx_cv_deviance_hash", G.data.data_ptr(), G.ldg, w.M, w.N, 77, 3, fold, Pd.data_ptr(),
                  Qd.data_ptr(), K, work.data_ptr(), out.data_ptr(), _lib.stream_handle(cuda))
        assert abs(out.item() - want) <= 1e-12 * abs(want)
