"""The batched and learned-start workloads at their full sizes.

C4 (batched): N=50,000 x M=200,000, K=2..12, Balding-Nichols F_st=0.1 P,
Dirichlet(0.3) Q, 0.5% missing (2.5 GB packed):
  * the graph-captured multi-K sweep, capped at 20 iterations, against
    optimize_parameters_gpu run K by K from the same starts (same checks,
    logL to 1e-9);
  * one K against the oracle's Adam-EM (em_adam_gpu.py:281-356) on a
    slice of the matrix the oracle finishes in seconds;
  * the 50,000-problem SQP batch (Q-side Hessians at K = 12 made SPD with
    condition number <= 1e4, simplex constraint, uneven active sets)
    against the oracle's sweep-operator QP (sqp_kernel.cu:6-123) on
    3,000 of the problems.

C5 evaluation (learned start): N=200,000 x M=600,000, K=12, Dirichlet(0.1)
Q, 1% missing (30 GB packed): the warm start's gate against the SVD + ALS
start by FP64 logL, and each precision form's EM statistics against the
FP64 pass at the start (the stated bounds: FP32 2e-4, mixed 1e-3 of the
largest entry).
"""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import multik, solvers
from adamixture_amd import em_adam_gpu as E
from adamixture_amd.em_adam_gpu import EMAdamOptimizer, em_pass, optimize_parameters_gpu
from adamixture_amd.genotypes import allele_frequencies, as_packed
from adamixture_amd.synth import WORKLOADS, make_genotypes

pytestmark = pytest.mark.gpu

HP = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, lr_decay=0.5, min_lr=1e-4, patience=3, tol_adam=0.1)


@pytest.fixture(scope="module")
def c4(cuda):
    w = WORKLOADS["multik"]
    print(f"case: C4 genotypes N={w.N} M={w.M}", flush=True)
    G, _, _ = make_genotypes(w, cuda)
    yield G
    del G
    torch.cuda.empty_cache()


def _starts(G, Ks, dev, seed=21):
    rng = np.random.default_rng(seed)
    out = {}
    for K in Ks:
        P0 = np.clip(rng.random((G.M, K)), 1e-5, 1 - 1e-5)
        Q0 = rng.dirichlet(np.ones(K), size=G.N)
        out[K] = (torch.from_numpy(P0).to(dev), torch.from_numpy(Q0).to(dev))
    return out


def test_c4_sweep_at_size_matches_k_by_k(cuda, c4):
    G = c4
    Ks = list(range(2, 13))
    starts = _starts(G, Ks, cuda)
    print("case: C4 sweep K=2..12, 20 iterations, f32", flush=True)
    res = multik.sweep(G, starts, "f32", graph=True, max_iter=20, **HP)
    saved = (E._WS.min_cap, E._WS.frac)
    try:
        # one-band workspace for the K-by-K runs too (the sweep's layout)
        E._WS.min_cap = 8 << 30
        for K, (P0, Q0) in starts.items():
            hist = []
            P, Q = optimize_parameters_gpu(G, P0.clone(), Q0.clone(), HP["lr"], HP["beta1"], HP["beta2"],
                                           HP["reg_adam"], 20, 5, K, G.M, G.N, HP["lr_decay"], HP["min_lr"],
                                           HP["patience"], HP["tol_adam"], cuda, 0, 0, precision="f32",
                                           on_check=lambda it, L: hist.append((it, L)))
            got = res[K]["checks"]
            assert [it for it, _ in got] == [it for it, _ in hist], K
            np.testing.assert_allclose([L for _, L in got], [L for _, L in hist], rtol=1e-9)
            np.testing.assert_allclose(res[K]["P"].cpu().numpy(), P.cpu().numpy(), atol=1e-7)
            print(f"K={K}: logL {hist[-1][1]:.1f} ok", flush=True)
    finally:
        E._WS.min_cap, E._WS.frac = saved
        E._WS.buf = None


def test_c4_one_k_matches_the_oracle_on_a_slice(cuda, c4):
    """K = 3, 20 Adam-EM iterations on the first 8,000 SNPs x 3,000 samples
    of the C4 matrix, FP64 form, against the oracle."""
    M, N, K = 8000, 3000, 3
    Gp = np.ascontiguousarray(c4.data[: M // 4, :N].cpu().numpy())
    G = as_packed(torch.from_numpy(Gp), M, cuda)
    rng = np.random.default_rng(31)
    P0 = np.clip(rng.random((M, K)), 1e-5, 1 - 1e-5)
    Q0 = rng.dirichlet(np.ones(K), size=N)
    hist = []
    P, Q = optimize_parameters_gpu(G, torch.from_numpy(P0).to(cuda), torch.from_numpy(Q0).to(cuda), HP["lr"],
                                   HP["beta1"], HP["beta2"], HP["reg_adam"], 20, 5, K, M, N, HP["lr_decay"],
                                   HP["min_lr"], HP["patience"], HP["tol_adam"], cuda, 0, 0, precision="f64",
                                   on_check=lambda it, L: hist.append((it, L)))
    ohist = []
    Po, Qo = O.adamem_gpu(Gp, P0, Q0, M, max_iter=20, history=ohist)
    assert [it for it, _ in hist] == [it for it, _ in ohist]
    np.testing.assert_allclose([L for _, L in hist], [L for _, L in ohist], rtol=1e-11)
    np.testing.assert_allclose(P.cpu().numpy(), Po, atol=1e-8)
    np.testing.assert_allclose(Q.cpu().numpy(), Qo, atol=1e-8)


def test_c4_qp_batch_of_50000(cuda, c4):
    G = c4
    K = 12
    rng = np.random.default_rng(41)
    P = torch.from_numpy(np.clip(rng.beta(0.5, 1.5, size=(G.M, K)), 1e-5, 1 - 1e-5)).to(cuda)
    Q = torch.from_numpy(rng.dirichlet(np.full(K, 0.3), size=G.N)).to(cuda)
    print(f"case: C4 QP batch {G.N} x K={K}", flush=True)
    H, z, cond = multik.qp_workload(G, P, Q, seed=2)
    assert H.shape == (50000, K, K) and float(cond.max()) <= 1e4 * (1 + 1e-9)
    vkk = solvers.v_kk(K, cuda)
    X, ms = multik.qp_batch(H, z, Q, vkk, graph=True, replays=3)
    Xn = X.cpu().numpy()
    assert np.all(Xn >= 1e-5 - 1e-12) and np.allclose(Xn.sum(1), 1.0, atol=1e-12)
    # uneven active sets: the number of coordinates left at the lower bound varies
    at_bound = (Xn <= 1e-5 + 1e-9).sum(1)
    sizes = np.bincount(at_bound, minlength=K + 1)
    print(f"QP batch: {ms:.3f} ms ({50000 / ms * 1e3:.3e} problems/s); problems by #coordinates at the bound: "
          f"{sizes.tolist()}", flush=True)
    assert (sizes > 0).sum() >= 3
    sel = slice(0, 3000)
    want = O.sqp_q(H[sel].cpu().numpy(), z[sel].cpu().numpy(), Q[sel].cpu().numpy(), O.v_kk(K))
    np.testing.assert_allclose(Xn[sel], want, atol=1e-9)


@pytest.fixture(scope="module")
def c5(cuda):
    w = WORKLOADS["learned_eval"]
    print(f"case: C5 evaluation genotypes N={w.N} M={w.M}", flush=True)
    G, _, _ = make_genotypes(w, cuda)
    yield G, w
    del G
    torch.cuda.empty_cache()


def test_c5_eval_learned_gate_and_forms(cuda, c5):
    from adamixture_amd.learned_init import WEIGHTS_PATH, WarmStartMLP, gate_band, gate_loglik, learned_warm_start
    from adamixture_amd.learned_train import plain_start
    G, w = c5
    K = w.K
    f = allele_frequencies(G)
    print("case: C5 randomized SVD (rank 2K) + ALS", flush=True)
    U, S, V, P_a, Q_a = plain_start(G, f, K, w.seed, cuda)
    ll_a = em_pass(G, P_a, Q_a, loglik=True, stats=False).item()
    model = WarmStartMLP.load(cuda)
    assert WEIGHTS_PATH.exists()
    P_l, Q_l, ll_gate, used = learned_warm_start(G, U, S, V, f, K, model, w.seed)
    ll_l = em_pass(G, P_l, Q_l, loglik=True, stats=False).item()
    print(f"C5: ALS start logL {ll_a:.1f}, learned start logL {ll_l:.1f}, gate used the network: {used} "
          f"(gate band {gate_band(G).M} of {G.M} SNPs)", flush=True)
    # the gate returns the better start by FP64 logL over its band, and a valid one
    ll_a_gate = gate_loglik(G, P_a, Q_a)
    assert ll_gate > ll_a_gate if used else ll_gate == pytest.approx(ll_a_gate, rel=1e-12)
    Ql = Q_l.cpu().numpy()
    # clamp to [1e-5, 1 - 1e-5] then rows divided by their sums (als_gpu.py:23-36):
    # an entry ends at least 1e-5 / (1 + K 1e-5)
    assert Ql.min() >= 1e-5 / (1 + K * 1e-5) * (1 - 1e-12) and np.allclose(Ql.sum(1), 1.0, atol=1e-12)
    assert gate_loglik(G, P_l, Q_l) == pytest.approx(ll_gate, rel=1e-13)
    # every form's EM statistics against the FP64 pass at the ALS start
    ref = EMAdamOptimizer(P_a.shape, Q_a.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision="f64")
    ref.accumulate(G, P_a, Q_a)
    want = [x.clone() for x in (ref.A_accum, ref.B_accum, ref.T_accum)]
    del ref
    for prec, bound in (("f32", 2e-4), ("bf16", 1e-3)):
        opt = EMAdamOptimizer(P_a.shape, Q_a.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision=prec)
        opt.accumulate(G, P_a, Q_a)
        errs = []
        for got, ref_t in zip((opt.A_accum, opt.B_accum, opt.T_accum), want):
            err = float((got.double() - ref_t).abs().max() / ref_t.abs().max())
            errs.append(err)
            assert err <= bound, (prec, err)
        print(f"C5 {prec}: max error / max entry A {errs[0]:.2e} B {errs[1]:.2e} T {errs[2]:.2e}", flush=True)
        del opt
