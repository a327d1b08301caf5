"""Batched forms: the multi-K Adam-EM sweep replayed from captured graphs,
the graph-replayed SQP batch, and the tile-shape autotuner.

Graph replays are checked against the same kernels enqueued eagerly (every
replay of the sweep is followed by host reads and eager copies, as the
schedule runs), against the single-K optimizer, and the SQP batch against
the oracle."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import autotune, multik, solvers
from adamixture_amd.em_adam_gpu import optimize_parameters_gpu
from adamixture_amd.synth import Workload, make_genotypes

pytestmark = pytest.mark.gpu

HP = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, lr_decay=0.5, min_lr=1e-4, patience=3, tol_adam=0.1)


def _data(cuda, N=1200, M=6000, K=4, seed=11):
    w = Workload("t", N, M, K, "bn", 0.3, 0.01, fst=0.1, seed=seed)
    G, _, _ = make_genotypes(w, cuda)
    return G


def _starts(G, Ks, cuda, seed=5):
    rng = np.random.default_rng(seed)
    out = {}
    for K in Ks:
        P0 = np.clip(rng.random((G.M, K)), 1e-5, 1 - 1e-5)
        Q0 = rng.dirichlet(np.ones(K), size=G.N)
        out[K] = (torch.from_numpy(P0).to(cuda), torch.from_numpy(Q0).to(cuda))
    return out


@pytest.mark.parametrize("precision", ["f32", "bf16", "f64"])
def test_sweep_graph_matches_eager(cuda, precision):
    G = _data(cuda)
    starts = _starts(G, [2, 3, 5], cuda)
    print(f"sweep graph vs eager {precision}", flush=True)
    a = multik.sweep(G, starts, precision, graph=True, max_iter=40, **HP)
    b = multik.sweep(G, starts, precision, graph=False, max_iter=40, **HP)
    for K in starts:
        assert [it for it, _ in a[K]["checks"]] == [it for it, _ in b[K]["checks"]]
        La = np.array([L for _, L in a[K]["checks"]])
        Lb = np.array([L for _, L in b[K]["checks"]])
        print(f"K={K}: {len(La)} checks, max rel dL {np.max(np.abs(La - Lb) / np.abs(Lb)):.2e}", flush=True)
        # the passes sum their partial slices in a fixed order: replaying the
        # graph reproduces the eager run bit for bit
        np.testing.assert_array_equal(La, Lb)
        np.testing.assert_array_equal(a[K]["P"].cpu().numpy(), b[K]["P"].cpu().numpy())
        np.testing.assert_array_equal(a[K]["Q"].cpu().numpy(), b[K]["Q"].cpu().numpy())


def test_sweep_matches_single_k_optimizer(cuda):
    """Each K of the sweep follows optimize_parameters_gpu from the same
    start (same check schedule and logL up to summation-order rounding)."""
    G = _data(cuda, seed=12)
    starts = _starts(G, [2, 4], cuda, seed=6)
    res = multik.sweep(G, starts, "f64", graph=True, max_iter=60, **HP)
    for K, (P0, Q0) in starts.items():
        hist = []
        P, Q = optimize_parameters_gpu(G, P0.clone(), Q0.clone(), HP["lr"], HP["beta1"], HP["beta2"],
                                       HP["reg_adam"], 60, 5, K, G.M, G.N, HP["lr_decay"], HP["min_lr"],
                                       HP["patience"], HP["tol_adam"], cuda, 0, 0, precision="f64",
                                       on_check=lambda it, L: hist.append((it, L)))
        got = res[K]["checks"]
        assert [it for it, _ in got] == [it for it, _ in hist]
        np.testing.assert_allclose([L for _, L in got], [L for _, L in hist], rtol=1e-9)
        np.testing.assert_allclose(res[K]["P"].cpu().numpy(), P.cpu().numpy(), atol=1e-6)


def test_sweep_runs_until_each_k_stops(cuda):
    """Without an iteration cap every K ends by the lr rule; finished K drop
    out of the re-captured graphs while the others continue."""
    G = _data(cuda, N=800, M=4000, seed=13)
    starts = _starts(G, [2, 3], cuda, seed=7)
    res = multik.sweep(G, starts, "f32", graph=True, max_iter=3000, **HP)
    for K in starts:
        assert res[K]["iterations"] < 3000
        L = [L for _, L in res[K]["checks"]]
        # the best is replaced only by a check better by more than tol_adam
        assert max(L) - HP["tol_adam"] <= res[K]["logL"] <= max(L)


def test_qp_batch_graph_matches_eager_and_oracle(cuda):
    G = _data(cuda, N=3000, M=4000, K=6, seed=14)
    K = 6
    rng = np.random.default_rng(3)
    P = torch.from_numpy(np.clip(rng.random((G.M, K)), 1e-5, 1 - 1e-5)).to(cuda)
    Q = torch.from_numpy(rng.dirichlet(np.full(K, 0.3), size=G.N)).to(cuda)
    H, z, cond = multik.qp_workload(G, P, Q, seed=1)
    assert float(cond.max()) <= 1e4 * (1 + 1e-9)
    vkk = solvers.v_kk(K, cuda)
    Xg, ms_g = multik.qp_batch(H, z, Q, vkk, graph=True, replays=3)
    Xe, ms_e = multik.qp_batch(H, z, Q, vkk, graph=False, replays=1)
    print(f"qp batch {G.N} x K={K}: graph {ms_g:.3f} ms, eager {ms_e:.3f} ms", flush=True)
    assert torch.equal(Xg, Xe)                      # one wavefront per problem: deterministic
    sel = slice(0, 500)
    want = O.sqp_q(H[sel].cpu().numpy(), z[sel].cpu().numpy(), Q[sel].cpu().numpy(), O.v_kk(K))
    np.testing.assert_allclose(Xg[sel].cpu().numpy(), want, atol=1e-9)


@pytest.mark.parametrize("precision", ["f32", "bf16", "f64"])
def test_autotune_picks_a_candidate_and_keeps_results(cuda, precision):
    from adamixture_amd.em_adam_gpu import EMAdamOptimizer
    G = _data(cuda, N=2000, M=9000, K=5, seed=15)
    K = 5
    rng = np.random.default_rng(4)
    P = torch.from_numpy(np.clip(rng.random((G.M, K)), 1e-5, 1 - 1e-5)).to(cuda)
    Q = torch.from_numpy(rng.dirichlet(np.ones(K), size=G.N)).to(cuda)
    default = autotune.get_tiles(precision)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, cuda, precision)
    PE0, QE0 = (x.clone() for x in opt.run_em_step(G, P, Q, G.M))
    try:
        res = autotune.tune_em(G, P, Q, precision, repeats=1)
        print(f"autotune {precision}: {res}", flush=True)
        assert res["best"] in [tuple(c) for c in autotune.CANDIDATES[precision]]
        assert autotune.get_tiles(precision) == res["best"]
        cands = autotune.CANDIDATES[precision]
        for tiles in cands[:2] + [c for c in cands if c[1] == 8][:2]:
            print(f"case: {precision} tiles {autotune.tiles_label(tiles)}", flush=True)
            autotune.set_tiles(precision, *tiles)
            PE, QE = opt.run_em_step(G, P, Q, G.M)
            tol = 1e-12 if precision == "f64" else 1e-4
            np.testing.assert_allclose(PE.cpu().numpy(), PE0.cpu().numpy(), rtol=tol, atol=tol)
            np.testing.assert_allclose(QE.cpu().numpy(), QE0.cpu().numpy(), rtol=tol, atol=tol)
    finally:
        autotune.set_tiles(precision, *default)
    with pytest.raises(RuntimeError):
        autotune.set_tiles(precision, 5, 4, 2048)
    with pytest.raises(RuntimeError):
        autotune.set_tiles(precision, 4, 6, 2048)


def test_loglik_independent_of_tile_shape(cuda):
    """The logL-only FP64 pass over every FP64 tile candidate (grids of
    64..65536 workgroups change only the summation order of the per-lane
    log-products): equal to 1e-13 relative, and the oracle's logL
    (em_adam_gpu.py:238-246) to 1e-12."""
    from adamixture_amd.em_adam_gpu import em_pass
    G = _data(cuda, N=1500, M=7001, K=6, seed=16)
    rng = np.random.default_rng(6)
    P = torch.from_numpy(np.clip(rng.random((G.M, 6)), 1e-5, 1 - 1e-5)).to(cuda)
    Q = torch.from_numpy(rng.dirichlet(np.ones(6), size=G.N)).to(cuda)
    want = O.loglik_gpu(G.data[:, :G.N].cpu().numpy(), P.cpu().numpy(), Q.cpu().numpy(), G.M)
    default = autotune.get_tiles("f64")
    vals = []
    try:
        for tiles in autotune.CANDIDATES["f64"]:
            print(f"case: logL tiles {autotune.tiles_label(tiles)}", flush=True)
            autotune.set_tiles("f64", *tiles)
            vals.append(em_pass(G, P, Q, loglik=True, stats=False).item())
    finally:
        autotune.set_tiles("f64", *default)
    assert max(vals) - min(vals) <= 1e-13 * abs(want), vals
    assert abs(vals[0] - want) <= 1e-12 * abs(want), (vals[0], want)
