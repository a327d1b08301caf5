"""Adam-EM trajectories on the GPU against the oracle and the reference's own
fixtures (adamixture/demo/outputs/gpu/adamem, CUDA FP64 run).

Over hundreds of Adam-EM iterations the trajectory amplifies rounding.  The
chaos floor is measured on the oracle itself: it is run again with every EM
target multiplied by (1 + 1e-16 z), z ~ N(0, 1) -- the size of the rounding
differences another FP64 summation order makes -- for two noise seeds.  On
the reference demo data this moves P by ~2e-3 and logL by ~4e-8 relative at
the end (perturbing only the initial P by 1e-15 moves it ~500x less and
underestimates the floor).  The response is linear in the noise amplitude,
so a GPU run within 10x of that floor injects at most ~1e-15 relative
rounding difference per step; it must also take the same check iterations.
Single steps are pinned to 1e-11 in test_gpu_kernels.py."""
import numpy as np
import pytest
import torch

import oracle as O
from adamixture_amd import genotypes as GT
from adamixture_amd import reader
from adamixture_amd.em_adam_gpu import loglikelihood_gpu, optimize_parameters_gpu
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu

HP = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, check=5, lr_decay=0.5, min_lr=1e-4, patience=3,
          tol_adam=0.1)


def _sensitivity(Gp, P0, Q0, M, max_iter, Po, Qo, ohist):
    """How far the oracle moves under FP64-rounding-sized noise on its EM
    targets (max over two noise seeds): P, Q at the end, and the largest
    relative logL difference over the checks both runs share."""
    kw = {} if max_iter is None else {"max_iter": max_iter}
    out = {"P": 0.0, "Q": 0.0, "ll": 0.0}
    L0 = np.array([L for _, L in ohist])
    for seed in (0, 1):
        h = []
        Pp, Qp = O.adamem_gpu(Gp, P0, Q0, M, history=h, noise=(seed, 1e-16), **kw)
        L1 = np.array([L for _, L in h])
        n = min(len(L0), len(L1))
        out["P"] = max(out["P"], np.abs(Pp - Po).max())
        out["Q"] = max(out["Q"], np.abs(Qp - Qo).max())
        out["ll"] = max(out["ll"], float(np.max(np.abs(L1[:n] - L0[:n]) / np.abs(L0[:n]))))
    return out


def _run(G, P0, Q0, M, N, K, dev, precision, max_iter=10000):
    hist = []
    P, Q = optimize_parameters_gpu(G, torch.from_numpy(P0).to(dev), torch.from_numpy(Q0).to(dev),
                                   HP["lr"], HP["beta1"], HP["beta2"], HP["reg_adam"], max_iter, HP["check"], K, M,
                                   N, HP["lr_decay"], HP["min_lr"], HP["patience"], HP["tol_adam"], dev, 8192, 256,
                                   precision=precision, on_check=lambda it, L: hist.append((it, L)))
    return P.cpu().numpy(), Q.cpu().numpy(), hist


def test_demo_adamem_f64_matches_oracle_and_reference_fixture(cuda, fixtures):
    Gt, N, M = reader.read_bed_packed(GOLDEN / "demo_data.bed")
    K = 7
    P0, Q0 = fixtures["gpu_adamem_als_P"], fixtures["gpu_adamem_als_Q"]
    P, Q, hist = _run(Gt, P0, Q0, M, N, K, cuda, "f64")
    ohist = []
    Po, Qo = O.adamem_gpu(fixtures["reader_G_packed"], P0, Q0, M, history=ohist)
    print("checks gpu/oracle", len(hist), len(ohist), "max|dP|", np.abs(P - Po).max(), "max|dQ|",
          np.abs(Q - Qo).max(), "last logL", hist[-1], ohist[-1])
    assert [it for it, _ in hist] == [it for it, _ in ohist]      # same lr schedule and stop
    floor = _sensitivity(fixtures["reader_G_packed"], P0, Q0, M, None, Po, Qo, ohist)
    print(f"oracle chaos floor: dP {floor['P']:.3e} dQ {floor['Q']:.3e} logL rel {floor['ll']:.3e}")
    np.testing.assert_allclose([L for _, L in hist], [L for _, L in ohist], rtol=max(1e-11, 10 * floor["ll"]))
    assert np.abs(P - Po).max() <= 10 * floor["P"] + 1e-12
    assert np.abs(Q - Qo).max() <= 10 * floor["Q"] + 1e-12
    # and the reference's own CUDA fixture at its tolerance (tests/config.py:24-26)
    Pe, Qe = fixtures["gpu_adamem_adamem_P"], fixtures["gpu_adamem_adamem_Q"]
    np.testing.assert_allclose(P, Pe, rtol=2e-2, atol=2e-2)
    np.testing.assert_allclose(Q, Qe, rtol=2e-2, atol=2e-2)
    G = fixtures["reader_G"]
    assert abs(O.loglik_cpu(G, P, Q) - O.loglik_cpu(G, Pe, Qe)) <= 50.0


@pytest.mark.parametrize("precision", ["f64", "f32", "bf16"])
def test_cpu_ref_workload_fixed_iterations(cuda, precision):
    """N=1,000 x M=20,000, K=3, 1% missing: 100 Adam-EM iterations from the
    same init on the GPU and on the oracle (FP64)."""
    from adamixture_amd.synth import WORKLOADS, draw_params, make_genotypes
    w = WORKLOADS["cpu_ref"]
    Ptrue, Qtrue = draw_params(w)
    G, _, _ = make_genotypes(w, cuda, params=(Ptrue, Qtrue))
    rng = np.random.default_rng(w.seed)
    P0 = np.clip(rng.random((w.M, w.K)), 1e-5, 1 - 1e-5)
    Q0 = rng.dirichlet(np.ones(w.K), size=w.N)
    P, Q, hist = _run(G, P0, Q0, w.M, w.N, w.K, cuda, precision, max_iter=100)
    ohist = []
    Po, Qo = O.adamem_gpu(G.to_host_packed(), P0, Q0, w.M, max_iter=100, history=ohist)
    dP, dQ = np.abs(P - Po).max(), np.abs(Q - Qo).max()
    rel_ll = abs(hist[-1][1] - ohist[-1][1]) / abs(ohist[-1][1])
    print(f"{precision}: max|dP| {dP:.3e} max|dQ| {dQ:.3e} logL rel {rel_ll:.3e}")
    if precision == "f64":
        assert [it for it, _ in hist] == [it for it, _ in ohist]
        floor = _sensitivity(G.to_host_packed(), P0, Q0, w.M, 100, Po, Qo, ohist)
        print(f"oracle chaos floor: dP {floor['P']:.3e} dQ {floor['Q']:.3e} logL rel {floor['ll']:.3e}")
        assert dP <= 10 * floor["P"] + 1e-12 and dQ <= 10 * floor["Q"] + 1e-12
        assert rel_ll <= max(1e-11, 10 * floor["ll"])
    else:
        assert dP < 5e-3 and dQ < 5e-3 and rel_ll < 1e-6
    Gl = G
    assert loglikelihood_gpu(Gl, torch.from_numpy(Po).to(cuda), torch.from_numpy(Qo).to(cuda), w.M, w.N) == \
        pytest.approx(O.loglik_gpu(G.to_host_packed(), Po, Qo, w.M), rel=1e-12)


def test_cpu_ref_workload_as_configured_within_1e5_relative(cuda):
    """The 1,000 x 20,000, K=3 workload as configured: randomized SVD + ALS
    init (seed 42, through the drop-in driver's initialize_k), then 100
    Adam-EM iterations with the reference defaults, FP64 form, on the GPU
    and on the oracle from that same init.  P, Q (elementwise) and logL
    (every check) within 1e-5 relative, the same check iterations."""
    from adamixture_amd.adamixture import initialize_k, setup
    from adamixture_amd.synth import WORKLOADS, draw_params, make_genotypes
    w = WORKLOADS["cpu_ref"]
    G, _, _ = make_genotypes(w, cuda, params=draw_params(w))
    dev, tpb, f, _, _, _, G = setup(G, w.N, w.M, w.K, w.seed, 5, 0.1, 8192, cuda, compute_svd=False)
    P0, Q0, _ = initialize_k(G, w.N, w.M, w.K, f, w.seed, 5, 0.1, 1000, 1e-4, 8192, dev, tpb)
    P0, Q0 = P0.cpu().numpy(), Q0.cpu().numpy()
    P, Q, hist = _run(G, P0, Q0, w.M, w.N, w.K, cuda, "f64", max_iter=100)
    ohist = []
    Po, Qo = O.adamem_gpu(G.to_host_packed(), P0, Q0, w.M, max_iter=100, history=ohist)
    rP = np.abs(P - Po) / np.abs(Po)
    rQ = np.abs(Q - Qo) / np.abs(Qo)
    rL = np.abs(np.array([L for _, L in hist]) - np.array([L for _, L in ohist])) / np.abs([L for _, L in ohist])
    print(f"C1 as configured: checks {len(hist)}, max rel dP {rP.max():.3e} (median {np.median(rP):.1e}), "
          f"max rel dQ {rQ.max():.3e}, max rel dlogL {rL.max():.3e}")
    assert [it for it, _ in hist] == [it for it, _ in ohist]
    assert rP.max() <= 1e-5 and rQ.max() <= 1e-5 and rL.max() <= 1e-5


def test_gpu_f32_workload_as_configured(cuda):
    """The 10,000 x 100,000, K=8 workload (Balding-Nichols P, Dirichlet(0.3)
    Q, 0.5% missing) as configured: bit-exact unpack against the oracle's
    generator, SVD + ALS init through the driver, then 20 Adam-EM iterations
    from that init in the FP64 form (P, Q, logL within 1e-5 relative of the
    oracle) and in the FP32 form (logL within 1e-6 relative; P and Q within
    2 lr: a parameter at its fixed point steps +-lr per Adam step, and the
    FP32 pass's rounding can flip the sign of its tiny EM displacement)."""
    from adamixture_amd import genotypes as GTm
    from adamixture_amd.adamixture import initialize_k, setup
    from adamixture_amd.synth import WORKLOADS, draw_params, make_genotypes
    w = WORKLOADS["gpu_f32"]
    Ptrue, Qtrue = draw_params(w)
    G, _, _ = make_genotypes(w, cuda, params=(Ptrue, Qtrue))
    Gp = G.to_host_packed()
    rows = slice(1000, 1250)
    assert np.array_equal(Gp[rows], O.synth_packed(Ptrue, Qtrue, w.seed, w.miss, row0=1000, rows=250))
    assert np.array_equal(GTm.unpack_chunk(G, 4000, 1000).cpu().numpy(), O.unpack(Gp, w.M, 4000, 1000))
    dev, tpb, f, _, _, _, G = setup(G, w.N, w.M, w.K, w.seed, 5, 0.1, 8192, cuda, compute_svd=False)
    P0, Q0, _ = initialize_k(G, w.N, w.M, w.K, f, w.seed, 5, 0.1, 1000, 1e-4, 8192, dev, tpb)
    P0, Q0 = P0.cpu().numpy(), Q0.cpu().numpy()
    ohist = []
    Po, Qo = O.adamem_gpu(Gp, P0, Q0, w.M, max_iter=20, history=ohist)
    Lo = np.array([L for _, L in ohist])
    for precision in ("f64", "f32"):
        P, Q, hist = _run(G, P0, Q0, w.M, w.N, w.K, cuda, precision, max_iter=20)
        L = np.array([L for _, L in hist])
        assert [it for it, _ in hist] == [it for it, _ in ohist]
        rP, rQ = np.abs(P - Po) / np.abs(Po), np.abs(Q - Qo) / np.abs(Qo)
        rL = np.abs(L - Lo) / np.abs(Lo)
        print(f"C2 {precision}: max rel dP {rP.max():.3e}, max |dP| {np.abs(P - Po).max():.3e}, max rel dQ "
              f"{rQ.max():.3e}, max rel dlogL {rL.max():.3e}")
        if precision == "f64":
            assert rP.max() <= 1e-5 and rQ.max() <= 1e-5 and rL.max() <= 1e-5
        else:
            assert rL.max() <= 1e-6 and np.abs(P - Po).max() <= 2 * HP["lr"] and np.abs(Q - Qo).max() <= 2 * HP["lr"]
