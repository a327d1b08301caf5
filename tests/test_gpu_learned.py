"""What the trained warm-start network (adamixture_amd/learned_init_weights.npz,
trained on the GPU by adamixture_amd.learned_train) gains over the SVD + ALS
start on a held-out synthetic data set (a seed outside the training range):
its Q is closer to the Q the plain method converges to, and the warm start's
log-likelihood is higher.  The number of Adam-EM iterations to a logL within
1e-6 of the converged one is printed, not asserted: from any start this
close, Adam's step size, not the start, sets it."""
import numpy as np
import pytest
import torch

from adamixture_amd.em_adam_gpu import em_pass
from adamixture_amd.genotypes import allele_frequencies
from adamixture_amd.learned_init import WEIGHTS_PATH, WarmStartMLP, features, gate_loglik, learned_warm_start
from adamixture_amd.learned_train import align_columns, converge_plain, first_reach, plain_start, run_to_stop
from adamixture_amd.synth import Workload, make_genotypes

pytestmark = pytest.mark.gpu


def test_trained_network_beats_the_als_start(cuda):
    assert WEIGHTS_PATH.exists(), "trained weights are part of the package"
    model = WarmStartMLP.load(cuda)
    K, seed = 8, 99123
    w = Workload("heldout", 5000, 50000, K, "bn", 0.3, 0.01, fst=0.1, seed=seed)
    G, _, _ = make_genotypes(w, cuda)
    f = allele_frequencies(G)
    U, S, V, P_a, Q_a = plain_start(G, f, K, seed, cuda)
    P_star, Q_star = converge_plain(G, P_a, Q_a, K, cuda, "f64")
    Q_star = Q_star[:, align_columns(Q_a, Q_star)]
    Q_w = model.infer(features(U, S, V, f, P_a, Q_a, K), K)
    mae_als = float((Q_a - Q_star).abs().sum(1).mean())
    mae_net = float((Q_w - Q_star).abs().sum(1).mean())
    P_l, Q_l, ll_gate, used = learned_warm_start(G, U, S, V, f, K, model, seed)
    ll_l = em_pass(G, P_l, Q_l, loglik=True, stats=False).item()
    ll_a = em_pass(G, P_a, Q_a, loglik=True, stats=False).item()
    L_star = em_pass(G, P_star, Q_star.contiguous(), loglik=True, stats=False).item()
    target = L_star - 1e-6 * abs(L_star)
    its = {}
    for name, (P0, Q0) in {"svd_als": (P_a, Q_a), "learned": (P_l, Q_l)}.items():
        _, _, ch = run_to_stop(G, P0, Q0, K, cuda, "f64", 0.0)
        its[name] = first_reach(ch, target)
    print(f"mean L1 to converged Q: ALS {mae_als:.4f}, network {mae_net:.4f}; start logL: ALS {ll_a:.1f}, "
          f"learned {ll_l:.1f} (used={used}); iterations to 1e-6: {its}", flush=True)
    assert mae_net < 0.8 * mae_als
    # the gate (over its SNP band) chose the network's start, which is also
    # the better one over all SNPs
    assert used and ll_gate > gate_loglik(G, P_a, Q_a) and ll_l > ll_a
