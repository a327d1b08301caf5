"""Hand-written closed forms the oracle and the kernels are checked against.

Adam (Kingma & Ba 2015, Algorithm 1) as the reference's Adam-EM uses it
(adamixture/model/em_adam_gpu.py:11-42, :201-224; the Cython twin
adamixture/src/utils_c/cython/em.pyx:92-209): the "gradient" is the EM
displacement d_t = target_t - theta_{t-1}, the step is added (ascent toward
the target), t is incremented before use, and after each step P and Q are
clamped to [1e-5, 1 - 1e-5] and Q rows are divided by their sums.

The moments are written as the closed-form sums
    m_t = (1 - b1) sum_{s<=t} b1^(t-s) d_s,   v_t = (1 - b2) sum_{s<=t} b2^(t-s) d_s^2
rather than the recursion the oracle and the kernels run, in pure-Python
scalar loops (small inputs only).  ``variant`` plants one plausible slip, so
a test can show that its check tells the slip apart from the right answer.
"""
from __future__ import annotations

import math

LO, HI = 1e-5, 1.0 - 1e-5

VARIANTS = ("exact", "no_bias_correction", "swapped_betas", "renorm_before_clamp", "eps_inside_sqrt",
            "t_from_zero_at_step_two")


def adam_em_steps(P0, Q0, P_targets, Q_targets, lr, b1, b2, eps, variant="exact"):
    """Run len(P_targets) Adam-EM steps from (P0, Q0) toward the given EM
    targets; returns (P, Q, mP, vP, mQ, vQ) as nested lists."""
    if variant not in VARIANTS:
        raise ValueError(variant)
    if variant == "swapped_betas":
        b1, b2 = b2, b1
    P = [[float(x) for x in row] for row in P0]
    Q = [[float(x) for x in row] for row in Q0]
    hist_P, hist_Q = [], []

    def moments(hist, r, c, t):
        m = (1.0 - b1) * sum(b1 ** (t - s) * hist[s - 1][r][c] for s in range(1, t + 1))
        v = (1.0 - b2) * sum(b2 ** (t - s) * hist[s - 1][r][c] ** 2 for s in range(1, t + 1))
        return m, v

    def adam(theta, target, hist, t):
        hist.append([[target[r][c] - theta[r][c] for c in range(len(theta[r]))] for r in range(len(theta))])
        tb = t
        if variant == "t_from_zero_at_step_two" and t >= 2:
            tb = t - 1
        for r in range(len(theta)):
            for c in range(len(theta[r])):
                m, v = moments(hist, r, c, t)
                if variant == "no_bias_correction":
                    mh, vh = m, v
                else:
                    mh, vh = m / (1.0 - b1 ** tb), v / (1.0 - b2 ** tb)
                den = math.sqrt(vh + eps) if variant == "eps_inside_sqrt" else math.sqrt(vh) + eps
                theta[r][c] += lr * mh / den

    def clamp(theta):
        for row in theta:
            for c in range(len(row)):
                row[c] = min(max(row[c], LO), HI)

    def renorm(theta):
        for row in theta:
            s = sum(row)
            for c in range(len(row)):
                row[c] = row[c] / s

    for t, (tp, tq) in enumerate(zip(P_targets, Q_targets), start=1):
        adam(P, tp, hist_P, t)
        adam(Q, tq, hist_Q, t)
        clamp(P)
        if variant == "renorm_before_clamp":
            renorm(Q)
            clamp(Q)
        else:
            clamp(Q)
            renorm(Q)
    T = len(P_targets)
    mv = []
    for theta, hist in ((P, hist_P), (Q, hist_Q)):
        m = [[moments(hist, r, c, T)[0] for c in range(len(theta[r]))] for r in range(len(theta))]
        v = [[moments(hist, r, c, T)[1] for c in range(len(theta[r]))] for r in range(len(theta))]
        mv.append((m, v))
    return P, Q, mv[0][0], mv[0][1], mv[1][0], mv[1][1]


def adam_em_case(seed: int = 0, M: int = 5, N: int = 4, K: int = 3, steps: int = 3):
    """A small Adam-EM case whose steps cross both clamp bounds: one P entry
    and one Q row start at the bounds with targets beyond them, so the clamp
    and its order against the Q renormalisation decide the result."""
    import numpy as np
    rng = np.random.default_rng(seed)
    P0 = rng.uniform(0.05, 0.95, size=(M, K))
    P0[0, 0] = HI - 1e-4
    P0[1, 1] = LO + 1e-4
    Q0 = rng.dirichlet(np.ones(K), size=N)
    Q0[0] = [1.0 - 2e-5 * (K - 1)] + [2e-5] * (K - 1)
    P_t, Q_t = [], []
    for s in range(steps):
        tp = np.clip(P0 + rng.normal(0.0, 0.1, size=(M, K)), 0.0, 1.0)
        tp[0, 0] = 1.0
        tp[1, 1] = 0.0
        tq = rng.dirichlet(np.ones(K), size=N)
        tq[0] = [1.0] + [0.0] * (K - 1)
        P_t.append(tp)
        Q_t.append(tq)
    return P0, Q0, P_t, Q_t
