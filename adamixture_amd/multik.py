"""Batched forms (project extension): a multi-K sweep in one captured HIP
graph, and a batch of per-sample SQP problems replayed from a graph.

The reference runs a K sweep one K after another (adamixture/src/main.py:
84-107 -> adamixture.py:503-602 per K: ALS start, then Adam-EM with a host
round trip at every logL check).  Here every K of the sweep is one ``_Run``
(its own P, Q, Adam state and EM accumulators on the device) and one
Adam-EM iteration of all active K is captured once as a graph
(``torch.cuda.CUDAGraph`` = hipGraph): prep / EM pass / finalize / Adam for
each K, with the Adam step count, lr and bias corrections in device memory
(adx_adam_step_sched) so that the same graph is replayed every iteration.
A second graph holds the FP64 log-likelihood passes of all active K, which
write one slot each of a device vector read back with a single copy every
``check`` iterations; the host then applies the reference's checkpoint rule
(em_adam_gpu.py:325-353) to each K.  When a K stops, the graphs are
re-captured without it.  The randomized SVD runs once at K_max and is
sliced per K, as adamixture.py:555-557 does.
"""
from __future__ import annotations

import logging
import time

import numpy as np
import torch

from . import _lib
from .genotypes import PackedGenotypes

log = logging.getLogger(__name__)


class _Run:
    """Device state of one K of the sweep."""

    def __init__(self, G: PackedGenotypes, P0: torch.Tensor, Q0: torch.Tensor, precision: str, hp: dict,
                 ll_slot: torch.Tensor, shared: dict):
        dev = G.device
        self.shared = shared
        self.G = G
        self.K = P0.shape[1]
        self.precision = precision
        self.hp = hp
        M, N, K = G.M, G.N, self.K
        z = lambda *s, dt=torch.float64: torch.zeros(*s, dtype=dt, device=dev)  # noqa: E731
        self.P = P0.to(dev, torch.float64).contiguous().clone()
        self.Q = Q0.to(dev, torch.float64).contiguous().clone()
        self.mP, self.vP, self.mQ, self.vQ = z(M, K), z(M, K), z(N, K), z(N, K)
        acc_dt = torch.float64 if precision == "f64" else torch.float32
        self.acc_tag = _lib.ADX_F64 if precision == "f64" else _lib.ADX_F32
        self.A, self.B, self.T = z(M, K, dt=acc_dt), z(M, K, dt=acc_dt), z(N, K, dt=acc_dt)
        self.P_EM, self.Q_EM = z(M, K), z(N, K)
        self.P_lo = self.Q_lo = None
        if precision == "f32":
            self.P_lo = self.P.to(torch.float32)
            self.Q_lo = self.Q.to(torch.float32)
        if precision == "bf16":
            nb = _lib.load().adx_em_bf16_operand_bytes(M, N, K, None)
            self.ops = torch.zeros(nb, dtype=torch.uint8, device=dev)
        self.sched = torch.tensor([0.0, hp["lr"], 0.0, 0.0], dtype=torch.float64, device=dev)
        self.ll_slot = ll_slot
        self.ws = None
        # host-side checkpoint state (em_adam_gpu.py:325-353)
        self.lr = hp["lr"]
        self.L_best = None
        self.wait = 0
        self.done = False
        self.iters = 0
        self.checks: list[tuple[int, float]] = []
        self.P_best = self.P.clone()
        self.Q_best = self.Q.clone()

    def _workspace(self):
        """Log-likelihood partial sums of this K (K = 0: no T slices)."""
        need = _lib.load().adx_em_workspace_bytes(_lib.ADX_F64, self.G.M, self.G.N, 0)
        if self.ws is None or self.ws.numel() < need:
            self.ws = torch.empty(max(need, 8), dtype=torch.uint8, device=self.G.device)
        return self.ws, need

    def _stats_workspace(self, dtag: int):
        """T partial slices of the FP32/FP64 pass; one buffer shared by every
        K of the sweep (their passes run one after another on one stream)."""
        need = _lib.load().adx_em_workspace_bytes(dtag, self.G.M, self.G.N, self.K)
        buf = self.shared.get("stats")
        if buf is None or buf.numel() < need:
            self.shared["stats"] = None
            buf = self.shared["stats"] = torch.empty(need, dtype=torch.uint8, device=self.G.device)
        return buf, need

    def iterate(self) -> None:
        """Enqueue one Adam-EM iteration (no host synchronisation)."""
        G, K = self.G, self.K
        M, N = G.M, G.N
        st = _lib.stream_handle(G.device)
        if self.precision == "bf16":
            need = _lib.load().adx_em_bf16_operand_bytes(M, N, K, None)
            if self.ops.numel() < need:          # tile shape changed (never inside a capture: the
                self.ops = torch.empty(need, dtype=torch.uint8, device=G.device)  # priming ran first)
            nb = self.ops.numel()
            _lib.call("adx_em_prep_bf16", self.P.data_ptr(), self.Q.data_ptr(), M, N, K, self.ops.data_ptr(), nb, st)
            _lib.call("adx_em_pass_bf16", G.data.data_ptr(), G.ldg, M, N, K, self.ops.data_ptr(), nb,
                      self.A.data_ptr(), self.B.data_ptr(), self.T.data_ptr(), st)
        else:
            lowp = self.precision == "f32"
            dtag = _lib.ADX_F32 if lowp else _lib.ADX_F64
            Pp, Qp = (self.P_lo, self.Q_lo) if lowp else (self.P, self.Q)
            ws, need = self._stats_workspace(dtag)
            _lib.call("adx_em_pass", dtag, G.data.data_ptr(), G.ldg, M, N, K, Pp.data_ptr(), K, Qp.data_ptr(), K,
                      self.A.data_ptr(), self.B.data_ptr(), self.T.data_ptr(), _lib.ADX_EM_STATS, None,
                      ws.data_ptr(), need, st)
        _lib.call("adx_em_finalize", self.acc_tag, self.P.data_ptr(), self.Q.data_ptr(), self.A.data_ptr(),
                  self.B.data_ptr(), self.T.data_ptr(), G.qbat().data_ptr(), M, N, K, self.P_EM.data_ptr(),
                  self.Q_EM.data_ptr(), st)
        _lib.call("adx_adam_step_sched", self.P.data_ptr(), self.Q.data_ptr(), self.P_EM.data_ptr(),
                  self.Q_EM.data_ptr(), self.mP.data_ptr(), self.vP.data_ptr(), self.mQ.data_ptr(),
                  self.vQ.data_ptr(), M, N, K, self.sched.data_ptr(), self.hp["beta1"], self.hp["beta2"],
                  self.hp["reg_adam"], _lib.ptr(self.P_lo), _lib.ptr(self.Q_lo), K, st)

    def loglik(self) -> None:
        """Enqueue the FP64 log-likelihood pass at the current (P, Q) into
        this run's slot of the shared vector."""
        G = self.G
        ws, need = self._workspace()
        _lib.call("adx_em_pass", _lib.ADX_F64, G.data.data_ptr(), G.ldg, G.M, G.N, self.K, self.P.data_ptr(),
                  self.K, self.Q.data_ptr(), self.K, None, None, None, _lib.ADX_EM_LOGLIK, self.ll_slot.data_ptr(),
                  ws.data_ptr(), need, _lib.stream_handle(G.device))

    def decide(self, it: int, L: float) -> None:
        """Checkpoint rule of em_adam_gpu.py:325-353 (the first call is the
        post-priming evaluation)."""
        hp = self.hp
        self.checks.append((it, L))
        if self.L_best is None:
            self.L_best = L
            return
        if L > self.L_best + hp["tol_adam"]:
            self.L_best = L
            self.P_best.copy_(self.P)
            self.Q_best.copy_(self.Q)
            self.wait = 0
            return
        self.wait += 1
        if self.wait >= hp["patience"]:
            self.lr = max(self.lr * hp["lr_decay"], hp["min_lr"])
            self.sched[1].fill_(self.lr)
            if self.lr <= hp["min_lr"]:
                self.done = True
            self.wait = 0


def _capture(fn) -> torch.cuda.CUDAGraph:
    g = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.graph(g, stream=s):
        fn()
    torch.cuda.current_stream().wait_stream(s)
    return g


def sweep(G: PackedGenotypes, starts: dict, precision: str = "f32", *, graph: bool = True, max_iter: int = 10000,
          check: int = 5, lr: float = 0.005, beta1: float = 0.80, beta2: float = 0.88, reg_adam: float = 1e-8,
          lr_decay: float = 0.5, min_lr: float = 1e-4, patience: int = 3, tol_adam: float = 0.1,
          autotune: bool = False) -> dict:
    """Adam-EM for every K of ``starts`` ({K: (P0, Q0)}) at once.

    Returns {K: {"P", "Q", "logL", "iterations", "checks"}} with the best
    (P, Q) of each K by the reference's rule.  ``graph=False`` enqueues the
    same kernels eagerly (the reference for the graph replay).
    """
    hp = dict(lr=lr, beta1=beta1, beta2=beta2, reg_adam=reg_adam, lr_decay=lr_decay, min_lr=min_lr,
              patience=patience, tol_adam=tol_adam)
    dev = G.device
    G.qbat()                                       # computed once, outside any capture
    Ks = sorted(starts)
    ll_all = torch.zeros(len(Ks), dtype=torch.float64, device=dev)
    shared: dict = {}
    runs = [_Run(G, *starts[K], precision, hp, ll_all[i:i + 1], shared) for i, K in enumerate(Ks)]
    if autotune:
        from .autotune import ensure_tuned
        for r in runs:
            ensure_tuned(G, r.P, r.Q, precision)

    def launch_iter(active):
        for r in active:
            if autotune:
                from .autotune import ensure_tuned
                ensure_tuned(G, r.P, r.Q, precision)
            r.iterate()

    def launch_ll(active):
        for r in active:
            r.loglik()

    # priming iteration and post-priming evaluation, eager (loads every kernel)
    launch_iter(runs)
    for r in runs:                                 # best = post-priming state (em_adam_gpu.py:322-327)
        r.P_best.copy_(r.P)
        r.Q_best.copy_(r.Q)
    launch_ll(runs)
    L = ll_all.cpu().numpy()
    for i, r in enumerate(runs):
        r.iters = 1
        r.decide(0, float(L[i]))
    active = list(runs)
    g_iter = g_ll = None
    it = 0
    t0 = time.time()
    while active and it < max_iter:
        if graph and g_iter is None:
            g_iter = _capture(lambda: launch_iter(active))
            g_ll = _capture(lambda: launch_ll(active))
        n = min(check, max_iter - it)
        for _ in range(n):
            if graph:
                g_iter.replay()
            else:
                launch_iter(active)
        it += n
        if graph:
            g_ll.replay()
        else:
            launch_ll(active)
        L = ll_all.cpu().numpy()
        still = []
        for r in active:
            r.iters = it
            r.decide(it, float(L[runs.index(r)]))
            if not r.done:
                still.append(r)
        if len(still) != len(active):
            g_iter = g_ll = None                      # re-capture without the finished K
            log.info(f"    sweep it {it}: K done {[r.K for r in active if r.done]} ({time.time() - t0:.1f}s)")
        active = still
    return {r.K: {"P": r.P_best, "Q": r.Q_best, "logL": r.L_best, "iterations": r.iters, "checks": r.checks}
            for r in runs}


def qp_batch(XtX: torch.Tensor, Xtz: torch.Tensor, X: torch.Tensor, vkk: torch.Tensor, *, graph: bool = True,
             replays: int = 5) -> tuple[torch.Tensor, float]:
    """Solve the simplex-constrained SQP batch (adx_sqp_solve, persistent
    wavefront-per-problem kernel with an atomic work queue) ``replays``
    times from a captured graph; returns (Xn, mean ms per solve)."""
    nprob, K = X.shape
    dev = X.device
    Xn = torch.empty_like(X)
    ws = torch.zeros(8, dtype=torch.uint8, device=dev)
    neg = (-Xtz).contiguous()

    def solve():
        _lib.call("adx_sqp_solve", 1, XtX.data_ptr(), neg.data_ptr(), X.data_ptr(), vkk.data_ptr(), Xn.data_ptr(),
                  nprob, K, ws.data_ptr(), 8, _lib.stream_handle(dev))

    solve()                                        # eager (loads the kernel)
    g = _capture(solve) if graph else None
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(replays):
        if graph:
            g.replay()
        else:
            solve()
    e1.record()
    e1.synchronize()
    return Xn, e0.elapsed_time(e1) / replays


def qp_workload(G: PackedGenotypes, P: torch.Tensor, Q: torch.Tensor, seed: int, cond_max: float = 1e4):
    """C4 QP batch: per-sample Hessians and gradients of the Q-side SQP step
    at (P, Q) (adx_grad_hess), each Hessian made SPD with condition number
    <= cond_max by a random SPD perturbation scaled to its spectrum."""
    from .br_qn_gpu import grad_hess
    XtX, Xtz = grad_hess(G, Q, P, "q")
    N, K = Q.shape
    gen = torch.Generator(device=Q.device)
    gen.manual_seed(seed)
    R = torch.randn((N, K, K), dtype=torch.float64, device=Q.device, generator=gen)
    RR = R @ R.transpose(1, 2).contiguous()
    # E = I + RR / (2 lambda_max(RR)): SPD with spectrum in [1, 1.5]
    E = torch.eye(K, dtype=torch.float64, device=Q.device) + RR / (2.0 * torch.linalg.eigvalsh(RR)[:, -1:, None])
    lam_max = torch.linalg.eigvalsh(XtX)[:, -1].clamp_min(1e-12)
    c = lam_max / (cond_max - 1.5)        # lambda_min(H) >= c, lambda_max(H) <= lam_max + 1.5 c
    H = XtX + c[:, None, None] * E
    lamH = torch.linalg.eigvalsh(H)
    return H.contiguous(), Xtz.contiguous(), (lamH[:, -1] / lamH[:, 0])


def main(argv=None) -> int:
    """Batched workload (N=50,000 x M=200,000, K=2..12, Balding-Nichols
    F_st=0.1, Dirichlet(0.3), 0.5% missing): SVD once at K=12, ALS per K,
    the graph-captured sweep to convergence, an iteration-capped comparison
    with the K-by-K schedule, and the 50,000-problem SQP batch.  Prints one
    JSON line."""
    import argparse
    import json

    from .adamixture import _require_gpu
    from .als_gpu import ALS_gpu
    from .em_adam_gpu import optimize_parameters_gpu
    from .genotypes import allele_frequencies
    from .solvers import v_kk
    from .svd_gpu import SVD_gpu
    from .synth import WORKLOADS, make_genotypes

    ap = argparse.ArgumentParser(description="multi-K sweep and SQP batch on one GPU")
    ap.add_argument("--k-min", type=int, default=2)
    ap.add_argument("--k-max", type=int, default=12)
    ap.add_argument("--precision", default="f32", choices=["f32", "bf16", "f64"])
    ap.add_argument("--cap", type=int, default=50, help="iterations of the capped graph vs K-by-K comparison")
    ap.add_argument("--autotune", action="store_true")
    ap.add_argument("--no-full", action="store_true", help="skip the sweep to convergence")
    args = ap.parse_args(argv)
    dev = _require_gpu("cuda:0")
    w = WORKLOADS["multik"]
    t0 = time.time()
    G, _, _ = make_genotypes(w, dev, K=w.K)
    f = allele_frequencies(G)
    U, S, V = SVD_gpu(G, G.N, G.M, f, args.k_max, w.seed, 5, 0.1, 0, dev)
    torch.cuda.synchronize()
    t_svd = time.time() - t0
    starts = {}
    t1 = time.time()
    for K in range(args.k_min, args.k_max + 1):
        starts[K] = ALS_gpu(U[:, :K].contiguous(), S[:K].contiguous(), V[:, :K].contiguous(), f, w.seed, G.M, K,
                            1000, 1e-4, dev)
    torch.cuda.synchronize()
    t_als = time.time() - t1
    out = {"workload": f"{w.name}: N={w.N} M={w.M} K={args.k_min}..{args.k_max}", "precision": args.precision,
           "svd_s": round(t_svd, 3), "als_s": round(t_als, 3)}
    print(f"[multik] svd {t_svd:.1f}s, als {t_als:.1f}s", flush=True)
    hp = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, lr_decay=0.5, min_lr=1e-4, patience=3,
              tol_adam=0.1)
    # capped comparison: graph-captured sweep vs the K-by-K schedule
    torch.cuda.synchronize()
    t2 = time.time()
    multik_res = sweep(G, starts, args.precision, graph=True, max_iter=args.cap, autotune=args.autotune, **hp)
    torch.cuda.synchronize()
    t_graph = time.time() - t2
    t3 = time.time()
    for K, (P0, Q0) in starts.items():
        optimize_parameters_gpu(G, P0.clone(), Q0.clone(), hp["lr"], hp["beta1"], hp["beta2"], hp["reg_adam"],
                                args.cap, 5, K, G.M, G.N, hp["lr_decay"], hp["min_lr"], hp["patience"],
                                hp["tol_adam"], dev, 0, 0, precision=args.precision)
    torch.cuda.synchronize()
    t_seq = time.time() - t3
    nK = len(starts)
    out["capped"] = {"iterations_per_k": args.cap, "graph_s": round(t_graph, 3), "k_by_k_s": round(t_seq, 3),
                     "graph_em_iters_per_s": round(nK * args.cap / t_graph, 2),
                     "k_by_k_em_iters_per_s": round(nK * args.cap / t_seq, 2)}
    print(f"[multik] capped: {json.dumps(out['capped'])}", flush=True)
    del multik_res
    if not args.no_full:
        t4 = time.time()
        res = sweep(G, starts, args.precision, graph=True, autotune=args.autotune, **hp)
        torch.cuda.synchronize()
        out["full"] = {"seconds": round(time.time() - t4, 3),
                       "per_k": {K: {"logL": r["logL"], "iterations": r["iterations"]} for K, r in res.items()}}
        print(f"[multik] full sweep {out['full']['seconds']:.1f}s", flush=True)
        del res
    # SQP batch: Q-side problems of the K_max start
    K = args.k_max
    P0, Q0 = starts[K]
    H, z, cond = qp_workload(G, P0, Q0, seed=w.seed)
    vkk = v_kk(K, dev)
    Xn, ms = qp_batch(H, z, Q0.contiguous(), vkk, graph=True, replays=10)
    _, ms_eager = qp_batch(H, z, Q0.contiguous(), vkk, graph=False, replays=3)
    active = ((Xn > 1.5e-5).sum(dim=1)).float()
    out["sqp_batch"] = {"problems": int(Q0.shape[0]), "K": K, "cond_max": float(cond.max()),
                        "ms_graph": round(ms, 4), "ms_eager": round(ms_eager, 4),
                        "problems_per_s": round(Q0.shape[0] / (ms * 1e-3), 1),
                        "free_vars_min_mean_max": [float(active.min()), round(float(active.mean()), 2),
                                                   float(active.max())]}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
