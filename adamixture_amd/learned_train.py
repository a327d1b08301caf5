"""Training of the learned Q warm start on synthetic runs that the plain
method converges (project extension, no counterpart in the reference).

Each training set is drawn from the Balding-Nichols admixture model
(synth.draw_params) with N samples, M SNPs, K in [k_lo, k_hi], F_st in
[fst_lo, fst_hi] and Dirichlet(alpha) ancestry, alpha log-uniform in
[alpha_lo, alpha_hi]; genotypes are generated on the GPU.  The plain method
(randomized SVD + ALS, then Adam-EM with the reference defaults,
adamixture/entry.py:35-47,63-72, until its own stopping rule) gives the
target Q*; the ALS start keeps the label order of Q*, so (features, Q*) are
aligned row by row.  Components are permuted at random per row during
training (the network must not depend on the arbitrary cluster order).

    python -m adamixture_amd.learned_train --datasets 200 --out weights.npz
"""
from __future__ import annotations

import argparse
import json
import logging
import math
import time

import numpy as np
import torch

from .adamixture import _require_gpu
from .als_gpu import ALS_gpu
from .em_adam_gpu import em_pass, optimize_parameters_gpu
from .genotypes import allele_frequencies
from .learned_init import ALS_ITERS, D_OUT, WarmStartMLP, features
from .svd_gpu import SVD_gpu
from .synth import Workload, make_genotypes

log = logging.getLogger(__name__)

# Reference defaults of the Adam-EM optimizer (adamixture/entry.py:35-47,63-72).
ADAM_DEFAULTS = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, lr_decay=0.5, min_lr=1e-4, patience=3,
                     tol_adam=0.1, max_iter=10000, check=5)
SVD_DEFAULTS = dict(power=5, tol_svd=0.1, max_als=1000, tol_als=1e-4)


def draw_config(rng: np.random.Generator, k_lo=4, k_hi=12, fst_lo=0.02, fst_hi=0.2, alpha_lo=0.1, alpha_hi=1.0):
    K = int(rng.integers(k_lo, k_hi + 1))
    fst = float(rng.uniform(fst_lo, fst_hi))
    alpha = float(math.exp(rng.uniform(math.log(alpha_lo), math.log(alpha_hi))))
    return K, fst, alpha


def plain_start(G, f, K: int, seed: int, device, *, k_svd: int | None = None):
    """Randomized SVD at rank k_svd (default 2K) and ALS on its first K
    components: returns (U, S, V, P_als, Q_als)."""
    k_svd = 2 * K if k_svd is None else k_svd
    U, S, V = SVD_gpu(G, G.N, G.M, f, k_svd, seed, SVD_DEFAULTS["power"], SVD_DEFAULTS["tol_svd"], 0, device)
    P_als, Q_als = ALS_gpu(U[:, :K].contiguous(), S[:K].contiguous(), V[:, :K].contiguous(), f, seed, G.M, K,
                           SVD_DEFAULTS["max_als"], SVD_DEFAULTS["tol_als"], device)
    return U, S, V, P_als, Q_als


def converge_plain(G, P, Q, K: int, device, precision: str = "f64", on_check=None):
    """Adam-EM from (P, Q) with the reference defaults until it stops."""
    a = ADAM_DEFAULTS
    return optimize_parameters_gpu(G, P.clone(), Q.clone(), a["lr"], a["beta1"], a["beta2"], a["reg_adam"],
                                   a["max_iter"], a["check"], K, G.M, G.N, a["lr_decay"], a["min_lr"], a["patience"],
                                   a["tol_adam"], device, 0, 0, precision, on_check)


def align_columns(Q_ref: torch.Tensor, Q: torch.Tensor) -> list[int]:
    """Column permutation of Q that best matches Q_ref (max sum of column
    overlaps Q_ref^T Q, Hungarian method on the K x K overlap matrix)."""
    from scipy.optimize import linear_sum_assignment
    C = (Q_ref.T @ Q).cpu().numpy()
    rows, cols = linear_sum_assignment(-C)
    return [int(c) for _, c in sorted(zip(rows, cols))]


def make_training_set(seed: int, N: int, M: int, K: int, fst: float, alpha: float, miss: float, device,
                      precision: str = "f64", als_iters: int | None = ALS_ITERS):
    """(X, T, info) of one synthetic data set: features N x 64 from the ALS
    start (to tolerance, or ``als_iters`` iterations), targets N x 16 = the
    Q the plain method converges to (SVD + ALS to tolerance + Adam-EM),
    zero-padded, its columns aligned to the feature start's clusters."""
    dev = torch.device(device)
    w = Workload(f"train{seed}", N, M, K, "bn", alpha, miss, fst=fst, seed=seed)
    t0 = time.time()
    G, _, _ = make_genotypes(w, dev)
    f = allele_frequencies(G)
    U, S, V, P_als, Q_als = plain_start(G, f, K, seed, dev)
    if als_iters is None:
        P_s, Q_s = P_als, Q_als
    else:
        P_s, Q_s = ALS_gpu(U[:, :K].contiguous(), S[:K].contiguous(), V[:, :K].contiguous(), f, seed, G.M, K,
                           als_iters, SVD_DEFAULTS["tol_als"], dev)
    X = features(U, S, V, f, P_s, Q_s, K)
    n_checks = []
    P_star, Q_star = converge_plain(G, P_als, Q_als, K, dev, precision, on_check=lambda it, ll: n_checks.append(it))
    perm = align_columns(Q_s, Q_star)
    T = torch.zeros((N, D_OUT), dtype=torch.float32, device=dev)
    T[:, :K] = Q_star[:, perm].to(torch.float32)
    ll_als = em_pass(G, P_als, Q_als, loglik=True, stats=False).item()
    ll_short = em_pass(G, P_s.contiguous(), Q_s.contiguous(), loglik=True, stats=False).item()
    ll_star = em_pass(G, P_star, Q_star, loglik=True, stats=False).item()
    info = dict(seed=seed, K=K, fst=fst, alpha=alpha, ll_als=ll_als, ll_short_als=ll_short, ll_star=ll_star,
                permuted=perm != list(range(K)), iters=n_checks[-1] if n_checks else 0,
                seconds=time.time() - t0)
    del G
    return X, T, info


def _permute_components(X: torch.Tensor, T: torch.Tensor, kv: torch.Tensor, gen: torch.Generator):
    """Random permutation of the first K components of each row, applied to
    the three K-indexed feature blocks and the target together."""
    B = X.shape[0]
    cols = torch.arange(D_OUT, device=X.device).expand(B, D_OUT)
    key = torch.rand((B, D_OUT), device=X.device, generator=gen) + (cols >= kv[:, None]).to(torch.float32) * 2.0
    perm = torch.argsort(key, dim=1)
    X = X.clone()
    for b in range(3):  # ALS Q, least-squares Q, log ALS Q
        X[:, b * D_OUT:(b + 1) * D_OUT] = torch.gather(X[:, b * D_OUT:(b + 1) * D_OUT], 1, perm)
    return X, torch.gather(T, 1, perm)


def soft_xent(Qp: torch.Tensor, T: torch.Tensor) -> float:
    """Mean over rows of -sum t log q (zero entries of q skipped where t = 0)."""
    lq = torch.log(Qp.clamp_min(1e-30))
    return float(-(T * lq).sum(dim=1).mean())


def evaluate(model: WarmStartMLP, X: torch.Tensor, T: torch.Tensor, kv: torch.Tensor) -> dict:
    """Cross-entropy and mean |Q - Q*| of the ALS prior, the FP32 network and
    the BF16 inference kernel on (X, T)."""
    prior = X[:, :D_OUT].clamp_min(1e-5) * (torch.arange(D_OUT, device=X.device)[None, :] < kv[:, None])
    prior = prior / prior.sum(dim=1, keepdim=True)
    out = {"xent_prior": soft_xent(prior, T), "mae_prior": float((prior - T).abs().sum(1).mean())}
    P32 = torch.cat([model.probs(X[i:i + 65536], kv[i:i + 65536]) for i in range(0, X.shape[0], 65536)])
    out["xent_fp32"] = soft_xent(P32, T)
    out["mae_fp32"] = float((P32 - T).abs().sum(1).mean())
    mae_bf16, xent_bf16 = [], []
    for K in torch.unique(kv).tolist():
        sel = kv == K
        Qb = model.infer(X[sel], int(K)).to(torch.float32)
        Tk = T[sel][:, :K]
        mae_bf16.append(float((Qb - Tk).abs().sum(1).sum()))
        xent_bf16.append(float(-(Tk * torch.log(Qb.clamp_min(1e-30))).sum()))
    out["mae_bf16"] = sum(mae_bf16) / X.shape[0]
    out["xent_bf16"] = sum(xent_bf16) / X.shape[0]
    return out


def train(model: WarmStartMLP, X: torch.Tensor, T: torch.Tensor, kv: torch.Tensor, *, epochs: int = 40,
          batch: int = 2048, lr: float = 1e-3, seed: int = 0, log_every: int = 5) -> list[dict]:
    """Adam on the mean soft cross-entropy, cosine learning-rate decay,
    random component permutations per row."""
    gen = torch.Generator(device=X.device)
    gen.manual_seed(seed)
    n = X.shape[0]
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    hist = []
    step = 0
    for ep in range(epochs):
        order = torch.randperm(n, device=X.device, generator=gen)
        acc = torch.zeros((), dtype=torch.float32, device=X.device)
        for s in range(steps_per_epoch):
            idx = order[s * batch:(s + 1) * batch]
            xb, tb = _permute_components(X[idx], T[idx], kv[idx], gen)
            acc += model.loss_and_grad(xb, tb, kv[idx].contiguous())
            lr_t = lr * 0.5 * (1.0 + math.cos(math.pi * step / total))
            model.adam_step(lr_t)
            step += 1
        rec = {"epoch": ep + 1, "loss": float(acc) / steps_per_epoch}
        hist.append(rec)
        if (ep + 1) % log_every == 0 or ep == 0:
            print(f"[learned] epoch {ep + 1}/{epochs} loss {rec['loss']:.5f}", flush=True)
    return hist


def _sync_time(device) -> float:
    torch.cuda.synchronize(device)
    return time.perf_counter()


def run_to_stop(G, P0, Q0, K: int, device, precision: str, t_origin: float):
    """Adam-EM from (P0, Q0) with the reference defaults until its stopping
    rule; returns (P, Q, checks) with checks = [(iteration, logL, seconds
    since t_origin)] (each check reads logL back, so its time is synced)."""
    checks = []
    P, Q = converge_plain(G, P0, Q0, K, device, precision,
                          on_check=lambda it, ll: checks.append((it, ll, time.perf_counter() - t_origin)))
    return P, Q, checks


def first_reach(checks, target: float):
    """(iteration, seconds) of the first check with logL >= target, or None."""
    for it, ll, t in checks:
        if ll >= target:
            return it, t
    return None


class _Reached(Exception):
    pass


def race(G, f, K: int, seed: int, device, model: WarmStartMLP | None, *, plain_precision: str = "f32",
         fast_precision: str = "bf16", rel: float = 1e-6, extra_plain: tuple = (), omega: str = "device") -> dict:
    """Time from the start of initialisation to the first logL check within
    ``rel`` of L*, where L* is the FP64 logL of the converged plain GPU run
    (randomized SVD + ALS start, Adam-EM in ``plain_precision``).  Races:
    plain start in ``plain_precision`` (defines L*), plain start in
    ``fast_precision``, and (with a model) the learned start in
    ``fast_precision``.  Each precision in ``extra_plain`` (e.g. "f64", the
    reference's arithmetic) races the plain start too, stopped at the target.
    ``omega``: where every start's SVD draws its test matrix (svd_gpu.SVD_gpu;
    "device" as the reference GPU path, the same for all starts)."""
    from functools import partial

    from .svd_gpu import SVD_gpu
    _svd = partial(SVD_gpu, omega=omega)
    from .learned_init import learned_warm_start
    dev = torch.device(device)
    out = {"K": K, "rel": rel}
    # Untimed warm-up of both initialisations: the first call of a kernel in
    # a process loads its code object and the first workspace allocations
    # grow the caching allocator.  Without it that one-time cost would land
    # in the first timed start only (the plain one, which every plain form
    # reuses).
    U, S, V = _svd(G, G.N, G.M, f, K, seed, SVD_DEFAULTS["power"], SVD_DEFAULTS["tol_svd"], 0, dev)
    P_w, Q_w = ALS_gpu(U, S, V, f, seed, G.M, K, SVD_DEFAULTS["max_als"], SVD_DEFAULTS["tol_als"], dev)
    from .em_adam_gpu import EMAdamOptimizer
    for prec in dict.fromkeys((plain_precision, fast_precision) + tuple(extra_plain)):
        a = ADAM_DEFAULTS
        EMAdamOptimizer(P_w.shape, Q_w.shape, a["lr"], a["beta1"], a["beta2"], a["reg_adam"], dev,
                        precision=prec).run_em_step(G, P_w.contiguous(), Q_w.contiguous(), G.M)
    del P_w, Q_w
    if model is not None:
        U2, S2, V2 = _svd(G, G.N, G.M, f, 2 * K, seed, SVD_DEFAULTS["power"], SVD_DEFAULTS["tol_svd"], 0, dev)
        learned_warm_start(G, U2, S2, V2, f, K, model, seed)
        del U2, S2, V2
    del U, S, V
    t0 = _sync_time(dev)
    U, S, V = _svd(G, G.N, G.M, f, K, seed, SVD_DEFAULTS["power"], SVD_DEFAULTS["tol_svd"], 0, dev)
    P_als, Q_als = ALS_gpu(U, S, V, f, seed, G.M, K, SVD_DEFAULTS["max_als"], SVD_DEFAULTS["tol_als"], dev)
    t_init = _sync_time(dev) - t0
    ll_als = em_pass(G, P_als, Q_als, loglik=True, stats=False).item()
    t1 = _sync_time(dev)
    P1, Q1, ch_plain = run_to_stop(G, P_als, Q_als, K, dev, plain_precision, t1)
    L_star = em_pass(G, P1, Q1, loglik=True, stats=False).item()
    target = L_star - rel * abs(L_star)
    out["L_star"] = L_star
    out["init_logL_svd_als"] = ll_als

    def summary(name, t_init_s, checks, Pf, Qf, extra=None):
        hit = first_reach(checks, target)
        L_final = em_pass(G, Pf, Qf, loglik=True, stats=False).item()
        rec = {"init_s": round(t_init_s, 3), "iterations_run": checks[-1][0] if checks else 0,
               "converged_logL": L_final, "reached": hit is not None}
        if hit is not None:
            rec["iterations_to_target"] = hit[0]
            rec["seconds_to_target"] = round(t_init_s + hit[1], 3)
        rec["seconds_to_stop"] = round(t_init_s + (checks[-1][2] if checks else 0.0), 3)
        if checks and checks[-1][0] > 0:
            rec["ms_per_iteration"] = round(1e3 * checks[-1][2] / checks[-1][0], 3)
        if extra:
            rec.update(extra)
        out[name] = rec
        print(f"[race] {name}: {json.dumps(rec)}", flush=True)

    summary(f"plain_{plain_precision}", t_init, ch_plain, P1, Q1)
    if fast_precision != plain_precision:
        t2 = _sync_time(dev)
        P2, Q2, ch_fast = run_to_stop(G, P_als, Q_als, K, dev, fast_precision, t2)
        summary(f"plain_{fast_precision}", t_init, ch_fast, P2, Q2)
    if model is not None:
        t3 = _sync_time(dev)
        U2, S2, V2 = _svd(G, G.N, G.M, f, 2 * K, seed, SVD_DEFAULTS["power"], SVD_DEFAULTS["tol_svd"], 0, dev)
        P_w, Q_w, _, used = learned_warm_start(G, U2, S2, V2, f, K, model, seed)
        t_init_l = _sync_time(dev) - t3
        ll_w = em_pass(G, P_w, Q_w, loglik=True, stats=False).item()      # reported, untimed
        t4 = _sync_time(dev)
        P3, Q3, ch_l = run_to_stop(G, P_w, Q_w, K, dev, fast_precision, t4)
        summary(f"learned_{fast_precision}", t_init_l, ch_l, P3, Q3,
                {"init_logL": ll_w, "used_learned": bool(used)})
    for prec in extra_plain:
        checks = []

        def on_check(it, ll, checks=checks, t_origin=_sync_time(dev)):
            checks.append((it, ll, time.perf_counter() - t_origin))
            if ll >= target:
                raise _Reached

        try:
            converge_plain(G, P_als, Q_als, K, dev, prec, on_check)
        except _Reached:
            pass
        hit = first_reach(checks, target)
        rec = {"init_s": round(t_init, 3), "reached": hit is not None, "stopped_at_target": True}
        if hit is not None:
            rec["iterations_to_target"] = hit[0]
            rec["seconds_to_target"] = round(t_init + hit[1], 3)
        if checks and checks[-1][0] > 0:
            rec["ms_per_iteration"] = round(1e3 * checks[-1][2] / checks[-1][0], 3)
        out[f"plain_{prec}"] = rec
        print(f"[race] plain_{prec}: {json.dumps(rec)}", flush=True)
    # L* is the plain FP32 run's converged logL (the metric's definition);
    # the best converged logL over the raced forms is reported beside it
    finals = [r["converged_logL"] for r in out.values() if isinstance(r, dict) and "converged_logL" in r]
    if finals:
        out["best_converged_logL"] = max(finals)
        out["L_star_rel_gap_to_best"] = (max(finals) - L_star) / abs(L_star)
    base = out[f"plain_{plain_precision}"].get("seconds_to_target")
    for name in list(out):
        if isinstance(out[name], dict) and out[name].get("seconds_to_target") and base:
            out[name]["speedup_vs_plain"] = round(base / out[name]["seconds_to_target"], 3)
    for prec in extra_plain:
        ref = out[f"plain_{prec}"].get("seconds_to_target")
        for name in list(out):
            if isinstance(out[name], dict) and out[name].get("seconds_to_target") and ref:
                out[name][f"speedup_vs_plain_{prec}"] = round(ref / out[name]["seconds_to_target"], 3)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--datasets", type=int, default=200)
    ap.add_argument("--val", type=int, default=20, help="held-out data sets (from the total)")
    ap.add_argument("--N", type=int, default=5000)
    ap.add_argument("--M", type=int, default=50000)
    ap.add_argument("--miss", type=float, default=0.01)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--batch", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1000)
    ap.add_argument("--precision", default="f64", choices=["f64", "f32", "bf16"])
    ap.add_argument("--budget-s", type=float, default=0.0, help="stop generating data sets after this many seconds")
    ap.add_argument("--eval-sets", type=int, default=3, help="held-out data sets raced after training")
    ap.add_argument("--weights", default=None, help="skip training: race these weights on --eval-workload")
    ap.add_argument("--eval-workload", default="learned_eval", help="synth.WORKLOADS entry raced with --weights")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    dev = _require_gpu("cuda:0")
    if args.weights:
        from .synth import WORKLOADS
        model = WarmStartMLP.load(dev, args.weights)
        w = WORKLOADS[args.eval_workload]
        t0 = time.time()
        G, _, _ = make_genotypes(w, dev)
        f = allele_frequencies(G)
        print(f"[learned] {w.name}: N={w.N} M={w.M} K={w.K} generated in {time.time() - t0:.1f}s", flush=True)
        r = race(G, f, w.K, w.seed, dev, model, plain_precision="f32", fast_precision="bf16")
        r["workload"] = w.name
        print(json.dumps(r), flush=True)
        return 0
    if not args.out:
        ap.error("--out is required when training")
    rng = np.random.default_rng(args.seed)
    Xs, Ts, Ks, infos = [], [], [], []
    t_start = time.time()
    for d in range(args.datasets):
        K, fst, alpha = draw_config(rng)
        X, T, info = make_training_set(args.seed + d, args.N, args.M, K, fst, alpha, args.miss, dev, args.precision)
        Xs.append(X)
        Ts.append(T)
        Ks.append(torch.full((args.N,), K, dtype=torch.int32, device=dev))
        infos.append(info)
        print(f"[learned] set {d + 1}/{args.datasets} K={K} fst={fst:.3f} alpha={alpha:.2f} "
              f"ll_als={info['ll_als']:.1f} ll*={info['ll_star']:.1f} iters={info['iters']} "
              f"{info['seconds']:.1f}s", flush=True)
        if args.budget_s and time.time() - t_start > args.budget_s:
            print(f"[learned] data budget reached after {d + 1} sets", flush=True)
            break
    n_sets = len(Xs)
    n_val = min(args.val, max(1, n_sets // 10))
    tr = slice(0, n_sets - n_val)
    va = slice(n_sets - n_val, n_sets)
    Xtr, Ttr, Ktr = torch.cat(Xs[tr]), torch.cat(Ts[tr]), torch.cat(Ks[tr])
    Xva, Tva, Kva = torch.cat(Xs[va]), torch.cat(Ts[va]), torch.cat(Ks[va])
    model = WarmStartMLP(dev, seed=args.seed)
    before = evaluate(model, Xva, Tva, Kva)
    t0 = time.time()
    hist = train(model, Xtr, Ttr, Ktr, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    after = evaluate(model, Xva, Tva, Kva)
    print(f"[learned] validation before {json.dumps(before)}", flush=True)
    print(f"[learned] validation after  {json.dumps(after)}", flush=True)
    meta = dict(n_sets=n_sets, n_val=n_val, N=args.N, M=args.M, miss=args.miss, epochs=args.epochs,
                batch=args.batch, lr=args.lr, seed=args.seed, precision=args.precision, train_seconds=t_train,
                val_before=before, val_after=after, history=hist[-5:], sets=infos)
    model.save(args.out, meta)
    del Xtr, Ttr, Ktr, Xs, Ts, Ks
    races = []
    for e in range(args.eval_sets):
        K, fst, alpha = draw_config(rng)
        seed = args.seed + 100000 + e
        w = Workload(f"eval{seed}", args.N, args.M, K, "bn", alpha, args.miss, fst=fst, seed=seed)
        G, _, _ = make_genotypes(w, dev)
        f = allele_frequencies(G)
        r = race(G, f, K, seed, dev, model, plain_precision="f32", fast_precision="bf16")
        r.update(fst=fst, alpha=alpha, seed=seed)
        races.append(r)
        del G
    meta["races"] = races
    model.save(args.out, meta)
    print(f"[learned] wrote {args.out} ({t_train:.1f}s training, {n_sets} sets)", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
