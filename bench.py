"""Adam-EM likelihood path on one MI355X, biobank workload (N=100,000 samples
x M=500,000 SNPs, K=10; Balding-Nichols Fst=0.1 P, Dirichlet(0.2) Q, 1%
missing; 12.5 GB of 2-bit genotypes resident in HBM).

Headline (``value``, seconds, lower is better): time from the start of
initialisation to the first FP64 logL check within 1e-6 relative of L*, the
FP64 logL of the converged plain FP32 run from the randomized-SVD + ALS start,
for the learned start + mixed-precision (BF16/FP16 MFMA) form.  The plain
FP32, plain BF16 and plain FP64 (the reference's arithmetic) forms race
beside it ("time_to_target").  Every form is judged by FP64 logL, so the
figure is precision-fair.

Per form ("forms": bf16, f32, f64): EM iterations/s, one step = one Adam-EM
iteration exactly as optimize_parameters_gpu runs it (the fused decode +
EM-statistics pass, EM finalize, Adam step, and every 5th step the FP64
log-likelihood check with its host read-back), W untimed warm-up steps then
K timed steps bracketed by device synchronisation; the dominant kernel's
average duration from HIP events on the launch stream, with its roofline.
``steps``, ``warmup`` and ``ms_per_step`` are the BF16 form's loop.

Also: one BR-QN iteration and its gradient / Hessian passes ("brqn"), and
the CPU baseline: the reference CPU EM iteration (the oracle's restatement
of em.pyx P_step + Q_step) on the host's cores and on one core, and the
full 100-iteration CPU Adam-EM of the 1,000 x 20,000, K=3 workload from
SVD + ALS init beside the same run on the GPU.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0
PEAK_MFMA_TFLOPS = {"f32": 157.3, "f64": 78.6, "bf16": 2500.0}
SIMDS, CLOCK_HZ = 1024, 2.4e9          # 256 CUs x 4 SIMDs; one wave64 VALU instruction per 4 cycles
KERNEL = {"bf16": "em_pass_bf16_kernel", "f32": "em_pass_kernel<float>", "f64": "em_pass_kernel<double>"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="biobank")
    ap.add_argument("--forms", default="bf16,f32,f64", help="comma list of forms to time (first: the headline's)")
    ap.add_argument("--cpu-rows", type=int, default=None,
                    help="SNPs in the bounded CPU-baseline sample (default: ~10-30 s of CPU work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-c1", action="store_true", help="skip the 100-iteration C1 CPU/GPU runs")
    ap.add_argument("--no-brqn", action="store_true")
    ap.add_argument("--copy-probe-gb", type=float, default=0.0,
                    help="run one device copy of this many GB first (calibrates HBM counters under rocprofv3 --pmc)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default EM tile shape instead of timing the candidates first")
    ap.add_argument("--tiles", default=None, help="fixed EM tile shape 'cw/waves/target_wgs' (implies --no-autotune)")
    ap.add_argument("--no-ttt", action="store_true",
                    help="skip the time-to-target race (the headline; ~40 s per seed)")
    ap.add_argument("--ttt-seeds", default="42,43,44,45,46", help="initialisation seeds of the race (median reported)")
    return ap.parse_args()


def host_cpu():
    """Cores this process may use (the affinity set, capped by OMP_NUM_THREADS
    when set, as the GPU host shares its CPUs), all logical CPUs, and the
    CPU model."""
    avail = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS")
    threads = min(avail, int(omp)) if omp and omp.isdigit() and int(omp) > 0 else avail
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return max(threads, 1), os.cpu_count(), model


def cpu_em_rate(G, P0, Q0, w, rows, threads):
    """One reference CPU EM iteration (em.pyx P_step + Q_step, FP64, OpenMP,
    built with the reference's -O3 -ffast-math -fno-wrapv -fopenmp) on the
    first `rows` SNPs x all samples, unpacked as the reference's CPU path
    holds them; returns (EM iterations/s of the full workload, seconds)."""
    import oracle as O
    Gs = O.unpack(G.data[: (rows + 3) // 4, : w.N].cpu().numpy(), min(rows, w.M))
    Ps = np.ascontiguousarray(P0[:rows])
    Q = np.ascontiguousarray(Q0)
    P1, Q1, T, qb = np.zeros_like(Ps), np.zeros_like(Q), np.zeros_like(Q), np.zeros(w.N)
    t = time.perf_counter()
    O.em_step_cpu(Gs, Ps, Q, T, P1, Q1, qb, nthreads=threads, fast=True)
    dt = time.perf_counter() - t
    return (rows / w.M) / dt, dt


def c1_runs(threads, dev):
    """The 1,000 x 20,000, K=3 workload (uniform P, Dirichlet(0.5) Q, 1%
    missing, seed 42): randomized SVD + ALS, then Adam-EM with the reference
    defaults capped at 100 iterations, on the host (the oracle's CPU
    restatement of svd.py / als.py / em_adam.py, EM statistics on `threads`
    cores) and on the GPU (FP64 form)."""
    import torch

    import oracle as O
    from adamixture_amd.adamixture import initialize_k, setup
    from adamixture_amd.em_adam_gpu import optimize_parameters_gpu
    from adamixture_amd.genotypes import as_packed
    from adamixture_amd.synth import WORKLOADS, draw_params
    w = WORKLOADS["cpu_ref"]
    P, Q = draw_params(w)
    Gp = O.synth_packed(P, Q, w.seed, w.miss)
    G = O.unpack(Gp, w.M)
    f = O.freq_cpu(G)

    def rows(s, n):
        X = G[s:s + n].astype(np.float32) - 2.0 * f[s:s + n, None].astype(np.float32)
        X[G[s:s + n] == 3] = 0.0
        return X

    t0 = time.perf_counter()
    U, S, V = O.rsvd(rows, w.M, w.N, w.K, w.seed, 5, 0.1, 8192)
    P0, Q0 = O.als(U, S, V, f, w.seed, w.M, w.N, w.K, 1000, 1e-4)
    t_init = time.perf_counter() - t0
    hist = []
    t0 = time.perf_counter()
    Pc, Qc = O.adamem_cpu(G, P0, Q0, max_iter=100, history=hist, nthreads=threads, fast=True)
    t_opt = time.perf_counter() - t0
    iters = hist[-1][0] if hist else 0
    cpu = {"init_s": round(t_init, 3), "optimize_s": round(t_opt, 3), "iterations": iters,
           "em_iters_per_s": round(iters / t_opt, 3) if t_opt > 0 else None,
           "total_s": round(t_init + t_opt, 3), "final_logL": O.loglik_cpu(G, Pc, Qc), "threads": threads}
    # the same run on the GPU, through the drop-in driver
    Gd = as_packed(torch.from_numpy(Gp), w.M, dev)
    # one untimed initialisation first: the first call of each kernel in the
    # process loads its code object (0.1-0.2 s here), as the race warms up too
    dw, tw, fw, _, _, _, Gw = setup(Gd, w.N, w.M, w.K, w.seed, 5, 0.1, 8192, dev, compute_svd=False)
    initialize_k(Gw, w.N, w.M, w.K, fw, w.seed, 5, 0.1, 1000, 1e-4, 8192, dw, tw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_obj, tpb, fd, _, _, _, Gd = setup(Gd, w.N, w.M, w.K, w.seed, 5, 0.1, 8192, dev, compute_svd=False)
    Pg, Qg, _ = initialize_k(Gd, w.N, w.M, w.K, fd, w.seed, 5, 0.1, 1000, 1e-4, 8192, dev_obj, tpb)
    torch.cuda.synchronize()
    t_init_g = time.perf_counter() - t0
    checks = []
    t0 = time.perf_counter()
    Pg, Qg = optimize_parameters_gpu(Gd, Pg, Qg, 0.005, 0.8, 0.88, 1e-8, 100, 5, w.K, w.M, w.N, 0.5, 1e-4, 3, 0.1,
                                     dev_obj, 8192, tpb, precision="f64", on_check=lambda it, ll: checks.append(it))
    torch.cuda.synchronize()
    t_opt_g = time.perf_counter() - t0
    it_g = checks[-1] if checks else 0
    gpu = {"init_s": round(t_init_g, 3), "optimize_s": round(t_opt_g, 3), "iterations": it_g,
           "em_iters_per_s": round(it_g / t_opt_g, 1) if t_opt_g > 0 else None,
           "total_s": round(t_init_g + t_opt_g, 3), "final_logL": O.loglik_cpu(G, Pg.cpu().numpy(), Qg.cpu().numpy())}
    return {"workload": f"cpu_ref: N={w.N} M={w.M} K={w.K}, uniform P, Dirichlet({w.alpha}) Q, {w.miss:.0%} missing",
            "cpu": cpu, "gpu_f64": gpu,
            "speedup_total": round(cpu["total_s"] / gpu["total_s"], 1) if gpu["total_s"] > 0 else None}


def cpu_baseline(G, P0, Q0, w, rows, dev, c1):
    threads, nproc, model = host_cpu()
    rate, dt = cpu_em_rate(G, P0, Q0, w, rows, threads)
    rows1 = max(4, rows // threads)
    rate1, dt1 = cpu_em_rate(G, P0, Q0, w, rows1, 1)
    out = {"value": round(rate, 6), "unit": "EM iters/s", "cores": threads, "kind": "port",
           "sample": f"one EM iteration (P_step+Q_step, FP64, OpenMP) on {rows} of {w.M} SNPs x {w.N} samples, "
                     f"{dt:.2f} s on {threads} threads, scaled by M/{rows}",
           "cpu_model": model, "host_logical_cpus": nproc,
           "one_core": {"value": round(rate1, 6), "unit": "EM iters/s",
                        "sample": f"{rows1} SNPs x {w.N} samples, {dt1:.2f} s, scaled by M/{rows1}"}}
    if c1:
        out["c1_full_run"] = c1_runs(threads, dev)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the EM pass would shard by "
              "SNPs with one all-reduce of T and q_bat per iteration)", file=sys.stderr)
        sys.exit(2)
    import torch

    from adamixture_amd import _lib
    from adamixture_amd.em_adam_gpu import EMAdamOptimizer, em_pass
    from adamixture_amd.synth import WORKLOADS, draw_params, make_genotypes

    dev = torch.device("cuda:0")
    if args.copy_probe_gb > 0:
        n = int(args.copy_probe_gb * 1e9) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy probe: {4 * n} bytes read, {4 * n} bytes written", file=sys.stderr, flush=True)
        del src, dst
    w = WORKLOADS[args.workload]
    Ptrue, Qtrue = draw_params(w)
    G, _, _ = make_genotypes(w, dev, params=(Ptrue, Qtrue))
    rng = np.random.default_rng(w.seed + 1)
    P0 = np.clip(rng.random((w.M, w.K)), 1e-5, 1 - 1e-5)
    Q0 = rng.dirichlet(np.ones(w.K), size=w.N)
    G.qbat()
    M, N, K = w.M, w.N, w.K
    forms = [f for f in args.forms.split(",") if f]
    check = 5
    from adamixture_amd.autotune import get_tiles, set_tiles, tune_em
    default_tiles = {p: get_tiles(p) for p in ("bf16", "f32", "f64")}

    def time_form(prec):
        P = torch.from_numpy(P0).to(dev)
        Q = torch.from_numpy(Q0).to(dev)
        tuning = None
        if args.tiles:
            set_tiles(prec, *(int(x) for x in args.tiles.split("/")))
        elif not args.no_autotune and prec != "f64":
            tuning = tune_em(G, P, Q, prec)
        opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, dev, precision=prec)
        ev = []

        def iteration(i, timed):
            need_ll = (i % check) == 0
            e0 = e1 = None
            if timed:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            if prec == "f64":
                if timed:
                    e0.record()
                ll = opt.accumulate(G, P, Q, loglik=need_ll)
                if timed:
                    e1.record()
            elif prec == "bf16":
                st = _lib.stream_handle(dev)
                nb = opt.operands.numel()
                _lib.call("adx_em_prep_bf16", P.data_ptr(), Q.data_ptr(), M, N, K, opt.operands.data_ptr(), nb, st)
                if timed:
                    e0.record()
                _lib.call("adx_em_pass_bf16", G.data.data_ptr(), G.ldg, G.M, G.N, K, opt.operands.data_ptr(), nb,
                          opt.A_accum.data_ptr(), opt.B_accum.data_ptr(), opt.T_accum.data_ptr(), st)
                if timed:
                    e1.record()
                ll = em_pass(G, P, Q, dtype=_lib.ADX_F64, loglik=True, stats=False) if need_ll else None
            else:
                opt.P_lo.copy_(P)
                opt.Q_lo.copy_(Q)
                if timed:
                    e0.record()
                em_pass(G, opt.P_lo, opt.Q_lo, opt.A_accum, opt.B_accum, opt.T_accum, dtype=_lib.ADX_F32)
                if timed:
                    e1.record()
                ll = em_pass(G, P, Q, dtype=_lib.ADX_F64, loglik=True, stats=False) if need_ll else None
            if timed:
                ev.append((e0, e1, need_ll))
            if ll is not None:
                ll.item()
            opt.finalize(G, P, Q)
            opt.step(P, Q, opt.P_EM, opt.Q_EM)

        for i in range(args.warmup):
            iteration(i + 1, False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            iteration(i + 1, True)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        # the FP64 form's check iterations run the statistics pass and the
        # logL pass between the same events: the other iterations' statistics
        # passes are the kernel figure
        ms = [a.elapsed_time(b) for a, b, ll in ev if not (prec == "f64" and ll)]
        kern_ms = float(np.mean(ms))
        rec = {"em_iters_per_s": round(args.steps / elapsed, 4), "ms_per_step": round(1e3 * elapsed / args.steps, 3),
               "tiles": dict(zip(("cw", "waves", "target_wgs"), get_tiles(prec))),
               "roofline": roofline(prec, kern_ms, w)}
        if prec == "f64":
            ll_ms = [a.elapsed_time(b) for a, b, ll in ev if ll]
            if ll_ms:
                rec["pass_with_logL_ms"] = round(float(np.mean(ll_ms)), 3)
        if tuning is not None:
            rec["autotune_ms"] = {k: round(v, 3) for k, v in tuning["ms"].items()}
        del opt, P, Q
        torch.cuda.empty_cache()
        return rec, elapsed

    results = {}
    head_elapsed = None
    for prec in forms:
        print(f"[bench] timing form {prec}", file=sys.stderr, flush=True)
        results[prec], el = time_form(prec)
        if head_elapsed is None:
            head_elapsed = el
    head = forms[0]

    line = {
        "metric": "time to logL within 1e-6 rel of plain FP32 (Adam-EM, N=100k M=500k K=10), learned start + "
                  "mixed precision; EM iters/s per form beside it",
        "value": None,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": results[head]["ms_per_step"],
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": head,
        "data": "synthetic (counter-based binomial genotypes on device; numpy-seeded P, Q)",
        "config": {"workload": f"{w.name}: N={N} M={M} K={K}, P Balding-Nichols Fst={w.fst}, "
                               f"Q Dirichlet({w.alpha}), {w.miss:.1%} missing, 2-bit packed G "
                               f"{((M + 3) // 4) * N / 1e9:.1f} GB in HBM",
                   "logl_check_every": check, "timed_forms_init": "random (seeded)",
                   "race_init": "randomized SVD + ALS (plain) or learned warm start"},
        "roofline": results[head]["roofline"],
        "forms": results,
    }
    if not args.no_brqn:
        print("[bench] BR-QN", file=sys.stderr, flush=True)
        line["brqn"] = brqn_bench(G, P0, Q0, w, dev)
    if not args.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        rows = args.cpu_rows or max(4, min(M, int(4e9 / N)))
        line["cpu_baseline"] = cpu_baseline(G, P0, Q0, w, rows, dev, not args.no_c1)
    if not args.no_ttt:
        print("[bench] time-to-target race", file=sys.stderr, flush=True)
        ttt = time_to_target(G, w, dev, [int(x) for x in args.ttt_seeds.split(",") if x], default_tiles)
        line["time_to_target"] = ttt
        fast = ttt.get("learned_bf16") or ttt.get("plain_bf16") or {}
        line["value"] = fast.get("seconds_to_target")
        line["config"]["headline_form"] = "learned_bf16" if "learned_bf16" in ttt else "plain_bf16"
    else:
        # without the race the headline falls back to the EM rate of the first form
        line.update({"metric": f"EM iters/s ({head}, Adam-EM, N=100k M=500k K=10)",
                     "value": results[head]["em_iters_per_s"], "unit": "EM iters/s", "higher_is_better": True})
    print(json.dumps(line), flush=True)


def roofline(prec, kern_ms, w):
    """The pass's algorithmic flops (8 K N M) and bytes per launch against the
    form's MFMA peak and HBM, measured traffic and the bound named from the
    committed counter run of the same pass (VALU issue vs MFMA busy)."""
    M, N, K = w.M, w.N, w.K
    s = 8 if prec == "f64" else 4
    flops = 8.0 * K * N * M
    bytes_alg = ((M + 3) // 4) * N + (M * K + N * K) * s + (2 * M * K + N * K) * s
    traffic = None
    tfile = ROOT / "profiles" / f"em_pass_{prec}_{w.name}_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
    achieved = flops / (kern_ms * 1e-3) / 1e12
    peak = PEAK_MFMA_TFLOPS[prec]
    bound, valu = "mfma", None
    sqfile = ROOT / "profiles" / f"em_pass_{prec}_{w.name}_sq_counters.json"
    if sqfile.exists():
        sq = json.loads(sqfile.read_text())
        n_valu = sq["SQ_INSTS_VALU"] - sq["SQ_INSTS_MFMA"]
        cyc = SIMDS * CLOCK_HZ * kern_ms * 1e-3
        valu = {"instructions": n_valu, "issue_capacity": round(cyc / 4.0), "frac": round(n_valu / (cyc / 4.0), 4),
                "mfma_busy_frac": round(sq["SQ_VALU_MFMA_BUSY_CYCLES"] / cyc, 4)}
        bound = "valu" if valu["frac"] > valu["mfma_busy_frac"] else "mfma"
    return {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
            "frac": round(achieved / peak, 4), "traffic": traffic, "kernel": KERNEL[prec],
            "kernel_ms": round(kern_ms, 4), "algorithmic_flops": flops, "algorithmic_bytes": bytes_alg,
            "hbm_GBs_algorithmic": round(bytes_alg / (kern_ms * 1e-3) / 1e9, 1),
            "hbm_frac_algorithmic": round(bytes_alg / (kern_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4),
            "traffic_over_algorithmic": round(traffic / bytes_alg, 3) if traffic else None,
            "bound_basis": "SQ counters of the pass (VALU issue vs MFMA busy)" if valu else "form's matrix unit",
            "valu_issue": valu}


def brqn_bench(G, P0, Q0, w, dev):
    """The reference's default optimizer (SQP block relaxation + ZAL
    quasi-Newton): its Q- and P-side gradient / Hessian passes timed with HIP
    events (algorithmic flops per genotype: H 2K, gradient 2K, Hessian
    K(K+1) on both sides, plus 2K for the Q side's rank-one terms sum w2 p,
    against the FP64 peak), and two whole iterations from the seeded random
    start."""
    import torch

    from adamixture_amd.br_qn_gpu import grad_hess, optimize_original_gpu
    M, N, K = w.M, w.N, w.K
    P = torch.from_numpy(P0).to(dev)
    Q = torch.from_numpy(Q0).to(dev)
    out = {}
    for side, per in (("q", 2 * K + 2 * K + 2 * K + K * (K + 1)), ("p", 2 * K + 2 * K + K * (K + 1))):
        grad_hess(G, Q, P, side)
        ms = []
        for _ in range(2):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            grad_hess(G, Q, P, side)
            e1.record()
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        t = float(np.mean(ms))
        fl = float(per) * N * M
        out[f"grad_hess_{side}"] = {"ms": round(t, 2), "algorithmic_flops": fl,
                                    "achieved": round(fl / (t * 1e-3) / 1e12, 2), "peak": 78.6, "unit": "TFLOP/s",
                                    "frac": round(fl / (t * 1e-3) / 1e12 / 78.6, 4), "bound": "fp64 valu"}
    iters = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    optimize_original_gpu(G, P.clone(), Q.clone(), 2, K, M, N, 0.1, 3, 3, dev, 8192, 256,
                          on_iter=lambda it, ll: iters.append((it, time.perf_counter())))
    torch.cuda.synchronize()
    total = time.perf_counter() - t0
    out["iterations"] = len(iters)
    out["s_per_iteration"] = round(total / max(len(iters), 1), 3)
    out["note"] = "per iteration: 4 gradient/Hessian passes, 4 batched SQP solves, ZAL step, 1-2 FP64 logL passes"
    return out


def time_to_target(G, w, dev, seeds, default_tiles):
    """Seconds from the start of initialisation to the first logL check
    within 1e-6 relative of L* (FP64 logL of the converged plain FP32 run
    from the randomized-SVD + ALS start): plain FP32, plain BF16, the
    learned start + BF16 (when trained weights are present), and plain FP64
    (the reference's arithmetic; first seed only, stopped at the target).

    The iteration count to the target moves by whole checks (5 iterations)
    with the rounding of each form's trajectory, so the race is run for
    several initialisation seeds (randomized-SVD test matrix and ALS start)
    and the median is the figure; every race uses the library's default
    tile shapes, so it does not depend on the autotuner's pick."""
    from adamixture_amd.autotune import set_tiles
    from adamixture_amd.genotypes import allele_frequencies
    from adamixture_amd.learned_init import WEIGHTS_PATH, WarmStartMLP
    from adamixture_amd.learned_train import race
    for prec, tiles in default_tiles.items():
        set_tiles(prec, *tiles)
    model = WarmStartMLP.load(dev) if WEIGHTS_PATH.exists() else None
    f = allele_frequencies(G)
    runs = []
    for i, seed in enumerate(seeds):
        print(f"[bench] race seed {seed}", file=sys.stderr, flush=True)
        runs.append(race(G, f, w.K, seed, dev, model, plain_precision="f32", fast_precision="bf16",
                         extra_plain=("f64",) if i == 0 else ()))
    out = {"seeds": list(seeds), "rel": runs[0]["rel"], "L_star_definition":
           "FP64 logL of the converged plain FP32 run (randomized SVD + ALS start) of the same seed",
           "tiles": {k: dict(zip(("cw", "waves", "target_wgs"), v)) for k, v in default_tiles.items()},
           "per_seed": runs}
    for form in ("plain_f32", "plain_bf16", "learned_bf16", "plain_f64"):
        recs = [r[form] for r in runs if form in r and r[form].get("seconds_to_target")]
        if not recs:
            continue
        sec = [r["seconds_to_target"] for r in recs]
        agg = {"seconds_to_target": round(float(np.median(sec)), 3), "seconds_min": min(sec), "seconds_max": max(sec),
               "iterations_to_target": [r["iterations_to_target"] for r in recs],
               "init_s": round(float(np.median([r["init_s"] for r in recs])), 3),
               "ms_per_iteration": round(float(np.median([r["ms_per_iteration"] for r in recs
                                                          if "ms_per_iteration" in r] or [0.0])), 3),
               "races": len(recs)}
        sp = [r["speedup_vs_plain"] for r in recs if "speedup_vs_plain" in r]
        if sp:
            agg["speedup_vs_plain_f32"] = round(float(np.median(sp)), 3)
            agg["speedup_vs_plain_f32_per_seed"] = sp
        sp64 = [r["speedup_vs_plain_f64"] for r in recs if "speedup_vs_plain_f64" in r]
        if sp64:
            agg["speedup_vs_plain_f64"] = round(float(np.median(sp64)), 3)
        out[form] = agg
    lb = [r["learned_bf16"]["seconds_to_target"] <= r["plain_bf16"]["seconds_to_target"] for r in runs
          if r.get("learned_bf16", {}).get("seconds_to_target") and r.get("plain_bf16", {}).get("seconds_to_target")]
    if lb:
        out["learned_not_slower_than_plain_bf16"] = f"{sum(lb)}/{len(lb)} seeds"
    return out


if __name__ == "__main__":
    main()
