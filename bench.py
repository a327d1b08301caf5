"""Adam-EM throughput on one MI355X: EM iterations/s on the biobank workload
(N=100,000 samples x M=500,000 SNPs, K=10; Balding-Nichols Fst=0.1 P,
Dirichlet(0.2) Q, 1% missing; 12.5 GB of 2-bit genotypes resident in HBM).

A step is one Adam-EM iteration exactly as optimize_parameters_gpu runs it:
the fused decode+EM-statistics pass, EM finalize, Adam step, and every 5th
step the FP64 log-likelihood check (fused into the next pass in the f64
form, a separate FP64 pass in the f32 form) with its host read-back.

Then, unless --no-ttt, the workload's headline race: seconds from the start
of initialisation (randomized SVD + ALS, or the learned warm start) to the
first FP64 logL check within 1e-6 relative of the converged plain FP32
result, for the plain FP32, plain BF16 and learned-start BF16 forms
("time_to_target" in the JSON line).

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="biobank")
    ap.add_argument("--precision", default="bf16", choices=["bf16", "f32", "f64"])
    ap.add_argument("--cpu-rows", type=int, default=None,
                    help="SNPs in the bounded CPU-baseline sample (default: ~10-30 s of CPU work)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-gb", type=float, default=0.0,
                    help="run one device copy of this many GB first (calibrates HBM counters under rocprofv3 --pmc)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default EM tile shape instead of timing the candidates first")
    ap.add_argument("--tiles", default=None, help="fixed EM tile shape 'cw/target_wgs' (implies --no-autotune)")
    ap.add_argument("--no-ttt", action="store_true",
                    help="skip the time-to-target race (plain FP32 vs BF16 vs learned start + BF16, ~1 min)")
    return ap.parse_args()


def cpu_baseline(G, P0, Q0, w, rows):
    """The reference CPU EM iteration (em.pyx P_step + Q_step restated in
    oracle/adx_oracle.c, compiled with the reference's -O3 -ffast-math
    -fno-wrapv -fopenmp) on the first `rows` SNPs x all samples, unpacked
    as the reference's CPU path holds them; scaled to EM iterations/s of the
    full workload."""
    import oracle as O
    threads = min(16, os.cpu_count() or 1)
    Gs = O.unpack(G.data[: (rows + 3) // 4, : w.N].cpu().numpy(), min(rows, w.M))
    Ps = np.ascontiguousarray(P0[:rows])
    Q = np.ascontiguousarray(Q0)
    P1 = np.zeros_like(Ps)
    Q1 = np.zeros_like(Q)
    T = np.zeros_like(Q)
    qb = np.zeros(w.N)
    t = time.perf_counter()
    O.em_step_cpu(Gs, Ps, Q, T, P1, Q1, qb, nthreads=threads, fast=True)
    dt = time.perf_counter() - t
    return {"value": round((rows / w.M) / dt, 6), "unit": "EM iters/s", "cores": threads, "kind": "port",
            "sample": f"one EM iteration (P_step+Q_step, FP64, OpenMP) on {rows} of {w.M} SNPs x {w.N} "
                      f"samples, {dt:.2f} s, scaled by M/{rows}"}


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
    P = torch.from_numpy(P0).to(dev)
    Q = torch.from_numpy(Q0).to(dev)
    G.qbat()
    tuning = None
    if args.tiles:
        from adamixture_amd.autotune import set_tiles
        set_tiles(args.precision, *(int(x) for x in args.tiles.split("/")))
    elif not args.no_autotune:
        from adamixture_amd.autotune import tune_em
        tuning = tune_em(G, P, Q, args.precision)
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, dev, precision=args.precision)
    check = 5
    ev = []

    def iteration(i, timed):
        need_ll = (i % check) == 0
        if timed:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        if args.precision == "f64":
            ll = opt.accumulate(G, P, Q, loglik=need_ll)
            if timed:
                e1.record()
        elif args.precision == "bf16":
            st = _lib.stream_handle(dev)
            nb = opt.operands.numel()
            _lib.call("adx_em_prep_bf16", P.data_ptr(), Q.data_ptr(), w.M, w.N, w.K, opt.operands.data_ptr(), nb, st)
            if timed:
                e0.record()
            _lib.call("adx_em_pass_bf16", G.data.data_ptr(), G.ldg, G.M, G.N, w.K, opt.operands.data_ptr(), nb,
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
            ev.append((e0, e1))
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
    kern_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))

    s = 8 if args.precision == "f64" else 4      # bytes per P/Q/A/B/T element (FP32 accumulators)
    M, N, K = w.M, w.N, w.K
    flops = 8.0 * K * N * M
    bytes_alg = ((M + 3) // 4) * N + (M * K + N * K) * s + (2 * M * K + N * K) * s
    traffic = None
    tfile = ROOT / "profiles" / f"em_pass_{args.precision}_{w.name}_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
    achieved = flops / (kern_ms * 1e-3) / 1e12
    peak = PEAK_MFMA_TFLOPS[args.precision]
    # The pass is bound by VALU issue, not by MFMA or HBM: its vector
    # instruction count per dispatch (SQ_INSTS_VALU - SQ_INSTS_MFMA, from the
    # committed counter run of this workload and build) against the issue
    # capacity of the measured kernel time (1024 SIMDs x 2.4 GHz, one wave64
    # VALU instruction per 4 cycles).
    valu = None
    sqfile = ROOT / "profiles" / f"em_pass_{args.precision}_{w.name}_sq_counters.json"
    if sqfile.exists():
        sq = json.loads(sqfile.read_text())
        n_valu = sq["SQ_INSTS_VALU"] - sq["SQ_INSTS_MFMA"]
        cap = 1024 * 2.4e9 * kern_ms * 1e-3 / 4.0
        valu = {"instructions": n_valu, "issue_capacity": round(cap), "frac": round(n_valu / cap, 4),
                "mfma_busy_frac": round(sq["SQ_VALU_MFMA_BUSY_CYCLES"] / (1024 * 2.4e9 * kern_ms * 1e-3), 4)}
    line = {
        "metric": "EM iters/s (Adam-EM, N=100k M=500k K=10)",
        "value": round(args.steps / elapsed, 4),
        "unit": "EM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.precision,
        "data": "synthetic (counter-based binomial genotypes on device; numpy-seeded P, Q)",
        "config": {"workload": f"{w.name}: N={N} M={M} K={K}, P Balding-Nichols Fst={w.fst}, "
                               f"Q Dirichlet({w.alpha}), {w.miss:.1%} missing, 2-bit packed G "
                               f"{((M + 3) // 4) * N / 1e9:.1f} GB in HBM",
                   "precision": args.precision, "logl_check_every": check, "init": "random (seeded)"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": traffic,
                     "kernel": "em_pass_bf16_kernel" if args.precision == "bf16" else "em_pass_kernel",
                     "kernel_ms": round(kern_ms, 4),
                     "algorithmic_flops": flops, "algorithmic_bytes": bytes_alg,
                     "hbm_GBs_algorithmic": round(bytes_alg / (kern_ms * 1e-3) / 1e9, 1),
                     "valu_issue": valu},
    }
    from adamixture_amd.autotune import get_tiles
    line["config"]["tiles"] = dict(zip(("cw", "target_wgs"), get_tiles(args.precision)))
    if tuning is not None:
        line["config"]["autotune_ms"] = {k: round(v, 3) for k, v in tuning["ms"].items()}
    if not args.no_cpu_baseline:
        rows = args.cpu_rows or max(4, min(M, int(4e9 / N)))
        line["cpu_baseline"] = cpu_baseline(G, P0, Q0, w, rows)
    if not args.no_ttt:
        del opt, P, Q
        line["time_to_target"] = time_to_target(G, w, dev)
    print(json.dumps(line), flush=True)


def time_to_target(G, w, dev):
    """Seconds from the start of initialisation to the first logL check
    within 1e-6 relative of L* (FP64 logL of the converged plain FP32 run
    from the randomized-SVD + ALS start): plain FP32, plain BF16, the
    learned start + BF16 (when trained weights are present), and plain FP64
    (the reference's arithmetic; stopped once it reaches the target)."""
    from adamixture_amd.genotypes import allele_frequencies
    from adamixture_amd.learned_init import WEIGHTS_PATH, WarmStartMLP
    from adamixture_amd.learned_train import race
    model = WarmStartMLP.load(dev) if WEIGHTS_PATH.exists() else None
    f = allele_frequencies(G)
    return race(G, f, w.K, w.seed, dev, model, plain_precision="f32", fast_precision="bf16", extra_plain=("f64",))


if __name__ == "__main__":
    main()
