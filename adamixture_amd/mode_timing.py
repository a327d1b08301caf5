"""Timing of the projection, supervised and cross-validation modes, of
the SVD + ALS initialisation and of the file readers on synthetic workloads
(``python -m adamixture_amd.mode_timing --mode cv``).

Prints one JSON line per mode: wall seconds per iteration (Adam-EM modes,
iterations after the priming step, logL checks included) or per fold (CV:
mask + 5 polish iterations + deviance).  Genotypes are generated on the
device by the counter-based generator of ``synth``.
"""
from __future__ import annotations

import argparse
import json
import time

import numpy as np
import torch

from .synth import WORKLOADS, make_genotypes

ADAM = dict(lr=0.005, beta1=0.80, beta2=0.88, reg_adam=1e-8, check=5, lr_decay=0.5, min_lr=1e-4, patience=3,
            tol_adam=0.1)


def _sync(dev) -> float:
    torch.cuda.synchronize(dev)
    return time.perf_counter()


def _start(w, P0, Q0, seed):
    rng = np.random.default_rng(seed)
    P = np.clip(P0 + rng.normal(0.0, 0.02, P0.shape), 1e-5, 1 - 1e-5)
    Q = rng.random(Q0.shape)
    return P, Q / Q.sum(axis=1, keepdims=True)


def time_projection(G, P0, Q0, w, dev, precision: str, iters: int) -> dict:
    from .projection import optimize_projection_gpu
    Pf = torch.from_numpy(np.clip(P0, 1e-5, 1 - 1e-5)).to(dev)
    _, Q = _start(w, P0, Q0, 1)

    def run(n):
        Qd = torch.from_numpy(Q).to(dev)
        t0 = _sync(dev)
        optimize_projection_gpu(G, Pf, Qd, ADAM["lr"], ADAM["beta1"], ADAM["beta2"], ADAM["reg_adam"], n,
                                ADAM["check"], w.M, ADAM["lr_decay"], ADAM["min_lr"], ADAM["patience"],
                                ADAM["tol_adam"], dev, 0, 0, precision=precision)
        return _sync(dev) - t0

    run(ADAM["check"])                                   # warm-up (code load, autotune-free defaults)
    t_short, t_long = run(ADAM["check"]), run(ADAM["check"] + iters)
    return {"mode": "projection_adamem", "precision": precision, "iterations": iters,
            "s_per_iteration": round((t_long - t_short) / iters, 5)}


def time_supervised(G, P0, Q0, w, dev, precision: str, iters: int) -> dict:
    from .supervised import init_q_supervised, label_freq_gpu, optimize_supervised_gpu
    rng = np.random.default_rng(3)
    y = np.where(rng.random(w.N) < 0.5, Q0.argmax(axis=1) + 1, 0).astype(np.int32)
    t0 = _sync(dev)
    P = label_freq_gpu(G, y, w.K)
    t_init = _sync(dev) - t0
    _, Q = _start(w, P0, Q0, 2)
    init_q_supervised(Q, y, w.K)

    def run(n):
        t0 = _sync(dev)
        optimize_supervised_gpu(G, P.clone(), torch.from_numpy(Q).to(dev), y, ADAM["lr"], ADAM["beta1"],
                                ADAM["beta2"], ADAM["reg_adam"], n, ADAM["check"], w.M, ADAM["lr_decay"],
                                ADAM["min_lr"], ADAM["patience"], ADAM["tol_adam"], dev, 0, 0, precision=precision)
        return _sync(dev) - t0

    run(ADAM["check"])
    t_short, t_long = run(ADAM["check"]), run(ADAM["check"] + iters)
    return {"mode": "supervised_adamem", "precision": precision, "iterations": iters,
            "s_per_iteration": round((t_long - t_short) / iters, 5), "label_freq_s": round(t_init, 5),
            "labelled": int((y > 0).sum())}


def time_cv(G, P0, Q0, w, dev, folds: int, timed_folds: int) -> dict:
    """Folds 0 .. timed_folds-1 of a ``folds``-fold CV on the hashed path."""
    from .cv import FoldScorer
    P, Q = _start(w, P0, Q0, 4)
    scorer = FoldScorer(G, torch.from_numpy(P).to(dev), torch.from_numpy(Q).to(dev), folds, 42, streaming=True)
    t0 = _sync(dev)
    sums = [scorer.fold_sum(f) for f in range(timed_folds)]
    dt = _sync(dev) - t0
    return {"mode": "cv", "folds": folds, "timed_folds": timed_folds, "path": "hashed",
            "s_per_fold": round(dt / timed_folds, 4), "fold_deviance_sums": sums}


def time_init(G, w, dev) -> dict:
    """Randomized SVD (power 5, tol 0.1) and ALS (to tol 1e-4) as the plain
    method initialises; the second of two runs is reported."""
    from .als_gpu import ALS_gpu
    from .genotypes import allele_frequencies
    from .svd_gpu import SVD_gpu
    f = allele_frequencies(G)
    out = {}
    for _ in range(2):
        t0 = _sync(dev)
        U, S, V = SVD_gpu(G, w.N, w.M, f, w.K, 42, 5, 0.1, 0, dev)
        t1 = _sync(dev)
        ALS_gpu(U, S, V, f, 42, w.M, w.K, 1000, 1e-4, dev)
        t2 = _sync(dev)
        out = {"mode": "init", "svd_s": round(t1 - t0, 4), "als_s": round(t2 - t1, 4)}
    return out


def _write_vcf(path, G: np.ndarray) -> None:
    """Unpacked M x N codes as a minimal VCF (GT only; 3 = './.')."""
    M, N = G.shape
    gt = np.frombuffer(b"0/0\t0/1\t1/1\t./.\t", dtype=np.uint8).reshape(4, 4)
    head = "##fileformat=VCFv4.3\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\t"
    head += "\t".join(f"s{i}" for i in range(N)) + "\n"
    with open(path, "wb") as fh:
        fh.write(head.encode())
        for m0 in range(0, M, 1024):
            blk = G[m0:m0 + 1024]
            body = gt[blk].reshape(blk.shape[0], -1)
            body[:, -1] = ord("\n")
            for r in range(blk.shape[0]):
                fh.write(b"1\t%d\t.\tA\tC\t.\t.\t.\tGT\t" % (m0 + r + 1))
                fh.write(body[r].tobytes())


def time_readers(dev, N: int = 10_000, M_bed: int = 100_000, M_vcf: int = 10_000) -> list[dict]:
    """File -> device packed matrix: PLINK .bed (N x M_bed) and VCF text
    (N x M_vcf) from one synthetic genotype matrix; the second of two reads
    is reported (file in the page cache), with the read-back checked against
    the source codes."""
    import tempfile
    from pathlib import Path
    from .formats import read_vcf_gpu
    from .genotypes import unpack_chunk
    from .reader import read_bed_gpu, write_bed
    from .synth import Workload
    w = Workload("readers", N, M_bed, 4, "bn", 0.3, 0.01, fst=0.1, seed=11)
    G, _, _ = make_genotypes(w, dev)
    host = unpack_chunk(G, 0, M_bed).cpu().numpy()
    out = []
    with tempfile.TemporaryDirectory() as tmp:
        base = Path(tmp) / "r"
        write_bed(base, host)
        vcf = Path(tmp) / "r.vcf"
        _write_vcf(vcf, host[:M_vcf])
        for kind, path, M, reader in (("bed", Path(f"{base}.bed"), M_bed, read_bed_gpu),
                                      ("vcf", vcf, M_vcf, read_vcf_gpu)):
            for _ in range(2):
                t0 = _sync(dev)
                Gr, n, m = reader(path, dev)
                dt = _sync(dev) - t0
            same = bool(np.array_equal(unpack_chunk(Gr, 0, M).cpu().numpy(), host[:M]))
            size = path.stat().st_size
            out.append({"mode": f"reader_{kind}", "samples": n, "snps": m, "file_MB": round(size / 1e6, 1),
                        "seconds": round(dt, 4), "file_GB_per_s": round(size / dt / 1e9, 3),
                        "genotypes_per_s": round(n * m / dt, 1), "matches_source": same})
            del Gr
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--mode", choices=["projection", "supervised", "cv", "init", "readers", "all"], default="all")
    ap.add_argument("--workload", default="biobank", choices=sorted(WORKLOADS))
    ap.add_argument("--precision", default="bf16", choices=["f64", "f32", "bf16"])
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--folds", type=int, default=5)
    ap.add_argument("--timed-folds", type=int, default=1)
    args = ap.parse_args(argv)
    dev = torch.device("cuda:0")
    if args.mode == "readers":
        for rec in time_readers(dev):
            print(json.dumps(rec), flush=True)
        return 0
    w = WORKLOADS[args.workload]
    G, P0, Q0 = make_genotypes(w, dev)
    base = {"workload": f"{w.name}: N={w.N} M={w.M} K={w.K}"}
    if args.mode in ("projection", "all"):
        print(json.dumps({**base, **time_projection(G, P0, Q0, w, dev, args.precision, args.iters)}), flush=True)
    if args.mode in ("supervised", "all"):
        print(json.dumps({**base, **time_supervised(G, P0, Q0, w, dev, args.precision, args.iters)}), flush=True)
    if args.mode in ("cv", "all"):
        print(json.dumps({**base, **time_cv(G, P0, Q0, w, dev, args.folds, args.timed_folds)}), flush=True)
    if args.mode in ("init", "all"):
        print(json.dumps({**base, **time_init(G, w, dev)}), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
