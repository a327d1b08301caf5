"""Checks at the full sizes of the benchmark workloads, where the oracle
cannot run: N=100,000 x M=500,000, K=10 (12.5 GB packed; the headline
workload) and N=10,000 x M=100,000, K=8 (the FP32 comparison workload).
Genotypes come from the device generator (bit-exact with the oracle's at
small sizes, test_gpu_kernels.py::test_synth_bit_exact).  The properties
do not depend on size:

* every precision form's EM statistics against the FP64 pass at the same
  (P, Q), within the bound each form states (FP32 2e-4, mixed BF16/FP16
  1e-3, relative to the largest entry, as the small-size tests);
* the passes are deterministic (bit-identical statistics on a repeat);
* the log-likelihood is additive over SNP blocks (the pass's partial-sum
  tree against separate launches on byte-row slices of G);
* the log-product logL of a 256-SNP slice across all N samples against a
  direct FP64 sum of g log H + (2-g) log(1-H) (em_adam_gpu.py:238-246).
"""
import numpy as np
import pytest
import torch

from adamixture_amd.em_adam_gpu import EMAdamOptimizer, em_pass
from adamixture_amd.genotypes import PackedGenotypes, unpack_chunk
from adamixture_amd.synth import WORKLOADS, make_genotypes

pytestmark = pytest.mark.gpu


def _workload(name, cuda):
    w = WORKLOADS[name]
    print(f"case: full-size workload {name} N={w.N} M={w.M} K={w.K}", flush=True)
    G, P0, Q0 = make_genotypes(w, cuda)
    rng = np.random.default_rng(7)
    P = np.clip(P0 + rng.normal(0.0, 0.02, P0.shape), 1e-5, 1 - 1e-5)
    Q = Q0 + 0.05
    Q /= Q.sum(axis=1, keepdims=True)
    return w, G, torch.from_numpy(P).to(cuda), torch.from_numpy(Q).to(cuda)


@pytest.fixture(scope="module")
def biobank(cuda):
    data = _workload("biobank", cuda)
    yield data
    del data
    torch.cuda.empty_cache()


@pytest.fixture(scope="module")
def gpu_f32(cuda):
    data = _workload("gpu_f32", cuda)
    yield data
    del data
    torch.cuda.empty_cache()


def _stats(G, P, Q, precision):
    opt = EMAdamOptimizer(P.shape, Q.shape, 0.005, 0.8, 0.88, 1e-8, P.device, precision=precision)
    opt.accumulate(G, P, Q)
    out = [x.double() for x in (opt.A_accum, opt.B_accum, opt.T_accum)]
    del opt
    return out


def _check_forms(G, P, Q, forms):
    ref = _stats(G, P, Q, "f64")
    for prec, bound in forms:
        got = _stats(G, P, Q, prec)
        for name, g, want in zip("ABT", got, ref):
            scale = want.abs().max()
            err = ((g - want).abs().max() / scale).item()
            print(f"{prec} {name}: max |err| / max |ref| = {err:.2e}", flush=True)
            assert bool(((g - want).abs() <= bound * want.abs() + bound * scale).all()), (prec, name, err)
        again = _stats(G, P, Q, prec)
        assert all(torch.equal(a, b) for a, b in zip(got, again)), f"{prec} pass is not deterministic"


def test_biobank_forms_match_fp64(biobank):
    _, G, P, Q = biobank
    _check_forms(G, P, Q, (("f32", 2e-4), ("bf16", 1e-3)))


def test_gpu_f32_workload_forms_match_fp64(gpu_f32):
    _, G, P, Q = gpu_f32
    _check_forms(G, P, Q, (("f32", 2e-4), ("bf16", 1e-3)))


@pytest.mark.parametrize("name", ["biobank", "gpu_f32"])
def test_loglik_additive_over_snp_blocks(request, name):
    w, G, P, Q = request.getfixturevalue(name)
    full = em_pass(G, P, Q, loglik=True, stats=False).item()
    rows = G.data.shape[0]
    cuts = [0, 1, rows // 3, rows // 2 + 5, rows]             # byte rows (4 SNPs each); uneven blocks
    parts = []
    for r0, r1 in zip(cuts[:-1], cuts[1:]):
        m0, m1 = 4 * r0, min(4 * r1, w.M)
        Gc = PackedGenotypes(G.data[r0:r1], m1 - m0, w.N)
        parts.append(em_pass(Gc, P[m0:m1].contiguous(), Q, loglik=True, stats=False).item())
    total = sum(parts)
    print(f"{name}: logL {full:.6f}, sum over blocks {total:.6f}", flush=True)
    assert abs(full - total) <= 1e-12 * abs(full)


@pytest.mark.parametrize("name", ["biobank", "gpu_f32"])
def test_loglik_slice_matches_direct_sum(request, name):
    w, G, P, Q = request.getfixturevalue(name)
    s, c = 4 * 1000, 256                                      # SNPs s .. s+255, all N samples
    Gc = PackedGenotypes(G.data[s // 4:(s + c) // 4], c, w.N)
    got = em_pass(Gc, P[s:s + c].contiguous(), Q, loglik=True, stats=False).item()
    g = unpack_chunk(G, s, c).to(torch.float64)              # c x N, 3 = missing
    H = (P[s:s + c] @ Q.T).clamp(1e-10, 1 - 1e-10)
    obs = g != 3
    want = torch.where(obs, g * torch.log(H) + (2 - g) * torch.log1p(-H), torch.zeros_like(H)).sum().item()
    print(f"{name}: slice logL {got:.9f} direct {want:.9f}", flush=True)
    assert abs(got - want) <= 1e-12 * abs(want)
