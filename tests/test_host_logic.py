"""Host-side logic of the drop-in layer that needs no GPU: the memory
manager's placement decision (device, host-streamed, or out of memory) with
a faked free-memory value (the reference's
own approach to testing it, adamixture/src/utils.py:311-429), the OOM
classifier and chunk halving (utils.py:19-89), the .P / .Q writer
format (utils.py:252-273), and the bit identities behind the EM passes'
select-free genotype weights and log-likelihood factors."""
import numpy as np
import pytest
import torch

from adamixture_amd import utils


def test_manage_gpu_memory_places_or_raises(monkeypatch):
    M, N, K = 500_000, 100_000, 10
    G = torch.zeros(1, dtype=torch.uint8).expand((M + 3) // 4, N)   # host tensor of the full shape, no storage
    placed = []
    monkeypatch.setattr(utils, "as_packed", lambda g, m, dev: placed.append((m, str(dev))) or "device-copy")
    need = ((M + 3) // 4) * ((N + 63) // 64 * 64) + utils.estimate_gpu_bytes(M, N, K, "adamem")
    monkeypatch.setattr(utils, "get_free_gpu_memory", lambda dev: need / 2**20 / 0.95 * 1.01)
    assert utils.manage_gpu_memory(G, torch.device("cuda:0"), M, N, K, 8192, "adamem") == "device-copy"
    assert placed == [(M, "cuda:0")]
    # G does not fit, the buffers and a staging buffer do: host-streamed
    streamed = []
    monkeypatch.setattr(utils.HostPackedGenotypes, "from_packed",
                        classmethod(lambda cls, g, m, dev, staging_cap: streamed.append((m, staging_cap)) or "host"))
    monkeypatch.setattr(utils, "get_free_gpu_memory", lambda dev: need / 2**20 / 0.95 * 0.99)
    assert utils.manage_gpu_memory(G, torch.device("cuda:0"), M, N, K, 8192, "adamem") == "host"
    assert streamed == [(M, utils.STREAM_STAGING_BYTES)]
    # not even the buffers fit: out of memory, as the reference classifies it
    buffers = utils.estimate_gpu_bytes(M, N, K, "adamem")
    monkeypatch.setattr(utils, "get_free_gpu_memory", lambda dev: buffers / 2**20 / 0.95 * 0.99)
    with pytest.raises(RuntimeError, match="out of memory") as ei:
        utils.manage_gpu_memory(G, torch.device("cuda:0"), M, N, K, 8192, "adamem")
    assert utils.is_gpu_oom(ei.value)
    # a CPU device leaves G where it is
    assert utils.manage_gpu_memory(G, torch.device("cpu"), M, N, K, 8192, "adamem") is G


def test_estimate_gpu_bytes_orders_algorithms():
    M, N, K = 200_000, 50_000, 12
    a = utils.estimate_gpu_bytes(M, N, K, "adamem")
    b = utils.estimate_gpu_bytes(M, N, K, "brqn")
    assert a > 0 and b >= ((M + N) * K * K) * 8          # BR-QN keeps (M+N) K^2 FP64 Hessians
    # Adam-EM keeps the FP64 pass's partial sums, capped at max(G / 10, 256 MiB)
    g_bytes = ((M + 3) // 4) * ((N + 63) // 64 * 64)
    cap = max(g_bytes // 10, 256 << 20)
    assert a >= (8 * M * K + 7 * N * K) * 8 + cap
    assert a <= (8 * M * K + 7 * N * K + N) * 8 + 2 * (M + N) * K * 4 + cap + (M * 20 * 4 * 3)
    assert utils.estimate_gpu_bytes(M, N, K, "adamem", include_initialization=False) <= a
    assert utils.estimate_gpu_bytes(2 * M, N, K, "adamem") > a
    with pytest.raises(ValueError):
        utils.estimate_gpu_bytes(M, N, K, "nope")


def test_oom_classifier_and_chunk_halving():
    assert utils.is_gpu_oom(RuntimeError("HIP error: out of memory"))
    assert not utils.is_gpu_oom(RuntimeError("invalid argument"))
    exc = RuntimeError("out of memory")
    assert utils.reduce_gpu_chunk_or_raise(8192, exc, "testing") == 4096
    assert utils.reduce_gpu_chunk_or_raise(3, exc, "testing") == utils.MIN_GPU_CHUNK_SIZE
    with pytest.raises(RuntimeError):
        utils.reduce_gpu_chunk_or_raise(utils.MIN_GPU_CHUNK_SIZE, exc, "testing")


def test_write_outputs_admixture_text_format(tmp_path):
    Q = np.array([[0.25, 0.75], [1e-5, 1 - 1e-5]])
    P = np.array([[0.1, 0.9]])
    utils.write_outputs(Q, "run", 2, tmp_path, P)
    lines = (tmp_path / "run.2.Q").read_text().splitlines()
    assert lines[0] == "2.500000000000000000e-01 7.500000000000000000e-01"
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "run.2.Q"), Q)
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "run.2.P", ndmin=2), P)


def test_supervised_label_host_logic(tmp_path):
    from adamixture_amd import supervised as S
    with pytest.raises(ValueError):
        S._check_labels(np.array([0, 1, 4]), 3, 3)
    with pytest.raises(ValueError):
        S._check_labels(np.array([0, 1]), 3, 3)
    (tmp_path / "pop.txt").write_text("x\n-\n\ny\nx\n")
    y, ids = S.labels_to_y(S.parse_labels_file(tmp_path / "pop.txt"))
    assert y.tolist() == [1, 0, 2, 1] and ids == {"x": 1, "y": 2} and y.dtype == np.int32
    Q = np.random.default_rng(0).random((4, 2))
    S.init_q_supervised(Q, y, 2)
    np.testing.assert_array_equal(Q[[0, 3]], [[1 - 1e-5, 1e-5], [1 - 1e-5, 1e-5]])
    np.testing.assert_array_equal(Q[2], [1e-5, 1 - 1e-5])
    np.testing.assert_allclose(Q.sum(axis=1), 1.0, rtol=1e-15)


def test_cv_pack_host_matches_oracle_pack():
    import oracle as O
    from adamixture_amd.cv import pack_host
    for M, N in ((1, 3), (5, 7), (64, 33)):
        G = np.random.default_rng(M).integers(0, 4, size=(M, N), dtype=np.uint8)
        np.testing.assert_array_equal(pack_host(G), O.pack(G))


def test_chromosome_filter_matches_reference_counts():
    """The reference's tests/test_reader.py expects 682 SNPs for chromosome
    1, 1454 for 1 and 2, the same through 'chr1', all 8451 when chrom_mode is
    'all', and none for chromosome 9999 (the CHROM column of the demo data)."""
    from adamixture_amd.formats import ChromFilter, normalize_specific_chrom
    from tests.conftest import GOLDEN
    chroms = [line.split()[0].encode() for line in (GOLDEN / "demo_data.bim").read_text().splitlines()]
    count = lambda f: sum(f(c) for c in chroms)                                  # noqa: E731
    assert count(ChromFilter("autosomes", 22, [1])) == 682
    assert count(ChromFilter("autosomes", 22, ["1", "2"])) == 1454
    assert count(ChromFilter("autosomes", 22, ["chr1"])) == 682
    assert count(ChromFilter("all", 22, [1])) == 8451
    assert count(ChromFilter("autosomes", 22, [9999])) == 0
    assert count(ChromFilter("autosomes", 22)) == 8451
    assert normalize_specific_chrom("chr1, 2 X") == [1, 2, "X"]
    f = ChromFilter("autosomes", 3)
    assert [f(t) for t in (b"1", b"chr3", b"Chr2", b"4", b"X", b"", b"chr")] == [True, True, True, False, False,
                                                                               False, False]
    g = ChromFilter("autosomes", 22, ["X", "chrY"])
    assert [g(t) for t in (b"X", b"chrX", b"Y", b"chrY", b"1")] == [True, True, True, True, False]
    with pytest.raises(ValueError):
        ChromFilter("some", 22)


def test_line_blocks_and_bcf_index():
    import gzip
    import io
    from adamixture_amd.formats import ChromFilter, _line_blocks, bcf_index
    from tests.conftest import GOLDEN
    text = b"a\tb\nccc\n\ndd"
    for size in (1, 2, 3, 100):
        blocks = list(_line_blocks(io.BytesIO(text), size))
        assert all(b.endswith(b"\n") for b in blocks) and b"".join(blocks) == text + b"\n"
    data = gzip.decompress((GOLDEN / "demo_data.bcf").read_bytes())
    N, rs, re_, go, gd, skipped = bcf_index(data, ChromFilter("all", 22))
    assert N == 105 and len(rs) == 8451 and skipped == 0 and np.all(gd >= 0)
    assert np.all(go >= rs) and np.all(go + 105 * (gd & 0xFFFF) * (gd >> 16) <= re_)
    N, rs, *_ = bcf_index(data, ChromFilter("autosomes", 22, ["1", "2"]))
    assert len(rs) == 1454


def test_pgen_layout_matches_oracle_header():
    import oracle as O
    from adamixture_amd.formats import pgen_layout
    from tests.conftest import GOLDEN
    data = (GOLDEN / "demo_data.pgen").read_bytes()
    for a, b in zip(pgen_layout(data), O.pgen_header(data)):
        np.testing.assert_array_equal(a, b)
    with pytest.raises(ValueError):
        pgen_layout(data[:2] + b"\x01" + data[3:])          # fixed-width storage mode
    with pytest.raises(ValueError):
        pgen_layout(data[:-1])                                # truncated records


def test_timing_vcf_writer_round_trips_through_oracle_parser(tmp_path):
    """The reader-throughput mode's synthetic VCF decodes (oracle parser) to
    the codes it was written from, missing included."""
    import oracle as O
    from adamixture_amd.mode_timing import _write_vcf
    G = np.random.default_rng(0).integers(0, 4, size=(1500, 37)).astype(np.uint8)
    _write_vcf(tmp_path / "x.vcf", G)
    got, N, M = O.vcf_to_unpacked((tmp_path / "x.vcf").read_bytes())
    assert (N, M) == (37, 1500)
    np.testing.assert_array_equal(got, G)


def test_packed_cache_is_keyed_by_tensor_identity():
    """Repeated setup calls with the same host matrix (the reference's
    multi-init loop, src/main.py:112-157) reuse one cache slot; another
    tensor gets its own; a dead source's slot is dropped."""
    import gc
    from adamixture_amd import genotypes as GT
    a = torch.zeros((3, 5), dtype=torch.uint8)
    b = torch.zeros((3, 5), dtype=torch.uint8)
    sa = GT._per_source(a)
    sa["x"] = 1
    assert GT._per_source(a) is sa and GT._per_source(b) is not sa
    n = len(GT._CACHE)
    del a
    gc.collect()
    assert len(GT._CACHE) == n - 1


def _code_masks(w):
    """Bitwise word masks of the log-likelihood pass: bit 2i of each is a flag of
    the 2-bit code in field i (observed; code != 2; code is 0 or 3)."""
    hi = w >> np.uint32(1)
    return ~(w & hi), ~(hi & ~w), ~(w ^ hi)


def test_select_free_loglik_factor_equals_the_select_form():
    """The log-likelihood pass forms each genotype's factor without selects:
    |H' - s1| |H' - s2| with s1 = [code != 2], s2 = [code in {0, 3}] read from
    word masks, and H' = H with its high word cleared for a missing code.  In
    IEEE FP64 (round to nearest) that equals the select form (1-H)^2, H (1-H),
    H^2 and 1 bit for bit, because |H - 1| is the rounded 1 - H and a
    subnormal H' gives H' - 1 = -1 exactly.  Checked on every code of random
    genotype words, with H at the domain's edges, at 1/2 and around it."""
    rng = np.random.default_rng(3)
    n = 200_000
    h = np.concatenate([rng.random(n - 8), [1e-10, 1 - 1e-10, 2e-10, 1 - 2e-10, 0.5,
                                            np.nextafter(0.5, 0), np.nextafter(0.5, 1), 0.25]])
    words = rng.integers(0, 2**32, size=n, dtype=np.uint64).astype(np.uint32)
    obs, m1, m2 = _code_masks(words)
    one = np.uint32(1)
    for p in range(0, 32, 2):
        code = (words >> np.uint32(p)) & np.uint32(3)
        u = 1.0 - h
        want = np.where(code == 3, 1.0, np.where(code == 2, h, u) * np.where(code == 0, u, h))
        keep = ((obs >> np.uint32(p)) & one).astype(np.uint64) * np.uint64(0xFFFFFFFF)
        hk = (h.view(np.uint64) & ((keep << np.uint64(32)) | np.uint64(0xFFFFFFFF))).view(np.float64)
        s1 = ((m1 >> np.uint32(p)) & one).astype(np.float64)
        s2 = ((m2 >> np.uint32(p)) & one).astype(np.float64)
        got = np.abs(hk - s1) * np.abs(hk - s2)
        np.testing.assert_array_equal(got.view(np.uint64), want.view(np.uint64))
    # a planted mistake: the masks of the wrong field give other factors
    code = words & np.uint32(3)
    s1_wrong = ((m1 >> np.uint32(2)) & one).astype(np.float64)
    wrong = np.abs(h - s1_wrong) * np.abs(h - ((m2 & one).astype(np.float64)))
    right = np.where(code == 2, h, 1.0 - h) * np.where(code == 0, 1.0 - h, h)
    assert np.mean((wrong != right)[code != 3]) > 0.3


def test_weight_words_give_g_and_two_minus_g():
    """The statistics passes read the EM weights g and 2 - g (0 for a missing
    code; em_adam_gpu.py:57-67 masks them) from two masked copies of each
    genotype word instead of converting the code and selecting: every 2-bit
    field of the first holds g, of the second 2 - g, on random words and on
    the four all-equal words."""
    rng = np.random.default_rng(4)
    words = np.concatenate([rng.integers(0, 2**32, size=100_000, dtype=np.uint64).astype(np.uint32),
                            np.array([0, 0x55555555, 0xAAAAAAAA, 0xFFFFFFFF], dtype=np.uint32)])
    m55 = np.uint32(0x55555555)
    lo, hi = words & m55, (words >> np.uint32(1)) & m55
    miss = lo & hi
    ga = words & ~(miss | (miss << np.uint32(1)))
    gb = ((~(lo | hi) & m55) << np.uint32(1)) | (lo & ~hi)
    for p in range(0, 32, 2):
        code = (words >> np.uint32(p)) & np.uint32(3)
        g_want = np.where(code == 3, 0, code)
        b_want = np.where(code == 3, 0, 2 - code.astype(np.int64))
        np.testing.assert_array_equal((ga >> np.uint32(p)) & np.uint32(3), g_want)
        np.testing.assert_array_equal((gb >> np.uint32(p)) & np.uint32(3), b_want)
