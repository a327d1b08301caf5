"""VCF / BCF readers on the GPU (adx_vcf_gt / adx_bcf_gt) against the
reference's reader fixture (demo_data.G.packed.expected, bit-exact; the
reference's tests/test_reader.py reads every format to it) and against the
oracle's parsers on hand-built files with the edge cases (haploid,
multi-allelic, missing alleles, short lines, int16 GT, vector end), small
transfer blocks, the chromosome filter and its zero-SNP exit."""
import gzip

import numpy as np
import pytest

import oracle as O
from adamixture_amd import formats as F
from tests.conftest import GOLDEN
from tests.test_oracle import _bcf_stream

pytestmark = pytest.mark.gpu


def _packed_host(G):
    return G.data[:, :G.N].cpu().numpy()


def _expect_packed(G_unpacked: np.ndarray) -> np.ndarray:
    """The reference's packed result: 2-bit codes, flipped when the mean
    allele frequency is >= 0.5."""
    Gp = O.pack((np.asarray(G_unpacked) & 3).astype(np.uint8))
    if O.mean_packed(Gp, G_unpacked.shape[0]) >= 0.5:
        O.flip_packed(Gp, G_unpacked.shape[0])
    return Gp


@pytest.fixture(scope="module")
def demo_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("formats")
    (d / "demo_data.vcf").write_bytes(gzip.decompress((GOLDEN / "demo_data.vcf.gz").read_bytes()))
    return {"vcf": d / "demo_data.vcf", "vcf.gz": GOLDEN / "demo_data.vcf.gz", "bcf": GOLDEN / "demo_data.bcf",
            "bcf.gz": GOLDEN / "demo_data.bcf.gz"}


@pytest.mark.parametrize("kind", ["vcf", "vcf.gz", "bcf", "bcf.gz"])
@pytest.mark.parametrize("block", [F.BLOCK_BYTES, 1 << 15])
def test_demo_formats_match_reader_fixture(cuda, fixtures, demo_files, kind, block):
    print(f"case: {kind} block {block}", flush=True)
    reader = F.read_vcf_gpu if kind.startswith("vcf") else F.read_bcf_gpu
    G, N, M = reader(demo_files[kind], cuda, "autosomes", 22, block_bytes=block)
    assert (N, M) == (105, 8451)
    np.testing.assert_array_equal(_packed_host(G), fixtures["reader_G_packed"])
    assert np.all(G.data[:, N:].cpu().numpy() == 0xFF)


@pytest.mark.parametrize("kind", ["vcf.gz", "bcf"])
def test_read_data_filters_and_unpacked(cuda, fixtures, demo_files, kind):
    """test_reader.py's chromosome cases through read_data (unpacked output),
    and its zero-SNP exit."""
    chroms = np.array([line.split()[0] for line in (GOLDEN / "demo_data.bim").read_text().splitlines()])
    for spec, M_want in (([1], 682), (["1", "2"], 1454), (["chr1"], 682)):
        G, N, M = F.read_data(str(demo_files[kind]), packed=False, chrom_mode="autosomes", specific_chrom=spec,
                              device=cuda, verbose=False)
        assert (N, M) == (105, M_want)
        keep = np.isin(chroms, ["1"] if M_want == 682 else ["1", "2"])
        want = O.unpack(_expect_packed(fixtures["reader_G"][keep]), M_want)
        np.testing.assert_array_equal(G, want)
    G, N, M = F.read_data(str(demo_files[kind]), packed=True, chrom_mode="all", specific_chrom=[1], device=cuda,
                          verbose=False)
    assert M == 8451
    with pytest.raises(SystemExit) as exc:
        F.read_data(str(demo_files[kind]), chrom_mode="autosomes", specific_chrom=[9999], device=cuda, verbose=False)
    assert exc.value.code == 1


def test_vcf_edge_cases_match_oracle(cuda, tmp_path):
    head = b"##fileformat=VCFv4.3\n#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT\ta\tb\tc\td\te\n"
    body = (b"1\t1\t.\tA\tC\t.\t.\t.\tGT\t0/0\t0|1\t1/1\t./.\t0/.\n"
            b"1\t2\t.\tA\tC\t.\t.\t.\tGT:DP\t1/2:3\t0/2:1\t.|1\t-\t\n"
            b"2\t3\t.\tA\tC\t.\t.\t.\tGT\t0\t1\t2\t10/0\t1:9\n"
            b"2\t4\t.\tA\tC\t.\t.\t.\tGT\t1/0\t0/0\n"
            b"X\t5\t.\tA\tC\t.\t.\t.\tGT\t0/1\t1/1\t0/0\t0/0\t0/0")
    rng = np.random.default_rng(0)
    calls = np.array([b"0/0", b"0|1", b"1/1", b"./.", b"1", b"0/2", b"2/1"])
    big = b"".join(b"3\t%d\t.\tA\tC\t.\t.\t.\tGT\t" % k + b"\t".join(calls[rng.integers(0, 7, 5)]) + b"\n"
                   for k in range(37))
    path = tmp_path / "edge.vcf"
    path.write_bytes(head + big + body)
    want, N, M = O.vcf_to_unpacked(head + big + body)
    for block in (64, 1 << 20):
        print(f"case: vcf edge block {block}", flush=True)
        G, N2, M2 = F.read_vcf_gpu(path, cuda, "all", 22, block_bytes=block)
        assert (N2, M2) == (N, M)
        np.testing.assert_array_equal(_packed_host(G), _expect_packed(want))


def test_bcf_edge_cases_match_oracle(cuda, tmp_path):
    enc = lambda a, ph=0: ((a + 1) << 1) | ph                                   # noqa: E731
    rng = np.random.default_rng(1)
    recs = []
    for k in range(23):
        n = 2 if k % 5 else 1
        dt = np.int16 if k % 3 == 0 else np.int8
        v = np.array([[enc(int(a)) for a in rng.integers(0, 3, n)] for _ in range(70)], dtype=dt)
        v[rng.random(v.shape) < 0.1] = 0
        if n == 2:
            v[rng.random(70) < 0.05, 1] = -127 if dt == np.int8 else -32767
        recs.append((k % 2, v))
    data = _bcf_stream([f"s{i}" for i in range(70)], recs)
    path = tmp_path / "edge.bcf.gz"
    path.write_bytes(gzip.compress(data))
    want, N, M = O.bcf_to_unpacked(data)
    for block in (256, 1 << 20):
        print(f"case: bcf edge block {block}", flush=True)
        G, N2, M2 = F.read_bcf_gpu(path, cuda, "all", 22, block_bytes=block)
        assert (N2, M2) == (N, M) == (70, 23)
        np.testing.assert_array_equal(_packed_host(G), _expect_packed(want))


@pytest.mark.parametrize("sidecar", ["plain", "gz"])
def test_pgen_matches_reader_fixture(cuda, fixtures, tmp_path, sidecar):
    """Record types 0, 1, 2, 3, 4 of the demo .pgen decoded on the GPU; the
    .pvar sidecar plain or gzipped; the chromosome filter gathers rows."""
    (tmp_path / "d.pgen").write_bytes((GOLDEN / "demo_data.pgen").read_bytes())
    pvar = (GOLDEN / "demo_data.pvar").read_bytes()
    if sidecar == "gz":
        (tmp_path / "d.pvar.gz").write_bytes(gzip.compress(pvar))
    else:
        (tmp_path / "d.pvar").write_bytes(pvar)
    print(f"case: pgen sidecar {sidecar}", flush=True)
    G, N, M = F.read_pgen_gpu(tmp_path / "d.pgen", cuda, "autosomes", 22)
    assert (N, M) == (105, 8451)
    np.testing.assert_array_equal(_packed_host(G), fixtures["reader_G_packed"])
    chroms = np.array([line.split()[0] for line in (GOLDEN / "demo_data.bim").read_text().splitlines()])
    G, N, M = F.read_data(str(tmp_path / "d.pgen"), packed=True, chrom_mode="autosomes", specific_chrom=["1", "2"],
                          device=cuda, verbose=False)
    assert M == 1454
    np.testing.assert_array_equal(_packed_host(G), _expect_packed(fixtures["reader_G"][np.isin(chroms, ["1", "2"])]))
    with pytest.raises(SystemExit):
        F.read_data(str(tmp_path / "d.pgen"), specific_chrom=[9999], device=cuda, verbose=False)
