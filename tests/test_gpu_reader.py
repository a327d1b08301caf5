"""GPU .bed reader (adx_bed_to_packed / adx_packed_allele_sum /
adx_flip_packed) against the reference's own reader fixture
(demo_data.G.packed.expected, bit-exact) and against the host reader on
synthetic files that exercise the flip, the chromosome filter, several
transfer blocks and a partial last byte row."""
import numpy as np
import pytest

from adamixture_amd import reader
from tests.conftest import GOLDEN

pytestmark = pytest.mark.gpu


def test_gpu_reader_matches_reference_fixture(cuda, fixtures):
    Gp, N, M = reader.read_bed_gpu(GOLDEN / "demo_data.bed", cuda)
    want = fixtures["reader_G_packed"]
    assert N == want.shape[1] and (M + 3) // 4 == want.shape[0]
    got = Gp.data.cpu().numpy()
    np.testing.assert_array_equal(got[:, :N], want)
    assert np.all(got[:, N:] == 0xFF)


@pytest.mark.parametrize("M,N,flip,block", [(1001, 77, True, 4 * 20), (4096, 130, False, 1 << 29),
                                            (7, 5, True, 4)])
def test_gpu_reader_matches_host_reader(cuda, tmp_path, M, N, flip, block):
    rng = np.random.default_rng(M + N)
    p = rng.uniform(0.6, 0.95, size=(M, 1)) if flip else rng.uniform(0.05, 0.4, size=(M, 1))
    G = rng.binomial(2, p, size=(M, N)).astype(np.uint8)
    G[rng.random((M, N)) < 0.05] = 3
    base = tmp_path / "x"
    reader.write_bed(base, G)
    # chromosomes 1..3 in the .bim so that a filter drops rows
    lines = (tmp_path / "x.bim").read_text().splitlines()
    (tmp_path / "x.bim").write_text("\n".join(f"{1 + (j % 3)}" + line[1:] for j, line in enumerate(lines)) + "\n")
    for mode, spec in (("all", None), ("autosomes", [2])):
        want, Nw, Mw = reader.read_bed_packed(str(base) + ".bed", chrom_mode=mode, specific_chrom=spec)
        Gp, Ng, Mg = reader.read_bed_gpu(str(base) + ".bed", cuda, chrom_mode=mode, specific_chrom=spec,
                                         block_bytes=block * ((N + 3) // 4))
        assert (Ng, Mg) == (Nw, Mw)
        np.testing.assert_array_equal(Gp.data.cpu().numpy()[:, :N], want.numpy())
