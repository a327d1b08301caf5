"""Copy the reference's own test data for the hot path into tests/golden/.

Run in the development container only (the reference tree is not present on
the GPU machine).  Inputs are data files that the reference's tests hold:

* adamixture/demo/data/demo_data.{bed,bim,fam}            (N=105, M=8451)
* adamixture/demo/data/demo_data.{vcf.gz,bcf,bcf.gz,pgen,pvar,psam} (the same genotypes;
  the reference's tests/test_reader.py reads every format to one fixture)
* adamixture/demo/outputs/reader/demo_data.G(.packed).expected
* adamixture/demo/outputs/{cpu,gpu}/{adamem,brqn}/demo_run.*.expected

The text fixtures (written with %.18e / %d, tests/helpers.py:13-18) are stored
losslessly as float64 / uint8 arrays in demo_fixtures.npz.
"""
from __future__ import annotations

import shutil
import sys
from pathlib import Path

import numpy as np

REF = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = Path(__file__).resolve().parent


def main() -> None:
    data = REF / "adamixture" / "demo" / "data"
    outs = REF / "adamixture" / "demo" / "outputs"
    for ext in ("bed", "bim", "fam", "vcf.gz", "bcf", "bcf.gz", "pgen", "pvar", "psam"):
        shutil.copyfile(data / f"demo_data.{ext}", OUT / f"demo_data.{ext}")
    arrays = {
        "reader_G": np.loadtxt(outs / "reader" / "demo_data.G.expected").astype(np.uint8),
        "reader_G_packed": np.loadtxt(outs / "reader" / "demo_data.G.packed.expected").astype(np.uint8),
    }
    for dev in ("cpu", "gpu"):
        for algo in ("adamem", "brqn"):
            d = outs / dev / algo
            if not d.is_dir():
                continue
            for f in sorted(d.glob("demo_run.*.expected")):
                key = f"{dev}_{algo}_" + f.name[len("demo_run."):-len(".expected")].replace(".", "_")
                arrays[key] = np.atleast_1d(np.loadtxt(f)).astype(np.float64)
    np.savez_compressed(OUT / "demo_fixtures.npz", **arrays)
    for k, v in arrays.items():
        print(k, v.shape, v.dtype)


if __name__ == "__main__":
    main()
