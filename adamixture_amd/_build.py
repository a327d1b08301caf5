"""In-tree build of libadx.so for gfx950 with hipcc."""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libadx.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-munsafe-fp-atomics",
    # packed FP32 VALU ops are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA results in VGPRs (no AGPR round trips in the elementwise epilogues)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
]


def sources() -> list[Path]:
    return sorted(CSRC.glob("*.hip"))


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(exe):
        raise RuntimeError("hipcc not found")
    return exe


def build_library(verbose: bool = False) -> Path:
    """Compile every csrc/*.hip into adamixture_amd/libadx.so (always rebuilds)."""
    tmp = OUT.with_name("libadx.so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{ROOT / 'include'}", "-o", str(tmp), *map(str, sources())]
    if verbose:
        print(" ".join(cmd))
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{res.stderr}")
    os.replace(tmp, OUT)
    return OUT
