"""Build libwoofrad.so (the gfx950 HIP radiation library) in-tree.

``python -m woof_nam_amd.build`` compiles every translation unit under ``csrc/`` with hipcc for
gfx950 and links ``woof_nam_amd/libwoofrad.so``.  Flags:

* ``-ffp-contract=off``: no FMA contraction, so the FP64 path evaluates exactly the operation
  sequence written (table indices and McICA bits are reproducible bit for bit);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libwoofrad.so"
OBJ = PKG / "_build"
ARCH = os.environ.get("WR_OFFLOAD_ARCH", "gfx950")

UNITS = ["wr_host.hip", "wr_fp64.hip", "wr_mixed.hip", "wr_mlp.hip", "wr_surr.hip"]
CFLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libwoofrad")


# The mixed-precision and surrogate units contract FP32 multiply-adds and use the hardware FP32
# reciprocal/sqrt; the FP64 index arithmetic inside them keeps contraction off by pragma.  The waves
# per SIMD of every band kernel come from per-band tables in the sources (wr_kernels.h, wr_nam.h).
# WR_LW_GSPLIT=2 (two lanes per column in LW, half of the g-points each) halves the LW registers but
# measured 9 % slower: the per-layer column work then runs twice.
_FP32_UNIT = ["-ffp-contract=fast", "-fno-hip-fp32-correctly-rounded-divide-sqrt"]
UNIT_FLAGS = {
    "wr_mixed.hip": _FP32_UNIT,
    "wr_surr.hip": _FP32_UNIT,
}


def _compile(src: Path, obj: Path) -> None:
    extra = os.environ.get("WR_CFLAGS_EXTRA", "").split()      # tuning builds, e.g. -DWR_BAND_WAVES=4
    cmd = [hipcc(), *CFLAGS, *UNIT_FLAGS.get(src.name, []), *extra, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")


def build(verbose: bool = False, out: Path | None = None) -> Path:
    out = Path(out or os.environ.get("WR_LIB_OUT", LIB))
    objdir = OBJ if out == LIB else out.parent / (out.stem + "_obj")
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (Path(u).stem + ".o") for u in UNITS]
    with ThreadPoolExecutor(max_workers=min(len(UNITS), 8)) as ex:
        list(ex.map(lambda p: _compile(CSRC / p[0], p[1]), zip(UNITS, objs)))
    tmp = out.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
