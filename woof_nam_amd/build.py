"""Build libwoofrad.so (the gfx950 HIP radiation library) in-tree.

``python -m woof_nam_amd.build`` compiles every translation unit under ``csrc/`` with hipcc for
gfx950 and links ``woof_nam_amd/libwoofrad.so``.  Flags:

* ``-ffp-contract=off``: no FMA contraction, so the FP64 path evaluates exactly the operation
  sequence written (table indices and McICA bits are reproducible bit for bit);
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: no packed FP32 instructions.

Every kernel's private segment (register spills and private arrays, "ScratchSize") is read from the
compiler's resource-usage remarks and the build fails when one exceeds ``SCRATCH_MAX`` bytes per lane:
the runtime backs the private segment of every queue that runs a kernel, for every lane in flight,
and large per-lane arrays have exhausted it (an HSA out-of-resources queue abort).  The per-kernel
table is written to ``_build/resource_usage.txt``.

The build also fails when a kernel that the host pass registers (by its host-side mangled name) has no
device symbol of that name: a kernel parameter whose type differs between the passes (a pointer
qualified with a device address space, declared only in the device pass) mangles differently, and
launching it aborts the process because the runtime finds no such kernel in the code object.
"""
from __future__ import annotations

import os
import re
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

UNITS = ["wr_host.hip", "wr_fp64.hip", "wr_fp64_ext.hip", "wr_mixed.hip", "wr_mixed_ext.hip", "wr_mlp.hip", "wr_surr.hip",
         "wr_nmmb.hip", "wr_solar.hip", "wr_aer.hip", "wr_serve.hip"]
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
    "wr_mixed_ext.hip": _FP32_UNIT,
    "wr_surr.hip": _FP32_UNIT,
}


# Largest private segment a kernel may use, bytes per lane (the two-wave band kernels spill up to
# about 1.1 KB; the private arrays behind the queue aborts were 13-45 KB)
SCRATCH_MAX = int(os.environ.get("WR_SCRATCH_MAX", "1536"))
_REMARK = re.compile(r"remark:\s+(Function Name|VGPRs|AGPRs|ScratchSize \[bytes/lane\]|Occupancy \[waves/SIMD\]):\s*(\S+)")


def resource_usage(stderr: str) -> list:
    """[(kernel, vgprs, agprs, scratch bytes/lane, waves/SIMD)] from -Rpass-analysis=kernel-resource-usage."""
    rows, cur = [], None
    for m in _REMARK.finditer(stderr):
        key, val = m.group(1), m.group(2)
        if key == "Function Name":
            cur = {"name": val}
            rows.append(cur)
        elif cur is not None:
            cur[key.split()[0]] = int(val)
    return [(r["name"], r.get("VGPRs", 0), r.get("AGPRs", 0), r.get("ScratchSize", 0), r.get("Occupancy", 0))
            for r in rows]


_REGISTER = re.compile(r"@__hipRegisterFunction\(ptr [^,]+, ptr (?:nonnull )?@\"?([\w.$]+)\"?,")


def registered_kernels(src: Path) -> list:
    """Kernel names the host pass of a unit registers with the runtime (its host-side mangling)."""
    extra = os.environ.get("WR_CFLAGS_EXTRA", "").split()
    cmd = [hipcc(), *CFLAGS, *UNIT_FLAGS.get(src.name, []), *extra, "--cuda-host-only", "-emit-llvm", "-S",
           "-o", "-", str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc (host pass) failed for {src.name}:\n{r.stderr}")
    return _REGISTER.findall(r.stdout)


def _compile(src: Path, obj: Path) -> list:
    extra = os.environ.get("WR_CFLAGS_EXTRA", "").split()      # tuning builds, e.g. -DWR_BAND_WAVES=4
    cmd = [hipcc(), *CFLAGS, *UNIT_FLAGS.get(src.name, []), *extra, "-Rpass-analysis=kernel-resource-usage",
           "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr}")
    rows = [(src.name, *row) for row in resource_usage(r.stderr)]
    device = {row[1] for row in rows}
    missing = [k for k in registered_kernels(src) if k not in device]
    if missing:
        raise RuntimeError(f"{src.name}: kernels registered by the host pass with no device symbol of the same "
                           f"name (a parameter type that differs between the passes): {missing}")
    return rows


def build(verbose: bool = False, out: Path | None = None) -> Path:
    out = Path(out or os.environ.get("WR_LIB_OUT", LIB))
    objdir = OBJ if out == LIB else out.parent / (out.stem + "_obj")
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (Path(u).stem + ".o") for u in UNITS]
    with ThreadPoolExecutor(max_workers=min(len(UNITS), 8)) as ex:
        usage = [row for rows in ex.map(lambda p: _compile(CSRC / p[0], p[1]), zip(UNITS, objs)) for row in rows]
    with open(objdir / "resource_usage.txt", "w") as f:
        f.write("unit kernel vgprs agprs scratch_bytes_per_lane waves_per_simd\n")
        for row in usage:
            f.write(" ".join(map(str, row)) + "\n")
    over = [row for row in usage if row[4] > SCRATCH_MAX]
    if over:
        raise RuntimeError("private segment above %d bytes/lane: %s" % (
            SCRATCH_MAX, "; ".join(f"{u}:{k} {sc} B" for u, k, _, _, sc, _ in over)))
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
