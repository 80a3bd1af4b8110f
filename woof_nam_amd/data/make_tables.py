"""Convert the RRTMG k-distribution, Planck, cloud-optics and solar-source tables that the
reference ships as numeric ``.npz`` data into this repository's own binary table file.

Run once in a checkout that has the reference tree beside it:

    python woof_nam_amd/data/make_tables.py /root/reference/namgpu/data

It writes ``rrtmg_tables.bin`` (the file the HIP library and the CPU oracle load) and
``rrtmg_tables.json`` (human-readable manifest: shape, dtype, SHA-256 of every table, and the
reference file each one came from).

Only data is read: every file goes through ``numpy.load(..., allow_pickle=False)``.

Canonical layouts (all k-tables are g-point fastest, so one (pressure, temperature, eta) row of
a band is a contiguous vector of that band's g-points):
  * LW/SW ``absa``/``absb``/``selfref``/``forref``  [index][g]
  * eta-interpolated minor tables ``ka_m*``/``kb_m*``  [minor-T index][eta][g]
  * 1-D minor tables ``ka_m*``/``kb_m*``               [minor-T index][g]
  * ``fracrefa``/``fracrefb``                          [eta][g]   (eta = 1 when not interpolated)
Values are the float32 values the reference uploads to the device (its kernels are float32);
band-5 LW data, kept as float64 in the reference tree, are rounded to float32 here exactly as
the reference's loader rounds them (phys_rrtm_lw_taugb05.py:86-94).

Binary format (little endian):
  8 bytes  magic b"WRTBL001"
  u32      number of entries
  entries: 48-byte name, u32 dtype (0 = float32, 1 = int32), u32 ndim, 4 x u32 dims,
           u64 byte offset from file start, u64 element count
  data (each array 64-byte aligned)
"""
from __future__ import annotations

import hashlib
import json
import struct
import sys
from pathlib import Path

import numpy as np

MAGIC = b"WRTBL001"
NAME_LEN = 48

LW_NG = (10, 12, 16, 14, 16, 8, 12, 8, 12, 6, 8, 8, 4, 2, 2, 2)
SW_NG = {16: 6, 17: 12, 18: 8, 19: 8, 20: 10, 21: 10, 22: 2, 23: 10, 24: 8, 25: 6, 26: 6,
         27: 8, 28: 6, 29: 12}

# LW band widths (cm-1), radlw_param.f:133 as held by phys_rrtm_lw_setcoef.py:64-65.
LW_DELWAVE = [340., 150., 130., 70., 120., 160., 100., 100., 210., 90., 320., 280., 170., 130.,
              220., 650.]


def _npz(path: Path) -> dict:
    with np.load(path, allow_pickle=False) as z:
        return {k: np.array(z[k]) for k in z.files}


class Writer:
    def __init__(self):
        self.items: list[tuple[str, np.ndarray, str]] = []

    def add(self, name: str, arr, src: str, dtype=np.float32):
        a = np.ascontiguousarray(np.asarray(arr, dtype=dtype))
        if a.ndim == 0:
            a = a.reshape(1)
        assert a.ndim <= 4 and len(name) < NAME_LEN, name
        assert np.all(np.isfinite(a)) if a.dtype == np.float32 else True, name
        self.items.append((name, a, src))

    def write(self, bin_path: Path, json_path: Path):
        header = 8 + 4 + len(self.items) * (NAME_LEN + 4 + 4 + 16 + 8 + 8)
        off = (header + 63) // 64 * 64
        entries, blobs, manifest = [], [], []
        for name, a, src in self.items:
            dims = list(a.shape) + [1] * (4 - a.ndim)
            dt = 0 if a.dtype == np.float32 else 1
            entries.append(struct.pack("<48sII4IQQ", name.encode(), dt, a.ndim, *dims, off, a.size))
            raw = a.tobytes()
            blobs.append((off, raw))
            manifest.append({"name": name, "dtype": "float32" if dt == 0 else "int32",
                             "shape": list(a.shape), "offset": off,
                             "sha256": hashlib.sha256(raw).hexdigest(), "source": src})
            off = (off + len(raw) + 63) // 64 * 64
        buf = bytearray(off)
        buf[0:8] = MAGIC
        buf[8:12] = struct.pack("<I", len(self.items))
        p = 12
        for e in entries:
            buf[p:p + len(e)] = e
            p += len(e)
        for o, raw in blobs:
            buf[o:o + len(raw)] = raw
        bin_path.write_bytes(bytes(buf))
        json_path.write_text(json.dumps({
            "format": "WRTBL001", "file": bin_path.name,
            "sha256": hashlib.sha256(bytes(buf)).hexdigest(),
            "notice": ("RRTMG tables: Copyright (c) 2002-2020 Atmospheric and Environmental Research, "
                       "Inc. (AER), BSD-3-Clause. NCEP's NMMB modifications (NAM nam.v4.2.11) are US "
                       "Government work. Values converted from the numeric .npz data of "
                       "recastsystems/woof-nam namgpu/data (Apache-2.0 AND BSD-3-Clause)."),
            "tables": manifest}, indent=1))


def build(data: Path, out_dir: Path):
    w = Writer()
    lw = data / "rrtm_lw"
    sw = data / "rrtm_sw"

    # ---- LW general tables -------------------------------------------------------------
    ref = _npz(lw / "ref.npz")
    w.add("lw/preflog", ref["preflog"], "rrtm_lw/ref.npz")
    w.add("lw/tref", ref["tref"], "rrtm_lw/ref.npz")
    w.add("lw/chi_mls", ref["chi_mls"], "rrtm_lw/ref.npz (chi_mls(gas,ref) at [ref][gas])")
    w.add("lw/totplnk", _npz(data / "rrtm_lw_avplank.npz")["totplnk"], "rrtm_lw_avplank.npz [band][T]")
    w.add("lw/delwave", LW_DELWAVE, "phys_rrtm_lw_setcoef.py:64-65 (radlw_param.f:133)")
    cp = _npz(data / "rrtm_lw_cldprlw.npz")
    for k in ("absliq1", "absice1", "absice2", "absice3"):
        w.add("lw/" + k, cp[k], "rrtm_lw_cldprlw.npz")
    w.add("lw/ipat", cp["ipat"], "rrtm_lw_cldprlw.npz", dtype=np.int32)

    # ---- LW band k-tables ([ind][g] etc. already in the npz except band 5) ---------------
    for b in range(1, 17):
        ng = LW_NG[b - 1]
        if b == 5:
            z = _npz(data / "rrtm_lw_kgb05.npz")
            src = "rrtm_lw_kgb05.npz (float64 [g][..] -> float32 [..][g])"
            tabs = {"absa": z["absa"].T, "absb": z["absb"].T, "selfref": z["selfref"].T,
                    "forref": z["forref"].T, "fracrefa": z["fracrefa"].T, "fracrefb": z["fracrefb"].T,
                    "ka_mo3": z["ka_mo3"].transpose(2, 1, 0), "ccl4": z["ccl4"]}
        else:
            z = _npz(lw / f"kgb{b:02d}.npz")
            src = f"rrtm_lw/kgb{b:02d}.npz"
            tabs = {k: v for k, v in z.items()
                    if k not in ("source", "chi_mls", "chi_mls_co2") and not k.startswith("nmmb_")}
        for k, v in tabs.items():
            v = np.asarray(v, dtype=np.float64).astype(np.float32)
            if k in ("fracrefa", "fracrefb") and v.ndim == 1:
                v = v.reshape(1, ng)
            assert v.shape[-1] == ng, (b, k, v.shape)
            w.add(f"lw/b{b:02d}/{k}", v, src)

    # ---- SW general tables -----------------------------------------------------------------
    sref = _npz(sw / "ref.npz")
    w.add("sw/preflog", sref["preflog"], "rrtm_sw/ref.npz")
    w.add("sw/tref", sref["tref"], "rrtm_sw/ref.npz")
    sf = _npz(sw / "sflux.npz")
    for k in ("sfluxref01", "sfluxref02", "sfluxref03", "strrat", "specwt", "scalekur"):
        w.add("sw/" + k, sf[k], "rrtm_sw/sflux.npz [slot][solar-cycle][g]")
    for k in ("layreffr", "ix1", "ix2", "ibx"):
        w.add("sw/" + k, sf[k], "rrtm_sw/sflux.npz", dtype=np.int32)
    cps = _npz(data / "rrtm_sw_cldprtb.npz")
    for k in ("extliq1", "ssaliq1", "asyliq1", "extice2", "ssaice2", "asyice2",
              "extice3", "ssaice3", "asyice3"):
        w.add("sw/" + k, cps[k], "rrtm_sw_cldprtb.npz [band][index]")

    # SW band tables. g_first lists the npz arrays stored [g][ind] (transposed here); the layout of
    # every array follows the reference kernel's own indexing of it (phys_rrtm_sw_taumolNN.cu).
    g_first = {16: ("absa", "absb", "selfref", "forref"), 20: ("absa", "absb", "selfref", "forref"),
               24: ("absa", "absb", "selfref", "forref"), 25: ("absa",), 28: ("absa", "absb")}
    for b in range(16, 30):
        ng = SW_NG[b]
        z = _npz(sw / f"kgb{b}.npz")
        for k, v in z.items():
            if k == "source":
                continue
            v = np.asarray(v, dtype=np.float32)
            if k in g_first.get(b, ()):
                assert v.shape[0] == ng, (b, k, v.shape)
                v = v.T
            if v.ndim >= 1 and v.size > 1:
                assert v.shape[-1] == ng or k == "rayla", (b, k, v.shape)
            w.add(f"sw/b{b}/{k}", v, f"rrtm_sw/kgb{b}.npz")

    out_dir.mkdir(parents=True, exist_ok=True)
    w.write(out_dir / "rrtmg_tables.bin", out_dir / "rrtmg_tables.json")
    return w


if __name__ == "__main__":
    src = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/namgpu/data")
    wr = build(src, Path(__file__).resolve().parent)
    tot = sum(a.nbytes for _, a, _ in wr.items)
    print(f"{len(wr.items)} tables, {tot / 1024:.1f} KiB of data")
