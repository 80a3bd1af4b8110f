"""Convert the NAM's climatological ozone (global_o3clim.txt, the file gas_init reads when the model
carries no ozone tracer, ioznflg = 0) into this repository's data file.

Run once in a checkout that has the reference tree beside it:

    python woof_nam_amd/data/make_ozone.py /root/reference/namgpu/data/fix/global_o3clim.txt

File layout (radiation_gases_nmmb.f gas_init, formatted branch for 17 levels; the reference's reader
phys_rrtm_gas_init.py:134-149): 17 lines of pressure levels (f10.3, hPa, top first), then for each
of the 12 months and each of the 18 latitudes (-85 to 85 degrees by 10) one line (i2, i4, 10f6.2)
with the month, the latitude and levels 1..10, and one line (6x, 10f6.2) with levels 11..17, in ppmv.

Writes ``ozone_clim.json``: pressure levels (hPa), the ozone in ppmv as [month][level][latitude],
the latitude grid (first latitude and spacing) and the SHA-256 of the source file.
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
LOZ, JMR, NMONTH = 17, 18, 12


def field(s: str, d: int) -> float:
    """One Fortran Fw.d input field: blanks are zero, a field without a point has d implied decimals."""
    s = s.strip()
    if not s:
        return 0.0
    if "." in s or any(c in s for c in "eEdD"):
        return float(s.replace("d", "e").replace("D", "E"))
    return float(s) / 10 ** d


def main(src: str) -> None:
    p = Path(src)
    raw = p.read_bytes()
    lines = raw.decode().splitlines()
    pos = 0
    pstr = []
    for _ in range(LOZ):
        pstr.append(field(lines[pos][0:10], 3))
        pos += 1
    o3 = [[[0.0] * JMR for _ in range(LOZ)] for _ in range(NMONTH)]
    lats = []
    for imo in range(NMONTH):
        for j in range(JMR):
            hdr, body = lines[pos].ljust(66), lines[pos + 1].ljust(66)
            pos += 2
            mon, lat = int(hdr[0:2]), int(hdr[2:6])
            if mon != imo + 1:
                raise SystemExit(f"line {pos - 1}: month {mon}, expected {imo + 1}")
            if imo == 0:
                lats.append(lat)
            for m in range(10):
                o3[imo][m][j] = field(hdr[6 + 6 * m:12 + 6 * m], 2)
            for m in range(LOZ - 10):
                o3[imo][10 + m][j] = field(body[6 + 6 * m:12 + 6 * m], 2)
    blte, dlte = float(lats[0]), float(lats[1] - lats[0])
    if any(abs(lats[j] - (blte + j * dlte)) > 0 for j in range(JMR)):
        raise SystemExit(f"latitudes not on a regular grid: {lats}")
    doc = {"source": p.name, "sha256": hashlib.sha256(raw).hexdigest(),
           "what": "climatological ozone, ppmv, 12 months x 17 pressure levels x 18 latitudes (the NAM's "
                   "global_o3clim.txt read by gas_init when ioznflg = 0)",
           "pstr_hpa": pstr, "blte": blte, "dlte": dlte, "loz": LOZ, "jmr": JMR,
           "o3_ppmv_month_level_lat": o3}
    (HERE / "ozone_clim.json").write_text(json.dumps(doc) + "\n")
    print(f"wrote ozone_clim.json: {LOZ} levels {pstr[0]}..{pstr[-1]} hPa, latitudes {blte}..{lats[-1]}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/namgpu/data/fix/global_o3clim.txt")
