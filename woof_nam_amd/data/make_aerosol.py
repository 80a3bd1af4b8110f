"""Convert the NAM's climatological aerosol data (OPAC, the iaer = 11 scheme) into this repository's
data file.

Run once in a checkout that has the reference tree beside it:

    python woof_nam_amd/data/make_aerosol.py /root/reference/namgpu/parm/nam_climaeropac_global.txt \\
        /root/reference/namgpu/kernels/phys_rrtm_aer_init.cu

Sources (only numbers are read):
  * nam_climaeropac_global.txt: 12 monthly global distributions on the 5-degree grid (72 longitudes
    from 0, 37 latitudes from 90 N; per cell up to 5 (component, mixing ratio) pairs, the profile type
    and the number density of the mixing layer; record layout of phys_rrtm_aer_update.py:109-136), then
    the optical tables (phys_rrtm_clim_aerinit.py:103-133): the 61 aerosol spectral bands' ending
    wavenumbers, the profile heights haer and reference pressures prsref (7 profiles x 5 domains), the
    extinction / scattering / single-scattering albedo / asymmetry of the 6 humidity-independent and
    the 4 humidity-dependent species (8 relative humidities) per spectral band, and the stratospheric
    background extinction;
  * phys_rrtm_aer_init.cu: the solar spectrum's DATA statements (NWVNS0, S0INTV: 151 intervals of the
    57,600 one-wavenumber bins and their flux).

Writes ``aerosol_opac.npz`` (numpy, no pickles) with the SHA-256 of both sources.
"""
from __future__ import annotations

import hashlib
import re
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
IMXAE, JMXAE, NXC = 72, 37, 5
NAERBND, NRHLEV, NCM1, NCM2, NDM, NAE = 61, 8, 6, 4, 5, 7
MONTH_LINES = 2 + IMXAE * JMXAE


def ffloat(s: str) -> float:
    """A Fortran F / E input field: blank is zero, an exponent may come without its letter."""
    s = s.strip()
    if not s:
        return 0.0
    try:
        return float(s)
    except ValueError:
        k = max(s.rfind("+"), s.rfind("-"))
        return float(s[:k] + "E" + s[k:])


def block(lines, pos, n, per_rec, width, conv):
    """A table: one header record, then records of per_rec fields of the given width."""
    pos += 1
    vals = []
    while len(vals) < n:
        line = lines[pos]
        pos += 1
        for j in range(min(per_rec, n - len(vals))):
            vals.append(conv(line[j * width:(j + 1) * width]))
    return pos, vals


def main(aer_file: str, init_cu: str) -> None:
    raw = Path(aer_file).read_bytes()
    lines = raw.decode().splitlines()
    idx = np.zeros((12, JMXAE, IMXAE, NXC), np.int8)
    cmix = np.zeros((12, JMXAE, IMXAE, NXC), np.float64)
    kprf = np.zeros((12, JMXAE, IMXAE), np.int8)
    denn = np.zeros((12, JMXAE, IMXAE), np.float64)
    for m in range(12):
        base = m * MONTH_LINES + 2
        for j in range(JMXAE):
            for i in range(IMXAE):
                line = lines[base + j * IMXAE + i]
                for k in range(NXC):
                    b = 13 * k
                    idx[m, j, i, k] = int(line[b:b + 2])
                    cmix[m, j, i, k] = float(line[b + 2:b + 13])
                kprf[m, j, i] = int(line[65:67])
                denn[m, j, i] = float(line[67:75])
    pos = 12 * MONTH_LINES
    if not lines[pos].startswith("ENDING WAVENUMBER"):
        raise SystemExit(f"optical tables not at line {pos + 1}")
    pos, iendwv = block(lines, pos, NAERBND, 13, 6, int)
    pos, haer = block(lines, pos, NAE * NDM, 20, 4, ffloat)
    pos, prsref = block(lines, pos, NAE * NDM, 10, 7, ffloat)
    rhid = []
    for _ in range(4):
        pos, v = block(lines, pos, NAERBND * NCM1, 8, 10, ffloat)
        rhid.append(np.array(v, np.float32).reshape(NCM1, NAERBND))
    rhdp = []
    for _ in range(4):
        pos, v = block(lines, pos, NAERBND * NRHLEV * NCM2, 8, 10, ffloat)
        rhdp.append(np.array(v, np.float32).reshape(NCM2, NRHLEV, NAERBND))
    if not lines[pos].startswith("STRATOSPERIC"):
        raise SystemExit(f"stratospheric table not at line {pos + 1}")
    pos, straext = block(lines, pos, NAERBND, 8, 10, ffloat)
    cu = Path(init_cu).read_text()

    def cu_array(name):
        m = re.search(r"%s\[\w+\]\s*=\s*\{([^}]*)\}" % name, cu)
        return [float(v.rstrip("f")) for v in re.findall(r"[-+]?\d*\.?\d+(?:[eE][-+]?\d+)?f?", m.group(1))]

    nwvns0 = np.array(cu_array("AER_NWVNS0"), np.int32)
    s0intv = np.array(cu_array("AER_S0INTV"), np.float32)
    if nwvns0.size != 151 or s0intv.size != 151 or nwvns0.sum() != 57600:
        raise SystemExit("solar spectrum tables malformed")
    np.savez_compressed(
        HERE / "aerosol_opac.npz",
        idxc=idx, cmix=cmix, kprf=kprf, denn=denn,
        iendwv=np.array(iendwv, np.int32), haer=np.array(haer, np.float32).reshape(NAE, NDM),
        prsref=np.array(prsref, np.float32).reshape(NAE, NDM),
        rhidext0=rhid[0], rhidsca0=rhid[1], rhidssa0=rhid[2], rhidasy0=rhid[3],
        rhdpext0=rhdp[0], rhdpsca0=rhdp[1], rhdpssa0=rhdp[2], rhdpasy0=rhdp[3],
        straext0=np.array(straext, np.float32), nwvns0=nwvns0, s0intv=s0intv,
        sha256_aer=np.array(hashlib.sha256(raw).hexdigest()),
        sha256_init=np.array(hashlib.sha256(cu.encode()).hexdigest()))
    print("wrote aerosol_opac.npz")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/namgpu/parm/nam_climaeropac_global.txt",
         sys.argv[2] if len(sys.argv) > 2 else "/root/reference/namgpu/kernels/phys_rrtm_aer_init.cu")
