"""Climatological aerosol of the radiation call (the NAM's iaer = 11: OPAC aerosol climatology).

Host preparation of the tables, once per process and once per month (single precision, the
reference's order of operations):
* ``spectrum``: aer_init's set_spectrum (phys_rrtm_aer_init.cu:105-128): the solar flux of each of the
  57,600 one-wavenumber bins (151 tabulated intervals) and the Planck emission at 273.15 K of the
  first 4,000 wavenumbers (the LW weights);
* ``band_means``: clim_aerinit's set_aercoef and optavg (phys_rrtm_clim_aerinit.cu:65-284): the 61
  aerosol spectral bands mapped onto the 14 SW and 16 LW radiation bands, weighted by those fluxes,
  and each species' band-mean extinction, scattering, single-scattering albedo (through the two-stream
  reflectance) and asymmetry, plus the stratospheric background extinction;
* ``month``: aer_update's fill of one month's global distribution (phys_rrtm_aer_update.cu:4-29):
  the components and mixing ratios of each 5-degree cell (the last one topped up so that they sum
  to 1), the profile type and the number densities of the mixing layer and the mineral layer.
The optics of every column and layer (aer_property with radclimaer) run on the device in
wr_aer_optics; ``Climatology.struct`` hands it the tables.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np
import torch

from . import _lib, abi

DATA = Path(__file__).resolve().parent / "data" / "aerosol_opac.npz"
NWVTOT, NWVTIR, NAERBND = 57600, 4000, 61
NCM1, NCM2, NRHLEV = 6, 4, 8
RHLEV = np.array([0.0, 0.5, 0.7, 0.8, 0.9, 0.95, 0.98, 0.99], np.float32)
# radiation band limits (wavenumbers, cm-1): radsw_param (bands 16..29) and radlw_param (bands 1..16)
WVNSW1 = (2600., 3251., 4001., 4651., 5151., 6151., 7701., 8051., 12851., 16001., 22651., 29001., 38001., 820.)
WVNSW2 = (3250., 4000., 4650., 5150., 6150., 7700., 8050., 12850., 16000., 22650., 29000., 38000., 50000., 2600.)
WVNLW1 = (10., 251., 501., 631., 701., 821., 981., 1081., 1181., 1391., 1481., 1801., 2081., 2251., 2381., 2601.)
WVNLW2 = (250., 500., 630., 700., 820., 980., 1080., 1180., 1390., 1480., 1800., 2080., 2250., 2380., 2600., 3000.)
F = np.float32


def load(path: Path = DATA) -> dict:
    with np.load(path, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def spectrum(d: dict) -> tuple[np.ndarray, np.ndarray]:
    """(solfwv [57600], eirfwv [4000]) float32: the solar flux per wavenumber bin and the Planck
    emission 2 pi h c^2 nu^3 / (exp(h c nu / (k T0)) - 1) at T0 = 273.15 K, nu = 100 x wavenumber."""
    solfwv = np.repeat(d["s0intv"].astype(F), d["nwvns0"])
    pi, c, h, k, t0 = F(3.1415926535897931), F(2.99792458e8), F(6.6260693e-34), F(1.3806505e-23), F(273.15)
    tmp1 = F(2.0) * pi * h * (c * c)
    tmp2 = h * c / (k * t0)
    tmp3 = F(100.0) * np.arange(1, NWVTIR + 1, dtype=F)
    with np.errstate(over="ignore"):
        eirfwv = (tmp1 * (tmp3 * tmp3 * tmp3)) / (np.exp(tmp2 * tmp3).astype(F) - F(1.0))
    return solfwv.astype(F), eirfwv.astype(F)


def _map(fwv: np.ndarray, iendwv: np.ndarray, w1: float, w2: float):
    """set_aercoef for one radiation band: the flux of each aerosol band inside it (summed in wavenumber
    order), the band's total, and the first and last aerosol bands (1-based)."""
    iw1, iw2 = int(round(w1)), int(round(w2))
    ii = 1
    while ii != NAERBND and iw1 > iendwv[ii - 1]:
        ii += 1
    waer = np.zeros(NAERBND, F)
    tot, part = F(0.0), F(0.0)
    n1 = ii
    for iw in range(iw1, iw2 + 1):
        f = fwv[iw - 1]
        tot = F(tot + f)
        part = F(part + f)
        if iw == iendwv[ii - 1]:
            waer[ii - 1] = part
            if ii < NAERBND:
                part, ii = F(0.0), ii + 1
    if iw2 != iendwv[ii - 1]:
        waer[ii - 1] = part
    return waer, tot, n1, ii


def _optavg(w, tot, n1, n2, ext, sca, ssa, asy):
    """optavg of one species in one band: flux-weighted extinction and scattering, the asymmetry
    weighted by scattering, and the single-scattering albedo recovered from the mean two-stream
    reflectance of a semi-infinite layer."""
    rnorm = F(1.0) / tot
    sumk = sums = sumok = sumokg = sumreft = F(0.0)
    for ni in range(n1, n2 + 1):
        wi, s, g, e, sc = w[ni - 1], ssa[ni - 1], asy[ni - 1], ext[ni - 1], sca[ni - 1]
        sp = F(np.sqrt(F((F(1.0) - s) / (F(1.0) - s * g))))
        reft = F((F(1.0) - sp) / (F(1.0) + sp))
        sumreft = F(sumreft + reft * wi)
        sumk = F(sumk + e * wi)
        sums = F(sums + sc * wi)
        sumok = F(sumok + s * wi * e)
        sumokg = F(sumokg + s * wi * e * g)
    refb = F(sumreft * rnorm)
    asyo = F(sumokg / (sumok + F(1.0e-10)))
    ssao = F(F(4.0) * refb / ((F(1.0) + refb) * (F(1.0) + refb) - asyo * ((F(1.0) - refb) * (F(1.0) - refb))))
    return F(sumk * rnorm), F(sums * rnorm), ssao, asyo


def band_means(d: dict, solfwv: np.ndarray, eirfwv: np.ndarray) -> dict:
    """The per-radiation-band optical tables: rows 0..13 SW bands 16..29, rows 14..29 LW bands 1..16."""
    iendwv = d["iendwv"]
    bands = [(solfwv, a, b) for a, b in zip(WVNSW1, WVNSW2)] + [(eirfwv, a, b) for a, b in zip(WVNLW1, WVNLW2)]
    nb = len(bands)
    out = {k: np.zeros((nb, NCM1), F) for k in ("extrhi", "scarhi", "ssarhi", "asyrhi")}
    out.update({k: np.zeros((nb, NCM2, NRHLEV), F) for k in ("extrhd", "scarhd", "ssarhd", "asyrhd")})
    out["extstra"] = np.zeros(nb, F)
    for ib, (fwv, a, b) in enumerate(bands):
        w, tot, n1, n2 = _map(fwv, iendwv, a, b)
        for nc in range(NCM1):
            v = _optavg(w, tot, n1, n2, d["rhidext0"][nc], d["rhidsca0"][nc], d["rhidssa0"][nc], d["rhidasy0"][nc])
            out["extrhi"][ib, nc], out["scarhi"][ib, nc], out["ssarhi"][ib, nc], out["asyrhi"][ib, nc] = v
        for nc in range(NCM2):
            for nh in range(NRHLEV):
                v = _optavg(w, tot, n1, n2, d["rhdpext0"][nc, nh], d["rhdpsca0"][nc, nh], d["rhdpssa0"][nc, nh],
                            d["rhdpasy0"][nc, nh])
                (out["extrhd"][ib, nc, nh], out["scarhd"][ib, nc, nh], out["ssarhd"][ib, nc, nh],
                 out["asyrhd"][ib, nc, nh]) = v
        s = F(0.0)
        for ni in range(n1, n2 + 1):
            s = F(s + d["straext0"][ni - 1] * w[ni - 1])
        out["extstra"][ib] = F(s * (F(1.0) / tot))
    return out


def month(d: dict, imon: int) -> dict:
    """aer_update's fill of month imon (1..12): idxcg, cmixg [37][72][5], denng [37][72][2], kprfg [37][72]."""
    m = imon - 1
    idx = d["idxc"][m].astype(np.int32)
    cmix = d["cmix"][m].astype(np.float64)
    kprf = d["kprf"][m].astype(np.int32)
    cm = cmix.copy()
    cm[..., 4] = 1.0 - cmix[..., 0] - cmix[..., 1] - cmix[..., 2] - cmix[..., 3]   # components add to 1
    denn = np.zeros(kprf.shape + (2,), np.float64)
    denn[..., 0] = d["denn"][m]
    denn[..., 1] = np.where(kprf >= 6, cmix[..., 4], 0.0)
    return {"idxcg": idx, "cmixg": cm.astype(F), "denng": denn.astype(F), "kprfg": kprf}


_STATIC: dict = {}


def static_tables(path: Path = DATA) -> dict:
    """The month-independent tables of wr_aer_tables (host float32, computed once per process)."""
    key = str(path)
    if key not in _STATIC:
        d = load(path)
        if d["kprf"].min() < 1 or d["kprf"].max() > 7 or d["idxc"].min() < 0 or d["idxc"].max() > 10:
            raise ValueError(f"{path}: profile types outside 1..7 or species outside 0..10")
        sol, eir = spectrum(d)
        _STATIC[key] = (d, {"haer": d["haer"], "prsref": d["prsref"], "sigref": (F(0.001) * d["prsref"]).astype(F),
                            "rhlev": RHLEV, **band_means(d, sol, eir)})
    return _STATIC[key][1]


def host_tables(imon: int, path: Path = DATA) -> dict:
    """Every wr_aer_tables member for month imon (1..12) as host arrays."""
    if not 1 <= imon <= 12:
        raise ValueError(f"aerosol climatology: month {imon} outside 1..12")
    st = static_tables(path)
    return {**st, **month(_STATIC[str(path)][0], imon)}


class Climatology:
    """The device tables of wr_aer_optics for one month (the monthly part replaced when the month
    changes, as aer_update does on the first step of a month)."""

    def __init__(self, device, path: Path = DATA):
        self.device = torch.device(device)
        self.path = path
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(self.device) for k, v in static_tables(path).items()}
        self.imon = None

    def set_month(self, imon: int) -> None:
        if imon != self.imon:
            tb = host_tables(imon, self.path)
            for k in ("idxcg", "cmixg", "denng", "kprfg"):
                self.t[k] = torch.from_numpy(np.ascontiguousarray(tb[k])).to(self.device)
            self.imon = imon

    def struct(self) -> abi.WrAerTables:
        if self.imon is None:
            raise RuntimeError("aerosol climatology: set_month first")
        p = lambda k, ct=C.c_float: C.cast(C.c_void_p(self.t[k].data_ptr()), C.POINTER(ct))
        return abi.WrAerTables(*[p(k) for k in abi.AER_TABLE_FLOAT], *[p(k, C.c_int32) for k in abi.AER_TABLE_INT])


def optics(clim: Climatology, cols: dict, lat: torch.Tensor, lon: torch.Tensor, slmsk: torch.Tensor,
           stream=None) -> dict:
    """wr_aer_optics on radiation columns (p_lev, t_lay, q_lay of the column call, float64 device tensors;
    lat, lon radians, slmsk 0 sea / 1 land / 2 sea ice, [ncol]): {"aer_sw", "aer_lw"} float32
    [3][bands][nlay+1][ncol] (optical depth, single-scattering albedo, asymmetry)."""
    t = cols["t_lay"]
    nlay, ncol = int(t.shape[0]), int(t.shape[1])
    dev = t.device
    for k, shp in (("p_lev", (nlay + 1, ncol)), ("t_lay", (nlay, ncol)), ("q_lay", (nlay, ncol))):
        a = cols[k]
        if tuple(a.shape) != shp or a.dtype != torch.float64 or not a.is_contiguous() or a.device != dev:
            raise ValueError(f"aerosol optics: {k} must be contiguous float64 {shp} on {dev}")
    out = {k: torch.empty(abi.aer_shape(k, ncol, nlay), dtype=torch.float32, device=dev) for k in abi.AER_FIELDS}
    ci = abi.WrColumns()
    ci.ncol, ci.nlay = ncol, nlay
    for k in ("p_lev", "t_lay", "q_lay"):
        setattr(ci, k, C.cast(C.c_void_p(cols[k].data_ptr()), C.POINTER(C.c_double)))
    la = lat.to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
    lo = lon.to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
    sl = slmsk.to(device=dev, dtype=torch.float32).reshape(-1).contiguous()
    if la.numel() != ncol or lo.numel() != ncol or sl.numel() != ncol:
        raise ValueError(f"aerosol optics: lat, lon and slmsk need {ncol} columns")
    s = stream or torch.cuda.current_stream(dev)
    if stream is not None:   # the inputs made here stay allocated until the launch has run
        for a in (la, lo, sl):
            a.record_stream(stream)
    tb = clim.struct()
    rc = _lib.load().wr_aer_optics(C.byref(tb), C.byref(ci), C.c_void_p(lo.data_ptr()), C.c_void_p(la.data_ptr()),
                                   C.c_void_p(sl.data_ptr()), C.c_void_p(out["aer_sw"].data_ptr()),
                                   C.c_void_p(out["aer_lw"].data_ptr()), ncol, C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_aer_optics")
    return out
