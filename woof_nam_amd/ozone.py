"""Climatological ozone of the radiation call (the NAM carries no ozone tracer, ntoz = 0).

* ``Climatology``: the monthly zonal-mean ozone gas_init reads from global_o3clim.txt
  (phys_rrtm_gas_init.py:134-149; ``data/ozone_clim.json``), as the device tables getozn interpolates:
  o3r [12][17][18] mass mixing ratio = ppmv x 1.655e-6 in single precision (phys_rrtm_gas_init.cu:19-21)
  and the pressure levels pstr [17] in hPa.
* ``TimeFactors``: gas_update's climatology-ozone time factors (phys_rrtm_gas_update.py:78-100): the two
  months interpolated between and the weight facoz, which move on at the middle of each month.
The interpolation itself (latitude, month, Exner function of pressure) runs on the device in
wr_nmmb_prep (getozn, phys_rrtm_getozn.cu:57-110).
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import abi

DATA = Path(__file__).resolve().parent / "data" / "ozone_clim.json"
O3_SCALE = np.float32(1.655e-6)  # ppmv -> mass mixing ratio, gas_init's factor
_DAYS = (31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31)


def host_tables(path: Path = DATA) -> dict:
    """The climatology as host arrays: o3r float32 [12][loz][jmr] (mass mixing ratio), pstr float32
    [loz] (hPa), blte, dlte (degrees)."""
    d = json.loads(Path(path).read_text())
    ppmv = np.asarray(d["o3_ppmv_month_level_lat"], dtype=np.float32)
    return {"o3r": ppmv * O3_SCALE, "pstr": np.asarray(d["pstr_hpa"], dtype=np.float32),
            "blte": float(d["blte"]), "dlte": float(d["dlte"])}


class Climatology:
    """The climatology's device tables (kept for the process on one device)."""

    def __init__(self, device, path: Path = DATA):
        h = host_tables(path)
        self.device = torch.device(device)
        self.o3r = torch.from_numpy(h["o3r"]).to(self.device).contiguous()
        self.pstr = torch.from_numpy(h["pstr"]).to(self.device).contiguous()
        self.blte, self.dlte = h["blte"], h["dlte"]
        self.loz, self.jmr = int(self.o3r.shape[1]), int(self.o3r.shape[2])

    def struct(self, t: "TimeFactors") -> abi.WrOzone:
        f = C.POINTER(C.c_float)
        return abi.WrOzone(C.cast(C.c_void_p(self.o3r.data_ptr()), f), C.cast(C.c_void_p(self.pstr.data_ptr()), f),
                           self.loz, self.jmr, self.blte, self.dlte, t.k1oz, t.k2oz, float(t.facoz))


def _md(month: int) -> int:
    return _DAYS[month - 1]


@dataclass
class TimeFactors:
    """gas_update's saved ozone state: the months k1oz, k2oz (1..12), the mid-month days midm, midp
    that bracket the date, and the weight facoz of k2oz (single precision, as there)."""
    k1oz: int = 0
    k2oz: int = 0
    midm: int = 15
    midp: int = 45
    facoz: np.float32 = np.float32(0.0)
    started: bool = False

    def update(self, month: int, day: int, hour: int) -> "TimeFactors":
        """The factors for a date: the months move on at hour 0 of the middle day of a month (day
        days(month) / 2 + 1) and at the first call; facoz is the position of the day between the two
        months' middle days."""
        if not 1 <= month <= 12 or not 1 <= day <= 31:
            raise ValueError(f"ozone time factors: bad date month {month} day {day}")
        midmon = _md(month) // 2 + 1
        if not self.started or (day == midmon and hour == 0):
            if day < midmon:
                self.k1oz = (month + 10) % 12 + 1
                self.midm = _md(self.k1oz) // 2 + 1
                self.k2oz = month
                self.midp = _md(self.k1oz) + midmon
            else:
                self.k1oz = month
                self.midm = midmon
                self.k2oz = month % 12 + 1
                self.midp = _md(self.k2oz) // 2 + 1 + _md(self.k1oz)
            self.started = True
        ident = day + _md(self.k1oz) if day < midmon else day
        self.facoz = np.float32(ident - self.midm) / np.float32(self.midp - self.midm)
        return self
