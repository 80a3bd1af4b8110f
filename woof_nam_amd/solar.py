"""Solar geometry of the radiation hook, on the library's restatement of the reference's astronomy.

* ``sol_update``: the sun's position and the solar constant for a date (sol_update_nmmb,
  phys_rrtm_sol_update.cu:16-160; the reference's radupdate calls it on radiation steps); solc0 is
  con_solr_old = 1366 W/m2 (isolar = 0, phys_rrtm_sol_init.cu:30).
* ``coszen``: the mean cosine of the zenith angle of each column over the SW interval, over its sunlit
  instants (coszmn_nmmb, phys_rrtm_coszmn.cu:18-62, NRADS instants); the radiation's cos z and CZMEAN.
* ``rdtemp``: the consumer of the heating rates every model step (RDTEMP, drv_rdtemp.cu).
The NMMB date array JDAT is (year, month, day, time-zone offset, hour, minute, second, ms); times are
JDAT plus the elapsed seconds.
"""
from __future__ import annotations

import ctypes as C
import datetime as _dt

import torch

from . import _lib, abi

SOLC0 = 1366.0


def current(jdat, seconds: float = 0.0) -> _dt.datetime:
    y, mo, d, _tz, h, mi, s = (int(x) for x in list(jdat)[:7])
    return _dt.datetime(y, mo, d, h, mi, s) + _dt.timedelta(seconds=float(seconds))


def sol_update(jdat, seconds: float, deltsw: float, deltim: float, solc0: float = SOLC0) -> abi.WrSol:
    """wr_sol_update at JDAT + seconds for a SW interval of deltsw seconds and model step deltim."""
    t = current(jdat, seconds)
    o = abi.WrSol()
    rc = _lib.load().wr_sol_update(t.year, t.month, t.day, t.hour, t.minute, t.second, float(deltsw), float(deltim),
                                   float(solc0), C.byref(o))
    _lib.check(rc, "wr_sol_update")
    return o


def solcon(jdat, seconds: float = 0.0, deltsw: float = 3600.0, deltim: float = 60.0) -> float:
    """The solar constant at the earth-sun distance of the date, W/m2."""
    return float(sol_update(jdat, seconds, deltsw, deltim).solcon)


def coszen(jdat, seconds: float, nrads: int, dt: float, glat: torch.Tensor, glon: torch.Tensor,
           stream=None) -> tuple[torch.Tensor, torch.Tensor]:
    """(coszen, coszdg) of the columns (glat, glon in radians, device tensors of one shape) over the SW
    interval of nrads steps of dt seconds that starts at JDAT + seconds: float64 tensors."""
    sol = sol_update(jdat, seconds, nrads * dt, dt)
    t = current(jdat, seconds)
    solhr = (t.hour + t.minute / 60.0) % 24.0
    lat = glat.reshape(-1).to(torch.float64).contiguous()
    lon = glon.reshape(-1).to(device=lat.device, dtype=torch.float64).contiguous()
    cz = torch.empty_like(lat)
    cd = torch.empty_like(lat)
    s = stream or torch.cuda.current_stream(lat.device)
    rc = _lib.load().wr_coszmn(lat.numel(), C.c_void_p(lat.data_ptr()), C.c_void_p(lon.data_ptr()), float(solhr),
                               float(nrads * dt / 3600.0), int(nrads), C.byref(sol), C.c_void_p(cz.data_ptr()),
                               C.c_void_p(cd.data_ptr()), C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_coszmn")
    return cz.reshape(glat.shape), cd.reshape(glat.shape)


def rdtemp(ntsd: int, dt: float, start: _dt.datetime, glat, glon, czmean, rswtt, rlwtt, t, czen, rows=None,
           cols=None, stream=None) -> None:
    """RDTEMP on float32 device state arrays: czen [jm][im] of step ntsd and T += (RSWTT czen / CZMEAN
    + RLWTT) dt on the inner columns (default: all but the outermost row and column of each side)."""
    lm, jm, im = (int(x) for x in t.shape)
    j_lo, j_hi = rows if rows is not None else (1, jm - 2)
    i_lo, i_hi = cols if cols is not None else (1, im - 2)
    julday = start.timetuple().tm_yday
    leap = int(start.year % 4 == 0)
    s = stream or torch.cuda.current_stream(t.device)
    ptr = lambda a: C.c_void_p(a.data_ptr())
    rc = _lib.load().wr_rdtemp(im, jm, lm, i_lo, i_hi, j_lo, j_hi, int(ntsd), float(dt), julday, start.year,
                               start.hour, leap, ptr(glat), ptr(glon), ptr(czmean), ptr(rswtt), ptr(rlwtt), ptr(t),
                               ptr(czen), C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_rdtemp")
