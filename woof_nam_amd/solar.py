"""Solar geometry for the NMMB radiation hook: cosine of the solar zenith angle.

Declination and equation of time from Spencer's Fourier series (J. W. Spencer, "Fourier series
representation of the position of the sun", Search 2(5), 1971); the hour angle from UTC, the
equation of time and longitude.  ``coszen`` averages max(cos z, 0) over the shortwave interval
(the role of coszmn in the reference's driver, phys_rrtm_driver_prep / radupdate); its exact
sub-sampling of the interval is the reference's own and is not reproduced (parity unpinned).
The NMMB date array JDAT is (year, month, day, time-zone offset, hour, minute, second, ms).
"""
from __future__ import annotations

import datetime as _dt
import math

import torch

S0 = 1368.22


def day_angle(jdat, seconds: float = 0.0):
    """(day angle in radians, UTC hours) at JDAT + seconds."""
    y, mo, d, _tz, h, mi, s = (int(x) for x in list(jdat)[:7])
    t = _dt.datetime(y, mo, d, h, mi, s) + _dt.timedelta(seconds=float(seconds))
    doy = t.timetuple().tm_yday
    ndays = 366 if (t.year % 4 == 0 and (t.year % 100 != 0 or t.year % 400 == 0)) else 365
    hours = t.hour + t.minute / 60.0 + t.second / 3600.0
    return 2.0 * math.pi * (doy - 1 + hours / 24.0) / ndays, hours


def declination(gam: float) -> float:
    return (0.006918 - 0.399912 * math.cos(gam) + 0.070257 * math.sin(gam) - 0.006758 * math.cos(2 * gam)
            + 0.000907 * math.sin(2 * gam) - 0.002697 * math.cos(3 * gam) + 0.00148 * math.sin(3 * gam))


def equation_of_time_minutes(gam: float) -> float:
    return 229.18 * (0.000075 + 0.001868 * math.cos(gam) - 0.032077 * math.sin(gam)
                     - 0.014615 * math.cos(2 * gam) - 0.040849 * math.sin(2 * gam))


def cosz(jdat, seconds, glat: torch.Tensor, glon: torch.Tensor) -> torch.Tensor:
    """cos(zenith) at JDAT + seconds; glat/glon in radians (east positive), any shape."""
    gam, hours = day_angle(jdat, seconds)
    dec = declination(gam)
    tst = hours * 60.0 + equation_of_time_minutes(gam)      # true solar time at Greenwich, minutes
    ha = torch.deg2rad((tst / 4.0 - 180.0) + torch.rad2deg(glon.double()))
    lat = glat.double()
    return torch.sin(lat) * math.sin(dec) + torch.cos(lat) * math.cos(dec) * torch.cos(ha)


def coszen(jdat, seconds, interval, glat, glon, nsub: int = 12) -> torch.Tensor:
    """Mean of max(cos z, 0) over [seconds, seconds + interval) (nsub sub-samples, midpoints)."""
    acc = torch.zeros_like(glat, dtype=torch.float64)
    for j in range(nsub):
        acc += torch.clamp(cosz(jdat, seconds + (j + 0.5) * interval / nsub, glat, glon), min=0.0)
    return acc / nsub


def solcon(jdat, seconds=0.0) -> float:
    """Solar constant at the Earth-Sun distance of the date (Spencer's eccentricity factor)."""
    gam, _ = day_angle(jdat, seconds)
    e0 = (1.000110 + 0.034221 * math.cos(gam) + 0.001280 * math.sin(gam) + 0.000719 * math.cos(2 * gam)
          + 0.000077 * math.sin(2 * gam))
    return S0 * e0
