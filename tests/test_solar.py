"""Solar geometry (CPU): the library's sol_update against the oracle and against what the sun does on
known dates (equinoxes, solstices, perihelion and aphelion, the equation of time), and the oracle's
coszmn and rdtemp against closed forms."""
import ctypes as C
import datetime as dt
import math

import numpy as np
import pytest

from oracle import pyoracle
from woof_nam_amd import _lib, abi, solar


def sol(y, mo, d, h=12, mi=0):
    return solar.sol_update([y, mo, d, 0, h, mi, 0, 0], 0.0, 3600.0, 60.0)


def test_library_equals_oracle():
    for args in ((2025, 3, 20, 9, 1), (2024, 12, 31, 23, 59), (2026, 7, 4, 0, 0), (1999, 2, 28, 11, 30)):
        s = sol(*args)
        r = pyoracle.sol_update(*args, 0, 3600.0, 60.0, solar.SOLC0)
        for f in ("sollag", "sindec", "cosdec", "anginc", "solcon"):
            assert getattr(s, f) == pytest.approx(getattr(r, f), rel=1e-13, abs=1e-15), f
        assert s.nstp == r.nstp


def test_declination_at_equinoxes_and_solstices():
    obl = math.sin(math.radians(23.44))
    assert abs(sol(2025, 3, 20, 9, 1).sindec) < 0.004          # March equinox 09:01 UTC
    assert abs(sol(2025, 9, 22, 18, 19).sindec) < 0.004        # September equinox 18:19 UTC
    assert sol(2025, 6, 21, 2, 42).sindec == pytest.approx(obl, abs=0.002)
    assert sol(2025, 12, 21, 15, 3).sindec == pytest.approx(-obl, abs=0.002)
    s = sol(2025, 5, 1)
    assert s.sindec ** 2 + s.cosdec ** 2 == pytest.approx(1.0, abs=1e-14)


def test_solar_constant_follows_the_earth_sun_distance():
    # perihelion 4 January (r = 0.98333 AU), aphelion 3 July (r = 1.01664 AU)
    assert sol(2025, 1, 4).solcon / solar.SOLC0 == pytest.approx(1 / 0.98333 ** 2, abs=0.002)
    assert sol(2025, 7, 3).solcon / solar.SOLC0 == pytest.approx(1 / 1.01664 ** 2, abs=0.002)


def test_equation_of_time():
    # the solar lag is the equation of time in radians: about +16.4 min on 3 November, -14.2 min on
    # 11 February
    assert sol(2025, 11, 3).sollag * 720.0 / math.pi == pytest.approx(16.4, abs=0.6)
    assert sol(2025, 2, 11).sollag * 720.0 / math.pi == pytest.approx(-14.2, abs=0.6)


def test_interval_steps():
    s = solar.sol_update([2025, 3, 20, 0, 0, 0, 0, 0], 0.0, 3600.0, 25.0)   # 1 h SW interval of 25 s steps
    assert s.nstp == 13                                   # nn = clamp(nint(3600 / 25), 6, 12) = 12 per hour, + 1
    s = solar.sol_update([2025, 3, 20, 0, 0, 0, 0, 0], 0.0, 1200.0, 60.0)   # shorter: nswr = 20 + 1
    assert s.nstp == 21
    with pytest.raises(ValueError):
        solar.sol_update([2025, 13, 1, 0, 0, 0, 0, 0], 0.0, 3600.0, 60.0)


def test_coszmn_closed_forms():
    s = pyoracle.sol_update(2025, 3, 20, 12, 0, 0, 60.0, 60.0, solar.SOLC0)
    # one instant (nrads = 1): cos z = sin(lat) sin(dec) + cos(lat) cos(dec) cos(hour angle)
    lat = np.radians(np.array([0.0, 45.0, -30.0, 80.0]))
    lon = np.radians(np.array([0.0, 90.0, -100.0, 180.0]))
    cz, cd = pyoracle.coszmn(lat, lon, 12.0, 1.0 / 60.0, 1, s)
    ha = s.sollag + lon
    expect = np.sin(lat) * s.sindec + np.cos(lat) * s.cosdec * np.cos(ha)
    assert np.allclose(cd * 60.0, np.maximum(expect, 0.0), rtol=1e-14, atol=1e-15)
    assert cz[0] == pytest.approx(expect[0], rel=1e-14) and cz[0] > 0.99        # sub-solar point at noon
    assert cz[3] == 0.0                                                           # night side
    # over a 3 h interval of 12 instants: the mean of the sunlit instants
    cz, cd = pyoracle.coszmn(lat, lon, 6.0, 3.0, 12, s)
    t = np.radians(15.0 * (6.0 - 12.0 + np.arange(12) * 0.25))
    inst = np.sin(lat)[:, None] * s.sindec + np.cos(lat)[:, None] * s.cosdec * np.cos(t[None, :] + s.sollag + lon[:, None])
    up = inst > 1e-4
    for i in range(4):
        pos = np.maximum(inst[i], 0.0)
        assert cd[i] == pytest.approx(pos.sum() * 0.25, rel=1e-12, abs=1e-15)
        assert cz[i] == pytest.approx(pos.sum() / max(1, up[i].sum()) if up[i].any() else pos.sum(), rel=1e-12, abs=1e-15)


def test_rdtemp_heating_and_zenith():
    im, jm, lm = 6, 5, 4
    rng = np.random.default_rng(1)
    glat = np.radians(rng.uniform(10, 60, (jm, im))).astype(np.float32)
    glon = np.radians(rng.uniform(-130, -60, (jm, im))).astype(np.float32)
    czmean = rng.uniform(0.0, 0.8, (jm, im)).astype(np.float32)
    czmean[0, 0] = 0.0
    rsw = rng.uniform(0, 2e-4, (lm, jm, im)).astype(np.float32)
    rlw = rng.uniform(-2e-4, 0, (lm, jm, im)).astype(np.float32)
    t0 = rng.uniform(200, 300, (lm, jm, im)).astype(np.float32)
    t = t0.copy()
    czen = np.zeros((jm, im), np.float32)
    pyoracle.rdtemp(t, czen, glat, glon, czmean, rsw, rlw, ntsd=30, dt=60.0, start=dt.datetime(2025, 6, 21, 12),
                    box=(1, im - 2, 1, jm - 2))
    factr = np.where(czmean > 0, czen.astype(np.float64) / np.maximum(czmean, 1e-30), 0.0)
    expect = t0 + (rsw * factr[None] + rlw) * 60.0
    inner = (slice(None), slice(1, jm - 1), slice(1, im - 1))
    assert np.allclose(t[inner], expect[inner], rtol=0, atol=2e-5)
    assert np.array_equal(t[:, 0, :], t0[:, 0, :]) and np.array_equal(t[:, :, -1], t0[:, :, -1])
    assert (czen >= 0).all() and (czen <= 1).all() and czen.max() > 0.5          # summer noon over the US
