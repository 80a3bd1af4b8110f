"""Seeded synthetic radiation columns (the RRTMG benchmark workloads).

``columns(ncol, seed)`` -- independent columns (4,096 x 60 layers is the CPU-reference workload):
  * surface pressure U(950, 1040) hPa; 60 NAM hybrid sigma-pressure layers to a 2 hPa top
    (PT = 200 Pa, PTSGM = 30000 Pa, the 61 coord_levs of namgpu/parm/nam_namelist.nps:54-64,
    split into pressure/sigma parts as namgpu/grid.py:132-143 does);
  * temperature: 6.5 K/km lapse from 288.15 K to an 11 km tropopause (in pressure, the ISA form
    T = T0 (p/ps)^(R Gamma/g)), isothermal 216.65 K above, plus N(0, 3 K) noise per layer;
  * water vapour from relative humidity U(0.05, 0.95) (Bolton saturation), mixing ratio capped at
    3e-6 kg/kg above 300 hPa;
  * ozone: analytic mid-latitude profile, 8 ppmv peak at 10 hPa (Gaussian in ln p, width 1.1) on a
    30 ppbv background;
  * clouds: 0-3 decks per column of 1-6 contiguous layers below 150 hPa, each with one fraction
    U(0.05, 1) (maximum overlap inside a deck, random between decks); liquid U(0, 0.5) g/kg where
    T > 253 K and ice U(0, 0.2) g/kg where T < 273 K inside cloud;
  * surface albedo U(0.05, 0.8), emissivity U(0.9, 1.0), skin T = lowest-layer T + U(-5, 5) K,
    cos zenith U(-1, 1) (about half the columns are night).

``grid(nx, ny, seed)`` -- the 12 km NAM domain (954 x 835): the same construction with the
temperature and RH perturbations drawn as Gaussian random fields of 300 km length scale (FFT on the
12 km grid) and cos zenith from a 12 UTC equinox sun over 15-70 N, 150-50 W.

Everything is float64 numpy, column fastest ([k][column]), index 0 at the model top.
"""
from __future__ import annotations

import numpy as np

PT = 200.0
PTSGM = 30000.0
P0 = 101325.0
COORD_LEVS = np.array([
    0.0, 0.010994, 0.022004, 0.033054, 0.044144, 0.055294, 0.066594, 0.078094, 0.089794, 0.101894,
    0.114294, 0.127094, 0.140494, 0.154394, 0.169294, 0.185494, 0.203594, 0.223694, 0.245894,
    0.270694, 0.297694, 0.326694, 0.356994, 0.388094, 0.419694, 0.451894, 0.484394, 0.517114,
    0.549714, 0.582114, 0.614214, 0.645814, 0.676614, 0.706714, 0.735314, 0.762114, 0.786714,
    0.809114, 0.829314, 0.847114, 0.862914, 0.876814, 0.888811, 0.899507, 0.908404, 0.916402,
    0.923599, 0.930397, 0.936895, 0.943093, 0.948991, 0.954689, 0.960187, 0.965486, 0.970684,
    0.975782, 0.980781, 0.985679, 0.990479, 0.995253, 1.0])
NLAY = 60
RD, G, LAPSE = 287.05, 9.80665, 0.0065
AMO3, AMD = 47.9982, 28.9644

FIELDS = ("p_lev", "t_lay", "q_lay", "o3_lay", "cldf", "qc", "qi", "t_sfc", "emis", "alb", "cosz")


def _sigma():
    pref = PT + COORD_LEVS * (P0 - PT)
    pdtop = PTSGM - PT
    sg1 = np.where(pref <= PTSGM, (pref - PT) / pdtop, 1.0)
    sg2 = np.where(pref <= PTSGM, 0.0, (pref - PTSGM) / (P0 - PTSGM))
    return sg1, sg2, pdtop


def _qsat(t, p):
    es = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))
    es = np.minimum(es, 0.5 * p)
    return 0.622 * es / (p - 0.378 * es)


def _profiles(ps, tnoise, rh, rng):
    """Column profiles from surface pressure [n], T noise [60, n], RH [60, n]."""
    n = ps.shape[0]
    sg1, sg2, pdtop = _sigma()
    pd = ps - PT - pdtop
    p_lev = sg1[:, None] * pdtop + sg2[:, None] * pd[None, :] + PT
    p_lay = 0.5 * (p_lev[1:] + p_lev[:-1])
    expo = RD * LAPSE / G
    t = np.maximum(216.65, 288.15 * (p_lay / ps[None, :]) ** expo) + tnoise
    q = rh * _qsat(t, p_lay)
    q = np.where(p_lay < 30000.0, np.minimum(q, 3.0e-6), q)
    lnp = np.log(p_lay / 1000.0)
    o3vmr = 30e-9 + 8.0e-6 * np.exp(-0.5 * (lnp / 1.1) ** 2)
    o3 = o3vmr * (AMO3 / AMD)
    # cloud decks
    cldf = np.zeros((NLAY, n))
    ndeck = rng.integers(0, 4, size=n)
    ktop_min = int(np.argmax(p_lay[:, 0] > 15000.0))
    for d in range(3):
        has = ndeck > d
        top = rng.integers(ktop_min, NLAY - 1, size=n)
        thick = rng.integers(1, 7, size=n)
        frac = rng.uniform(0.05, 1.0, size=n)
        k = np.arange(NLAY)[:, None]
        inside = has[None, :] & (k >= top[None, :]) & (k < np.minimum(top + thick, NLAY)[None, :])
        cldf = np.where(inside, np.maximum(cldf, frac[None, :]), cldf)
    cloudy = cldf > 0.0
    qc = np.where(cloudy & (t > 253.0), rng.uniform(0.0, 0.5e-3, size=(NLAY, n)), 0.0)
    qi = np.where(cloudy & (t < 273.0), rng.uniform(0.0, 0.2e-3, size=(NLAY, n)), 0.0)
    return p_lev, t, q, o3, cldf, qc, qi


def columns(ncol: int, seed: int = 0) -> dict:
    rng = np.random.default_rng(seed)
    ps = rng.uniform(95000.0, 104000.0, size=ncol)
    tnoise = rng.normal(0.0, 3.0, size=(NLAY, ncol))
    rh = rng.uniform(0.05, 0.95, size=(NLAY, ncol))
    p_lev, t, q, o3, cldf, qc, qi = _profiles(ps, tnoise, rh, rng)
    return _finish(p_lev, t, q, o3, cldf, qc, qi, rng, rng.uniform(-1.0, 1.0, size=ncol))


def _finish(p_lev, t, q, o3, cldf, qc, qi, rng, cosz):
    n = t.shape[1]
    return {
        "p_lev": np.ascontiguousarray(p_lev), "t_lay": np.ascontiguousarray(t),
        "q_lay": np.ascontiguousarray(q), "o3_lay": np.ascontiguousarray(o3),
        "cldf": np.ascontiguousarray(cldf), "qc": np.ascontiguousarray(qc),
        "qi": np.ascontiguousarray(qi),
        "t_sfc": t[-1] + rng.uniform(-5.0, 5.0, size=n),
        "emis": rng.uniform(0.9, 1.0, size=n),
        "alb": rng.uniform(0.05, 0.8, size=n),
        "cosz": np.ascontiguousarray(cosz),
    }


def _grf(rng, nx, ny, dx_km, scale_km, nfields):
    """Unit-variance Gaussian random fields [nfields, ny, nx] with a Gaussian spectrum of the given
    length scale, drawn by FFT filtering of white noise.  The scale is capped at a quarter of the
    domain's shorter side: a longer one leaves only the mean, a field of zero variance."""
    scale_km = min(scale_km, 0.25 * min(nx, ny) * dx_km)
    kx = np.fft.rfftfreq(nx, d=dx_km)
    ky = np.fft.fftfreq(ny, d=dx_km)
    k2 = ky[:, None] ** 2 + kx[None, :] ** 2
    filt = np.exp(-2.0 * (np.pi * scale_km) ** 2 * k2)
    out = np.empty((nfields, ny, nx))
    for f in range(nfields):
        w = rng.standard_normal((ny, nx))
        s = np.fft.irfft2(np.fft.rfft2(w) * filt, s=(ny, nx))
        out[f] = s / s.std()
    return out


def _ncdf(x):
    from scipy.special import ndtr
    return ndtr(x)


def grid(nx: int = 954, ny: int = 835, seed: int = 0, dx_km: float = 12.0, scale_km: float = 300.0) -> dict:
    rng = np.random.default_rng(seed)
    n = nx * ny
    psf = _grf(rng, nx, ny, dx_km, scale_km, 1)[0].reshape(n)
    ps = 95000.0 + 9000.0 * _ncdf(psf)
    tn = 3.0 * _grf(rng, nx, ny, dx_km, scale_km, NLAY).reshape(NLAY, n)
    rh = 0.05 + 0.9 * _ncdf(_grf(rng, nx, ny, dx_km, scale_km, NLAY).reshape(NLAY, n))
    p_lev, t, q, o3, cldf, qc, qi = _profiles(ps, tn, rh, rng)
    lat = np.deg2rad(np.linspace(15.0, 70.0, ny))[:, None] * np.ones((1, nx))
    lon = np.deg2rad(np.linspace(-150.0, -50.0, nx))[None, :] * np.ones((ny, 1))
    cosz = (np.cos(lat) * np.cos(lon)).reshape(n)   # 12 UTC, declination 0: hour angle = longitude
    return _finish(p_lev, t, q, o3, cldf, qc, qi, rng, cosz)
