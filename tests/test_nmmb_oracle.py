"""Pins of the oracle's NMMB-state prep and post (CPU): driver prep, the Ferrier condensate split,
progcld2's cloud fraction, paths and radii, gethml's cloud cover, the subgrid cloud and driver post,
each against what the reference's formulas fix in closed form or by invariants."""
import numpy as np
import pytest

from oracle import pyoracle
from woof_nam_amd import abi, gen

LM = 60


def state(im, jm, seed=0, t=None, cw=None, f_ice=None, f_rain=None, q=None, lat_deg=40.0):
    """A float32 NMMB state of im x jm columns on the NAM hybrid coordinate (gen's profiles)."""
    c = gen.columns(im * jm, seed=seed)
    sg1, sg2, pdtop = gen._sigma()
    pd = (c["p_lev"][-1] - gen.PT - pdtop).astype(np.float32)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    st = {"pd": f32(pd.reshape(jm, im)), "pdsg1": f32(np.diff(sg1) * pdtop), "dsg2": f32(np.diff(sg2)),
          "psgml1": f32(0.5 * (sg1[1:] + sg1[:-1]) * pdtop + gen.PT), "sgml2": f32(0.5 * (sg2[1:] + sg2[:-1])),
          "t": f32(c["t_lay"].reshape(LM, jm, im)), "q": f32(c["q_lay"].reshape(LM, jm, im)),
          "o3": f32(c["o3_lay"].reshape(LM, jm, im)), "cw": f32(np.zeros((LM, jm, im))),
          "f_ice": f32(np.zeros((LM, jm, im))), "f_rain": f32(np.zeros((LM, jm, im))),
          "glat": f32(np.full((jm, im), np.radians(lat_deg)))}
    for k, v in (("t", t), ("cw", cw), ("f_ice", f_ice), ("f_rain", f_rain), ("q", q)):
        if v is not None:
            st[k][...] = v
    return st


def dims(st, j0=0, j1=None):
    lm, jm, im = st["t"].shape
    return (im, jm, lm, j0, jm if j1 is None else j1, np.float32(gen.PT))


def test_driver_prep_interfaces_and_layer_pressures():
    st = state(7, 5, seed=1)
    d = dims(st, 1, 4)
    cols = pyoracle.nmmb_prep(st, d, (10, 1, 0))
    im = 7
    for c in range(3 * im):
        j, i = 1 + c // im, c % im
        p = np.float32(gen.PT)
        ref = [p]
        for k in range(LM):                     # P8W(K+1) = P8W(K) + PDSG1(K) + DSG2(K) PD, single precision
            p = np.float32(np.float32(p + st["pdsg1"][k]) + np.float32(st["dsg2"][k] * st["pd"][j, i]))
            ref.append(p)
        assert np.array_equal(cols["p_lev"][:, c], np.array(ref, dtype=np.float64))
        pphy = np.float32(st["sgml2"] * st["pd"][j, i]) + st["psgml1"]
        assert np.array_equal(cols["p_phy"][:, j, i], pphy)
    assert (cols["p_phy"][:, 0, :] == -1.0).all() and (cols["p_phy"][:, 4, :] == -1.0).all()   # other rows
    assert np.all(np.diff(cols["p_lev"], axis=0) > 0)


def one_layer(k, clw, t=None, f_ice=0.0, f_rain=0.0, seed=3):
    st = state(1, 1, seed=seed)
    st["cw"][k, 0, 0] = clw
    st["f_ice"][k, 0, 0] = f_ice
    st["f_rain"][k, 0, 0] = f_rain
    if t is not None:
        st["t"][k, 0, 0] = t
    cols = pyoracle.nmmb_prep(st, dims(st), (10, 1, 0))
    return st, cols


def test_condensate_split_conserves_mass_and_excludes_rain_from_cover():
    k = 50
    for f_ice, f_rain in ((0.0, 0.0), (1.0, 0.0), (0.3, 0.5), (0.0, 1.0), (1.7, -0.2)):
        st, cols = one_layer(k, 2.0e-4, t=280.0, f_ice=f_ice, f_rain=f_rain)
        cpath = (cols["p_lev"][k + 1, 0] - cols["p_lev"][k, 0]) * 10.0e-3 * (100000.0 / 9.80665)
        fi = min(1.0, max(0.0, float(np.float32(f_ice))))
        fr = min(1.0, max(0.0, float(np.float32(f_rain))))
        clw = float(np.float32(2.0e-4))
        cwp, cip, crp = cols["cwp"][k, 0], cols["cip"][k, 0], cols["crp"][k, 0]
        if (1 - fi) * fr == 1.0:                                # all rain: no cloud, and no paths with it
            assert cwp == cip == crp == cols["cldf"][k, 0] == 0.0
            continue
        assert cwp + cip + crp == pytest.approx(cpath * clw, rel=1e-12)
        assert cip == pytest.approx(cpath * clw * fi, rel=1e-12, abs=1e-300)
        assert crp == pytest.approx(cpath * clw * (1 - fi) * fr, rel=1e-12, abs=1e-300)
        pl = 0.5 * (cols["p_lev"][k, 0] + cols["p_lev"][k + 1, 0]) * 1e-2
        clw2 = clw * (1.0 - (1 - fi) * fr)                      # rain is not cloud
        assert cols["cldf"][k, 0] == pytest.approx(min(1.0, 1.0e5 * clw2 * pl * 1e-3), rel=1e-12)


def test_below_minus_40_all_condensate_is_ice():
    k = 20
    _, cols = one_layer(k, 5.0e-6, t=220.0, f_ice=0.0, f_rain=1.0)
    assert cols["cwp"][k, 0] == 0.0 and cols["crp"][k, 0] == 0.0 and cols["cip"][k, 0] > 0.0


def test_thin_cloud_is_dropped_and_cover_saturates():
    k = 45
    pl = None
    _, cols = one_layer(k, 5.0e-9, t=280.0)     # 1e5 clw p/1000 < 0.001 with p ~ 800 hPa
    assert cols["cldf"][k, 0] == 0.0 and cols["cwp"][k, 0] == 0.0
    _, cols = one_layer(k, 1.0e-3, t=280.0)
    assert cols["cldf"][k, 0] == 1.0
    assert (cols["cldf"][np.arange(LM) != k, 0] == 0.0).all()
    del pl


def test_liquid_radius_scales_with_the_cube_root_of_water_content():
    k = 55
    r = []
    for clw in (2.0e-4, 1.6e-3):                      # 8x the water: twice the radius above the 5 micron floor
        _, cols = one_layer(k, clw, t=285.0)
        r.append(cols["rew"][k, 0])
    assert r[1] / r[0] == pytest.approx(2.0, rel=1e-6)
    _, cols = one_layer(k, 1.0e-7, t=285.0)           # tiny water content: the 5 micron floor
    assert cols["rew"][k, 0] == 5.0
    # rew = 620.35 / TNW^(1/3) (rho q)^(1/3) with TNW = 200: rho q = 200 (r / 620.35)^3 gives r
    st, cols = one_layer(k, 1.0e-3, t=285.0)
    pl = 0.5 * (cols["p_lev"][k, 0] + cols["p_lev"][k + 1, 0]) * 1e-2
    q = max(1e-7, float(st["q"][k, 0, 0]))
    rho = 100.0 * pl / (287.05 * 285.0 * (1.0 + (461.5 / 287.05 - 1.0) * q))
    qc = float(np.float32(1.0e-3))
    assert cols["rew"][k, 0] == pytest.approx(620.35 * (rho * qc / 200.0) ** (1.0 / 3.0), rel=1e-12)


def test_ice_radius_regimes_and_known_answer():
    k = 30
    st, cols = one_layer(k, 1.0e-5, t=250.0, f_ice=1.0)
    assert cols["rei"][k, 0] >= 10.0
    # at an ice water content of 1 g/m3 every power law is 1: rei = 1250 / 9.387 * 0.6495 above -30 C
    p_lo, p_hi = cols["p_lev"][k, 0], cols["p_lev"][k + 1, 0]
    pl = 0.5 * (p_lo + p_hi) * 1e-2
    q = max(1e-7, float(st["q"][k, 0, 0]))
    t = 250.0
    tvly = t * (1.0 + (461.5 / 287.05 - 1.0) * q)
    dens = 9.80665 * pl / (287.05 * (p_hi - p_lo) * 1e-2)            # tem2d: g p / (Rd dp)
    cpath = (p_hi - p_lo) * 1e-2 * (100000.0 / 9.80665)
    clw = tvly / (dens * cpath)                                       # tem3 = dens * cip / tvly = 1
    for tk, re in ((250.0, 1250.0 / 9.387), (238.0, 1250.0 / 9.208), (228.0, 1250.0 / 9.337), (215.0, 1250.0 / 9.917)):
        st2 = state(1, 1, seed=3)
        st2["cw"][k, 0, 0] = clw
        st2["f_ice"][k, 0, 0] = 1.0
        st2["t"][k, 0, 0] = tk
        c2 = pyoracle.nmmb_prep(st2, dims(st2), (10, 1, 0))
        tv2 = tk * (1.0 + (461.5 / 287.05 - 1.0) * q)
        tem3 = dens * c2["cip"][k, 0] / tv2
        expect = max(10.0, re * tem3 ** {1250.0 / 9.387: 0.031, 1250.0 / 9.208: 0.055, 1250.0 / 9.337: 0.08,
                                        1250.0 / 9.917: 0.109}[re] * 0.6495)
        assert c2["rei"][k, 0] == pytest.approx(expect, rel=1e-9)


def gethml_case(cfs, iovr=1, lat=40.0):
    """clds, mtop, mbot for a column with the given {model layer: cloud fraction} (liquid water chosen to
    give exactly that fraction)."""
    st = state(1, 1, seed=4, lat_deg=lat)
    st["t"][...] = np.maximum(st["t"], 260.0)
    cols0 = pyoracle.nmmb_prep(st, dims(st), (10, iovr, 0))
    for k, f in cfs.items():
        pl = 0.5 * (cols0["p_lev"][k, 0] + cols0["p_lev"][k + 1, 0]) * 1e-2
        st["cw"][k, 0, 0] = f / (1.0e5 * pl * 1e-3)
    cols = pyoracle.nmmb_prep(st, dims(st), (10, iovr, 0))
    return cols


def test_gethml_cloud_cover_by_domain_and_overlap():
    # model layer 55 is near 950 hPa (low), 30 near 500-650 hPa (middle), 15 above 400 hPa (high)
    c = gethml_case({55: 0.4})
    assert c["clds"][0, 0] == pytest.approx(0.4, rel=1e-6) and c["clds"][1, 0] == 0 and c["clds"][2, 0] == 0
    assert c["clds"][3, 0] == pytest.approx(0.4, rel=1e-6)
    assert c["mtop"][0, 0] == c["mbot"][0, 0] == 55 + 2          # radiation layer, 1-based, added top layer = 1
    c = gethml_case({54: 0.3, 55: 0.6})                           # adjacent: maximum overlap
    assert c["clds"][0, 0] == pytest.approx(0.6, rel=1e-6) and c["clds"][3, 0] == pytest.approx(0.6, rel=1e-6)
    c = gethml_case({52: 0.3, 55: 0.6})                           # separated by clear layers: random overlap
    assert c["clds"][3, 0] == pytest.approx(1 - 0.7 * 0.4, rel=1e-6)
    c = gethml_case({54: 0.3, 55: 0.6}, iovr=0)                    # random overlap everywhere
    assert c["clds"][3, 0] == pytest.approx(1 - 0.7 * 0.4, rel=1e-6)
    c = gethml_case({15: 0.5, 55: 0.2})
    assert c["clds"][2, 0] == pytest.approx(0.5, rel=1e-6) and c["clds"][0, 0] == pytest.approx(0.2, rel=1e-6)
    assert c["clds"][1, 0] == 0.0


def test_subgrid_cloud_water_of_humid_layers():
    k = 40
    st = state(1, 1, seed=5)
    t = float(st["t"][k, 0, 0])
    cols0 = pyoracle.nmmb_prep(st, dims(st), (10, 1, 0))
    p_kpa = 0.5 * (cols0["p_lev"][k, 0] + cols0["p_lev"][k + 1, 0]) * 1e-3
    # saturation over water / ice (Bolton / Murphy-Koop within 1 %), mixing-ratio humidity 97 %
    es = 0.6112 * np.exp(17.67 * (t - 273.15) / (t - 29.65)) if t >= 273.16 else \
        np.exp(9.550426 - 5723.265 / t + 3.53068 * np.log(t) - 0.00728332 * t) * 1e-3
    qsat = (287.04 / 461.6) * es / (p_kpa - es)
    w = 0.97 * qsat
    st["q"][k, 0, 0] = w / (1.0 + w)
    off = pyoracle.nmmb_prep(st, dims(st), (10, 1, 0))
    on = pyoracle.nmmb_prep(st, dims(st), (10, 1, 1))
    assert off["cwp"][k, 0] + off["cip"][k, 0] == 0.0 and off["cldf"][k, 0] == 0.0
    cw = min(1.0e-5, 0.02 * qsat)
    cpath = (on["p_lev"][k + 1, 0] - on["p_lev"][k, 0]) * 10.0e-3 * (100000.0 / 9.80665)
    got = (on["cwp"][k, 0] + on["cip"][k, 0]) / cpath
    assert got == pytest.approx(cw, rel=0.03)
    assert (on["cip"][k, 0] > 0) == (t <= 243.15)


def test_post_outputs_and_accumulators():
    im, jm = 6, 4
    st = state(im, jm, seed=6)
    st["cw"][30:40] = 2.0e-4
    st["cw"][:, 2, 3] = 0.0                                       # one column without condensate
    d = dims(st, 1, 3)
    cols = pyoracle.nmmb_prep(st, d, (10, 1, 0))
    ncol = 2 * im
    rng = np.random.default_rng(0)
    rad = {k: rng.uniform(1, 2, s) for k, s in abi.output_shapes(ncol, LM).items()}
    cz = rng.uniform(0, 1, ncol)
    out = {k: np.full((LM, jm, im), -3.0, np.float32) for k in ("rlwtt", "rswtt", "cldfra")}
    out.update({k: np.full((jm, im), 0.5, np.float32) for k in abi.NMMB_OUT_FIELDS if k not in out})
    out.update({k: np.full((jm, im), 2, np.int32) for k in ("ncfrst", "ncfrcv")})
    pyoracle.nmmb_post(st, d, cols, rad, cz, True, False, out)
    c = 1 * im + 2                                                # row 2, column 2
    assert out["rlwtt"][7, 2, 2] == np.float32(rad["hr_lw"][7, c])
    assert out["rlwin"][2, 2] == np.float32(rad["flw_dn"][LM + 1, c]) and out["rlwtoa"][2, 2] == np.float32(rad["flw_up"][0, c])
    tl = float(st["t"][LM - 1, 2, 2])
    assert out["sigt4"][2, 2] == pytest.approx(5.67051e-8 * tl ** 4, rel=1e-6)
    assert (out["rswtt"] == -3.0).all() and (out["rswin"] == 0.5).all()      # not a SW step
    assert out["czmean"][2, 2] == np.float32(cz[c])
    assert np.array_equal(out["cldfra"][:, 2, 2], cols["cldf"][:, c].astype(np.float32))
    assert (out["cldfra"][:, 2, 3] == 0.0).all()                            # no condensate: no cover
    assert out["cfracl"][2, 2] == cols["clds"][0, c]
    assert out["acfrst"][2, 2] == np.float32(0.5) + cols["clds"][3, c]
    assert out["ncfrst"][2, 2] == 3 and out["ncfrcv"][2, 2] == 3 and out["acfrcv"][2, 2] == np.float32(0.5)
    assert (out["cldfra"][:, 0, :] == -3.0).all() and out["ncfrst"][0, 0] == 2   # rows outside J_S..J_E


def payne(cz):
    """The Payne open-water albedo at transmission 0.537 by piecewise-linear interpolation of the table
    (ocean_albedo.json) in the zenith angle, numpy's interp standing in for the driver's node search."""
    import json
    from pathlib import Path
    d = json.loads((Path(__file__).resolve().parents[1] / "woof_nam_amd" / "data" / "ocean_albedo.json").read_text())
    za = np.array(d["ZA"])[::-1]
    tab = np.array(d["ALBD_zenith_by_transmission"])[::-1]
    iq = int(20.0 * 0.537 + 1.0)
    zen = (180.0 / 3.14159265) * np.arccos(np.maximum(cz, 0.0))
    a1, a2 = np.interp(zen, za, tab[:, iq - 1]), np.interp(zen, za, tab[:, iq])
    return a1 + 20.0 * (a2 - a1) * (0.537 - d["TRN"][iq - 1])


def test_nam_albedo_payne_table_over_water_in_daylight():
    rng = np.random.default_rng(21)
    jm, im = 6, 50
    alb = rng.uniform(0.05, 0.8, (jm, im)).astype(np.float32)
    sm = (rng.uniform(size=(jm, im)) < 0.6).astype(np.float32)          # 1 = sea
    j0, j1 = 1, jm - 1
    nc = (j1 - j0) * im
    cz = rng.uniform(-0.3, 1.0, nc)
    cz[:20] = np.cos(np.radians(np.array([90, 88, 86, 84, 82, 80, 78, 76, 74, 70, 66, 62, 58, 54, 50, 40, 30, 20, 10,
                                          0.0]) * 3.14159265 / np.pi))   # the table's own zenith nodes
    got = pyoracle.nmmb_albedo(alb, sm, cz, (j0, j1))
    a = alb[j0:j1].reshape(-1).astype(np.float64)
    s = sm[j0:j1].reshape(-1)
    water_day = (s > 0.5) & (cz > 0.0)
    assert np.array_equal(got[~water_day], a[~water_day])                 # land, or the sun down: ALBEDO
    assert np.allclose(got[water_day], payne(cz[water_day]), rtol=0, atol=1e-12)
    assert water_day.sum() > 50
    # the curve: 0.05-0.6, higher the lower the sun
    z = np.linspace(0.0, 89.9, 400)
    cz2 = np.cos(np.radians(z * 3.14159265 / np.pi))
    one = pyoracle.nmmb_albedo(np.zeros((1, 400), np.float32), np.ones((1, 400), np.float32), cz2, (0, 1))
    assert np.all(np.diff(one) >= -1e-15) and 0.045 < one.min() and one.max() < 0.6


def _clim(fn, pstr=None):
    """A synthetic ozone climatology o3r[m][k][j] = fn(m, k, j) on the real pressure levels."""
    from woof_nam_amd import ozone
    h = ozone.host_tables()
    pstr = h["pstr"] if pstr is None else pstr
    m, k, j = np.meshgrid(np.arange(12), np.arange(pstr.size), np.arange(18), indexing="ij")
    return {"o3r": fn(m, k, j).astype(np.float32), "pstr": pstr, "blte": h["blte"], "dlte": h["dlte"]}


def test_ozone_month_and_latitude_interpolation():
    """A climatology independent of pressure and linear in latitude (a_m + beta * latitude_j): every
    layer's ozone is (1 - facoz) a_k1 + facoz a_k2 + beta * latitude, inside the latitude range, and the
    edge latitude's value outside it."""
    a = np.linspace(1.0e-6, 2.2e-6, 12)
    beta = 1.0e-9
    lat_j = lambda j: -85.0 + 10.0 * j
    clim = _clim(lambda m, k, j: a[m] + beta * lat_j(j))
    for lat, k1, k2, f in ((37.3, 3, 4, 0.25), (-62.0, 12, 1, 0.8), (88.0, 6, 7, 0.5), (-89.5, 1, 2, 0.0)):
        st = state(2, 1, seed=4, lat_deg=lat)
        oz = {**clim, "k1oz": k1, "k2oz": k2, "facoz": f}
        cols = pyoracle.nmmb_prep(st, dims(st), (10, 1, 0), ozone=oz)
        deglat = float(st["glat"][0, 0]) * 57.29578
        latc = min(max(deglat, -85.0), 85.0)
        expect = (1.0 - np.float32(f)) * a[k1 - 1] + np.float32(f) * a[k2 - 1] + beta * latc
        assert np.allclose(cols["o3_lay"], expect, rtol=1e-6, atol=0), lat
        assert np.allclose(cols["o3_top"], expect, rtol=1e-6, atol=0), lat


def test_ozone_is_linear_in_the_exner_function():
    """A climatology proportional to the Exner function of its own levels, gamma (p / 1e5 Pa)^(R/cp):
    the layers inside the climatology's pressure range get gamma times their own Exner function (the
    driver call's (P_PHY / 1e5)^(R/cp)); the layers below its lowest level get the lowest level's value."""
    from woof_nam_amd import ozone
    h = ozone.host_tables()
    rocp = float(np.float32(287.05) / np.float32(1004.6))
    pk = (h["pstr"].astype(np.float64) * 100.0 / 1.0e5) ** rocp
    gamma = 3.0e-6
    clim = _clim(lambda m, k, j: gamma * pk[k] + 0.0 * m + 0.0 * j)
    st = state(3, 2, seed=6)
    cols = pyoracle.nmmb_prep(st, dims(st), (10, 1, 0), ozone={**clim, "k1oz": 5, "k2oz": 6, "facoz": 0.3})
    pphy = (st["sgml2"][:, None, None] * st["pd"][None] + st["psgml1"][:, None, None]).astype(np.float32)
    wk = ((pphy / np.float32(1000.0)).astype(np.float64) * 0.01) ** float(np.float32(287.04) / np.float32(1004.6))
    wk = wk.reshape(LM, -1)
    inside = (pphy.reshape(LM, -1) / 100.0 < h["pstr"][-1]) & (pphy.reshape(LM, -1) / 100.0 > h["pstr"][0])
    below = pphy.reshape(LM, -1) / 100.0 > h["pstr"][-1] * 1.001
    assert inside.sum() > 20 and below.sum() > 20
    assert np.allclose(cols["o3_lay"][inside], gamma * wk[inside], rtol=2e-6)
    assert np.allclose(cols["o3_lay"][below], np.float32(gamma * pk[-1]), rtol=1e-6)


def test_ozone_time_factors():
    """gas_update's time factors: facoz runs 0 -> 1 between two months' middle days, the months move
    on at hour 0 of the middle day, and the interpolated value is continuous across that change."""
    from woof_nam_amd import ozone
    days = (31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31)
    t = ozone.TimeFactors()
    prev = None
    for month in range(1, 13):
        for day in range(1, days[month - 1] + 1):
            for hour in (0, 12):
                t.update(month, day, hour)
                assert 0.0 <= t.facoz < 1.0 and t.k2oz == t.k1oz % 12 + 1
                mid = days[month - 1] // 2 + 1
                assert (t.k1oz == month) == (day >= mid)
                w = np.zeros(13)                               # the months' weights
                w[t.k1oz] += 1.0 - t.facoz
                w[t.k2oz] += t.facoz
                if prev is not None:
                    assert np.abs(w - prev).max() <= 1.0 / 28 + 1e-6    # at most one day's step
                prev = w
                if day == mid:
                    assert t.facoz == 0.0 and t.k1oz == month     # a month's middle day is its own value
