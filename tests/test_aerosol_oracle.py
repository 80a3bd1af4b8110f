"""Climatological aerosol (the NAM's iaer = 11), CPU only.

* The host tables (woof_nam_amd.aerosol): the solar spectrum integrates to the solar constant, the
  Planck weights to sigma T^4, each radiation band's aerosol-band weights to the band's flux; a
  spectrally flat species keeps its optics through the band means (the two-stream reflectance
  inversion returns the single-scattering albedo exactly); the monthly fill against the data file's
  own values.
* The oracle's column optics (oracle_aer.c) on synthetic tables with closed forms: hypsometric heights
  of an isothermal column, the domains from the reference sigmas, the effective thickness summed
  over a domain, the column optical depth of each band, a bilinear field reproduced inside a cell, the
  humidity interpolation of a table linear in humidity, the mineral-overlay rule, the smoothing at a
  domain boundary, and the exact doubling of optical depth with the extinction tables.
* Copies of the oracle with a planted mistake each must fail one of those pins.
"""
import numpy as np
import pytest

from oracle import pyoracle
from test_oracle_nodes import _build_copy
from woof_nam_amd import aerosol, gen

F = np.float32
SIGMA_SB = 5.670374419e-8
RD, G, RV = 287.05, 9.80665, 461.5
ROVG = 0.001 * RD / G
FVIRT = RV / RD - 1.0


# ---- host tables --------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def data():
    return aerosol.load()


@pytest.fixture(scope="module")
def spec(data):
    return aerosol.spectrum(data)


def test_solar_spectrum_is_the_solar_constant(spec):
    sol, _ = spec
    assert sol.shape == (aerosol.NWVTOT,) and (sol > 0).all()
    assert 1350.0 < sol.astype(np.float64).sum() < 1370.0      # W/m2 over 1..57600 cm-1


def test_planck_weights_integrate_to_sigma_t4(spec):
    _, eir = spec
    t0 = 273.15
    total = 100.0 * eir.astype(np.float64).sum()   # 1 cm-1 bins of a per-m-1 spectral flux
    assert abs(total / (SIGMA_SB * t0 ** 4) - 1.0) < 1e-3
    nu = 100.0 * np.arange(1, 4001)
    assert np.argmax(eir) + 1 == pytest.approx(2.821 * 1.380649e-23 * t0 / (6.62607015e-34 * 2.99792458e8) / 100.0,
                                               abs=1.0)   # Wien's displacement law in wavenumber
    assert nu.size == eir.size


def test_band_weights_partition_each_band(data, spec):
    sol, eir = spec
    iend = data["iendwv"]
    bands = [(sol, a, b) for a, b in zip(aerosol.WVNSW1, aerosol.WVNSW2)] + \
            [(eir, a, b) for a, b in zip(aerosol.WVNLW1, aerosol.WVNLW2)]
    for fwv, a, b in bands:
        w, tot, n1, n2 = aerosol._map(fwv, iend, a, b)
        assert tot == pytest.approx(float(fwv[int(a) - 1:int(b)].astype(np.float64).sum()), rel=1e-4)  # float32 sums
        assert w[:n1 - 1].sum() == 0 and w[n2:].sum() == 0
        assert float(w[n1 - 1:n2].astype(np.float64).sum()) == pytest.approx(float(tot), rel=1e-4)
        assert (n1 == 1 or iend[n1 - 2] < a <= iend[n1 - 1]) and b <= iend[n2 - 1] or n2 == aerosol.NAERBND


def test_flat_species_keeps_its_optics(data, spec):
    rng = np.random.default_rng(0)
    d = dict(data)
    ext = rng.uniform(1e-4, 1e-2, aerosol.NCM1).astype(F)
    ssa = rng.uniform(0.3, 0.99, aerosol.NCM1).astype(F)
    asy = rng.uniform(0.1, 0.8, aerosol.NCM1).astype(F)
    d["rhidext0"] = np.repeat(ext[:, None], aerosol.NAERBND, 1)
    d["rhidsca0"] = np.repeat((ext * ssa)[:, None], aerosol.NAERBND, 1)
    d["rhidssa0"] = np.repeat(ssa[:, None], aerosol.NAERBND, 1)
    d["rhidasy0"] = np.repeat(asy[:, None], aerosol.NAERBND, 1)
    bm = aerosol.band_means(d, *spec)
    sol, eir = spec
    tots = [float(sol[int(a) - 1:int(b)].astype(np.float64).sum()) for a, b in zip(aerosol.WVNSW1, aerosol.WVNSW2)] + \
           [float(eir[int(a) - 1:int(b)].astype(np.float64).sum()) for a, b in zip(aerosol.WVNLW1, aerosol.WVNLW2)]
    for ib in range(30):
        # float32 sums: the band's weights and its total agree to about 1e-4
        np.testing.assert_allclose(bm["extrhi"][ib], ext, rtol=2e-4)
        np.testing.assert_allclose(bm["scarhi"][ib], ext * ssa, rtol=2e-4)
        sok = ssa.astype(np.float64) * ext * tots[ib]     # the asymmetry's weight, regularised by 1e-10
        np.testing.assert_allclose(bm["asyrhi"][ib], asy * sok / (sok + 1.0e-10), rtol=2e-4)
        # 4R / ((1+R)^2 - g (1-R)^2) with R = (1-s)/(1+s), s^2 = (1-w)/(1-wg) is w itself
        np.testing.assert_allclose(bm["ssarhi"][ib], ssa, rtol=1e-3)


def test_band_means_of_the_data_are_physical(data, spec):
    bm = aerosol.band_means(data, *spec)
    for k in ("extrhi", "scarhi", "extrhd", "scarhd", "extstra"):
        assert (bm[k] >= 0).all(), k
    assert (bm["scarhi"] <= bm["extrhi"] * (1 + 1e-5)).all()
    for k in ("ssarhi", "ssarhd"):
        assert (bm[k] > 0).all() and (bm[k] <= 1.0 + 1e-5).all(), k
    for k in ("asyrhi", "asyrhd"):
        assert (bm[k] >= 0).all() and (bm[k] < 1.0).all(), k
    # humidity-dependent species swell: extinction grows with relative humidity in every band
    assert (np.diff(bm["extrhd"].astype(np.float64), axis=2) >= -1e-9).all()


def test_month_fill_matches_the_data_file(data):
    m = aerosol.month(data, 1)
    # January, the cell at 90 N 0 E, as nam_climaeropac_global.txt line 3 holds it
    assert m["idxcg"][0, 0].tolist() == [1, 7, 8, 0, 2] and m["kprfg"][0, 0] == 2
    assert m["denng"][0, 0].tolist() == [5675.0, 0.0]
    np.testing.assert_allclose(m["cmixg"][0, 0][:3], [4.434e-06, 0.2461, 5.438e-04], rtol=1e-4)
    for imon in (1, 7, 12):
        m = aerosol.month(data, imon)
        s = m["cmixg"].astype(np.float64).sum(-1)
        np.testing.assert_allclose(s, 1.0, atol=1e-6)     # the last component tops the sum up to 1
        raw = data["cmix"][imon - 1][..., 4]
        np.testing.assert_array_equal(m["denng"][..., 1], np.where(m["kprfg"] >= 6, raw, 0.0).astype(F))
    with pytest.raises(ValueError):
        aerosol.host_tables(13)


# ---- oracle column optics on synthetic tables -----------------------------------------------------

SIGREF = (0.8, 0.6, 0.3, 0.01, 0.0)


def synth(kprf=1, rh_linear=False, seed=0):
    """Tables with one profile type everywhere: domain scale heights 2 km (mixing layer) and 6 km (free
    troposphere), reference sigmas 0.8 / 0.6 / 0.3 / 0.01; species 1 (60 %) and 7 (40 %) in every cell,
    number densities 1000 and 50."""
    rng = np.random.default_rng(seed)
    t = {}
    t["haer"] = np.zeros((7, 5), F)
    t["haer"][:, 0], t["haer"][:, 2] = 2.0, 6.0
    t["haer"][5, 0], t["haer"][6, 0] = 3.0, 5.0
    t["prsref"] = np.tile(np.array([1000.0 * s for s in SIGREF], F), (7, 1))
    t["sigref"] = (F(0.001) * t["prsref"]).astype(F)
    t["rhlev"] = aerosol.RHLEV
    e = rng.uniform(1e-4, 1e-3, (30, 6)).astype(F)
    w = rng.uniform(0.5, 0.99, (30, 6)).astype(F)
    t["extrhi"], t["ssarhi"], t["asyrhi"] = e, w, rng.uniform(0.3, 0.8, (30, 6)).astype(F)
    t["scarhi"] = (e * w).astype(F)
    if rh_linear:
        rl = aerosol.RHLEV.astype(np.float64)
        ed = np.repeat(rng.uniform(1e-4, 1e-3, (30, 4, 1)), 8, 2)
        t["ssarhd"] = (0.5 + 0.4 * rl[None, None, :] + 0.0 * ed).astype(F)
        t["asyrhd"] = (0.3 + 0.5 * rl[None, None, :] + 0.0 * ed).astype(F)
    else:
        ed = np.repeat(rng.uniform(1e-4, 1e-3, (30, 4, 1)), 8, 2)
        t["ssarhd"] = np.repeat(rng.uniform(0.5, 0.99, (30, 4, 1)), 8, 2).astype(F)
        t["asyrhd"] = np.repeat(rng.uniform(0.3, 0.8, (30, 4, 1)), 8, 2).astype(F)
    t["extrhd"] = ed.astype(F)
    t["scarhd"] = (ed * t["ssarhd"]).astype(F)
    t["extstra"] = rng.uniform(1e-6, 1e-5, 30).astype(F)
    t["idxcg"] = np.zeros((37, 72, 5), np.int32)
    t["idxcg"][..., 0], t["idxcg"][..., 1] = 1, 7
    t["cmixg"] = np.zeros((37, 72, 5), F)
    t["cmixg"][..., 0], t["cmixg"][..., 1] = 0.6, 0.4
    t["denng"] = np.zeros((37, 72, 2), F)
    t["denng"][..., 0], t["denng"][..., 1] = 1000.0, 50.0
    t["kprfg"] = np.full((37, 72), kprf, np.int32)
    return t


def cols(n=64, seed=0, iso=None):
    c = gen.columns(n, seed=seed)
    rng = np.random.default_rng(seed + 100)
    lon = rng.uniform(0.01, 6.2, n).astype(F)
    lat = rng.uniform(-1.5, 1.5, n).astype(F)
    sl = rng.choice([0.0, 1.0, 2.0], n).astype(F)
    if iso is not None:
        c["t_lay"][:] = iso
        c["q_lay"][:] = 0.0
    return c, lon, lat, sl


def _col(L, tab, c, lon, lat, sl, i):
    return pyoracle.aer_column(tab, c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl, i, L=L)


def _levels(p_lev, i):
    pl = np.concatenate([[1.0e-6], 10.0 * (p_lev[:, i] * 1.0e-3)])
    return pl


def check_hypsometric_heights(L):
    tab = synth()
    c, lon, lat, sl = cols(8, seed=1, iso=250.0)
    for i in range(8):
        a = _col(L, tab, c, lon, lat, sl, i)
        pl = _levels(c["p_lev"], i)
        h = ROVG * 250.0 * (1.0 + FVIRT * 1.0e-7)
        np.testing.assert_allclose(a["hz"][1:], h * np.log(pl[-1] / pl[1:]), rtol=1e-12, atol=1e-12)
        assert a["hz"][0] - a["hz"][1] == pytest.approx(2.0 * h * np.log(2.0), rel=1e-12)


def check_relative_humidity(L):
    tab = synth()
    c, lon, lat, sl = cols(8, seed=10)
    t = 290.0
    c["t_lay"][:] = t
    p = 0.5 * (c["p_lev"][:-1] + c["p_lev"][1:])               # layer pressure, Pa
    es = 611.2 * np.exp(17.67 * (t - 273.15) / (t - 29.65))   # Bolton (1980), over water
    e = 0.5 * es
    c["q_lay"][:] = 0.622 * e / (p - 0.378 * e)                # specific humidity at 50 %
    for i in range(8):
        a = _col(L, tab, c, lon, lat, sl, i)
        ok = p[:, i] > 20.0 * es
        assert ok.sum() > 10
        np.testing.assert_allclose(a["rh"][1:][ok], 0.5, atol=0.01)
        assert a["rh"][0] == a["rh"][1]


def check_domains_from_reference_sigma(L):
    tab = synth()
    c, lon, lat, sl = cols(16, seed=2)
    sig = tab["sigref"][0].astype(np.float64)
    for i in range(16):
        a = _col(L, tab, c, lon, lat, sl, i)
        pl = _levels(c["p_lev"], i)
        x = pl[:-1] * (1.0 / pl[-1])    # each layer's top level over the surface pressure
        want = 1 + (x[:, None] < sig[None, :4]).sum(1)
        np.testing.assert_array_equal(a["idom"], want)


def _domain_top(a, d):
    """Height of the top of domain d (and everything below it)."""
    ls = np.nonzero(a["idom"] <= d)[0]
    return a["hz"][ls.min()] if ls.size else 0.0


def check_effective_thickness_telescopes(L):
    tab = synth()
    c, lon, lat, sl = cols(16, seed=3)
    for i in range(16):
        a = _col(L, tab, c, lon, lat, sl, i)
        z1, z2, z3, z4 = (_domain_top(a, d) for d in (1, 2, 3, 4))
        dl = a["delz"]
        assert dl[a["idom"] == 1].sum() == pytest.approx(2.0 * (1.0 - np.exp(-z1 / 2.0)), rel=1e-12)
        assert dl[a["idom"] == 2].sum() == pytest.approx(z2 - z1, rel=1e-12)
        assert dl[a["idom"] == 3].sum() == pytest.approx(6.0 * (np.exp(-z2 / 6.0) - np.exp(-z3 / 6.0)), rel=1e-12)
        assert dl[a["idom"] == 4].sum() == pytest.approx(z4 - z3, rel=1e-12)


def check_column_optical_depth(L):
    tab = synth()
    c, lon, lat, sl = cols(32, seed=4)
    sw, lw = pyoracle.aer_optics(tab, c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl, L=L)
    tau = np.concatenate([sw[0], lw[0]], axis=0).astype(np.float64)     # [30][nrl][ncol]
    e = tab["extrhi"].astype(np.float64)
    ed = tab["extrhd"][:, 0, 0].astype(np.float64)
    ext1 = 0.6 * e[:, 0] + 0.4 * ed
    ext3 = 0.17e-3 * e[:, 0] + 0.4 * e[:, 1] + 0.59983 * ed
    for i in range(32):
        a = _col(L, tab, c, lon, lat, sl, i)
        z1, z2, z3, z4 = (_domain_top(a, d) for d in (1, 2, 3, 4))
        want = (ext1 * 1000.0 * 2.0 * (1.0 - np.exp(-z1 / 2.0)) + ext3 * 730.0 * 6.0 * (np.exp(-z2 / 6.0) - np.exp(-z3 / 6.0))
                + tab["extstra"].astype(np.float64) * (z4 - z3))
        np.testing.assert_allclose(tau[:, :, i].sum(1), want, rtol=2e-6)


def check_bilinear_in_the_cell(L):
    tab = synth()
    ii, jj = np.meshgrid(np.arange(1, 73), np.arange(1, 38))
    tab["denng"][..., 0] = (100.0 + 3.0 * ii + 7.0 * jj).astype(F)
    tab["cmixg"][..., 0] = (0.001 * ii + 0.002 * jj).astype(F)
    c, _, _, sl = cols(24, seed=5)
    rng = np.random.default_rng(5)
    lon = rng.uniform(0.01, np.deg2rad(354.0), 24).astype(F)
    lat = rng.uniform(np.deg2rad(-89.0), np.deg2rad(89.0), 24).astype(F)
    for i in range(24):
        a = _col(L, tab, c, lon, lat, sl, i)
        alon, alat = float(lon[i]) * (180.0 / np.pi), float(lat[i]) * (180.0 / np.pi)
        fi, fj = 1.0 + alon / 5.0, 1.0 + (90.0 - alat) / 5.0       # fractional cell indices
        assert a["denn"][1] == pytest.approx(100.0 + 3.0 * fi + 7.0 * fj, rel=1e-12)
        assert a["cmix"][1] == pytest.approx(0.001 * fi + 0.002 * fj, rel=1e-6)
        assert a["cmix"][7] == pytest.approx(0.4, rel=1e-6)


def check_humidity_interpolation(L):
    tab = synth(rh_linear=True)
    tab["idxcg"][..., 0] = 0          # species 7 alone: ssa and asy are its own, interpolated in humidity
    c, lon, lat, sl = cols(24, seed=6)
    sw, lw = pyoracle.aer_optics(tab, c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl, L=L)
    for i in range(24):
        a = _col(L, tab, c, lon, lat, sl, i)
        d1 = a["idom"] == 1
        rh = np.minimum(a["rh"][d1], 0.99)
        assert d1.any() and (a["rh"] >= 0).all() and (a["rh"] <= 1).all()
        np.testing.assert_allclose(sw[1][:, d1, i], np.broadcast_to(0.5 + 0.4 * rh, (14, d1.sum())), rtol=2e-6)
        np.testing.assert_allclose(lw[2][:, d1, i], np.broadcast_to(0.3 + 0.5 * rh, (16, d1.sum())), rtol=2e-6)


def check_mineral_overlay(L):
    tab = synth()
    tab["kprfg"][2, 1] = 6            # the cell at 80 N, 5 E: an ocean profile with mineral transport
    p = np.deg2rad
    c, _, _, _ = cols(4, seed=7)
    lon = np.array([p(2.0), p(2.0), p(100.0), p(2.0)], F)
    lat = np.array([p(82.0), p(82.0), p(82.0), p(88.0)], F)
    sl = np.array([1.0, 0.0, 1.0, 1.0], F)
    kps = [_col(L, tab, c, lon, lat, sl, i)["kp"] for i in range(4)]
    assert kps == [7, 6, 1, 1]        # land point next to it: land profile; sea: the ocean one; far: none
    assert _col(L, tab, c, lon, lat, sl, 0)["denn"][2] == pytest.approx(50.0, rel=1e-12)
    assert _col(L, tab, c, lon, lat, sl, 2)["denn"][2] == 0.0


def check_smoothing_at_the_stratosphere(L):
    tab = synth()
    tab["extstra"][:] = 1.0e-9        # stratosphere far thinner than the free troposphere below it
    c, lon, lat, sl = cols(8, seed=8)
    sw, lw = pyoracle.aer_optics(tab, c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl, L=L)
    e = tab["extrhi"].astype(np.float64)
    ext3 = 0.17e-3 * e[:, 0] + 0.4 * e[:, 1] + 0.59983 * tab["extrhd"][:, 0, 0].astype(np.float64)
    tau = np.concatenate([sw[0], lw[0]], axis=0).astype(np.float64)
    for i in range(8):
        a = _col(L, tab, c, lon, lat, sl, i)
        l = int(np.nonzero(a["idom"] == 3)[0].min())       # top layer of the free troposphere
        assert a["idom"][l - 1] == 4
        up, lo = 1.0e-9 * a["delz"][l - 1], ext3 * 730.0 * a["delz"][l]
        np.testing.assert_allclose(tau[:, l - 1, i], 0.2 * (up + lo), rtol=1e-6)
        np.testing.assert_allclose(tau[:, l, i], 0.8 * (up + lo), rtol=1e-6)


def check_extinction_doubling_is_exact(L):
    tab = aerosol.host_tables(7)
    c, lon, lat, sl = cols(48, seed=9)
    sw, lw = pyoracle.aer_optics(tab, c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl, L=L)
    t2 = dict(tab)
    for k in ("extrhi", "extrhd", "extstra"):
        t2[k] = (2.0 * tab[k]).astype(F)
    sw2, lw2 = pyoracle.aer_optics(t2, c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl, L=L)
    for a, b in ((sw, sw2), (lw, lw2)):
        np.testing.assert_array_equal(b[0], 2.0 * a[0])
        np.testing.assert_array_equal(b[1:], a[1:])
    assert (sw[0] >= 0).all() and (sw[0][:, 1:].sum(1) > 0).all()


ALL = [check_hypsometric_heights, check_relative_humidity, check_domains_from_reference_sigma, check_effective_thickness_telescopes,
       check_column_optical_depth, check_bilinear_in_the_cell, check_humidity_interpolation, check_mineral_overlay,
       check_smoothing_at_the_stratosphere, check_extinction_doubling_is_exact]


@pytest.mark.parametrize("check", ALL, ids=lambda f: f.__name__)
def test_aerosol_oracle_pin(check):
    check(pyoracle.lib())


def test_saturation_vapour_pressure_known_points():
    assert pyoracle.aer_fpvs(273.16) == pytest.approx(610.78, rel=1e-9)     # the table's triple-point node
    assert pyoracle.aer_fpvs(330.0) == pytest.approx(17210.0, rel=0.02)     # over water at 56.85 C
    assert pyoracle.aer_fpvs(373.15) == pyoracle.aer_fpvs(330.0)             # the table ends at 330 K
    assert pyoracle.aer_fpvs(293.15) == pytest.approx(2339.0, rel=0.01)     # over water at 20 C
    assert pyoracle.aer_fpvs(253.15) == pytest.approx(103.3, rel=0.02)      # over ice at -20 C
    assert 259.9 < pyoracle.aer_fpvs(263.16) < 286.5                        # -10 C: between ice and water


MUTATIONS = [
    ("oracle_aer.c", "a->delz[l] = tmp1 * (om_exp(-a->hz[l + 1] * tmp2) - om_exp(-a->hz[l] * tmp2));",
     "a->delz[l] = tmp1 * (om_exp(-a->hz[l] * tmp2) - om_exp(-a->hz[l + 1] * tmp2));", "effective thickness sign"),
    ("oracle_aer.c", "if (ii <= A_NDM && a->pl[l] * rps < SIGREF(ii, kp)) {",
     "if (ii <= A_NDM && a->pl[l + 1] * rps < SIGREF(ii, kp)) {", "domain test on the layer's bottom level"),
    ("oracle_aer.c", "a->dz[0] = 2.0 * a->dz[0];", "a->dz[0] = 1.0 * a->dz[0];", "added top layer not doubled"),
    ("oracle_aer.c", "tt1 = 0.2 * tt0", "tt1 = 0.25 * tt0", "smoothing fraction"),
    ("oracle_aer.c", "tau[l] = ext1 * 730.0 * dzk;", "tau[l] = ext1 * 700.0 * dzk;", "free-troposphere density"),
    ("oracle_aer.c", "w12 = (1.0 - wi) * wj, w21 = wi * (1.0 - wj)", "w12 = wi * (1.0 - wj), w21 = (1.0 - wi) * wj",
     "bilinear weights on swapped corners"),
    ("oracle_aer.c", "a->rdrh[l] = (ih1 == ih2) ? 0.0 : drh1 / drh0;", "a->rdrh[l] = (ih1 == ih2) ? 0.0 : drh1 / 0.1;",
     "humidity weight over a fixed step"),
    ("oracle_aer.c", "            if (slm > 0.0) {\n                kp = 7;", "            if (slm >= 0.0) {\n                kp = 7;",
     "sea point taken as land in the mineral overlay"),
    ("oracle_aer.c", "tau[l] = ext1 * a->denn[1] * dzk;", "tau[l] = sca1 * a->denn[1] * dzk;",
     "mixing-layer depth from scattering"),
    ("oracle_aer.c", "const double es = fmin(pl_cb, 0.001 * a_fpvs(t));", "const double es = fmin(pl_cb, a_fpvs(t));",
     "saturation pressure in Pa against kPa"),
]


@pytest.mark.parametrize("fname,old,new,what", MUTATIONS, ids=[m[3] for m in MUTATIONS])
def test_planted_mistake_fails_an_aerosol_pin(tmp_path, fname, old, new, what):
    L = pyoracle.load(_build_copy(tmp_path, fname, old, new))
    failed = []
    for check in ALL:
        try:
            check(L)
        except AssertionError as e:
            failed.append(f"{check.__name__}: {e}")
    assert failed, f"no pin catches the planted mistake: {what}"
