"""Pins of the CPU FP64 oracle: known-answer vectors, closed forms and invariants.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a wrong index, sign or
operand order) fails: the Philox known-answer vectors of Random123, the look-up tables against
their closed forms evaluated by numpy, the McICA invariants of the reference's own check
(tools/check-phys-rrtm-lw-mcica-subcol.py:59-71), the CAL_CLDFRA truth table
(phys_rad_cal_cldfra.cu:43-53) and the band tables of phys_rrtm_lw_rtrnmr.py:83-95.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle
from woof_nam_amd import gen


def ulps(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return np.abs(a.view(np.int64) - b.view(np.int64))


def test_exp_log_pow_against_libm():
    L = pyoracle.lib()
    rng = np.random.default_rng(0)
    xs = np.concatenate([rng.uniform(-700, 700, 4000), rng.uniform(-1, 1, 4000), [-745.0, 0.0, 1e-300, 88.0]])
    e = np.array([L.oracle_exp(float(x)) for x in xs])
    assert ulps(e, np.exp(xs)).max() <= 1
    ys = np.concatenate([np.exp(rng.uniform(-700, 700, 4000)), rng.uniform(0.5, 2.0, 4000), [1.0, 2.0, 1e-310]])
    lg = np.array([L.oracle_log(float(y)) for y in ys])
    assert ulps(lg, np.log(ys)).max() <= 1
    assert L.oracle_log(1.0) == 0.0 and L.oracle_exp(0.0) == 1.0
    bs = rng.uniform(0.01, 100, 2000)
    ps = rng.uniform(-3, 3, 2000)
    pw = np.array([L.oracle_pow(float(b), float(p)) for b, p in zip(bs, ps)])
    # pow = exp(y * log x) in plain double: the rounding of y*log(x) (|y log x| < 14 here) is magnified
    # by |y log x| in the result, so the bound is a few tens of ulps rather than one
    assert ulps(pw, np.power(bs, ps)).max() <= 32


@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox4x32_10_known_answers(ctr, key, expect):
    assert tuple(int(v) for v in pyoracle.philox(ctr, key)) == expect


def test_lw_lookup_tables_closed_form():
    e, f, t = pyoracle.lw_tables()
    i = np.arange(1, 10000, dtype=np.float64)
    tau = (1.0 / 0.278) * (i / (10000.0 - i))
    assert np.array_equal(t[1:10000], tau)
    ex = np.where(tau >= 87.0, 1.0e-20, np.exp(-tau))
    assert ulps(e[1:10000], ex).max() <= 1
    tfn = np.where(tau < 0.06, tau / 6.0, 1.0 - 2.0 * (1.0 / tau - ex / (1.0 - ex)))
    rel = np.abs(f[1:10000] - tfn) / np.maximum(np.abs(tfn), 1e-300)
    assert rel.max() < 1e-10          # cancellation near tau = 0.06 magnifies 1-ulp exp differences
    e1 = e[1:10000]                   # with the table's own exp the formula is reproduced exactly
    assert np.array_equal(f[1:10000], np.where(tau < 0.06, tau / 6.0, 1.0 - 2.0 * (1.0 / tau - e1 / (1.0 - e1))))
    assert (e[0], t[0], f[0]) == (1.0, 0.0, 0.0)
    assert (e[10000], t[10000], f[10000]) == (1.0e-20, 1.0e10, 1.0)
    # tfn is the Pade source function: increasing from 0 towards 1
    assert np.all(np.diff(f) >= 0.0)


def test_sw_lookup_table_closed_form():
    e = pyoracle.sw_table()
    i = np.arange(1, 10000, dtype=np.float64)
    assert ulps(e[1:10000], np.exp(-(1.0 / 0.278) * (i / (10000.0 - i)))).max() <= 1
    assert e[0] == 1.0 and e[10000] == 1.0e-20


def test_mcica_invariants():
    rng = np.random.default_rng(12345)
    nl = 61
    for kind, ng in ((0, 140), (1, 112)):
        for iovr in (0, 1, 2):
            assert not pyoracle.mcica(3, 7, kind, ng, iovr, np.zeros(nl)).any()        # clear: no bits
            full = pyoracle.mcica(3, 7, kind, ng, iovr, np.ones(nl))                   # overcast: all bits
            bits = np.unpackbits(full.view(np.uint8), bitorder="little").reshape(nl, -1)[:, :ng]
            assert bits.all()
    # mean sub-column cloud fraction within 0.01 of cldf (1000 columns x 140 g)
    tot, want = 0.0, 0.0
    for col in range(300):
        cf = rng.random(nl)
        b = pyoracle.mcica(col, 99, 0, 140, 1, cf)
        bits = np.unpackbits(b.view(np.uint8), bitorder="little").reshape(nl, -1)[:, :140]
        tot += bits.mean()
        want += cf.mean()
    assert abs(tot / 300 - want / 300) < 0.01


def test_mcica_overlap_rules():
    nl = 61
    cf = np.full(nl, 0.4)
    # maximum-random with equal fractions in adjacent layers: cloud below implies cloud above
    b = pyoracle.mcica(11, 5, 0, 140, 1, cf)
    bits = np.unpackbits(b.view(np.uint8), bitorder="little").reshape(nl, -1)[:, :140].astype(bool)
    assert np.all(~bits[:-1] | bits[1:])
    # maximum overlap (iovr 2): one draw per g-point, so cloudy layers nest by fraction
    cf2 = np.linspace(0.05, 0.95, nl)
    b2 = pyoracle.mcica(11, 5, 1, 112, 2, cf2)
    bits2 = np.unpackbits(b2.view(np.uint8), bitorder="little").reshape(nl, -1)[:, :112].astype(bool)
    assert np.all(~bits2[:-1] | bits2[1:])
    # random overlap (iovr 0): adjacent layers independent, joint frequency ~ cf^2
    joint = 0.0
    for col in range(200):
        b0 = pyoracle.mcica(col, 1, 0, 140, 0, cf)
        bb = np.unpackbits(b0.view(np.uint8), bitorder="little").reshape(nl, -1)[:, :140].astype(bool)
        joint += (bb[:-1] & bb[1:]).mean()
    assert abs(joint / 200 - 0.16) < 0.01
    # distinct columns and seeds give distinct streams
    assert not np.array_equal(pyoracle.mcica(1, 5, 0, 140, 1, cf), pyoracle.mcica(2, 5, 0, 140, 1, cf))
    assert not np.array_equal(pyoracle.mcica(1, 5, 0, 140, 1, cf), pyoracle.mcica(1, 6, 0, 140, 1, cf))


def test_cal_cldfra_truth_table():
    lm, jm, im = 3, 4, 5
    qc = np.zeros((lm, jm, im), np.float32)
    qi = np.zeros((lm, jm, im), np.float32)
    thr = np.float32(1.0e-6)
    qc[0] = np.nextafter(thr, np.float32(1))       # just above
    qc[1] = thr                                    # equal: not cloudy
    qc[2] = np.float32(6e-7)
    qi[2] = np.float32(6e-7)                       # sum above only with qi counted
    out = np.full((lm, jm, im), -1.0, np.float32)
    pyoracle.cal_cldfra(out, qc, qi, 1, im - 2, 1, jm - 2, True, True)
    inner = out[:, 1:jm - 1, 1:im - 1]
    assert np.all(inner[0] == 1.0) and np.all(inner[1] == 0.0) and np.all(inner[2] == 1.0)
    assert np.all(out[:, 0, :] == -1.0) and np.all(out[:, :, 0] == -1.0)
    assert np.all(out[:, -1, :] == -1.0) and np.all(out[:, :, -1] == -1.0)
    pyoracle.cal_cldfra(out, qc, qi, 1, im - 2, 1, jm - 2, True, False)       # qc only
    assert np.all(out[2, 1:jm - 1, 1:im - 1] == 0.0) and np.all(out[0, 1:jm - 1, 1:im - 1] == 1.0)
    pyoracle.cal_cldfra(out, qc, qi, 1, im - 2, 1, jm - 2, False, False)      # no cloud water species
    assert np.all(out[:, 1:jm - 1, 1:im - 1] == 0.0)


def test_band_tables():
    from woof_nam_amd.abi import NGPT_LW, NGPT_SW
    ns = [0, 10, 22, 38, 52, 68, 76, 88, 96, 108, 114, 122, 130, 134, 136, 138]
    ng = [10, 12, 16, 14, 16, 8, 12, 8, 12, 6, 8, 8, 4, 2, 2, 2]
    assert sum(ng) == NGPT_LW and all(ns[b] + ng[b] == (ns[b + 1] if b < 15 else NGPT_LW) for b in range(16))
    assert sum([6, 12, 8, 8, 10, 10, 2, 10, 8, 6, 6, 8, 6, 12]) == NGPT_SW


# ---- physics invariants of whole columns -------------------------------------------------------
HEATFAC = 9.80665 * 1.0e-2 / 1.0046e3
SIGMA = 5.670374419e-8


def _run(cols, **kw):
    from woof_nam_amd.abi import make_params
    out, _ = pyoracle.run(cols, params=make_params(**kw), nthreads=4)
    return out


def test_heating_rate_is_flux_divergence():
    cols = gen.columns(64, seed=4)
    out = _run(cols)
    plev_hpa = cols["p_lev"] * 1.0e-2
    dp = plev_hpa[1:] - plev_hpa[:-1]                                  # [60][n] model layers, top first
    for band, up, dn in (("hr_lw", "flw_up", "flw_dn"), ("hr_sw", "fsw_up", "fsw_dn")):
        net = out[up] - out[dn]                                        # levels top first, [62][n]
        hr = (net[2:] - net[1:-1]) * HEATFAC / dp if band == "hr_lw" else (net[2:] - net[1:-1]) * HEATFAC / dp
        np.testing.assert_allclose(out[band], hr, rtol=1e-9, atol=1e-15)


def test_toa_solar_flux_and_night():
    cols = gen.columns(200, seed=6)
    out = _run(cols)
    cz = cols["cosz"]
    day = cz >= 1.0e-4
    toa = out["fsw_dn"][0]
    np.testing.assert_allclose(toa[day], 1368.22 * cz[day], rtol=2e-3)
    assert np.all(out["fsw_dn"][:, ~day] == 0.0) and np.all(out["hr_sw"][:, ~day] == 0.0)
    assert np.all(out["fsw_up"][:, day] >= 0.0) and np.all(out["fsw_dn"][:, day] >= 0.0)
    # no flux above the added layer is reflected back as more than arrives
    assert np.all(out["fsw_up"][0, day] < toa[day])


def test_transparent_atmosphere_emits_sigma_t4():
    cols = gen.columns(32, seed=8)
    cols["q_lay"][:] = 1.0e-12
    cols["o3_lay"][:] = 0.0
    cols["cldf"][:] = 0.0
    cols["emis"][:] = 1.0
    out = _run(cols, vmr=(0.0,) * 10)
    olr = out["flw_up"][0]
    st4 = SIGMA * cols["t_sfc"] ** 4
    # black surface: the 16 band Planck integrals (delwave x totplnk, x pi x 1e4) recover sigma T^4
    assert np.all(np.abs(out["flw_up"][-1] - st4) / st4 < 1.0e-3)
    # what is left to absorb (the 1e-7 kg/kg humidity floor in the water rotation band, N2 continuum)
    # takes well under a tenth of it
    ratio = olr / out["flw_up"][-1]
    assert np.all(ratio > 0.9) and np.all(ratio < 1.0), (ratio.min(), ratio.max())
    # adding CO2 can only lower the outgoing flux
    out2 = _run(cols, vmr=(410.0e-6,) + (0.0,) * 9)
    assert np.all(out2["flw_up"][0] < olr - 5.0)


def test_isothermal_atmosphere_upward_flux_is_black_body():
    """Black surface under an isothermal atmosphere: each layer re-emits what it absorbs, so the
    upward flux is sigma T^4 at every level (to the spread of the per-layer Planck fractions),
    clear or overcast, while space above draws the downward flux at the top to zero."""
    cols = gen.columns(32, seed=9)
    cols["t_lay"][:] = 250.0
    cols["t_sfc"][:] = 250.0
    cols["emis"][:] = 1.0
    st4 = SIGMA * 250.0 ** 4
    for cf in (0.0, 1.0):
        cols["cldf"][:] = cf
        out = _run(cols)
        assert np.abs(out["flw_up"] / st4 - 1.0).max() < 0.01
        assert np.all(out["flw_dn"][0] == 0.0)
        assert np.all(out["flw_dn"][-1] > out["flw_dn"][1])       # emission accumulates downward
