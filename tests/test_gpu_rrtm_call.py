"""The NMMB side of the call on the GPU against the oracle, bit for bit, on a Ferrier-style state:
total condensate CW with ice-only layers (F_ICE = 1), rain (F_RAIN > 0), layers below -40 C and humid
layers without condensate (the subgrid cloud).

* wr_nmmb_prep (driver prep, driver call, grrad setup, progcld2 with gethml) against oracle_nmmb_prep;
* wr_nmmb_post (driver post) against oracle_nmmb_post;
* the hook: rad_driver.run -> rrtm_call.run against the oracle chain (prep, the FP64 column call on
  the path-form columns, post), every state output;
* the driver's cadence and per-step zeroing."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import pyoracle
from woof_nam_amd import abi, aerosol, gen, ozone, rad_driver, rrtm_call, solar

pytestmark = pytest.mark.gpu

IM, JM, LM = 24, 20, 60


def ferrier_state(seed=0):
    """Host float32 NMMB state: generated profiles with Ferrier condensate fields."""
    rng = np.random.default_rng(seed)
    c = gen.columns(IM * JM, seed=77 + seed)
    sg1, sg2, pdtop = gen._sigma()
    pd = c["p_lev"][-1] - gen.PT - pdtop
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    lay = lambda a: f32(np.asarray(a).reshape(LM, JM, IM))
    t = lay(c["t_lay"])
    cw = lay(c["qc"] + c["qi"])
    cw[rng.uniform(size=cw.shape) < 0.05] = 3.0e-4                # extra cloudy layers
    f_ice = f32(rng.uniform(0, 1, cw.shape))
    f_ice[rng.uniform(size=cw.shape) < 0.3] = 1.0                  # ice-only layers
    f_rain = f32(rng.uniform(0, 1, cw.shape) * (rng.uniform(size=cw.shape) < 0.4))
    q = lay(c["q_lay"])
    humid = (cw == 0) & (rng.uniform(size=cw.shape) < 0.1)          # RH near saturation, no condensate
    q[humid] = q[humid] * 3.0
    return {"T": t, "Q": q, "O3": lay(c["o3_lay"]), "CW": cw, "F_ICE": f_ice, "F_RAIN": f_rain,
            "PD": f32(pd.reshape(JM, IM)), "GLAT": f32(np.radians(np.repeat(np.linspace(20, 60, JM), IM)).reshape(JM, IM)),
            "GLON": f32(np.radians(np.tile(np.linspace(-120, -60, IM), JM)).reshape(JM, IM)),
            "PDSG1": f32(np.diff(sg1) * pdtop), "DSG2": f32(np.diff(sg2)),
            "PSGML1": f32(0.5 * (sg1[1:] + sg1[:-1]) * pdtop + gen.PT), "SGML2": f32(0.5 * (sg2[1:] + sg2[:-1])),
            "TSKIN": f32(c["t_sfc"].reshape(JM, IM)), "ALBEDO": f32(c["alb"].reshape(JM, IM)),
            "SM": f32(np.zeros((JM, IM)))}   # land: the albedo is ALBEDO (the open-water table has its own test)


def oracle_state(h):
    m = {"pd": "PD", "pdsg1": "PDSG1", "dsg2": "DSG2", "psgml1": "PSGML1", "sgml2": "SGML2", "t": "T", "q": "Q",
         "o3": "O3", "cw": "CW", "f_ice": "F_ICE", "f_rain": "F_RAIN", "glat": "GLAT"}
    return {k: h[v] for k, v in m.items()}


def make_ctx(device, h, nrads=3, nradl=3):
    st = {k: torch.from_numpy(v).to(device) for k, v in h.items() if k not in ("PDSG1", "PSGML1")}
    for k in ("RLWTT", "RSWTT", "CLDFRA", "P_PHY"):
        st[k] = torch.full((LM, JM, IM), 7.0, dtype=torch.float32, device=device)
    for k in ("RLWIN", "RSWIN", "RSWINC", "RSWOUT", "RLWTOA", "RSWTOA", "CZMEAN", "SIGT4", "CFRACL", "CFRACM",
              "CFRACH", "ACFRST", "ACFRCV"):
        st[k] = torch.full((JM, IM), 7.0, dtype=torch.float32, device=device)
    for k in ("NCFRST", "NCFRCV"):
        st[k] = torch.full((JM, IM), 5, dtype=torch.int32, device=device)
    grid = SimpleNamespace(im=IM, jm=JM, lm=LM, pt=gen.PT, pdsg1=h["PDSG1"], psgml1=h["PSGML1"])
    cfg = SimpleNamespace(nrads=nrads, nradl=nradl, nphs=1, shortwave="rrtm", longwave="rrtm", microphysics="fer_hires",
                          cldfraction="default", rrtm_precision="fp64")
    return SimpleNamespace(cfg=cfg, st=st, grid=grid, dt=60.0, ntimestep_rad=3,
                           tables={"JDAT": [2025, 3, 20, 0, 12, 0, 0, 0]}, stream=None)


def dev_args(h, device):
    a = {k: torch.from_numpy(v).to(device) for k, v in h.items()}
    a["PT"] = float(np.float32(gen.PT))
    return a


def ozone_case(month, day, hour):
    """The climatology with gas_update's time factors of a date: (host dict for the oracle, WrOzone)."""
    tf = ozone.TimeFactors().update(month, day, hour)
    host = {**ozone.host_tables(), "k1oz": tf.k1oz, "k2oz": tf.k2oz, "facoz": float(tf.facoz)}
    return host, tf


@pytest.mark.parametrize("with_ozone", [False, True])
def test_prep_matches_oracle(gpu_device, with_ozone):
    print(f"case nmmb prep ozone={with_ozone}", flush=True)
    h = ferrier_state(1)
    llyr = rrtm_call.boundary_layer_top(rrtm_call.sigma_interfaces(dev_args(h, gpu_device), LM))
    args = dev_args(h, gpu_device)
    args["P_PHY"] = torch.full((LM, JM, IM), -1.0, dtype=torch.float32, device=gpu_device)
    oz_host, oz_dev = None, None
    if with_ozone:
        oz_host, tf = ozone_case(3, 7, 6)
        clim = ozone.Climatology(gpu_device)
        oz_dev = clim.struct(tf)
    cols, _ = rrtm_call.prep(args, gpu_device, (1, JM - 1), llyr, iovr=1, subgrid=True, ozone=oz_dev)
    torch.cuda.synchronize()
    ref = pyoracle.nmmb_prep(oracle_state(h), (IM, JM, LM, 1, JM - 1, np.float32(gen.PT)), (llyr, 1, 1), ozone=oz_host)
    if with_ozone:
        assert not np.array_equal(ref["o3_lay"], np.ascontiguousarray(h["O3"][:, 1:JM - 1].reshape(LM, -1), dtype=np.float64))
    for k in abi.NMMB_COL_FIELDS + ("clds", "mtop", "mbot", "o3_top"):
        a = cols[k].cpu().numpy()
        assert np.array_equal(a, ref[k]), f"{k}: {np.count_nonzero(a != ref[k])} elements differ"
    assert np.array_equal(args["P_PHY"].cpu().numpy(), ref["p_phy"])
    # the state exercises every branch: rain, ice-only, subgrid cloud
    assert (ref["crp"] > 0).any() and ((ref["cip"] > 0) & (ref["cwp"] == 0)).any() and (ref["cldf"] > 0).mean() > 0.05
    assert (ref["clds"][3] > 0).mean() > 0.3


def test_post_matches_oracle(gpu_device):
    print("case nmmb post", flush=True)
    h = ferrier_state(2)
    h["CW"][:, 5, 7] = 0.0                                       # a column without condensate
    args = dev_args(h, gpu_device)
    cols, _ = rrtm_call.prep(args, gpu_device, (1, JM - 1), 10)
    ncol = (JM - 2) * IM
    rng = np.random.default_rng(3)
    rad_h = {k: rng.uniform(-1, 300, s) for k, s in abi.output_shapes(ncol, LM).items()}
    rad_d = {k: torch.from_numpy(v).to(gpu_device) for k, v in rad_h.items()}
    cz = rng.uniform(0, 1, ncol)
    out_h = {k: np.full((LM, JM, IM), 7.0, np.float32) for k in ("rlwtt", "rswtt", "cldfra")}
    out_h.update({k: np.full((JM, IM), 7.0, np.float32) for k in abi.NMMB_OUT_FIELDS if k not in out_h})
    out_h.update({k: np.full((JM, IM), 5, np.int32) for k in ("ncfrst", "ncfrcv")})
    for k, v in out_h.items():
        args[k.upper()] = torch.from_numpy(v.copy()).to(gpu_device)
    for lw_step, sw_step in ((True, False), (False, True), (True, True)):
        rrtm_call.post(args, gpu_device, (1, JM - 1), cols, rad_d, torch.from_numpy(cz).to(gpu_device), lw_step, sw_step)
        ref_cols = {k: v.cpu().numpy() for k, v in cols.items() if k != "_keep"}
        pyoracle.nmmb_post(oracle_state(h), (IM, JM, LM, 1, JM - 1, np.float32(gen.PT)), ref_cols, rad_h, cz, lw_step,
                           sw_step, out_h)
        torch.cuda.synchronize()
        for k, v in out_h.items():
            a = args[k.upper()].cpu().numpy()
            assert np.array_equal(a, v), f"{k} (lw {lw_step}, sw {sw_step}): {np.count_nonzero(a != v)} differ"
    assert out_h["ncfrst"][3, 3] == 8 and (out_h["cldfra"][:, 5, 7] == 0).all()


def setaer_grid(seed):
    """Aerosol optics over the whole grid in setaer's layout [3][bands][LM+1][JM][IM] float32."""
    rng = np.random.default_rng(seed)
    out = {}
    for key, nb in (("rrtm_faersw", 14), ("rrtm_faerlw", 16)):
        a = np.zeros((3, nb, LM + 1, JM, IM), np.float32)
        a[0] = rng.uniform(0.0, 0.3, a[0].shape)
        a[1] = rng.uniform(0.6, 1.0, a[1].shape)
        a[2] = rng.uniform(0.5, 0.8, a[2].shape)
        out[key] = a
    return out


@pytest.mark.parametrize("aerosol_mode", ["none", "tables", "climatology"])
def test_hook_matches_oracle_chain(gpu_device, aerosol_mode):
    """rad_driver.run -> rrtm_call.run against the oracle chain on the same inputs: without aerosol, with
    setaer's optics left in ctx.tables, and with the OPAC climatology (the default)."""
    print(f"case rrtm_call hook vs oracle chain aerosol={aerosol_mode}", flush=True)
    h = ferrier_state(3)
    if aerosol_mode == "climatology":
        h["SICE"] = (np.random.default_rng(11).uniform(size=(JM, IM)) < 0.2).astype(np.float32)
    ctx = make_ctx(gpu_device, h)
    ctx.cfg.rrtm_aerosol = aerosol_mode
    aer = setaer_grid(9) if aerosol_mode == "tables" else {}
    for key, a in aer.items():
        ctx.tables[key] = torch.from_numpy(a).to(gpu_device)
    rad_driver.run(ctx, rrtm_call=rrtm_call.run)
    torch.cuda.synchronize()
    st = ctx.st
    # the oracle chain on the interior rows (J_S = 2, J_E = JM - 1)
    d = (IM, JM, LM, 1, JM - 1, np.float32(gen.PT))
    args = dev_args(h, gpu_device)
    llyr = rrtm_call.boundary_layer_top(rrtm_call.sigma_interfaces(args, LM))
    t = solar.current(ctx.tables["JDAT"], 0.0)        # JDAT is the time of the step
    cols = pyoracle.nmmb_prep(oracle_state(h), d, (llyr, 1, 1), ozone=ozone_case(t.month, t.day, t.hour)[0])
    c0, c1 = IM, (JM - 1) * IM
    glat = args["GLAT"].reshape(-1)[c0:c1]
    glon = args["GLON"].reshape(-1)[c0:c1]
    cz = solar.coszen(ctx.tables["JDAT"], 0.0, 3, 60.0, glat, glon)[0].cpu().numpy()
    flat = lambda a: np.ascontiguousarray(a.reshape(-1)[c0:c1], dtype=np.float64)
    rc = {k: cols[k] for k in abi.NMMB_COL_FIELDS + ("o3_top",)}
    rc.update({"cosz": cz, "t_sfc": flat(h["TSKIN"]), "alb": flat(h["ALBEDO"]), "emis": np.ones(c1 - c0),
               "col_id": np.arange(c0, c1, dtype=np.uint64)})
    for k, key in rrtm_call.AER_KEYS.items():
        if key in aer:
            a = aer[key]
            rc[k] = np.ascontiguousarray(a.reshape(3, a.shape[1], LM + 1, JM * IM)[..., c0:c1])
    if aerosol_mode == "climatology":
        slmsk = np.where(h["SICE"] > 0.5, 2.0, 1.0 - h["SM"]).astype(np.float32).reshape(-1)[c0:c1]
        rc["aer_sw"], rc["aer_lw"] = pyoracle.aer_optics(
            aerosol.host_tables(int(ctx.tables["JDAT"][1])), cols["p_lev"], cols["t_lay"], cols["q_lay"],
            h["GLON"].reshape(-1)[c0:c1], h["GLAT"].reshape(-1)[c0:c1], slmsk)
        assert (rc["aer_sw"][0].sum((0, 1)) > 0).all()
    from woof_nam_amd.radiation import Radiation
    solcon = ctx.tables["rrtm_call_flags"]["solcon"]
    rad, _ = pyoracle.run(rc, params=Radiation.default_params(mcica=0, solcon=solcon))
    out = {k: np.full((LM, JM, IM), 7.0, np.float32) for k in ("rlwtt", "rswtt", "cldfra")}
    out.update({k: np.full((JM, IM), 7.0, np.float32) for k in abi.NMMB_OUT_FIELDS if k not in out})
    out.update({k: np.full((JM, IM), 5, np.int32) for k in ("ncfrst", "ncfrcv")})
    for k in ("cldfra", "cfracl", "cfracm", "cfrach", "czmean", "sigt4"):
        out[k][...] = 0.0                                         # the driver zeroes these every step
    pyoracle.nmmb_post(oracle_state(h), d, cols, rad, cz, True, True, out)
    for k, v in out.items():
        a = st[k.upper()].cpu().numpy()
        assert np.array_equal(a, v), f"{k.upper()}: {np.count_nonzero(a != v)} elements differ from the oracle chain"
    assert np.array_equal(st["P_PHY"].cpu().numpy()[:, 1:JM - 1], cols["p_phy"][:, 1:JM - 1])
    # what the state shows: heating where there is cloud, ice-only and rainy layers radiated
    assert (st["CLDFRA"].cpu().numpy()[:, 1:-1] > 0).mean() > 0.05
    assert float(st["RSWTT"].abs().max()) > 0 and float(st["RLWTT"].abs().max()) > 0


def test_albedo_matches_oracle(gpu_device):
    """wr_nmmb_albedo against oracle_nmmb_albedo: ALBEDO over land and at night, the Payne open-water
    table over sea in daylight.  The zenith angle comes from the device's and the host's acos, which
    may differ in the last unit, so the albedo agrees to 1e-13."""
    print("case nmmb albedo", flush=True)
    rng = np.random.default_rng(8)
    h = ferrier_state(5)
    h["SM"] = (rng.uniform(size=(JM, IM)) < 0.5).astype(np.float32)
    args = dev_args(h, gpu_device)
    nc = (JM - 2) * IM
    cz = rng.uniform(-0.2, 1.0, nc)
    got = rrtm_call.albedo(args, gpu_device, (1, JM - 1), torch.from_numpy(cz).to(gpu_device))
    torch.cuda.synchronize()
    ref = pyoracle.nmmb_albedo(h["ALBEDO"], h["SM"], cz, (1, JM - 1))
    assert np.abs(got.cpu().numpy() - ref).max() <= 1e-13
    water_day = (h["SM"][1:JM - 1].reshape(-1) > 0.5) & (cz > 0)
    assert water_day.sum() > 20 and np.any(ref[water_day] != h["ALBEDO"][1:JM - 1].reshape(-1)[water_day])


def test_cadence_and_zeroing(gpu_device):
    h = ferrier_state(4)
    ctx = make_ctx(gpu_device, h, nrads=3, nradl=3)
    calls = []
    ctx.ntimestep_rad = 4                       # not a radiation step
    rad_driver.run(ctx, rrtm_call=lambda c, a: calls.append(a))
    assert not calls
    assert torch.all(ctx.st["SIGT4"] == 0.0) and torch.all(ctx.st["RLWTT"] == 7.0)
    ctx.ntimestep_rad = 1                       # first step always radiates
    rad_driver.run(ctx, rrtm_call=lambda c, a: calls.append(a))
    assert len(calls) == 1 and calls[0]["J_S"] == 2 and calls[0]["J_E"] == JM - 1
    with pytest.raises(KeyError):
        a = dict(calls[0])
        del a["PD"]
        rrtm_call.run(ctx, {**a, "NTIMESTEP": 3})
