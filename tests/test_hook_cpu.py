"""Host logic of the NMMB hook on CPU tensors (no GPU): the driver's cadence and per-step zeroing, the
boundary-layer cloud top, the checks of the state arguments, the tile sized from free memory and the
refused McICA route."""
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from woof_nam_amd import gen, rad_driver, rrtm_call

IM, JM, LM = 12, 9, 60
CPU = torch.device("cpu")


def make_ctx(nrads=3, nradl=3):
    c = gen.columns(IM * JM, seed=5)
    sg1, sg2, pdtop = gen._sigma()
    pd = c["p_lev"][-1] - gen.PT - pdtop
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32))
    lay = lambda a: f32(a.reshape(LM, JM, IM))
    sfc = lambda a: f32(a.reshape(JM, IM))
    st = {"T": lay(c["t_lay"]), "Q": lay(c["q_lay"]), "O3": lay(c["o3_lay"]), "QC": lay(c["qc"]), "QI": lay(c["qi"]),
          "CW": lay(c["qc"] + c["qi"]), "F_ICE": lay(np.zeros(LM * IM * JM)), "F_RAIN": lay(np.zeros(LM * IM * JM)),
          "PD": sfc(pd), "TSKIN": sfc(c["t_sfc"]), "ALBEDO": sfc(c["alb"]), "SM": sfc(np.ones(IM * JM)),
          "GLAT": sfc(np.radians(np.repeat(np.linspace(20, 60, JM), IM))),
          "GLON": sfc(np.radians(np.tile(np.linspace(-120, -60, IM), JM))),
          "DSG2": f32(np.diff(sg2)), "SGML2": f32(0.5 * (sg2[1:] + sg2[:-1]))}
    for k in ("RLWTT", "RSWTT", "CLDFRA"):
        st[k] = torch.full((LM, JM, IM), 7.0)
    for k in ("RLWIN", "RSWIN", "RSWINC", "RSWOUT", "RLWTOA", "RSWTOA", "CZMEAN", "SIGT4", "CFRACL", "CFRACM",
              "CFRACH", "ACFRST", "ACFRCV"):
        st[k] = torch.full((JM, IM), 7.0)
    for k in ("NCFRST", "NCFRCV"):
        st[k] = torch.zeros((JM, IM), dtype=torch.int32)
    grid = SimpleNamespace(im=IM, jm=JM, lm=LM, pt=gen.PT, pdsg1=np.diff(sg1) * pdtop, psgml1=np.zeros(LM))
    cfg = SimpleNamespace(nrads=nrads, nradl=nradl, nphs=1, shortwave="rrtm", longwave="rrtm", microphysics="gfs",
                          cldfraction="default")
    return SimpleNamespace(cfg=cfg, st=st, grid=grid, dt=60.0, ntimestep_rad=3,
                           tables={"JDAT": [2025, 3, 20, 0, 12, 0, 0, 0]}, stream=None)


def test_cadence_and_zeroing():
    ctx = make_ctx(nrads=3, nradl=3)
    calls = []
    ctx.ntimestep_rad = 4                       # not a radiation step: zeroing only
    rad_driver.run(ctx, rrtm_call=lambda c, a: calls.append(a))
    assert not calls
    for k in rad_driver.ZEROED:
        assert torch.all(ctx.st[k] == 0.0), k
    assert torch.all(ctx.st["RLWTT"] == 7.0)
    for n in (1, 3, 6):                          # the first step and every NRADS-th step radiate
        ctx.ntimestep_rad = n
        rad_driver.run(ctx, rrtm_call=lambda c, a: calls.append(a))
    assert [a["NTIMESTEP"] for a in calls] == [1, 3, 6]
    a = calls[0]
    assert (a["J_S"], a["J_E"], a["I_S"], a["I_E"]) == (2, JM - 1, 1, IM)
    assert a["T"] is ctx.st["T"]                 # aliased to the live state, not copied


def test_unsupported_schemes_raise():
    ctx = make_ctx()
    ctx.cfg.shortwave = "gfdl"
    with pytest.raises(NotImplementedError):
        rad_driver.run(ctx, rrtm_call=lambda c, a: None)
    ctx = make_ctx()
    del ctx.st["TSKIN"]
    ctx.ntimestep_rad = 1
    with pytest.raises(KeyError):
        rad_driver.run(ctx, rrtm_call=lambda c, a: None)


def test_boundary_layer_top_rule():
    """cld_init's llyr (radiation_clouds_nmmb.f:327-344, ivflip = 0): the last k = LM..2 before the first
    sigma interface below 0.9, counting from the surface."""
    si = np.linspace(0.0, 1.0, 61)                   # si(k) = (k - 1) / 60
    assert rrtm_call.boundary_layer_top(si) == 54    # si(54) = 0.883 < 0.9 <= si(55)
    assert rrtm_call.boundary_layer_top(np.full(61, 0.95)) == 2     # none below 0.9: the loop ends at 2
    ctx = make_ctx()
    ctx.ntimestep_rad = 1
    got = []
    rad_driver.run(ctx, rrtm_call=lambda c, a: got.append(a))
    si = rrtm_call.sigma_interfaces(got[0], LM)
    assert si[0] == 0.0 and si[-1] == pytest.approx(1.0) and np.all(np.diff(si) > 0)
    llyr = rrtm_call.boundary_layer_top(si)
    assert si[llyr - 1] < 0.9 <= si[llyr] and 40 < llyr < LM


def test_state_arguments_checked_before_any_device_work():
    ctx = make_ctx()
    ctx.ntimestep_rad = 1
    got = []
    rad_driver.run(ctx, rrtm_call=lambda c, a: got.append(a))
    args = got[0]
    st, keep, dims = rrtm_call.nmmb_state(args, CPU, (1, JM - 1))
    assert dims == (IM, JM, LM) and (st.j0, st.j1) == (1, JM - 1)
    with pytest.raises(ValueError):
        rrtm_call.nmmb_state({**args, "CW": args["CW"].double()}, CPU, (1, JM - 1))     # dtype
    with pytest.raises(ValueError):
        rrtm_call.nmmb_state({**args, "PD": args["PD"][:-1]}, CPU, (1, JM - 1))          # shape
    with pytest.raises(KeyError):
        rrtm_call.nmmb_state({k: v for k, v in args.items() if k != "F_RAIN"}, CPU, (1, JM - 1))


def test_tile_sized_from_free_memory():
    """The hook's context tile fits the free device memory beside the call's own column arrays
    (the reference sizes its batches from free bytes, rrtm_call.py:183-203)."""
    from woof_nam_amd import _lib
    L = _lib.load()
    ncol, lm = 796590, 60
    per_call = rrtm_call.call_bytes_per_column(lm) * ncol
    whole = L.wr_context_bytes(ncol)
    assert whole > 20e9                                 # tens of KB per column of work space
    assert rrtm_call.tile_columns(ncol, lm, int((whole + per_call) / 0.9) + 1) == ncol
    free = int(0.5 * (whole + per_call))                # half of what the whole grid needs
    m = rrtm_call.tile_columns(ncol, lm, free)
    assert 4096 <= m < ncol
    assert L.wr_context_bytes(m) + per_call <= 0.9 * free
    assert L.wr_context_bytes(m + 2048) + per_call > 0.9 * free   # within the bisection step
    with pytest.raises(MemoryError):
        rrtm_call.tile_columns(ncol, lm, int(per_call / 0.9))
    assert L.wr_context_bytes(0) == 0


def test_mcica_route_refused_by_the_hook():
    ctx = make_ctx()
    ctx.cfg.rrtm_mcica = 1
    args = rad_driver.rrtm_args(ctx, ctx.tables["JDAT"], 3, 60.0, rad_driver.MP_FLAGS["gfs"], 1361.0)
    with pytest.raises(NotImplementedError, match="RSWINC"):
        rrtm_call.run(ctx, args)


def test_restart_state_round_trip():
    """A run resumed from the hook's saved state takes the same ozone time factors as the run that went
    on: gas_update's months change only at mid-month, so the saved months matter between change days."""
    import json
    from types import SimpleNamespace
    from woof_nam_amd import ozone
    ctx = SimpleNamespace(tables={})
    tf = ozone.TimeFactors()
    for day in range(1, 13):                          # the run so far: January 1 to 12
        tf.update(1, day, 6)
    ctx.tables[rrtm_call.KEY_OZONE] = (None, tf)
    ctx.tables[rrtm_call.KEY_LLYR] = 57
    saved = json.loads(json.dumps(rrtm_call.state_dict(ctx)))
    resumed = SimpleNamespace(tables={})
    rrtm_call.load_state_dict(resumed, saved)
    a, b = tf, resumed.tables[rrtm_call.KEY_OZONE][1]
    for month, day in ((1, 13), (1, 15), (1, 16), (1, 30), (2, 14), (2, 15)):
        a.update(month, day, 6)
        b.update(month, day, 6)
        assert (a.k1oz, a.k2oz, a.facoz) == (b.k1oz, b.k2oz, b.facoz)
    assert resumed.tables[rrtm_call.KEY_LLYR] == 57
    # without the saved state a resumed run would pick other months: with no call at hour 0 of the
    # middle day, the running state keeps December / January where a first call takes January / February
    fresh = ozone.TimeFactors().update(1, 30, 6)
    assert (fresh.k1oz, fresh.k2oz) == (1, 2) and (b.k1oz, b.k2oz) == (12, 1)
    with pytest.raises(ValueError):
        rrtm_call.load_state_dict(resumed, {"format": 9})


def test_clock_gives_the_step_time():
    """JDAT handed to the hook is the time at the start of the step (phys_rad_driver.py:76-94): start +
    NTIMESTEP DT with a clock, the context's own JDAT as given without one."""
    import datetime as dt
    ctx = make_ctx()
    calls = []
    ctx.ntimestep_rad = 3
    rad_driver.run(ctx, rrtm_call=lambda c, a: calls.append(a))
    assert calls[-1]["JDAT"] == [2025, 3, 20, 0, 12, 0, 0, 0]
    ctx.start = dt.datetime(2025, 3, 31, 23, 0, 0)
    ctx.ntimestep_rad = 90
    rad_driver.run(ctx, rrtm_call=lambda c, a: calls.append(a))
    assert calls[-1]["JDAT"] == [2025, 4, 1, 0, 0, 30, 0, 0] == ctx.tables["JDAT"]


def test_aerosol_mode_checked():
    ctx = make_ctx()
    ctx.cfg.rrtm_aerosol = "volcanic"
    with pytest.raises(ValueError, match="rrtm_aerosol"):
        rrtm_call.aerosols(ctx, {}, {}, LM, JM, IM, 0, IM, CPU)
    ctx.cfg.rrtm_aerosol = "tables"
    with pytest.raises(KeyError, match="rrtm_faersw"):
        rrtm_call.aerosols(ctx, {}, {}, LM, JM, IM, 0, IM, CPU)
    ctx.cfg.rrtm_aerosol = "none"
    assert rrtm_call.aerosols(ctx, {}, {}, LM, JM, IM, 0, IM, CPU) == {}
