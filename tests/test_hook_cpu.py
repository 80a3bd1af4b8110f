"""Host logic of the NMMB hook on CPU tensors (no GPU): the driver's cadence and per-step zeroing,
and the state -> column mapping of rrtm_call.columns (interface pressures, interior rows, the NAM
condensate cloud fraction, column ids)."""
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
          "PD": sfc(pd), "TSKIN": sfc(c["t_sfc"]), "ALBEDO": sfc(c["alb"]), "SM": sfc(np.ones(IM * JM)),
          "GLAT": sfc(np.radians(np.repeat(np.linspace(20, 60, JM), IM))),
          "GLON": sfc(np.radians(np.tile(np.linspace(-120, -60, IM), JM))),
          "DSG2": f32(np.diff(sg2)), "SGML2": f32(0.5 * (sg2[1:] + sg2[:-1]))}
    for k in ("RLWTT", "RSWTT", "CLDFRA"):
        st[k] = torch.full((LM, JM, IM), 7.0)
    for k in ("RLWIN", "RSWIN", "RSWINC", "RSWOUT", "RLWTOA", "RSWTOA", "CZMEAN", "SIGT4", "CFRACL", "CFRACM",
              "CFRACH"):
        st[k] = torch.full((JM, IM), 7.0)
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


def test_columns_mapping():
    ctx = make_ctx()
    ctx.ntimestep_rad = 1
    got = []
    rad_driver.run(ctx, rrtm_call=lambda c, a: got.append(a))
    args = got[0]
    cols, (c0, c1), (im, jm, lm) = rrtm_call.columns(args, CPU)
    assert (c0, c1) == (IM, (JM - 1) * IM) and (im, jm, lm) == (IM, JM, LM)
    # interface pressures: PT + cumulative (PDSG1 + DSG2 PD) over the interior rows
    pd = ctx.st["PD"].double().reshape(-1)[c0:c1].numpy()
    pdsg1 = np.asarray(args["PDSG1"].double())
    dsg2 = ctx.st["DSG2"].double().numpy()
    ref = gen.PT + np.concatenate([np.zeros((1, c1 - c0)), np.cumsum(pdsg1[:, None] + dsg2[:, None] * pd[None, :], 0)])
    assert np.allclose(cols["p_lev"].numpy(), ref, rtol=0, atol=1e-6)
    assert np.all(np.diff(cols["p_lev"].numpy(), axis=0) > 0)
    # NAM condensate cloud fraction (progcld2, ivflip = 0): min(1, 1e5 clw p/1000) where clw > 1e-8 p/1000
    pl = 0.5 * (ref[1:] + ref[:-1]) * 1.0e-2
    clw = (ctx.st["QC"].double().reshape(LM, -1)[:, c0:c1] + ctx.st["QI"].double().reshape(LM, -1)[:, c0:c1]).numpy()
    cf = np.where(clw > 1.0e-8 * (pl * 1.0e-3), np.minimum(1.0e5 * clw * pl * 1.0e-3, 1.0), 0.0)
    assert np.allclose(cols["cldf"].numpy(), cf, rtol=1e-12, atol=0)
    assert 0.0 < float(cols["cldf"].mean()) < 1.0
    # layer fields are the interior rows of the state, in float64, top first
    t = ctx.st["T"].double().reshape(LM, -1)[:, c0:c1]
    assert torch.equal(cols["t_lay"], t)
    assert torch.equal(cols["col_id"], torch.arange(c0, c1))
    assert torch.all(cols["emis"] == 1.0)


def test_columns_rejects_bad_rows_and_shapes():
    ctx = make_ctx()
    ctx.ntimestep_rad = 1
    got = []
    rad_driver.run(ctx, rrtm_call=lambda c, a: got.append(a))
    args = dict(got[0])
    with pytest.raises(ValueError):
        rrtm_call.columns({**args, "J_S": JM + 3}, CPU)
    with pytest.raises(ValueError):
        rrtm_call.columns({**args, "PD": args["PD"][:-1]}, CPU)


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
