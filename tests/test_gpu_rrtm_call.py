"""The NMMB hook: rad_driver.run -> rrtm_call.run(ctx, args) on a small NMMB-shaped state.

Checks the state -> column mapping (interface pressures from PT/PD/PDSG1/DSG2, interior rows only),
the outputs written against a direct column call on the same columns, the cadence and the
per-step zeroing of the driver."""
import math
from types import SimpleNamespace

import numpy as np
import pytest
import torch

from woof_nam_amd import gen, rad_driver, rrtm_call, solar
from woof_nam_amd.radiation import Radiation

pytestmark = pytest.mark.gpu

IM, JM, LM = 24, 20, 60


def make_ctx(device, nrads=3, nradl=3):
    rng = np.random.default_rng(0)
    c = gen.columns(IM * JM, seed=77)
    sg1, sg2, pdtop = gen._sigma()
    ps = c["p_lev"][-1]
    pd = ps - gen.PT - pdtop
    f32 = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(device)
    lay = lambda a: f32(a.reshape(LM, JM, IM))
    sfc = lambda a: f32(a.reshape(JM, IM))
    st = {"T": lay(c["t_lay"]), "Q": lay(c["q_lay"]), "O3": lay(c["o3_lay"]), "QC": lay(c["qc"]), "QI": lay(c["qi"]),
          "PD": sfc(pd), "TSKIN": sfc(c["t_sfc"]), "ALBEDO": sfc(c["alb"]), "SM": sfc(np.ones(IM * JM)),
          "GLAT": sfc(np.radians(np.repeat(np.linspace(20, 60, JM), IM))),
          "GLON": sfc(np.radians(np.tile(np.linspace(-120, -60, IM), JM))),
          "DSG2": f32(np.diff(sg2)), "SGML2": f32(0.5 * (sg2[1:] + sg2[:-1]))}
    for k in ("RLWTT", "RSWTT", "CLDFRA"):
        st[k] = torch.full((LM, JM, IM), 7.0, dtype=torch.float32, device=device)
    for k in ("RLWIN", "RSWIN", "RSWINC", "RSWOUT", "RLWTOA", "RSWTOA", "CZMEAN", "SIGT4", "CFRACL", "CFRACM",
              "CFRACH"):
        st[k] = torch.full((JM, IM), 7.0, dtype=torch.float32, device=device)
    grid = SimpleNamespace(im=IM, jm=JM, lm=LM, pt=gen.PT, pdsg1=np.diff(sg1) * pdtop, psgml1=np.zeros(LM))
    cfg = SimpleNamespace(nrads=nrads, nradl=nradl, nphs=1, shortwave="rrtm", longwave="rrtm", microphysics="gfs",
                          cldfraction="default", rrtm_precision="fp64")
    return SimpleNamespace(cfg=cfg, st=st, grid=grid, dt=60.0, ntimestep_rad=3,
                           tables={"JDAT": [2025, 3, 20, 0, 12, 0, 0, 0]}, stream=None)


def test_hook_matches_column_call(gpu_device):
    print("case rrtm_call hook", flush=True)
    ctx = make_ctx(gpu_device)
    captured = {}

    def hook(c, args):
        captured["args"] = args
        rrtm_call.run(c, args)

    rad_driver.run(ctx, rrtm_call=hook)
    torch.cuda.synchronize()
    args = captured["args"]
    cols, (c0, c1), _ = rrtm_call.columns(args, gpu_device)
    assert (c0, c1) == (IM, (JM - 1) * IM)
    sg1, sg2, pdtop = gen._sigma()
    pd = args["PD"].double().reshape(-1)[c0:c1]
    direct = torch.from_numpy(sg1).to(gpu_device)[:, None] * pdtop + torch.from_numpy(sg2).to(gpu_device)[:, None] \
        * pd[None, :] + gen.PT
    # the driver hands PDSG1/DSG2 over as float32 vectors (as the reference's driver does), so the
    # cumulative interface pressures agree with the float64 sigma formula to float32 rounding
    err = float((cols["p_lev"] - direct).abs().max())
    assert err < 1.0e-2, f"interface pressure mismatch {err:.3e} Pa"
    glat = args["GLAT"].reshape(-1)[c0:c1]
    glon = args["GLON"].reshape(-1)[c0:c1]
    cols["cosz"] = solar.coszen(args["JDAT"], 3 * 60.0, 3 * 60.0, glat, glon)
    rad = Radiation(0, max_cols=4096)
    out, _ = rad.run(cols, params=Radiation.default_params(mcica=0, solcon=args["SOLCON"]))
    torch.cuda.synchronize()
    st = ctx.st
    inner = lambda a: a.reshape(a.shape[0], -1)[:, c0:c1] if a.dim() == 3 else a.reshape(-1)[c0:c1]
    assert torch.equal(inner(st["RLWTT"]), out["hr_lw"].float())
    assert torch.equal(inner(st["RSWTT"]), out["hr_sw"].float())
    assert torch.equal(inner(st["RLWIN"]), out["flw_dn"][-1].float())
    assert torch.equal(inner(st["RSWOUT"]), out["fsw_up"][-1].float())
    assert torch.equal(inner(st["RSWINC"]), out["fsw_dn_sfc_clr"].float())
    assert torch.equal(inner(st["RLWTOA"]), out["flw_up"][0].float())
    # boundary rows untouched, zeroed fields zero outside the interior
    assert torch.all(st["RLWTT"][:, 0] == 7.0) and torch.all(st["RLWTT"][:, -1] == 7.0)
    assert torch.all(st["CFRACL"] == 0.0) and torch.all(st["CZMEAN"][0] == 0.0)
    # daytime: clear-sky surface SW at least the all-sky value where the sun is up
    day = inner(st["CZMEAN"]) > 0.05
    assert bool(day.any()) and torch.all(inner(st["RSWINC"])[day] >= inner(st["RSWIN"])[day] - 1e-3)
    assert float(inner(st["SIGT4"]).min()) > 200.0
    rad.close()


def test_cadence_and_zeroing(gpu_device):
    ctx = make_ctx(gpu_device, nrads=3, nradl=3)
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
