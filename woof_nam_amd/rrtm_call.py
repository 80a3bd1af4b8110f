"""The NMMB radiation hook over the HIP library: ``run(ctx, args)``.

Same call contract as the reference's ``namgpu.physics.rrtm.rrtm_call.run(ctx, args)``
(namgpu/physics/rrtm/rrtm_call.py:246, called by phys_rad_driver.py:255-262): ``args`` is the
dict of Fortran RRTM argument names that the radiation driver builds (phys_rad_driver.py:108-168),
holding device arrays aliased to the model state ([LM, JM, IM] float32, level 1 at the model top,
or [JM, IM]); the hook writes the radiation outputs into the output arrays of ``args`` in place and
returns None.  The input species arrays are only read.

What it does on a call:
  * cadence: SW on NTIMESTEP % NRADS == 0, LW on NTIMESTEP % NRADL == 0 (rrtm_call.py:253-254);
  * driver prep: interface pressures P8W = PT + cumulative (PDSG1 + DSG2 * PD)
    (phys_rrtm_driver_prep.cu:82), the interior rows J_S..J_E only (phys_rad_driver.py:165-166);
  * clouds: with CLD_FRACTION = 0 the NAM condensate cloud fraction of progcld2
    (phys_rrtm_progcld2.cu:218-231, ivflip = 0: min(1, 1e5 clw p/1000) where clw > 1e-8 p/1000,
    p in hPa) from cloud water QCW and ice QCI; otherwise the CLDFRA field.  Effective radii are
    the call parameters (re_liq, re_ice), not progcld2's diagnosed radii;
  * solar geometry: cos zenith averaged over the SW interval (woof_nam_amd.solar);
  * the radiation call (woof_nam_amd.radiation, the NAM route mcica = 0 unless ctx.cfg says
    otherwise; FP64 unless ctx.cfg.rrtm_precision = "mixed");
  * driver post (phys_rrtm_driver_post.cu:71-150): RLWTT, RSWTT (K/s), RLWIN, RSWIN, RSWOUT,
    RSWINC, RLWTOA, RSWTOA, CZMEAN, and SIGT4 = sigma T(LM)^4 on LW steps.
The cloud-cover diagnostics (CFRACL/M/H and their accumulators) are not produced.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib, abi, solar
from .radiation import Radiation

KEY_CTX = "woof_nam_amd.radiation"
STBOLT = 5.67051e-8
SW_OUT = ("RSWTT", "RSWIN", "RSWOUT", "RSWINC", "RSWTOA", "CZMEAN")
LW_OUT = ("RLWTT", "RLWIN", "RLWTOA", "SIGT4")


def _arr(args, name, shape=None):
    if name not in args:
        raise KeyError(f"rrtm-call: argument {name!r} missing (its producer has not run)")
    a = args[name]
    if not isinstance(a, torch.Tensor):
        a = torch.as_tensor(np.asarray(a))
    if shape is not None and tuple(a.shape) != tuple(shape):
        raise ValueError(f"rrtm-call: {name} has shape {tuple(a.shape)}, expected {tuple(shape)}")
    return a


def _cfg(ctx, name, default):
    cfg = getattr(ctx, "cfg", None)
    return getattr(cfg, name, default) if cfg is not None else default


# device bytes per column the hook itself holds during a call besides the library context: the FP64
# column inputs (6 layer fields + interfaces + 4 surface values) and outputs (2 heating rates, 4 flux
# profiles, 1 surface value) at LM layers, with the progcld2 paths and radii (5 layer fields)
def call_bytes_per_column(lm: int) -> int:
    return 8 * ((6 + 5) * lm + (lm + 1) + 4 + 2 * lm + 4 * (lm + 2) + 1)


def tile_columns(ncol: int, lm: int, free_bytes: int, cap: int = 1048576, margin: float = 0.9) -> int:
    """Largest tile (columns per library context) that fits the free device memory beside the call's
    own arrays, as namgpu/physics/rrtm/rrtm_call.py:183-203 (batch_count) sizes its batches from free
    bytes.  Raises MemoryError when not even a 4,096-column tile fits."""
    L = _lib.load()
    avail = int(margin * free_bytes) - call_bytes_per_column(lm) * ncol
    hi = max(1, min(ncol, cap))
    if L.wr_context_bytes(hi) <= avail:
        return hi
    lo = min(4096, hi)
    if L.wr_context_bytes(lo) > avail:
        raise MemoryError(f"rrtm-call: {ncol} columns need {call_bytes_per_column(lm) * ncol / 1e9:.1f} GB of "
                          f"column arrays and a {lo}-column context of {L.wr_context_bytes(lo) / 1e9:.1f} GB; "
                          f"{free_bytes / 1e9:.1f} GB free")
    while hi - lo > 1024:                       # context bytes grow with the tile: bisect
        mid = (lo + hi) // 2
        if L.wr_context_bytes(mid) <= avail:
            lo = mid
        else:
            hi = mid
    return lo


def _context(ctx, device, ncol, lm):
    tables = ctx.tables
    rad = tables.get(KEY_CTX)
    if rad is None or rad.device != device:
        free, _ = torch.cuda.mem_get_info(device)
        m = tile_columns(max(ncol, 1), lm, free, cap=int(_cfg(ctx, "rrtm_tile", 1048576)))
        rad = Radiation(device.index or 0, max_cols=m)
        tables[KEY_CTX] = rad
    return rad


def columns(args, device):
    """The column-API inputs of the interior rows, float64 [k][ncol], plus (rows slice, im, jm, lm)."""
    T = _arr(args, "T")
    if T.dim() != 3:
        raise ValueError(f"rrtm-call: T must be [LM, JM, IM], got {tuple(T.shape)}")
    lm, jm, im = (int(x) for x in T.shape)
    j0, j1 = int(args.get("J_S", 2)) - 1, int(args.get("J_E", jm - 1))
    if not (0 <= j0 < j1 <= jm):
        raise ValueError(f"rrtm-call: rows J_S..J_E = {j0 + 1}..{j1} outside 1..{jm}")
    c0, c1 = j0 * im, j1 * im
    f64 = dict(dtype=torch.float64, device=device)

    def lay(name):
        a = _arr(args, name, (lm, jm, im))
        return a.reshape(lm, jm * im)[:, c0:c1].to(**f64).contiguous()

    def sfc(name):
        a = _arr(args, name, (jm, im))
        return a.reshape(jm * im)[c0:c1].to(**f64).contiguous()

    pd = sfc("PD")
    pdsg1 = _arr(args, "PDSG1", (lm,)).to(**f64)
    dsg2 = _arr(args, "DSG2", (lm,)).to(**f64)
    p_lev = torch.empty((lm + 1, c1 - c0), **f64)
    p_lev[0] = float(args["PT"])
    p_lev[1:] = float(args["PT"]) + torch.cumsum(pdsg1[:, None] + dsg2[:, None] * pd[None, :], dim=0)
    t = lay("T")
    q = lay("Q")
    o3 = lay("O3")
    qc = lay("QCW") if args.get("F_QC", True) else torch.zeros_like(t)
    qi = lay("QCI") if args.get("F_QI", False) else torch.zeros_like(t)
    if int(args.get("CLD_FRACTION", 0)) == 0:
        pl = 0.5 * (p_lev[1:] + p_lev[:-1]) * 1.0e-2                 # layer pressure, hPa
        clw = qc + qi
        cldf = torch.where(clw > 1.0e-8 * (pl * 1.0e-3), torch.clamp(1.0e5 * clw * pl * 1.0e-3, max=1.0),
                           torch.zeros_like(clw))
    else:
        cldf = lay("CLDFRA")
    cols = {"p_lev": p_lev.contiguous(), "t_lay": t, "q_lay": q, "o3_lay": o3, "cldf": cldf.contiguous(),
            "qc": qc, "qi": qi, "t_sfc": sfc("TSKIN"), "alb": sfc("ALBEDO")}
    cols["emis"] = sfc("EMISS") if "EMISS" in args else torch.ones_like(cols["t_sfc"])
    cols["col_id"] = torch.arange(c0, c1, dtype=torch.int64, device=device)
    return cols, (c0, c1), (im, jm, lm)


def run(ctx, args: dict, *, stamp=None, rows=None, sw: bool = True) -> None:
    """The RRTM call of one radiation step (see the module docstring); writes args' outputs in place."""
    nts = int(args["NTIMESTEP"])
    dt = float(args["DT"])
    nrads, nradl = int(args["NRADS"]), int(args["NRADL"])
    lsswr = nts % nrads == 0
    lslwr = nts % nradl == 0
    ctx.tables["rrtm_call_flags"] = {"lsswr": lsswr, "lslwr": lslwr, "ntimestep": nts,
                                     "solcon": float(args["SOLCON"]), "sw_run": bool(sw and lsswr)}
    if not (lsswr or lslwr):
        return
    mcica = int(_cfg(ctx, "rrtm_mcica", 0))
    if mcica != 0:
        # the McICA band kernels do not compute the clear-sky surface SW flux that RSWINC takes
        raise NotImplementedError("rrtm-call: rrtm_mcica = 1 leaves RSWINC (clear-sky surface SW) uncomputed; "
                                  "the NAM route (rrtm_mcica = 0, isubcsw = isubclw = 0) is the hook's route")
    if stamp is not None:
        stamp("driver-prep")
    T = _arr(args, "T")
    device = T.device if T.is_cuda else torch.device("cuda", 0)
    cols, (c0, c1), (im, jm, lm) = columns(args, device)
    glat = _arr(args, "GLAT", (jm, im)).reshape(-1)[c0:c1].to(device)
    glon = _arr(args, "GLON", (jm, im)).reshape(-1)[c0:c1].to(device)
    cz = solar.coszen(args["JDAT"], nts * dt, nrads * dt, glat, glon)
    cols["cosz"] = cz if (sw and lsswr) else torch.full_like(cz, -1.0)
    params = Radiation.default_params(mcica=mcica, solcon=float(args["SOLCON"]), seed=int(_cfg(ctx, "rrtm_seed", 0)))
    rad = _context(ctx, device, c1 - c0, lm)
    if stamp is not None:
        stamp("grrad")
    out, _ = rad.run(cols, params=params, precision=str(_cfg(ctx, "rrtm_precision", "fp64")))
    if stamp is not None:
        stamp("driver-post")

    def put3(name, val):
        a = _arr(args, name, (lm, jm, im))
        a.view(lm, jm * im)[:, c0:c1] = val.to(a.dtype)

    def put2(name, val):
        a = _arr(args, name, (jm, im))
        a.view(jm * im)[c0:c1] = val.to(a.dtype)

    if lslwr:
        put3("RLWTT", out["hr_lw"])
        put2("RLWIN", out["flw_dn"][-1])
        put2("RLWTOA", out["flw_up"][0])
        tl = cols["t_lay"][-1]
        put2("SIGT4", STBOLT * tl * tl * tl * tl)
    if sw and lsswr:
        put3("RSWTT", out["hr_sw"])
        put2("RSWIN", out["fsw_dn"][-1])
        put2("RSWOUT", out["fsw_up"][-1])
        put2("RSWINC", out["fsw_dn_sfc_clr"])
        put2("RSWTOA", out["fsw_up"][0])
        put2("CZMEAN", cz)
