"""The NMMB radiation driver step: cadence, per-step zeroing and the RRTM argument dict.

Mirrors namgpu/physics/driver/phys_rad_driver.py:171-262 (``run(ctx, radupdate=, rrtm_call=)``)
for the rrtm/rrtm scheme with the default cloud branch:
  1. every step: zero CLDFRA, CFRACH, CFRACL, CFRACM, CZMEAN and SIGT4 (phys_rad_driver.cu:53);
  2. every step: radupdate -> SOLCON (default: woof_nam_amd.solar.solcon for the date);
  3. on radiation steps (NTIMESTEP == 1 or NTIMESTEP % NRADS == 0): build the RRTM argument dict
     (phys_rad_driver.py:108-168) and call the hook (default woof_nam_amd.rrtm_call.run).
ctx: .cfg (nrads, nradl, nphs, shortwave, longwave, microphysics), .st (state arrays by name),
.grid (im, jm, lm, pt, pdsg1, psgml1), .ntimestep_rad, .dt, .tables, and the clock: ctx.start (a
datetime, the forecast start) from which each step's JDAT is formed, or else the date array JDAT of
the current step in ctx.tables["JDAT"] (phys_rad_driver.py:76-94).  JDAT handed to the hook is the
time at the start of the step.
"""
from __future__ import annotations

import numpy as np
import torch

from . import solar

MP_FLAGS = {  # per microphysics scheme (module_SOLVER_INTERNAL_STATE.F90:868-940)
    "fer": dict(num_water=4, F_QC=1, F_QI=0, F_QS=1, F_QR=1, F_QG=1, F_NI=0),
    "fer_hires": dict(num_water=4, F_QC=1, F_QI=0, F_QS=1, F_QR=1, F_QG=1, F_NI=0),
    "wsm6": dict(num_water=5, F_QC=1, F_QI=1, F_QS=1, F_QR=1, F_QG=1, F_NI=0),
    "thompson": dict(num_water=7, F_QC=1, F_QI=1, F_QS=1, F_QR=1, F_QG=1, F_NI=1),
    "gfs": dict(num_water=2, F_QC=1, F_QI=1, F_QS=0, F_QR=0, F_QG=0, F_NI=0),
}
ZEROED = ("CLDFRA", "CFRACH", "CFRACL", "CFRACM", "CZMEAN", "SIGT4")


def default_radupdate(ctx, *, jdat, dtsw, dtx, lssw):
    """The solar constant at the date of the step (sol_update_nmmb at JDAT)."""
    return {"SOLCON": solar.solcon(jdat, 0.0, dtsw, dtx)}


def clock_jdat(ctx) -> list:
    """JDAT (year, month, day, tz, hour, minute, second, ms) at the start of the step: ctx.start +
    NTIMESTEP DT when the context has a start time (kept in ctx.tables["JDAT"]), else ctx.tables["JDAT"]
    as given (phys_rad_driver.py:76-94)."""
    import datetime as _dt
    start = getattr(ctx, "start", None)
    if start is None:
        jd = ctx.tables.get("JDAT")
        if jd is None:
            raise RuntimeError("rad-driver: no clock (ctx.start) and no ctx.tables['JDAT']")
        return list(jd)
    t = start + _dt.timedelta(seconds=float(ctx.ntimestep_rad) * float(ctx.dt))
    jd = [t.year, t.month, t.day, 0, t.hour, t.minute, t.second, 0]
    ctx.tables["JDAT"] = jd
    return jd


def rrtm_args(ctx, jdat, nts, dt, flags, solcon):
    """The RRTM call's arguments keyed by the Fortran names (phys_rad_driver.py:108-168)."""
    st, g = ctx.st, ctx.grid
    im, jm, lm = int(g.im), int(g.jm), int(g.lm)
    dev = st["T"].device

    def vec(v):
        return torch.as_tensor(np.asarray(v, dtype=np.float32), device=dev)

    zero3 = torch.zeros((lm, jm, im), dtype=torch.float32, device=dev)
    a = {"NTIMESTEP": int(nts), "DT": float(dt), "JDAT": list(jdat), "NPHS": int(getattr(ctx.cfg, "nphs", 1)),
         "NRADS": int(ctx.cfg.nrads), "NRADL": int(ctx.cfg.nradl), "PT": float(g.pt),
         "PDSG1": vec(g.pdsg1), "PSGML1": vec(g.psgml1), "SOLCON": float(solcon),
         "F_QC": bool(flags["F_QC"]), "F_QI": bool(flags["F_QI"]), "F_QS": bool(flags["F_QS"]),
         "F_QR": bool(flags["F_QR"]), "F_QG": bool(flags["F_QG"]), "F_NI": bool(flags["F_NI"]),
         "NUM_WATER": int(flags["num_water"]), "CLD_FRACTION": 0,
         "IMS": 1, "IME": im, "JMS": 1, "JME": jm, "I_S": 1, "I_E": im, "J_S": 2, "J_E": jm - 1, "LM": lm}
    for k in ("GLAT", "GLON", "DSG2", "SGML2", "PD", "T", "Q", "O3", "CW", "F_ICE", "F_RAIN", "ALBEDO", "TSKIN",
              "SM", "CLDFRA", "RLWTT", "RSWTT", "RLWIN", "RSWIN", "RSWINC", "RSWOUT", "RLWTOA", "RSWTOA", "CZMEAN",
              "SIGT4", "CFRACL", "CFRACM", "CFRACH", "ACFRST", "ACFRCV", "NCFRST", "NCFRCV"):
        if k not in st:
            raise KeyError(f"rad-driver: state field {k!r} missing (its producer has not run)")
        a[k] = st[k]
    a["QCW"] = st["QC"] if "QC" in st else zero3
    a["QCI"] = st["QI"] if (flags["F_QI"] and "QI" in st) else zero3
    for k in ("EMISS", "P_PHY", "SICE"):
        if k in st:
            a[k] = st[k]
    return a


def run(ctx, *, radupdate=None, rrtm_call=None, mype: int = 0) -> None:
    cfg, st = ctx.cfg, ctx.st
    sw, lw = str(cfg.shortwave).strip(), str(cfg.longwave).strip()
    if sw != "rrtm" or lw != "rrtm":
        raise NotImplementedError(f"rad-driver runs shortwave=rrtm longwave=rrtm; got {sw!r}/{lw!r}")
    if str(getattr(cfg, "cldfraction", "default")).strip() != "default":
        raise NotImplementedError("rad-driver: only the default cloud-fraction branch")
    mp = str(cfg.microphysics).strip()
    if mp not in MP_FLAGS:
        raise NotImplementedError(f"rad-driver: no flag table for microphysics {mp!r}")
    if rrtm_call is None:
        from .rrtm_call import run as rrtm_call
    radupdate = radupdate or default_radupdate
    nts = int(ctx.ntimestep_rad)
    dt = float(ctx.dt)
    nrads = int(cfg.nrads)
    jdat = clock_jdat(ctx)
    for k in ZEROED:                                     # 1. every step
        st[k].zero_()
    solcon = float(radupdate(ctx, jdat=jdat, dtsw=nrads * dt, dtx=dt, lssw=(nts % nrads == 0))["SOLCON"])
    if nts == 1 or nts % nrads == 0:                     # 3. radiation step
        rrtm_call(ctx, rrtm_args(ctx, jdat, nts, dt, MP_FLAGS[mp], solcon))
