"""Error budget of the mixed-precision path against FP64, in W/m2 and K/day.

The FP64 HIP path is the reference here: it reproduces the CPU FP64 oracle bit for bit (the GPU
parity tests assert it), so the same numbers hold against the oracle.
"""
from __future__ import annotations

import numpy as np
import torch

from .radiation import Radiation

KDAY = 86400.0
LW_BANDS = list(range(1, 17))
SW_BANDS = list(range(16, 30))


def _rms(d):
    return float(torch.sqrt(torch.mean(d * d)))


def report(rad: Radiation, cols: dict, per_band: bool = True) -> dict:
    """Overall, per-layer and (optionally) per-band RMS errors of `mixed` against `fp64`.

    cols: float64 device columns.  Returns a dict of plain floats/lists.
    """
    ref, _ = rad.run(cols, precision="fp64")
    mix, _ = rad.run(cols, precision="mixed")
    torch.cuda.synchronize()
    rep = {"columns": int(cols["t_lay"].shape[1])}
    for k in ("hr_lw", "hr_sw"):
        d = mix[k] - ref[k]
        rep[k + "_rms_kday"] = _rms(d) * KDAY
        rep[k + "_max_kday"] = float(d.abs().max()) * KDAY
        rep[k + "_rms_by_layer_kday"] = [float(x) * KDAY for x in torch.sqrt(torch.mean(d * d, dim=1)).cpu()]
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        d = mix[k] - ref[k]
        rep[k + "_rms_wm2"] = _rms(d)
        rep[k + "_max_wm2"] = float(d.abs().max())
    if per_band:
        lw, sw = {}, {}
        for b in range(30):
            r = rad.run_band(cols, b, precision="fp64")
            m = rad.run_band(cols, b, precision="mixed")
            torch.cuda.synchronize()
            if b < 16:
                lw[LW_BANDS[b]] = {"flux_rms_wm2": max(_rms(m["flw_up"] - r["flw_up"]), _rms(m["flw_dn"] - r["flw_dn"])),
                                   "hr_rms_kday": _rms(m["hr_lw"] - r["hr_lw"]) * KDAY}
            else:
                sw[SW_BANDS[b - 16]] = {"flux_rms_wm2": max(_rms(m["fsw_up"] - r["fsw_up"]),
                                                            _rms(m["fsw_dn"] - r["fsw_dn"])),
                                        "hr_rms_kday": _rms(m["hr_sw"] - r["hr_sw"]) * KDAY}
        rep["lw_by_band"] = lw
        rep["sw_by_band"] = sw
    return rep


def format_report(rep: dict) -> str:
    """Human-readable tables of a report() result."""
    out = [f"columns: {rep['columns']}"]
    for k in ("hr_lw", "hr_sw"):
        out.append(f"{k}: RMS {rep[k + '_rms_kday']:.3e} K/day, max {rep[k + '_max_kday']:.3e} K/day")
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        out.append(f"{k}: RMS {rep[k + '_rms_wm2']:.3e} W/m2, max {rep[k + '_max_wm2']:.3e} W/m2")
    if "lw_by_band" in rep:
        out.append("band  flux RMS (W/m2)  HR RMS (K/day)")
        for b, v in list(rep["lw_by_band"].items()) + list(rep["sw_by_band"].items()):
            kind = "LW" if b in rep["lw_by_band"] and v is rep["lw_by_band"][b] else "SW"
            out.append(f"{kind}{b:02d}  {v['flux_rms_wm2']:.3e}        {v['hr_rms_kday']:.3e}")
    out.append("layer (top first)  LW HR RMS (K/day)  SW HR RMS (K/day)")
    for i, (a, b) in enumerate(zip(rep["hr_lw_rms_by_layer_kday"], rep["hr_sw_rms_by_layer_kday"])):
        out.append(f"{i:3d}  {a:.3e}  {b:.3e}")
    return "\n".join(out)


def _np(x):
    return np.asarray(x)
