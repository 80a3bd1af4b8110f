"""The NMMB radiation hook over the HIP library: ``run(ctx, args)``.

Same call contract as the reference's ``namgpu.physics.rrtm.rrtm_call.run(ctx, args)``
(namgpu/physics/rrtm/rrtm_call.py:246, called by phys_rad_driver.py:255-262): ``args`` is the
dict of Fortran RRTM argument names that the radiation driver builds (phys_rad_driver.py:108-168),
holding device arrays aliased to the model state ([LM, JM, IM] float32, level 1 at the model top,
or [JM, IM]); the hook writes the radiation outputs into the output arrays of ``args`` in place and
returns None.  The input species arrays are only read.

What it does on a call, all of it on the device in library kernels:
  * cadence: SW on NTIMESTEP % NRADS == 0, LW on NTIMESTEP % NRADL == 0 (rrtm_call.py:253-254);
  * climatological ozone, the NAM's (ntoz = 0): gas_init's monthly zonal means with gas_update's time
    factors, interpolated by getozn inside wr_nmmb_prep (ctx.cfg.rrtm_ozone = "state" takes O3 instead);
  * wr_nmmb_prep on the interior rows J_S..J_E (phys_rad_driver.py:165-166): driver prep (P8W,
    P_PHY, phys_rrtm_driver_prep.cu:82), the driver call's condensate and subgrid cloud
    (phys_rrtm_driver_call.cu:200-275), grrad setup, and progcld2 with gethml
    (phys_rrtm_progcld2.cu:92-450) in the NAM configuration: Ferrier total condensate CW split by
    F_ICE / F_RAIN into cloud water, ice and rain, the condensate cloud fraction, paths and diagnosed
    effective radii, low / middle / high cloud cover;
  * solar geometry: the sun's position for the date (sol_update) and each column's cos zenith averaged
    over the sunlit instants of the SW interval (coszmn), on the library (woof_nam_amd.solar);
  * climatological aerosol, the NAM's (iaer = 11): wr_aer_optics maps the OPAC climatology of the
    month onto the columns (setaer, aer_property, radclimaer) and gives each layer's optical depth,
    single-scattering albedo and asymmetry per band (ctx.cfg.rrtm_aerosol = "tables" takes setaer's
    outputs from ctx.tables instead, "none" runs without aerosol);
  * the surface albedo (wr_nmmb_albedo: setalb with IALB = 2, the Payne open-water table over sea in
    daylight, the state's ALBEDO elsewhere; emissivity 1, setemis with IEMS = 0, unless EMISS is given);
  * the radiation call on the path-form columns (woof_nam_amd.radiation, the NAM route mcica = 0;
    FP64 unless ctx.cfg.rrtm_precision = "mixed"), in tiles sized from free device memory;
  * wr_nmmb_post (phys_rrtm_driver_post.cu:71-150): RLWTT, RSWTT (K/s), RLWIN, RSWIN, RSWOUT,
    RSWINC, RLWTOA, RSWTOA, CZMEAN, SIGT4 = sigma T(LM)^4 on LW steps, and on every call CLDFRA,
    CFRACL/M/H and the ACFRST/NCFRST (ACFRCV/NCFRCV) accumulators.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib, abi, aerosol, ozone, solar
from .radiation import Radiation

KEY_CTX = "woof_nam_amd.radiation"
KEY_LLYR = "woof_nam_amd.llyr"
STATE_3D = ("T", "Q", "O3", "CW", "F_ICE", "F_RAIN")
OUT_3D = ("RLWTT", "RSWTT", "CLDFRA")
OUT_2D = ("RLWIN", "RLWTOA", "SIGT4", "RSWIN", "RSWOUT", "RSWINC", "RSWTOA", "CZMEAN", "CFRACL", "CFRACM",
          "CFRACH", "ACFRST", "ACFRCV")


def _arr(args, name, shape=None, dtype=None):
    if name not in args:
        raise KeyError(f"rrtm-call: argument {name!r} missing (its producer has not run)")
    a = args[name]
    if not isinstance(a, torch.Tensor):
        a = torch.as_tensor(np.asarray(a))
    if shape is not None and tuple(a.shape) != tuple(shape):
        raise ValueError(f"rrtm-call: {name} has shape {tuple(a.shape)}, expected {tuple(shape)}")
    if dtype is not None and a.dtype != dtype:
        raise ValueError(f"rrtm-call: {name} has dtype {a.dtype}, expected {dtype}")
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


def boundary_layer_top(si) -> int:
    """llyr of cld_init (radiation_clouds_nmmb.f:327-344, ivflip = 0; phys_rrtm_cld_init.cu): with the
    sigma interfaces si (LM + 1 values, top first), the last k of k = LM, LM-1, ..., 2 before the first
    si(k) < 0.9 (1-based); 2 when none is."""
    si = np.asarray(si, dtype=np.float32)
    nlay = si.size - 1
    kl = 1
    for k in range(nlay, 1, -1):
        kl = k
        if si[k - 1] < np.float32(0.9):
            break
    return kl


def sigma_interfaces(args, lm):
    """Interface sigma of the hybrid coordinate, (p - PT) / (p_sfc - PT), at the standard surface
    pressure 101325 Pa (the NMMB's SGM when the grid does not carry it)."""
    pt = float(args["PT"])
    pdsg1 = _arr(args, "PDSG1", (lm,)).double().cpu().numpy()
    dsg2 = _arr(args, "DSG2", (lm,)).double().cpu().numpy()
    pdtop = float(np.sum(pdsg1))
    pd = 101325.0 - pt - pdtop
    p = pt + np.concatenate([[0.0], np.cumsum(pdsg1 + dsg2 * pd)])
    return (p - pt) / (p[-1] - pt)


def _llyr(ctx, args, lm):
    v = _cfg(ctx, "rrtm_llyr", None)
    if v is not None:
        return int(v)
    if KEY_LLYR not in ctx.tables:
        sgm = getattr(getattr(ctx, "grid", None), "sgm", None)
        ctx.tables[KEY_LLYR] = boundary_layer_top(sgm if sgm is not None else sigma_interfaces(args, lm))
    return ctx.tables[KEY_LLYR]


def _fptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_float)) if t is not None else None


def _dptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double)) if t is not None else None


def _iptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_int32)) if t is not None else None


def nmmb_state(args, device, rows):
    """WrNmmbState over the state arrays of ``args`` (float32, contiguous, on the device) and the arrays
    it points to (kept alive by the caller)."""
    T = _arr(args, "T")
    if T.dim() != 3:
        raise ValueError(f"rrtm-call: T must be [LM, JM, IM], got {tuple(T.shape)}")
    lm, jm, im = (int(x) for x in T.shape)
    j0, j1 = rows
    keep = {}
    for k in STATE_3D:
        keep[k] = _arr(args, k, (lm, jm, im), torch.float32)
    for k in ("PD", "GLAT"):
        keep[k] = _arr(args, k, (jm, im), torch.float32)
    for k in ("PDSG1", "DSG2", "PSGML1", "SGML2"):
        keep[k] = _arr(args, k, (lm,)).to(device=device, dtype=torch.float32)
    for k, v in keep.items():
        if v.device != device or not v.is_contiguous():
            raise ValueError(f"rrtm-call: {k} must be a contiguous array on {device}")
    st = abi.WrNmmbState(im, jm, lm, j0, j1, float(args["PT"]),
                         *[_fptr(keep[k]) for k in ("PD", "PDSG1", "DSG2", "PSGML1", "SGML2", "T", "Q", "O3", "CW",
                                                    "F_ICE", "F_RAIN", "GLAT")])
    return st, keep, (im, jm, lm)


def prep(args, device, rows, llyr, iovr=1, subgrid=True, stream=None, ozone=None):
    """wr_nmmb_prep: the path-form columns of the radiated rows (float64 device tensors) with gethml's
    clds [5][ncol], mtop / mbot [3][ncol] and the added top layer's ozone o3_top [ncol]; P_PHY is
    written when args holds it.  ozone: an abi.WrOzone (climatological ozone) or None (the state's O3)."""
    st, keep, (im, jm, lm) = nmmb_state(args, device, rows)
    ncol = (rows[1] - rows[0]) * im
    cols = {k: torch.empty(s, dtype=torch.float64, device=device) for k, s in abi.nmmb_col_shapes(ncol, lm).items()}
    cols["clds"] = torch.empty((5, ncol), dtype=torch.float32, device=device)
    cols["mtop"] = torch.empty((3, ncol), dtype=torch.int32, device=device)
    cols["mbot"] = torch.empty((3, ncol), dtype=torch.int32, device=device)
    cols["o3_top"] = torch.empty(ncol, dtype=torch.float64, device=device)
    pphy = args.get("P_PHY")
    if pphy is not None:
        pphy = _arr(args, "P_PHY", (lm, jm, im), torch.float32)
    cst = abi.WrNmmbCols(*[_dptr(cols[k]) for k in abi.NMMB_COL_FIELDS], _fptr(cols["clds"]), _iptr(cols["mtop"]),
                         _iptr(cols["mbot"]), _fptr(pphy), _dptr(cols["o3_top"]))
    s = stream or torch.cuda.current_stream(device)
    opts = abi.WrNmmbOpts(int(llyr), int(iovr), int(bool(subgrid)), C.pointer(ozone) if ozone is not None else None)
    rc = _lib.load().wr_nmmb_prep(C.byref(st), C.byref(opts), C.byref(cst), C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_nmmb_prep")
    cols["_keep"] = keep
    return cols, (im, jm, lm)


def albedo(args, device, rows, coszen, stream=None):
    """wr_nmmb_albedo: the NAM surface albedo of the radiated rows (ALBEDO; over open water in daylight the
    Payne open-water table by zenith angle), float64 [ncol] on the device."""
    st, keep, (im, jm, lm) = nmmb_state(args, device, rows)
    a = _arr(args, "ALBEDO", (jm, im), torch.float32)
    sm = _arr(args, "SM", (jm, im), torch.float32)
    cz = coszen.to(device=device, dtype=torch.float64).contiguous()
    out = torch.empty((rows[1] - rows[0]) * im, dtype=torch.float64, device=device)
    s = stream or torch.cuda.current_stream(device)
    rc = _lib.load().wr_nmmb_albedo(C.byref(st), C.c_void_p(a.data_ptr()), C.c_void_p(sm.data_ptr()),
                                    C.c_void_p(cz.data_ptr()), C.c_void_p(out.data_ptr()), C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_nmmb_albedo")
    return out


def post(args, device, rows, cols, out, czmean, lw_step, sw_step, stream=None):
    """wr_nmmb_post: the radiation outputs and cloud diagnostics into the state arrays of ``args``."""
    st, keep, (im, jm, lm) = nmmb_state(args, device, rows)
    o = {k: _arr(args, k, (lm, jm, im), torch.float32) for k in OUT_3D}
    o.update({k: _arr(args, k, (jm, im), torch.float32) for k in OUT_2D})
    o.update({k: _arr(args, k, (jm, im), torch.int32) for k in ("NCFRST", "NCFRCV")})
    cst = abi.WrNmmbCols(*[_dptr(cols[k]) for k in abi.NMMB_COL_FIELDS], _fptr(cols["clds"]), None, None, None, None)
    ro = abi.WrOutputs(*[_dptr(out[k]) for k in abi.OUTPUT_FIELDS])
    oo = abi.WrNmmbOut(*[_fptr(o[k.upper()]) for k in abi.NMMB_OUT_FIELDS], _iptr(o["NCFRST"]), _iptr(o["NCFRCV"]))
    s = stream or torch.cuda.current_stream(device)
    rc = _lib.load().wr_nmmb_post(C.byref(st), C.byref(cst), C.byref(ro), _dptr(czmean), int(bool(lw_step)),
                                  int(bool(sw_step)), C.byref(oo), C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_nmmb_post")


# setaer's outputs in the reference's table keys (phys_rrtm_grrad_aerosols.py: faersw [3, 14, LM+1, JM, IM],
# faerlw [3, 16, LM+1, JM, IM] float32 on the device)
AER_KEYS = {"aer_sw": "rrtm_faersw", "aer_lw": "rrtm_faerlw"}
KEY_AER = "woof_nam_amd.aerosol"


def _aer_tables(ctx, lm, jm, im, c0, c1, device) -> dict:
    out = {}
    for k, key in AER_KEYS.items():
        a = ctx.tables.get(key) if hasattr(ctx, "tables") else None
        if a is None:
            raise KeyError(f"rrtm-call: rrtm_aerosol = 'tables' but ctx.tables[{key!r}] is missing")
        if not isinstance(a, torch.Tensor):
            a = torch.as_tensor(np.asarray(a))
        shp = (3, abi.AER_BANDS[k], lm + 1, jm, im)
        if tuple(a.shape) != shp or a.dtype != torch.float32:
            raise ValueError(f"rrtm-call: {key} must be float32 {shp}, got {a.dtype} {tuple(a.shape)}")
        out[k] = a.to(device).contiguous().reshape(3, shp[1], lm + 1, jm * im)[..., c0:c1]
    return out


def aerosols(ctx, args, cols, lm, jm, im, c0, c1, device) -> dict:
    """The aerosol optics of the radiated columns (the column call's aer_sw / aer_lw), by
    ctx.cfg.rrtm_aerosol:
      "climatology" (default; the NAM's iaer = 11): wr_aer_optics over the OPAC climatology of the month of
          JDAT, as the reference's call runs setaer on every radiation call (rrtm_call.py:354,
          phys_rrtm_grrad_aerosols.py:91-124) after aer_update loaded the month (phys_rrtm_radupdate.py:
          119-121); the surface type is the driver call's SLMSK: 2 on sea ice (SICE > 0.5), else 1 - SM;
      "tables": setaer's outputs already in ctx.tables (rrtm_faersw / rrtm_faerlw over the whole grid),
          taken as column slices;
      "none": no aerosol (iaer = 0)."""
    mode = str(_cfg(ctx, "rrtm_aerosol", "climatology"))
    if mode == "none":
        return {}
    if mode == "tables":
        return _aer_tables(ctx, lm, jm, im, c0, c1, device)
    if mode != "climatology":
        raise ValueError(f"rrtm-call: rrtm_aerosol must be 'climatology', 'tables' or 'none', got {mode!r}")
    clim = ctx.tables.get(KEY_AER)
    if clim is None or clim.device != device:
        clim = aerosol.Climatology(device)
        ctx.tables[KEY_AER] = clim
    clim.set_month(int(args["JDAT"][1]))
    sm = _arr(args, "SM", (jm, im), torch.float32).reshape(-1)[c0:c1].to(device)
    if "SICE" in args:
        sice = _arr(args, "SICE", (jm, im), torch.float32).reshape(-1)[c0:c1].to(device)
        slmsk = torch.where(sice > 0.5, torch.full_like(sm, 2.0), 1.0 - sm)
    else:
        slmsk = 1.0 - sm
    glat = _arr(args, "GLAT", (jm, im), torch.float32).reshape(-1)[c0:c1]
    glon = _arr(args, "GLON", (jm, im), torch.float32).reshape(-1)[c0:c1]
    return aerosol.optics(clim, cols, glat, glon, slmsk)


KEY_OZONE = "woof_nam_amd.ozone"


def ozone_struct(ctx, args, device, seconds):
    """The climatological ozone of the step (the NAM's ntoz = 0: getozn over gas_init's climatology with
    gas_update's time factors at JDAT + seconds), or None when ctx.cfg.rrtm_ozone = "state" (the
    state's O3)."""
    if str(_cfg(ctx, "rrtm_ozone", "climatology")) == "state":
        return None
    st = ctx.tables.get(KEY_OZONE)
    if st is None or st[0] is None or st[0].device != device:
        st = (ozone.Climatology(device), st[1] if st is not None else ozone.TimeFactors())
        ctx.tables[KEY_OZONE] = st
    clim, tf = st
    t = solar.current(args["JDAT"], seconds)
    tf.update(t.month, t.day, t.hour)
    return clim.struct(tf)


def state_dict(ctx) -> dict:
    """The hook's own state between calls, for the model's restart sets (namgpu/restart.py:1-12 saves
    every array reachable from the state; this is what the hook keeps outside it): gas_update's ozone
    time factors (their saved months and mid-month days carry across the change days) and the
    boundary-layer cloud top llyr.  The device context and tables are rebuilt on resume.  JSON-able."""
    out = {"format": 1}
    if KEY_LLYR in ctx.tables:
        out["llyr"] = int(ctx.tables[KEY_LLYR])
    oz = ctx.tables.get(KEY_OZONE)
    if oz is not None:
        tf = oz[1]
        out["ozone_time"] = {"k1oz": tf.k1oz, "k2oz": tf.k2oz, "midm": tf.midm, "midp": tf.midp,
                             "facoz": float(tf.facoz), "started": bool(tf.started)}
    return out


def load_state_dict(ctx, state: dict) -> None:
    """Restore what state_dict saved (before the first call of the resumed run)."""
    if int(state.get("format", 0)) != 1:
        raise ValueError(f"rrtm-call: unknown state format {state.get('format')!r}")
    if "llyr" in state:
        ctx.tables[KEY_LLYR] = int(state["llyr"])
    t = state.get("ozone_time")
    if t is not None:
        tf = ozone.TimeFactors(k1oz=int(t["k1oz"]), k2oz=int(t["k2oz"]), midm=int(t["midm"]), midp=int(t["midp"]),
                               facoz=np.float32(t["facoz"]), started=bool(t["started"]))
        ctx.tables[KEY_OZONE] = (None, tf)        # the climatology tables are loaded again on the next call


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
    if T.dim() != 3:
        raise ValueError(f"rrtm-call: T must be [LM, JM, IM], got {tuple(T.shape)}")
    lm, jm, im = (int(x) for x in T.shape)
    device = T.device if T.is_cuda else torch.device("cuda", 0)
    j0, j1 = int(args.get("J_S", 2)) - 1, int(args.get("J_E", jm - 1))
    if not (0 <= j0 < j1 <= jm):
        raise ValueError(f"rrtm-call: rows J_S..J_E = {j0 + 1}..{j1} outside 1..{jm}")
    oz = ozone_struct(ctx, args, device, 0.0)
    cols, _ = prep(args, device, (j0, j1), _llyr(ctx, args, lm), iovr=1, subgrid=bool(_cfg(ctx, "rrtm_subgrid", True)),
                   ozone=oz)
    c0, c1 = j0 * im, j1 * im
    glat = _arr(args, "GLAT", (jm, im)).reshape(-1)[c0:c1]
    glon = _arr(args, "GLON", (jm, im)).reshape(-1)[c0:c1].to(device)
    cz = solar.coszen(args["JDAT"], 0.0, nrads, dt, glat.to(device), glon)[0].contiguous()
    sw_step = bool(sw and lsswr)
    f64 = dict(dtype=torch.float64, device=device)
    cols["cosz"] = cz if sw_step else torch.full_like(cz, -1.0)
    cols["t_sfc"] = _arr(args, "TSKIN", (jm, im)).reshape(-1)[c0:c1].to(**f64).contiguous()
    cols["alb"] = albedo(args, device, (j0, j1), cz)
    cols["emis"] = (_arr(args, "EMISS", (jm, im)).reshape(-1)[c0:c1].to(**f64).contiguous() if "EMISS" in args
                    else torch.ones(c1 - c0, **f64))
    cols["col_id"] = torch.arange(c0, c1, dtype=torch.int64, device=device)
    cols.update(aerosols(ctx, args, cols, lm, jm, im, c0, c1, device))
    params = Radiation.default_params(mcica=mcica, solcon=float(args["SOLCON"]), seed=int(_cfg(ctx, "rrtm_seed", 0)))
    rad = _context(ctx, device, c1 - c0, lm)
    if stamp is not None:
        stamp("grrad")
    out, _ = rad.run(cols, params=params, precision=str(_cfg(ctx, "rrtm_precision", "fp64")))
    if stamp is not None:
        stamp("driver-post")
    post(args, device, (j0, j1), cols, out, cz.contiguous(), lslwr, sw_step)
