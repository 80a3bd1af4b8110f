"""Python binding of the CPU FP64 oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU baseline
as the checker.  The product (woof_nam_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
from pathlib import Path

import numpy as np

from woof_nam_amd import abi

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
TABLES = HERE.parent / "woof_nam_amd" / "data" / "rrtmg_tables.bin"

_lib = None


def build(force: bool = False) -> Path:
    """Compile liboracle.so with its Makefile (gcc, -ffp-contract=off)."""
    if force and LIB.exists():
        LIB.unlink()
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def _bind(path):
    """Load one build of the oracle (the default one, or a copy built by a test) and load its tables."""
    L = C.CDLL(str(path))
    L.oracle_init.argtypes = [C.c_char_p]
    L.oracle_run.argtypes = [C.POINTER(abi.WrColumns), C.POINTER(abi.WrParams), C.POINTER(abi.WrOutputs),
                             C.POINTER(abi.WrDiag), C.c_int]
    for f in ("oracle_exp", "oracle_log"):
        getattr(L, f).argtypes = [C.c_double]
        getattr(L, f).restype = C.c_double
    L.oracle_pow.argtypes = [C.c_double, C.c_double]
    L.oracle_pow.restype = C.c_double
    dp = C.POINTER(C.c_double)
    L.oracle_lw_setcoef_layer.argtypes = [C.c_double, C.c_double, dp, C.c_double, C.c_double, dp]
    L.oracle_lw_band_layer.argtypes = [dp, C.c_int, dp, dp]
    L.oracle_sw_setcoef_layer.argtypes = [C.c_double, C.c_double, C.c_double, dp, dp]
    L.oracle_sw_band_layer.argtypes = [dp, C.c_int, dp, dp]
    L.oracle_lw_cldprop_layer.argtypes = [C.c_double] * 6 + [dp]
    L.oracle_sw_cldprop_layer.argtypes = [C.c_double] * 6 + [dp, dp, dp]
    L.oracle_sw_expm.argtypes = [C.c_double, C.c_int]
    L.oracle_sw_expm.restype = C.c_double
    L.oracle_sw_layer_opt.argtypes = [C.c_double] * 4 + [dp]
    L.oracle_sw_gpoint.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, dp, dp, dp, dp, dp, C.c_int, dp, dp,
                                   dp, dp, dp]
    rc = L.oracle_init(str(TABLES).encode())
    if rc != 0:
        raise RuntimeError(f"oracle: table load failed ({rc})")
    return L


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = _bind(LIB)
    return _lib


def load(path):
    """A separately built copy of the oracle (for tests that plant a mistake in it)."""
    return _bind(path)


_tables = None


def tables() -> dict:
    """The table file as numpy arrays keyed by name (float32 entries widened to float64, as the
    oracle reads them)."""
    global _tables
    if _tables is None:
        meta = json.loads(TABLES.with_suffix(".json").read_text())
        buf = TABLES.read_bytes()
        out = {}
        for t in meta["tables"]:
            dt = np.float32 if t["dtype"] == "float32" else np.int32
            n = int(np.prod(t["shape"])) if t["shape"] else 1
            a = np.frombuffer(buf, dtype=dt, count=n, offset=t["offset"]).reshape(t["shape"])
            out[t["name"]] = a.astype(np.float64) if dt == np.float32 else a.astype(np.int64)
        _tables = out
    return _tables


def _ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


def _nmmb_state(st: dict, im, jm, lm, j0, j1, pt):
    """WrNmmbState over host float32 arrays (kept alive by the caller's dict)."""
    fp = C.c_float
    return abi.WrNmmbState(im, jm, lm, j0, j1, pt, *[_ptr(st[k], fp) for k in
                                                      ("pd", "pdsg1", "dsg2", "psgml1", "sgml2", "t", "q", "o3", "cw",
                                                       "f_ice", "f_rain", "glat")])


def nmmb_prep(st: dict, dims, opts, ozone: dict | None = None):
    """oracle_nmmb_prep on host float32 state arrays; dims = (im, jm, lm, j0, j1, pt), opts = (llyr, iovr,
    subgrid); ozone: the climatology (o3r float32 [12][loz][jmr] mass mixing ratio, pstr float32 [loz]
    hPa, blte, dlte, k1oz, k2oz, facoz) or None.  Returns the column dict (float64 [k][ncol]) with clds,
    mtop, mbot, p_phy and o3_top."""
    im, jm, lm, j0, j1, pt = dims
    ncol = (j1 - j0) * im
    cols = {k: np.zeros(s) for k, s in abi.nmmb_col_shapes(ncol, lm).items()}
    cols["clds"] = np.zeros((5, ncol), dtype=np.float32)
    cols["mtop"] = np.zeros((3, ncol), dtype=np.int32)
    cols["mbot"] = np.zeros((3, ncol), dtype=np.int32)
    cols["p_phy"] = np.full((lm, jm, im), -1.0, dtype=np.float32)
    cols["o3_top"] = np.zeros(ncol)
    cst = abi.WrNmmbCols(*[_ptr(cols[k]) for k in abi.NMMB_COL_FIELDS], _ptr(cols["clds"], C.c_float),
                         _ptr(cols["mtop"], C.c_int32), _ptr(cols["mbot"], C.c_int32), _ptr(cols["p_phy"], C.c_float),
                         _ptr(cols["o3_top"]))
    oz = None
    if ozone is not None:
        o3r = np.ascontiguousarray(ozone["o3r"], dtype=np.float32)
        pstr = np.ascontiguousarray(ozone["pstr"], dtype=np.float32)
        oz = abi.WrOzone(_ptr(o3r, C.c_float), _ptr(pstr, C.c_float), pstr.size, o3r.shape[2], ozone["blte"],
                         ozone["dlte"], ozone["k1oz"], ozone["k2oz"], ozone["facoz"])
    o = abi.WrNmmbOpts(*opts, C.pointer(oz) if oz is not None else None)
    L = lib()
    L.oracle_nmmb_prep.argtypes = [C.POINTER(abi.WrNmmbState), C.POINTER(abi.WrNmmbOpts), C.POINTER(abi.WrNmmbCols)]
    rc = L.oracle_nmmb_prep(C.byref(_nmmb_state(st, *dims)), C.byref(o), C.byref(cst))
    if rc != 0:
        raise ValueError(f"oracle_nmmb_prep: bad arguments ({rc})")
    return cols


def nmmb_post(st: dict, dims, cols: dict, rad: dict, czmean, lw_step: bool, sw_step: bool, out: dict):
    """oracle_nmmb_post: writes the state outputs in `out` (float32 / int32 host arrays) in place."""
    fp = C.c_float
    cst = abi.WrNmmbCols(*[_ptr(cols[k]) for k in abi.NMMB_COL_FIELDS], _ptr(cols["clds"], fp), None, None, None)
    ro = abi.WrOutputs(*[_ptr(rad[k]) for k in abi.OUTPUT_FIELDS])
    oo = abi.WrNmmbOut(*[_ptr(out[k], fp) for k in abi.NMMB_OUT_FIELDS], _ptr(out["ncfrst"], C.c_int32),
                       _ptr(out["ncfrcv"], C.c_int32))
    L = lib()
    L.oracle_nmmb_post.argtypes = [C.POINTER(abi.WrNmmbState), C.POINTER(abi.WrNmmbCols), C.POINTER(abi.WrOutputs),
                                   C.POINTER(C.c_double), C.c_int, C.c_int, C.POINTER(abi.WrNmmbOut)]
    cz = np.ascontiguousarray(czmean, dtype=np.float64)
    rc = L.oracle_nmmb_post(C.byref(_nmmb_state(st, *dims)), C.byref(cst), C.byref(ro), _ptr(cz), int(lw_step),
                            int(sw_step), C.byref(oo))
    if rc != 0:
        raise ValueError(f"oracle_nmmb_post: bad arguments ({rc})")


def nmmb_albedo(albedo, sm, coszen, rows):
    """oracle_nmmb_albedo: the NAM surface albedo of rows j0..j1 ([jm][im] float32 albedo and sea mask,
    coszen [ncol] float64)."""
    jm, im = albedo.shape
    a = np.ascontiguousarray(albedo, dtype=np.float32)
    s = np.ascontiguousarray(sm, dtype=np.float32)
    cz = np.ascontiguousarray(coszen, dtype=np.float64)
    out = np.zeros(cz.size)
    L = lib()
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    L.oracle_nmmb_albedo.argtypes = [C.c_int] * 4 + [fp, fp, dp, dp]
    if L.oracle_nmmb_albedo(im, jm, rows[0], rows[1], _ptr(a, C.c_float), _ptr(s, C.c_float), _ptr(cz), _ptr(out)) != 0:
        raise ValueError("oracle_nmmb_albedo: bad arguments")
    return out


def sol_update(yr, mon, day, hr, mn, sec, deltsw, deltim, solc0):
    o = abi.WrSol()
    L = lib()
    L.oracle_sol_update.argtypes = [C.c_int] * 6 + [C.c_double] * 3 + [C.POINTER(abi.WrSol)]
    if L.oracle_sol_update(yr, mon, day, hr, mn, sec, deltsw, deltim, solc0, C.byref(o)) != 0:
        raise ValueError("oracle_sol_update: bad arguments")
    return o


def coszmn(glat, glon, solhr, dtswav, nrads, sol):
    lat = np.ascontiguousarray(glat, dtype=np.float64)
    lon = np.ascontiguousarray(glon, dtype=np.float64)
    cz, cd = np.zeros_like(lat), np.zeros_like(lat)
    L = lib()
    dp = C.POINTER(C.c_double)
    L.oracle_coszmn.argtypes = [C.c_int, dp, dp, C.c_double, C.c_double, C.c_int, C.POINTER(abi.WrSol), dp, dp]
    L.oracle_coszmn(lat.size, _ptr(lat), _ptr(lon), solhr, dtswav, nrads, C.byref(sol), _ptr(cz), _ptr(cd))
    return cz, cd


def rdtemp(t, czen, glat, glon, czmean, rswtt, rlwtt, ntsd, dt, start, box):
    """oracle_rdtemp in place on host float32 arrays; start: datetime of the run start; box = (i_lo, i_hi,
    j_lo, j_hi)."""
    lm, jm, im = t.shape
    fp = C.POINTER(C.c_float)
    L = lib()
    L.oracle_rdtemp.argtypes = [C.c_int] * 8 + [C.c_double] + [C.c_int] * 4 + [fp] * 7
    rc = L.oracle_rdtemp(im, jm, lm, box[0], box[1], box[2], box[3], ntsd, dt, start.timetuple().tm_yday, start.year,
                         start.hour, int(start.year % 4 == 0), _ptr(glat, C.c_float), _ptr(glon, C.c_float),
                         _ptr(czmean, C.c_float), _ptr(rswtt, C.c_float), _ptr(rlwtt, C.c_float), _ptr(t, C.c_float),
                         _ptr(czen, C.c_float))
    if rc != 0:
        raise ValueError("oracle_rdtemp: bad arguments")


def run(cols: dict, params: abi.WrParams | None = None, nthreads: int | None = None, diag: bool = False, L=None):
    """Run the oracle on host float64 columns (gen.py layout; optional aerosol optics aer_sw / aer_lw
    [3][bands][nlay+1][ncol] float32).  L: a separately loaded oracle (default: this one).  Returns
    (outputs, diag|None)."""
    L = L or lib()
    params = params or abi.make_params()
    t = np.ascontiguousarray(cols["t_lay"], dtype=np.float64)
    nlay, ncol = t.shape
    path = cols.get("cwp") is not None
    arrs = {k: np.ascontiguousarray(cols[k], dtype=np.float64) for k in abi.INPUT_FIELDS + abi.PATH_FIELDS
            if cols.get(k) is not None}
    ids = cols.get("col_id")
    ids = None if ids is None else np.ascontiguousarray(ids, dtype=np.uint64)
    aer = {k: np.ascontiguousarray(cols[k], dtype=np.float32) for k in abi.AER_FIELDS if cols.get(k) is not None}
    o3t = None if cols.get("o3_top") is None else np.ascontiguousarray(cols["o3_top"], dtype=np.float64)
    for k, a in aer.items():
        if a.shape != abi.aer_shape(k, ncol, nlay):
            raise ValueError(f"{k}: shape {a.shape}, expected {abi.aer_shape(k, ncol, nlay)}")
    ci = abi.WrColumns(ncol, nlay, *[_ptr(arrs.get(k)) for k in abi.INPUT_FIELDS],
                       _ptr(ids, C.c_uint64) if ids is not None else None, None, None,
                       *[_ptr(arrs.get(k)) if path else None for k in abi.PATH_FIELDS],
                       *[_ptr(aer[k], C.c_float) if k in aer else None for k in abi.AER_FIELDS], 0,
                       _ptr(o3t) if o3t is not None else None)
    outs = {k: np.zeros(s) for k, s in abi.output_shapes(ncol, nlay).items()}
    co = abi.WrOutputs(*[_ptr(outs[k]) for k in abi.OUTPUT_FIELDS])
    dg = None
    cd = None
    if diag:
        shp = abi.diag_shapes(ncol, nlay)
        dg = {k: np.zeros(s, dtype=np.uint32 if k.endswith("mask") else np.int32) for k, s in shp.items()}
        cd = abi.WrDiag(*[_ptr(dg[k], C.c_uint32 if k.endswith("mask") else C.c_int32) for k in abi.DIAG_FIELDS])
    n = nthreads or int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    rc = L.oracle_run(C.byref(ci), C.byref(params), C.byref(co), C.byref(cd) if cd is not None else None, n)
    if rc != 0:
        raise RuntimeError(f"oracle_run failed ({rc})")
    return outs, dg


def lw_tables():
    e, f, t = (np.zeros(10001) for _ in range(3))
    lib().oracle_lw_tables(_ptr(e), _ptr(f), _ptr(t))
    return e, f, t


def sw_table():
    e = np.zeros(10001)
    lib().oracle_sw_table(_ptr(e))
    return e


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox(_ptr(c, C.c_uint32), _ptr(k, C.c_uint32), _ptr(o, C.c_uint32))
    return o


def mcica(col: int, seed: int, kind: int, ngpt: int, iovr: int, cf):
    cf = np.ascontiguousarray(cf, dtype=np.float64)
    nl = cf.shape[0]
    words = (ngpt + 31) // 32
    bits = np.zeros((nl, words), dtype=np.uint32)
    L = lib()
    L.oracle_mcica.argtypes = [C.c_uint64, C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.c_int]
    L.oracle_mcica(col, seed, kind, ngpt, nl, iovr, _ptr(cf), _ptr(bits, C.c_uint32), words)
    return bits


def cal_cldfra(cldfra, qc, qi, i_lo, i_hi, j_lo, j_hi, f_qc=True, f_qi=True):
    """CAL_CLDFRA on host float32 [lm][jm][im] arrays, in place."""
    lm, jm, im = cldfra.shape
    L = lib()
    fp = C.POINTER(C.c_float)
    L.oracle_cal_cldfra.argtypes = [C.c_int] * 9 + [fp, fp, fp]
    L.oracle_cal_cldfra(im, jm, lm, i_lo, i_hi, j_lo, j_hi, int(f_qc), int(f_qi), _ptr(cldfra, C.c_float),
                        _ptr(qc, C.c_float), _ptr(qi, C.c_float))


def _aer_tables(tab: dict):
    """wr_aer_tables over host arrays; the arrays it points into are returned with it."""
    keep, vals = [], []
    for k in abi.AER_TABLE_FLOAT:
        a = np.ascontiguousarray(tab[k], dtype=np.float32)
        keep.append(a)
        vals.append(_ptr(a, C.c_float))
    for k in abi.AER_TABLE_INT:
        a = np.ascontiguousarray(tab[k], dtype=np.int32)
        keep.append(a)
        vals.append(_ptr(a, C.c_int32))
    return abi.WrAerTables(*vals), keep


def _aer_cols(p_lev, t_lay, q_lay, xlon, xlat, slmsk):
    f64 = [np.ascontiguousarray(a, dtype=np.float64) for a in (p_lev, t_lay, q_lay)]
    f32 = [np.ascontiguousarray(a, dtype=np.float32) for a in (xlon, xlat, slmsk)]
    nlay, ncol = f64[1].shape
    if f64[0].shape != (nlay + 1, ncol) or f64[2].shape != (nlay, ncol) or any(a.shape != (ncol,) for a in f32):
        raise ValueError("aerosol oracle: p_lev [nlay+1][ncol], t_lay/q_lay [nlay][ncol], xlon/xlat/slmsk [ncol]")
    return f64, f32, nlay, ncol


def aer_optics(tab: dict, p_lev, t_lay, q_lay, xlon, xlat, slmsk, L=None):
    """oracle_aer_optics: (aer_sw float32 [3][14][nlay+1][ncol], aer_lw float32 [3][16][nlay+1][ncol])."""
    (p, t, q), (lo, la, sm), nlay, ncol = _aer_cols(p_lev, t_lay, q_lay, xlon, xlat, slmsk)
    tb, keep = _aer_tables(tab)
    sw = np.zeros((3, abi.NBANDS_SW, nlay + 1, ncol), np.float32)
    lw = np.zeros((3, abi.NBANDS_LW, nlay + 1, ncol), np.float32)
    L = L or lib()
    fp, dp = C.POINTER(C.c_float), C.POINTER(C.c_double)
    L.oracle_aer_optics.argtypes = [C.POINTER(abi.WrAerTables), C.c_int, C.c_int, dp, dp, dp, fp, fp, fp, fp, fp,
                                    C.c_int64]
    if L.oracle_aer_optics(C.byref(tb), ncol, nlay, _ptr(p), _ptr(t), _ptr(q), _ptr(lo, C.c_float),
                           _ptr(la, C.c_float), _ptr(sm, C.c_float), _ptr(sw, C.c_float), _ptr(lw, C.c_float),
                           ncol) != 0:
        raise ValueError("oracle_aer_optics: bad arguments")
    del keep
    return sw, lw


def aer_column(tab: dict, p_lev, t_lay, q_lay, xlon, xlat, slmsk, c: int, L=None) -> dict:
    """oracle_aer_column: one column's intermediate fields (dz, hz, rh, delz, rdrh, idom, ih1, cmix, denn,
    kp)."""
    (p, t, q), _, nlay, ncol = _aer_cols(p_lev, t_lay, q_lay, xlon, xlat, slmsk)
    tb, keep = _aer_tables(tab)
    nrl = nlay + 1
    o = {k: np.zeros(nrl) for k in ("dz", "rh", "delz", "rdrh")}
    o["hz"] = np.zeros(nrl + 1)
    o["idom"], o["ih1"] = np.zeros(nrl, np.int32), np.zeros(nrl, np.int32)
    o["cmix"], o["denn"] = np.zeros(11), np.zeros(3)
    L = L or lib()
    dp, ip = C.POINTER(C.c_double), C.POINTER(C.c_int)
    L.oracle_aer_column.argtypes = [C.POINTER(abi.WrAerTables), C.c_int, C.c_int, dp, dp, dp, C.c_double, C.c_double,
                                    C.c_double, C.c_int] + [dp] * 5 + [ip, ip, dp, dp]
    kp = L.oracle_aer_column(C.byref(tb), ncol, nlay, _ptr(p), _ptr(t), _ptr(q), float(np.float32(xlon[c])),
                             float(np.float32(xlat[c])), float(np.float32(slmsk[c])), int(c),
                             *[_ptr(o[k]) for k in ("dz", "hz", "rh", "delz", "rdrh")],
                             _ptr(o["idom"], C.c_int), _ptr(o["ih1"], C.c_int), _ptr(o["cmix"]), _ptr(o["denn"]))
    if kp < 0:
        raise ValueError("oracle_aer_column: bad arguments")
    o["kp"] = kp
    del keep
    return o


def aer_fpvs(t: float, L=None) -> float:
    L = L or lib()
    L.oracle_aer_fpvs.argtypes = [C.c_double]
    L.oracle_aer_fpvs.restype = C.c_double
    return L.oracle_aer_fpvs(float(t))
