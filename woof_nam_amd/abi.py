"""ctypes mirror of include/woofrad.h (structs, status codes, defaults).

Only layouts and constants live here; the compute entry points are bound in ``_lib.py``.
"""
from __future__ import annotations

import ctypes as C

NBANDS_LW, NGPT_LW, NBANDS_SW, NGPT_SW = 16, 140, 14, 112

WR_OK, WR_ERR_ARG, WR_ERR_TABLES, WR_ERR_HIP, WR_ERR_NOMEM, WR_ERR_UNSUPPORTED, WR_ERR_TOO_LARGE = range(7)
NOALLOC_MAX_COLS = 4096  # WR_NOALLOC_MAX_COLS

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint32)


class WrColumns(C.Structure):
    _fields_ = [("ncol", C.c_int32), ("nlay", C.c_int32),
                ("p_lev", _dp), ("t_lay", _dp), ("q_lay", _dp), ("o3_lay", _dp), ("cldf", _dp),
                ("qc", _dp), ("qi", _dp), ("t_sfc", _dp), ("emis", _dp), ("alb", _dp), ("cosz", _dp),
                ("col_id", C.POINTER(C.c_uint64)), ("seed_dev", C.POINTER(C.c_uint64)),
                ("ncol_dev", C.POINTER(C.c_int32)),
                ("cwp", _dp), ("cip", _dp), ("crp", _dp), ("rew", _dp), ("rei", _dp),
                ("aer_sw", C.POINTER(C.c_float)), ("aer_lw", C.POINTER(C.c_float)), ("aer_ld", C.c_int64),
                ("o3_top", _dp)]


# aerosol optics inputs: [3][bands][nlay+1][ncol] float32 (optical depth, single-scattering albedo,
# asymmetry per radiation layer, the added top layer first)
AER_FIELDS = ("aer_sw", "aer_lw")
AER_BANDS = {"aer_sw": NBANDS_SW, "aer_lw": NBANDS_LW}


def aer_shape(field: str, ncol: int, nlay: int) -> tuple:
    return (3, AER_BANDS[field], nlay + 1, ncol)


# wr_aer_tables: the climatological aerosol's tables (device pointers, float32 then int32 members)
AER_TABLE_FLOAT = ("haer", "prsref", "sigref", "rhlev", "extrhi", "scarhi", "ssarhi", "asyrhi", "extrhd", "scarhd",
                   "ssarhd", "asyrhd", "extstra", "cmixg", "denng")
AER_TABLE_INT = ("idxcg", "kprfg")


class WrAerTables(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_float)) for n in AER_TABLE_FLOAT] + [(n, _ip) for n in AER_TABLE_INT]


class WrParams(C.Structure):
    _fields_ = [("mcica", C.c_int32), ("iovr", C.c_int32), ("seed", C.c_uint64),
                ("solcon", C.c_double), ("re_liq", C.c_double), ("re_ice", C.c_double),
                ("vmr", C.c_double * 10)]


class WrOutputs(C.Structure):
    _fields_ = [("hr_lw", _dp), ("hr_sw", _dp), ("flw_up", _dp), ("flw_dn", _dp),
                ("fsw_up", _dp), ("fsw_dn", _dp), ("fsw_dn_sfc_clr", _dp)]


class WrDiag(C.Structure):
    _fields_ = [("lw_jp", _ip), ("lw_jt", _ip), ("lw_jt1", _ip), ("lw_laytrop", _ip),
                ("sw_jp", _ip), ("sw_jt", _ip), ("sw_jt1", _ip), ("sw_laytrop", _ip),
                ("lw_mask", _up), ("sw_mask", _up)]


class WrSurrogate(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("w1", "w2", "w3", "b1", "b2", "b3", "fmean", "fstd", "fmin", "fmax",
                                           "tmean", "tstd")] + [("bands", C.c_uint32)]


INPUT_FIELDS = ("p_lev", "t_lay", "q_lay", "o3_lay", "cldf", "qc", "qi", "t_sfc", "emis", "alb", "cosz")
# the optional path form of the clouds (progcld2's paths and radii; cwp set: qc/qi are not read)
PATH_FIELDS = ("cwp", "cip", "crp", "rew", "rei")

_fp = C.POINTER(C.c_float)


class WrSol(C.Structure):
    _fields_ = [("sollag", C.c_double), ("sindec", C.c_double), ("cosdec", C.c_double), ("anginc", C.c_double),
                ("solcon", C.c_double), ("nstp", C.c_int32)]


class WrNmmbState(C.Structure):
    _fields_ = [("im", C.c_int32), ("jm", C.c_int32), ("lm", C.c_int32), ("j0", C.c_int32), ("j1", C.c_int32),
                ("pt", C.c_float)] + [(n, _fp) for n in ("pd", "pdsg1", "dsg2", "psgml1", "sgml2", "t", "q", "o3",
                                                          "cw", "f_ice", "f_rain", "glat")]


class WrOzone(C.Structure):
    _fields_ = [("o3r", C.POINTER(C.c_float)), ("pstr", C.POINTER(C.c_float)), ("loz", C.c_int32), ("jmr", C.c_int32),
                ("blte", C.c_float), ("dlte", C.c_float), ("k1oz", C.c_int32), ("k2oz", C.c_int32),
                ("facoz", C.c_float)]


class WrNmmbOpts(C.Structure):
    _fields_ = [("llyr", C.c_int32), ("iovr", C.c_int32), ("subgrid", C.c_int32), ("ozone", C.POINTER(WrOzone))]


NMMB_COL_FIELDS = ("p_lev", "t_lay", "q_lay", "o3_lay", "cldf", "cwp", "cip", "crp", "rew", "rei")


class WrNmmbCols(C.Structure):
    _fields_ = [(n, _dp) for n in NMMB_COL_FIELDS] + [("clds", _fp), ("mtop", _ip), ("mbot", _ip), ("p_phy", _fp),
                                                       ("o3_top", _dp)]


NMMB_OUT_FIELDS = ("rlwtt", "rswtt", "cldfra", "rlwin", "rlwtoa", "sigt4", "rswin", "rswout", "rswinc", "rswtoa",
                   "czmean", "cfracl", "cfracm", "cfrach", "acfrst", "acfrcv")


class WrNmmbOut(C.Structure):
    _fields_ = [(n, _fp) for n in NMMB_OUT_FIELDS] + [("ncfrst", _ip), ("ncfrcv", _ip)]


def nmmb_col_shapes(ncol: int, lm: int) -> dict:
    s = {"p_lev": (lm + 1, ncol)}
    for k in NMMB_COL_FIELDS[1:]:
        s[k] = (lm, ncol)
    return s


OUTPUT_FIELDS = ("hr_lw", "hr_sw", "flw_up", "flw_dn", "fsw_up", "fsw_dn", "fsw_dn_sfc_clr")
DIAG_FIELDS = ("lw_jp", "lw_jt", "lw_jt1", "lw_laytrop", "sw_jp", "sw_jt", "sw_jt1", "sw_laytrop",
               "lw_mask", "sw_mask")

# NAM defaults (phys_rrtm_getgases.cu:56-65 minor-gas VMRs; CO2/N2O/CH4 of the benchmark workload)
DEFAULT_VMR = (410.0e-6, 0.33e-6, 1.9e-6, 0.209, 1.50e-8, 3.520e-10, 6.358e-10, 1.500e-10,
               1.397e-10, 8.2000e-11)


def make_params(mcica: int = 1, iovr: int = 1, seed: int = 0, solcon: float = 1368.22,
                re_liq: float = 10.0, re_ice: float = 50.0, vmr=DEFAULT_VMR) -> WrParams:
    p = WrParams()
    p.mcica, p.iovr, p.seed = int(mcica), int(iovr), int(seed)
    p.solcon, p.re_liq, p.re_ice = float(solcon), float(re_liq), float(re_ice)
    for i, v in enumerate(vmr):
        p.vmr[i] = float(v)
    return p


def output_shapes(ncol: int, nlay: int) -> dict:
    return {"hr_lw": (nlay, ncol), "hr_sw": (nlay, ncol), "flw_up": (nlay + 2, ncol),
            "flw_dn": (nlay + 2, ncol), "fsw_up": (nlay + 2, ncol), "fsw_dn": (nlay + 2, ncol),
            "fsw_dn_sfc_clr": (ncol,)}


def diag_shapes(ncol: int, nlay: int) -> dict:
    nrl = nlay + 1
    return {"lw_jp": (nrl, ncol), "lw_jt": (nrl, ncol), "lw_jt1": (nrl, ncol), "lw_laytrop": (ncol,),
            "sw_jp": (nrl, ncol), "sw_jt": (nrl, ncol), "sw_jt1": (nrl, ncol), "sw_laytrop": (ncol,),
            "lw_mask": (nrl, 5, ncol), "sw_mask": (nrl, 4, ncol)}
