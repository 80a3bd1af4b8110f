"""The C-ABI library loads and exports every entry point include/woofrad.h declares (no GPU calls)."""
import ctypes as C
import re
from pathlib import Path

from woof_nam_amd import _lib, abi

HEADER = Path(__file__).resolve().parents[1] / "include" / "woofrad.h"


def declared():
    txt = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(wr_[a-z0-9_]+)\s*\(", txt)))


def test_every_declared_symbol_is_exported():
    L = _lib.load()
    names = declared()
    assert len(names) >= 11
    assert set(names) == set(_lib.EXPORTS)
    for n in names:
        assert hasattr(L, n), n


def test_struct_layouts_match_header():
    # x86-64 SysV layout of the header structs
    assert C.sizeof(abi.WrColumns) == 8 + 23 * 8
    assert C.sizeof(abi.WrNmmbState) == 6 * 4 + 12 * 8
    assert C.sizeof(abi.WrNmmbOpts) == 3 * 4 + 4 + 8
    assert C.sizeof(abi.WrOzone) == 2 * 8 + 7 * 4 + 4
    assert C.sizeof(abi.WrNmmbCols) == 15 * 8
    assert C.sizeof(abi.WrAerTables) == 17 * 8
    assert C.sizeof(abi.WrSurrogate) == 12 * 8 + 8
    assert C.sizeof(abi.WrNmmbOut) == 18 * 8
    assert C.sizeof(abi.WrParams) == 4 + 4 + 8 + 3 * 8 + 10 * 8
    assert C.sizeof(abi.WrOutputs) == 7 * 8
    assert C.sizeof(abi.WrDiag) == 10 * 8


def test_default_params_and_strerror():
    L = _lib.load()
    p = abi.WrParams()
    L.wr_default_params(C.byref(p))
    assert (p.mcica, p.iovr, p.seed) == (1, 1, 0)
    assert p.solcon == 1368.22 and p.re_liq == 10.0 and p.re_ice == 50.0
    assert tuple(p.vmr) == abi.DEFAULT_VMR
    assert L.wr_strerror(abi.WR_ERR_UNSUPPORTED).decode().startswith("option")
    assert L.wr_kernel_name(3).decode() == "k_lw_b01"
    assert L.wr_kernel_name(32).decode() == "k_sw_b29"


def test_null_arguments_rejected_without_gpu():
    L = _lib.load()
    assert L.wr_run_fp64(None, None, None, None, None, None) == abi.WR_ERR_ARG
    assert L.wr_timing(None, 1) == abi.WR_ERR_ARG
    assert L.wr_mcica_masks(-1, 61, None, None, 0, 1, None, None, None) == abi.WR_ERR_ARG


def test_aer_optics_rejects_bad_arguments_without_gpu():
    L = _lib.load()
    t, cols = abi.WrAerTables(), abi.WrColumns()
    assert L.wr_aer_optics(None, None, None, None, None, None, None, 0, None) == abi.WR_ERR_ARG
    # missing tables, then 64 radiation layers (63 model layers is the most)
    assert L.wr_aer_optics(C.byref(t), C.byref(cols), 8, 8, 8, 8, 8, 1, None) == abi.WR_ERR_ARG
    f = C.cast(C.c_void_p(8), C.POINTER(C.c_float))
    t = abi.WrAerTables(*([f] * len(abi.AER_TABLE_FLOAT)), *([C.cast(C.c_void_p(8), C.POINTER(C.c_int32))] * 2))
    d = C.cast(C.c_void_p(8), C.POINTER(C.c_double))
    cols.ncol, cols.nlay, cols.p_lev, cols.t_lay, cols.q_lay = 1, 64, d, d, d
    assert L.wr_aer_optics(C.byref(t), C.byref(cols), 8, 8, 8, 8, 8, 1, None) == abi.WR_ERR_ARG
    cols.nlay = 60
    assert L.wr_aer_optics(C.byref(t), C.byref(cols), 8, 8, 8, 8, 8, 0, None) == abi.WR_ERR_ARG   # aer_ld < ncol
    cols.ncol = 0
    assert L.wr_aer_optics(C.byref(t), C.byref(cols), 8, 8, 8, 8, 8, 0, None) == abi.WR_OK       # nothing to do
