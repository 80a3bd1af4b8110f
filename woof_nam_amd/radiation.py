"""Column API of the MI355X radiation path over torch device tensors.

``Radiation(device, max_cols).run(cols, params, precision)`` calls the HIP library on the current
torch stream.  Inputs are float64 device tensors in the column-fastest layout of include/woofrad.h
([nk][ncol], index 0 = model top); outputs are allocated on the same device (or passed in).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib, abi


def _dptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_double)) if t is not None else None


def _check_field(name, t, shape, dtype, device):
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name}: dtype {t.dtype}, expected {dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(t.shape)}, expected {tuple(shape)}")
    if t.device != device:
        raise ValueError(f"{name}: on {t.device}, expected {device}")
    if not t.is_contiguous():
        raise ValueError(f"{name}: must be contiguous")


def input_shapes(ncol: int, nlay: int, path: bool = False) -> dict:
    """Shapes of the column inputs: qc/qi, or with path=True the path form (cwp, cip, crp, rew, rei)."""
    s = {"p_lev": (nlay + 1, ncol)}
    for k in ("t_lay", "q_lay", "o3_lay", "cldf") + (abi.PATH_FIELDS if path else ("qc", "qi")):
        s[k] = (nlay, ncol)
    for k in ("t_sfc", "emis", "alb", "cosz"):
        s[k] = (ncol,)
    return s


def _is_path(cols: dict) -> bool:
    return cols.get("cwp") is not None


def _colstruct(cols, ncol, nlay, device) -> abi.WrColumns:
    """WrColumns over device tensors; optional col_id [ncol] and seed_dev [1] are int64 tensors,
    optional ncol_dev [1] (valid columns, for a replayed graph of a longer tile) an int32 tensor."""
    ids, sd, nd = cols.get("col_id"), cols.get("seed_dev"), cols.get("ncol_dev")
    if ids is not None:
        _check_field("col_id", ids, (ncol,), torch.int64, device)
    if sd is not None:
        _check_field("seed_dev", sd, (1,), torch.int64, device)
    if nd is not None:
        _check_field("ncol_dev", nd, (1,), torch.int32, device)
    u64 = C.POINTER(C.c_uint64)
    f32 = C.POINTER(C.c_float)
    aer, ld = {}, 0
    for k in abi.AER_FIELDS:
        a = cols.get(k)
        if a is not None:
            p, ldk = _aer_ptr(k, a, ncol, nlay, device)
            if ld and ldk != ld:
                raise ValueError("aer_sw and aer_lw must share their column stride")
            aer[k], ld = C.cast(C.c_void_p(p), f32), ldk
    return abi.WrColumns(ncol, nlay, *[_dptr(cols.get(k)) for k in abi.INPUT_FIELDS],
                         C.cast(C.c_void_p(ids.data_ptr()), u64) if ids is not None else None,
                         C.cast(C.c_void_p(sd.data_ptr()), u64) if sd is not None else None,
                         C.cast(C.c_void_p(nd.data_ptr()), C.POINTER(C.c_int32)) if nd is not None else None,
                         *[_dptr(cols.get(k)) for k in abi.PATH_FIELDS], aer.get("aer_sw"), aer.get("aer_lw"), ld,
                         _dptr(_o3_top(cols, ncol, device)))


def _o3_top(cols, ncol, device):
    a = cols.get("o3_top")
    if a is not None:
        _check_field("o3_top", a, (ncol,), torch.float64, device)
    return a


def _aer_ptr(name, a, ncol, nlay, device):
    """(data pointer, column stride) of aerosol optics [3][bands][nlay+1][ncol] float32 on the device:
    contiguous, or a column slice [..., c0:c0+ncol] of a wider [3][bands][nlay+1][ld] array (the
    layout setaer writes over the whole grid)."""
    shp = abi.aer_shape(name, ncol, nlay)
    if not isinstance(a, torch.Tensor) or a.dtype != torch.float32 or a.device != device:
        raise ValueError(f"{name} must be a float32 tensor on {device}")
    if tuple(a.shape) != shp:
        raise ValueError(f"{name} has shape {tuple(a.shape)}, expected {shp}")
    ld = a.stride(2)
    if a.stride(3) != 1 or ld < ncol or a.stride(1) != shp[2] * ld or a.stride(0) != shp[1] * shp[2] * ld:
        raise ValueError(f"{name} must be contiguous or a column slice of a contiguous grid array")
    return a.data_ptr(), ld


class Radiation:
    """A device-resident table set and work space for tiles of up to ``max_cols`` columns."""

    def __init__(self, device: int | torch.device = 0, max_cols: int = 131072, tables=None):
        L = _lib.load()
        self.device = torch.device("cuda", device if isinstance(device, int) else device.index or 0)
        self.max_cols = int(max_cols)
        h = C.c_void_p()
        path = str(tables or _lib.TABLES).encode()
        _lib.check(L.wr_create(C.byref(h), path, self.device.index, self.max_cols), "wr_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            _lib.load().wr_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run_band(self, cols: dict, band: int, params: abi.WrParams | None = None, precision: str = "fp64",
                 out: dict | None = None, stream=None) -> dict:
        """Fluxes and heating rates of one band alone: 0..15 = LW bands 1..16, 16..29 = SW bands 16..29."""
        t = cols["t_lay"]
        nlay, ncol = int(t.shape[0]), int(t.shape[1])
        for k, shp in input_shapes(ncol, nlay).items():
            _check_field(k, cols[k], shp, torch.float64, self.device)
        if out is None:
            out = {k: torch.empty(s, dtype=torch.float64, device=self.device)
                   for k, s in abi.output_shapes(ncol, nlay).items()}
        ci = _colstruct(cols, ncol, nlay, self.device)
        co = abi.WrOutputs(*[_dptr(out[k]) for k in abi.OUTPUT_FIELDS])
        params = params or self.default_params()
        s = stream or torch.cuda.current_stream(self.device)
        rc = _lib.load().wr_run_band(self._h, C.byref(ci), C.byref(params), C.byref(co), int(band),
                                     int(precision == "mixed"), C.c_void_p(s.cuda_stream))
        _lib.check(rc, "wr_run_band")
        return out

    def attach_surrogate(self, mlp) -> None:
        """Use a trained surrogate.MLP for precision="surrogate" calls (its tensors must stay alive); its
        ``bands`` mask (0: all) says which bands take the network's optics, the others the tables."""
        if mlp.norm is None:
            raise ValueError("surrogate has no normalisation: train it (surrogate.fit) or load it first")
        n = mlp.norm
        self._sur_keep = {k: v.float().contiguous() for k, v in n.items()}
        w = abi.WrSurrogate(*[C.c_void_p(t.data_ptr()) for t in (
            mlp.wb["W1"], mlp.wb["W2"], mlp.wb["W3"], mlp.w["b1"], mlp.w["b2"], mlp.w["b3"],
            self._sur_keep["fmean"], self._sur_keep["fstd"], self._sur_keep["fmin"], self._sur_keep["fmax"],
            self._sur_keep["tmean"], self._sur_keep["tstd"])], int(getattr(mlp, "bands", 0)))
        self._sur_mlp = mlp
        _lib.check(_lib.load().wr_surrogate_attach(self._h, C.byref(w)), "wr_surrogate_attach")

    def timing(self, enable: bool = True) -> None:
        """Bracket every later kernel launch with HIP events (per-launch durations, see timing_read)."""
        _lib.check(_lib.load().wr_timing(self._h, int(bool(enable))), "wr_timing")

    def timing_read(self, cap: int = 8192) -> list:
        """[(kernel name, ms), ...] of the launches recorded since timing() or the last read."""
        L = _lib.load()
        kid = (C.c_int32 * cap)()
        ms = (C.c_float * cap)()
        n = C.c_int32()
        _lib.check(L.wr_timing_read(self._h, kid, ms, cap, C.byref(n)), "wr_timing_read")
        return [(L.wr_kernel_name(kid[j]).decode(), float(ms[j])) for j in range(n.value)]

    @staticmethod
    def default_params(**kw) -> abi.WrParams:
        p = abi.WrParams()
        _lib.load().wr_default_params(C.byref(p))
        for k, v in kw.items():
            if k == "vmr":
                for i, x in enumerate(v):
                    p.vmr[i] = float(x)
            else:
                setattr(p, k, v)
        return p

    def run(self, cols: dict, params: abi.WrParams | None = None, precision: str = "fp64",
            out: dict | None = None, diag: bool = False, stream: torch.cuda.Stream | None = None):
        """Radiation call on device columns.  Returns (outputs, diag or None); asynchronous on the stream."""
        if precision not in ("fp64", "mixed", "surrogate"):
            raise ValueError(f"precision must be 'fp64', 'mixed' or 'surrogate', got {precision!r}")
        t = cols["t_lay"]
        nlay, ncol = int(t.shape[0]), int(t.shape[1])
        for k, shp in input_shapes(ncol, nlay, _is_path(cols)).items():
            if k == "crp" and cols.get(k) is None:
                continue
            if k not in cols:
                raise KeyError(f"missing input field {k!r}")
            _check_field(k, cols[k], shp, torch.float64, self.device)
        if out is None:
            out = {k: torch.empty(s, dtype=torch.float64, device=self.device)
                   for k, s in abi.output_shapes(ncol, nlay).items()}
        else:
            for k, s in abi.output_shapes(ncol, nlay).items():
                _check_field(k, out[k], s, torch.float64, self.device)
        dg = cd = None
        if diag:
            dg = {}
            for k, s in abi.diag_shapes(ncol, nlay).items():
                dg[k] = torch.zeros(s, dtype=torch.int32, device=self.device)
            cd = abi.WrDiag(*[C.cast(C.c_void_p(dg[k].data_ptr()),
                                     C.POINTER(C.c_uint32 if k.endswith("mask") else C.c_int32))
                              for k in abi.DIAG_FIELDS])
        ci = _colstruct(cols, ncol, nlay, self.device)
        co = abi.WrOutputs(*[_dptr(out[k]) for k in abi.OUTPUT_FIELDS])
        params = params or self.default_params()
        s = stream or torch.cuda.current_stream(self.device)
        L = _lib.load()
        if precision == "surrogate":
            if diag:
                raise ValueError("diagnostics are not available on the surrogate path")
            flag = torch.zeros(ncol, dtype=torch.int32, device=self.device)
            rc = L.wr_run_surrogate(self._h, C.byref(ci), C.byref(params), C.byref(co), C.c_void_p(flag.data_ptr()),
                                    C.c_void_p(s.cuda_stream))
            _lib.check(rc, "wr_run_surrogate")
            return out, {"fallback": flag}
        fn = L.wr_run_fp64 if precision == "fp64" else L.wr_run_mixed
        rc = fn(self._h, C.byref(ci), C.byref(params), C.byref(co), C.byref(cd) if cd is not None else None,
                C.c_void_p(s.cuda_stream))
        _lib.check(rc, f"wr_run_{precision}")
        return out, dg  # mask words are returned as int32 bit patterns

    def run_noalloc(self, cols: dict, params: abi.WrParams | None = None, out: dict | None = None, stream=None) -> dict:
        """wr_run_mixed_noalloc: the mixed-precision call with the library's static work space (at most
        abi.NOALLOC_MAX_COLS columns; calls on one device must not overlap)."""
        t = cols["t_lay"]
        nlay, ncol = int(t.shape[0]), int(t.shape[1])
        for k, shp in input_shapes(ncol, nlay).items():
            _check_field(k, cols[k], shp, torch.float64, self.device)
        if out is None:
            out = {k: torch.empty(s, dtype=torch.float64, device=self.device)
                   for k, s in abi.output_shapes(ncol, nlay).items()}
        ci = _colstruct(cols, ncol, nlay, self.device)
        co = abi.WrOutputs(*[_dptr(out[k]) for k in abi.OUTPUT_FIELDS])
        params = params or self.default_params()
        s = stream or torch.cuda.current_stream(self.device)
        rc = _lib.load().wr_run_mixed_noalloc(self._h, C.byref(ci), C.byref(params), C.byref(co),
                                              C.c_void_p(s.cuda_stream))
        _lib.check(rc, "wr_run_mixed_noalloc")
        return out


def cal_cldfra(cldfra: torch.Tensor, qc: torch.Tensor, qi: torch.Tensor | None, i_lo: int, i_hi: int,
               j_lo: int, j_hi: int, f_qc: bool = True, f_qi: bool = True, stream=None) -> None:
    """CAL_CLDFRA on [lm][jm][im] float32 device fields (in place on the interior points)."""
    lm, jm, im = cldfra.shape
    for nm, t in (("cldfra", cldfra), ("qc", qc)) + ((("qi", qi),) if qi is not None else ()):
        _check_field(nm, t, (lm, jm, im), torch.float32, cldfra.device)
    s = stream or torch.cuda.current_stream(cldfra.device)
    rc = _lib.load().wr_cal_cldfra(im, jm, lm, i_lo, i_hi, j_lo, j_hi, int(bool(f_qc)), int(bool(f_qi)),
                                   cldfra.data_ptr(), qc.data_ptr(), qi.data_ptr() if qi is not None else None,
                                   C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_cal_cldfra")


def mcica_masks(cldf: torch.Tensor, col_id: torch.Tensor | None, seed: int, iovr: int, stream=None):
    """McICA bits of the 140 LW and 112 SW g-points for a surface-first cloud fraction [nrl][ncol]."""
    nrl, ncol = cldf.shape
    _check_field("cldf", cldf, (nrl, ncol), torch.float64, cldf.device)
    if col_id is not None:
        _check_field("col_id", col_id, (ncol,), torch.int64, cldf.device)
    lw = torch.zeros((nrl, 5, ncol), dtype=torch.int32, device=cldf.device)
    sw = torch.zeros((nrl, 4, ncol), dtype=torch.int32, device=cldf.device)
    s = stream or torch.cuda.current_stream(cldf.device)
    rc = _lib.load().wr_mcica_masks(ncol, nrl, cldf.data_ptr(), col_id.data_ptr() if col_id is not None else None,
                                    C.c_uint64(seed), iovr, lw.data_ptr(), sw.data_ptr(), C.c_void_p(s.cuda_stream))
    _lib.check(rc, "wr_mcica_masks")
    return lw, sw
