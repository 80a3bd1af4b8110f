"""Learned gas-optics surrogate: a small MLP trained on layers the plain RRTMG path generates.

Network (per radiation layer): 5 normalised inputs (ln p, T, ln q_H2O, ln q_O3, dp) -> 128 -> 128
-> 504 outputs (140 LW ln tau, 140 Planck fractions, 112 SW ln tau, 112 SW single-scattering
albedo), SiLU activations.  Every matrix product runs in the library's BF16 MFMA kernels with FP32
accumulation (forward, backward and the Adam step are the library's HIP kernels); PyTorch only
holds the device buffers and drives the loop (batch sampling, normalisation statistics).

Inputs leaving the training envelope (per-feature min/max of the training set) mark the column
for the plain path.
"""
from __future__ import annotations

import ctypes as C
import math
import time

import numpy as np
import torch

from . import _lib, abi
from .radiation import Radiation, _colstruct, input_shapes

K_IN, K_PAD, HIDDEN, N_OUT = 5, 8, 128, 504
SLICES = {"lw_lntau": (0, 140), "lw_frac": (140, 280), "sw_lntau": (280, 392), "sw_ssa": (392, 504)}
# layer optical depths outside [1e-5, 100] change no flux the radiation resolves (a transmittance of
# 1 - 1e-5 or of e^-100): the training targets are clamped to that range, so the network spends its
# precision where the heating rates depend on it
LNTAU_MIN, LNTAU_MAX = math.log(1.0e-5), math.log(100.0)
ALL_BANDS = (1 << 30) - 1


def clamp_targets(targ):
    """ln tau targets (LW and SW) clamped to [ln 1e-5, ln 100], in place."""
    for name in ("lw_lntau", "sw_lntau"):
        a, b = SLICES[name]
        targ[:, a:b].clamp_(LNTAU_MIN, LNTAU_MAX)
    return targ
EPI_STORE, EPI_BIAS, EPI_BIAS_SILU, EPI_DSILU = 0, 1, 2, 3


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def gemm(M, N, K, A, lda_m, lda_k, B, ldb_k, ldb_n, C_=None, Cb=None, ldc_m=0, ldc_n=1, bias=None, Zout=None,
         Zin=None, epi=EPI_STORE, splits=1, split_stride=0):
    """C[M x N] = A[M x K] B[K x N] on the library's BF16 MFMA kernel (strides in elements)."""
    rc = _lib.load().wr_mlp_gemm(M, N, K, _p(A), int(A.dtype == torch.bfloat16), lda_m, lda_k, _p(B),
                                 int(B.dtype == torch.bfloat16), ldb_k, ldb_n, _p(C_), _p(Cb), ldc_m, ldc_n,
                                 _p(bias), _p(Zout), _p(Zin), epi, splits, split_stride, _stream())
    _lib.check(rc, "wr_mlp_gemm")


class MLP:
    """Weights in one flat FP32 buffer (views W1, b1, W2, b2, W3, b3) with a BF16 copy and Adam moments."""

    SHAPES = (("W1", (HIDDEN, K_PAD)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)), ("b2", (HIDDEN,)),
              ("W3", (N_OUT, HIDDEN)), ("b3", (N_OUT,)))

    def __init__(self, device="cuda:0", seed=0):
        self.device = torch.device(device)
        n = sum(int(np.prod(s)) for _, s in self.SHAPES)
        self.flat = torch.zeros(n, dtype=torch.float32, device=self.device)
        self.flat_b = torch.zeros(n, dtype=torch.bfloat16, device=self.device)
        self.grad = torch.zeros(n, dtype=torch.float32, device=self.device)
        self.m1 = torch.zeros_like(self.flat)
        self.m2 = torch.zeros_like(self.flat)
        self.t = 0
        self.w, self.wb, self.g = {}, {}, {}
        o = 0
        for name, shp in self.SHAPES:
            k = int(np.prod(shp))
            self.w[name] = self.flat[o:o + k].view(shp)
            self.wb[name] = self.flat_b[o:o + k].view(shp)
            self.g[name] = self.grad[o:o + k].view(shp)
            o += k
        rng = np.random.default_rng(seed)
        for name, shp in self.SHAPES:
            if name.startswith("W"):
                fan_in = K_IN if name == "W1" else shp[1]
                w = rng.uniform(-1, 1, size=shp) * math.sqrt(3.0 / fan_in)
                if name == "W1":
                    w[:, K_IN:] = 0.0
                self.w[name].copy_(torch.from_numpy(w.astype(np.float32)))
        self.sync_bf16()
        self.norm = None      # feature/target normalisation and envelope (set by fit)
        self.bands = 0        # bands that use the network (bit b: LW band b + 1, b >= 16: SW band b); 0: all
        self._ws = {}

    def sync_bf16(self):
        _lib.check(_lib.load().wr_mlp_to_bf16(_p(self.flat), _p(self.flat_b), self.flat.numel(), _stream()),
                   "wr_mlp_to_bf16")

    def params_numpy(self):
        return tuple(self.w[n].detach().cpu().numpy().astype(np.float64) for n, _ in self.SHAPES)

    def _buf(self, name, shape, dtype):
        t = self._ws.get(name)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = torch.empty(shape, dtype=dtype, device=self.device)
            self._ws[name] = t
        return t

    def forward(self, X, train=False):
        """X: [B][8] float32 (normalised features, columns 5..7 zero).  Returns Y [B][504] float32."""
        B = int(X.shape[0])
        bf = torch.bfloat16
        Z1 = self._buf("Z1", (B, HIDDEN), torch.float32)
        H1 = self._buf("H1", (B, HIDDEN), bf)
        Z2 = self._buf("Z2", (B, HIDDEN), torch.float32)
        H2 = self._buf("H2", (B, HIDDEN), bf)
        Y = self._buf("Y", (B, N_OUT), torch.float32) if train else torch.empty((B, N_OUT), dtype=torch.float32,
                                                                                   device=self.device)
        gemm(B, HIDDEN, K_PAD, X, K_PAD, 1, self.wb["W1"], 1, K_PAD, None, H1, HIDDEN, 1, self.w["b1"], Z1,
             epi=EPI_BIAS_SILU)
        gemm(B, HIDDEN, HIDDEN, H1, HIDDEN, 1, self.wb["W2"], 1, HIDDEN, None, H2, HIDDEN, 1, self.w["b2"], Z2,
             epi=EPI_BIAS_SILU)
        gemm(B, N_OUT, HIDDEN, H2, HIDDEN, 1, self.wb["W3"], 1, HIDDEN, Y, None, N_OUT, 1, self.w["b3"],
             epi=EPI_BIAS)
        return Y

    def forward_fused(self, X, ldy=None):
        """The inference kernel (wr_mlp_forward, the three layers in one kernel): X [B][8] normalised
        features -> Y bfloat16 [504][ldy] output-major (rows >= B of each output untouched)."""
        B = int(X.shape[0])
        ldy = ldy or B
        Y = torch.empty((N_OUT, ldy), dtype=torch.bfloat16, device=self.device)
        w = abi.WrSurrogate(*[C.c_void_p(t.data_ptr()) for t in (self.wb["W1"], self.wb["W2"], self.wb["W3"],
                                                                   self.w["b1"], self.w["b2"], self.w["b3"])])
        X = X.contiguous()
        _lib.check(_lib.load().wr_mlp_forward(C.byref(w), B, _p(X), _p(Y), ldy, _stream()), "wr_mlp_forward")
        return Y

    def _wgrad(self, dZ, H, n_out, n_in, out):
        """out[n_out][n_in] = dZ^T H over the batch (split-K partials, fixed-order sum)."""
        B = int(dZ.shape[0])
        splits = max(1, min(64, B // 1024))
        part = self._buf(f"part{n_out}x{n_in}", (splits, n_out, n_in), torch.float32)
        gemm(n_out, n_in, B, dZ, 1, n_out, H, n_in, 1, part, None, n_in, 1, splits=splits,
             split_stride=n_out * n_in)
        _lib.check(_lib.load().wr_mlp_splitk_reduce(_p(part), splits, n_out * n_in, _p(out), n_out * n_in,
                                                     _stream()), "wr_mlp_splitk_reduce")

    def _colsum(self, X, out):
        M, N = int(X.shape[0]), int(X.shape[1])
        _lib.check(_lib.load().wr_mlp_colsum(_p(X), M, N, N, _p(out), _stream()), "wr_mlp_colsum")

    def loss_and_grad(self, X, T):
        """Forward + MSE + backward on one batch; gradients land in self.grad.  Returns the loss (tensor)."""
        B = int(X.shape[0])
        Y = self.forward(X, train=True)
        dY = self._buf("dY", (B, N_OUT), torch.float32)
        nb = 256
        part = self._buf("msepart", (nb,), torch.float32)
        _lib.check(_lib.load().wr_mlp_mse(_p(Y), _p(T), B * N_OUT, 1.0 / (B * N_OUT), _p(dY), _p(part), nb,
                                          _stream()), "wr_mlp_mse")
        H1, H2 = self._ws["H1"], self._ws["H2"]
        Z1, Z2 = self._ws["Z1"], self._ws["Z2"]
        self._wgrad(dY, H2, N_OUT, HIDDEN, self.g["W3"])
        self._colsum(dY, self.g["b3"])
        dZ2 = self._buf("dZ2", (B, HIDDEN), torch.float32)
        gemm(B, HIDDEN, N_OUT, dY, N_OUT, 1, self.wb["W3"], HIDDEN, 1, dZ2, None, HIDDEN, 1, Zin=Z2, epi=EPI_DSILU)
        self._wgrad(dZ2, H1, HIDDEN, HIDDEN, self.g["W2"])
        self._colsum(dZ2, self.g["b2"])
        dZ1 = self._buf("dZ1", (B, HIDDEN), torch.float32)
        gemm(B, HIDDEN, HIDDEN, dZ2, HIDDEN, 1, self.wb["W2"], HIDDEN, 1, dZ1, None, HIDDEN, 1, Zin=Z1,
             epi=EPI_DSILU)
        self._wgrad(dZ1, X, HIDDEN, K_PAD, self.g["W1"])
        self._colsum(dZ1, self.g["b1"])
        return part.sum() / (B * N_OUT)

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        _lib.check(_lib.load().wr_mlp_adam(_p(self.flat), _p(self.flat_b), _p(self.grad), _p(self.m1), _p(self.m2),
                                           self.flat.numel(), lr, b1, b2, eps, self.t, _stream()), "wr_mlp_adam")

    # ---- persistence ---------------------------------------------------------------------------
    def save(self, path):
        d = {n: self.w[n].detach().cpu().numpy() for n, _ in self.SHAPES}
        for k, v in (self.norm or {}).items():
            d["norm_" + k] = v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v)
        d["bands"] = np.uint32(self.bands)
        np.savez(path, **d)

    @classmethod
    def load(cls, path, device="cuda:0"):
        z = np.load(path, allow_pickle=False)
        m = cls(device)
        for n, _ in cls.SHAPES:
            m.w[n].copy_(torch.from_numpy(z[n]))
        m.sync_bf16()
        m.norm = {k[5:]: torch.from_numpy(z[k]).to(m.device) for k in z.files if k.startswith("norm_")}
        m.bands = int(z["bands"]) if "bands" in z.files else 0
        return m


def layer_optics(rad: Radiation, cols: dict, params=None, targets=True):
    """(feat [S][8], targ [S][504] or None) of every radiation layer, S = (nlay+1) * ncol."""
    nlay, ncol = (int(x) for x in cols["t_lay"].shape)
    S = (nlay + 1) * ncol
    feat = torch.empty((S, K_PAD), dtype=torch.float32, device=rad.device)
    targ = torch.empty((S, N_OUT), dtype=torch.float32, device=rad.device) if targets else None
    ci = _colstruct(cols, ncol, nlay, rad.device)
    params = params or rad.default_params()
    rc = _lib.load().wr_layer_optics(rad._h, C.byref(ci), C.byref(params), _p(feat), _p(targ), _stream())
    _lib.check(rc, "wr_layer_optics")
    return feat, targ


def normalisation(feat, targ):
    """Per-feature / per-output mean and std, and the training envelope (raw feature min/max)."""
    f = feat[:, :K_IN]
    norm = {"fmean": f.mean(0), "fstd": f.std(0).clamp_min(1e-6), "fmin": f.min(0).values, "fmax": f.max(0).values,
            "tmean": targ.mean(0), "tstd": targ.std(0).clamp_min(1e-6)}
    return norm


def apply_feat_norm(feat, norm):
    out = torch.zeros_like(feat)
    out[:, :K_IN] = (feat[:, :K_IN] - norm["fmean"]) / norm["fstd"]
    return out


def fit(mlp: MLP, feat, targ, steps=2000, batch=8192, lr=2e-3, lr_end=1e-4, seed=0, log_every=200, log=print):
    """Train on (feat, targ) (raw, as layer_optics returns them; the ln tau targets are clamped in
    place).  Returns the loss history."""
    clamp_targets(targ)
    norm = normalisation(feat, targ)
    mlp.norm = norm
    X = apply_feat_norm(feat, norm)
    T = (targ - norm["tmean"]) / norm["tstd"]
    S = int(X.shape[0])
    g = torch.Generator(device=X.device)
    g.manual_seed(seed)
    hist = []
    t0 = time.perf_counter()
    for it in range(1, steps + 1):
        idx = torch.randint(0, S, (batch,), device=X.device, generator=g)
        xb = X.index_select(0, idx).contiguous()
        tb = T.index_select(0, idx).contiguous()
        loss = mlp.loss_and_grad(xb, tb)
        frac = (it - 1) / max(1, steps - 1)
        mlp.adam(lr=lr_end + 0.5 * (lr - lr_end) * (1 + math.cos(math.pi * frac)))
        if it == 1 or it % log_every == 0 or it == steps:
            lv = float(loss)
            hist.append((it, lv))
            log(f"step {it}: loss {lv:.4e} ({time.perf_counter() - t0:.1f} s)")
    return hist


def predict(mlp: MLP, feat):
    """Raw outputs (de-normalised) for raw features; also the per-sample in-envelope flag."""
    n = mlp.norm
    X = apply_feat_norm(feat, n)
    Y = mlp.forward(X)
    Y = Y * n["tstd"] + n["tmean"]
    f = feat[:, :K_IN]
    inside = ((f >= n["fmin"]) & (f <= n["fmax"])).all(1)
    return Y, inside
