"""BF16 MFMA kernels of the surrogate: GEMM against float64 products of the bf16-rounded operands,
gradients against the CPU oracle (oracle/mlp.py), a short training run that must learn, and the
trained network (woof_nam_amd/data/surrogate_rrtmg.npz) on the surrogate path: fallback, the
accuracy bars and the trade against the table paths."""
import time
from pathlib import Path

import numpy as np
import pytest
import torch

from oracle import mlp as omlp
from woof_nam_amd import abi, gen, surrogate as S
from woof_nam_amd.radiation import Radiation

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (100, 37, 29), (513, 130, 200), (7, 504, 128)])
@pytest.mark.parametrize("abf,bbf", [(False, False), (True, False), (False, True), (True, True)])
def test_gemm_matches_bf16_product(gpu_device, M, N, K, abf, bbf):
    print(f"case gemm {M}x{N}x{K} a_bf16={abf} b_bf16={bbf}", flush=True)
    rng = np.random.default_rng(M * 7 + N)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    bias = rng.normal(size=N).astype(np.float32)
    ref = omlp.to_bf16(A) @ omlp.to_bf16(B) + bias
    dA = torch.from_numpy(A).to(gpu_device)
    dBt = torch.from_numpy(np.ascontiguousarray(B.T)).to(gpu_device)   # B stored transposed [N][K]
    if abf:
        dA = dA.to(torch.bfloat16)
    if bbf:
        dBt = dBt.to(torch.bfloat16)
    C = torch.empty((M, N), dtype=torch.float32, device=gpu_device)
    S.gemm(M, N, K, dA, K, 1, dBt, 1, K, C, None, N, 1, torch.from_numpy(bias).to(gpu_device), epi=S.EPI_BIAS)
    torch.cuda.synchronize()
    got = C.cpu().numpy()
    err = np.abs(got - ref).max() / (np.abs(ref).max() + 1e-30)
    assert err < 2e-6, err


def test_gemm_split_k_and_silu(gpu_device):
    rng = np.random.default_rng(3)
    M, N, K = 96, 80, 3000
    A = rng.normal(size=(K, M)).astype(np.float32)      # A^T stored: element (m,k) at A[k][m]
    B = rng.normal(size=(K, N)).astype(np.float32)
    ref = omlp.to_bf16(A).T @ omlp.to_bf16(B)
    dA, dB = torch.from_numpy(A).to(gpu_device), torch.from_numpy(B).to(gpu_device)
    splits = 6
    part = torch.empty((splits, M, N), dtype=torch.float32, device=gpu_device)
    S.gemm(M, N, K, dA, 1, M, dB, N, 1, part, None, N, 1, splits=splits, split_stride=M * N)
    out = torch.empty((M, N), dtype=torch.float32, device=gpu_device)
    from woof_nam_amd import _lib
    _lib.check(_lib.load().wr_mlp_splitk_reduce(S._p(part), splits, M * N, S._p(out), M * N, S._stream()), "reduce")
    torch.cuda.synchronize()
    assert np.abs(out.cpu().numpy() - ref).max() / np.abs(ref).max() < 2e-6
    # bias + SiLU epilogue keeps the pre-activation and writes a bf16 copy
    bias = torch.zeros(N, device=gpu_device)
    Z = torch.empty((M, N), dtype=torch.float32, device=gpu_device)
    H = torch.empty((M, N), dtype=torch.bfloat16, device=gpu_device)
    S.gemm(M, N, K, dA, 1, M, dB, N, 1, None, H, N, 1, bias, Z, epi=S.EPI_BIAS_SILU)
    torch.cuda.synchronize()
    z = Z.cpu().numpy()
    assert np.abs(z - ref).max() / np.abs(ref).max() < 2e-6
    np.testing.assert_allclose(H.float().cpu().numpy(), omlp.to_bf16(omlp.silu(z.astype(np.float64))), rtol=1e-2,
                               atol=1e-2)


def test_gradients_match_cpu_oracle(gpu_device):
    print("case mlp gradients vs oracle", flush=True)
    m = S.MLP(gpu_device, seed=4)
    rng = np.random.default_rng(5)
    B = 2048
    x = np.zeros((B, S.K_PAD), np.float32)
    x[:, :S.K_IN] = rng.normal(size=(B, S.K_IN))
    t = rng.normal(size=(B, S.N_OUT)).astype(np.float32)
    loss = float(m.loss_and_grad(torch.from_numpy(x).to(gpu_device), torch.from_numpy(t).to(gpu_device)))
    torch.cuda.synchronize()
    P = m.params_numpy()
    lref, gref = omlp.loss_and_grads(P, x.astype(np.float64), t.astype(np.float64), bf16=True)
    assert abs(loss - lref) <= 1e-4 * lref
    for (name, _), g in zip(S.MLP.SHAPES, gref):
        got = m.g[name].cpu().numpy().astype(np.float64)
        rel = np.linalg.norm(got - g) / (np.linalg.norm(g) + 1e-30)
        print(name, "relative gradient error", rel)
        assert rel < 5e-3, (name, rel)


@pytest.mark.parametrize("M,pad", [(5000, 0), (4999, 5), (64, 0), (1, 7)])
def test_fused_forward_equals_layer_chain(gpu_device, M, pad):
    """wr_mlp_forward (the inference kernel) against the three wr_mlp_gemm layers: the same BF16
    outputs bit for bit, with aligned (vector) and unaligned output strides and a ragged last block."""
    print(f"case fused forward M={M} pad={pad}", flush=True)
    m = S.MLP.load(TRAINED, gpu_device)
    rng = np.random.default_rng(M + pad)
    x = np.zeros((M, S.K_PAD), np.float32)
    x[:, :S.K_IN] = rng.normal(size=(M, S.K_IN))
    X = torch.from_numpy(x).to(gpu_device)
    chain = m.forward(X).to(torch.bfloat16)               # [M][504]
    Y = m.forward_fused(X, ldy=M + pad)                    # [504][M + pad]
    torch.cuda.synchronize()
    assert torch.equal(Y[:, :M].t(), chain), int((Y[:, :M].t() != chain).sum())


def test_training_reduces_loss_on_rrtmg_layers(gpu_device):
    print("case mlp short training", flush=True)
    rad = Radiation(0, max_cols=4096)
    cols = gen.columns(3000, seed=21)
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(gpu_device) for k, v in cols.items()}
    feat, targ = S.layer_optics(rad, dev)
    torch.cuda.synchronize()
    assert torch.isfinite(feat).all() and torch.isfinite(targ).all()
    m = S.MLP(gpu_device, seed=1)
    hist = S.fit(m, feat, targ, steps=400, batch=4096, lr=3e-3, log_every=100)
    assert hist[-1][1] < 0.2 * hist[0][1], hist
    Y, inside = S.predict(m, feat)
    assert bool(inside.all())
    rad.close()


TRAINED = Path(S.__file__).resolve().parent / "data" / "surrogate_rrtmg.npz"
BAR_HR_KDAY, BAR_FLUX_WM2 = 0.05, 0.5


def _timed(rad, dev, prec, reps=3):
    out, info = rad.run(dev, precision=prec)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        rad.run(dev, precision=prec, out=out)
    torch.cuda.synchronize()
    return out, info, (time.perf_counter() - t0) / reps


def _errs(out, ref):
    rms = lambda k: float(torch.sqrt(torch.mean((out[k] - ref[k]) ** 2)))
    return {"hr_lw": rms("hr_lw") * 86400, "hr_sw": rms("hr_sw") * 86400,
            "flux_lw": max(rms("flw_up"), rms("flw_dn")), "flux_sw": max(rms("fsw_up"), rms("fsw_dn"))}


def test_surrogate_path_runs_and_falls_back(gpu_device):
    """The surrogate call with the trained network (woof_nam_amd/data/surrogate_rrtmg.npz): finite
    outputs, table fallback for a column outside the training envelope (that column then follows the
    plain mixed path to FP32 rounding: the two kernels contract multiply-adds independently), and
    heating rates and fluxes within the bars against FP64."""
    print("case surrogate path", flush=True)
    rad = Radiation(0, max_cols=2048)
    m = S.MLP.load(TRAINED, gpu_device)
    assert m.bands != 0 and not (m.bands >> 3) & 1          # LW band 4 keeps its tables
    rad.attach_surrogate(m)
    cols = gen.columns(1500, seed=42)
    cols["t_lay"][:, 7] = 345.0                     # far outside the training temperatures
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(gpu_device) for k, v in cols.items()}
    out, info = rad.run(dev, precision="surrogate")
    mix, _ = rad.run(dev, precision="mixed")
    ref, _ = rad.run(dev, precision="fp64")
    torch.cuda.synchronize()
    flag = info["fallback"].cpu().numpy()
    assert flag[7] == 1 and flag.sum() < 20, (int(flag[7]), int(flag.sum()))
    for k in abi.OUTPUT_FIELDS:
        assert torch.isfinite(out[k]).all(), k
        tol = 1e-3 / 86400 if k.startswith("hr") else 1e-2      # 1e-3 K/day, 1e-2 W/m2
        d = float((out[k][..., 7] - mix[k][..., 7]).abs().max())
        assert d <= tol, (k, d, tol)
    e = _errs(out, ref)
    print("surrogate errors vs FP64 (K/day, W/m2):", e)
    assert e["hr_lw"] <= BAR_HR_KDAY and e["hr_sw"] <= BAR_HR_KDAY, e
    assert e["flux_lw"] <= BAR_FLUX_WM2 and e["flux_sw"] <= BAR_FLUX_WM2, e
    rad.close()


def test_trained_network_against_table_path(gpu_device):
    """What the trained network gives against the methods without it, on 100,000 columns of a grid it
    was not trained or selected on (generator seed 2):
    * against the plain FP64 table path (the oracle's arithmetic): faster, with heating rates and
      fluxes within the bars (0.05 K/day, 0.5 W/m2);
    * against the mixed-precision table path: slower, and less accurate (the trade it is);
    * against the same network with its output layer zeroed (every gas optical property fixed at
      its training mean, no dependence on the layer's state): far more accurate."""
    print("case surrogate against the table path", flush=True)
    grid = gen.grid(400, 250, seed=2)
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(gpu_device) for k, v in grid.items()}
    rad = Radiation(0, max_cols=100000)
    m = S.MLP.load(TRAINED, gpu_device)
    rad.attach_surrogate(m)
    ref, _, t64 = _timed(rad, dev, "fp64")
    mix, _, tmx = _timed(rad, dev, "mixed")
    out, info, tsu = _timed(rad, dev, "surrogate")
    e, emx = _errs(out, ref), _errs(mix, ref)
    n = int(dev["t_lay"].shape[1])
    print(f"columns/s: fp64 {n / t64:.4g}, mixed {n / tmx:.4g}, surrogate {n / tsu:.4g} "
          f"(fallback columns {int(info['fallback'].sum())})", flush=True)
    print("errors vs FP64: surrogate", e, "mixed", emx, flush=True)
    assert e["hr_lw"] <= BAR_HR_KDAY and e["hr_sw"] <= BAR_HR_KDAY, e
    assert e["flux_lw"] <= BAR_FLUX_WM2 and e["flux_sw"] <= BAR_FLUX_WM2, e
    assert t64 > 1.1 * tsu, (t64, tsu)                   # the network path beats the FP64 tables
    assert emx["hr_lw"] < e["hr_lw"] and tmx < tsu       # the mixed tables are faster and closer
    base = S.MLP.load(TRAINED, gpu_device)
    base.w["W3"].zero_()
    base.w["b3"].zero_()
    base.sync_bf16()
    rad.attach_surrogate(base)
    mean_out, _ = rad.run(dev, precision="surrogate")
    eb = _errs(mean_out, ref)
    print("errors vs FP64 with mean optics:", eb, flush=True)
    assert e["hr_lw"] < 0.1 * eb["hr_lw"] and e["hr_sw"] < 0.1 * eb["hr_sw"], (e, eb)
    rad.close()
