"""BF16 MFMA kernels of the surrogate: GEMM against float64 products of the bf16-rounded operands,
gradients against the CPU oracle (oracle/mlp.py), and a short training run that must learn."""
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


def test_surrogate_path_runs_and_falls_back(gpu_device):
    """The surrogate call: finite outputs, table fallback for a column outside the training envelope
    (that column then follows the plain mixed path to FP32 rounding: the two kernels contract
    multiply-adds independently), heating rates close to FP64."""
    print("case surrogate path", flush=True)
    rad = Radiation(0, max_cols=2048)
    train = {k: torch.from_numpy(np.ascontiguousarray(v)).to(gpu_device) for k, v in gen.columns(3000, seed=41).items()}
    feat, targ = S.layer_optics(rad, train)
    m = S.MLP(gpu_device, seed=2)
    S.fit(m, feat, targ, steps=600, batch=4096, lr=3e-3, log_every=200)
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
    hr = float(torch.sqrt(torch.mean((out["hr_lw"] - ref["hr_lw"]) ** 2))) * 86400
    print("surrogate LW HR RMS K/day (short training):", hr)
    assert hr < 5.0, hr
    rad.close()


def test_trained_network_gains_over_mean_optics(gpu_device):
    """What the learned part gains over the method without it.  The same surrogate call once with the
    trained network and once with its output layer zeroed, which leaves every output at its training
    mean: gas optics of each g-point fixed at their average over the training layers, with no
    dependence on pressure, temperature or absorber amounts.  Heating-rate RMS against FP64."""
    print("case surrogate gain over mean optics", flush=True)
    rad = Radiation(0, max_cols=2048)
    train = {k: torch.from_numpy(np.ascontiguousarray(v)).to(gpu_device) for k, v in gen.columns(3000, seed=51).items()}
    feat, targ = S.layer_optics(rad, train)
    m = S.MLP(gpu_device, seed=3)
    S.fit(m, feat, targ, steps=600, batch=4096, lr=3e-3, log_every=200)
    base = S.MLP(gpu_device, seed=3)
    base.norm = m.norm
    base.w["W3"].zero_()
    base.w["b3"].zero_()
    base.sync_bf16()
    dev = {k: torch.from_numpy(np.ascontiguousarray(v)).to(gpu_device) for k, v in gen.columns(1500, seed=52).items()}
    ref, _ = rad.run(dev, precision="fp64")
    err = {}
    for name, net in (("trained", m), ("mean", base)):
        rad.attach_surrogate(net)
        out, info = rad.run(dev, precision="surrogate")
        torch.cuda.synchronize()
        err[name] = [float(torch.sqrt(torch.mean((out[k] - ref[k]) ** 2))) * 86400 for k in ("hr_lw", "hr_sw")]
    print("heating-rate RMS K/day [LW, SW]: trained", err["trained"], "mean optics", err["mean"])
    assert err["trained"][0] < 0.5 * err["mean"][0], err
    assert err["trained"][1] < 0.5 * err["mean"][1], err
    rad.close()
