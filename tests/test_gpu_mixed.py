"""Mixed-precision path against the FP64 oracle.

Bar (stated in W/m2 and K/day): heating-rate RMS <= 0.05 K/day and flux RMS <= 0.5 W/m2 over all
levels and columns; every column's worst heating-rate error is also bounded.  Table indices and
McICA bits come from FP64 arithmetic in both paths and must be bit-identical.
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle
from woof_nam_amd import abi, gen
from woof_nam_amd.radiation import Radiation

pytestmark = pytest.mark.gpu

KDAY = 86400.0
HR_RMS_KDAY = 0.05
FLUX_RMS_WM2 = 0.5


@pytest.fixture(scope="module")
def rad(gpu_device):
    r = Radiation(0, max_cols=4096)
    yield r
    r.close()


def errors(out, ref):
    e = {}
    for k in ("hr_lw", "hr_sw"):
        d = out[k].cpu().numpy() - ref[k]
        e[k] = (float(np.sqrt(np.mean(d * d)) * KDAY), float(np.abs(d).max() * KDAY))
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        d = out[k].cpu().numpy() - ref[k]
        e[k] = (float(np.sqrt(np.mean(d * d))), float(np.abs(d).max()))
    return e


@pytest.mark.parametrize("seed", [31, 32])
def test_mixed_within_tolerance(rad, seed):
    cols = gen.columns(3000, seed=seed)
    print(f"case mixed ncol=3000 seed={seed}", flush=True)
    ref, rdg = pyoracle.run(cols, diag=True)
    dev = {k: torch.from_numpy(v).to(rad.device) for k, v in cols.items()}
    out, dg = rad.run(dev, precision="mixed", diag=True)
    torch.cuda.synchronize()
    e = errors(out, ref)
    print("RMS / max error (K/day, W/m2):", e)
    for k in ("hr_lw", "hr_sw"):
        assert e[k][0] <= HR_RMS_KDAY, (k, e[k])
        assert e[k][1] <= 20 * HR_RMS_KDAY, (k, e[k])
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        assert e[k][0] <= FLUX_RMS_WM2, (k, e[k])
    for k in abi.DIAG_FIELDS:
        assert np.array_equal(dg[k].cpu().numpy().view(np.uint32), rdg[k].view(np.uint32)), k


def test_mixed_grid_sample(rad):
    """Grid columns (smooth fields, 12 UTC sun) including the low-sun edge of the domain."""
    cols = gen.grid(120, 100, seed=3)
    print("case mixed grid 120x100", flush=True)
    ref, _ = pyoracle.run(cols)
    dev = {k: torch.from_numpy(v).to(rad.device) for k, v in cols.items()}
    out, _ = rad.run(dev, precision="mixed")
    torch.cuda.synchronize()
    e = errors(out, ref)
    print("RMS / max error (K/day, W/m2):", e)
    for k in ("hr_lw", "hr_sw"):
        assert e[k][0] <= HR_RMS_KDAY, (k, e[k])
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        assert e[k][0] <= FLUX_RMS_WM2, (k, e[k])


def test_mixed_nam_route_within_tolerance(rad):
    cols = gen.columns(2500, seed=71)
    print("case mixed nam route", flush=True)
    p = Radiation.default_params(mcica=0)
    ref, _ = pyoracle.run(cols, params=p)
    dev = {k: torch.from_numpy(v).to(rad.device) for k, v in cols.items()}
    out, _ = rad.run(dev, params=p, precision="mixed")
    torch.cuda.synchronize()
    e = errors(out, ref)
    print("RMS / max error (K/day, W/m2):", e)
    for k in ("hr_lw", "hr_sw"):
        assert e[k][0] <= HR_RMS_KDAY, (k, e[k])
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        assert e[k][0] <= FLUX_RMS_WM2, (k, e[k])


@pytest.mark.parametrize("precision,envs", [("mixed", ("WR_LW_LDS", "WR_SW_LDS")),
                                            ("fp64", ("WR_LW_LDS64", "WR_SW_LDS64"))])
def test_lds_staged_band_kernels_match_global_tables(gpu_device, monkeypatch, precision, envs):
    """The band kernels with the band tables staged in LDS (k_lwl / k_swl for FP32, k_lwl64 / k_swl
    for FP64) do the arithmetic of the kernels that gather from global memory (k_lwc / k_sw, k_lw /
    k_sw), on grid columns with several 64-column blocks per wave and a partly filled last block.
    FP64 (no contraction) and the FP32 LW outputs are bit-identical.  The FP32 SW kernels are
    compiled with multiply-add contraction, which the compiler may form differently in the two
    instances: their outputs agree to FP32 rounding (fluxes within 1e-6 of the field's largest value,
    heating rates within 1e-4 K/day, 1/500 of the mixed path's bar)."""
    cols = gen.grid(150, 131, seed=5)  # 19,650 columns
    print(f"case LDS-staged band kernels vs global tables, {precision}, grid 150x131", flush=True)
    dev = {k: torch.from_numpy(v).to(gpu_device) for k, v in cols.items()}
    outs = {}
    for mask in ("0", "ffff"):
        for env in envs:
            monkeypatch.setenv(env, mask)
        r = Radiation(0, max_cols=20000)
        try:
            out, _ = r.run(dev, precision=precision)
            torch.cuda.synchronize()
            outs[mask] = {k: v.cpu().numpy().copy() for k, v in out.items()}
        finally:
            r.close()
    for k, v in outs["0"].items():
        w = outs["ffff"][k]
        if precision == "mixed" and k.startswith("hr_sw"):
            assert float(np.abs(v - w).max()) * KDAY <= 1e-4, k
        elif precision == "mixed" and "sw" in k:
            assert float(np.abs(v - w).max()) <= 1e-6 * max(float(np.abs(v).max()), 1e-30), k
        else:
            assert np.array_equal(v.view(np.uint64), w.view(np.uint64)), k


@pytest.mark.parametrize("precision,envs", [("mixed", ("WR_LWR_LDS",)), ("fp64", ("WR_LWR_LDS64", "WR_SWR_LDS64"))])
def test_lds_staged_nam_route_kernels_match_global_tables(gpu_device, monkeypatch, precision, envs):
    """The NAM route's kernels with the band tables in LDS (k_lwrl, and k_swrl in FP64) against those
    that gather from global memory (k_lwr, k_swr): bit-identical outputs."""
    cols = gen.grid(150, 131, seed=6)
    print(f"case NAM-route LDS-staged LW kernels vs global tables, {precision}, grid 150x131", flush=True)
    dev = {k: torch.from_numpy(v).to(gpu_device) for k, v in cols.items()}
    p = Radiation.default_params(mcica=0)
    outs = {}
    for mask in ("0", "ffff"):
        for env in envs:
            monkeypatch.setenv(env, mask)
        r = Radiation(0, max_cols=20000)
        try:
            out, _ = r.run(dev, params=p, precision=precision)
            torch.cuda.synchronize()
            outs[mask] = {k: v.cpu().numpy().copy() for k, v in out.items()}
        finally:
            r.close()
    for k, v in outs["0"].items():
        assert np.array_equal(v.view(np.uint64), outs["ffff"][k].view(np.uint64)), k
