"""FP64 HIP path against the CPU FP64 oracle on the same seeded columns.

Bar: table indices (jp, jt, jt1, laytrop) and McICA bits bit-exact; fluxes and heating rates within
1e-10 relative (|gpu - oracle| <= 1e-10 * (|oracle| + rms(oracle field))), and in practice bit
for bit, since both evaluate the same IEEE operation sequence without contraction.
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle
from woof_nam_amd import abi, gen
from woof_nam_amd.radiation import Radiation

pytestmark = pytest.mark.gpu

RTOL = 1.0e-10


def to_dev(cols, dev):
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in cols.items() if k != "col_id"}
    if cols.get("col_id") is not None:
        d["col_id"] = torch.from_numpy(cols["col_id"].astype(np.int64)).to(dev)
    return d


def compare(out, ref, tag):
    worst = {}
    for k in abi.OUTPUT_FIELDS:
        a = out[k].cpu().numpy()
        b = ref[k]
        scale = float(np.sqrt(np.mean(b * b))) if b.size else 0.0
        err = np.abs(a - b)
        bad = err > RTOL * (np.abs(b) + scale)
        worst[k] = (float(err.max()) if err.size else 0.0, int(np.count_nonzero(a != b)))
        assert not bad.any(), f"{tag}: {k} exceeds {RTOL} rel at {np.argwhere(bad)[:5].tolist()}, max err {err.max():.3e}"
    return worst


def compare_diag(dg, rdg, tag):
    for k in abi.DIAG_FIELDS:
        a = dg[k].cpu().numpy().view(np.uint32 if k.endswith("mask") else np.int32)
        b = rdg[k].view(np.uint32 if k.endswith("mask") else np.int32)
        mism = np.argwhere(a != b)
        assert mism.size == 0, f"{tag}: diag {k} differs at {mism[:5].tolist()} ({len(mism)} entries)"


@pytest.fixture(scope="module")
def rad(gpu_device):
    r = Radiation(0, max_cols=777)  # several tiles, ragged last tile
    yield r
    r.close()


def run_both(rad, cols, params=None):
    print(f"case ncol={cols['t_lay'].shape[1]} params={params and (params.iovr, params.seed)}", flush=True)
    ref, rdg = pyoracle.run(cols, params=params, diag=True)
    out, dg = rad.run(to_dev(cols, rad.device), params=params, diag=True)
    torch.cuda.synchronize()
    return out, dg, ref, rdg


def test_fp64_matches_oracle(rad):
    cols = gen.columns(2000, seed=11)
    out, dg, ref, rdg = run_both(rad, cols)
    compare_diag(dg, rdg, "default")
    worst = compare(out, ref, "default")
    print("max abs err / differing elements:", worst)


def test_fp64_bit_identical(rad):
    cols = gen.columns(600, seed=5)
    out, _, ref, _ = run_both(rad, cols)
    for k in abi.OUTPUT_FIELDS:
        a = out[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), \
            f"{k}: {np.count_nonzero(a != ref[k])} elements differ, max {np.abs(a - ref[k]).max():.3e}"


def test_glane_lw_kernels(gpu_device, monkeypatch):
    """The opt-in g-point-lane LW kernels (WR_LW_GLANE=1, read when a context is created): FP64
    bit-identical to the oracle; in the mixed path the LW fluxes within 0.01 W/m2 and the LW heating
    rates within 0.01 K/day of the lane-per-column mixed kernels at every level (both FP32 optics and
    sums in g order; the lane-per-column kernel keeps the up sweep's transmittance and source from
    the down sweep where the g-lane kernel recomputes them)."""
    cols = gen.columns(1500, seed=17)
    plain = Radiation(0, max_cols=2048)
    monkeypatch.setenv("WR_LW_GLANE", "1")
    r = Radiation(0, max_cols=2048)
    try:
        out, _, ref, _ = run_both(r, cols)
        for k in abi.OUTPUT_FIELDS:
            a = out[k].cpu().numpy()
            assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), \
                f"{k}: {np.count_nonzero(a != ref[k])} elements differ, max {np.abs(a - ref[k]).max():.3e}"
        dcols = to_dev(cols, r.device)
        mg, _ = r.run(dcols, precision="mixed")
        mc, _ = plain.run(dcols, precision="mixed")
        torch.cuda.synchronize()
        for k in ("flw_up", "flw_dn"):
            assert (mg[k] - mc[k]).abs().max().item() <= 0.01, k
        assert (mg["hr_lw"] - mc["hr_lw"]).abs().max().item() * 86400.0 <= 0.01
    finally:
        r.close()
        plain.close()


def aerosol(ncol, nlay, seed, ld=None):
    """Aerosol optics [3][bands][nlay+1][ld] float32: optical depth, single-scattering albedo, asymmetry."""
    rng = np.random.default_rng(seed)
    ld = ld or ncol
    out = {}
    for k in abi.AER_FIELDS:
        a = np.zeros((3, abi.AER_BANDS[k], nlay + 1, ld), np.float32)
        a[0] = rng.uniform(0.0, 0.5, a[0].shape) * (rng.uniform(size=a[0].shape) < 0.7)
        a[1] = rng.uniform(0.5, 1.0, a[1].shape)
        a[2] = rng.uniform(0.5, 0.8, a[2].shape)
        out[k] = a
    return out


@pytest.mark.parametrize("mcica", [1, 0])
def test_aerosol_matches_oracle(gpu_device, mcica):
    """Aerosol optics on both routes: FP64 bit-identical to the oracle, also when the aerosol arrays
    are a column slice of a wider grid array (setaer's whole-grid layout); mixed within the bar."""
    print(f"case aerosol mcica={mcica}", flush=True)
    ncol, c0, ld = 700, 150, 1100
    cols = gen.columns(ncol, seed=40 + mcica)
    nlay = cols["t_lay"].shape[0]
    wide = aerosol(ncol, nlay, seed=50 + mcica, ld=ld)
    host = {k: np.ascontiguousarray(v[..., c0:c0 + ncol]) for k, v in wide.items()}
    p = Radiation.default_params(mcica=mcica, seed=3)
    ref, _ = pyoracle.run({**cols, **host}, params=p)
    plain, _ = pyoracle.run(cols, params=p)
    assert np.abs(ref["flw_dn"] - plain["flw_dn"]).max() > 1.0 and np.abs(ref["fsw_dn"] - plain["fsw_dn"]).max() > 1.0
    r = Radiation(0, max_cols=1024)
    try:
        dcols = to_dev(cols, r.device)
        dwide = {k: torch.from_numpy(v).to(r.device) for k, v in wide.items()}
        for aer in ({k: torch.from_numpy(v).to(r.device) for k, v in host.items()},
                    {k: v[..., c0:c0 + ncol] for k, v in dwide.items()}):
            out, _ = r.run({**dcols, **aer}, params=p)
            torch.cuda.synchronize()
            for k in abi.OUTPUT_FIELDS:
                a = out[k].cpu().numpy()
                assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), \
                    f"{k}: {np.count_nonzero(a != ref[k])} elements differ, max {np.abs(a - ref[k]).max():.3e}"
        mx, _ = r.run({**dcols, **aer}, params=p, precision="mixed")
        torch.cuda.synchronize()
        for k in ("hr_lw", "hr_sw"):
            d = (mx[k].cpu().numpy() - ref[k]) * 86400.0
            assert np.sqrt(np.mean(d * d)) <= 0.05, k
        for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
            d = mx[k].cpu().numpy() - ref[k]
            assert np.sqrt(np.mean(d * d)) <= 0.5, k
    finally:
        r.close()


def test_day_list_over_several_segments(gpu_device):
    """One 7,000-column tile: the day list is compacted from four 2,048-column segments, in order."""
    cols = gen.columns(7000, seed=13)
    nday = int(np.count_nonzero(cols["cosz"] > 0.0))
    assert 0 < nday < 7000
    r = Radiation(0, max_cols=8192)
    try:
        out, _, ref, _ = run_both(r, cols)
    finally:
        r.close()
    for k in abi.OUTPUT_FIELDS:
        a = out[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), \
            f"{k}: {np.count_nonzero(a != ref[k])} elements differ, max {np.abs(a - ref[k]).max():.3e}"


@pytest.mark.parametrize("iovr", [0, 1, 2])
def test_overlap_and_seed(rad, iovr):
    cols = gen.columns(500, seed=20 + iovr)
    cols["col_id"] = (np.arange(500, dtype=np.uint64) * np.uint64(7919) + np.uint64(2**33 + 5))
    p = Radiation.default_params(iovr=iovr, seed=0x9E3779B97F4A7C15)
    out, dg, ref, rdg = run_both(rad, cols, p)
    compare_diag(dg, rdg, f"iovr={iovr}")
    compare(out, ref, f"iovr={iovr}")


def test_edge_columns(rad):
    """night-only, overcast, clear, ice-only and sun-at-horizon columns."""
    cols = gen.columns(320, seed=3)
    n = 320
    cols["cosz"][:64] = -0.5                  # night
    cols["cosz"][64:96] = 1.0e-4              # exactly at the day threshold
    cols["cosz"][96:128] = 0.99999e-4         # just below it
    cols["cldf"][:, 128:192] = 1.0            # overcast
    cols["cldf"][:, 192:256] = 0.0            # clear
    cols["qc"][:, 256:] = 0.0                 # ice only
    cols["cldf"][:, 256:] = 0.0009            # below the 0.001 cloud threshold in half the layers
    cols["cldf"][::2, 256:] = 0.6
    out, dg, ref, rdg = run_both(rad, cols)
    compare_diag(dg, rdg, "edge")
    compare(out, ref, "edge")
    assert float(out["fsw_dn"][:, :64].abs().max()) == 0.0


def test_single_and_empty(rad):
    cols = gen.columns(1, seed=9)
    out, dg, ref, rdg = run_both(rad, cols)
    compare(out, ref, "single")
    cols0 = {k: v[..., :0] for k, v in gen.columns(2, seed=1).items()}
    o0, _ = rad.run(to_dev(cols0, rad.device))
    torch.cuda.synchronize()
    assert o0["hr_lw"].shape == (60, 0)


def test_bad_arguments(rad):
    cols = to_dev(gen.columns(4, seed=1), rad.device)
    bad = dict(cols)
    bad["t_lay"] = bad["t_lay"].float()
    with pytest.raises(ValueError):
        rad.run(bad)
    with pytest.raises(ValueError):
        rad.run(cols, params=Radiation.default_params(mcica=2))
    with pytest.raises(ValueError):
        rad.run(cols, params=Radiation.default_params(iovr=3))


def test_band_runs_sum_to_full_call(rad):
    """wr_run_band: the 16 LW and 14 SW single-band calls add up to the full call."""
    cols = gen.columns(300, seed=44)
    dev = to_dev(cols, rad.device)
    full, _ = rad.run(dev)
    acc = {k: torch.zeros_like(v) for k, v in full.items()}
    for b in range(30):
        o = rad.run_band(dev, b)
        for k in acc:
            acc[k] += o[k]
    torch.cuda.synchronize()
    for k in abi.OUTPUT_FIELDS:
        a, b = acc[k].cpu().numpy(), full[k].cpu().numpy()
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=1e-9 * (np.abs(b).max() + 1e-300), err_msg=k)


@pytest.mark.parametrize("seed", [61, 62])
def test_nam_route_matches_oracle(rad, seed):
    """mcica = 0: rtrnmr (maximum-random fractional clouds) and cldfrc-mixed spcvrtc with zcf1/zcf0
    weighting, the route NAM runs; bit for bit against the oracle, indices bit-exact, masks zero."""
    cols = gen.columns(1500, seed=seed)
    cols["cldf"][:, :100] = 1.0                   # overcast columns: zcf0 = 0
    cols["cldf"][:, 100:200] = 0.0                # clear columns: zcf1 = 0, SW clear pass only
    cols["cldf"][20:30, 200:300] = 0.0            # cloud blocks separated by clear layers
    p = Radiation.default_params(mcica=0)
    out, dg, ref, rdg = run_both(rad, cols, p)
    compare_diag(dg, rdg, "nam")
    assert not dg["lw_mask"].any() and not dg["sw_mask"].any()
    compare(out, ref, "nam")
    for k in abi.OUTPUT_FIELDS:
        a = out[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), \
            f"{k}: {np.count_nonzero(a != ref[k])} elements differ, max {np.abs(a - ref[k]).max():.3e}"


@pytest.mark.parametrize("mcica", [1, 0])
def test_valid_column_count_skips_the_rest(rad, mcica):
    """ncol_dev = m < ncol (a replayed graph of a longer tile): the first m columns equal a call on
    those m columns alone, bit for bit, and the outputs of the others are left as they were."""
    print(f"case ncol_dev mcica={mcica}", flush=True)
    cols = to_dev(gen.columns(700, seed=71), rad.device)
    m = 413
    params = Radiation.default_params(mcica=mcica)
    out = {k: torch.full(s, -7.0, dtype=torch.float64, device=rad.device)
           for k, s in abi.output_shapes(700, 60).items()}
    part = dict(cols, ncol_dev=torch.tensor([m], dtype=torch.int32, device=rad.device))
    rad.run(part, params=params, out=out)
    alone, _ = rad.run({k: v[..., :m].contiguous() for k, v in cols.items()}, params=params)
    torch.cuda.synchronize()
    for k in abi.OUTPUT_FIELDS:
        assert torch.equal(out[k][..., :m], alone[k]), k
        assert bool((out[k][..., m:] == -7.0).all()), k
