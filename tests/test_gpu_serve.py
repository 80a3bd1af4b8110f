"""Served radiation: continuous batching, HBM-budget admission, graph replays (FP64, mixed, and the
surrogate path with the trained network's inference inside the served graphs).

Each request's result must be the same alone and in a mix of other requests (bit for bit: every
column depends only on its inputs, column id and seed), and the same as a direct call.
"""
import numpy as np
import pytest
import torch

from pathlib import Path

from woof_nam_amd import abi, gen
from woof_nam_amd import surrogate as S
from woof_nam_amd.radiation import Radiation
from woof_nam_amd.serve import Request, Server, poisson_run

TRAINED = Path(S.__file__).resolve().parent / "data" / "surrogate_rrtmg.npz"

pytestmark = pytest.mark.gpu


def dev_cols(n, seed, device):
    c = gen.columns(n, seed=seed)
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in c.items()}


def same(a, b):
    return all(torch.equal(a[k], b[k]) for k in abi.OUTPUT_FIELDS)


@pytest.mark.parametrize("precision", ["mixed", "fp64", "surrogate"])
def test_alone_equals_mix_equals_direct(gpu_device, precision):
    print(f"case serve alone-vs-mix {precision}", flush=True)
    a_cols = dev_cols(1500, 1, gpu_device)
    ids = torch.arange(10_000, 11_500, dtype=torch.int64, device=gpu_device)
    net = S.MLP.load(TRAINED, gpu_device) if precision == "surrogate" else None
    srv = Server(0, tile=1024, slots=3, precision=precision, surrogate=net)
    a1 = Request(rid=0, kind="a", cols=a_cols, seed=5, col_id=ids)
    srv.submit(a1)
    srv.drain()
    mix = [Request(rid=1, kind="b", cols=dev_cols(5000, 2, gpu_device), seed=9),
           Request(rid=2, kind="a", cols=a_cols, seed=5, col_id=ids),
           Request(rid=3, kind="c", cols=dev_cols(300, 3, gpu_device), seed=1),
           Request(rid=4, kind="a", cols=a_cols, seed=6, col_id=ids)]
    for r in mix:
        srv.submit(r)
    srv.drain()
    torch.cuda.synchronize()
    assert same(a1.out, mix[1].out)
    assert not torch.equal(a1.out["hr_sw"], mix[3].out["hr_sw"]) or not torch.equal(a1.out["hr_lw"], mix[3].out["hr_lw"])
    rad = Radiation(0, max_cols=4096)
    if net is not None:
        rad.attach_surrogate(net)
    direct, _ = rad.run({**a_cols, "col_id": ids}, params=Radiation.default_params(seed=5), precision=precision)
    torch.cuda.synchronize()
    assert same(a1.out, direct)
    rad.close()
    srv.close()


def test_graph_replays_match_eager(gpu_device):
    """Several replays with eager work and synchronisations in between all equal the eager call."""
    print("case serve graph replays", flush=True)
    cols = dev_cols(2048, 4, gpu_device)
    srv_g = Server(0, tile=2048, slots=1, precision="mixed", graphs=True)
    srv_e = Server(0, tile=2048, slots=1, precision="mixed", graphs=False)
    ref = Request(rid=0, kind="x", cols=cols, seed=3)
    srv_e.submit(ref)
    srv_e.drain()
    for rep in range(4):
        r = Request(rid=rep, kind="x", cols=cols, seed=3)
        srv_g.submit(r)
        srv_g.drain()
        torch.cuda.synchronize()
        assert same(r.out, ref.out), f"replay {rep} differs"
        _ = (cols["t_lay"] * 2.0).sum().item()      # eager work between replays
    srv_g.close()
    srv_e.close()


def test_budget_admission(gpu_device):
    print("case serve budget", flush=True)
    cols = dev_cols(3000, 5, gpu_device)
    one = 3000 * 8 * (6 * 60 + 61 + 4 + 2 * 60 + 4 * 62)
    srv = Server(0, tile=1024, slots=2, budget_gb=1.5 * one / 1e9, precision="mixed")
    reqs = [Request(rid=j, kind="x", cols=cols, seed=j) for j in range(3)]
    for r in reqs:
        srv.submit(r)
    peak = 0
    while not srv.idle():
        srv.poll()
        peak = max(peak, srv.inflight_bytes)
    assert peak <= 1.5 * one
    assert all(r.done is not None for r in reqs)
    assert reqs[0].done <= reqs[1].admitted + 1e-9     # the second waited for the first to leave
    with pytest.raises(MemoryError):
        srv.submit(Request(rid=9, kind="big", cols=dev_cols(6000, 6, gpu_device)))
    srv.close()


def test_poisson_stream_reports_latency(gpu_device):
    print("case serve poisson", flush=True)
    kinds = {"small": (dev_cols(700, 7, gpu_device), None), "large": (dev_cols(3000, 8, gpu_device), None)}
    srv = Server(0, tile=1024, slots=2, precision="mixed")
    res = poisson_run(srv, kinds, rate=50.0, duration=1.0, seed=1)
    srv.close()
    assert res["requests"] > 10
    assert res["p99_ms"] >= res["p50_ms"] > 0
    assert res["columns_per_s"] > 0


def test_autotune_picks_a_measured_shape(gpu_device):
    print("case autotune", flush=True)
    from woof_nam_amd.serve import autotune
    r = autotune(0, tiles=(4096, 8192), slots=(1, 2), ncol=20000, precision="mixed", nreq=5)
    assert len(r["all"]) == 4
    assert all(x["columns_per_s"] > 0 for x in r["all"])
    assert r["best"] == max(r["all"], key=lambda x: x["columns_per_s"])


def grid_dev(nx, ny, seed, dx_km, device):
    c = gen.grid(nx, ny, seed=seed, dx_km=dx_km)
    assert all(np.isfinite(v).all() for v in c.values()), "synthetic grid has non-finite values"
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in c.items()}, c


def test_config3_request_kinds_mix(gpu_device):
    """The request kinds of the served workload at reduced size: a parent-grid sample (12 km, seed 0),
    a 3 km and a 6 km nest, and the 8 ensemble members of the parent (McICA seeds 1..8), mixed
    precision, arriving while earlier ones run.  Every result equals a direct call; the direct FP64
    call of a member equals the CPU oracle on a column subsample, and the member's mixed result is
    within the error bars of the FP64 one."""
    print("case serve config-3 kinds", flush=True)
    from oracle import pyoracle
    parent, parent_h = grid_dev(120, 100, 0, 12.0, gpu_device)
    nest3, _ = grid_dev(70, 60, 11, 3.0, gpu_device)
    nest6, _ = grid_dev(50, 40, 12, 6.0, gpu_device)
    pid = torch.arange(parent["t_lay"].shape[1], dtype=torch.int64, device=gpu_device)
    n3 = torch.arange(10_000_000, 10_000_000 + nest3["t_lay"].shape[1], dtype=torch.int64, device=gpu_device)
    n6 = torch.arange(20_000_000, 20_000_000 + nest6["t_lay"].shape[1], dtype=torch.int64, device=gpu_device)
    kinds = [("parent", parent, pid, 0), ("nest3km", nest3, n3, 0), ("nest6km", nest6, n6, 0)]
    kinds += [("member", parent, pid, m) for m in range(1, 9)]
    srv = Server(0, tile=4096, slots=3, precision="mixed")
    reqs = [Request(rid=j, kind=k, cols=c, col_id=ids, seed=sd) for j, (k, c, ids, sd) in enumerate(kinds)]
    for r in reqs[:4]:
        srv.submit(r)
    for _ in range(3):                     # the rest arrive while the first tiles run
        srv.poll()
    for r in reqs[4:]:
        srv.submit(r)
    srv.drain()
    torch.cuda.synchronize()
    rad = Radiation(0, max_cols=16384)
    for r in reqs:
        direct, _ = rad.run({**r.cols, "col_id": r.col_id}, params=Radiation.default_params(seed=r.seed),
                            precision="mixed")
        torch.cuda.synchronize()
        assert same(r.out, direct), f"request {r.rid} ({r.kind}, seed {r.seed}) differs from the direct call"
    # member 3: FP64 direct call against the oracle on a subsample, and the mixed error bars
    m3 = next(r for r in reqs if r.kind == "member" and r.seed == 3)
    f64, _ = rad.run({**parent, "col_id": pid}, params=Radiation.default_params(seed=3), precision="fp64")
    torch.cuda.synchronize()
    idx = np.arange(0, parent_h["t_lay"].shape[1], 20)
    sub = {k: np.ascontiguousarray(v[..., idx]) for k, v in parent_h.items()}
    sub["col_id"] = idx.astype(np.uint64)
    ref, _ = pyoracle.run(sub, params=Radiation.default_params(seed=3))
    for k in abi.OUTPUT_FIELDS:
        a = f64[k].cpu().numpy()[..., idx]
        assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), f"FP64 {k} differs from the oracle"
    for k in ("hr_lw", "hr_sw"):
        rms = float(torch.sqrt(torch.mean((m3.out[k] - f64[k]) ** 2))) * 86400.0
        assert rms <= 0.05, f"{k} RMS {rms:.3g} K/day"
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        rms = float(torch.sqrt(torch.mean((m3.out[k] - f64[k]) ** 2)))
        assert rms <= 0.5, f"{k} RMS {rms:.3g} W/m2"
    rad.close()
    srv.close()
