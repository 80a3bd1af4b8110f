"""Radiation benchmark: LW+SW RRTMG columns/s on the 954 x 835 x 60 NAM 12 km grid on one MI355X.

One step = one radiation call (LW + SW, McICA) over all 796,590 columns with the inputs already
resident in HBM.  Prints ONE JSON line (rank 0).  Also reported beside the throughput:
  * latency_ms: per-call p50/p99 from HIP events on the call's stream;
  * accuracy: heating-rate (K/day) and flux (W/m2) RMS against the CPU FP64 oracle on a sample;
  * roofline: the dominant kernel's algorithmic HBM bytes / its mean launch duration (HIP events
    on the launching stream, recorded by the library around every launch);
  * cpu_baseline: the CPU FP64 oracle on the 4,096-column CPU-reference workload, on this host's
    cores (median of 5 after 1 warm-up), with nproc and the CPU model;
  * aerosol: the climatological aerosol optics (wr_aer_optics) on the same columns, ms per call and
    its HBM roofline (the headline call runs without aerosol, as the configuration states none).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "LW+SW radiation columns/s (60 lev) + p99 latency, at stated heating-rate error"
HBM_PEAK_GBS = 8000.0
KDAY = 86400.0

# Algorithmic HBM bytes per column of each kernel (FP64 inputs it must read, outputs it must write).
# Column API inputs: 6 layer fields x 60 + p_lev 61 + 4 surface scalars; outputs: 2 HR x 60 + 4 x 62 levels.
NL = 60
IN_LW = (6 * NL + (NL + 1) + 2) * 8          # T, q, O3, cldf, qc, qi, p_lev, T_s, emissivity
IN_SW = (6 * NL + (NL + 1) + 3) * 8          # + albedo, cos z (day columns only)
OUT_ALL = (2 * NL + 4 * (NL + 2)) * 8


# Measured HBM bytes per column (LW) or per day column (SW) of the band kernels: rocprofv3 FETCH_SIZE
# and WRITE_SIZE on a 160,000-column call, HBM bytes = 2 x FETCH + WRITE (calibrated on a 1 GiB
# device copy); the tables and their commands are profiles/hbm_traffic_{fp64,mixed}_grid400x400.md.
TRAFFIC_FP64 = {
    "k_lw_b01": 11692, "k_lw_b02": 11410, "k_lw_b03": 7700, "k_lw_b04": 8552, "k_lw_b05": 8665,
    "k_lw_b06": 9819, "k_lw_b07": 8433, "k_lw_b08": 11028, "k_lw_b09": 7461, "k_lw_b10": 9091,
    "k_lw_b11": 9852, "k_lw_b12": 9941, "k_lw_b13": 8001, "k_lw_b14": 7667, "k_lw_b15": 8471, "k_lw_b16": 7477,
    "k_sw_b16": 21335, "k_sw_b17": 33790, "k_sw_b18": 25222, "k_sw_b19": 25176, "k_sw_b20": 29498,
    "k_sw_b21": 29547, "k_sw_b22": 12241, "k_sw_b23": 29526, "k_sw_b24": 26168, "k_sw_b25": 21989,
    "k_sw_b26": 19979, "k_sw_b27": 26218, "k_sw_b28": 22202, "k_sw_b29": 33784,
}


TRAFFIC_MIXED = {
    "k_lw_b01": 18810, "k_lw_b02": 21082, "k_lw_b03": 27694, "k_lw_b04": 24335, "k_lw_b05": 27629,
    "k_lw_b06": 15479, "k_lw_b07": 21765, "k_lw_b08": 15993, "k_lw_b09": 21276, "k_lw_b10": 13113,
    "k_lw_b11": 15546, "k_lw_b12": 15549, "k_lw_b13": 10506, "k_lw_b14": 7274, "k_lw_b15": 7757,
    "k_lw_b16": 7272, "k_sw_b16": 25803, "k_sw_b17": 48094, "k_sw_b18": 33012, "k_sw_b19": 33008,
    "k_sw_b20": 40600, "k_sw_b21": 40484, "k_sw_b22": 11371, "k_sw_b23": 40595, "k_sw_b24": 33465,
    "k_sw_b25": 26281, "k_sw_b26": 25289, "k_sw_b27": 33459, "k_sw_b28": 26265, "k_sw_b29": 48029,
}


def kernel_traffic(name: str, precision: str, cols_per_launch: float, days_per_launch: float):
    if precision == "surrogate":      # measured for the table paths only
        return None
    table = TRAFFIC_FP64 if precision == "fp64" else TRAFFIC_MIXED
    if name not in table:
        return None
    return table[name] * (cols_per_launch if name.startswith("k_lw") else days_per_launch)


def kernel_bytes(name: str, ncol: int, nday: int) -> float:
    if name.startswith("k_lw"):
        return IN_LW * ncol
    if name.startswith("k_sw"):
        return IN_SW * nday
    if name == "k_post":
        return (NL + 1) * 8 * ncol + OUT_ALL * ncol
    if name == "k_colprep":
        return (6 * NL + NL + 1 + 1) * 8 * ncol
    return 0.0


# Dense vector peaks of MI355X (TFLOP/s): FP64 78.6, FP32 157.3
VALU_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}

# FLOPs per column (LW) or per day column (SW) of the band kernels, from rocprofv3 SQ_INSTS_VALU_{ADD,MUL,
# FMA,TRANS}_F64 / _F32 counts (wave instructions x 64 lanes, FMA counted twice, so lanes masked off in
# divergent code count too: an upper bound) on the 400 x 400 grid (160,000 columns, 64,000 day columns).
# FP64 path: FP64 FLOPs.  Mixed path: FP32-equivalent FLOPs, FP32 + 2 x FP64 (FP64 issues at half the
# FP32 rate), against the FP32 peak.
FLOPS_FP64 = {
    "k_lw_b01": 145539, "k_lw_b02": 152887, "k_lw_b03": 255119, "k_lw_b04": 208076, "k_lw_b05": 246546,
    "k_lw_b06": 114218, "k_lw_b07": 188830, "k_lw_b08": 125292, "k_lw_b09": 195105, "k_lw_b10": 91067,
    "k_lw_b11": 116786, "k_lw_b12": 159111, "k_lw_b13": 105033, "k_lw_b14": 58000, "k_lw_b15": 76799,
    "k_lw_b16": 69924, "k_sw_b16": 267695, "k_sw_b17": 516824, "k_sw_b18": 346071, "k_sw_b19": 347459,
    "k_sw_b20": 417488, "k_sw_b21": 428415, "k_sw_b22": 106142, "k_sw_b23": 411511, "k_sw_b24": 350855,
    "k_sw_b25": 239041, "k_sw_b26": 226762, "k_sw_b27": 327658, "k_sw_b28": 261471, "k_sw_b29": 503154
}
FLOPS_MIXED = {
    "k_lw_b01": 89332, "k_lw_b02": 88151, "k_lw_b03": 167792, "k_lw_b04": 129832, "k_lw_b05": 163931,
    "k_lw_b06": 67145, "k_lw_b07": 121752, "k_lw_b08": 74139, "k_lw_b09": 124769, "k_lw_b10": 55205,
    "k_lw_b11": 72732, "k_lw_b12": 81436, "k_lw_b13": 58433, "k_lw_b14": 33661, "k_lw_b15": 43539,
    "k_lw_b16": 39158, "k_sw_b16": 136168, "k_sw_b17": 256402, "k_sw_b18": 166713, "k_sw_b19": 168384,
    "k_sw_b20": 198289, "k_sw_b21": 196579, "k_sw_b22": 59655, "k_sw_b23": 213413, "k_sw_b24": 158891,
    "k_sw_b25": 109955, "k_sw_b26": 249383, "k_sw_b27": 138230, "k_sw_b28": 121776, "k_sw_b29": 250155
}


def kernel_flops(name: str, precision: str, ncol: int, nday: int):
    if precision == "surrogate":
        return None
    table = FLOPS_FP64 if precision == "fp64" else FLOPS_MIXED
    if name not in table:
        return None
    return table[name] * (ncol if name.startswith("k_lw") else nday)


def serial_pass(d, precision, max_cols):
    """One call with the band kernels on the caller's stream (a context without band streams): the
    duration of each kernel alone on the GPU."""
    import torch

    from woof_nam_amd.radiation import Radiation
    prev = os.environ.get("WR_BAND_STREAMS")
    os.environ["WR_BAND_STREAMS"] = "0"
    try:
        rad = Radiation(0, max_cols=max_cols)
    finally:
        if prev is None:
            os.environ.pop("WR_BAND_STREAMS")
        else:
            os.environ["WR_BAND_STREAMS"] = prev
    out, _ = rad.run(d, precision=precision)                # warm-up
    torch.cuda.synchronize()
    rad.timing(True)
    t0 = time.perf_counter()
    rad.run(d, precision=precision, out=out)
    torch.cuda.synchronize()
    call_ms = (time.perf_counter() - t0) * 1e3
    ms = {}
    for name, t in rad.timing_read():
        ms[name] = ms.get(name, 0.0) + t
    rad.timing(False)
    rad.close()
    del out
    print(f"# serial pass: {call_ms:.1f} ms per call with the band kernels on one stream", file=sys.stderr, flush=True)
    return {"call_ms": call_ms, "ms": ms}


def served_block(d, dev, ncol, seconds, precision="mixed"):
    """Config-3 serving: Poisson arrivals at 40 req/s of the parent grid, 3 km / 6 km nests and
    ensemble members (one member per request, seeds 1..8), 200 GB HBM budget; mixed precision, or the
    surrogate path with the trained network (woof_nam_amd/data/surrogate_rrtmg.npz) in the served
    graphs."""
    import torch

    from woof_nam_amd import gen
    from woof_nam_amd.serve import Server, autotune, poisson_run
    net = None
    shapes = ((65536, 4), (131072, 2), (131072, 4), (262144, 2), (262144, 4), (524288, 2))
    if precision == "surrogate":
        from woof_nam_amd import surrogate as S
        net = S.MLP.load(os.path.join(ROOT, "woof_nam_amd", "data", "surrogate_rrtmg.npz"), dev)
        shapes = ((65536, 4), (131072, 2), (131072, 4), (262144, 2))   # 25 GB of network work space per 262,144-column slot

    def dev_grid(nx, ny, seed, id0):
        c = gen.grid(nx, ny, seed=seed)
        g = {k: torch.from_numpy(v).to(dev) for k, v in c.items()}
        return g, torch.arange(id0, id0 + nx * ny, dtype=torch.int64, device=dev)

    pid = torch.arange(ncol, dtype=torch.int64, device=dev)
    kinds = {"parent": (d, pid), "nest3km": dev_grid(333, 333, 11, 10_000_000),
             "nest6km": dev_grid(200, 200, 12, 20_000_000), "member": (d, pid)}
    t0 = time.perf_counter()
    # 524,288-column tiles only with 2 slots: 4 such slots beside this run's whole-grid context and the
    # requests' outputs could pass the 288 GB of HBM
    tune = autotune(0, ncol=786432, precision=precision, nreq=12, shapes=shapes, surrogate=net)
    best = tune["best"]
    print(f"# autotune {time.perf_counter() - t0:.1f} s: tile {best['tile']} slots {best['slots']}",
          file=sys.stderr, flush=True)
    srv = Server(0, tile=best["tile"], slots=best["slots"], budget_gb=200.0, precision=precision,
                 filler={k: v[..., :best["tile"]] for k, v in d.items()}, surrogate=net)
    r = poisson_run(srv, kinds, rate=40.0, duration=seconds, seed=0, progress=30.0 if seconds > 30 else 0.0)
    srv.close()
    r["autotune"] = tune["all"]
    r["tile"], r["slots"], r["precision"] = best["tile"], best["slots"], precision
    r["arrivals"] = f"Poisson 40 req/s for {seconds} s; kinds uniform over parent 796,590 / nest 110,889 / " \
                    "nest 40,000 / one ensemble member of the parent (own seed); offered load exceeds capacity, " \
                    "so p99 includes queueing"
    print(f"# served: {r['requests']} requests, {r['columns_per_s']:.4g} columns/s, p50 {r['p50_ms']:.0f} ms, "
          f"p99 {r['p99_ms']:.0f} ms", file=sys.stderr, flush=True)
    return r


MFMA_PEAK_TFLOPS_BF16 = 2500.0   # dense BF16 MFMA peak of MI355X


def surrogate_block(rad, d, out, tot, cnt, ncol, max_cols, net):
    """The surrogate path's network against the MFMA roofline (the three layers of one tile per
    launch: 2 x (8 x 128 + 128 x 128 + 128 x 504) FLOPs per radiation layer, inputs padded to 8), its
    share of the call, and the call's errors against the FP64 path on the same columns."""
    import torch
    lay = (NL + 1) * min(ncol, max_cols)
    fl = 2.0 * lay * (8 * 128 + 128 * 128 + 128 * 504)
    mlp_ms = tot.get("mlp_gemm", 0.0) / max(1, cnt.get("mlp_gemm", 1))
    feat_ms = tot.get("k_features", 0.0) / max(1, cnt.get("k_features", 1))
    ach = fl / (mlp_ms * 1e-3) / 1e12 if mlp_ms > 0 else None
    ref, _ = rad.run(d, precision="fp64")
    torch.cuda.synchronize()
    rms = {k: float(torch.sqrt(torch.mean((out[k] - ref[k]) ** 2))) for k in ref}
    del ref
    return {"mlp_ms_per_tile": mlp_ms, "features_ms_per_tile": feat_ms,
            "roofline": {"bound": "mfma", "achieved": ach, "peak": MFMA_PEAK_TFLOPS_BF16, "unit": "TFLOP/s",
                         "frac": ach / MFMA_PEAK_TFLOPS_BF16 if ach else None, "flops_per_launch": fl},
            "network_bands": {"lw": [b + 1 for b in range(16) if net.bands >> b & 1],
                              "sw": [16 + b for b in range(14) if net.bands >> (16 + b) & 1]},
            "accuracy_vs_fp64": {"hr_lw_rms_kday": rms["hr_lw"] * KDAY, "hr_sw_rms_kday": rms["hr_sw"] * KDAY,
                                 "flux_lw_rms_wm2": max(rms["flw_up"], rms["flw_dn"]),
                                 "flux_sw_rms_wm2": max(rms["fsw_up"], rms["fsw_dn"])}}


def aerosol_block(d, dev, nx, ny, steps):
    """wr_aer_optics (the OPAC aerosol climatology of January) on the grid's columns: ms per call from
    HIP events on the launching stream, and its HBM roofline.  Algorithmic bytes per column: p_lev (61),
    T and q (60 each) in FP64 and longitude, latitude and surface type in FP32 read; optical depth,
    single-scattering albedo and asymmetry of 30 bands x 61 layers written in FP32."""
    import torch

    from woof_nam_amd import aerosol
    ncol = int(d["t_lay"].shape[1])
    clim = aerosol.Climatology(dev)
    clim.set_month(1)
    f32 = dict(dtype=torch.float32, device=dev)
    lat = torch.deg2rad(torch.linspace(15.0, 70.0, ny, **f32))[:, None].expand(ny, nx).reshape(-1).contiguous()
    lon = torch.deg2rad(torch.linspace(-150.0, -50.0, nx, **f32))[None, :].expand(ny, nx).reshape(-1).contiguous()
    sl = (torch.arange(ncol, device=dev) % 3).to(torch.float32)     # sea, land and sea ice in turn
    cols = {k: d[k] for k in ("p_lev", "t_lay", "q_lay")}
    out = aerosol.optics(clim, cols, lat, lon, sl)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream(dev)
    ms = []
    for _ in range(max(steps, 3)):
        b, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        del out
        b.record(stream)
        out = aerosol.optics(clim, cols, lat, lon, sl)
        e.record(stream)
        torch.cuda.synchronize()
        ms.append(b.elapsed_time(e))
    aod = float(out["aer_sw"][0, 9].sum(0).mean())
    del out
    torch.cuda.empty_cache()
    t = float(np.median(ms))
    per_col = (NL + 1 + 2 * NL) * 8 + 3 * 4 + 3 * 30 * (NL + 1) * 4
    ach = per_col * ncol / (t * 1e-3) / 1e9
    return {"ms_per_call": t, "columns_per_s": ncol / (t * 1e-3), "mean_aod_band_0p44_0p63um": aod,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
                         "algorithmic_bytes_per_column": per_col},
            "what": "climatological aerosol optics (setaer + aer_property + radclimaer) for every column, "
                    "January OPAC tables, median of the calls"}


def cpu_info() -> dict:
    """Host cores and CPU model (nproc and /proc/cpuinfo), and the threads this run may use: the
    OMP_NUM_THREADS share when the host sets one, else every core."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = os.cpu_count() or 1
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return {"nproc": nproc, "cpu_model": model, "threads": share if share > 0 else nproc}


def cpu_baseline(cols, out, ncol):
    """The CPU FP64 oracle (the C restatement, OpenMP over columns) on the 4,096-column CPU-reference
    workload (gen.columns(4096, seed=0), McICA): 1 warm-up, then the median of 5 timed runs on this
    host's cores.  The full-grid figure is extrapolated from it.  Beside it, the GPU FP64 outputs
    against the oracle on the first columns of the grid."""
    from oracle import pyoracle
    from woof_nam_amd import gen
    info = cpu_info()
    threads = info["threads"]
    c0 = gen.columns(4096, seed=0)
    pyoracle.run(c0, nthreads=threads)                      # warm-up
    times = []
    for _ in range(5):
        t1 = time.perf_counter()
        pyoracle.run(c0, nthreads=threads)
        times.append(time.perf_counter() - t1)
    med = float(np.median(times))
    base = {"value": 4096 / med, "unit": "columns/s", "cores": threads, "kind": "port",
            "sample": "4,096 independent columns x 60 layers (the CPU-reference workload, seed 0), LW+SW McICA, "
                      f"CPU FP64 oracle (C, OpenMP); median of 5 runs after 1 warm-up: {med * 1e3:.1f} ms",
            "runs_ms": [t * 1e3 for t in times], "nproc": info["nproc"], "cpu_model": info["cpu_model"],
            "grid_extrapolated_s": ncol / (4096 / med),
            "grid_note": f"the {ncol}-column grid is extrapolated from the 4,096-column rate, not run"}
    ns = min(ncol, 8192)
    sub = {k: np.ascontiguousarray(v[..., :ns]) for k, v in cols.items()}
    ref, _ = pyoracle.run(sub, nthreads=threads)
    acc = {}
    for k in ("hr_lw", "hr_sw"):
        diff = out[k][:, :ns].cpu().numpy() - ref[k]
        acc[k + "_rms_kday"] = float(np.sqrt(np.mean(diff ** 2)) * KDAY)
    for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
        diff = out[k][:, :ns].cpu().numpy() - ref[k]
        acc[k + "_rms_wm2"] = float(np.sqrt(np.mean(diff ** 2)))
    acc["vs"] = f"CPU FP64 oracle on the first {ns} columns of the grid"
    return base, acc


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=("fp64", "mixed", "surrogate"), default="fp64",
                    help="surrogate: the mixed path with the trained network's gas optics (data/surrogate_rrtmg.npz)")
    ap.add_argument("--max-cols", type=int, default=0, help="tile columns (0 = the whole grid in one tile)")
    ap.add_argument("--nx", type=int, default=954)
    ap.add_argument("--ny", type=int, default=835)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--acc-cols", type=int, default=65536, help="mixed-precision error-budget sample")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision comparison block")
    ap.add_argument("--served-seconds", type=float, default=3.0,
                    help="seconds of Poisson arrivals (40 req/s) for the served block (0 = skip)")
    ap.add_argument("--no-serial", action="store_true", help="skip the pass with the band kernels on one stream")
    ap.add_argument("--no-aerosol", action="store_true", help="skip the climatological-aerosol block")
    ap.add_argument("--served-precision", choices=("mixed", "surrogate"), default="mixed",
                    help="the served block's path (surrogate: the trained network inside the served graphs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB between two device buffers (reference for HBM byte counters)")
    a = ap.parse_args(argv)
    if a.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {a.gpus}); "
              "the radiation columns are independent and would shard by contiguous column ranges, "
              "but no multi-GPU run is part of it", file=sys.stderr)
        return 2

    import torch

    from woof_nam_amd import gen
    from woof_nam_amd.radiation import Radiation

    if a.copy_calibration:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    t0 = time.time()
    cols = gen.grid(a.nx, a.ny, seed=0)
    ncol = cols["t_lay"].shape[1]
    nday = int(np.count_nonzero(cols["cosz"] >= 1.0e-4))
    print(f"# grid {a.nx}x{a.ny}x{NL}: {ncol} columns, {nday} day, generated in {time.time() - t0:.1f} s",
          file=sys.stderr, flush=True)
    dev = torch.device("cuda", 0)
    d = {k: torch.from_numpy(v).to(dev) for k, v in cols.items()}
    a.max_cols = a.max_cols or ncol
    rad = Radiation(0, max_cols=a.max_cols)
    if a.precision == "surrogate":
        from woof_nam_amd import surrogate as S
        net = S.MLP.load(os.path.join(ROOT, "woof_nam_amd", "data", "surrogate_rrtmg.npz"), dev)
        rad.attach_surrogate(net)
        a.no_serial = True          # the serial pass's context carries no network
    out, _ = rad.run(d, precision=a.precision)
    torch.cuda.synchronize()
    for _ in range(a.warmup):
        rad.run(d, precision=a.precision, out=out)
    torch.cuda.synchronize()
    print("# warmup done", file=sys.stderr, flush=True)

    stream = torch.cuda.current_stream(dev)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    rad.timing(True)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for s in range(a.steps):
        evs[s][0].record(stream)
        rad.run(d, precision=a.precision, out=out)
        evs[s][1].record(stream)
    torch.cuda.synchronize()
    t_el = time.perf_counter() - t_start
    launches = rad.timing_read()
    rad.timing(False)
    lat = np.array([b.elapsed_time(e) for b, e in ((x, y) for x, y in evs)])
    ms_per_step = 1e3 * t_el / a.steps
    value = ncol * a.steps / t_el
    print(f"# {a.steps} steps: {ms_per_step:.1f} ms/step, {value:.4g} columns/s", file=sys.stderr, flush=True)

    # dominant kernel and its roofline
    tot, cnt = {}, {}
    for name, ms in launches:
        tot[name] = tot.get(name, 0.0) + ms
        cnt[name] = cnt.get(name, 0) + 1
    dom = max(tot, key=tot.get)
    k_ms = tot[dom] / cnt[dom]
    bytes_per_launch = kernel_bytes(dom, ncol, nday) * a.steps / cnt[dom]
    achieved = bytes_per_launch / (k_ms * 1e-3) / 1e9
    traffic = kernel_traffic(dom, a.precision, ncol * a.steps / cnt[dom], nday * a.steps / cnt[dom])
    shares = sorted(((v / sum(tot.values()), k) for k, v in tot.items()), reverse=True)[:6]
    print("# kernel time shares: " + ", ".join(f"{k} {100 * f:.1f}%" for f, k in shares), file=sys.stderr, flush=True)
    serial = serial_pass(d, a.precision, a.max_cols) if not a.no_serial else None

    res = {
        "metric": METRIC, "value": value, "unit": "columns/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "mixed": "f32+f64", "surrogate": "bf16+f32+f64"}[a.precision],
        "data": "synthetic (seeded generator: NAM sigma levels, ISA T + 300 km GRF, RH GRF, clouds, 12 UTC sun)",
        "config": {"workload": f"NAM 12 km grid {a.nx}x{a.ny} = {ncol} columns x {NL} layers, LW+SW RRTMG, "
                               f"McICA 140/112 g, {a.precision}",
                   "columns": ncol, "day_columns": nday, "layers": NL, "precision": a.precision,
                   "tile_columns": a.max_cols},
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "per_call": "HIP events around each call on its stream"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "launch_ms": k_ms, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "band_streams": "the 30 band kernels of a tile run on 4 streams at once: launch_ms includes "
                                     "the time a kernel shares the CUs with the others",
                     "call_achieved_gbs": (6312.0 if a.precision == "fp64" else 3600.0) * ncol / (ms_per_step * 1e-3) / 1e9},
    }
    table = {"fp64": FLOPS_FP64, "mixed": FLOPS_MIXED}.get(a.precision, {})
    if len(table) == 30:
        # the whole call's VALU work against the vector peak (all 30 band kernels, one call)
        call_fl = sum(v * (ncol if k.startswith("k_lw") else nday) for k, v in table.items())
        peak = VALU_PEAK_TFLOPS["fp64" if a.precision == "fp64" else "fp32"]
        ach = call_fl / (ms_per_step * 1e-3) / 1e12
        res["roofline"]["call_valu"] = {"achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                                        "flops_per_call": call_fl}
    if serial is not None:
        rf = res["roofline"]
        rf["serial_launch_ms"] = serial["ms"].get(dom)
        rf["serial_frac"] = (bytes_per_launch / (rf["serial_launch_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS
                             if rf["serial_launch_ms"] else None)
        rf["serial_note"] = "the same kernel with the band kernels on one stream (each alone on the GPU)"
        fl = kernel_flops(dom, a.precision, ncol, nday)
        if fl is not None and rf["serial_launch_ms"]:
            peak = VALU_PEAK_TFLOPS["fp64" if a.precision == "fp64" else "fp32"]
            ach = fl / (rf["serial_launch_ms"] * 1e-3) / 1e12
            rf["valu"] = {"achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                          "flops_per_launch": fl, "from": "rocprofv3 SQ_INSTS_VALU_* counts per column x 64 lanes (an upper bound: masked lanes count)"}
        res["serial"] = {"ms_per_call": serial["call_ms"], "kernel_ms": serial["ms"]}

    if not a.no_aerosol:
        res["aerosol"] = aerosol_block(d, dev, a.nx, a.ny, a.steps)
        print(f"# aerosol optics: {res['aerosol']['ms_per_call']:.2f} ms/call", file=sys.stderr, flush=True)

    if a.precision == "surrogate":
        res["surrogate"] = surrogate_block(rad, d, out, tot, cnt, ncol, a.max_cols, net)

    if a.precision == "fp64" and not a.no_mixed:
        # the same workload on the mixed-precision path: throughput, speedup, errors against FP64
        mout, _ = rad.run(d, precision="mixed")
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(a.steps):
            rad.run(d, precision="mixed", out=mout)
        torch.cuda.synchronize()
        tm = (time.perf_counter() - t1) / a.steps
        mixed = {"value": ncol / tm, "ms_per_step": tm * 1e3, "speedup_vs_fp64": ms_per_step / (tm * 1e3)}
        for k in ("hr_lw", "hr_sw"):
            mixed[k + "_rms_kday"] = float(torch.sqrt(torch.mean((mout[k] - out[k]) ** 2))) * KDAY
        for k in ("flw_up", "flw_dn", "fsw_up", "fsw_dn"):
            mixed[k + "_rms_wm2"] = float(torch.sqrt(torch.mean((mout[k] - out[k]) ** 2)))
        mixed["vs"] = "FP64 path on all columns (the FP64 path equals the CPU oracle bit for bit)"
        res["mixed"] = mixed
        del mout
        print(f"# mixed: {tm * 1e3:.1f} ms/step, {ncol / tm:.4g} columns/s", file=sys.stderr, flush=True)

    if a.precision == "mixed" and a.acc_cols > 0:
        # error budget against the FP64 path (bit-identical to the CPU oracle) on a column sample,
        # per layer and per band
        from woof_nam_amd import accuracy
        ns = min(ncol, a.acc_cols)
        idx = torch.linspace(0, ncol - 1, ns, device=dev).long()
        sub = {k: v[..., idx].contiguous() for k, v in d.items()}
        rep = accuracy.report(rad, sub, per_band=True)
        with open(os.path.join(ROOT, "gpurun_out", "accuracy_mixed.txt") if os.path.isdir(
                os.path.join(ROOT, "gpurun_out")) else os.devnull, "w") as f:
            f.write(accuracy.format_report(rep) + "\n")
        worst_lw = max(rep["lw_by_band"].items(), key=lambda kv: kv[1]["flux_rms_wm2"])
        worst_sw = max(rep["sw_by_band"].items(), key=lambda kv: kv[1]["flux_rms_wm2"])
        lay_lw = int(np.argmax(rep["hr_lw_rms_by_layer_kday"]))
        lay_sw = int(np.argmax(rep["hr_sw_rms_by_layer_kday"]))
        res["accuracy_vs_fp64"] = {
            "sample_columns": ns,
            **{k: rep[k] for k in rep if k.endswith("_rms_kday") or k.endswith("_rms_wm2") or k.endswith("_max_kday")},
            "worst_lw_band": {"band": worst_lw[0], **worst_lw[1]},
            "worst_sw_band": {"band": worst_sw[0], **worst_sw[1]},
            "worst_layer_hr_lw": {"layer": lay_lw, "rms_kday": rep["hr_lw_rms_by_layer_kday"][lay_lw]},
            "worst_layer_hr_sw": {"layer": lay_sw, "rms_kday": rep["hr_sw_rms_by_layer_kday"][lay_sw]},
        }

    if not a.no_cpu:
        res["cpu_baseline"], res["accuracy"] = cpu_baseline(cols, out, ncol)
    else:
        res["cpu_baseline"] = None
    rad.close()                     # the served run's slots and requests get the device memory
    if a.served_seconds > 0:
        torch.cuda.empty_cache()
        res["served"] = served_block(d, dev, ncol, a.served_seconds, a.served_precision)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
