"""Train the gas-optics surrogate on layers of the plain FP64 path and judge it against FP64.

    python -m woof_nam_amd.surrogate_train --layers 20000000 --out surrogate.npz --report report.json

Training data: radiation layers of columns drawn from the 954 x 835 NAM-grid generator (seed 0),
their features and targets produced by the library's FP64 layer-optics kernel.  Band selection: on
a sample of a separate grid (seed 1), each band's heating-rate error with the network in that band
alone, then bands added in order of their own error while the set's measured LW (SW) heating-rate
RMS stays within the bar; the other bands keep the tables.  Validation on other columns of that
grid: heating-rate and flux RMS of the surrogate path (selected bands, and every band) and of the
plain mixed path against the FP64 path (which the CPU oracle reproduces bit for bit), and columns/s
of each path.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import gen
from . import surrogate as S
from .radiation import Radiation

KDAY = 86400.0


def _dev(cols, device, idx=None):
    out = {}
    for k, v in cols.items():
        a = v if idx is None else v[..., idx]
        out[k] = torch.from_numpy(np.ascontiguousarray(a)).to(device)
    return out


def _rms(a, b):
    return float(torch.sqrt(torch.mean((a - b) ** 2)))


def band_errors(rad, mlp, cols, ref):
    """Heating-rate RMS (K/day) against FP64 with the network in one band at a time: 16 LW, 14 SW."""
    lw, sw = [], []
    for b in range(30):
        mlp.bands = 1 << b
        rad.attach_surrogate(mlp)
        out, _ = rad.run(cols, precision="surrogate")
        key = "hr_lw" if b < 16 else "hr_sw"
        (lw if b < 16 else sw).append(_rms(out[key], ref[key]) * KDAY)
    return lw, sw


def select_bands(rad, mlp, cols, target=0.045):
    """The bands that keep the network: per spectrum, bands in order of their own error, each kept if
    the set's measured heating-rate RMS stays <= target (K/day).  Sets mlp.bands; returns the report."""
    ref, _ = rad.run(cols, precision="fp64")
    lw, sw = band_errors(rad, mlp, cols, ref)
    chosen = 0
    steps = []
    for errs, off, key in ((lw, 0, "hr_lw"), (sw, 16, "hr_sw")):
        for i in np.argsort(errs):
            bit = 1 << (off + int(i))
            mlp.bands = chosen | bit
            rad.attach_surrogate(mlp)
            out, _ = rad.run(cols, precision="surrogate")
            e = _rms(out[key], ref[key]) * KDAY
            keep = e <= target
            steps.append({"band": int(i) + 1 if off == 0 else 16 + int(i), "set_rms_kday": e, "kept": bool(keep)})
            if keep:
                chosen |= bit
    if chosen == 0:
        raise RuntimeError("surrogate: no band meets the heating-rate bar with the network")
    mlp.bands = chosen
    rad.attach_surrogate(mlp)
    return {"target_kday": target, "mask": chosen, "lw_alone_kday": lw, "sw_alone_kday": sw, "greedy": steps,
            "lw_bands": [b + 1 for b in range(16) if chosen >> b & 1],
            "sw_bands": [16 + b for b in range(14) if chosen >> (16 + b) & 1]}


def evaluate(rad, cols, reps=3, mlp=None):
    """Errors against FP64 and throughput of fp64 / mixed / surrogate (and, given the network, the
    surrogate in every band) on device columns."""
    res = {}
    ref, _ = rad.run(cols, precision="fp64")
    torch.cuda.synchronize()
    ncol = int(cols["t_lay"].shape[1])
    runs = [("fp64", "fp64", None), ("mixed", "mixed", None), ("surrogate", "surrogate", None)]
    if mlp is not None:
        runs.append(("surrogate_all_bands", "surrogate", S.ALL_BANDS))
    for name, prec, bands in runs:
        if bands is not None:
            keep = mlp.bands
            mlp.bands = bands
            rad.attach_surrogate(mlp)
        out, info = rad.run(cols, precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            rad.run(cols, precision=prec, out=out)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / reps
        r = {"columns_per_s": ncol / dt, "ms_per_call": dt * 1e3,
             "hr_lw_rms_kday": _rms(out["hr_lw"], ref["hr_lw"]) * KDAY,
             "hr_sw_rms_kday": _rms(out["hr_sw"], ref["hr_sw"]) * KDAY,
             "flux_lw_rms_wm2": max(_rms(out["flw_up"], ref["flw_up"]), _rms(out["flw_dn"], ref["flw_dn"])),
             "flux_sw_rms_wm2": max(_rms(out["fsw_up"], ref["fsw_up"]), _rms(out["fsw_dn"], ref["fsw_dn"]))}
        if prec == "surrogate":
            r["fallback_columns"] = int(info["fallback"].sum())
        res[name] = r
        if bands is not None:
            mlp.bands = keep
            rad.attach_surrogate(mlp)
    return res


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--layers", type=int, default=20_000_000)
    ap.add_argument("--steps", type=int, default=12000)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--val-cols", type=int, default=200_000)
    ap.add_argument("--select-cols", type=int, default=65_536)
    ap.add_argument("--target", type=float, default=0.045, help="heating-rate RMS bar of the band selection, K/day")
    ap.add_argument("--out", default="surrogate.npz")
    ap.add_argument("--report", default=None)
    a = ap.parse_args(argv)
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    grid = gen.grid(seed=0)
    ncol_all = grid["t_lay"].shape[1]
    nl = grid["t_lay"].shape[0] + 1
    ncol = min(ncol_all, (a.layers + nl - 1) // nl)
    idx = np.sort(np.random.default_rng(0).choice(ncol_all, size=ncol, replace=False))
    rad = Radiation(0, max_cols=131072)
    feats, targs = [], []
    for c0 in range(0, ncol, 100_000):
        sub = _dev(grid, dev, idx[c0:c0 + 100_000])
        f, t = S.layer_optics(rad, sub)
        feats.append(f)
        targs.append(t)
        del sub
    feat = torch.cat(feats)
    targ = torch.cat(targs)
    del feats, targs
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    print(f"# data: {feat.shape[0]} layers from {ncol} columns in {t_data:.1f} s", file=sys.stderr, flush=True)
    mlp = S.MLP(dev, seed=0)
    t1 = time.perf_counter()
    hist = S.fit(mlp, feat, targ, steps=a.steps, batch=a.batch, lr=a.lr, log_every=max(1, a.steps // 20),
                 log=lambda m: print("# " + m, file=sys.stderr, flush=True))
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t1
    del feat, targ
    torch.cuda.empty_cache()
    val = gen.grid(seed=1)
    nv = val["t_lay"].shape[1]
    perm = np.random.default_rng(1).permutation(nv)
    sidx = np.sort(perm[:a.select_cols])
    vidx = np.sort(perm[a.select_cols:a.select_cols + min(a.val_cols, nv - a.select_cols)])
    rad.close()
    rad = Radiation(0, max_cols=262144)
    scols = _dev(val, dev, sidx)
    sel = select_bands(rad, mlp, scols, target=a.target)
    print(f"# bands on the network: LW {sel['lw_bands']} SW {sel['sw_bands']}", file=sys.stderr, flush=True)
    del scols
    mlp.save(a.out)
    vcols = _dev(val, dev, vidx)
    ev = evaluate(rad, vcols, mlp=mlp)
    rep = {"training_layers": int(ncol * nl), "training_columns": int(ncol), "steps": a.steps, "batch": a.batch,
           "data_s": t_data, "train_s": t_train, "loss_first": hist[0][1], "loss_last": hist[-1][1],
           "band_selection": sel, "selection_columns": int(len(sidx)),
           "validation_columns": int(len(vidx)), "validation": ev}
    print(json.dumps(rep))
    if a.report:
        with open(a.report, "w") as f:
            json.dump(rep, f, indent=1)
    rad.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
