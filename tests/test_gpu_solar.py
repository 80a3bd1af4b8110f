"""wr_coszmn and wr_rdtemp on the GPU against the oracle.  Both evaluate in FP64; the device's
trigonometric functions and the host C library's may differ in the last unit, so cos z agrees to
1e-13 and the float32 temperatures and zenith cosines to one unit in the last place."""
import datetime as dt

import numpy as np
import pytest
import torch

from oracle import pyoracle
from woof_nam_amd import solar

pytestmark = pytest.mark.gpu


def test_coszmn_matches_oracle(gpu_device):
    print("case coszmn", flush=True)
    rng = np.random.default_rng(5)
    lat = np.radians(rng.uniform(-89, 89, 5000))
    lon = np.radians(rng.uniform(-180, 180, 5000))
    jdat = [2025, 6, 21, 0, 11, 30, 0, 0]
    for nrads, dtm in ((48, 25.0), (12, 300.0), (1, 60.0)):
        cz, cd = solar.coszen(jdat, 0.0, nrads, dtm, torch.from_numpy(lat).to(gpu_device),
                              torch.from_numpy(lon).to(gpu_device))
        torch.cuda.synchronize()
        s = pyoracle.sol_update(2025, 6, 21, 11, 30, 0, nrads * dtm, dtm, solar.SOLC0)
        rz, rd = pyoracle.coszmn(lat, lon, 11.5, nrads * dtm / 3600.0, nrads, s)
        assert np.abs(cz.cpu().numpy() - rz).max() <= 1e-13
        assert np.abs(cd.cpu().numpy() - rd).max() <= 1e-13
        assert (rz > 0).mean() > 0.3 and (rz == 0).mean() > 0.3


def test_rdtemp_matches_oracle(gpu_device):
    print("case rdtemp", flush=True)
    im, jm, lm = 95, 83, 60
    rng = np.random.default_rng(6)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)
    glat = f32(np.radians(np.linspace(15, 70, jm))[:, None] * np.ones((1, im)))
    glon = f32(np.radians(np.linspace(-150, -50, im))[None, :] * np.ones((jm, 1)))
    czmean = f32(rng.uniform(0, 0.9, (jm, im)))
    czmean[rng.uniform(size=(jm, im)) < 0.2] = 0.0
    rsw = f32(rng.uniform(0, 3e-4, (lm, jm, im)))
    rlw = f32(rng.uniform(-3e-4, 1e-5, (lm, jm, im)))
    t0 = f32(rng.uniform(200, 305, (lm, jm, im)))
    start = dt.datetime(2025, 3, 20, 6)
    d = {k: torch.from_numpy(v.copy()).to(gpu_device) for k, v in
         (("glat", glat), ("glon", glon), ("czmean", czmean), ("rsw", rsw), ("rlw", rlw), ("t", t0))}
    czen = torch.zeros((jm, im), dtype=torch.float32, device=gpu_device)
    t_h, cz_h = t0.copy(), np.zeros((jm, im), np.float32)
    for step in (1, 200, 1000):
        solar.rdtemp(step, 25.0, start, d["glat"], d["glon"], d["czmean"], d["rsw"], d["rlw"], d["t"], czen)
        pyoracle.rdtemp(t_h, cz_h, glat, glon, czmean, rsw, rlw, ntsd=step, dt=25.0, start=start,
                        box=(1, im - 2, 1, jm - 2))
    torch.cuda.synchronize()
    tg, cg = d["t"].cpu().numpy(), czen.cpu().numpy()
    ulp = lambda a: np.spacing(np.abs(a).astype(np.float32))
    assert (np.abs(tg - t_h) <= 3 * ulp(t_h)).all()
    assert (np.abs(cg - cz_h) <= ulp(np.maximum(cz_h, 1e-30))).all() or np.abs(cg - cz_h).max() < 1e-6
    assert np.array_equal(tg[:, 0, :], t0[:, 0, :])          # boundary rows keep their temperature
    assert (cz_h > 0).any() and (cz_h == 0).any()
