"""Entry points and workloads checked against the CPU oracle on the GPU.

* wr_cal_cldfra against oracle_cal_cldfra (phys_rad_cal_cldfra.cu:29-55), every flag combination,
  values straddling the 1e-6 threshold, boundary points untouched (tools/check-phys-rad-cal-cldfra.py);
* wr_mcica_masks against the oracle's McICA generator, bit for bit, for the three overlap rules and
  column ids above 2^32;
* the plain FP64 call on columns of the 954 x 835 grid workload (GRF fields, 12 UTC sun): the low-sun
  edge of the domain and a random sample, bit-identical to the oracle with indices and bits;
* the mixed-precision path's own table indices (its FP32 instantiation of the layer setup), bit-exact.
"""
import numpy as np
import pytest
import torch

from oracle import pyoracle
from woof_nam_amd import abi, gen
from woof_nam_amd.radiation import Radiation, cal_cldfra, mcica_masks

pytestmark = pytest.mark.gpu


def to_dev(cols, dev):
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in cols.items()}


@pytest.mark.parametrize("flags", [(1, 1), (1, 0), (0, 0), (0, 1)])
def test_cal_cldfra_matches_oracle(gpu_device, flags):
    f_qc, f_qi = flags
    print(f"case cal_cldfra f_qc={f_qc} f_qi={f_qi}", flush=True)
    rng = np.random.default_rng(17 + 2 * f_qc + f_qi)
    lm, jm, im = 9, 13, 21
    qc = rng.uniform(0.0, 2.0e-6, (lm, jm, im)).astype(np.float32)
    qi = rng.uniform(0.0, 2.0e-6, (lm, jm, im)).astype(np.float32)
    # exact threshold cases: qc alone at 1e-6 (not cloudy), qc + qi summing to just above
    qc[0, 3, 4], qi[0, 3, 4] = np.float32(1.0e-6), np.float32(0.0)
    qc[1, 5, 6], qi[1, 5, 6] = np.float32(5.0e-7), np.float32(5.0000006e-7)
    qc[2, 7, 8], qi[2, 7, 8] = np.float32(0.0), np.float32(3.0e-6)
    ref = np.full((lm, jm, im), -1.0, dtype=np.float32)
    pyoracle.cal_cldfra(ref, qc, qi, 1, im - 2, 1, jm - 2, bool(f_qc), bool(f_qi))
    dev = torch.full((lm, jm, im), -1.0, dtype=torch.float32, device=gpu_device)
    cal_cldfra(dev, torch.from_numpy(qc).to(gpu_device), torch.from_numpy(qi).to(gpu_device),
               1, im - 2, 1, jm - 2, bool(f_qc), bool(f_qi))
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert (got[:, 0, :] == -1.0).all() and (got[:, :, -1] == -1.0).all()     # boundary untouched
    inner = got[:, 1:-1, 1:-1]
    assert ((inner == 0.0) | (inner == 1.0)).all()
    if f_qc and f_qi:
        assert inner.mean() > 0.2


@pytest.mark.parametrize("iovr", [0, 1, 2])
def test_mcica_masks_match_oracle(gpu_device, iovr):
    print(f"case mcica_masks iovr={iovr}", flush=True)
    rng = np.random.default_rng(40 + iovr)
    nrl, ncol = 61, 192
    cf = rng.uniform(0.0, 1.0, (nrl, ncol))
    cf[rng.uniform(size=(nrl, ncol)) < 0.4] = 0.0
    cf[:, 0] = 0.0                                       # clear column
    cf[:, 1] = 1.0                                       # overcast column
    ids = (np.arange(ncol, dtype=np.uint64) * np.uint64(2654435761) + np.uint64(1 << 33)).astype(np.uint64)
    seed = 0x1234_5678_9ABC
    lw, sw = mcica_masks(torch.from_numpy(cf).to(gpu_device), torch.from_numpy(ids.view(np.int64)).to(gpu_device),
                         seed, iovr)
    torch.cuda.synchronize()
    lw = lw.cpu().numpy().view(np.uint32)
    sw = sw.cpu().numpy().view(np.uint32)
    for i in range(ncol):
        rl = pyoracle.mcica(int(ids[i]), seed, 0, 140, iovr, cf[:, i])
        rs = pyoracle.mcica(int(ids[i]), seed, 1, 112, iovr, cf[:, i])
        assert np.array_equal(lw[:, :, i], rl), f"LW bits of column {i} differ"
        assert np.array_equal(sw[:, :, i], rs), f"SW bits of column {i} differ"
    assert not lw[:, :, 0].any() and not sw[:, :, 0].any()
    assert (lw[:, :4, 1] == 0xFFFFFFFF).all() and (lw[:, 4, 1] == 0xFFF).all()


@pytest.fixture(scope="module")
def grid_sample():
    """Columns of the 954 x 835 grid workload: every 8th column within the low-sun edge (cos z in
    [-0.01, 0.03], through the 1e-4 day threshold) and 2,000 random ones."""
    c = gen.grid(954, 835, seed=0)
    cz = c["cosz"]
    edge = np.flatnonzero((cz > -0.01) & (cz < 0.03))[::8]
    rest = np.random.default_rng(3).choice(cz.size, 2000, replace=False)
    idx = np.unique(np.concatenate([edge, rest]))
    sub = {k: np.ascontiguousarray(v[..., idx]) for k, v in c.items()}
    sub["col_id"] = idx.astype(np.uint64)
    return sub


def test_fp64_grid_columns_bit_identical(gpu_device, grid_sample):
    cols = grid_sample
    n = cols["t_lay"].shape[1]
    nday = int(np.count_nonzero(cols["cosz"] >= 1.0e-4))
    print(f"case grid sample fp64: {n} columns, {nday} day", flush=True)
    assert 0.2 * n < nday < 0.8 * n
    ref, rdg = pyoracle.run(cols, diag=True)
    rad = Radiation(0, max_cols=1024)
    d = to_dev({k: v for k, v in cols.items() if k != "col_id"}, gpu_device)
    d["col_id"] = torch.from_numpy(cols["col_id"].view(np.int64)).to(gpu_device)
    out, dg = rad.run(d, diag=True)
    torch.cuda.synchronize()
    for k in abi.OUTPUT_FIELDS:
        a = out[k].cpu().numpy()
        assert np.array_equal(a.view(np.uint64), ref[k].view(np.uint64)), \
            f"{k}: {np.count_nonzero(a != ref[k])} elements differ"
    for k in abi.DIAG_FIELDS:
        a = dg[k].cpu().numpy().view(np.uint32)
        assert np.array_equal(a, rdg[k].view(np.uint32)), f"diag {k} differs"
    # the low-sun columns: SW zero at night, positive TOA downward flux by day
    day = cols["cosz"] >= 1.0e-4
    fsw = out["fsw_dn"].cpu().numpy()
    assert (fsw[:, ~day] == 0.0).all() and (fsw[0, day] > 0.0).all()
    rad.close()


def test_mixed_path_indices_bit_exact(gpu_device, grid_sample):
    """The mixed call's diagnostics come from the FP32 instantiation of the layer setup (the code
    the FP32 band kernels run); its indices and McICA bits must equal the oracle's."""
    cols = grid_sample
    print(f"case grid sample mixed indices: {cols['t_lay'].shape[1]} columns", flush=True)
    _, rdg = pyoracle.run(cols, diag=True)
    rad = Radiation(0, max_cols=4096)
    d = to_dev({k: v for k, v in cols.items() if k != "col_id"}, gpu_device)
    d["col_id"] = torch.from_numpy(cols["col_id"].view(np.int64)).to(gpu_device)
    _, dg = rad.run(d, precision="mixed", diag=True)
    torch.cuda.synchronize()
    for k in abi.DIAG_FIELDS:
        a = dg[k].cpu().numpy().view(np.uint32)
        assert np.array_equal(a, rdg[k].view(np.uint32)), f"mixed-path diag {k} differs from the oracle"
    rad.close()
