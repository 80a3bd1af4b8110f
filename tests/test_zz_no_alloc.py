"""wr_run_mixed_noalloc (the mixed-precision call on the library's static work space) against
wr_run_mixed: bit for bit, on two calls in a row, the second with fewer columns, on both routes."""
import numpy as np
import pytest
import torch

from woof_nam_amd import abi, gen
from woof_nam_amd.radiation import Radiation

pytestmark = pytest.mark.gpu


def dev_cols(n, seed, device):
    c = gen.columns(n, seed=seed)
    return {k: torch.from_numpy(np.ascontiguousarray(v)).to(device) for k, v in c.items()}


@pytest.mark.parametrize("mcica", [1, 0])
def test_noalloc_matches_mixed(gpu_device, mcica):
    rad = Radiation(0, max_cols=abi.NOALLOC_MAX_COLS)
    params = Radiation.default_params(mcica=mcica, seed=7)
    for n, seed in ((abi.NOALLOC_MAX_COLS, 31), (1500, 32)):       # the second call is smaller
        print(f"case noalloc mcica={mcica} ncol={n}", flush=True)
        cols = dev_cols(n, seed, gpu_device)
        ref, _ = rad.run(cols, params=params, precision="mixed")
        got = {k: torch.full_like(v, -7.0) for k, v in ref.items()}
        got = rad.run_noalloc(cols, params=params, out=got)
        torch.cuda.synchronize()
        for k in abi.OUTPUT_FIELDS:
            a, b = got[k].cpu().numpy(), ref[k].cpu().numpy()
            assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), \
                f"{k}: {np.count_nonzero(a != b)} elements differ (ncol {n})"
    rad.close()


def test_noalloc_rejects_more_columns_than_its_work_space(gpu_device):
    rad = Radiation(0, max_cols=128)
    cols = dev_cols(abi.NOALLOC_MAX_COLS + 1, 33, gpu_device)
    with pytest.raises(ValueError, match="static work space"):
        rad.run_noalloc(cols)
    rad.close()
