"""wr_aer_optics on the GPU against the oracle (oracle_aer.c), bit for bit: the OPAC climatology of two
months on synthetic columns over the whole globe (poles, the date line, negative longitudes, sea, land
and sea ice), ragged column counts (1, 63, 64, 65 and a few thousand), a padded leading dimension, the
synthetic tables of the CPU pins, and the hook's aerosol inputs."""
import numpy as np
import pytest
import torch

from oracle import pyoracle
from test_aerosol_oracle import synth
from woof_nam_amd import aerosol, gen

pytestmark = pytest.mark.gpu
F = np.float32


def _cases(n, seed):
    c = gen.columns(n, seed=seed)
    rng = np.random.default_rng(seed)
    lon = rng.uniform(-np.pi, 2.0 * np.pi, n)
    lat = rng.uniform(-np.pi / 2, np.pi / 2, n)
    edge = [(0.0, np.pi / 2), (0.0, -np.pi / 2), (2.0 * np.pi, 0.0), (-1e-7, 0.3), (np.pi, -np.pi / 2 + 1e-7)]
    for i, (lo, la) in enumerate(edge[:n]):
        lon[i], lat[i] = lo, la
    sl = rng.choice([0.0, 1.0, 2.0], n)
    return c, lon.astype(F), lat.astype(F), sl.astype(F)


def _device(clim, c, lon, lat, sl, dev, ld=None):
    cols = {k: torch.from_numpy(np.ascontiguousarray(c[k])).to(dev) for k in ("p_lev", "t_lay", "q_lay")}
    t = lambda a: torch.from_numpy(a).to(dev)
    if ld is None:
        out = aerosol.optics(clim, cols, t(lat), t(lon), t(sl))
        torch.cuda.synchronize()
        return out["aer_sw"].cpu().numpy(), out["aer_lw"].cpu().numpy()
    # padded leading dimension through the C-ABI directly
    import ctypes as C
    from woof_nam_amd import _lib, abi
    nlay, ncol = c["t_lay"].shape
    sw = torch.full((3, 14, nlay + 1, ld), -7.0, dtype=torch.float32, device=dev)
    lw = torch.full((3, 16, nlay + 1, ld), -7.0, dtype=torch.float32, device=dev)
    ci = abi.WrColumns()
    ci.ncol, ci.nlay = ncol, nlay
    for k in ("p_lev", "t_lay", "q_lay"):
        setattr(ci, k, C.cast(C.c_void_p(cols[k].data_ptr()), C.POINTER(C.c_double)))
    la, lo, s_ = t(lat), t(lon), t(sl)
    rc = _lib.load().wr_aer_optics(C.byref(clim.struct()), C.byref(ci), C.c_void_p(lo.data_ptr()),
                                   C.c_void_p(la.data_ptr()), C.c_void_p(s_.data_ptr()), C.c_void_p(sw.data_ptr()),
                                   C.c_void_p(lw.data_ptr()), ld, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc, "wr_aer_optics")
    torch.cuda.synchronize()
    return sw.cpu().numpy(), lw.cpu().numpy()


class _HostClim(aerosol.Climatology):
    """Device tables from a host dict (synthetic tables)."""

    def __init__(self, tab, dev):
        self.device = torch.device(dev)
        self.t = {k: torch.from_numpy(np.ascontiguousarray(v)).to(self.device) for k, v in tab.items()}
        self.imon = 0


@pytest.mark.parametrize("imon,n,seed", [(1, 4099, 0), (7, 1, 1), (7, 63, 2), (12, 64, 3), (4, 65, 4)])
def test_climatology_matches_oracle(gpu_device, imon, n, seed):
    print(f"aerosol climatology month {imon} ncol {n} seed {seed}", flush=True)
    clim = aerosol.Climatology(gpu_device)
    clim.set_month(imon)
    c, lon, lat, sl = _cases(n, seed)
    sw, lw = _device(clim, c, lon, lat, sl, gpu_device)
    osw, olw = pyoracle.aer_optics(aerosol.host_tables(imon), c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl)
    np.testing.assert_array_equal(sw, osw)
    np.testing.assert_array_equal(lw, olw)
    assert (sw[0] >= 0).all() and (sw[0].sum(1) > 0).all()


def test_padded_leading_dimension(gpu_device):
    print("aerosol padded ld 300 ncol 257", flush=True)
    clim = aerosol.Climatology(gpu_device)
    clim.set_month(3)
    c, lon, lat, sl = _cases(257, 5)
    sw, lw = _device(clim, c, lon, lat, sl, gpu_device, ld=300)
    osw, olw = pyoracle.aer_optics(aerosol.host_tables(3), c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl)
    np.testing.assert_array_equal(sw[..., :257], osw)
    np.testing.assert_array_equal(lw[..., :257], olw)
    assert (sw[..., 257:] == -7.0).all() and (lw[..., 257:] == -7.0).all()   # the padding is not written


@pytest.mark.parametrize("kw", [dict(), dict(rh_linear=True)], ids=["synthetic", "synthetic-rh-linear"])
def test_synthetic_tables_match_oracle(gpu_device, kw):
    print(f"aerosol synthetic tables {kw}", flush=True)
    tab = synth(**kw)
    tab["kprfg"][2, 1] = 6
    tab["kprfg"][20, 40] = 7
    c, lon, lat, sl = _cases(1000, 6)
    sw, lw = _device(_HostClim(tab, gpu_device), c, lon, lat, sl, gpu_device)
    osw, olw = pyoracle.aer_optics(tab, c["p_lev"], c["t_lay"], c["q_lay"], lon, lat, sl)
    np.testing.assert_array_equal(sw, osw)
    np.testing.assert_array_equal(lw, olw)
