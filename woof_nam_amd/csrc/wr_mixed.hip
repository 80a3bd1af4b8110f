// wr_mixed.hip -- the mixed-precision path: kernel instantiations for R = float and the tile launch.
// FP32: gas-optics interpolation, Planck source, cloud optics, McICA-masked two-stream sweeps and
// the per-band g-point sums.  FP64: pressure/temperature index arithmetic (layer prep and setcoef),
// the SW layer R/T where the single-scattering albedo exceeds 1 - 1e-6, the level-flux
// accumulators and the flux differencing into heating rates.
#include "wr_launch.h"

using namespace wr;

int wr_launch_mixed(wr_ctx *c, const ColIn &in, const RunPar &par, const OutPtrs &out, int col0, int n, int band,
                   hipStream_t s)
{
    if (has_ext(in)) return wr_launch_mixed_ext(c, in, par, out, col0, n, band, s);
    if (par.mcica == 0) return launch_tile_nam<float, false>(c, c->tabs_f, c->sflux_f, in, par, out, col0, n, band, s);
    return launch_tile<float, false, false>(c, c->tabs_f, c->sflux_f, in, par, out, col0, n, band, s);
}

int wr_launch_diag_mixed(wr_ctx *c, const ColIn &in, const RunPar &par, const wr_diag *dg, int col0, int n, hipStream_t s)
{
    return launch_diag<float>(c, c->tabs_f, in, par, dg, col0, n, s);
}
