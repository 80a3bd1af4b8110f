// wr_fp64_ext.hip -- the plain FP64 path for calls that carry optional column inputs (path-form
// clouds with radii and rain, aerosol optics, the added top layer's ozone): the band kernels that
// read them, R = double.  Compiled with -ffp-contract=off like wr_fp64.hip.
#include "wr_launch.h"

using namespace wr;

int wr_launch_fp64_ext(wr_ctx *c, const ColIn &in, const RunPar &par, const OutPtrs &out, int col0, int n, int band,
                       hipStream_t s)
{
    if (par.mcica == 0) return launch_tile_nam<double, true>(c, c->tabs_d, c->sflux_d, in, par, out, col0, n, band, s);
    return launch_tile<double, false, true>(c, c->tabs_d, c->sflux_d, in, par, out, col0, n, band, s);
}
