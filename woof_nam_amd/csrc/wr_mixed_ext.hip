// wr_mixed_ext.hip -- the mixed-precision path for calls that carry optional column inputs
// (path-form clouds with radii and rain, aerosol optics, the added top layer's ozone): the band
// kernels that read them, R = float.  Compiled with the flags of wr_mixed.hip.
#include "wr_launch.h"

using namespace wr;

int wr_launch_mixed_ext(wr_ctx *c, const ColIn &in, const RunPar &par, const OutPtrs &out, int col0, int n, int band,
                        hipStream_t s)
{
    if (par.mcica == 0) return launch_tile_nam<float, true>(c, c->tabs_f, c->sflux_f, in, par, out, col0, n, band, s);
    return launch_tile<float, false, true>(c, c->tabs_f, c->sflux_f, in, par, out, col0, n, band, s);
}
