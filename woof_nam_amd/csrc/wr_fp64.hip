// wr_fp64.hip -- the plain FP64 path: kernel instantiations for R = double and the tile launch.
// Compiled with -ffp-contract=off: every operation rounds as written, so table indices and McICA
// bits are reproducible bit for bit against a sequential IEEE evaluation.
#include "wr_launch.h"

namespace {
using namespace wr;

// rlwinit / rswinit look-up tables (phys_rrtm_lw_rtrnmr.py:97-132, phys_rrtm_sw_rswinit.cu:310-345)
__global__ void k_luts(double *exp_tbl, double *tfn_tbl, double *tau_tbl, double *sw_exp_tbl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > NTBL) return;
    const double bpade = 1.0 / 0.278;
    if (i == 0) {
        exp_tbl[0] = 1.0, tau_tbl[0] = 0.0, tfn_tbl[0] = 0.0, sw_exp_tbl[0] = 1.0;
    } else if (i == NTBL) {
        exp_tbl[NTBL] = 1.0e-20, tau_tbl[NTBL] = 1.0e10, tfn_tbl[NTBL] = 1.0, sw_exp_tbl[NTBL] = 1.0e-20;
    } else {
        const double tfn = (double)i / (double)(NTBL - i);
        const double tau = bpade * tfn;
        const double e = (tau >= 87.0) ? 1.0e-20 : exp_d(-tau);
        tau_tbl[i] = tau;
        exp_tbl[i] = e;
        tfn_tbl[i] = (tau < 0.06) ? tau / 6.0 : 1.0 - 2.0 * ((1.0 / tau) - (e / (1.0 - e)));
        sw_exp_tbl[i] = exp_d(-(bpade * tfn));
    }
}
}  // namespace

int wr_init_luts(wr_ctx *c, hipStream_t s)
{
    hipLaunchKernelGGL(k_luts, dim3((wr::NTBL + 256) / 256), dim3(256), 0, s, (double *)c->host_d.exp_tbl,
                       (double *)c->host_d.tfn_tbl, (double *)c->host_d.tau_tbl, (double *)c->host_d.sw_exp_tbl);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

int wr_launch_fp64(wr_ctx *c, const ColIn &in, const RunPar &par, const OutPtrs &out, int col0, int n, int band,
                   hipStream_t s)
{
    if (has_ext(in)) return wr_launch_fp64_ext(c, in, par, out, col0, n, band, s);
    if (par.mcica == 0) return launch_tile_nam<double, false>(c, c->tabs_d, c->sflux_d, in, par, out, col0, n, band, s);
    return launch_tile<double, false, false>(c, c->tabs_d, c->sflux_d, in, par, out, col0, n, band, s);
}

int wr_launch_diag_fp64(wr_ctx *c, const ColIn &in, const RunPar &par, const wr_diag *dg, int col0, int n, hipStream_t s)
{
    return launch_diag<double>(c, c->tabs_d, in, par, dg, col0, n, s);
}
