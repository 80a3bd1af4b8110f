// wr_fp64.hip -- the plain FP64 path: kernel instantiations for R = double and the tile launch.
// Compiled with -ffp-contract=off: every operation rounds as written, so table indices and McICA
// bits are reproducible bit for bit against a sequential IEEE evaluation.
#include "wr_internal.h"

namespace {
using namespace wr;

template <int B> void lw_bands(wr_ctx *c, const ColIn &in, const RunPar &par, int col0, int n, ColScal cs, Acc acc,
                               hipStream_t s)
{
    if constexpr (B < 16) {
        tbeg(c, KID_LW0 + B, s);
        hipLaunchKernelGGL((k_lw<double, B>), dim3((n + 63) / 64), dim3(64), 0, s, c->tabs_d, in, par, col0, n, cs, acc);
        tend(c, s);
        lw_bands<B + 1>(c, in, par, col0, n, cs, acc, s);
    }
}
template <int B> void sw_bands(wr_ctx *c, const ColIn &in, const RunPar &par, int col0, int n, ColScal cs, Acc acc,
                               hipStream_t s)
{
    if constexpr (B < 14) {
        tbeg(c, KID_SW0 + B, s);
        hipLaunchKernelGGL((k_sw<double, B>), dim3((n + 63) / 64), dim3(64), 0, s, c->tabs_d, in, par, col0, n, cs,
                           (const double *)c->sflux_d, acc);
        tend(c, s);
        sw_bands<B + 1>(c, in, par, col0, n, cs, acc, s);
    }
}

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

int wr_launch_fp64(wr_ctx *c, const ColIn &in, const RunPar &par, const OutPtrs &out, int col0, int n, hipStream_t s)
{
    const int nrl = in.nlay + 1;
    ColScal cs = colscal(c);
    Acc acc = tile_acc(c, n, nrl);
    const size_t nacc = (size_t)4 * (nrl + 1) * n;
    tbeg(c, KID_ZERO, s);
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, c->acc, nacc);
    tend(c, s);
    tbeg(c, KID_COLPREP, s);
    hipLaunchKernelGGL((k_colprep<double>), dim3((n + 255) / 256), dim3(256), 0, s, c->tabs_d, in, par, col0, n, cs,
                       c->sflux_d);
    tend(c, s);
    tbeg(c, KID_DAYLIST, s);
    hipLaunchKernelGGL(k_daylist, dim3(1), dim3(1024), 0, s, (const int *)cs.day, n, cs.dlist, cs.nday);
    tend(c, s);
    lw_bands<0>(c, in, par, col0, n, cs, acc, s);
    sw_bands<0>(c, in, par, col0, n, cs, acc, s);
    tbeg(c, KID_POST, s);
    hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, s, in, col0, n, cs, acc, out);
    tend(c, s);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}
