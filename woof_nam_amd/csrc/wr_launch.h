// wr_launch.h -- the tile launch sequence, shared by the FP64 and the mixed-precision units.
#pragma once
#include "wr_internal.h"
#include "wr_nam.h"

namespace wr {

// band: -1 = all bands; 0..15 = LW band 1..16 only; 16..29 = SW band 16..29 only (diagnostics)
template <typename R, int B, bool SURR = false>
void lw_bands(wr_ctx *c, const Tabs<R> *tabs, const ColIn &in, const RunPar &par, int col0, int n, ColScal cs, Acc acc,
              int band, hipStream_t s, const Surr &sg = Surr{})
{
    if constexpr (B < 16) {
        if (band < 0 || band == B) {
            tbeg(c, KID_LW0 + B, s);
            hipLaunchKernelGGL((k_lw<R, B, SURR>), dim3((n + 63) / 64), dim3(64), 0, s, tabs, in, par, col0, n, cs, acc,
                               sg);
            tend(c, s);
        }
        lw_bands<R, B + 1, SURR>(c, tabs, in, par, col0, n, cs, acc, band, s, sg);
    }
}

template <typename R, int B, bool SURR = false>
void sw_bands(wr_ctx *c, const Tabs<R> *tabs, const R *sflux, const ColIn &in, const RunPar &par, int col0, int n,
              ColScal cs, Acc acc, int band, hipStream_t s, const Surr &sg = Surr{})
{
    if constexpr (B < 14) {
        if (band < 0 || band == 16 + B) {
            tbeg(c, KID_SW0 + B, s);
            hipLaunchKernelGGL((k_sw<R, B, SURR>), dim3((n + 63) / 64), dim3(64), 0, s, tabs, in, par, col0, n, cs,
                               sflux, acc, sg);
            tend(c, s);
        }
        sw_bands<R, B + 1, SURR>(c, tabs, sflux, in, par, col0, n, cs, acc, band, s, sg);
    }
}

// One tile of n columns from col0: accumulators cleared, column scalars, day list, 16 LW and 14 SW
// band kernels, FP64 post.  Everything on stream s, no host synchronisation.
template <typename R, bool SURR = false>
int launch_tile(wr_ctx *c, const Tabs<R> *tabs, R *sflux, const ColIn &in, const RunPar &par, const OutPtrs &out,
                int col0, int n, int band, hipStream_t s, const Surr &sg = Surr{})
{
    const int nrl = in.nlay + 1;
    ColScal cs = colscal(c);
    Acc acc = tile_acc(c, n, nrl);
    const size_t nacc = (size_t)(4 * (nrl + 1) + 1) * n;
    tbeg(c, KID_ZERO, s);
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, c->acc, nacc);
    tend(c, s);
    tbeg(c, KID_COLPREP, s);
    hipLaunchKernelGGL((k_colprep<R>), dim3((n + 255) / 256), dim3(256), 0, s, tabs, in, par, col0, n, cs, sflux);
    tend(c, s);
    tbeg(c, KID_DAYLIST, s);
    hipLaunchKernelGGL(k_daylist, dim3(1), dim3(1024), 0, s, (const int *)cs.day, n, cs.dlist, cs.nday);
    tend(c, s);
    lw_bands<R, 0, SURR>(c, tabs, in, par, col0, n, cs, acc, band, s, sg);
    sw_bands<R, 0, SURR>(c, tabs, (const R *)sflux, in, par, col0, n, cs, acc, band, s, sg);
    tbeg(c, KID_POST, s);
    hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, s, in, col0, n, cs, acc, out);
    tend(c, s);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

// The NAM route (mcica = 0): overlap factors once per tile, then rtrnmr and cldfrc-mixed spcvrtc
// band kernels.
template <typename R, int B>
void lwr_bands(wr_ctx *c, const Tabs<R> *tabs, const ColIn &in, const RunPar &par, int col0, int n, ColScal cs, Acc acc,
               Rtr<R> rt, int band, hipStream_t s)
{
    if constexpr (B < 16) {
        if (band < 0 || band == B) {
            tbeg(c, KID_LW0 + B, s);
            hipLaunchKernelGGL((k_lwr<R, B>), dim3((n + 63) / 64), dim3(64), 0, s, tabs, in, par, col0, n, cs, acc, rt);
            tend(c, s);
        }
        lwr_bands<R, B + 1>(c, tabs, in, par, col0, n, cs, acc, rt, band, s);
    }
}

template <typename R, int B>
void swr_bands(wr_ctx *c, const Tabs<R> *tabs, const R *sflux, const ColIn &in, const RunPar &par, int col0, int n,
               ColScal cs, Acc acc, int band, hipStream_t s)
{
    if constexpr (B < 14) {
        if (band < 0 || band == 16 + B) {
            tbeg(c, KID_SW0 + B, s);
            hipLaunchKernelGGL((k_swr<R, B>), dim3((n + 63) / 64), dim3(64), 0, s, tabs, in, par, col0, n, cs, sflux, acc,
                               c->swc);
            tend(c, s);
        }
        swr_bands<R, B + 1>(c, tabs, sflux, in, par, col0, n, cs, acc, band, s);
    }
}

template <typename R>
int launch_tile_nam(wr_ctx *c, const Tabs<R> *tabs, R *sflux, const ColIn &in, const RunPar &par, const OutPtrs &out,
                    int col0, int n, int band, hipStream_t s)
{
    const int nrl = in.nlay + 1;
    ColScal cs = colscal(c);
    Acc acc = tile_acc(c, n, nrl);
    const size_t nacc = (size_t)(4 * (nrl + 1) + 1) * n;
    Rtr<R> rt;
    rt.f = (R *)c->rtr;
    rt.lstu = (uint64_t *)((char *)c->rtr + sizeof(double) * 12 * (NL + 3) * (size_t)c->max_cols);
    rt.lstd = rt.lstu + c->max_cols;
    rt.n = n;
    tbeg(c, KID_ZERO, s);
    hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, c->acc, nacc);
    tend(c, s);
    tbeg(c, KID_COLPREP, s);
    hipLaunchKernelGGL((k_colprep<R>), dim3((n + 255) / 256), dim3(256), 0, s, tabs, in, par, col0, n, cs, sflux);
    tend(c, s);
    tbeg(c, KID_DAYLIST, s);
    hipLaunchKernelGGL(k_daylist, dim3(1), dim3(1024), 0, s, (const int *)cs.day, n, cs.dlist, cs.nday);
    tend(c, s);
    tbeg(c, KID_RTRFAC, s);
    hipLaunchKernelGGL((k_rtr_fac<R>), dim3((n + 63) / 64), dim3(64), 0, s, in, col0, n, rt);
    tend(c, s);
    lwr_bands<R, 0>(c, tabs, in, par, col0, n, cs, acc, rt, band, s);
    swr_bands<R, 0>(c, tabs, (const R *)sflux, in, par, col0, n, cs, acc, band, s);
    tbeg(c, KID_POST, s);
    hipLaunchKernelGGL(k_post, dim3((n + 255) / 256), dim3(256), 0, s, in, col0, n, cs, acc, out);
    tend(c, s);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

}  // namespace wr
