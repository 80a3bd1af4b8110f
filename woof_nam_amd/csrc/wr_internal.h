// wr_internal.h -- host-side context of libwoofrad (shared by the host ABI and the launch units).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/woofrad.h"
#include "wr_kernels.h"

#define WR_NAUX 4  // auxiliary streams the band kernels of a tile are spread over
// upper bound of the device memory of the two table blobs (FP64 and FP32 tables, about 3.5 MB with
// their look-up tables) and their structs
#define WR_TABLE_BYTES (8ull << 20)

struct wr_ctx {
    int device;
    int max_cols;
    int nl_max;                      // radiation layers the work space is sized for
    void *tab_buf_d, *tab_buf_f;     // device table storage (FP64, FP32)
    wr::Tabs<double> *tabs_d;        // device copies of the table structs
    wr::Tabs<float> *tabs_f;
    wr::Tabs<double> host_d;         // host images (device pointers inside)
    wr::Tabs<float> host_f;
    // tile work space
    double *acc;                     // band flux planes: 60 x [NL + 1][max_cols] + 14 x [max_cols]
    int *iws;                        // 7 x [max_cols]: 6 column words, then day-list segment counts
    double *dws;                     // (4 + 2 NL) x [max_cols]: 3 column words, layer logs, level temperatures
    void *rtr;                       // maximum-random overlap factors of the NAM LW route
    double *sflux_d;                 // [112][max_cols]
    float *sflux_f;
    void *lwrec;                     // LW layer records of a tile (g-lane LW kernels), sized for FP64
    int lw_glane;                    // 1: McICA LW bands as g-lane kernels over the layer records
    uint32_t lw_lds;                 // bit B: the FP32 McICA kernel of LW band B + 1 stages its tables in LDS
    uint32_t lw_lds64;               // the same for the FP64 McICA kernels
    uint32_t sw_lds, sw_lds64;       // bit B: the McICA SW kernel of band B + 16 stages its tables in LDS
    uint32_t lwr_lds, lwr_lds64;     // bit B: the NAM-route LW kernel of band B + 1 stages its tables in LDS
    uint32_t swr_lds, swr_lds64;     // bit B: the NAM-route SW kernel of band B + 16 stages its tables in LDS
    int n_cu, lds_max;               // compute units; LDS bytes per workgroup
    // learned gas optics (wr_surrogate_attach): network weights and tile work space
    int has_sur;
    wr_surrogate sur;
    float *sx;                       // features [S][8]
    void *sy;                        // network outputs, bfloat16 [504][S]
    int *sflag;                      // [max_cols] per-column table-optics flag
    // per-stream cache pools of the band kernels (pool[0]: the caller's stream and aux stream 0;
    // pool[q]: aux stream q), pool_bytes each, all allocated by wr_create
    char *pool[WR_NAUX];
    size_t pool_bytes;
    // band-kernel streams (default priority, non-blocking) and their fork/join events
    int naux;
    hipStream_t aux[WR_NAUX];
    hipEvent_t fork, join[WR_NAUX];
    // optional per-launch event timing (wr_timing)
    int timing, ev_cap, ev_n;
    hipEvent_t *ev;                  // 2 x ev_cap
    int *ev_kid;
};

// kernel ids reported by wr_timing_read
enum { KID_ZERO = 0, KID_COLPREP = 1, KID_DAYLIST = 2, KID_LW0 = 3, KID_SW0 = 19, KID_POST = 33, KID_DIAG = 34,
       KID_FEAT = 35, KID_MLP = 36, KID_RTRFAC = 37, KID_LWREC = 38, KID_COUNT = 39 };

inline void tbeg(wr_ctx *c, int kid, hipStream_t s)
{
    if (c->timing && c->ev_n < c->ev_cap) {
        (void)hipEventRecord(c->ev[2 * c->ev_n], s);
        c->ev_kid[c->ev_n] = kid;
    }
}
inline void tend(wr_ctx *c, hipStream_t s)
{
    if (c->timing && c->ev_n < c->ev_cap) {
        (void)hipEventRecord(c->ev[2 * c->ev_n + 1], s);
        c->ev_n++;
    }
}

namespace wr {
// The device view of a call's column inputs.  Every ColIn is made here: value-initialised, then every
// field of wr_columns copied, so that no field a kernel tests (the optional path form, aerosol, top
// ozone) is ever left indeterminate.
inline ColIn to_colin(const wr_columns *in)
{
    ColIn ci{};
    ci.p_lev = gptr(in->p_lev), ci.t_lay = gptr(in->t_lay), ci.q_lay = gptr(in->q_lay), ci.o3_lay = gptr(in->o3_lay);
    ci.cldf = gptr(in->cldf), ci.qc = gptr(in->qc), ci.qi = gptr(in->qi), ci.t_sfc = gptr(in->t_sfc);
    ci.emis = gptr(in->emis), ci.alb = gptr(in->alb), ci.cosz = gptr(in->cosz), ci.col_id = gptr(in->col_id);
    ci.seed_dev = gptr(in->seed_dev), ci.ncol_dev = gptr(in->ncol_dev);
    ci.cwp = gptr(in->cwp), ci.cip = gptr(in->cip), ci.crp = gptr(in->crp), ci.rew = gptr(in->rew), ci.rei = gptr(in->rei);
    ci.aer_sw = gptr(in->aer_sw), ci.aer_lw = gptr(in->aer_lw);
    ci.aer_ld = in->aer_ld > 0 ? (long)in->aer_ld : (long)in->ncol;
    ci.o3_top = gptr(in->o3_top);
    ci.ncol = in->ncol, ci.nlay = in->nlay;
    return ci;
}
inline ColScal colscal(wr_ctx *c)
{
    const size_t m = (size_t)c->max_cols;
    ColScal s;
    s.lw_laytrop = gptr(c->iws + 0 * m);
    s.lcf1 = gptr(c->iws + 1 * m);
    s.day = gptr(c->iws + 2 * m);
    s.sw_laytrop = gptr(c->iws + 3 * m);
    s.dlist = gptr(c->iws + 4 * m);
    s.nday = gptr(c->iws + 5 * m);
    s.pwv = gptr(c->dws + 0 * m);
    s.zcf1 = gptr(c->dws + 1 * m);
    s.zcf0 = gptr(c->dws + 2 * m);
    s.la = gptr(c->dws + 3 * m);                 // [nrl][n], nrl <= NL
    s.tz = gptr(c->dws + (3 + (size_t)NL) * m);  // [nrl + 1][n]
    return s;
}
}  // namespace wr

// launch units
int wr_launch_fp64(wr_ctx *c, const wr::ColIn &in, const wr::RunPar &par, const wr::OutPtrs &out, int col0, int n,
                   int band, hipStream_t s);
int wr_launch_mixed(wr_ctx *c, const wr::ColIn &in, const wr::RunPar &par, const wr::OutPtrs &out, int col0, int n,
                    int band, hipStream_t s);
// the same tiles with the band kernels that read the optional column inputs (path-form clouds, aerosol,
// top-layer ozone); their own units, so both instance sets compile in parallel
int wr_launch_fp64_ext(wr_ctx *c, const wr::ColIn &in, const wr::RunPar &par, const wr::OutPtrs &out, int col0, int n,
                       int band, hipStream_t s);
int wr_launch_mixed_ext(wr_ctx *c, const wr::ColIn &in, const wr::RunPar &par, const wr::OutPtrs &out, int col0, int n,
                        int band, hipStream_t s);
int wr_launch_diag_fp64(wr_ctx *c, const wr::ColIn &in, const wr::RunPar &par, const wr_diag *dg, int col0, int n,
                        hipStream_t s);
int wr_launch_diag_mixed(wr_ctx *c, const wr::ColIn &in, const wr::RunPar &par, const wr_diag *dg, int col0, int n,
                         hipStream_t s);
int wr_init_luts(wr_ctx *c, hipStream_t s);
int wr_launch_surr(wr_ctx *c, const wr::ColIn &in, const wr::RunPar &par, const wr::OutPtrs &out, int col0, int n,
                   int32_t *flag_out, hipStream_t s);
