// wr_launch.h -- the tile launch sequence, shared by the FP64 and the mixed-precision units.
//
// A tile runs the column kernels on the caller's stream, then the 30 band kernels (16 LW, 14 SW),
// then the FP64 post.  Every band kernel writes its own flux planes (BandAcc), so the band kernels
// are independent of one another: they are spread over the context's auxiliary streams (fork and
// join through events, also under graph capture once the streams' pools exist), which fills the CUs
// that a finishing band kernel leaves idle with the next band's waves.  k_post adds the planes in
// band order, so the result does not depend on which stream ran which band.
#pragma once
#include <algorithm>
#include <atomic>

#include "wr_internal.h"
#include "wr_glw.h"
#include "wr_nam.h"

namespace wr {

// Band kernel id: 0..15 = LW band 1..16, 16..29 = SW band 16..29.
// Static longest-processing-time plan of the 30 band kernels over WR_NAUX streams (from mixed-path
// kernel times measured on 262,144-column tiles; -1 ends a row).  A single band kernel fills the CUs,
// so the streams only fill the tails of kernels: 2 % (FP64) and 5 % (mixed) against one stream on the
// whole grid, and the plan's balance matters little.
constexpr int BAND_PLAN[WR_NAUX][9] = {
    {2, 17, 11, 20, 19, 27, 15, 22, -1},
    {4, 0, 7, 23, 24, 16, 25, -1, -1},
    {8, 1, 10, 5, 9, 18, 14, 13, -1},
    {6, 3, 29, 21, 12, 26, 28, -1, -1},
};

constexpr bool band_plan_complete()
{
    int seen[30] = {};
    for (int q = 0; q < WR_NAUX; ++q)
        for (int r = 0; r < 9 && BAND_PLAN[q][r] >= 0; ++r) {
            const int id = BAND_PLAN[q][r];
            if (id >= 30 || seen[id]++) return false;
        }
    for (int id = 0; id < 30; ++id)
        if (seen[id] != 1) return false;
    return true;
}
static_assert(band_plan_complete(), "every band kernel appears exactly once in BAND_PLAN");

// Day list of a tile: segment counts into the words after nday (cs.nday + 1 .. + nseg; the integer
// work space holds 2 max_cols - 1 words there, nseg <= max_cols), then their scan, then the ordered
// scatter.
inline void launch_daylist(const ColScal &cs, int n, hipStream_t s)
{
    const int nseg = (n + DAY_SEG - 1) / DAY_SEG;
    int *seg = (int *)cs.nday + 1;
    hipLaunchKernelGGL(k_daycount, dim3(nseg), dim3(DAY_T), 0, s, (const int *)cs.day, n, seg);
    hipLaunchKernelGGL(k_dayscan, dim3(1), dim3(1024), 0, s, seg, nseg, (int *)cs.nday);
    hipLaunchKernelGGL(k_dayscatter, dim3(nseg), dim3(DAY_T), 0, s, (const int *)cs.day, n, (const int *)seg,
                       (int *)cs.dlist);
}

inline BandAcc tile_planes(wr_ctx *c, int n, int nrl, int band)
{
    BandAcc ba;
    ba.base = gptr(c->acc);
    ba.P = (long long)(nrl + 1) * n;
    ba.n = n;
    ba.lw_mask = band < 0 ? 0xffffu : (band < 16 ? 1u << band : 0u);
    ba.sw_mask = band < 0 ? 0x3fffu : (band >= 16 ? 1u << (band - 16) : 0u);
    return ba;
}

// Everything a band launch needs besides the band id.
template <typename R>
struct BandJob {
    wr_ctx *c;
    const Tabs<R> *tabs;
    const R *sflux;
    ColIn in;
    RunPar par;
    int col0, n;
    ColScal cs;
    BandAcc ba;
    Surr sg;
    Rtr<R> rt;
    LwRec<R> rec;
};

// The LW layer records of a tile in the context's record space: [nk][n] index words, then
// [REC_R][nk][n] values of precision R, then [REC_D][nk][n] FP64 values (256-byte aligned parts).
template <typename R>
inline LwRec<R> tile_lwrec(wr_ctx *c, int n, int nk)
{
    LwRec<R> r;
    char *b = (char *)c->lwrec;
    const size_t a = ((size_t)nk * n * sizeof(int32_t) + 255) & ~(size_t)255;
    const size_t v = ((size_t)REC_R * nk * n * sizeof(R) + 255) & ~(size_t)255;
    r.idx = gptr((int32_t *)b);
    r.r = gptr((R *)(b + a));
    r.d = gptr((double *)(b + a + v));
    r.nk = nk;
    r.n = n;
    return r;
}

// McICA route band kernels (k_lw / k_sw), dispatched from a runtime id to the template instance.
// Grid of a band kernel: one 64-lane block per 64 columns, at most one block per pool slot (the
// kernels loop over the remaining blocks).
template <typename R>
inline unsigned band_grid(const BandJob<R> &j, int nv, int ng)
{
    const size_t slot = cache_slot_bytes(nv, ng, j.in.nlay + 2, (int)sizeof(R));
    const size_t nslot = j.c->pool_bytes / slot;
    const size_t nblk = (size_t)(j.n + 63) / 64;
    return (unsigned)(nblk < nslot ? nblk : nslot);
}

template <typename R, bool SURR, bool EXT, int B = 0>
void launch_band(const BandJob<R> &j, int id, hipStream_t s, char *pool)
{
    if constexpr (B < 30) {
        if (id != B) return launch_band<R, SURR, EXT, B + 1>(j, id, s, pool);
        const AccT<R> acc = band_acc<R>(j.ba, B);
        tbeg(j.c, (B < 16 ? KID_LW0 + B : KID_SW0 + B - 16), s);
        if constexpr (B < 16) {
            if (j.c->lw_glane) {
                constexpr int CPW = 64 / pow2_at_least(LW_NG[B]);
                hipLaunchKernelGGL((k_glw<R, B, SURR>), dim3((j.n + CPW - 1) / CPW), dim3(64), 0, s, j.tabs, j.in,
                                   j.par, j.col0, j.n, j.cs, acc, j.sg, j.rec);
                tend(j.c, s);
                return;
            }
        }
        if constexpr (B < 16 && sizeof(R) == 4 && lw_gsplit<R>() == 1) {
            // k_lwl: as many persistent workgroups as the CUs hold at once (the runtime's occupancy for
            // the kernel's registers and LDS), each wave with its own pool slot; at most as many as the
            // pool has slots for (none fit: k_lwc)
            constexpr unsigned WPG = lwl_wpg<B>();
            const size_t lds = (size_t)j.c->host_f.lw_img[B].n * sizeof(float);
            static std::atomic<int> per_cu{-1};  // workgroups per CU, once per instance (one table set per process)
            if (per_cu < 0 && ((j.c->lw_lds >> B) & 1)) {
                int nb = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_lwl<B, SURR, EXT>, 64 * WPG, lds) !=
                    hipSuccess)
                    nb = 0;
                per_cu = nb;
            }
            const size_t slot = cache_slot_bytes(LW_NV_F32, LW_NG[B], j.in.nlay + 2, (int)sizeof(R));
            const size_t nwg_pool = j.c->pool_bytes / slot / WPG;
            const size_t nwg_cols = ((size_t)j.n + 64 * WPG - 1) / (64 * WPG);
            size_t nwg = (size_t)(j.c->n_cu > 0 ? j.c->n_cu : 1) * (size_t)std::max(per_cu.load(), 0);
            nwg = nwg < nwg_pool ? nwg : nwg_pool;
            nwg = nwg < nwg_cols ? nwg : nwg_cols;
            if (((j.c->lw_lds >> B) & 1) && nwg > 0) {
                hipLaunchKernelGGL((k_lwl<B, SURR, EXT>), dim3((unsigned)nwg), dim3(64 * WPG), lds, s, j.tabs, j.in, j.par,
                                   j.col0, j.n, j.cs, acc, j.sg, pool);
            } else {
                hipLaunchKernelGGL((k_lwc<B, SURR, EXT>), dim3(band_grid(j, LW_NV_F32, LW_NG[B])), dim3(64), 0, s, j.tabs,
                                   j.in, j.par, j.col0, j.n, j.cs, acc, j.sg, pool);
            }
        } else if constexpr (B < 16 && sizeof(R) == 8 && lw_gsplit<R>() == 1) {
            // k_lwl64: persistent workgroups, as many as the CUs hold at once
            constexpr unsigned WPG = lwl64_wpg<R, B>();
            const size_t lds = (size_t)j.c->host_d.lw_img[B].n * sizeof(double);
            static std::atomic<int> per_cu{-1};  // workgroups per CU, once per instance (one table set per process)
            if (per_cu < 0 && ((j.c->lw_lds64 >> B) & 1)) {
                int nb = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_lwl64<B, SURR, EXT>, 64 * WPG, lds) !=
                    hipSuccess)
                    nb = 0;
                per_cu = nb;
            }
            const size_t nwg_cols = ((size_t)j.n + 64 * WPG - 1) / (64 * WPG);
            size_t nwg = (size_t)(j.c->n_cu > 0 ? j.c->n_cu : 1) * (size_t)std::max(per_cu.load(), 0);
            nwg = nwg < nwg_cols ? nwg : nwg_cols;
            if (((j.c->lw_lds64 >> B) & 1) && nwg > 0)
                hipLaunchKernelGGL((k_lwl64<B, SURR, EXT>), dim3((unsigned)nwg), dim3(64 * WPG), lds, s, j.tabs, j.in, j.par,
                                   j.col0, j.n, j.cs, acc, j.sg);
            else
                hipLaunchKernelGGL((k_lw<R, B, SURR, EXT>), dim3((j.n + 63) / 64), dim3(64), 0, s, j.tabs, j.in, j.par, j.col0,
                                   j.n, j.cs, acc, j.sg);
        } else if constexpr (B < 16)
            hipLaunchKernelGGL((k_lw<R, B, SURR, EXT>), dim3((j.n * lw_gsplit<R>() + 63) / 64), dim3(64), 0, s, j.tabs, j.in,
                               j.par, j.col0, j.n, j.cs, acc, j.sg);
        else {
            // k_swl: persistent workgroups, as many as the CUs hold at once and the pool has slots for
            constexpr int S = B < 16 ? 0 : B - 16;
            constexpr unsigned WPG = swl_wpg<R, S>();
            const uint32_t on = ((sizeof(R) == 4 ? j.c->sw_lds : j.c->sw_lds64) >> S) & 1u;
            int img_n = 0;
            if constexpr (sizeof(R) == 4)
                img_n = j.c->host_f.sw_img[S].n;
            else
                img_n = j.c->host_d.sw_img[S].n;
            const size_t lds = (size_t)img_n * sizeof(R);
            static std::atomic<int> per_cu{-1};  // workgroups per CU, once per instance (one table set per process)
            if (per_cu < 0 && on) {
                int nb = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_swl<R, S, SURR, EXT>, 64 * WPG, lds) !=
                    hipSuccess)
                    nb = 0;
                per_cu = nb;
            }
            const size_t slot = cache_slot_bytes(sw_nv_mc<R>(), SW_NG[S], j.in.nlay + 2, (int)sizeof(R));
            const size_t nwg_pool = j.c->pool_bytes / slot / WPG;
            const size_t nwg_cols = ((size_t)j.n + 64 * WPG - 1) / (64 * WPG);
            size_t nwg = (size_t)(j.c->n_cu > 0 ? j.c->n_cu : 1) * (size_t)std::max(per_cu.load(), 0);
            nwg = nwg < nwg_pool ? nwg : nwg_pool;
            nwg = nwg < nwg_cols ? nwg : nwg_cols;
            if (on && nwg > 0)
                hipLaunchKernelGGL((k_swl<R, S, SURR, EXT>), dim3((unsigned)nwg), dim3(64 * WPG), lds, s, j.tabs, j.in, j.par,
                                   j.col0, j.n, j.cs, j.sflux, acc, j.sg, pool);
            else
                hipLaunchKernelGGL((k_sw<R, S, SURR, EXT>), dim3(band_grid(j, sw_nv_mc<R>(), SW_NG[S])), dim3(64), 0, s, j.tabs,
                                   j.in, j.par, j.col0, j.n, j.cs, j.sflux, acc, j.sg, pool);
        }
        tend(j.c, s);
    }
}

// NAM route band kernels (k_lwr / k_swr).
template <typename R, bool EXT, int B = 0>
void launch_band_nam(const BandJob<R> &j, int id, hipStream_t s, char *pool)
{
    if constexpr (B < 30) {
        if (id != B) return launch_band_nam<R, EXT, B + 1>(j, id, s, pool);
        const AccT<R> acc = band_acc<R>(j.ba, B);
        tbeg(j.c, (B < 16 ? KID_LW0 + B : KID_SW0 + B - 16), s);
        if constexpr (B < 16) {
            // k_lwrl: persistent workgroups with the band's tables in LDS, as many as the CUs hold
            constexpr unsigned WPG = lwrl_wpg<R, B>();
            const uint32_t on = ((sizeof(R) == 4 ? j.c->lwr_lds : j.c->lwr_lds64) >> B) & 1u;
            int img_n = 0;
            if constexpr (sizeof(R) == 4)
                img_n = j.c->host_f.lw_img[B].n;
            else
                img_n = j.c->host_d.lw_img[B].n;
            const size_t lds = (size_t)img_n * sizeof(R);
            static std::atomic<int> per_cu{-1};  // workgroups per CU, once per instance (one table set per process)
            if (per_cu < 0 && on) {
                int nb = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_lwrl<R, B, EXT>, 64 * WPG, lds) !=
                    hipSuccess)
                    nb = 0;
                per_cu = nb;
            }
            const size_t nwg_cols = ((size_t)j.n + 64 * WPG - 1) / (64 * WPG);
            size_t nwg = (size_t)(j.c->n_cu > 0 ? j.c->n_cu : 1) * (size_t)std::max(per_cu.load(), 0);
            nwg = nwg < nwg_cols ? nwg : nwg_cols;
            if (on && nwg > 0)
                hipLaunchKernelGGL((k_lwrl<R, B, EXT>), dim3((unsigned)nwg), dim3(64 * WPG), lds, s, j.tabs, j.in, j.par,
                                   j.col0, j.n, j.cs, acc, j.rt);
            else
                hipLaunchKernelGGL((k_lwr<R, B, EXT>), dim3((j.n + 63) / 64), dim3(64), 0, s, j.tabs, j.in, j.par, j.col0,
                                   j.n, j.cs, acc, j.rt);
        } else {
            // k_swrl: persistent workgroups with the band's tables in LDS, within the pool's slots
            constexpr int S = B < 16 ? 0 : B - 16;
            constexpr unsigned WPG = swrl_wpg<R, S>();
            const uint32_t on = ((sizeof(R) == 4 ? j.c->swr_lds : j.c->swr_lds64) >> S) & 1u;
            int img_n = 0;
            if constexpr (sizeof(R) == 4)
                img_n = j.c->host_f.sw_img[S].n;
            else
                img_n = j.c->host_d.sw_img[S].n;
            const size_t lds = (size_t)img_n * sizeof(R);
            static std::atomic<int> per_cu{-1};  // workgroups per CU, once per instance (one table set per process)
            if (per_cu < 0 && on) {
                int nb = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_swrl<R, S, EXT>, 64 * WPG, lds) !=
                    hipSuccess)
                    nb = 0;
                per_cu = nb;
            }
            const size_t slot = cache_slot_bytes(sw_nv_mc<R>(), SW_NG[S], j.in.nlay + 2, (int)sizeof(R));
            const size_t nwg_pool = j.c->pool_bytes / slot / WPG;
            const size_t nwg_cols = ((size_t)j.n + 64 * WPG - 1) / (64 * WPG);
            size_t nwg = (size_t)(j.c->n_cu > 0 ? j.c->n_cu : 1) * (size_t)std::max(per_cu.load(), 0);
            nwg = nwg < nwg_pool ? nwg : nwg_pool;
            nwg = nwg < nwg_cols ? nwg : nwg_cols;
            if (on && nwg > 0)
                hipLaunchKernelGGL((k_swrl<R, S, EXT>), dim3((unsigned)nwg), dim3(64 * WPG), lds, s, j.tabs, j.in, j.par,
                                   j.col0, j.n, j.cs, j.sflux, acc, pool);
            else
                hipLaunchKernelGGL((k_swr<R, S, EXT>), dim3(band_grid(j, sw_nv_mc<R>(), SW_NG[S])), dim3(64), 0, s, j.tabs,
                                   j.in, j.par, j.col0, j.n, j.cs, j.sflux, acc, pool);
        }
        tend(j.c, s);
    }
}

// True when the band kernels may leave stream s: the context has auxiliary streams (their pools are
// allocated by wr_create).  Under graph capture the fork and join events bring the auxiliary streams
// into the graph, so a captured tile holds the same band concurrency as an eager one.
inline bool spread_bands(wr_ctx *c, hipStream_t s)
{
    (void)s;
    if (c->naux <= 0) return false;
    for (int q = 1; q < c->naux; ++q)
        if (!c->pool[q]) return false;
    return true;
}

// Launches the band kernels of a tile: one band (band >= 0) or all of them, in band order on s, or
// by BAND_PLAN over the auxiliary streams with fork and join events around them.
template <typename F>
int run_bands(wr_ctx *c, int band, hipStream_t s, F &&one)
{
    if (band >= 0) {
        one(band, s, c->pool[0]);
        return WR_OK;
    }
    if (!spread_bands(c, s)) {
        for (int id = 0; id < 30; ++id) one(id, s, c->pool[0]);
        return WR_OK;
    }
    if (hipEventRecord(c->fork, s) != hipSuccess) return WR_ERR_HIP;
    for (int q = 0; q < c->naux; ++q)
        if (hipStreamWaitEvent(c->aux[q], c->fork, 0) != hipSuccess) return WR_ERR_HIP;
    // aux stream 0 shares pool 0 with the caller's stream: band kernels on s never overlap the
    // auxiliary ones (fork and join events order them)
    for (int q = 0; q < WR_NAUX; ++q) {
        const int a = q % c->naux;
        for (int r = 0; r < 9 && BAND_PLAN[q][r] >= 0; ++r) one(BAND_PLAN[q][r], c->aux[a], c->pool[a]);
    }
    for (int q = 0; q < c->naux; ++q) {
        if (hipEventRecord(c->join[q], c->aux[q]) != hipSuccess) return WR_ERR_HIP;
        if (hipStreamWaitEvent(s, c->join[q], 0) != hipSuccess) return WR_ERR_HIP;
    }
    return WR_OK;
}

// One tile of n columns from col0 (McICA route): column scalars, day list, band kernels, FP64 post.
// No host synchronisation.  EXT selects the band kernels that read the optional column inputs (the
// caller passes EXT = has_ext(in) or true).
template <typename R, bool SURR = false, bool EXT = true>
int launch_tile(wr_ctx *c, const Tabs<R> *tabs, R *sflux, const ColIn &in, const RunPar &par, const OutPtrs &out,
                int col0, int n, int band, hipStream_t s, const Surr &sg = Surr{})
{
    const int nrl = in.nlay + 1;
    BandJob<R> j{c, tabs, sflux, in, par, col0, n, colscal(c), tile_planes(c, n, nrl, band), sg, Rtr<R>{},
                 tile_lwrec<R>(c, n, nrl)};
    tbeg(c, KID_COLPREP, s);
    hipLaunchKernelGGL((k_colprep<R>), dim3((n + 255) / 256), dim3(256), 0, s, tabs, in, par, col0, n, j.cs, sflux);
    tend(c, s);
    tbeg(c, KID_DAYLIST, s);
    launch_daylist(j.cs, n, s);
    tend(c, s);
    if (c->lw_glane && (band < 0 || band < 16)) {
        tbeg(c, KID_LWREC, s);
        hipLaunchKernelGGL((k_lwrec<R>), dim3((n + 63) / 64), dim3(64), 0, s, tabs, in, par, col0, n, j.cs, j.rec);
        tend(c, s);
    }
    int rc = run_bands(c, band, s, [&](int id, hipStream_t q, char *pool) { launch_band<R, SURR, EXT>(j, id, q, pool); });
    if (rc != WR_OK) return rc;
    tbeg(c, KID_POST, s);
    hipLaunchKernelGGL(k_post<R>, dim3((n + 255) / 256), dim3(256), 0, s, in, col0, n, j.cs, j.ba, out);
    tend(c, s);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

// The NAM route (mcica = 0): overlap factors once per tile, then rtrnmr and cldfrc-mixed spcvrtc
// band kernels.
template <typename R, bool EXT = true>
int launch_tile_nam(wr_ctx *c, const Tabs<R> *tabs, R *sflux, const ColIn &in, const RunPar &par, const OutPtrs &out,
                    int col0, int n, int band, hipStream_t s)
{
    const int nrl = in.nlay + 1;
    Rtr<R> rt;
    rt.f = gptr((R *)c->rtr);
    rt.lstu = gptr((uint64_t *)((char *)c->rtr + sizeof(double) * 12 * (NL + 3) * (size_t)c->max_cols));
    rt.lstd = rt.lstu + c->max_cols;
    rt.n = n;
    BandJob<R> j{c, tabs, sflux, in, par, col0, n, colscal(c), tile_planes(c, n, nrl, band), Surr{}, rt, LwRec<R>{}};
    tbeg(c, KID_COLPREP, s);
    hipLaunchKernelGGL((k_colprep<R>), dim3((n + 255) / 256), dim3(256), 0, s, tabs, in, par, col0, n, j.cs, sflux);
    tend(c, s);
    tbeg(c, KID_DAYLIST, s);
    launch_daylist(j.cs, n, s);
    tend(c, s);
    tbeg(c, KID_RTRFAC, s);
    hipLaunchKernelGGL((k_rtr_fac<R>), dim3((n + 63) / 64), dim3(64), 0, s, in, col0, n, rt);
    tend(c, s);
    int rc = run_bands(c, band, s, [&](int id, hipStream_t q, char *pool) { launch_band_nam<R, EXT>(j, id, q, pool); });
    if (rc != WR_OK) return rc;
    tbeg(c, KID_POST, s);
    hipLaunchKernelGGL(k_post<R>, dim3((n + 255) / 256), dim3(256), 0, s, in, col0, n, j.cs, j.ba, out);
    tend(c, s);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

// ---- diagnostics: setcoef indices and McICA words of one tile ---------------------------------
// Runs after the tile, on its stream.  The layer setup is the instantiation of the precision that ran
// (lw_layer<R>, sw_layer<R>: the code of that path's band kernels) on the layer log pressures that
// k_colprep left in the work space for this tile.
template <typename R>
__global__ void __launch_bounds__(64) k_diag(const Tabs<R> *__restrict__ T, ColIn in, RunPar par, int col0, int n,
                                             ColScal cs, wr_diag dg)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= valid_cols(in, col0, n)) return;
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    const size_t nc = (size_t)in.ncol;
    const int day = cs.day[t];
    for (int k = 0; k < nrl; ++k) {
        const int kin = nrl - 1 - k;
        const double la = cs.la[(size_t)kin * n + t];
        LwLay<R> L;
        lw_layer<R>(T, rc, par, kin, la, L);
        if (dg.lw_jp) dg.lw_jp[k * nc + i] = L.jp;
        if (dg.lw_jt) dg.lw_jt[k * nc + i] = L.jt;
        if (dg.lw_jt1) dg.lw_jt1[k * nc + i] = L.jt1;
        SwLay<R> S;
        if (day) sw_layer<R>(T, rc, par, kin, la, S);
        if (dg.sw_jp) dg.sw_jp[k * nc + i] = day ? S.jp : -1;
        if (dg.sw_jt) dg.sw_jt[k * nc + i] = day ? S.jt : -1;
        if (dg.sw_jt1) dg.sw_jt1[k * nc + i] = day ? S.jt1 : -1;
    }
    if (dg.lw_laytrop) dg.lw_laytrop[i] = cs.lw_laytrop[t];
    if (dg.sw_laytrop) dg.sw_laytrop[i] = day ? cs.sw_laytrop[t] : -1;
    const uint64_t col = in.col_id ? in.col_id[i] : (uint64_t)i;
    const uint64_t seed = in.seed_dev ? *in.seed_dev : par.seed;
    if (dg.lw_mask) {
        const int lcf1 = cs.lcf1[t];
        for (int k = 0; k < nrl; ++k)
            for (int w = 0; w < 5; ++w) dg.lw_mask[((size_t)k * 5 + w) * nc + i] = 0u;
        for (int g = 0; g < (par.mcica ? 140 : 0); ++g) {
            const uint64_t mb = mcica_g(col, seed, 0, g, nrl, par.iovr,
                                        [&](int k) { return lcf1 ? rc.cfrac(nrl - 1 - k) : 0.0; });
            for (int k = 0; k < nrl; ++k)
                if ((mb >> k) & 1) dg.lw_mask[((size_t)k * 5 + (g >> 5)) * nc + i] |= 1u << (g & 31);
        }
    }
    if (dg.sw_mask) {
        for (int k = 0; k < nrl; ++k)
            for (int w = 0; w < 4; ++w) dg.sw_mask[((size_t)k * 4 + w) * nc + i] = 0u;
        if (day && par.mcica) {
            const double cf1 = cs.zcf1[t];
            for (int g = 0; g < 112; ++g) {
                const uint64_t mb = mcica_g(col, seed, 1, g, nrl, par.iovr, [&](int k) {
                    const double ct = rc.cfrac(nrl - 1 - k);
                    return (cf1 > 0.0) ? ((ct < FTINY) ? 0.0 : ct) : 0.0;
                });
                for (int k = 0; k < nrl; ++k)
                    if ((mb >> k) & 1) dg.sw_mask[((size_t)k * 4 + (g >> 5)) * nc + i] |= 1u << (g & 31);
            }
        }
    }
}

template <typename R>
int launch_diag(wr_ctx *c, const Tabs<R> *tabs, const ColIn &in, const RunPar &par, const wr_diag *dg, int col0, int n,
                hipStream_t s)
{
    tbeg(c, KID_DIAG, s);
    hipLaunchKernelGGL(k_diag<R>, dim3((n + 63) / 64), dim3(64), 0, s, tabs, in, par, col0, n, colscal(c), *dg);
    tend(c, s);
    return hipGetLastError() == hipSuccess ? WR_OK : WR_ERR_HIP;
}

}  // namespace wr
