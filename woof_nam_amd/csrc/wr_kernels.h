// wr_kernels.h -- kernels of the MI355X RRTMG radiation path (templated on the optics precision R).
//
// Per column tile (n columns starting at col0):
//   k_colprep   column scalars: LW laytrop, precipitable water, cloud flag; SW day flag, column cloud
//               cover (max-random), SW laytrop and the solar source of the 112 g-points
//   k_day*      ordered compaction of the day columns (count, scan, scatter; deterministic order)
//   k_lw<B>     one thread per column, all g-points of band B in registers: McICA masks, then the
//               rtrnmc down sweep and up sweep (optics recomputed per sweep); the band's level fluxes
//               go to its own FP64 planes (FP64, and FP32 with two lanes per column)
//   k_lwc<B>    FP32: as k_lw, with each layer's upward transmittance and source kept in the pool slot
//               by the down sweep, so that the up sweep needs no optics
//   k_sw<B>     one thread per day column: for each g-point of band B the two-stream adding method
//               as one bottom-up sweep (reflectance of the atmosphere below each level, to the pool
//               slot; FP32 also the layer R/T) and one top-down sweep (direct/diffuse transmission
//               and level fluxes)
//   k_post      FP64 level fluxes -> heating rates, top-first outputs
#pragma once
#include <type_traits>

#include "wr_lw.h"
#include "wr_sw.h"

// Default minimum waves per SIMD of the LW band kernels (register budget 512 / waves).  One
// keeps everything in registers (AGPRs included), with no scratch (private segment) that the runtime
// must back on every HIP queue running the kernel.  The per-band tables below raise it to two where
// that measured faster; the spills this costs stay at or below about 1.1 KB of scratch per lane.
#ifndef WR_LW_WAVES
#define WR_LW_WAVES 1
#endif

namespace wr {

// Level fluxes of ONE band for one tile ([nrl+1][n] planes, written once per level; sw_clr0 [n]), in
// the band kernel's precision P: FP64 planes for the FP64 kernels; FP32 planes for the FP32 kernels,
// whose band sums are FP32 values, so that nothing is lost and k_post reads half the bytes.
template <typename P> struct AccT {
    WR_G P *lw_up, *lw_dn, *sw_up, *sw_dn, *sw_clr0;
};
// All band planes of a tile: LW band b at planes 2b, 2b+1; SW band b at planes 32+2b, 33+2b; the
// clear-sky surface SW of SW band b at base + 60 P + b n (in elements of the plane type; the space is
// sized for FP64).  k_post sums the bands in band order in FP64, which reproduces the sequential
// accumulation of the reference bit for bit.
struct BandAcc {
    WR_G double *base;
    long long P;  // (nrl + 1) * n
    int n;
    unsigned lw_mask, sw_mask;  // bands present (bit b)
};
template <typename P>
__host__ __device__ inline AccT<P> band_acc(const BandAcc &ba, int id)  // id 0..15 LW, 16..29 SW
{
    AccT<P> a = {nullptr, nullptr, nullptr, nullptr, nullptr};
    WR_G P *base = (WR_G P *)ba.base;
    if (id < 16) {
        a.lw_up = base + (size_t)(2 * id) * ba.P;
        a.lw_dn = base + (size_t)(2 * id + 1) * ba.P;
    } else {
        const int b = id - 16;
        a.sw_up = base + (size_t)(32 + 2 * b) * ba.P;
        a.sw_dn = base + (size_t)(33 + 2 * b) * ba.P;
        a.sw_clr0 = base + (size_t)60 * ba.P + (size_t)b * ba.n;
    }
    return a;
}
struct ColScal {
    WR_G int *lw_laytrop, *lcf1, *day, *sw_laytrop, *dlist, *nday;
    WR_G double *pwv, *zcf1, *zcf0;
    // per tile column, written by k_colprep: la[kin * n + t] = log of the layer pressure (layer kin,
    // top-first); tz[L * n + t] = LW level temperature, L = 0 the top layer's temperature, L = kin + 1
    // the interface below layer kin, L = nrl the skin temperature
    WR_G double *la, *tz;
};

template <typename R>
__global__ void __launch_bounds__(256) k_colprep(const Tabs<R> *__restrict__ T, ColIn in, RunPar par, int col0, int n,
                                                 ColScal cs, R *__restrict__ sflux)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    if (t >= valid_cols(in, col0, n)) {  // padding column of a shorter tile: not in the day list
        cs.day[t] = 0;
        return;
    }
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    const double tem1g = 100.0 * CON_G;
    const double tem2c = 1.0e-20 * 1.0e3 * CON_AVGD;
    double tem1 = 0.0, tem2 = 0.0, z0 = 1.0, z1 = 1.0;
    int lt = 0, lcf = 0, swlt = nrl;
    int jpk[NL];
    for (int k = 0; k < nrl; ++k) {
        const int kin = nrl - 1 - k;
        const double dlp = rc.plvl(kin + 1) - rc.plvl(kin);
        const double qv = rc.qlyr(kin);
        const double h2ovmr = fmax(0.0, qv * AMDW / (1.0 - qv));
        const double tem0 = (1.0 - h2ovmr) * CON_AMD + h2ovmr * CON_AMW;
        const double cdry = tem2c * dlp / (tem1g * tem0 * (1.0 + h2ovmr));
        const double ca0 = fmax(0.0, cdry * h2ovmr);
        tem1 = tem1 + cdry + ca0;
        tem2 = tem2 + ca0;
        const double la = log_d(rc.plyr(kin));
        cs.la[(size_t)kin * n + t] = la;
        if (la > 4.56) {
            lt += 1;
            swlt = k + 1;
        }
        jpk[k] = imax(1, imin(58, (int)(36.0 - 5.0 * (la + 0.04))));
        double ct, cwp, cip;
        rc.cloud(kin, ct, cwp, cip);
        if (ct > EPS) lcf = 1;
        if (ct > FTINY) {
            z1 = fmin(z1, 1.0 - ct);
        } else if (z1 < 1.0) {
            z0 = z0 * z1;
            z1 = 1.0;
        }
    }
    cs.lw_laytrop[t] = lt;
    cs.lcf1[t] = lcf;
    // LW level temperatures, the operations the band kernels' sweeps used to repeat per band
    cs.tz[t] = rc.tlyr(0);
    for (int kin = 0; kin + 1 < nrl; ++kin)
        cs.tz[(size_t)(kin + 1) * n + t] = tlvl_v(rc.tlyr(kin + 1), rc.tlyr(kin), log_d(rc.plvl(kin + 1)),
                                                  cs.la[(size_t)(kin + 1) * n + t], cs.la[(size_t)kin * n + t]);
    cs.tz[(size_t)nrl * n + t] = rc.tskn();
    const double tem0p = 10.0 * tem2 / (AMDW * tem1 * CON_G);
    cs.pwv[t] = tem0p * rc.plvl(nrl);
    const double cz = in.cosz[i];
    const int day = cz >= 1.0e-4;
    cs.day[t] = day;
    if (!day) return;
    z0 = z0 * z1;
    if (z0 <= FTINY) z0 = 0.0;
    if (z0 > ONEMINUS) z0 = 1.0;
    cs.zcf1[t] = 1.0 - z0;
    cs.zcf0[t] = z0;
    cs.sw_laytrop[t] = swlt;
    // solar source per g-point (phys_rrtm_sw_taumol.cu:59-120)
    for (int b = 0; b < 14; ++b) {
        const int jb = b + 16, ibd = T->ibx[b], njb = SW_NG[b], ns = SW_NS[b], lr = T->layreffr[b];
        if (jb == 16 || jb == 20 || jb == 23 || jb == 25 || jb == 26 || jb == 29) {
            for (int j = 0; j < njb; ++j) sflux[(size_t)(ns + j) * n + t] = T->sf01[(ibd - 1) * 16 + j];
        } else if (jb == 27) {
            for (int j = 0; j < njb; ++j) sflux[(size_t)(ns + j) * n + t] = T->scalekur * T->sf01[(ibd - 1) * 16 + j];
        } else {
            int ks, mfs;
            const WR_G R *tab;
            if (jb == 17 || jb == 28) {
                ks = nrl;
                for (int k = swlt; k <= nrl - 1; ++k)
                    if (jpk[k - 1] < lr && jpk[k] >= lr) {
                        ks = k + 1;
                        break;
                    }
                tab = T->sf02;
                mfs = 5;
            } else {
                ks = swlt;
                for (int k = 1; k <= swlt - 1; ++k)
                    if (jpk[k - 1] < lr && jpk[k] >= lr) {
                        ks = k + 1;
                        break;
                    }
                tab = T->sf03;
                mfs = 9;
            }
            SwLay<R> L;
            const int kin = nrl - 1 - (ks - 1);
            sw_layer<R>(T, rc, par, kin, log_d(rc.plyr(kin)), L);
            const R colm1 = L.colamt[T->ix1[b] - 1];
            const R colm2 = L.colamt[T->ix2[b] - 1];
            const R speccomb = colm1 + T->sw_strrat[b] * colm2;
            const R specmult = T->sw_specwt[b] * fmin((R)0.999999, colm1 / speccomb);
            const int jsi = (int)specmult;
            const int js = 1 + jsi;
            const R fs = specmult - (R)jsi;
            const int base = ((ibd - 1) * mfs) * 16;
            for (int j = 0; j < njb; ++j) {
                const R a = tab[base + (js - 1) * 16 + j];
                const R c = tab[base + js * 16 + j];
                sflux[(size_t)(ns + j) * n + t] = a + fs * (c - a);
            }
        }
    }
}

// ---- ordered compaction of the day columns of a tile --------------------------------------------
// Three small kernels, all reads coalesced: per-segment day counts (DAY_SEG columns per 256-thread
// block), one workgroup's exclusive scan of the counts, then each segment writes its day columns in
// column order.  The result is the same ordered list whatever the launch shape.
constexpr int DAY_T = 256, DAY_IT = 8, DAY_SEG = DAY_T * DAY_IT;

// exclusive prefix of v over the 256 threads of the block (4 waves); total = the block sum
__device__ __forceinline__ int block_excl_scan256(int v, int *lds, int &total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int q = 0; q < DAY_T / 64; ++q) {
        const int c = lds[q];
        before += (q < w) ? c : 0;
        total += c;
    }
    __syncthreads();  // lds is reused by the next call
    return before + x - v;
}

static __global__ void __launch_bounds__(DAY_T) k_daycount(const int *__restrict__ day, int n, int *__restrict__ seg)
{
    __shared__ int lds[DAY_T / 64];
    const int j0 = blockIdx.x * DAY_SEG + threadIdx.x;
    int c = 0;
#pragma unroll
    for (int it = 0; it < DAY_IT; ++it) {
        const int j = j0 + it * DAY_T;
        c += (j < n && day[j] != 0) ? 1 : 0;
    }
    int tot;
    (void)block_excl_scan256(c, lds, tot);
    if (threadIdx.x == 0) seg[blockIdx.x] = tot;
}

// in place: segment counts -> exclusive segment offsets; *nday = the total
static __global__ void __launch_bounds__(1024) k_dayscan(int *__restrict__ seg, int nseg, int *__restrict__ nday)
{
    __shared__ int part[1024];
    const int tid = threadIdx.x;
    int carry = 0;
    for (int base = 0; base < nseg; base += 1024) {
        const int v = (base + tid < nseg) ? seg[base + tid] : 0;
        part[tid] = v;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
            const int u = (tid >= off) ? part[tid - off] : 0;
            __syncthreads();
            part[tid] += u;
            __syncthreads();
        }
        if (base + tid < nseg) seg[base + tid] = carry + part[tid] - v;
        const int tot = part[1023];
        __syncthreads();
        carry += tot;
    }
    if (tid == 0) *nday = carry;
}

static __global__ void __launch_bounds__(DAY_T) k_dayscatter(const int *__restrict__ day, int n, const int *__restrict__ seg,
                                                         int *__restrict__ dlist)
{
    __shared__ int lds[DAY_T / 64];
    const int j0 = blockIdx.x * DAY_SEG + threadIdx.x;
    int run = seg[blockIdx.x];
    for (int it = 0; it < DAY_IT; ++it) {
        const int j = j0 + it * DAY_T;
        const int f = (j < n && day[j] != 0) ? 1 : 0;
        int tot;
        const int r = block_excl_scan256(f, lds, tot);
        if (f) dlist[run + r] = j;
        run += tot;
    }
}

// ---- LW band kernel (rtrnmc, McICA) -------------------------------------------------------------
constexpr double LW_A0[16] = {1.66, 1.55, 1.58, 1.66, 1.54, 1.454, 1.89, 1.33, 1.668, 1.66, 1.66, 1.66, 1.66, 1.66, 1.66, 1.66};
constexpr double LW_A1[16] = {0.00, 0.25, 0.22, 0.00, 0.13, 0.446, -0.10, 0.40, -0.006, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00};
constexpr double LW_A2[16] = {0.00, -12.0, -11.7, 0.00, -0.72, -0.243, 0.19, -0.062, 0.414, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00};

template <typename R> struct GasT {  // rtrnmc gas transmittance and source factor of one g, one layer
    R trng, atrgas, gasfac, odepth;
};
template <typename R> __device__ __forceinline__ GasT<R> lw_gas(const Tabs<R> *__restrict__ T, R sdif, R tau)
{
    GasT<R> q;
    R od = fmax((R)0, sdif * tau);
    if (od <= (R)0.06) {
        q.atrgas = od - (R)0.5 * od * od;
        q.trng = (R)1 - q.atrgas;
        q.gasfac = (R)0.166667 * od;
    } else {
        const R tblind = od / ((R)(1.0 / 0.278) + od);
        const int itgas = imin(NTBL, (int)((R)10000 * tblind + (R)0.5));
        if constexpr (sizeof(R) == 4) {
            const LutF l = lut_f32(itgas, true);
            q.trng = l.e, q.gasfac = l.tfn, od = l.tau;
        } else {
            q.trng = T->exp_tbl[itgas];
            q.gasfac = T->tfn_tbl[itgas];
            od = T->tau_tbl[itgas];
        }
        q.atrgas = (R)1 - q.trng;
    }
    q.odepth = od;
    return q;
}
template <typename R> __device__ __forceinline__ void lw_tot(const Tabs<R> *__restrict__ T, R odtot, R &totfac, R &atrtot)
{
    if (odtot < (R)0.06) {
        totfac = (R)0.166667 * odtot;
        atrtot = odtot - (R)0.5 * odtot * odtot;
    } else {
        const R tblind = odtot / ((R)(1.0 / 0.278) + odtot);
        const int ittot = imin(NTBL, (int)((R)10000 * tblind + (R)0.5));
        if constexpr (sizeof(R) == 4) {
            const LutF l = lut_f32(ittot, true);
            totfac = l.tfn, atrtot = (R)1 - l.e;
        } else {
            totfac = T->tfn_tbl[ittot];
            atrtot = (R)1 - T->exp_tbl[ittot];
        }
    }
}

// Learned gas optics of one tile (surrogate path): normalised network outputs Y[j * S + k * n + t] (BF16;
// S the padded row stride)
// (j < 504 outputs, surface-first layer k, tile column t), de-normalised as y * tstd[j] + tmean[j];
// flag[t] != 0 sends column t back to the table optics (inputs outside the training envelope);
// bands: bit b set when band b uses the network (LW bands at bits 0..15, SW bands at 16..29).
struct Surr {
    const WR_G __bf16 *Y;
    const WR_G float *tmean, *tstd;
    const WR_G int *flag;
    long long S;
    uint32_t bands;
    __device__ __forceinline__ int table(int t, int bit) const { return ((bands >> bit) & 1u) ? flag[t] : 1; }
};

__device__ __forceinline__ float surr_out(const Surr &sg, int j, size_t s)
{
    return (float)sg.Y[(size_t)j * sg.S + s] * sg.tstd[j] + sg.tmean[j];
}

template <typename R, int B, int NGL = LW_NG[B]>
__device__ __forceinline__ void surr_lw(const Surr &sg, size_t s, R *tau, R *fr, int g0 = 0)
{
#pragma unroll
    for (int jj = 0; jj < NGL; ++jj) {
        const int g = g0 + jj;
        tau[jj] = (R)__expf(surr_out(sg, LW_NS[B] + g, s));
        fr[jj] = (R)fminf(1.0f, fmaxf(0.0f, surr_out(sg, 140 + LW_NS[B] + g, s)));
    }
}

// the layer quantities the LW transfer needs besides the gas optics
template <typename R> __device__ __forceinline__ void lw_layer_cloud(const LayV &v, LwLay<R> &L)
{
    L.ct = v.ct;
    L.clwp = v.cwp * 1.5;
    L.ciwp = v.cip * 1.5;
    L.tavel = v.t;
    L.delp = v.pb - v.pt;
}

// ---- per-wave level cache of the band kernels -------------------------------------------------------
// The SW adding method keeps two reflectances per g-point and level from the bottom-up sweep for the
// top-down sweep (the FP32 kernels also keep each layer's R/T, and the FP32 LW kernel its upward
// transmittance and source).  They live in a slot of the context's pool, not in per-lane private
// arrays (which would need kilobytes of scratch per lane, backed by the runtime on every HIP queue):
// one slot per resident 64-lane workgroup (the kernels loop over column blocks, so the slot count
// bounds the memory), laid out [level][value][g][lane] so that every access is one coalesced 64-lane
// row.  Level-major order makes the offset of every (value, g) row within a level a compile-time
// constant; with the level innermost instead, the compiler kept one 64-bit address per row live
// across the level loops, and the SW kernels spilled up to 150 registers more (8 % slower).
// A slot holds nk = nlay + 2 levels (layers 0..nrl-1, or levels 0..nrl) of nv values per g-point.
template <typename R> struct LaneCache {
    WR_G R *p;  // this lane's element 0
    int nv;
    __device__ __forceinline__ R &at(int v, int ng, int g, int k) const
    {
        return p[(((size_t)k * nv + v) * ng + g) * 64];
    }
};
__host__ __device__ constexpr size_t cache_slot_bytes(int nv, int ng, int nk, int rbytes)
{
    return (size_t)64 * nv * ng * nk * rbytes;
}
template <typename R>
__device__ __forceinline__ LaneCache<R> lane_cache(char *pool, int nv, int ng, int nk)
{
    LaneCache<R> c;
    c.p = (WR_G R *)(pool + (size_t)blockIdx.x * cache_slot_bytes(nv, ng, nk, (int)sizeof(R))) + threadIdx.x;
    c.nv = nv;
    return c;
}
constexpr int SW_NV = 2;  // per level and g: zrupb, zrupd (reflectances of the medium below)
// The FP32 SW kernels also keep each layer's R/T (rb, rd, tb, td, ld) for the top-down sweep instead
// of recomputing optics and two-stream there (3 % faster); in FP64 the extra HBM traffic outweighs it.
template <typename R> constexpr int sw_nv_mc() { return sizeof(R) == 4 ? SW_NV + 5 : SW_NV; }

// Lanes per column of the LW band kernels: one by default.  WR_LW_GSPLIT=2 gives the FP32
// instantiations two (each lane takes half of the band's g-points, which halves the registers a lane
// holds; the two partial g-sums meet through a lane swap), but the per-layer column work then runs
// twice and it measured 9 % slower.  FP64 always uses one (sums over g in the reference's order).
#ifndef WR_LW_GSPLIT
#define WR_LW_GSPLIT 1
#endif
template <typename R> constexpr int lw_gsplit() { return sizeof(R) == 4 ? WR_LW_GSPLIT : 1; }

template <int GS, typename R> __device__ __forceinline__ R gsum(R s)
{
    if constexpr (GS == 2) return s + __shfl_xor(s, 1);
    return s;
}

// Optics chunks and wave bounds of the LW band kernels, per band.  These kernels wait on table gathers
// most of the time, so a second resident wave per SIMD pays wherever the registers allow it:
// * chunk: g-points per optics chunk (0 = the whole band at once).  A scheduling barrier between
//   chunks keeps the gathers of one chunk from being hoisted above the arithmetic of the previous
//   one, which bounds the registers the gathered values hold.  The arithmetic of every g-point is
//   unchanged, so the FP64 results stay bit-identical;
// * waves: waves per SIMD the compiler must fit.
// FP32 cached kernel (k_lwc): the two-key-species bands 3, 4, 5, 7, 9 spill 18-80 registers at two
// waves and still run 25-30 % faster.  The other bands reach two waves by their register count
// (bands 6, 8, 12 through the chunks); for them the bound itself made the schedule 15-20 % slower,
// and chunking bands 1, 2, 10, 11 only lowered the loads in flight.  The two-g-point bands 15 and 16
// run 11-19 % faster at a three-wave bound (6-10 spilled registers).
// FP64 recompute kernel (k_lw<double>): two-g-point chunks bring bands 1, 2, 6, 11 to two waves
// without spills (26-33 % faster; band 1 needs the bound as well since the column body became a
// function shared with the LDS-staged kernel); bands 8 and 12 gain 6-10 % from a two-wave bound
// with spills.
// The other bands stay at one wave unchunked: chunked at one wave they lost half their speed (fewer
// gathers in flight, no second wave to cover the waits).
// WR_LWC_WAVES / WR_LW64_WAVES > 0 or WR_LW_CHUNK / WR_LW64_CHUNK >= 0 override the tables for every
// band (tuning builds).
constexpr int LWC_WAVES[16] = {1, 1, 2, 2, 2, 1, 2, 1, 2, 1, 1, 1, 1, 1, 3, 3};
constexpr int LWC_CHUNK[16] = {0, 0, 4, 4, 4, 4, 4, 4, 4, 0, 0, 4, 0, 0, 0, 0};
constexpr int LW64_WAVES[16] = {2, 1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 2, 1, 1, 1, 1};
constexpr int LW64_CHUNK[16] = {2, 2, 0, 0, 0, 2, 0, 4, 0, 0, 2, 2, 0, 0, 0, 0};
#ifndef WR_LWC_WAVES
#define WR_LWC_WAVES 0
#endif
#ifndef WR_LW_CHUNK
#define WR_LW_CHUNK -1
#endif
#ifndef WR_LW64_WAVES
#define WR_LW64_WAVES 0
#endif
#ifndef WR_LW64_CHUNK
#define WR_LW64_CHUNK -1
#endif
template <int B> constexpr int lwc_waves() { return WR_LWC_WAVES > 0 ? WR_LWC_WAVES : LWC_WAVES[B]; }
template <typename R, int B> constexpr int lw_waves()
{
    return sizeof(R) == 8 ? (WR_LW64_WAVES > 0 ? WR_LW64_WAVES : LW64_WAVES[B]) : WR_LW_WAVES;
}
// the smallest divisor of ngl that is >= want (ngl when want <= 0)
constexpr int chunk_of(int ngl, int want)
{
    if (want <= 0) return ngl;
    for (int c = want; c < ngl; ++c)
        if (ngl % c == 0) return c;
    return ngl;
}
template <int B> constexpr int lw_chunk() { return chunk_of(LW_NG[B], WR_LW_CHUNK >= 0 ? WR_LW_CHUNK : LWC_CHUNK[B]); }
template <typename R, int B, int NGL> constexpr int lw_chunk_r()
{
    return sizeof(R) == 8 ? chunk_of(NGL, WR_LW64_CHUNK >= 0 ? WR_LW64_CHUNK : LW64_CHUNK[B]) : NGL;
}
// lw_optics of g-points g0 .. g0 + NGL - 1 in chunks of CH, from the band tables Bt
template <typename R, int B, int NGL, int CH, typename BT>
__device__ __forceinline__ void lw_optics_ch(const Tabs<R> *__restrict__ T, const BT &Bt, const LwLay<R> &L, R *tau, R *fr,
                                             int g0)
{
    if constexpr (CH >= NGL) {
        lw_optics_t<R, B, NGL>(T, Bt, L, tau, fr, g0);
    } else {
#pragma unroll
        for (int c = 0; c < NGL; c += CH) {
            lw_optics_t<R, B, CH>(T, Bt, L, tau + c, fr + c, g0 + c);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}
template <typename R, int B, int NGL, int CH>
__device__ __forceinline__ void lw_optics_ch(const Tabs<R> *__restrict__ T, const LwLay<R> &L, R *tau, R *fr, int g0)
{
    lw_optics_ch<R, B, NGL, CH>(T, T->lw[B], L, tau, fr, g0);
}

// One column of LW band B (GS lanes per column; gt = the lane's index over the tile), optics from the
// band tables Bt: down sweep, surface, up sweep with the optics recomputed.
template <typename R, int B, bool SURR, typename BT>
__device__ __forceinline__ void lw_column(const Tabs<R> *__restrict__ T, const BT &Bt, const ColIn &in, const RunPar &par,
                                          int col0, int n, const ColScal &cs, const AccT<R> &acc, const Surr &sg, int gt)
{
    constexpr int GS = lw_gsplit<R>();
    constexpr int NG = LW_NG[B];
    constexpr int NGL = NG / GS;
    static_assert(NG % GS == 0, "g-points split evenly over the lanes of a column");
    constexpr int NS = LW_NS[B];
    const int t = gt / GS;
    const int g0 = (gt % GS) * NGL;
    const bool lead = (gt % GS) == 0;  // writes the column's level fluxes
    if (t >= valid_cols(in, col0, n)) return;  // both lanes of a column leave together
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    const int lcf1 = cs.lcf1[t];
    const int table = SURR ? sg.table(t, B) : 1;  // 0: learned gas optics for this column
    const double pwv = cs.pwv[t];
    const R sdif = (R)((B == 0 || B == 3 || B == 9) ? 1.66 : fmin(1.80, fmax(1.50, LW_A0[B] + LW_A1[B] * exp_d(LW_A2[B] * pwv))));
    const double sfe = in.emis[i];
    const R semis = (R)((sfe > EPS && sfe <= 1.0) ? sfe : 1.0);
    const uint64_t col = in.col_id ? in.col_id[i] : (uint64_t)i;
    const uint64_t seed = in.seed_dev ? *in.seed_dev : par.seed;
    uint64_t mb[NGL];
    for (int jj = 0; jj < NGL; ++jj)
        mb[jj] = mcica_g(col, seed, 0, NS + g0 + jj, nrl, par.iovr, [&](int k) { return lcf1 ? rc.cfrac(nrl - 1 - k) : 0.0; });

    if (lead) acc.lw_dn[(size_t)nrl * n + t] = (R)0;  // no downward LW at the top
    // ---- down sweep, top to surface ----
    R rad[NGL];
#pragma unroll
    for (int jj = 0; jj < NGL; ++jj) rad[jj] = (R)0;
    // layer inputs: cur = layer kin, nxt = kin + 1 (loaded an iteration ahead), nn = kin + 2
    LayV cur = rc.lay(0), nxt = cur;
    if (nrl > 1) nxt = rc.lay(1);
    double tz_top = cs.tz[t], la_nx = cs.la[t], tz_nx = 0.0;
    if constexpr (sizeof(R) != 8) tz_nx = cs.tz[(size_t)n + t];
    for (int kin = 0; kin < nrl; ++kin) {
        const int k = nrl - 1 - kin;
        LayV nn = nxt;
        if (kin + 2 < nrl) nn = rc.lay(kin + 2);
        // this layer's log pressure and bottom level temperature: FP32 kernels load k_colprep's values an
        // iteration ahead; FP64 kernels evaluate them here by the same operations (measured faster there)
        double la_c, tz_bot;
        if constexpr (sizeof(R) == 8) {
            la_c = la_nx;
            if (kin + 1 < nrl) {
                la_nx = log_d(nxt.pl);
                tz_bot = tlvl_v(nxt.t, cur.t, log_d(cur.pb), la_nx, la_c);
            } else {
                tz_bot = rc.tskn();
            }
        } else {
            la_c = la_nx, tz_bot = tz_nx;
            if (kin + 1 < nrl) la_nx = cs.la[(size_t)(kin + 1) * n + t], tz_nx = cs.tz[(size_t)(kin + 2) * n + t];
        }
        LwLay<R> L;
        R tau[NGL], fr[NGL];
        if (SURR && !table) {
            lw_layer_cloud<R>(cur, L);
            surr_lw<R, B, NGL>(sg, (size_t)k * n + t, tau, fr, g0);
        } else {
            lw_layer<R>(T, cur, par, la_c, L);
            lw_optics_ch<R, B, NGL, lw_chunk_r<R, B, NGL>()>(T, Bt, L, tau, fr, g0);
        }
        const R blay = planck<R>(T, B, L.tavel);
        const R dplnkd = planck<R>(T, B, tz_bot) - blay;
        const R taucld = lw_taucld_l<R>(T, B, lcf1, L, rc, kin, par);
        const R taer = (R)rc.aer_lw_tau(B, kin);  // tautot = taug + tauaer
        R s = (R)0;
#pragma unroll
        for (int jj = 0; jj < NGL; ++jj) {
            const GasT<R> q = lw_gas<R>(T, sdif, tau[jj] + taer);
            const R plfrac = fr[jj];
            const R bbdgas = plfrac * (blay + dplnkd * q.gasfac);
            const R gassrcd = bbdgas * q.atrgas;
            const R clfm = ((mb[jj] >> k) & 1) ? (R)1 : (R)0;
            if (clfm >= (R)EPS) {
                const R odcld = sdif * taucld;
                const R efclrfr = (R)1 - ((R)1 - M<R>::exp(-odcld)) * clfm;
                R totfac, atrtot;
                lw_tot<R>(T, q.odepth + odcld, totfac, atrtot);
                const R bbdtot = plfrac * (blay + dplnkd * totfac);
                const R totsrcd = bbdtot * atrtot;
                rad[jj] = rad[jj] * q.trng * efclrfr + gassrcd + clfm * (totsrcd - gassrcd);
            } else {
                rad[jj] = rad[jj] * q.trng + gassrcd;
            }
            s = s + rad[jj];
        }
        s = gsum<GS>(s);
        if (lead) acc.lw_dn[(size_t)k * n + t] = s;
        tz_top = tz_bot;
        cur = nxt, nxt = nn;
    }
    // ---- up sweep, surface to top ----
    const R reflct = (R)1 - semis;
    double tz_bot = cs.tz[(size_t)nrl * n + t];
    // layer inputs: cur = layer kin, nxt = kin - 1 (loaded an iteration ahead), nn = kin - 2
    cur = rc.lay(nrl - 1), nxt = cur;
    if (nrl > 1) nxt = rc.lay(nrl - 2);
    if constexpr (sizeof(R) == 8)
        la_nx = log_d(cur.pl);
    else
        la_nx = cs.la[(size_t)(nrl - 1) * n + t], tz_nx = cs.tz[(size_t)(nrl - 1) * n + t];
    for (int kin = nrl - 1; kin >= 0; --kin) {
        const int k = nrl - 1 - kin;
        LayV nn = nxt;
        if (kin >= 2) nn = rc.lay(kin - 2);
        const double la_c = la_nx;
        if constexpr (sizeof(R) == 8) {
            if (kin > 0) {
                la_nx = log_d(nxt.pl);
                tz_top = tlvl_v(cur.t, nxt.t, log_d(cur.pt), la_c, la_nx);
            } else {
                tz_top = cur.t;
            }
        } else {
            tz_top = tz_nx;
            if (kin > 0) la_nx = cs.la[(size_t)(kin - 1) * n + t], tz_nx = cs.tz[(size_t)(kin - 1) * n + t];
        }
        LwLay<R> L;
        R tau[NGL], fr[NGL];
        if (SURR && !table) {
            lw_layer_cloud<R>(cur, L);
            surr_lw<R, B, NGL>(sg, (size_t)k * n + t, tau, fr, g0);
        } else {
            lw_layer<R>(T, cur, par, la_c, L);
            lw_optics_ch<R, B, NGL, lw_chunk_r<R, B, NGL>()>(T, Bt, L, tau, fr, g0);
        }
        const R blay = planck<R>(T, B, L.tavel);
        const R dplnku = planck<R>(T, B, tz_top) - blay;
        if (k == 0) {  // surface emission and reflection
            const R pk0 = planck<R>(T, B, rc.tskn());
            R s0 = (R)0;
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const R rad0 = semis * fr[jj] * pk0;
                rad[jj] = rad0 + reflct * rad[jj];
                s0 = s0 + rad[jj];
            }
            s0 = gsum<GS>(s0);
            if (lead) acc.lw_up[t] = s0;
        }
        const R taucld = lw_taucld_l<R>(T, B, lcf1, L, rc, kin, par);
        const R taer = (R)rc.aer_lw_tau(B, kin);  // tautot = taug + tauaer
        R s = (R)0;
#pragma unroll
        for (int jj = 0; jj < NGL; ++jj) {
            const GasT<R> q = lw_gas<R>(T, sdif, tau[jj] + taer);
            const R plfrac = fr[jj];
            const R bbugas = plfrac * (blay + dplnku * q.gasfac);
            const R gassrcu = bbugas * q.atrgas;
            const R clfm = ((mb[jj] >> k) & 1) ? (R)1 : (R)0;
            if (clfm > (R)EPS) {
                const R odcld = sdif * taucld;
                const R efclrfr = (R)1 - ((R)1 - M<R>::exp(-odcld)) * clfm;
                R totfac, atrtot;
                lw_tot<R>(T, q.odepth + odcld, totfac, atrtot);
                const R bbutot = plfrac * (blay + dplnku * totfac);
                const R totsrcu = bbutot * atrtot;
                rad[jj] = rad[jj] * q.trng * efclrfr + gassrcu + clfm * (totsrcu - gassrcu);
            } else {
                rad[jj] = rad[jj] * q.trng + gassrcu;
            }
            s = s + rad[jj];
        }
        s = gsum<GS>(s);
        if (lead) acc.lw_up[(size_t)(k + 1) * n + t] = s;
        tz_bot = tz_top;
        cur = nxt, nxt = nn;
    }
}

template <typename R, int B, bool SURR = false, bool EXT = true>
__global__ void __launch_bounds__(64, (lw_waves<R, B>())) k_lw(const Tabs<R> *__restrict__ T, ColIn in_, RunPar par, int col0,
                                           int n, ColScal cs, AccT<R> acc, Surr sg)
{
    const ColIn in = band_in<EXT>(in_);
    lw_column<R, B, SURR>(T, T->lw[B], in, par, col0, n, cs, acc, sg, blockIdx.x * blockDim.x + threadIdx.x);
}

// The band tables of an LDS image (the LwImg order), absb from PB: the image's when the image holds it,
// the global table otherwise.  A table the band does not have points at the image start and is never
// read.  The offsets are multiples of 4 elements (the & ~3 states it, so that 4 g-points of a row
// load as one 16-byte read).
template <typename R, typename PB>
__device__ __forceinline__ LwBandT<R, const WR_L R *, PB> lds_band(const WR_L R *lt, const LwImg<R> &im, const LwBand<R> &G,
                                                                   PB absb)
{
    LwBandT<R, const WR_L R *, PB> Bl;
    auto at = [&](int j) { return lt + (imax(0, im.off[j]) & ~3); };
    Bl.absa = at(0), Bl.absb = absb, Bl.selfref = at(2), Bl.forref = at(3), Bl.fracrefa = at(4), Bl.fracrefb = at(5);
    Bl.ka_m[0] = at(6), Bl.ka_m[1] = at(7), Bl.ka_m[2] = at(8);
    Bl.kb_m[0] = at(9), Bl.kb_m[1] = at(10), Bl.kb_m[2] = at(11);
    Bl.x1 = at(12), Bl.x2 = at(13);
    Bl.refrat_planck_a = G.refrat_planck_a, Bl.refrat_planck_b = G.refrat_planck_b;
    Bl.refrat_m_a = G.refrat_m_a, Bl.refrat_m_b = G.refrat_m_b, Bl.refrat_m_a3 = G.refrat_m_a3;
    return Bl;
}
// copies the image into the workgroup's LDS (16-byte words) and waits for the whole workgroup
template <typename R, typename I> __device__ __forceinline__ void lds_fill(float4 *dst, const I &im)
{
    const WR_G float4 *src = (const WR_G float4 *)im.base;
    const int n4 = (int)((size_t)im.n * sizeof(R) / 16);
    for (int q = threadIdx.x; q < n4; q += blockDim.x) dst[q] = src[q];
    __syncthreads();
}

// k_lw (FP64, one lane per column) with the band's tables other than absb in LDS (absb, the tables of
// the layers above laytrop, stays in global memory: the FP64 image of every band then fits in LDS).
// Persistent workgroups of 4 x lw_waves waves, as many as the CUs hold; each copies the image into LDS
// once and its waves then run 64-column blocks.  Same arithmetic as k_lw.
template <typename R, int B> constexpr int lwl64_wpg() { return 4 * lw_waves<R, B>(); }
template <int B, bool SURR = false, bool EXT = true>
__global__ void __launch_bounds__((64 * lwl64_wpg<double, B>()), (lw_waves<double, B>()))
    k_lwl64(const Tabs<double> *__restrict__ T, ColIn in_, RunPar par, int col0, int n, ColScal cs, AccT<double> acc, Surr sg)
{
    static_assert(lw_gsplit<double>() == 1, "one lane per column");
    extern __shared__ float4 lwl64_tab[];
    const ColIn in = band_in<EXT>(in_);
    const LwImg<double> &im = T->lw_img[B];
    lds_fill<double>(lwl64_tab, im);
    const auto Bl = lds_band<double>((const WR_L double *)lwl64_tab, im, T->lw[B], T->lw[B].absb);
    constexpr int WPG = lwl64_wpg<double, B>();
    const int lane = threadIdx.x % 64;
    for (int blk = blockIdx.x * WPG + threadIdx.x / 64; blk * 64 < n; blk += gridDim.x * WPG)
        lw_column<double, B, SURR>(T, Bl, in, par, col0, n, cs, acc, sg, blk * 64 + lane);
}

// FP32 LW band kernel with a cached upward sweep.  The down sweep also forms, per layer and g-point,
// the upward transmittance a = trng * efclrfr and source b (gas plus cloud term) and keeps them in the
// lane's pool slot (8 bytes per layer and g-point); the up sweep is then rad = rad * a + b, with no
// optics.  The FP64 kernels recompute the upward optics instead: their sums follow the reference's
// operation order, and twice the bytes would not pay.
constexpr int LW_NV_F32 = 2;
// the smallest divisor of the band's g-point count that is >= the chunk asked for
template <int B, bool SURR, typename BT>
__device__ __forceinline__ void lw_column_f32c(const Tabs<float> *__restrict__ T, const BT &Bt, const ColIn &in,
                                               const RunPar &par, int col0, int n, const ColScal &cs, const AccT<float> &acc,
                                               const Surr &sg, int t, const LaneCache<float> &cc)
{
    using R = float;
    constexpr int NG = LW_NG[B];
    constexpr int NS = LW_NS[B];
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    const int lcf1 = cs.lcf1[t];
    const int table = SURR ? sg.table(t, B) : 1;
    const double pwv = cs.pwv[t];
    const R sdif = (R)((B == 0 || B == 3 || B == 9) ? 1.66 : fmin(1.80, fmax(1.50, LW_A0[B] + LW_A1[B] * exp_d(LW_A2[B] * pwv))));
    const double sfe = in.emis[i];
    const R semis = (R)((sfe > EPS && sfe <= 1.0) ? sfe : 1.0);
    const R reflct = (R)1 - semis;
    const uint64_t col = in.col_id ? in.col_id[i] : (uint64_t)i;
    const uint64_t seed = in.seed_dev ? *in.seed_dev : par.seed;
    uint64_t mb[NG];
    for (int g = 0; g < NG; ++g)
        mb[g] = mcica_g(col, seed, 0, NS + g, nrl, par.iovr, [&](int k) { return lcf1 ? rc.cfrac(nrl - 1 - k) : 0.0; });

    acc.lw_dn[(size_t)nrl * n + t] = (R)0;  // no downward LW at the top
    R rad[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) rad[g] = (R)0;
    LayV cur = rc.lay(0);  // layer kin; the next one is loaded an iteration ahead
    double tz_top = cs.tz[t], la_nx = cs.la[t], tz_nx = cs.tz[(size_t)n + t];
    for (int kin = 0; kin < nrl; ++kin) {
        const int k = nrl - 1 - kin;
        const LayV nxt = (kin + 1 < nrl) ? rc.lay(kin + 1) : cur;
        // this layer's log pressure and bottom level temperature, loaded an iteration ahead
        const double la_c = la_nx, tz_bot = tz_nx;
        if (kin + 1 < nrl) la_nx = cs.la[(size_t)(kin + 1) * n + t], tz_nx = cs.tz[(size_t)(kin + 2) * n + t];
        LwLay<R> L;
        R tau[NG], fr[NG];
        if (SURR && !table) {
            lw_layer_cloud<R>(cur, L);
            surr_lw<R, B>(sg, (size_t)k * n + t, tau, fr);
        } else {
            lw_layer<R>(T, cur, par, la_c, L);
            lw_optics_ch<R, B, NG, lw_chunk<B>()>(T, Bt, L, tau, fr, 0);
        }
        const R blay = planck<R>(T, B, L.tavel);
        const R dplnkd = planck<R>(T, B, tz_bot) - blay;
        const R dplnku = planck<R>(T, B, tz_top) - blay;
        const R taucld = lw_taucld_l<R>(T, B, lcf1, L, rc, kin, par);
        const R taer = (R)rc.aer_lw_tau(B, kin);  // tautot = taug + tauaer
        R s = (R)0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const GasT<R> q = lw_gas<R>(T, sdif, tau[g] + taer);
            const R plfrac = fr[g];
            const R gassrcd = plfrac * (blay + dplnkd * q.gasfac) * q.atrgas;
            const R gassrcu = plfrac * (blay + dplnku * q.gasfac) * q.atrgas;
            const R clfm = ((mb[g] >> k) & 1) ? (R)1 : (R)0;
            R a = q.trng, b = gassrcu;
            if (clfm >= (R)EPS) {
                const R odcld = sdif * taucld;
                const R efclrfr = (R)1 - ((R)1 - M<R>::exp(-odcld)) * clfm;
                R totfac, atrtot;
                lw_tot<R>(T, q.odepth + odcld, totfac, atrtot);
                const R totsrcd = plfrac * (blay + dplnkd * totfac) * atrtot;
                const R totsrcu = plfrac * (blay + dplnku * totfac) * atrtot;
                rad[g] = rad[g] * q.trng * efclrfr + gassrcd + clfm * (totsrcd - gassrcd);
                a = q.trng * efclrfr;
                b = gassrcu + clfm * (totsrcu - gassrcu);
            } else {
                rad[g] = rad[g] * q.trng + gassrcd;
            }
            s = s + rad[g];
            cc.at(0, NG, g, k) = a;
            cc.at(1, NG, g, k) = b;
        }
        acc.lw_dn[(size_t)k * n + t] = s;
        if (k == 0) {  // surface emission and reflection start the upward sweep
            const R pk0 = planck<R>(T, B, rc.tskn());
            R s0 = (R)0;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                rad[g] = semis * fr[g] * pk0 + reflct * rad[g];
                s0 = s0 + rad[g];
            }
            acc.lw_up[t] = s0;
        }
        tz_top = tz_bot;
        cur = nxt;
    }
    for (int k = 0; k < nrl; ++k) {  // up sweep over the cached layers
        R s = (R)0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            rad[g] = rad[g] * cc.at(0, NG, g, k) + cc.at(1, NG, g, k);
            s = s + rad[g];
        }
        acc.lw_up[(size_t)(k + 1) * n + t] = s;
    }
}

template <int B, bool SURR = false, bool EXT = true>
__global__ void __launch_bounds__(64, lwc_waves<B>()) k_lwc(const Tabs<float> *__restrict__ T, ColIn in_, RunPar par,
                                                         int col0, int n, ColScal cs, AccT<float> acc, Surr sg,
                                                         char *__restrict__ pool)
{
    const ColIn in = band_in<EXT>(in_);
    const LaneCache<float> cc = lane_cache<float>(pool, LW_NV_F32, LW_NG[B], in.nlay + 2);
    for (int blk = blockIdx.x; blk * 64 < n; blk += gridDim.x) {
        const int t = blk * 64 + threadIdx.x;
        if (t < valid_cols(in, col0, n)) lw_column_f32c<B, SURR>(T, T->lw[B], in, par, col0, n, cs, acc, sg, t, cc);
    }
}

// k_lwc with the band's tables in LDS.  Persistent workgroups of one wave per SIMD times the band's
// wave bound (LWC_WAVES; 1 where the registers alone set the waves, so a CU then holds several); each
// copies the band's packed table image (T->lw_img[B], up to about 130 KB) into LDS once, and each wave
// then runs 64-column blocks with its own pool slot, so the optics' table reads are LDS reads instead
// of L2/HBM gathers.  Same arithmetic as k_lwc.
template <int B> constexpr int lwl_wpg() { return 4 * lwc_waves<B>(); }
template <int B, bool SURR = false, bool EXT = true>
__global__ void __launch_bounds__(64 * lwl_wpg<B>(), lwc_waves<B>())
    k_lwl(const Tabs<float> *__restrict__ T, ColIn in_, RunPar par, int col0, int n, ColScal cs, AccT<float> acc, Surr sg,
          char *__restrict__ pool)
{
    extern __shared__ float4 lwl_tab[];
    const ColIn in = band_in<EXT>(in_);
    const LwImg<float> &im = T->lw_img[B];
    lds_fill<float>(lwl_tab, im);
    const WR_L float *lt = (const WR_L float *)lwl_tab;
    const auto Bl = lds_band<float>(lt, im, T->lw[B], lt + (imax(0, im.off[1]) & ~3));
    constexpr int WPG = lwl_wpg<B>();
    const int lane = threadIdx.x % 64, slot = blockIdx.x * WPG + threadIdx.x / 64;
    LaneCache<float> cc;
    cc.p = (WR_G float *)(pool + (size_t)slot * cache_slot_bytes(LW_NV_F32, LW_NG[B], in.nlay + 2, 4)) + lane;
    cc.nv = LW_NV_F32;
    for (int blk = slot; blk * 64 < n; blk += gridDim.x * WPG) {
        const int t = blk * 64 + lane;
        if (t < valid_cols(in, col0, n)) lw_column_f32c<B, SURR>(T, Bl, in, par, col0, n, cs, acc, sg, t, cc);
    }
}

// ---- SW band kernel (McICA two-stream) -------------------------------------------------------------
// single-scattering albedo above which the mixed path evaluates the SW layer R/T in FP64
constexpr double W0_FP64 = 0.999999;

template <typename R, int B> struct SwLayerRT {
    R rb[SW_NG[B]], rd[SW_NG[B]], tb[SW_NG[B]], td[SW_NG[B]], ld[SW_NG[B]];
};

// Layer R/T of one g-point from its gas (tg) and Rayleigh (tr) optical depths with the aerosol's
// optical depth, single-scattering albedo and asymmetry (ta, sa, ya), plus the cloud
// (tau, tau*w, tau*w*g) when cld.  FP32 evaluates in FP64 where FP32 loses the answer:
// near-conservative scattering (w0 > 1 - 1e-6: 1 - w0 and gamma1 - gamma2 cancel; an FP32 pre-test
// with a 10x wider margin skips the FP64 test for most layers) and near the two-stream singularity
// k mu = 1 (flagged by the FP32 evaluation); the optical properties are then summed in FP64.
template <typename R>
__device__ __forceinline__ void layer_rt_eval(const Tabs<R> *__restrict__ T, R tr, R tg, R ta, R sa, R ya, bool cld, R ctau,
                                              R cssa, R casy, R cz, R snt, R &rb, R &rd, R &tb, R &td, R &ld)
{
    // clear-sky optical parameters with the aerosol (phys_rrtm_sw_spcvrtc.cu:323-329)
    R ztau0 = fmax((R)FTINY, tr + tg + ta);
    R zssa0 = tr + ta * sa;
    R zasy0 = ya * sa * ta;
    if (cld) {
        ztau0 = ztau0 + ctau;
        zssa0 = zssa0 + cssa;
        zasy0 = zasy0 + casy;
    }
    if constexpr (sizeof(R) < sizeof(double)) {
        const double ta0 = (double)ta, sa0 = (double)sa, ya0 = (double)ya;
        const double t0 = fmax(FTINY, (double)tr + (double)tg + ta0) + (cld ? (double)ctau : 0.0);
        const double s0 = (double)tr + ta0 * sa0 + (cld ? (double)cssa : 0.0);
        const double g0 = ya0 * sa0 * ta0 + (cld ? (double)casy : 0.0);
        bool fp64 = zssa0 > (R)0.99999 * ztau0 && s0 > W0_FP64 * t0;
        if (!fp64) layer_rt<R>(ztau0, zssa0, zasy0, cz, snt, T->sw_exp_tbl, rb, rd, tb, td, ld, &fp64);
        if (fp64) {
            double b, d, e, f, h;
            layer_rt<double>(t0, s0, g0, (double)cz, (double)snt, T->sw_exp_tbl64, b, d, e, f, h);
            rb = (R)b, rd = (R)d, tb = (R)e, td = (R)f, ld = (R)h;
        }
    } else {
        layer_rt<R>(ztau0, zssa0, zasy0, cz, snt, T->sw_exp_tbl, rb, rd, tb, td, ld);
    }
}

template <typename R, int B, bool SURR, typename BT>
__device__ __forceinline__ void sw_layer_rt(const Tabs<R> *__restrict__ T, const BT &Bt, const RadCol &rc, const RunPar &par, int k,
                                            int lt, double cf1, const uint64_t *mb, R cz, R snt, SwLayerRT<R, B> &o,
                                            const Surr &sg, size_t s, int table, double la)
{
    constexpr int NG = SW_NG[B];
    const int nrl = rc.nrl;
    const int kin = nrl - 1 - k;
    SwLay<R> L;
    R tg[NG], tr[NG];
    if (SURR && !table) {
        rc.cloud(kin, L.ct, L.cwp, L.cip);
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const R tau = (R)__expf(surr_out(sg, 280 + SW_NS[B] + g, s));
            const R ssa = (R)fminf(1.0f, fmaxf(0.0f, surr_out(sg, 392 + SW_NS[B] + g, s)));
            tr[g] = tau * ssa;
            tg[g] = tau - tr[g];
        }
    } else {
        sw_layer<R>(T, rc, par, kin, la, L);
        sw_optics_t<R, B>(T, Bt, L, k + 1, lt, tg, tr);
    }
    double ta, sa, ya;
    rc.aer_sw3(B, kin, ta, sa, ya);
    R ctau = (R)0, cssa = (R)1, casy = (R)0;
    if (cf1 > 0.0 && L.ct > FTINY) sw_cloud_l<R>(T, B, L, rc, kin, par, ctau, cssa, casy);
#pragma unroll
    for (int g = 0; g < NG; ++g)
        layer_rt_eval<R>(T, tr[g], tg[g], (R)ta, (R)sa, (R)ya, (mb[g] >> k) & 1, ctau, cssa, casy, cz, snt, o.rb[g],
                         o.rd[g], o.tb[g],
                         o.td[g], o.ld[g]);
}

// One day column of one SW band (spcvrtc, McICA): bottom-up adding pass (reflectances of each level
// to the lane cache), then the top-down pass with the layer R/T recomputed and the level fluxes.
template <typename R, int B, bool SURR, typename BT>
__device__ __forceinline__ void sw_column(const Tabs<R> *__restrict__ T, const BT &Bt, const ColIn &in, const RunPar &par, int col0,
                                          int n, const ColScal &cs, const R *__restrict__ sflux, const AccT<R> &acc,
                                          const Surr &sg, int t, const LaneCache<R> &cc)
{
    constexpr bool CACHE_RT = sw_nv_mc<R>() > SW_NV;
    constexpr int NG = SW_NG[B];
    constexpr int NS = SW_NS[B];
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl, nlp1 = nrl + 1;
    const double czd = in.cosz[i];
    const R cz = (R)czd, snt = (R)(1.0 / czd);
    const double ssolar = (par.solcon / S0) * czd;
    const double alb = in.alb[i];
    const R asfc = (R)((SW_IDXSFC[B] != 0) ? alb : 0.5 * (alb + alb));
    const double cf1 = cs.zcf1[t];
    const int lt = cs.sw_laytrop[t];
    const int table = SURR ? sg.table(t, 16 + B) : 1;
    const uint64_t col = in.col_id ? in.col_id[i] : (uint64_t)i;
    const uint64_t seed = in.seed_dev ? *in.seed_dev : par.seed;
    uint64_t mb[NG];
    for (int g = 0; g < NG; ++g) {
        mb[g] = mcica_g(col, seed, 1, NS + g, nrl, par.iovr, [&](int k) {
            const double ct = rc.cfrac(nrl - 1 - k);
            return (cf1 > 0.0) ? ((ct < FTINY) ? 0.0 : ct) : 0.0;
        });
    }
    // bottom-up: reflectance of the medium below each level (swflux sweep 1); zrupb/zrupd of level
    // L (1-based, surface L = 1) sit at cache entry L - 1, the latest level also in pb/pd
    R pb[NG], pd[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        pb[g] = asfc, pd[g] = asfc;
        cc.at(0, NG, g, 0) = asfc;
        cc.at(1, NG, g, 0) = asfc;
    }
    for (int k = 0; k < nrl; ++k) {
        SwLayerRT<R, B> o;
        sw_layer_rt<R, B, SURR>(T, Bt, rc, par, k, lt, cf1, mb, cz, snt, o, sg, (size_t)k * n + t, table,
                                cs.la[(size_t)(nrl - 1 - k) * n + t]);
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const R zden1 = (R)1 / ((R)1 - pd[g] * o.rd[g]);
            const R nb = o.rb[g] + (o.td[g] * ((o.tb[g] - o.ld[g]) * pd[g] + o.ld[g] * pb[g])) * zden1;
            const R nd = o.rd[g] + o.td[g] * o.td[g] * pd[g] * zden1;
            pb[g] = nb, pd[g] = nd;
            cc.at(0, NG, g, k + 1) = nb;
            cc.at(1, NG, g, k + 1) = nd;
            if constexpr (CACHE_RT) {
                cc.at(2, NG, g, k) = o.rb[g];
                cc.at(3, NG, g, k) = o.rd[g];
                cc.at(4, NG, g, k) = o.tb[g];
                cc.at(5, NG, g, k) = o.td[g];
                cc.at(6, NG, g, k) = o.ld[g];
            }
        }
    }
    // top-down: direct beam, diffuse transmission and reflection from above, level fluxes
    R tdbt[NG], tdn[NG], rdnd[NG], zsolar[NG];
    R su = (R)0, sd = (R)0;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        zsolar[g] = (R)ssolar * sflux[(size_t)(NS + g) * n + t];
        tdbt[g] = (R)1, tdn[g] = (R)1, rdnd[g] = (R)0;
        const R zden1 = (R)1 / ((R)1 - rdnd[g] * pd[g]);
        const R zfu = (tdbt[g] * pb[g] + (tdn[g] - tdbt[g]) * pd[g]) * zden1;
        const R zfd = tdbt[g] + (tdn[g] - tdbt[g] + tdbt[g] * pb[g] * rdnd[g]) * zden1;
        su = su + zsolar[g] * zfu;
        sd = sd + zsolar[g] * zfd;
    }
    acc.sw_up[(size_t)nrl * n + t] = su;
    acc.sw_dn[(size_t)nrl * n + t] = sd;
    acc.sw_clr0[t] = (R)0;  // clear-sky surface flux: NAM route only
    for (int s = nrl; s >= 1; --s) {
        SwLayerRT<R, B> o;
        if constexpr (CACHE_RT) {
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                o.rb[g] = cc.at(2, NG, g, s - 1), o.rd[g] = cc.at(3, NG, g, s - 1), o.tb[g] = cc.at(4, NG, g, s - 1);
                o.td[g] = cc.at(5, NG, g, s - 1), o.ld[g] = cc.at(6, NG, g, s - 1);
            }
        } else {
            sw_layer_rt<R, B, SURR>(T, Bt, rc, par, s - 1, lt, cf1, mb, cz, snt, o, sg, (size_t)(s - 1) * n + t, table,
                                    cs.la[(size_t)(nrl - s) * n + t]);
        }
        su = (R)0, sd = (R)0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const R tdbt_s = o.ld[g] * tdbt[g];
            R tdn_s, rdnd_s;
            if (s == nrl) {
                tdn_s = o.tb[g];
                rdnd_s = o.rd[g];
            } else {
                const R zden1 = (R)1 / ((R)1 - o.rd[g] * rdnd[g]);
                tdn_s = tdbt[g] * o.tb[g] + (o.td[g] * ((tdn[g] - tdbt[g]) + tdbt[g] * o.rb[g] * rdnd[g])) * zden1;
                rdnd_s = o.rd[g] + o.td[g] * o.td[g] * rdnd[g] * zden1;
            }
            tdbt[g] = tdbt_s, tdn[g] = tdn_s, rdnd[g] = rdnd_s;
            const R zb = cc.at(0, NG, g, s - 1), zd = cc.at(1, NG, g, s - 1);  // level s
            const R zden1 = (R)1 / ((R)1 - rdnd_s * zd);
            const R zfu = (tdbt_s * zb + (tdn_s - tdbt_s) * zd) * zden1;
            const R zfd = tdbt_s + (tdn_s - tdbt_s + tdbt_s * zb * rdnd_s) * zden1;
            su = su + zsolar[g] * zfu;
            sd = sd + zsolar[g] * zfd;
        }
        acc.sw_up[(size_t)(s - 1) * n + t] = su;
        acc.sw_dn[(size_t)(s - 1) * n + t] = sd;
    }
}

// Waves per SIMD the SW kernels are compiled for, per band, measured serialised on the 954x835x60
// grid:
// * FP64 (both routes): bands 16, 25, 26, 28 fit two waves with few spills (16-27 % faster each);
// * FP32 McICA route: two waves for the bands that spill at one; bands 16, 18, 19, 22, 24, 25, 26 fit
//   two waves by their register count, and the bound only slowed their schedule (8-14 %);
// * FP32 NAM route: two waves for every band but 22 (one wave 25-45 % slower).
// WR_SW64_WAVES / WR_SW32_WAVES > 0 override the tables for every band (tuning builds).
constexpr int SW64_WAVES[14] = {2, 1, 1, 1, 1, 1, 1, 1, 1, 2, 2, 1, 2, 1};
constexpr int SW32_WAVES[14] = {1, 2, 1, 1, 2, 2, 1, 2, 1, 1, 1, 2, 2, 2};
constexpr int SWR32_WAVES[14] = {2, 2, 2, 2, 2, 2, 1, 2, 2, 2, 2, 2, 2, 2};
#ifndef WR_SW64_WAVES
#define WR_SW64_WAVES 0
#endif
#ifndef WR_SW32_WAVES
#define WR_SW32_WAVES 0
#endif
template <typename R, int B, bool NAM = false> constexpr int sw_waves()
{
    if (sizeof(R) == 8) return WR_SW64_WAVES > 0 ? WR_SW64_WAVES : SW64_WAVES[B];
    return WR_SW32_WAVES > 0 ? WR_SW32_WAVES : (NAM ? SWR32_WAVES[B] : SW32_WAVES[B]);
}

// SW band kernel: blocks of 64 day-list entries, a grid no larger than the pool's slot count.
template <typename R, int B, bool SURR = false, bool EXT = true>
__global__ void __launch_bounds__(64, (sw_waves<R, B>())) k_sw(const Tabs<R> *__restrict__ T, ColIn in_, RunPar par, int col0,
                                           int n, ColScal cs, const R *__restrict__ sflux, AccT<R> acc, Surr sg,
                                           char *__restrict__ pool)
{
    const ColIn in = band_in<EXT>(in_);
    const LaneCache<R> cc = lane_cache<R>(pool, sw_nv_mc<R>(), SW_NG[B], in.nlay + 2);
    const int nday = *cs.nday;
    for (int blk = blockIdx.x; blk * 64 < nday; blk += gridDim.x) {
        const int tt = blk * 64 + threadIdx.x;
        if (tt < nday) sw_column<R, B, SURR>(T, T->sw[B], in, par, col0, n, cs, sflux, acc, sg, cs.dlist[tt], cc);
    }
}

// The SW band tables of an LDS image (the SwImg order); FP64 keeps absb in global memory.  A table the
// band does not have points at the image start and is never read; offsets are multiples of 4 elements.
template <typename R>
__device__ __forceinline__ SwBandT<R, const WR_L R *,
                                   typename std::conditional<sizeof(R) == 4, const WR_L R *, const WR_G R *>::type>
sw_lds_band(const WR_L R *lt, const SwImg<R> &im, const SwBand<R> &G)
{
    SwBandT<R, const WR_L R *, typename std::conditional<sizeof(R) == 4, const WR_L R *, const WR_G R *>::type> Bl;
    auto at = [&](int j) { return lt + (imax(0, im.off[j]) & ~3); };
    Bl.absa = at(0);
    if constexpr (sizeof(R) == 4)
        Bl.absb = at(1);
    else
        Bl.absb = G.absb;
    Bl.selfref = at(2), Bl.forref = at(3), Bl.rayl_v = at(4), Bl.rayla = at(5), Bl.raylb = at(6);
    Bl.abso3a = at(7), Bl.abso3b = at(8), Bl.absch4 = at(9), Bl.absh2o = at(10), Bl.absco2 = at(11);
    Bl.rayl = G.rayl, Bl.strrat = G.strrat;
    return Bl;
}

// k_sw with the band's tables in LDS (absb in global memory in FP64), as k_lwl / k_lwl64 do for LW:
// persistent workgroups of 4 x sw_waves waves, as many as the CUs hold at once; each copies the band's
// packed table image (T->sw_img[B]) into LDS once and each wave then runs blocks of 64 day-list
// entries with its own pool slot.  Same arithmetic as k_sw.
template <typename R, int B> constexpr int swl_wpg() { return 4 * sw_waves<R, B>(); }
template <typename R, int B, bool SURR = false, bool EXT = true>
__global__ void __launch_bounds__((64 * swl_wpg<R, B>()), (sw_waves<R, B>()))
    k_swl(const Tabs<R> *__restrict__ T, ColIn in_, RunPar par, int col0, int n, ColScal cs, const R *__restrict__ sflux,
          AccT<R> acc, Surr sg, char *__restrict__ pool)
{
    extern __shared__ float4 swl_tab[];
    const ColIn in = band_in<EXT>(in_);
    lds_fill<R>(swl_tab, T->sw_img[B]);
    const auto Bl = sw_lds_band<R>((const WR_L R *)swl_tab, T->sw_img[B], T->sw[B]);
    constexpr int WPG = swl_wpg<R, B>();
    const int lane = threadIdx.x % 64, slot = blockIdx.x * WPG + threadIdx.x / 64;
    LaneCache<R> cc;
    cc.p = (WR_G R *)(pool + (size_t)slot * cache_slot_bytes(sw_nv_mc<R>(), SW_NG[B], in.nlay + 2, (int)sizeof(R))) + lane;
    cc.nv = sw_nv_mc<R>();
    const int nday = *cs.nday;
    for (int blk = slot; blk * 64 < nday; blk += gridDim.x * WPG) {
        const int tt = blk * 64 + lane;
        if (tt < nday) sw_column<R, B, SURR>(T, Bl, in, par, col0, n, cs, sflux, acc, sg, cs.dlist[tt], cc);
    }
}

// ---- post: FP64 fluxes -> heating rates, top-first outputs -------------------------------------------
struct OutPtrs {
    WR_G double *hr_lw, *hr_sw, *flw_up, *flw_dn, *fsw_up, *fsw_dn, *fsw_dn_sfc_clr;
};
template <typename P>
__global__ void __launch_bounds__(256) k_post(ColIn in, int col0, int n, ColScal cs, BandAcc ba, OutPtrs out)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= valid_cols(in, col0, n)) return;
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl, nm = in.nlay;
    const size_t nc = (size_t)in.ncol;
    const double flxfac = 0.5 * FLUXFAC;
    const WR_G P *pl = (const WR_G P *)ba.base;
    // band sums in band order (the order the reference accumulates them).  Every plane is loaded and
    // the sum is selected per band, not branched on, so that the loads of a level are all in flight
    // at once; a band that is absent (single-band runs) keeps the sum unchanged, as skipping it would.
    auto lwsum = [&](int plane0, int k) {
        double v = 0.0;
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const double x = (double)pl[(size_t)(2 * b + plane0) * ba.P + (size_t)k * n + t];
            v = ((ba.lw_mask >> b) & 1) ? v + x : v;
        }
        return v;
    };
    auto swsum = [&](int plane0, int k) {
        double v = 0.0;
#pragma unroll
        for (int b = 0; b < 14; ++b) {
            const double x = (double)pl[(size_t)(32 + 2 * b + plane0) * ba.P + (size_t)k * n + t];
            v = ((ba.sw_mask >> b) & 1) ? v + x : v;
        }
        return v;
    };
    const double up0 = lwsum(0, 0), dn0 = lwsum(1, 0);
    double fprev = up0 * flxfac - dn0 * flxfac;
    out.flw_up[(size_t)nrl * nc + i] = up0 * flxfac;
    out.flw_dn[(size_t)nrl * nc + i] = dn0 * flxfac;
    for (int k = 1; k <= nrl; ++k) {
        const double fu = lwsum(0, k) * flxfac, fd = lwsum(1, k) * flxfac;
        out.flw_up[(size_t)(nrl - k) * nc + i] = fu;
        out.flw_dn[(size_t)(nrl - k) * nc + i] = fd;
        const double fnet = fu - fd;
        const int kin = nrl - k;  // surface-first layer k-1
        const double rfdelp = HEATFAC / (rc.plvl(kin + 1) - rc.plvl(kin));
        const double h = (fprev - fnet) * rfdelp;
        if (k - 1 < nm) out.hr_lw[(size_t)(nm - 1 - (k - 1)) * nc + i] = h;
        fprev = fnet;
    }
    if (cs.day[t]) {
        double fu0 = swsum(0, 0), fd0 = swsum(1, 0);
        out.fsw_up[(size_t)nrl * nc + i] = fu0;
        out.fsw_dn[(size_t)nrl * nc + i] = fd0;
        for (int k = 0; k < nrl; ++k) {
            const double fu1 = swsum(0, k + 1), fd1 = swsum(1, k + 1);
            out.fsw_up[(size_t)(nrl - 1 - k) * nc + i] = fu1;
            out.fsw_dn[(size_t)(nrl - 1 - k) * nc + i] = fd1;
            const int kin = nrl - 1 - k;
            const double rf = HEATFAC / (rc.plvl(kin + 1) - rc.plvl(kin));
            const double h = ((fd1 - fu1) - (fd0 - fu0)) * rf;
            if (k < nm) out.hr_sw[(size_t)(nm - 1 - k) * nc + i] = h;
            fu0 = fu1, fd0 = fd1;
        }
        if (out.fsw_dn_sfc_clr) {
            double v = 0.0;
#pragma unroll
            for (int b = 0; b < 14; ++b) {
                const double x = (double)pl[(size_t)60 * ba.P + (size_t)b * n + t];
                v = ((ba.sw_mask >> b) & 1) ? v + x : v;
            }
            out.fsw_dn_sfc_clr[i] = v;
        }
    } else {
        for (int L = 0; L <= nrl; ++L) out.fsw_up[(size_t)L * nc + i] = 0.0, out.fsw_dn[(size_t)L * nc + i] = 0.0;
        for (int k0 = 0; k0 < nm; ++k0) out.hr_sw[(size_t)k0 * nc + i] = 0.0;
        if (out.fsw_dn_sfc_clr) out.fsw_dn_sfc_clr[i] = 0.0;
    }
}

}  // namespace wr
