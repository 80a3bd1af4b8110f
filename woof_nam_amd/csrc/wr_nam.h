// wr_nam.h -- the NAM route (mcica = 0): LW transfer with maximum-random overlap of fractional
// clouds (rtrnmr, phys_rrtm_lw_rtrnmr_body.cuh:9-365) and the SW two-stream with cloud-fraction
// mixing and column-cover weighting (spcvrtc, phys_rrtm_sw_spcvrtc.cu:217-547).
#pragma once
#include "wr_kernels.h"

namespace wr {

// Overlap factors of one tile (rtrnmr:50-172): f[q][idx][t], idx 0..nl+2 as in the reference;
// q = 0 faccld1u, 1 faccld2u, 2 facclr1u, 3 facclr2u, 4 faccmb1u, 5 faccmb2u, 6..11 the same
// downward; lstu/lstd bit (idx - 1) set for idx = 1..nl.
template <typename R> struct Rtr {
    WR_G R *f;
    WR_G uint64_t *lstu, *lstd;
    int n;
    __device__ __forceinline__ R &at(int q, int idx, int t) const { return f[((size_t)q * (NL + 3) + idx) * n + t]; }
};
enum { FCLD1U = 0, FCLD2U, FCLR1U, FCLR2U, FCMB1U, FCMB2U, FCLD1D, FCLD2D, FCLR1D, FCLR2D, FCMB1D, FCMB2D };

// layer cloud fraction cldfrc(kk) of the LW route (1-based surface-first, cldfrc(0) = 1, (nl+1) = 0)
__device__ __forceinline__ double lw_cldfrc(const RadCol &rc, int kk)
{
    if (kk == 0) return 1.0;
    if (kk > rc.nrl) return 0.0;
    return rc.cfrac(rc.nrl - kk);
}

// The six factors of one index while the sweep that sets them is at it.
struct Fac6 {
    double cld1 = 0.0, cld2 = 0.0, clr1 = 0.0, clr2 = 0.0, cmb1 = 0.0, cmb2 = 0.0;
};

// Step k of a sweep sets the factors of the next index (k + 1 upward, k - 1 downward) and may clear
// clr2/cld2 of index k itself; no later step touches index k.  So the sweeps keep two indices in
// registers (cur = k, nxt) and three cloud fractions (cfm, cfk, cfp), and write each index as soon as
// its step is done: no per-lane arrays, no private memory.
template <typename R>
__global__ void __launch_bounds__(64) k_rtr_fac(ColIn in, int col0, int n, Rtr<R> rt)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= valid_cols(in, col0, n)) return;
    const RadCol rc{in, col0 + t, in.nlay + 1};
    const int nl = rc.nrl;
    const double eps = 1.0e-6;
    auto put = [&](int base, int idx, const Fac6 &f) {
        rt.at(base + 0, idx, t) = (R)f.cld1, rt.at(base + 1, idx, t) = (R)f.cld2, rt.at(base + 2, idx, t) = (R)f.clr1;
        rt.at(base + 3, idx, t) = (R)f.clr2, rt.at(base + 4, idx, t) = (R)f.cmb1, rt.at(base + 5, idx, t) = (R)f.cmb2;
    };
    // one step of either sweep: k's own fraction cfk, the one it moves to (cfn) and the one it came
    // from (cfo); cur = index k, nxt = the index the step sets
    auto step = [&](double cfo, double cfk, double cfn, bool lk, double &rat1, double &rat2, Fac6 &cur, Fac6 &nxt) {
        if (!(cfk > eps)) return;
        if (cfn >= cfk) {
            if (lk) {
                if (cfk < 1.0) nxt.clr2 = (cfn - cfk) / (1.0 - cfk);
                cur.clr2 = 0.0;
                cur.cld2 = 0.0;
            } else {
                const double fmx = fmax(cfk, cfo);
                if (cfn > fmx) {
                    nxt.clr1 = rat2;
                    nxt.clr2 = (cfn - fmx) / (1.0 - fmx);
                } else if (cfn < fmx) {
                    nxt.clr1 = (cfn - cfk) / (cfo - cfk);
                } else {
                    nxt.clr1 = rat2;
                }
            }
            if (nxt.clr1 > 0.0 || nxt.clr2 > 0.0) rat1 = 1.0, rat2 = 0.0;
            else rat1 = 0.0, rat2 = 0.0;
        } else {
            if (lk) {
                nxt.cld2 = (cfk - cfn) / cfk;
                cur.clr2 = 0.0;
                cur.cld2 = 0.0;
            } else {
                const double fmn = fmin(cfk, cfo);
                if (cfn <= fmn) {
                    nxt.cld1 = rat1;
                    nxt.cld2 = (fmn - cfn) / fmn;
                } else {
                    nxt.cld1 = (cfk - cfn) / (cfk - fmn);
                }
            }
            if (nxt.cld1 > 0.0 || nxt.cld2 > 0.0) rat1 = 0.0, rat2 = 1.0;
            else rat1 = 0.0, rat2 = 0.0;
        }
        nxt.cmb1 = nxt.clr1 * cur.cld2 * cfo;
        nxt.cmb2 = nxt.cld1 * cur.clr2 * (1.0 - cfo);
    };
    const Fac6 zero;
    // upward factors: index 0 stays zero, step k = 1..nl-1 sets index k + 1
    {
        uint64_t lst = 0;
        double cfm = lw_cldfrc(rc, 0), cfk = lw_cldfrc(rc, 1);
        bool lk = cfk > eps;
        if (lk) lst |= 1ull;
        double rat1 = 0.0, rat2 = 0.0;
        Fac6 cur;
        put(FCLD1U, 0, zero);
        for (int k = 1; k <= nl - 1; ++k) {
            const double cfp = lw_cldfrc(rc, k + 1);
            const bool lk1 = cfp > eps && cfk <= eps;
            if (lk1) lst |= 1ull << k;
            Fac6 nxt;
            step(cfm, cfk, cfp, lk, rat1, rat2, cur, nxt);
            put(FCLD1U, k, cur);
            cur = nxt;
            lk = lk1;
            cfm = cfk;
            cfk = cfp;
        }
        put(FCLD1U, nl, cur);
        put(FCLD1U, nl + 1, zero);
        put(FCLD1U, nl + 2, zero);
        rt.lstu[t] = lst;
    }
    // downward factors: indices nl+1, nl+2 stay zero, step k = nl..2 sets index k - 1
    {
        uint64_t lst = 0;
        double cfp = lw_cldfrc(rc, nl + 1), cfk = lw_cldfrc(rc, nl);
        bool lk = cfk > eps;
        if (lk) lst |= 1ull << (nl - 1);
        double rat1 = 0.0, rat2 = 0.0;
        Fac6 cur;
        put(FCLD1D, nl + 2, zero);
        put(FCLD1D, nl + 1, zero);
        for (int k = nl; k >= 2; --k) {
            const double cfm = lw_cldfrc(rc, k - 1);
            const bool lkm = cfm > eps && cfk <= eps;
            if (lkm) lst |= 1ull << (k - 2);
            Fac6 nxt;
            step(cfp, cfk, cfm, lk, rat1, rat2, cur, nxt);
            put(FCLD1D, k, cur);
            cur = nxt;
            lk = lkm;
            cfp = cfk;
            cfk = cfm;
        }
        put(FCLD1D, 1, cur);
        put(FCLD1D, 0, zero);
        rt.lstd[t] = lst;
    }
}

// total-sky transmittance, source factor and absorptance of odtot (rtrnmr's cloudy branch)
template <typename R>
__device__ __forceinline__ void lw_tot3(const Tabs<R> *__restrict__ T, R odtot, R &totfac, R &atrtot, R &trnt)
{
    if (odtot < (R)0.06) {
        totfac = (R)0.166667 * odtot;
        atrtot = odtot - (R)0.5 * odtot * odtot;
        trnt = (R)1 - atrtot;
    } else {
        const R tblind = odtot / ((R)(1.0 / 0.278) + odtot);
        const int ittot = imin(NTBL, (int)((R)10000 * tblind + (R)0.5));
        if constexpr (sizeof(R) == 4) {
            const LutF l = lut_f32(ittot, true);
            totfac = l.tfn, trnt = l.e;
        } else {
            totfac = T->tfn_tbl[ittot];
            trnt = T->exp_tbl[ittot];
        }
        atrtot = (R)1 - trnt;
    }
}

// ---- LW band kernel of the NAM route (rtrnmr) ---------------------------------------------------
// Per-band optics chunks (g-points, 0 = whole band) and wave bounds, by precision, as for the McICA
// route's LW kernels: a second resident wave covers the table-gather waits where the registers allow.
// Chosen per band from serialised timings on the 954x835x60 grid (chunks of 2 or 4, bound 1 or 2):
// FP32 bands 1, 2, 4, 7, 8, 9, 12 run 14-29 % faster at two waves, band 16 7 % faster at three; the
// FP32 bands 3 and 5 and most FP64 bands were fastest as they were (FP64 bands 6 and 11 gain 9 and
// 17 %).
// WR_LWR_WAVES > 0 or WR_LWR_CHUNK >= 0 override the tables for every band (tuning builds).
constexpr int LWR64_WAVES[16] = {1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1};
constexpr int LWR64_CHUNK[16] = {0, 0, 0, 0, 0, 2, 0, 0, 0, 0, 4, 0, 0, 0, 0, 0};
constexpr int LWR32_WAVES[16] = {2, 2, 1, 2, 1, 1, 2, 2, 2, 1, 1, 2, 1, 1, 1, 3};
constexpr int LWR32_CHUNK[16] = {4, 4, 0, 2, 0, 0, 2, 4, 2, 0, 0, 4, 0, 0, 0, 0};
#ifndef WR_LWR_WAVES
#define WR_LWR_WAVES 0
#endif
#ifndef WR_LWR_CHUNK
#define WR_LWR_CHUNK -1
#endif
template <typename R, int B> constexpr int lwr_waves()
{
    return WR_LWR_WAVES > 0 ? WR_LWR_WAVES : (sizeof(R) == 8 ? LWR64_WAVES[B] : LWR32_WAVES[B]);
}
template <typename R, int B> constexpr int lwr_chunk()
{
    return chunk_of(LW_NG[B], WR_LWR_CHUNK >= 0 ? WR_LWR_CHUNK : (sizeof(R) == 8 ? LWR64_CHUNK[B] : LWR32_CHUNK[B]));
}

// One column of LW band B on the NAM route (rtrnmr), optics from the band tables Bt.
template <typename R, int B, typename BT>
__device__ __forceinline__ void lwr_column(const Tabs<R> *__restrict__ T, const BT &Bt, const ColIn &in, const RunPar &par,
                                           int col0, int n, const ColScal &cs, const AccT<R> &acc, const Rtr<R> &rt, int t)
{
    constexpr int NG = LW_NG[B];
    if (t >= valid_cols(in, col0, n)) return;
    const int i = col0 + t;
    const RadCol rc{in, i, in.nlay + 1};
    const int nrl = rc.nrl;
    const int lcf1 = cs.lcf1[t];
    const double pwv = cs.pwv[t];
    const R sdif = (R)((B == 0 || B == 3 || B == 9) ? 1.66 : fmin(1.80, fmax(1.50, LW_A0[B] + LW_A1[B] * exp_d(LW_A2[B] * pwv))));
    const double sfe = in.emis[i];
    const R semis = (R)((sfe > EPS && sfe <= 1.0) ? sfe : 1.0);
    const uint64_t lstu = rt.lstu[t], lstd = rt.lstd[t];
    const R eps = (R)1.0e-6;

    acc.lw_dn[(size_t)nrl * n + t] = (R)0;  // no downward LW at the top
    // ---- down sweep, top to surface ----
    R radtotd[NG], totradd[NG], clrradd[NG], rad[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) radtotd[g] = totradd[g] = clrradd[g] = rad[g] = (R)0;
    // layer inputs: cur = layer kin, nxt = kin + 1 (loaded an iteration ahead), nn = kin + 2
    LayV cur = rc.lay(0), nxt = cur;
    if (nrl > 1) nxt = rc.lay(1);
    double tz_top = cs.tz[t], la_nx = cs.la[t], tz_nx = 0.0;
    if constexpr (sizeof(R) != 8) tz_nx = cs.tz[(size_t)n + t];
    for (int kin = 0; kin < nrl; ++kin) {
        const int kk = nrl - kin;  // 1-based surface-first layer
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
        lw_layer<R>(T, cur, par, la_c, L);
        R tau[NG], fr[NG];
        lw_optics_ch<R, B, NG, lwr_chunk<R, B>()>(T, Bt, L, tau, fr, 0);
        const R blay = planck<R>(T, B, L.tavel);
        const R dplnkd = planck<R>(T, B, tz_bot) - blay;
        const R taucld = lw_taucld_l<R>(T, B, lcf1, L, rc, kin, par);
        const R taer = (R)rc.aer_lw_tau(B, kin);  // tautot = taug + tauaer
        const R clfr = (R)lw_cldfrc(rc, kk);
        const bool top = (lstd >> (kk - 1)) & 1;
        const R c1 = rt.at(FCLR1D, kk - 1, t), d1 = rt.at(FCLD1D, kk - 1, t), m1 = rt.at(FCMB1D, kk - 1, t);
        const R m2 = rt.at(FCMB2D, kk - 1, t), c2 = rt.at(FCLR2D, kk - 1, t), d2 = rt.at(FCLD2D, kk - 1, t);
        R s = (R)0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const GasT<R> q = lw_gas<R>(T, sdif, tau[g] + taer);
            const R gassrcd = fr[g] * (blay + dplnkd * q.gasfac) * q.atrgas;
            if (top) {
                totradd[g] = clfr * radtotd[g];
                clrradd[g] = radtotd[g] - totradd[g];
                rad[g] = (R)0;
            }
            if (clfr >= eps) {
                const R odcld = sdif * taucld;
                R totfac, atrtot, trnt;
                lw_tot3<R>(T, q.odepth + odcld, totfac, atrtot, trnt);
                const R totsrcd = fr[g] * (blay + dplnkd * totfac) * atrtot;
                totradd[g] = totradd[g] * trnt + clfr * totsrcd;
                clrradd[g] = clrradd[g] * q.trng + ((R)1 - clfr) * gassrcd;
                radtotd[g] = totradd[g] + clrradd[g];
                s = s + radtotd[g];
                const R radmod = rad[g] * (c1 * q.trng + d1 * trnt) - m1 * gassrcd + m2 * totsrcd;
                rad[g] = -radmod + c2 * (clrradd[g] + radmod) - d2 * (totradd[g] - radmod);
                totradd[g] = totradd[g] + rad[g];
                clrradd[g] = clrradd[g] - rad[g];
            } else {
                radtotd[g] = radtotd[g] * q.trng + gassrcd;
                s = s + radtotd[g];
            }
        }
        acc.lw_dn[(size_t)(kk - 1) * n + t] = s;
        tz_top = tz_bot;
        cur = nxt, nxt = nn;
    }
    // ---- up sweep, surface to top ----
    const R reflct = (R)1 - semis;
    R radtotu[NG], totradu[NG], clrradu[NG];
    double tz_bot = cs.tz[(size_t)nrl * n + t];
    // layer inputs: cur = layer kin, nxt = kin - 1 (loaded an iteration ahead), nn = kin - 2
    cur = rc.lay(nrl - 1), nxt = cur;
    if (nrl > 1) nxt = rc.lay(nrl - 2);
    if constexpr (sizeof(R) == 8)
        la_nx = log_d(cur.pl);
    else
        la_nx = cs.la[(size_t)(nrl - 1) * n + t], tz_nx = cs.tz[(size_t)(nrl - 1) * n + t];
    for (int kin = nrl - 1; kin >= 0; --kin) {
        const int kk = nrl - kin;
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
        lw_layer<R>(T, cur, par, la_c, L);
        R tau[NG], fr[NG];
        lw_optics_ch<R, B, NG, lwr_chunk<R, B>()>(T, Bt, L, tau, fr, 0);
        const R blay = planck<R>(T, B, L.tavel);
        const R dplnku = planck<R>(T, B, tz_top) - blay;
        if (kk == 1) {  // surface emission and reflection
            const R pk0 = planck<R>(T, B, rc.tskn());
            R s0 = (R)0;
#pragma unroll
            for (int g = 0; g < NG; ++g) {
                const R rad0 = semis * fr[g] * pk0;
                radtotu[g] = rad0 + reflct * radtotd[g];
                totradu[g] = clrradu[g] = (R)0;
                rad[g] = (R)0;
                s0 = s0 + radtotu[g];
            }
            acc.lw_up[t] = s0;
        }
        const R taucld = lw_taucld_l<R>(T, B, lcf1, L, rc, kin, par);
        const R taer = (R)rc.aer_lw_tau(B, kin);  // tautot = taug + tauaer
        const R clfr = (R)lw_cldfrc(rc, kk);
        const bool bot = (lstu >> (kk - 1)) & 1;
        const R c1 = rt.at(FCLR1U, kk + 1, t), d1 = rt.at(FCLD1U, kk + 1, t), m1 = rt.at(FCMB1U, kk + 1, t);
        const R m2 = rt.at(FCMB2U, kk + 1, t), c2 = rt.at(FCLR2U, kk + 1, t), d2 = rt.at(FCLD2U, kk + 1, t);
        R s = (R)0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const GasT<R> q = lw_gas<R>(T, sdif, tau[g] + taer);
            const R gasu = fr[g] * (blay + dplnku * q.gasfac) * q.atrgas;
            if (bot) {
                totradu[g] = clfr * radtotu[g];
                clrradu[g] = radtotu[g] - totradu[g];
                rad[g] = (R)0;
            }
            if (clfr >= eps) {
                const R odcld = sdif * taucld;
                R totfac, atrtot, trnt;
                lw_tot3<R>(T, q.odepth + odcld, totfac, atrtot, trnt);
                const R totu = fr[g] * (blay + dplnku * totfac) * atrtot;
                totradu[g] = totradu[g] * trnt + clfr * totu;
                clrradu[g] = clrradu[g] * q.trng + ((R)1 - clfr) * gasu;
                radtotu[g] = totradu[g] + clrradu[g];
                s = s + radtotu[g];
                const R radmod = rad[g] * (c1 * q.trng + d1 * trnt) - m1 * gasu + m2 * totu;
                rad[g] = -radmod + c2 * (clrradu[g] + radmod) - d2 * (totradu[g] - radmod);
                totradu[g] = totradu[g] + rad[g];
                clrradu[g] = clrradu[g] - rad[g];
            } else {
                radtotu[g] = radtotu[g] * q.trng + gasu;
                s = s + radtotu[g];
            }
        }
        acc.lw_up[(size_t)kk * n + t] = s;
        tz_bot = tz_top;
        cur = nxt, nxt = nn;
    }
}

template <typename R, int B, bool EXT = true>
__global__ void __launch_bounds__(64, (lwr_waves<R, B>())) k_lwr(const Tabs<R> *__restrict__ T, ColIn in_, RunPar par, int col0,
                                            int n, ColScal cs, AccT<R> acc, Rtr<R> rt)
{
    const ColIn in = band_in<EXT>(in_);
    lwr_column<R, B>(T, T->lw[B], in, par, col0, n, cs, acc, rt, blockIdx.x * blockDim.x + threadIdx.x);
}

// k_lwr with the band's tables in LDS (the image of k_lwl / k_lwl64: FP64 keeps absb in global
// memory): persistent workgroups of 4 x lwr_waves waves, as many as the CUs hold at once, each
// copying the image into LDS once.  Same arithmetic as k_lwr.
template <typename R, int B> constexpr int lwrl_wpg() { return 4 * lwr_waves<R, B>(); }
template <typename R, int B, bool EXT = true>
__global__ void __launch_bounds__((64 * lwrl_wpg<R, B>()), (lwr_waves<R, B>()))
    k_lwrl(const Tabs<R> *__restrict__ T, ColIn in_, RunPar par, int col0, int n, ColScal cs, AccT<R> acc, Rtr<R> rt)
{
    extern __shared__ float4 lwrl_tab[];
    const ColIn in = band_in<EXT>(in_);
    const LwImg<R> &im = T->lw_img[B];
    lds_fill<R>(lwrl_tab, im);
    const WR_L R *lt = (const WR_L R *)lwrl_tab;
    constexpr int WPG = lwrl_wpg<R, B>();
    const int lane = threadIdx.x % 64;
    if constexpr (sizeof(R) == 4) {
        const auto Bl = lds_band<R>(lt, im, T->lw[B], lt + (imax(0, im.off[1]) & ~3));
        for (int blk = blockIdx.x * WPG + threadIdx.x / 64; blk * 64 < n; blk += gridDim.x * WPG)
            lwr_column<R, B>(T, Bl, in, par, col0, n, cs, acc, rt, blk * 64 + lane);
    } else {
        const auto Bl = lds_band<R>(lt, im, T->lw[B], T->lw[B].absb);
        for (int blk = blockIdx.x * WPG + threadIdx.x / 64; blk * 64 < n; blk += gridDim.x * WPG)
            lwr_column<R, B>(T, Bl, in, par, col0, n, cs, acc, rt, blk * 64 + lane);
    }
}

// ---- SW band kernel of the NAM route (spcvrtc with cldfrc mixing and zcf1/zcf0 weighting) -------
// R/T of layer k (surface-first): pass 0 clear sky; pass 1 total sky, the cloudy R/T mixed with the
// clear ones by the layer cloud fraction zc1 = cfrac / zcf1 (phys_rrtm_sw_spcvrtc.cu, cloudy block)
template <typename R, int B, typename BT>
__device__ __forceinline__ void swr_layer(const Tabs<R> *__restrict__ T, const BT &Bt, const RadCol &rc, const RunPar &par, int k,
                                          int lt, int pass, double cf1, R cz, R snt, SwLayerRT<R, B> &o, double la)
{
    constexpr int NG = SW_NG[B];
    const int kin = rc.nrl - 1 - k;
    SwLay<R> L;
    sw_layer<R>(T, rc, par, kin, la, L);
    R tg[NG], tr[NG];
    sw_optics_t<R, B>(T, Bt, L, k + 1, lt, tg, tr);
    double ta, sa, ya;
    rc.aer_sw3(B, kin, ta, sa, ya);
#pragma unroll
    for (int g = 0; g < NG; ++g)
        layer_rt_eval<R>(T, tr[g], tg[g], (R)ta, (R)sa, (R)ya, false, (R)0, (R)1, (R)0, cz, snt, o.rb[g], o.rd[g], o.tb[g],
                         o.td[g], o.ld[g]);
    if (pass == 0) return;
    const double zc1d = rc.cfrac(kin) / cf1;
    if (!(zc1d > FTINY)) return;
    const R zc1 = (R)zc1d, zc0 = (R)(1.0 - zc1d);
    R ctau = (R)0, cssa = (R)1, casy = (R)0;
    sw_cloud_l<R>(T, B, L, rc, kin, par, ctau, cssa, casy);
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        R rb, rd, tb, td, ld;
        layer_rt_eval<R>(T, tr[g], tg[g], (R)ta, (R)sa, (R)ya, true, ctau, cssa, casy, cz, snt, rb, rd, tb, td, ld);
        o.rb[g] = zc0 * o.rb[g] + zc1 * rb;
        o.rd[g] = zc0 * o.rd[g] + zc1 * rd;
        o.tb[g] = zc0 * o.tb[g] + zc1 * tb;
        o.td[g] = zc0 * o.td[g] + zc1 * td;
        o.ld[g] = zc0 * o.ld[g] + zc1 * ld;
    }
}

// One day column of one SW band on the NAM route: clear-sky pass, then (cloudy columns) the
// total-sky pass; each pass is the bottom-up adding sweep (level reflectances to the lane cache) and
// the top-down sweep with the layer R/T recomputed.
template <typename R, int B, typename BT>
__device__ __forceinline__ void swr_column(const Tabs<R> *__restrict__ T, const BT &Bt, const ColIn &in, const RunPar &par, int col0,
                                           int n, const ColScal &cs, const R *__restrict__ sflux, const AccT<R> &acc, int t,
                                           const LaneCache<R> &cc)
{
    constexpr bool CACHE_RT = sw_nv_mc<R>() > SW_NV;  // FP32: layer R/T kept for the top-down sweep
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
    const double cf1 = cs.zcf1[t], cf0 = cs.zcf0[t];
    const bool total = cf1 > EPS;
    const int lt = cs.sw_laytrop[t];
    R zsolar[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) zsolar[g] = (R)ssolar * sflux[(size_t)(NS + g) * n + t];

    for (int pass = 0; pass < (total ? 2 : 1); ++pass) {
        // bottom-up (swflux sweep 1); zrupb/zrupd of level L (surface L = 1) at cache entry L - 1,
        // the latest level also in pb/pd
        R pb[NG], pd[NG];
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            pb[g] = asfc, pd[g] = asfc;
            cc.at(0, NG, g, 0) = asfc;
            cc.at(1, NG, g, 0) = asfc;
        }
        for (int k = 0; k < nrl; ++k) {
            SwLayerRT<R, B> o;
            swr_layer<R, B>(T, Bt, rc, par, k, lt, pass, cf1, cz, snt, o, cs.la[(size_t)(nrl - 1 - k) * n + t]);
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
        // top-down and level fluxes (layer R/T recomputed); pass 0 leaves the clear-sky fluxes in the
        // band planes, which pass 1 reads back and mixes with the cloudy ones
        R tdbt[NG], tdn[NG], rdnd[NG];
        R su = (R)0, sd = (R)0;
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            tdbt[g] = (R)1, tdn[g] = (R)1, rdnd[g] = (R)0;
            const R zden1 = (R)1 / ((R)1 - rdnd[g] * pd[g]);
            const R zfu = (tdbt[g] * pb[g] + (tdn[g] - tdbt[g]) * pd[g]) * zden1;
            const R zfd = tdbt[g] + (tdn[g] - tdbt[g] + tdbt[g] * pb[g] * rdnd[g]) * zden1;
            su = su + zsolar[g] * zfu;
            sd = sd + zsolar[g] * zfd;
        }
        for (int s = nrl; s >= 0; --s) {
            if (s < nrl) {
                SwLayerRT<R, B> o;
                if constexpr (CACHE_RT) {
#pragma unroll
                    for (int g = 0; g < NG; ++g) {
                        o.rb[g] = cc.at(2, NG, g, s), o.rd[g] = cc.at(3, NG, g, s), o.tb[g] = cc.at(4, NG, g, s);
                        o.td[g] = cc.at(5, NG, g, s), o.ld[g] = cc.at(6, NG, g, s);
                    }
                } else {
                    swr_layer<R, B>(T, Bt, rc, par, s, lt, pass, cf1, cz, snt, o, cs.la[(size_t)(nrl - 1 - s) * n + t]);
                }
                su = (R)0, sd = (R)0;
#pragma unroll
                for (int g = 0; g < NG; ++g) {
                    const R tdbt_s = o.ld[g] * tdbt[g];
                    R tdn_s, rdnd_s;
                    if (s == nrl - 1) {
                        tdn_s = o.tb[g];
                        rdnd_s = o.rd[g];
                    } else {
                        const R zden1 = (R)1 / ((R)1 - o.rd[g] * rdnd[g]);
                        tdn_s = tdbt[g] * o.tb[g] + (o.td[g] * ((tdn[g] - tdbt[g]) + tdbt[g] * o.rb[g] * rdnd[g])) * zden1;
                        rdnd_s = o.rd[g] + o.td[g] * o.td[g] * rdnd[g] * zden1;
                    }
                    tdbt[g] = tdbt_s, tdn[g] = tdn_s, rdnd[g] = rdnd_s;
                    const R zb = cc.at(0, NG, g, s), zd = cc.at(1, NG, g, s);  // level s + 1
                    const R zden1 = (R)1 / ((R)1 - rdnd_s * zd);
                    const R zfu = (tdbt_s * zb + (tdn_s - tdbt_s) * zd) * zden1;
                    const R zfd = tdbt_s + (tdn_s - tdbt_s + tdbt_s * zb * rdnd_s) * zden1;
                    su = su + zsolar[g] * zfu;
                    sd = sd + zsolar[g] * zfd;
                }
            }
            const size_t o = (size_t)s * n + t;  // level s (surface-first)
            if (pass == 0 && s == 0) acc.sw_clr0[t] = sd;
            if (pass == 0) {
                acc.sw_up[o] = su;
                acc.sw_dn[o] = sd;
            } else {
                acc.sw_up[o] = (R)(cf1 * (double)su + cf0 * (double)acc.sw_up[o]);
                acc.sw_dn[o] = (R)(cf1 * (double)sd + cf0 * (double)acc.sw_dn[o]);
            }
        }
    }
}

// SW band kernel of the NAM route: blocks of 64 day-list entries looped over a pool-sized grid.
template <typename R, int B, bool EXT = true>
__global__ void __launch_bounds__(64, (sw_waves<R, B, true>())) k_swr(const Tabs<R> *__restrict__ T, ColIn in_, RunPar par, int col0,
                                            int n, ColScal cs, const R *__restrict__ sflux, AccT<R> acc,
                                            char *__restrict__ pool)
{
    const ColIn in = band_in<EXT>(in_);
    const LaneCache<R> cc = lane_cache<R>(pool, sw_nv_mc<R>(), SW_NG[B], in.nlay + 2);
    const int nday = *cs.nday;
    for (int blk = blockIdx.x; blk * 64 < nday; blk += gridDim.x) {
        const int tt = blk * 64 + threadIdx.x;
        if (tt < nday) swr_column<R, B>(T, T->sw[B], in, par, col0, n, cs, sflux, acc, cs.dlist[tt], cc);
    }
}

// k_swr with the band's tables in LDS (the images of k_swl): persistent workgroups of 4 x the band's
// waves, as many as the CUs hold at once, each wave with its own pool slot.  Same arithmetic as k_swr.
template <typename R, int B> constexpr int swrl_wpg() { return 4 * sw_waves<R, B, true>(); }
template <typename R, int B, bool EXT = true>
__global__ void __launch_bounds__((64 * swrl_wpg<R, B>()), (sw_waves<R, B, true>()))
    k_swrl(const Tabs<R> *__restrict__ T, ColIn in_, RunPar par, int col0, int n, ColScal cs, const R *__restrict__ sflux,
           AccT<R> acc, char *__restrict__ pool)
{
    extern __shared__ float4 swrl_tab[];
    const ColIn in = band_in<EXT>(in_);
    lds_fill<R>(swrl_tab, T->sw_img[B]);
    const auto Bl = sw_lds_band<R>((const WR_L R *)swrl_tab, T->sw_img[B], T->sw[B]);
    constexpr int WPG = swrl_wpg<R, B>();
    const int lane = threadIdx.x % 64, slot = blockIdx.x * WPG + threadIdx.x / 64;
    LaneCache<R> cc;
    cc.p = (WR_G R *)(pool + (size_t)slot * cache_slot_bytes(sw_nv_mc<R>(), SW_NG[B], in.nlay + 2, (int)sizeof(R))) + lane;
    cc.nv = sw_nv_mc<R>();
    const int nday = *cs.nday;
    for (int blk = slot; blk * 64 < nday; blk += gridDim.x * WPG) {
        const int tt = blk * 64 + lane;
        if (tt < nday) swr_column<R, B>(T, Bl, in, par, col0, n, cs, sflux, acc, cs.dlist[tt], cc);
    }
}

}  // namespace wr
