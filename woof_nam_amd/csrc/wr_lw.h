// wr_lw.h -- longwave per-layer physics and band optics of the MI355X RRTMG kernels.
//
// lw_layer:  lwrad_prep (phys_rrtm_lw_lwrad.cu:98-306) + setcoef (phys_rrtm_lw_setcoef.cu:82-244)
//            for one surface-first layer, from the raw column inputs.
// lw_optics: taugb01..16 (phys_rrtm_lw_taugbNN.cu) for the g-points of band B in one layer (all, or
//            a contiguous half when two lanes share a column).
// The operations are written in the order of the reference kernels; the FP64 instantiation is
// compiled without contraction, so it reproduces a sequential IEEE evaluation bit for bit.
#pragma once
#include "wr_dev.h"

namespace wr {

constexpr int LW_NG[16] = {10, 12, 16, 14, 16, 8, 12, 8, 12, 6, 8, 8, 4, 2, 2, 2};
constexpr int LW_NS[16] = {0, 10, 22, 38, 52, 68, 76, 88, 96, 108, 114, 122, 130, 134, 136, 138};

template <typename R> struct LwLay {
    int jp, jt, jt1, indself, indfor, indminor, lower;
    R fac00, fac01, fac10, fac11, selffac, selffrac, forfac, forfrac, minorfrac, scaleminor, scaleminorn2;
    R colamt[7], wx[4], coldry, colbrd, pavel;
    double tavel, delp, ct, clwp, ciwp;
};

// lwrad_prep + setcoef of surface-first layer k (top-first kin); la = log(plyr(kin)) already known.
template <typename R>
__device__ __forceinline__ void lw_layer(const Tabs<R> *__restrict__ T, const LayV &v, const RunPar &par, double la,
                                         LwLay<R> &L)
{
#pragma clang fp contract(off)
    const double tem1g = 100.0 * CON_G;
    const double tem2c = 1.0e-20 * 1.0e3 * CON_AVGD;
    const double pav = v.pl, tav = v.t;
    const double dlp = v.pb - v.pt;
    const double qv = v.q, ov = v.o3;
    const double h2ovmr = fmax(0.0, qv * AMDW / (1.0 - qv));
    const double o3vmr = fmax(0.0, ov * AMDO3);
    const double tem0 = (1.0 - h2ovmr) * CON_AMD + h2ovmr * CON_AMW;
    const double cdry = tem2c * dlp / (tem1g * tem0 * (1.0 + h2ovmr));
    const double temcol = 1.0e-12 * cdry;
    double ca[7];
    ca[0] = fmax(0.0, cdry * h2ovmr);
    ca[1] = fmax(temcol, cdry * par.vmr[0]);
    ca[2] = fmax(temcol, cdry * o3vmr);
    ca[3] = fmax(temcol, cdry * par.vmr[1]);
    ca[4] = fmax(temcol, cdry * par.vmr[2]);
    ca[5] = fmax(0.0, cdry * par.vmr[3]);
    ca[6] = fmax(0.0, cdry * par.vmr[4]);
    double summol = 0.0;
#pragma unroll
    for (int m = 1; m < 7; ++m) summol = summol + ca[m];
    const double colbrd = cdry - summol;
    const double h2o = ca[0] / fmax(cdry, 1e-30);  // lw_core.py:238
    L.ct = v.ct;
    L.clwp = v.cwp * 1.5;  // CPATHFAC4LW (np3d = 3), phys_rrtm_grrad_lw.cu:58
    L.ciwp = v.cip * 1.5;
    L.tavel = tav;
    L.delp = dlp;
    // setcoef
    const double plog = la;
    const int jp = imax(1, imin(58, (int)(36.0 - 5.0 * (plog + 0.04))));
    const int jp1 = jp + 1;
    const double fp = fmax(0.0, fmin(1.0, 5.0 * (T->preflog[jp - 1] - plog)));
    const double tem1t = (tav - T->tref[jp - 1]) / 15.0;
    const double tem2t = (tav - T->tref[jp1 - 1]) / 15.0;
    const int jt = imax(1, imin(4, (int)(3.0 + tem1t)));
    const int jt1 = imax(1, imin(4, (int)(3.0 + tem2t)));
    const double ft = fmax(-0.5, fmin(1.5, tem1t - (double)(jt - 3)));
    const double ft1 = fmax(-0.5, fmin(1.5, tem2t - (double)(jt1 - 3)));
    L.jp = jp, L.jt = jt, L.jt1 = jt1;
    const double tem1 = 1.0 - fp;
    L.fac10 = (R)(tem1 * ft);
    L.fac00 = (R)(tem1 * (1.0 - ft));
    L.fac11 = (R)(fp * ft1);
    L.fac01 = (R)(fp * (1.0 - ft1));
    const double forf = pav * STPFAC / (tav * (1.0 + h2o));
    const double selff = h2o * forf;
    L.scaleminor = (R)(pav / tav);
    L.scaleminorn2 = (R)((pav / tav) * (colbrd / (cdry + ca[0])));
    const double tem1m = (tav - 180.8) / 7.2;
    const int indmin = imin(18, imax(1, (int)tem1m));
    L.indminor = indmin;
    L.minorfrac = (R)(tem1m - (double)indmin);
    L.lower = plog > 4.56;
    if (L.lower) {
        const double tem1f = (332.0 - tav) / 36.0;
        const int indf = imin(2, imax(1, (int)tem1f));
        L.indfor = indf;
        L.forfrac = (R)(tem1f - (double)indf);
        const double tem1s = (tav - 188.0) / 7.2;
        const int inds = imin(9, imax(1, (int)tem1s - 7));
        L.indself = inds;
        L.selffrac = (R)(tem1s - (double)(inds + 7));
    } else {
        const double tem1f = (tav - 188.0) / 36.0;
        L.indfor = 3;
        L.forfrac = (R)(tem1f - 1.0);
        L.indself = 0;
        L.selffrac = (R)0;
    }
    L.selffac = (R)(ca[0] * selff);
    L.forfac = (R)(ca[0] * forf);
#pragma unroll
    for (int m = 0; m < 7; ++m) L.colamt[m] = (R)ca[m];
    L.wx[0] = (R)fmax(0.0, cdry * par.vmr[8]);
    L.wx[1] = (R)fmax(0.0, cdry * par.vmr[5]);
    L.wx[2] = (R)fmax(0.0, cdry * par.vmr[6]);
    L.wx[3] = (R)fmax(0.0, cdry * par.vmr[7]);
    L.coldry = (R)cdry;
    L.colbrd = (R)colbrd;
    L.pavel = (R)pav;
}

template <typename R>
__device__ __forceinline__ void lw_layer(const Tabs<R> *__restrict__ T, const RadCol &rc, const RunPar &par, int kin,
                                         double la, LwLay<R> &L)
{
    lw_layer<R>(T, rc.lay(kin), par, la, L);
}

// reference ratio chi_mls(a, lev) / chi_mls(b, lev) (setcoef rfrate, FP64)
__device__ __forceinline__ double chirat(const WR_G double *__restrict__ chi, int a, int b, int lev)
{
    return chi[(lev - 1) * 7 + (a - 1)] / chi[(lev - 1) * 7 + (b - 1)];
}

template <typename R> __device__ __forceinline__ R planck(const Tabs<R> *__restrict__ T, int b, double temp)
{
    const int ind = imin(180, imax(1, (int)(temp - 159.0)));
    const R fr = (R)(temp - (double)((int)temp));
    const WR_G R *tp = T->totplnk + b * 181;
    return T->delwave[b] * (tp[ind - 1] + fr * (tp[ind] - tp[ind - 1]));
}

// LW cloud optical depth of band b (phys_rrtm_lw_cldprop.cu:79, ilwcliq = 1, ilwcice = 1)
template <typename R>
__device__ __forceinline__ R lw_taucld(const Tabs<R> *__restrict__ T, int b, int lcf1, double ct, double cldliq,
                                       double cldice, double refliq, double refice, double cldran = 0.0)
{
    if (!lcf1 || !(ct > 0.0)) return (R)0;
    R tauliq = 0, tauice = 0;
    if (cldliq > 0.0) {
        const double factor = refliq - 1.5;
        const int index = imax(1, imin(57, (int)factor));
        const R fint = (R)(factor - (double)index);
        const R a1 = T->absliq1[b * 58 + (index - 1)];
        const R a2 = T->absliq1[b * 58 + index];
        tauliq = fmax((R)0, (R)cldliq * (a1 + fint * (a2 - a1)));
    }
    if (cldice > 0.0) {
        const R ri = (R)fmin(130.0, fmax(13.0, refice));
        const int ia0 = T->ipat[b] - 1;
        tauice = fmax((R)0, (R)cldice * (T->absice1[ia0 * 2 + 0] + T->absice1[ia0 * 2 + 1] / ri));
    }
    const R tauran = (R)(0.33e-3 * cldran);  // absrain (NCAR) x rain path
    return tauice + tauliq + tauran + (R)0;  // + tausnw (snow path 0)
}

// lw_taucld of radiation layer kin of column rc, with the layer's radii and rain path (path form)
// loaded only for a cloudy layer
template <typename R, typename L_T, typename RC_T, typename P_T>
__device__ __forceinline__ R lw_taucld_l(const Tabs<R> *__restrict__ T, int b, int lcf1, const L_T &L, const RC_T &rc,
                                         int kin, const P_T &par)
{
    if (!lcf1 || !(L.ct > 0.0)) return (R)0;
    double rl, ri, rp;
    rc.props(kin, par.re_liq, par.re_ice, rl, ri, rp);
    return lw_taucld<R>(T, b, 1, L.ct, L.clwp, L.ciwp, rl, ri, rp);
}

// ---- band optics ------------------------------------------------------------------------------
template <typename R> struct Sel {
    R fk0, fk1, fk2;
    int id[6];
};
template <typename R> __device__ __forceinline__ Sel<R> sel_one(R specparm, R fs, int ind)
{
    Sel<R> s;
    if (specparm < (R)0.125) {
        const R p = fs - (R)1;
        const R p4 = (p * p) * (p * p);
        s.fk0 = p4, s.fk1 = (R)1 - p - (R)2 * p4, s.fk2 = p + p4;
        s.id[0] = ind, s.id[1] = ind + 9, s.id[2] = ind + 1, s.id[3] = ind + 10, s.id[4] = ind + 2, s.id[5] = ind + 11;
    } else if (specparm > (R)0.875) {
        const R p = -fs;
        const R p4 = (p * p) * (p * p);
        s.fk0 = p4, s.fk1 = (R)1 - p - (R)2 * p4, s.fk2 = p + p4;
        s.id[0] = ind + 1, s.id[1] = ind + 10, s.id[2] = ind, s.id[3] = ind + 9, s.id[4] = ind - 1, s.id[5] = ind + 8;
    } else {
        s.fk0 = (R)1 - fs, s.fk1 = fs, s.fk2 = (R)0;
        s.id[0] = ind, s.id[1] = ind + 9, s.id[2] = ind + 1, s.id[3] = ind + 10, s.id[4] = ind, s.id[5] = ind;
    }
    return s;
}
template <typename R>
__device__ __forceinline__ void sel_joint(R sp0, R fs0, int i0, R sp1, R fs1, int i1, Sel<R> &s0, Sel<R> &s1)
{
    if (sp0 < (R)0.125 && sp1 < (R)0.125) {
        s0 = sel_one<R>((R)0, fs0, i0), s1 = sel_one<R>((R)0, fs1, i1);
    } else if (sp0 > (R)0.875 && sp1 > (R)0.875) {
        s0 = sel_one<R>((R)1, fs0, i0), s1 = sel_one<R>((R)1, fs1, i1);
    } else {
        s0 = sel_one<R>((R)0.5, fs0, i0), s1 = sel_one<R>((R)0.5, fs1, i1);
    }
}
template <typename R> struct Eta {
    R speccomb, specparm, fs;
    int js;
};
template <typename R> __device__ __forceinline__ Eta<R> eta(R a, R ratio, R b, R mult)
{
    Eta<R> e;
    e.speccomb = a + ratio * b;
    e.specparm = a / e.speccomb;
    const R specmult = mult * fmin(e.specparm, (R)ONEMINUS);
    e.js = 1 + (int)specmult;
    e.fs = specmult - (R)((int)specmult);
    return e;
}
#define T2(tab, ind, g, ng) (tab)[((ind)-1) * (ng) + (g)]
#define T3(tab, m, j, nj, g, ng) (tab)[(((m)-1) * (nj) + ((j)-1)) * (ng) + (g)]

template <typename R, typename P>
__device__ __forceinline__ R maj6(P __restrict__ tab, int ng, int g, const Sel<R> &s, R fa, R fb)
{
    return (s.fk0 * fa) * T2(tab, s.id[0], g, ng) + (s.fk0 * fb) * T2(tab, s.id[1], g, ng) +
           (s.fk1 * fa) * T2(tab, s.id[2], g, ng) + (s.fk1 * fb) * T2(tab, s.id[3], g, ng) +
           (s.fk2 * fa) * T2(tab, s.id[4], g, ng) + (s.fk2 * fb) * T2(tab, s.id[5], g, ng);
}
template <typename R, typename P>
__device__ __forceinline__ R maj4(P __restrict__ tab, int ng, int g, int i0, R fs, R fa, R fb)
{
    return (((R)1 - fs) * fa) * T2(tab, i0, g, ng) + (((R)1 - fs) * fb) * T2(tab, i0 + 5, g, ng) +
           (fs * fa) * T2(tab, i0 + 1, g, ng) + (fs * fb) * T2(tab, i0 + 6, g, ng);
}
template <typename R, typename P>
__device__ __forceinline__ R eta_minor(P __restrict__ tab, int nj, int ng, int g, int indm, const Eta<R> &em, R mf)
{
    const R m1 = T3(tab, indm, em.js, nj, g, ng) + em.fs * (T3(tab, indm, em.js + 1, nj, g, ng) - T3(tab, indm, em.js, nj, g, ng));
    const R m2 = T3(tab, indm + 1, em.js, nj, g, ng) +
                 em.fs * (T3(tab, indm + 1, em.js + 1, nj, g, ng) - T3(tab, indm + 1, em.js, nj, g, ng));
    return m1 + mf * (m2 - m1);
}

// g-points g0 .. g0 + NGL - 1 of band B from the band tables Bt (global or staged in LDS); tau/fr
// hold them from index 0
template <typename R, int B, int NGL, typename BT>
__device__ __forceinline__ void lw_optics_t(const Tabs<R> *__restrict__ T, const BT &Bt, const LwLay<R> &L, R *tau, R *fr,
                                            int g0)
{
    constexpr int ng = LW_NG[B];
    const int lower = L.lower;
    const int jp = L.jp, jt = L.jt, jt1 = L.jt1;
    const R f00 = L.fac00, f01 = L.fac01, f10 = L.fac10, f11 = L.fac11;
    const int inds = L.indself, indsp = inds + 1, indf = L.indfor, indfp = indf + 1;
    const int indm = L.indminor, indmp = indm + 1;
    const R sf = L.selffac, sfr = L.selffrac, ff = L.forfac, ffr = L.forfrac, mf = L.minorfrac;
    const R *ca = L.colamt;
    const WR_G double *chi = T->chi_mls;
#define SELF(g) (sf * (T2(Bt.selfref, inds, g, ng) + sfr * (T2(Bt.selfref, indsp, g, ng) - T2(Bt.selfref, inds, g, ng))))
#define FOR(g) (ff * (T2(Bt.forref, indf, g, ng) + ffr * (T2(Bt.forref, indfp, g, ng) - T2(Bt.forref, indf, g, ng))))
#define MIN1(tab, g) (T2(tab, indm, g, ng) + mf * (T2(tab, indmp, g, ng) - T2(tab, indm, g, ng)))
#define ABS4(tab, i0, i1, g) \
    (f00 * T2(tab, i0, g, ng) + f10 * T2(tab, (i0) + 1, g, ng) + f01 * T2(tab, i1, g, ng) + f11 * T2(tab, (i1) + 1, g, ng))
#define FRACA(e, g) (T2(Bt.fracrefa, (e).js, g, ng) + (e).fs * (T2(Bt.fracrefa, (e).js + 1, g, ng) - T2(Bt.fracrefa, (e).js, g, ng)))
#define FRACB(e, g) (T2(Bt.fracrefb, (e).js, g, ng) + (e).fs * (T2(Bt.fracrefb, (e).js + 1, g, ng) - T2(Bt.fracrefb, (e).js, g, ng)))
    const int l0 = (jp - 1) * 5 + (jt - 1), l1 = jp * 5 + (jt1 - 1);
    const int u0 = (jp - 13) * 5 + (jt - 1), u1 = (jp - 12) * 5 + (jt1 - 1);
    const int jpp = jp + 1;
    if constexpr (B == 0) {
        const R pp = L.pavel;
        const R scalen2 = L.colbrd * L.scaleminorn2;
        if (lower) {
            const R corradj = (pp < (R)250) ? ((R)1 - (R)0.15 * ((R)250 - pp) / (R)154.4) : (R)1;
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R taun2 = scalen2 * MIN1(Bt.ka_m[0], g);
                tau[jj] = corradj * (ca[0] * ABS4(Bt.absa, l0 + 1, l1 + 1, g) + SELF(g) + FOR(g) + taun2);
                fr[jj] = Bt.fracrefa[g];
            }
        } else {
            const R corradj = (R)1 - (R)0.15 * (pp / (R)95.6);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R taun2 = scalen2 * MIN1(Bt.kb_m[0], g);
                tau[jj] = corradj * (ca[0] * ABS4(Bt.absb, u0 + 1, u1 + 1, g) + FOR(g) + taun2);
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 1) {
        if (lower) {
            const R corradj = (R)1 - (R)0.05 * (L.pavel - (R)100) / (R)900;
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = corradj * (ca[0] * ABS4(Bt.absa, l0 + 1, l1 + 1, g) + SELF(g) + FOR(g));
                fr[jj] = Bt.fracrefa[g];
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[0] * ABS4(Bt.absb, u0 + 1, u1 + 1, g) + FOR(g);
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 2) {
        const R p = L.coldry * (R)chi[jp * 7 + 3];
        const R ratn2o = ca[3] / p;
        const R adjcoln2o = (ratn2o > (R)1.5) ? ((R)0.5 + M<R>::pow(ratn2o - (R)0.5, (R)0.65)) * p : ca[3];
        if (lower) {
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jp), ca[1], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jpp), ca[1], (R)8);
            const Eta<R> em = eta<R>(ca[0], Bt.refrat_m_a, ca[1], (R)8), ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[1], (R)8);
            const Sel<R> s0 = sel_one<R>(e0.specparm, e0.fs, l0 * 9 + e0.js);
            const Sel<R> s1 = sel_one<R>(e1.specparm, e1.fs, l1 * 9 + e1.js);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R absn2o = eta_minor<R>(Bt.ka_m[0], 9, ng, g, indm, em, mf);
                const R tmaj = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10);
                const R tmaj1 = e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11);
                tau[jj] = tmaj + tmaj1 + SELF(g) + FOR(g) + adjcoln2o * absn2o;
                fr[jj] = FRACA(ep, g);
            }
        } else {
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jp), ca[1], (R)4), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jpp), ca[1], (R)4);
            const Eta<R> em = eta<R>(ca[0], Bt.refrat_m_b, ca[1], (R)4), ep = eta<R>(ca[0], Bt.refrat_planck_b, ca[1], (R)4);
            const int i0 = u0 * 5 + e0.js, i1 = u1 * 5 + e1.js;
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R absn2o = eta_minor<R>(Bt.kb_m[0], 5, ng, g, indm, em, mf);
                const R tmaj = e0.speccomb * maj4<R>(Bt.absb, ng, g, i0, e0.fs, f00, f10);
                const R tmaj1 = e1.speccomb * maj4<R>(Bt.absb, ng, g, i1, e1.fs, f01, f11);
                tau[jj] = tmaj + tmaj1 + FOR(g) + adjcoln2o * absn2o;
                fr[jj] = FRACB(ep, g);
            }
        }
    } else if constexpr (B == 3) {
        if (lower) {
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jp), ca[1], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jpp), ca[1], (R)8);
            const Eta<R> ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[1], (R)8);
            const Sel<R> s0 = sel_one<R>(e0.specparm, e0.fs, l0 * 9 + e0.js);
            const Sel<R> s1 = sel_one<R>(e1.specparm, e1.fs, l1 * 9 + e1.js);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R tmaj = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10);
                const R tmaj1 = e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11);
                tau[jj] = tmaj + tmaj1 + SELF(g) + FOR(g);
                fr[jj] = FRACA(ep, g);
            }
        } else {
            const Eta<R> e0 = eta<R>(ca[2], (R)chirat(chi, 3, 2, jp), ca[1], (R)4), e1 = eta<R>(ca[2], (R)chirat(chi, 3, 2, jpp), ca[1], (R)4);
            const Eta<R> ep = eta<R>(ca[2], Bt.refrat_planck_b, ca[1], (R)4);
            const int i0 = u0 * 5 + e0.js, i1 = u1 * 5 + e1.js;
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                R tg = e0.speccomb * maj4<R>(Bt.absb, ng, g, i0, e0.fs, f00, f10) + e1.speccomb * maj4<R>(Bt.absb, ng, g, i1, e1.fs, f01, f11);
                if (g == 7) tg = tg * (R)0.92;
                if (g == 8) tg = tg * (R)0.88;
                if (g == 9) tg = tg * (R)1.07;
                if (g == 10) tg = tg * (R)1.1;
                if (g == 11) tg = tg * (R)0.99;
                if (g == 12) tg = tg * (R)0.88;
                if (g == 13) tg = tg * (R)0.943;
                tau[jj] = tg;
                fr[jj] = FRACB(ep, g);
            }
        }
    } else if constexpr (B == 4) {
        const R wxk = L.wx[0];
        if (lower) {
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jp), ca[1], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jpp), ca[1], (R)8);
            const Eta<R> em = eta<R>(ca[0], Bt.refrat_m_a, ca[1], (R)8), ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[1], (R)8);
            Sel<R> s0, s1;
            sel_joint<R>(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, s0, s1);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R abso3 = eta_minor<R>(Bt.ka_m[0], 9, ng, g, indm, em, mf);
                tau[jj] = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10) + e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11) +
                         SELF(g) + FOR(g) + abso3 * ca[2] + wxk * Bt.x1[g];
                fr[jj] = FRACA(ep, g);
            }
        } else {
            const Eta<R> e0 = eta<R>(ca[2], (R)chirat(chi, 3, 2, jp), ca[1], (R)4), e1 = eta<R>(ca[2], (R)chirat(chi, 3, 2, jpp), ca[1], (R)4);
            const Eta<R> ep = eta<R>(ca[2], Bt.refrat_planck_b, ca[1], (R)4);
            const int i0 = u0 * 5 + e0.js, i1 = u1 * 5 + e1.js;
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = e0.speccomb * maj4<R>(Bt.absb, ng, g, i0, e0.fs, f00, f10) + e1.speccomb * maj4<R>(Bt.absb, ng, g, i1, e1.fs, f01, f11) +
                         wxk * Bt.x1[g];
                fr[jj] = FRACB(ep, g);
            }
        }
    } else if constexpr (B == 5) {
        const R wx2 = L.wx[1], wx3 = L.wx[2];
        if (lower) {
            const R temp = L.coldry * (R)chi[jp * 7 + 1];
            const R ratco2 = ca[1] / temp;
            const R adjcolco2 = (ratco2 > (R)3) ? ((R)2 + M<R>::pow(ratco2 - (R)2, (R)0.77)) * temp : ca[1];
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[0] * ABS4(Bt.absa, l0 + 1, l1 + 1, g) + SELF(g) + FOR(g) + adjcolco2 * MIN1(Bt.ka_m[0], g) +
                         wx2 * Bt.x1[g] + wx3 * Bt.x2[g];
                fr[jj] = Bt.fracrefa[g];
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = wx2 * Bt.x1[g] + wx3 * Bt.x2[g];
                fr[jj] = Bt.fracrefa[g];
            }
        }
    } else if constexpr (B == 6) {
        const R temp = L.coldry * (R)chi[jp * 7 + 1];
        const R ratco2 = ca[1] / temp;
        if (lower) {
            const R adjcolco2 = (ratco2 > (R)3) ? ((R)3 + M<R>::pow(ratco2 - (R)3, (R)0.79)) * temp : ca[1];
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 3, jp), ca[2], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 3, jpp), ca[2], (R)8);
            const Eta<R> em = eta<R>(ca[0], Bt.refrat_m_a, ca[2], (R)8), ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[2], (R)8);
            Sel<R> s0, s1;
            sel_joint<R>(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, s0, s1);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R absco2 = eta_minor<R>(Bt.ka_m[0], 9, ng, g, indm, em, mf);
                tau[jj] = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10) + e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11) +
                         SELF(g) + FOR(g) + adjcolco2 * absco2;
                fr[jj] = FRACA(ep, g);
            }
        } else {
            const R adjcolco2 = (ratco2 > (R)3) ? ((R)2 + M<R>::pow(ratco2 - (R)2, (R)0.79)) * temp : ca[1];
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                R tg = ca[2] * ABS4(Bt.absb, u0 + 1, u1 + 1, g) + adjcolco2 * MIN1(Bt.kb_m[0], g);
                if (g == 5) tg = tg * (R)0.92;
                if (g == 6) tg = tg * (R)0.88;
                if (g == 7) tg = tg * (R)1.07;
                if (g == 8) tg = tg * (R)1.1;
                if (g == 9) tg = tg * (R)0.99;
                if (g == 10) tg = tg * (R)0.855;
                tau[jj] = tg;
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 7) {
        const R temp = L.coldry * (R)chi[jp * 7 + 1];
        const R ratco2 = ca[1] / temp;
        const R adjcolco2 = (ratco2 > (R)3) ? ((R)2 + M<R>::pow(ratco2 - (R)2, (R)0.65)) * temp : ca[1];
        const R w3 = L.wx[2], w4 = L.wx[3];
        if (lower) {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[0] * (f00 * Bt.absa[l0 * ng + g] + f10 * Bt.absa[(l0 + 1) * ng + g] + f01 * Bt.absa[l1 * ng + g] +
                                  f11 * Bt.absa[(l1 + 1) * ng + g]) +
                         SELF(g) + FOR(g) + adjcolco2 * MIN1(Bt.ka_m[0], g) + ca[2] * MIN1(Bt.ka_m[1], g) +
                         ca[3] * MIN1(Bt.ka_m[2], g) + w3 * Bt.x1[g] + w4 * Bt.x2[g];
                fr[jj] = Bt.fracrefa[g];
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[2] * (f00 * Bt.absb[u0 * ng + g] + f10 * Bt.absb[(u0 + 1) * ng + g] + f01 * Bt.absb[u1 * ng + g] +
                                  f11 * Bt.absb[(u1 + 1) * ng + g]) +
                         adjcolco2 * MIN1(Bt.kb_m[0], g) + ca[3] * MIN1(Bt.kb_m[2], g) + w3 * Bt.x1[g] + w4 * Bt.x2[g];
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 8) {
        const R temp = L.coldry * (R)chi[jp * 7 + 3];
        const R ratn2o = ca[3] / temp;
        const R adjcoln2o = (ratn2o > (R)1.5) ? ((R)0.5 + M<R>::pow(ratn2o - (R)0.5, (R)0.65)) * temp : ca[3];
        if (lower) {
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 6, jp), ca[4], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 6, jpp), ca[4], (R)8);
            const Eta<R> em = eta<R>(ca[0], Bt.refrat_m_a, ca[4], (R)8), ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[4], (R)8);
            Sel<R> s0, s1;
            sel_joint<R>(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, s0, s1);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R absn2o = eta_minor<R>(Bt.ka_m[0], 9, ng, g, indm, em, mf);
                tau[jj] = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10) + e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11) +
                         SELF(g) + FOR(g) + adjcoln2o * absn2o;
                fr[jj] = FRACA(ep, g);
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[4] * ABS4(Bt.absb, u0 + 1, u1 + 1, g) + adjcoln2o * MIN1(Bt.kb_m[0], g);
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 9) {
        if (lower) {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[0] * ABS4(Bt.absa, l0 + 1, l1 + 1, g) + SELF(g) + FOR(g);
                fr[jj] = Bt.fracrefa[g];
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[0] * ABS4(Bt.absb, u0 + 1, u1 + 1, g) + FOR(g);
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 10) {
        const R scaleo2 = ca[5] * L.scaleminor;
        if (lower) {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R tauo2 = scaleo2 * MIN1(Bt.ka_m[0], g);
                tau[jj] = ca[0] * ABS4(Bt.absa, l0 + 1, l1 + 1, g) + SELF(g) + FOR(g) + tauo2;
                fr[jj] = Bt.fracrefa[g];
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R tauo2 = scaleo2 * MIN1(Bt.kb_m[0], g);
                tau[jj] = ca[0] * ABS4(Bt.absb, u0 + 1, u1 + 1, g) + FOR(g) + tauo2;
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 11) {
        if (lower) {
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jp), ca[1], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 2, jpp), ca[1], (R)8);
            const Eta<R> ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[1], (R)8);
            Sel<R> s0, s1;
            sel_joint<R>(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, s0, s1);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10) + e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11) +
                         SELF(g) + FOR(g);
                fr[jj] = FRACA(ep, g);
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) tau[jj] = (R)0, fr[jj] = (R)0;
        }
    } else if constexpr (B == 12) {
        if (lower) {
            const R temp = L.coldry * (R)3.55e-4;
            const R ratco2 = ca[1] / temp;
            const R adjcolco2 = (ratco2 > (R)3) ? ((R)2 + M<R>::pow(ratco2 - (R)2, (R)0.68)) * temp : ca[1];
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 4, jp), ca[3], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 4, jpp), ca[3], (R)8);
            const Eta<R> emc2 = eta<R>(ca[0], Bt.refrat_m_a, ca[3], (R)8), emc = eta<R>(ca[0], Bt.refrat_m_a3, ca[3], (R)8);
            const Eta<R> ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[3], (R)8);
            Sel<R> s0, s1;
            sel_joint<R>(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, s0, s1);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R absco2 = eta_minor<R>(Bt.ka_m[0], 9, ng, g, indm, emc2, mf);
                const R absco = eta_minor<R>(Bt.ka_m[1], 9, ng, g, indm, emc, mf);
                tau[jj] = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10) + e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11) +
                         SELF(g) + FOR(g) + adjcolco2 * absco2 + ca[6] * absco;
                fr[jj] = FRACA(ep, g);
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[2] * MIN1(Bt.kb_m[0], g);
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 13) {
        if (lower) {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[1] * ABS4(Bt.absa, l0 + 1, l1 + 1, g) + SELF(g) + FOR(g);
                fr[jj] = Bt.fracrefa[g];
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[1] * ABS4(Bt.absb, u0 + 1, u1 + 1, g);
                fr[jj] = Bt.fracrefb[g];
            }
        }
    } else if constexpr (B == 14) {
        if (lower) {
            const R scalen2 = L.colbrd * L.scaleminor;
            const Eta<R> e0 = eta<R>(ca[3], (R)chirat(chi, 4, 2, jp), ca[1], (R)8), e1 = eta<R>(ca[3], (R)chirat(chi, 4, 2, jpp), ca[1], (R)8);
            const Eta<R> em = eta<R>(ca[3], Bt.refrat_m_a, ca[1], (R)8), ep = eta<R>(ca[3], Bt.refrat_planck_a, ca[1], (R)8);
            Sel<R> s0, s1;
            sel_joint<R>(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, s0, s1);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                const R taun2 = scalen2 * eta_minor<R>(Bt.ka_m[0], 9, ng, g, indm, em, mf);
                tau[jj] = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10) + e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11) +
                         SELF(g) + FOR(g) + taun2;
                fr[jj] = FRACA(ep, g);
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) tau[jj] = (R)0, fr[jj] = (R)0;
        }
    } else if constexpr (B == 15) {
        if (lower) {
            const Eta<R> e0 = eta<R>(ca[0], (R)chirat(chi, 1, 6, jp), ca[4], (R)8), e1 = eta<R>(ca[0], (R)chirat(chi, 1, 6, jpp), ca[4], (R)8);
            const Eta<R> ep = eta<R>(ca[0], Bt.refrat_planck_a, ca[4], (R)8);
            Sel<R> s0, s1;
            sel_joint<R>(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, s0, s1);
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = e0.speccomb * maj6<R>(Bt.absa, ng, g, s0, f00, f10) + e1.speccomb * maj6<R>(Bt.absa, ng, g, s1, f01, f11) +
                         SELF(g) + FOR(g);
                fr[jj] = FRACA(ep, g);
            }
        } else {
#pragma unroll
            for (int jj = 0; jj < NGL; ++jj) {
                const int g = g0 + jj;
                tau[jj] = ca[4] * (f00 * T2(Bt.absb, 1, g, ng) + f10 * T2(Bt.absb, 2, g, ng) + f01 * T2(Bt.absb, 1, g, ng) +
                                  f11 * T2(Bt.absb, 2, g, ng));
                fr[jj] = Bt.fracrefb[g];
            }
        }
    }
#undef SELF
#undef FOR
#undef MIN1
#undef ABS4
#undef FRACA
#undef FRACB
}
// the same from the band's tables in global memory (all g-points by default)
template <typename R, int B, int NGL = LW_NG[B]>
__device__ __forceinline__ void lw_optics(const Tabs<R> *__restrict__ T, const LwLay<R> &L, R *tau, R *fr, int g0 = 0)
{
    lw_optics_t<R, B, NGL>(T, T->lw[B], L, tau, fr, g0);
}

}  // namespace wr
