// wr_sw.h -- shortwave per-layer physics, band optics and two-stream layer solution.
//
// sw_layer:  swrad_prep (phys_rrtm_sw_swrad.cu:127-318) + SW setcoef (phys_rrtm_sw_setcoef.cu:60-130)
// sw_optics: taumol16..29 (phys_rrtm_sw_taumolNN.cu), the MINVAL(laytrop) group of bands
//            19/23/25/28 taken per column (one-column groups), so columns are independent.
// sw_cloud:  cldprop (phys_rrtm_sw_cldprop.cu:76), iswcliq = 1, iswcice = 3.
// layer_rt:  delta scaling, PIFM (iswmode 2) and swc_layer_rt (phys_rrtm_sw_spcvrtc.cu:112-173).
#pragma once
#include "wr_dev.h"

namespace wr {

constexpr int SW_NG[14] = {6, 12, 8, 8, 10, 10, 2, 10, 8, 6, 6, 8, 6, 12};
constexpr int SW_NS[14] = {0, 6, 18, 26, 34, 44, 54, 56, 66, 74, 80, 86, 94, 100};
constexpr int SW_NSPA[14] = {9, 9, 9, 9, 1, 9, 9, 1, 9, 1, 0, 1, 9, 1};
constexpr int SW_NSPB[14] = {1, 5, 1, 1, 1, 5, 1, 0, 1, 0, 0, 1, 5, 1};
constexpr int SW_IDXSFC[14] = {1, 1, 1, 1, 1, 1, 1, 1, 0, 2, 2, 2, 2, 1};

template <typename R> struct SwLay {
    int jp, jt, jt1, indself, indfor;
    R fac00, fac01, fac10, fac11, selffac, selffrac, forfac, forfrac;
    R colamt[6], colmol;
    double ct, cwp, cip, delp;
};

// swrad_prep + SW setcoef of surface-first layer (top-first kin), la = log(plyr(kin))
template <typename R>
__device__ __forceinline__ void sw_layer(const Tabs<R> *__restrict__ T, const RadCol &rc, const RunPar &par, int kin,
                                         double la, SwLay<R> &L)
{
#pragma clang fp contract(off)
    const double tem1 = 100.0 * CON_G;
    const double tem2 = 1.0e-20 * 1.0e3 * CON_AVGD;
    const double pav = rc.plyr(kin), tav = rc.tlyr(kin);
    const double dl = rc.plvl(kin + 1) - rc.plvl(kin);
    const double qv = rc.qlyr(kin), ov = rc.olyr(kin);
    const double h2o = fmax(0.0, qv * AMDW / (1.0 - qv));
    const double o3 = fmax(0.0, ov * AMDO3);
    const double tem0 = (1.0 - h2o) * CON_AMD + h2o * CON_AMW;
    const double cd = tem2 * dl / (tem1 * tem0 * (1.0 + h2o));
    const double tc = 1.0e-12 * cd;
    double ca[6];
    ca[0] = fmax(0.0, cd * h2o);
    ca[1] = fmax(tc, cd * par.vmr[0]);
    ca[2] = fmax(0.0, cd * o3);
    ca[3] = fmax(tc, cd * par.vmr[1]);
    ca[4] = fmax(tc, cd * par.vmr[2]);
    ca[5] = fmax(tc, cd * par.vmr[3]);
    L.colmol = (R)(cd + ca[0]);
#pragma unroll
    for (int m = 0; m < 6; ++m) L.colamt[m] = (R)ca[m];
    L.delp = dl;
    rc.cloud(kin, L.ct, L.cwp, L.cip);
    // setcoef
    const double forf = pav * STPFAC / (tav * (1.0 + h2o));
    L.forfac = (R)forf;
    const double plog = la;
    const int jp = imax(1, imin(58, (int)(36.0 - 5.0 * (plog + 0.04))));
    const int jp1 = jp + 1;
    const double fp = 5.0 * (T->sw_preflog[jp - 1] - plog);
    const double t1 = (tav - T->sw_tref[jp - 1]) / 15.0;
    const double t2 = (tav - T->sw_tref[jp1 - 1]) / 15.0;
    const int jt = imax(1, imin(4, (int)(3.0 + t1)));
    const int jt1 = imax(1, imin(4, (int)(3.0 + t2)));
    const double ft = t1 - (double)(jt - 3);
    const double ft1 = t2 - (double)(jt1 - 3);
    L.jp = jp, L.jt = jt, L.jt1 = jt1;
    const double fp1 = 1.0 - fp;
    L.fac10 = (R)(fp1 * ft);
    L.fac00 = (R)(fp1 * (1.0 - ft));
    L.fac11 = (R)(fp * ft1);
    L.fac01 = (R)(fp * (1.0 - ft1));
    if (plog > 4.56) {
        const double tem1f = (332.0 - tav) / 36.0;
        const int indf = imin(2, imax(1, (int)tem1f));
        L.indfor = indf;
        L.forfrac = (R)(tem1f - (double)indf);
        const double tem2s = (tav - 188.0) / 7.2;
        const int inds = imin(9, imax(1, (int)tem2s - 7));
        L.indself = inds;
        L.selffrac = (R)(tem2s - (double)(inds + 7));
        L.selffac = (R)(h2o * forf);
    } else {
        const double tem1f = (tav - 188.0) / 36.0;
        L.indfor = 3;
        L.forfrac = (R)(tem1f - 1.0);
        L.indself = 0;
        L.selffrac = (R)0;
        L.selffac = (R)0;
    }
}

#define S2(tab, ind, g, ng) (tab)[((ind)-1) * (ng) + (g)]

template <typename R> __device__ __forceinline__ void fac8(R fs, R f00, R f01, R f10, R f11, R *fc)
{
    const R fs1 = (R)1 - fs;
    fc[0] = fs1 * f00, fc[1] = fs * f00, fc[2] = fs1 * f10, fc[3] = fs * f10;
    fc[4] = fs1 * f01, fc[5] = fs * f01, fc[6] = fs1 * f11, fc[7] = fs * f11;
}
__device__ __forceinline__ void ind8(int id0, int id1, int js, int nrow, int d1, int d2, int *id)
{
    id[0] = imax(1, imin(nrow, id0 + js));
    id[1] = imin(nrow, id[0] + 1);
    id[2] = imin(nrow, id[0] + d1);
    id[3] = imin(nrow, id[0] + d2);
    id[4] = imax(1, imin(nrow, id1 + js));
    id[5] = imin(nrow, id[4] + 1);
    id[6] = imin(nrow, id[4] + d1);
    id[7] = imin(nrow, id[4] + d2);
}
template <typename R, typename P> __device__ __forceinline__ R sum8(P __restrict__ A, int ng, int g, const int *id, const R *fc)
{
    return fc[0] * S2(A, id[0], g, ng) + fc[1] * S2(A, id[1], g, ng) + fc[2] * S2(A, id[2], g, ng) +
           fc[3] * S2(A, id[3], g, ng) + fc[4] * S2(A, id[4], g, ng) + fc[5] * S2(A, id[5], g, ng) +
           fc[6] * S2(A, id[6], g, ng) + fc[7] * S2(A, id[7], g, ng);
}
template <typename R, typename P>
__device__ __forceinline__ R abs4(P __restrict__ A, int ng, int g, int id0, int id1, int nrow, R f00, R f01, R f10, R f11)
{
    const int i01 = imax(1, imin(nrow, id0 + 1)), i02 = imin(nrow, i01 + 1);
    const int i11 = imax(1, imin(nrow, id1 + 1)), i12 = imin(nrow, i11 + 1);
    return f00 * S2(A, i01, g, ng) + f10 * S2(A, i02, g, ng) + f01 * S2(A, i11, g, ng) + f11 * S2(A, i12, g, ng);
}
template <typename R> __device__ __forceinline__ int jsfs(R specmult, R &fs)
{
    const int jsi = (int)specmult;
    fs = specmult - (R)jsi;
    return 1 + jsi;
}

// gas (tg) and Rayleigh (tr) optical depths of band B (index 0..13 = bands 16..29), layer kk (1-based),
// from the band tables Bt (global or staged in LDS)
template <typename R, int B, typename BT>
__device__ __forceinline__ void sw_optics_t(const Tabs<R> *__restrict__ T, const BT &Bt, const SwLay<R> &L, int kk, int Lt,
                                            R *tg, R *tr)
{
    constexpr int ng = SW_NG[B];
    const int jp = L.jp, jt = L.jt, jt1 = L.jt1;
    const R f00 = L.fac00, f01 = L.fac01, f10 = L.fac10, f11 = L.fac11;
    const int inds = L.indself, indsp = inds + 1, indf = L.indfor, indfp = indf + 1;
    const R sf = L.selffac, sfr = L.selffrac, ff = L.forfac, ffr = L.forfrac;
    const R *ca = L.colamt;
    const R cm = L.colmol;
    int id0, id1;
    if (kk <= Lt) {
        id0 = ((jp - 1) * 5 + (jt - 1)) * SW_NSPA[B];
        id1 = (jp * 5 + (jt1 - 1)) * SW_NSPA[B];
    } else {
        id0 = ((jp - 13) * 5 + (jt - 1)) * SW_NSPB[B];
        id1 = ((jp - 12) * 5 + (jt1 - 1)) * SW_NSPB[B];
    }
#define SELFI(g) (S2(Bt.selfref, inds, g, ng) + sfr * (S2(Bt.selfref, indsp, g, ng) - S2(Bt.selfref, inds, g, ng)))
#define FORI(g) (S2(Bt.forref, indf, g, ng) + ffr * (S2(Bt.forref, indfp, g, ng) - S2(Bt.forref, indf, g, ng)))
    R fc[8];
    int id[8];
    R fs;
    if constexpr (B == 0) { /* band 16 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * Bt.rayl;
        if (kk <= Lt) {
            const R speccomb = ca[0] + Bt.strrat * ca[4];
            const int js = jsfs<R>((R)8 * fmin((R)ONEMINUS, ca[0] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        }
        if (kk >= Lt)
            for (int g = 0; g < ng; ++g) tg[g] = ca[4] * abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
    } else if constexpr (B == 1) { /* band 17 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * (R)6.86e-10;
        const R speccomb = ca[0] + Bt.strrat * ca[1];
        if (kk <= Lt) {
            const int js = jsfs<R>((R)8 * fmin((R)ONEMINUS, ca[0] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(((jp - 1) * 5 + (jt - 1)) * 9, (jp * 5 + (jt1 - 1)) * 9, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) {
                const R forpart = FORI(g);
                const R selfpart = (inds > 0) ? sf * SELFI(g) : (R)0;
                tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc) + ca[0] * (selfpart + ff * forpart);
            }
        } else {
            const int js = jsfs<R>((R)4 * fmin((R)ONEMINUS, ca[0] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(((jp - 13) * 5 + (jt - 1)) * 5, ((jp - 12) * 5 + (jt1 - 1)) * 5, js, 1175, 5, 6, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absb, ng, g, id, fc) + (ca[0] * ff) * FORI(g);
        }
    } else if constexpr (B == 2) { /* band 18 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * Bt.rayl;
        if (kk <= Lt) {
            const R speccomb = ca[0] + Bt.strrat * ca[4];
            const R ratio = (speccomb > (R)0) ? (ca[0] / speccomb) : (R)0;
            const int js = jsfs<R>((R)8 * fmin((R)ONEMINUS, ratio), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = ca[4] * abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
        }
    } else if constexpr (B == 3) { /* band 19 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * (R)2.29e-09;
        if (kk <= Lt) {
            const R speccomb = ca[0] + Bt.strrat * ca[1];
            const int js = jsfs<R>((R)8 * fmin((R)0.999999, ca[0] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = ca[1] * abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
        }
    } else if constexpr (B == 4) { /* band 20 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * Bt.rayl;
        if (kk <= Lt) {
            for (int g = 0; g < ng; ++g) {
                const R absterm = abs4<R>(Bt.absa, ng, g, id0, id1, 65, f00, f01, f10, f11);
                tg[g] = ca[0] * ((absterm) + sf * (SELFI(g)) + ff * (FORI(g))) + ca[4] * Bt.absch4[g];
            }
        } else {
            for (int g = 0; g < ng; ++g) {
                const R absterm = abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
                tg[g] = ca[0] * ((absterm) + ff * (FORI(g))) + ca[4] * Bt.absch4[g];
            }
        }
    } else if constexpr (B == 5) { /* band 21 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * (R)9.41e-09;
        const R speccomb = ca[0] + Bt.strrat * ca[1];
        if (kk < Lt) {
            const int js = jsfs<R>((R)8 * fmin((R)ONEMINUS, ca[0] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        }
        if (kk >= Lt) {
            const int js = jsfs<R>((R)4 * fmin((R)ONEMINUS, ca[0] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 1175, 5, 6, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absb, ng, g, id, fc) + ca[0] * ff * FORI(g);
        }
    } else if constexpr (B == 6) { /* band 22 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * (R)1.54e-08;
        const R o2cont = (R)(4.35e-4 / (350.0 * 2.0)) * ca[5];
        if (kk <= Lt) {
            const R speccomb = ca[0] + Bt.strrat * ca[5];
            const R ratio = (speccomb > (R)0) ? (ca[0] / speccomb) : (R)0;
            const int js = jsfs<R>((R)8 * fmin((R)ONEMINUS, ratio), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g)
                tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g)) + o2cont;
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = ca[5] * (R)1.6 * abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11) + o2cont;
        }
    } else if constexpr (B == 7) { /* band 23 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * Bt.rayl_v[g];
        if (kk <= Lt) {
            for (int g = 0; g < ng; ++g)
                tg[g] = ca[0] * ((R)1.029 * abs4<R>(Bt.absa, ng, g, id0, id1, 65, f00, f01, f10, f11) + sf * SELFI(g) + ff * FORI(g));
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = (R)0;
        }
    } else if constexpr (B == 8) { /* band 24 */
        if (kk <= Lt) {
            const R speccomb = ca[0] + Bt.strrat * ca[5];
            const int js = jsfs<R>((R)8 * fmin((R)ONEMINUS, ca[0] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) {
                tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc) + ca[2] * Bt.abso3a[g] + ca[0] * (sf * SELFI(g) + ff * FORI(g));
                const R ray0 = Bt.rayla[(js - 1) * 8 + g], ray1 = Bt.rayla[js * 8 + g];
                tr[g] = cm * (ray0 + fs * (ray1 - ray0));
            }
        } else {
            for (int g = 0; g < ng; ++g) {
                tg[g] = ca[5] * abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11) + ca[2] * Bt.abso3b[g];
                tr[g] = cm * Bt.raylb[g];
            }
        }
    } else if constexpr (B == 9) { /* band 25 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * Bt.rayl_v[g];
        if (kk <= Lt) {
            for (int g = 0; g < ng; ++g)
                tg[g] = ca[0] * abs4<R>(Bt.absa, ng, g, id0, id1, 65, f00, f01, f10, f11) + ca[2] * Bt.abso3a[g];
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = ca[2] * Bt.abso3b[g];
        }
    } else if constexpr (B == 10) { /* band 26 */
        for (int g = 0; g < ng; ++g) tg[g] = (R)0, tr[g] = cm * Bt.rayl_v[g];
    } else if constexpr (B == 11) { /* band 27 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * Bt.rayl_v[g];
        if (kk < Lt)
            for (int g = 0; g < ng; ++g) tg[g] = ca[2] * abs4<R>(Bt.absa, ng, g, id0, id1, 65, f00, f01, f10, f11);
        if (kk >= Lt)
            for (int g = 0; g < ng; ++g) tg[g] = ca[2] * abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
    } else if constexpr (B == 12) { /* band 28 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * Bt.rayl;
        const R speccomb = ca[2] + Bt.strrat * ca[5];
        if (kk <= Lt) {
            const int js = jsfs<R>((R)8 * fmin((R)ONEMINUS, ca[2] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absa, ng, g, id, fc);
        } else {
            const int js = jsfs<R>((R)4 * fmin((R)ONEMINUS, ca[2] / speccomb), fs);
            fac8<R>(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 1175, 5, 6, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8<R>(Bt.absb, ng, g, id, fc);
        }
    } else if constexpr (B == 13) { /* band 29 */
        for (int g = 0; g < ng; ++g) tr[g] = cm * (R)9.30e-11;
        if (kk <= Lt) {
            const int isp = imin(inds + 1, 10), ifp = imin(indf + 1, 4);
            for (int g = 0; g < ng; ++g) {
                const R selfv = S2(Bt.selfref, inds, g, ng) + sfr * (S2(Bt.selfref, isp, g, ng) - S2(Bt.selfref, inds, g, ng));
                const R forv = S2(Bt.forref, indf, g, ng) + ffr * (S2(Bt.forref, ifp, g, ng) - S2(Bt.forref, indf, g, ng));
                tg[g] = ca[0] * ((abs4<R>(Bt.absa, ng, g, id0, id1, 65, f00, f01, f10, f11)) + sf * selfv + ff * forv) + ca[1] * Bt.absco2[g];
            }
        } else {
            for (int g = 0; g < ng; ++g)
                tg[g] = ca[1] * abs4<R>(Bt.absb, ng, g, id0, id1, 235, f00, f01, f10, f11) + ca[0] * Bt.absh2o[g];
        }
    }
#undef SELFI
#undef FORI
}
// the same from the band's tables in global memory
template <typename R, int B>
__device__ __forceinline__ void sw_optics(const Tabs<R> *__restrict__ T, const SwLay<R> &L, int kk, int Lt, R *tg, R *tr)
{
    sw_optics_t<R, B>(T, T->sw[B], L, kk, Lt, tg, tr);
}

// cloud optical properties of band B for one layer (taucw, ssacw = tau*ssa, asycw = tau*ssa*g)
template <typename R>
__device__ __forceinline__ void sw_cloud(const Tabs<R> *__restrict__ T, int B, double cldliq, double refliq, double cldice,
                                         double refice, R &tau, R &ssa, R &asy, double cldran = 0.0)
{
    R tauliq = 0, ssaliq = 0, asyliq = 0, tauice = 0, ssaice = 0, asyice = 0;
    if (!(cldliq <= 0.0)) {
        const double factor = refliq - 1.5;
        const int index = imax(1, imin(57, (int)factor));
        const R fint = (R)(factor - (double)index);
        const WR_G R *e = T->extliq1 + B * 58, *s = T->ssaliq1 + B * 58, *a = T->asyliq1 + B * 58;
        const R extco = fmax((R)0, e[index - 1] + fint * (e[index] - e[index - 1]));
        const R ssaco = fmax((R)0, fmin((R)1, s[index - 1] + fint * (s[index] - s[index - 1])));
        const R asyco = fmax((R)0, fmin((R)1, a[index - 1] + fint * (a[index] - a[index - 1])));
        tauliq = (R)cldliq * extco;
        ssaliq = tauliq * ssaco;
        asyliq = ssaliq * asyco;
    }
    if (!(cldice <= 0.0)) {
        const double dgeice = fmax(5.0, fmin(140.0, 1.0315 * refice));
        const double factor = (dgeice - 2.0) / 3.0;
        const int index = imax(1, imin(45, (int)factor));
        const R fint = (R)(factor - (double)index);
        const WR_G R *e = T->extice3 + B * 46, *s = T->ssaice3 + B * 46, *a = T->asyice3 + B * 46;
        const R extco = fmax((R)0, e[index - 1] + fint * (e[index] - e[index - 1]));
        const R ssaco = fmax((R)0, fmin((R)1, s[index - 1] + fint * (s[index] - s[index - 1])));
        const R asyco = fmax((R)0, fmin((R)1, a[index - 1] + fint * (a[index] - a[index - 1])));
        tauice = (R)cldice * extco;
        ssaice = tauice * ssaco;
        asyice = ssaice * asyco;
    }
    // rain (a0r, b0r, c0r of radsw_datatb); snow path zero
    constexpr double B0R[14] = {0.466, 0.437, 0.416, 0.391, 0.374, 0.352, 0.183, 0.048, 0.012, 0.0, 0.0, 0.0, 0.0, 0.496};
    constexpr double C0R[14] = {0.975, 0.965, 0.960, 0.955, 0.952, 0.950, 0.944, 0.894, 0.884, 0.883, 0.883, 0.883, 0.883, 0.980};
    const double tauran = cldran * 3.07e-3;
    const double ssaran = tauran * (1.0 - B0R[B]);
    const double asyran = ssaran * C0R[B];
    tau = tauliq + tauice + (R)tauran + (R)0;
    ssa = ssaliq + ssaice + (R)ssaran + (R)0;
    asy = asyliq + asyice + (R)asyran + (R)0;
}

// sw_cloud of radiation layer kin of column rc with the layer's radii and rain path (path form)
template <typename R>
__device__ __forceinline__ void sw_cloud_l(const Tabs<R> *__restrict__ T, int B, const SwLay<R> &L, const RadCol &rc, int kin,
                                           const RunPar &par, R &tau, R &ssa, R &asy)
{
    double rl, ri, rp;
    rc.props(kin, par.re_liq, par.re_ice, rl, ri, rp);
    sw_cloud<R>(T, B, L.cwp, rl, L.cip, ri, tau, ssa, asy, rp);
}

template <typename R> __device__ __forceinline__ R expm(R zb, const WR_G R *__restrict__ tbl)
{
    if (zb <= (R)0.06) return (R)1 - zb + (R)0.5 * zb * zb;
    const R ftind = zb / ((R)(1.0 / 0.278) + zb);
    const int itind = imax(0, imin(10000, (int)((R)0.5 + (R)10000 * ftind)));
    if constexpr (sizeof(R) == 4) return lut_f32(itind, false).e;
    return tbl[itind];
}

// delta scaling + PIFM + layer R/T (zssaw < zcrit branch; the conservative branch is unreachable
// behind the 1 - 1e-6 clamp of zssaw, phys_rrtm_sw_spcvrtc.cu:54,117)
template <typename R>
__device__ __forceinline__ void layer_rt(R ztau0, R zssa0, R zasy0, R cz, R snt, const WR_G R *__restrict__ tbl, R &rb, R &rd,
                                         R &tb, R &td, R &zldbt, bool *ill = nullptr)
{
    const R zssaw = fmin((R)ONEMINUS, zssa0 / ztau0);
    const R zasyw = zasy0 / fmax((R)FTINY, zssa0);
    const R za1s = zasyw * zasyw;
    const R za2s = zssaw * za1s;
    const R ztau1 = ((R)1 - za2s) * ztau0;
    const R zssa1 = (zssaw - za2s) / ((R)1 - za2s);
    const R zasy1 = zasyw / ((R)1 + zasyw);
    const R zasy3 = (R)0.75 * zasy1;
    const R zgam1 = (R)2 - zssa1 * ((R)1.25 + zasy3);
    const R zgam2 = (R)0.75 * zssa1 * ((R)1 - zasy1);
    const R zgam3 = (R)0.5 - zasy3 * cz;
    const R zgam4 = (R)1 - zgam3;
    const R eps1 = (R)1.0e-8;
    const R za1 = zgam1 * zgam4 + zgam2 * zgam3;
    const R za2 = zgam1 * zgam3 + zgam2 * zgam4;
    const R zrk = sqrt((zgam1 - zgam2) * (zgam1 + zgam2));
    const R zrk2 = (R)2 * zrk;
    const R zrp = zrk * cz;
    const R zrp1 = (R)1 + zrp;
    const R zrm1 = (R)1 - zrp;
    R zrpp = (R)1 - zrp * zrp;
    // FP32 near the two-stream singularity k mu = 1: numerator and denominator of the direct-beam
    // terms both vanish there, so the caller re-evaluates the layer in FP64
    if (ill && fabs(zrpp) < (R)1.0e-2) {
        *ill = true;
        return;
    }
    if (fabs(zrpp) < eps1) zrpp = (zrpp < (R)0) ? -eps1 : eps1;
    const R zrkg1 = zrk + zgam1;
    const R zrkg3 = zrk * zgam3;
    const R zrkg4 = zrk * zgam4;
    const R zr1 = zrm1 * (za2 + zrkg3);
    const R zr2 = zrp1 * (za2 - zrkg3);
    const R zr3 = zrk2 * (zgam3 - za2 * cz);
    const R zr4 = zrpp * zrkg1;
    const R zr5 = zrpp * (zrk - zgam1);
    const R zt1 = zrp1 * (za1 + zrkg4);
    const R zt2 = zrm1 * (za1 - zrkg4);
    const R zt3 = zrk2 * (zgam4 + za1 * cz);
    // FP32: cap the exponents at 80 so exp(-zb) stays a normal number and 1/exp(-zb) finite (the
    // FP32 table underflows to 0 past about 103); the R/T expressions depend on the capped terms
    // only through factors below exp(-160)
    const R zcap = (sizeof(R) < sizeof(double)) ? (R)80 : (R)500;
    const R zb1 = fmin(zrk * ztau1, zcap);
    const R zexm1 = expm<R>(zb1, tbl);
    const R zexp1 = (R)1 / zexm1;
    const R zb2 = fmin(ztau1 * snt, zcap);
    const R zexm2 = expm<R>(zb2, tbl);
    const R zexp2 = (R)1 / zexm2;
    const R ze1r45 = zr4 * zexp1 + zr5 * zexm1;
    if (ze1r45 >= -eps1 && ze1r45 <= eps1) {
        rb = eps1;
        tb = zexm2;
    } else {
        const R zden1 = zssa1 / ze1r45;
        rb = fmax((R)0, fmin((R)1, (zr1 * zexp1 - zr2 * zexm1 - zr3 * zexm2) * zden1));
        tb = fmax((R)0, fmin((R)1, zexm2 * ((R)1 - (zt1 * zexp1 - zt2 * zexm1 - zt3 * zexp2) * zden1)));
    }
    const R zden1d = zr4 / (ze1r45 * zrkg1);
    rd = fmax((R)0, fmin((R)1, zgam2 * (zexp1 - zexm1) * zden1d));
    td = fmax((R)0, fmin((R)1, zrk2 * zden1d));
    zldbt = expm<R>(ztau1 * snt, tbl);
}

}  // namespace wr
