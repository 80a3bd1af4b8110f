/* oracle_sw.c -- CPU FP64 restatement of the reference's RRTMG shortwave path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP library; nothing in the product links it).
 * Layer index k is 0-based SURFACE-FIRST (swrad's internal order); kk = k + 1 is the Fortran
 * layer number and L the SW laytrop.  The MINVAL(laytrop) group coupling of taumol19/23/25/28
 * (sw_seam.py:28-37) is taken per column (group of one column), so every column is independent.
 */
#include "oracle.h"

/* ---- rswinit exp table (phys_rrtm_sw_rswinit.cu:310-345) ------------------------------------- */
void osw_table(double *exp_tbl)
{
    const double bpade = 1.0 / 0.278;
    exp_tbl[0] = 1.0;
    for (int i = 1; i < O_NTBL; ++i) {
        const double tfn = (double)i / (double)(O_NTBL - i);
        exp_tbl[i] = om_exp(-(bpade * tfn));
    }
    exp_tbl[O_NTBL] = 1.0e-20;
}

/* ---- swrad_prep (phys_rrtm_sw_swrad.cu:127-318) ----------------------------------------------- */
void osw_prep(const Rad *r, double alb, double cosz, double solcon, SwCol *c)
{
    const int nl = r->nrl;
    c->nlay = nl;
    const double s0fac = solcon / O_S0;
    c->cosz = cosz;
    c->sntz = 1.0 / cosz;
    c->ssolar = s0fac * cosz;
    c->albbm[0] = c->albbm[1] = c->albdf[0] = c->albdf[1] = alb;
    const double tem1 = 100.0 * O_CON_G;
    const double tem2 = 1.0e-20 * 1.0e3 * O_CON_AVGD;
    for (int k = 0; k < nl; ++k) {
        const int kk0 = nl - 1 - k;
        c->pavel[k] = r->plyr[kk0];
        c->tavel[k] = r->tlyr[kk0];
        for (int b = 0; b < 14; ++b) { /* aerosol optics (phys_rrtm_sw_swrad.cu:224-229) */
            c->tauae[b][k] = r->aersw[0][b][kk0];
            c->ssaae[b][k] = r->aersw[1][b][kk0];
            c->asyae[b][k] = r->aersw[2][b][kk0];
        }
        const double dl = r->plvl[kk0 + 1] - r->plvl[kk0];
        c->delp[k] = dl;
        const double qv = r->qlyr[kk0], ov = r->olyr[kk0];
        const double h2o = fmax(0.0, qv * O_AMDW / (1.0 - qv));
        const double o3 = fmax(0.0, ov * O_AMDO3);
        c->h2ovmr[k] = h2o;
        const double tem0 = (1.0 - h2o) * O_CON_AMD + h2o * O_CON_AMW;
        const double cd = tem2 * dl / (tem1 * tem0 * (1.0 + h2o));
        c->coldry[k] = cd;
        const double tc = 1.0e-12 * cd;
        c->colamt[k][0] = fmax(0.0, cd * h2o);
        c->colamt[k][1] = fmax(tc, cd * r->vmr[0]);
        c->colamt[k][2] = fmax(0.0, cd * o3);
        c->colmol[k] = cd + c->colamt[k][0];
        c->colamt[k][3] = fmax(tc, cd * r->vmr[1]);
        c->colamt[k][4] = fmax(tc, cd * r->vmr[2]);
        c->colamt[k][5] = fmax(tc, cd * r->vmr[3]);
        c->cfrac[k] = r->cld[kk0][0];
        c->cliqp[k] = r->cld[kk0][1];
        c->reliq[k] = r->cld[kk0][2];
        c->cicep[k] = r->cld[kk0][3];
        c->reice[k] = r->cld[kk0][4];
        c->cdat1[k] = r->cld[kk0][5];
        c->cdat3[k] = r->cld[kk0][7];
        c->cdat4[k] = r->cld[kk0][8];
    }
    /* iovrsw = 1: maximum-random column cloud cover */
    double z0 = 1.0, z1 = 1.0;
    for (int k = 0; k < nl; ++k) {
        const double cf = c->cfrac[k];
        if (cf > O_FTINY) {
            z1 = fmin(z1, 1.0 - cf);
        } else if (z1 < 1.0) {
            z0 = z0 * z1;
            z1 = 1.0;
        }
    }
    z0 = z0 * z1;
    if (z0 <= O_FTINY) z0 = 0.0;
    if (z0 > O_ONEMINUS) z0 = 1.0;
    c->zcf0 = z0;
    c->zcf1 = 1.0 - z0;
}

/* ---- SW setcoef (phys_rrtm_sw_setcoef.cu:60-130): fp, ft unclamped; laytrop = last lower layer */
void osw_setcoef(const OT *t, SwCol *c)
{
    const int nl = c->nlay;
    int lt = nl;
    for (int k = 0; k < nl; ++k) {
        const double pav = c->pavel[k], tav = c->tavel[k], h2o = c->h2ovmr[k];
        const double forf = pav * O_STPFAC / (tav * (1.0 + h2o));
        c->forfac[k] = forf;
        const double plog = om_log(pav);
        const int jp = imax(1, imin(58, (int)(36.0 - 5.0 * (plog + 0.04))));
        const int jp1 = jp + 1;
        const double fp = 5.0 * (t->sw_preflog[jp - 1] - plog);
        const double tem1 = (tav - t->sw_tref[jp - 1]) / 15.0;
        const double tem2 = (tav - t->sw_tref[jp1 - 1]) / 15.0;
        const int jt = imax(1, imin(4, (int)(3.0 + tem1)));
        const int jt1 = imax(1, imin(4, (int)(3.0 + tem2)));
        const double ft = tem1 - (double)(jt - 3);
        const double ft1 = tem2 - (double)(jt1 - 3);
        c->jp[k] = jp;
        c->jt[k] = jt;
        c->jt1[k] = jt1;
        const double fp1 = 1.0 - fp;
        c->fac10[k] = fp1 * ft;
        c->fac00[k] = fp1 * (1.0 - ft);
        c->fac11[k] = fp * ft1;
        c->fac01[k] = fp * (1.0 - ft1);
        if (plog > 4.56) {
            lt = k + 1;
            const double tem1f = (332.0 - tav) / 36.0;
            const int indf = imin(2, imax(1, (int)tem1f));
            c->indfor[k] = indf;
            c->forfrac[k] = tem1f - (double)indf;
            const double tem2s = (tav - 188.0) / 7.2;
            const int inds = imin(9, imax(1, (int)tem2s - 7));
            c->indself[k] = inds;
            c->selffrac[k] = tem2s - (double)(inds + 7);
            c->selffac[k] = h2o * forf;
        } else {
            const double tem1f = (tav - 188.0) / 36.0;
            c->indfor[k] = 3;
            c->forfrac[k] = tem1f - 1.0;
            c->indself[k] = 0;
            c->selffrac[k] = 0.0;
            c->selffac[k] = 0.0;
        }
    }
    c->laytrop = lt;
}

/* ---- solar source per g-point (phys_rrtm_sw_taumol.cu:59-120, rrtm_sw_bands.cuh) ------------- */
static const int SW_NGA[14] = {6, 12, 8, 8, 10, 10, 2, 10, 8, 6, 6, 8, 6, 12};
static const int SW_NSA[14] = {0, 6, 18, 26, 34, 44, 54, 56, 66, 74, 80, 86, 94, 100};
static const int SW_NSPA[14] = {9, 9, 9, 9, 1, 9, 9, 1, 9, 1, 0, 1, 9, 1};
static const int SW_NSPB[14] = {1, 5, 1, 1, 1, 5, 1, 0, 1, 0, 0, 1, 5, 1};

void osw_sfluxzen(const OT *t, SwCol *c)
{
    const int nl = c->nlay, lt = c->laytrop;
    for (int b = 0; b < 14; ++b) {
        const int jb = b + 16, ibd = t->ibx[b], njb = SW_NGA[b], ns = SW_NSA[b], lr = t->layreffr[b];
        if (jb == 16 || jb == 20 || jb == 23 || jb == 25 || jb == 26 || jb == 29) {
            for (int j = 0; j < njb; ++j) c->sfluxzen[ns + j] = t->sf01[(ibd - 1) * 16 + j];
        } else if (jb == 27) {
            for (int j = 0; j < njb; ++j) c->sfluxzen[ns + j] = t->scalekur * t->sf01[(ibd - 1) * 16 + j];
        } else {
            int ks;
            const double *tab;
            int mfs;
            if (jb == 17 || jb == 28) {
                ks = nl;
                for (int k = lt; k <= nl - 1; ++k)
                    if (c->jp[k - 1] < lr && c->jp[k] >= lr) {
                        ks = k + 1;
                        break;
                    }
                tab = t->sf02;
                mfs = 5;
            } else {
                ks = lt;
                for (int k = 1; k <= lt - 1; ++k)
                    if (c->jp[k - 1] < lr && c->jp[k] >= lr) {
                        ks = k + 1;
                        break;
                    }
                tab = t->sf03;
                mfs = 9;
            }
            const double colm1 = c->colamt[ks - 1][t->ix1[b] - 1];
            const double colm2 = c->colamt[ks - 1][t->ix2[b] - 1];
            const double speccomb = colm1 + t->sw_strrat[b] * colm2;
            const double specmult = t->sw_specwt[b] * fmin(0.999999, colm1 / speccomb);
            const int jsi = (int)specmult;
            const int js = 1 + jsi;
            const double fs = specmult - (double)jsi;
            const int base = ((ibd - 1) * mfs) * 16;
            for (int j = 0; j < njb; ++j) {
                const double a = tab[base + (js - 1) * 16 + j];
                const double cc = tab[base + js * 16 + j];
                c->sfluxzen[ns + j] = a + fs * (cc - a);
            }
        }
    }
}

/* ---- SW cldprop (phys_rrtm_sw_cldprop.cu:76), iswcliq = 1, iswcice = 3 (NAM) ------------------ */
void osw_cldprop(const OT *t, SwCol *c, int mcica)
{
    const int nl = c->nlay;
    static const double b0r[14] = {0.466, 0.437, 0.416, 0.391, 0.374, 0.352, 0.183, 0.048, 0.012, 0.0, 0.0, 0.0, 0.0, 0.496};
    static const double c0r[14] = {0.975, 0.965, 0.960, 0.955, 0.952, 0.950, 0.944, 0.894, 0.884, 0.883, 0.883, 0.883, 0.883, 0.980};
    const double a0r = 3.07e-3;
    if (!(c->zcf1 > 0.0)) {
        for (int k = 0; k < nl; ++k) {
            for (int b = 0; b < 14; ++b) c->taucw[b][k] = c->ssacw[b][k] = c->asycw[b][k] = 0.0;
            c->cldf[k] = 0.0;
            c->cldfrc[k] = 0.0;
        }
        return;
    }
    for (int k = 0; k < nl; ++k)
        for (int b = 0; b < 14; ++b) c->taucw[b][k] = 0.0, c->ssacw[b][k] = 1.0, c->asycw[b][k] = 0.0;
    for (int k = 0; k < nl; ++k) {
        const double cf = c->cfrac[k];
        if (!(cf > O_FTINY)) continue;
        /* rain and snow: snow path is zero on this path (tausnw = 0) */
        const double tauran = c->cdat1[k] * a0r;
        double ssaran[14], asyran[14], tauliq[14], ssaliq[14], asyliq[14], tauice[14], ssaice[14], asyice[14];
        for (int b = 0; b < 14; ++b) {
            ssaran[b] = tauran * (1.0 - b0r[b]);
            asyran[b] = ssaran[b] * c0r[b];
            tauliq[b] = ssaliq[b] = asyliq[b] = tauice[b] = ssaice[b] = asyice[b] = 0.0;
        }
        const double cldliq = c->cliqp[k], refliq = c->reliq[k], cldice = c->cicep[k], refice = c->reice[k];
        if (!(cldliq <= 0.0)) {
            const double factor = refliq - 1.5;
            const int index = imax(1, imin(57, (int)factor));
            const double fint = factor - (double)index;
            for (int b = 0; b < 14; ++b) {
                const double *e = t->extliq1 + b * 58, *s = t->ssaliq1 + b * 58, *a = t->asyliq1 + b * 58;
                const double extco = fmax(0.0, e[index - 1] + fint * (e[index] - e[index - 1]));
                const double ssaco = fmax(0.0, fmin(1.0, s[index - 1] + fint * (s[index] - s[index - 1])));
                const double asyco = fmax(0.0, fmin(1.0, a[index - 1] + fint * (a[index] - a[index - 1])));
                tauliq[b] = cldliq * extco;
                ssaliq[b] = tauliq[b] * ssaco;
                asyliq[b] = ssaliq[b] * asyco;
            }
        }
        if (!(cldice <= 0.0)) {
            const double dgeice = fmax(5.0, fmin(140.0, 1.0315 * refice));
            const double factor = (dgeice - 2.0) / 3.0;
            const int index = imax(1, imin(45, (int)factor));
            const double fint = factor - (double)index;
            for (int b = 0; b < 14; ++b) {
                const double *e = t->extice3 + b * 46, *s = t->ssaice3 + b * 46, *a = t->asyice3 + b * 46;
                const double extco = fmax(0.0, e[index - 1] + fint * (e[index] - e[index - 1]));
                const double ssaco = fmax(0.0, fmin(1.0, s[index - 1] + fint * (s[index] - s[index - 1])));
                const double asyco = fmax(0.0, fmin(1.0, a[index - 1] + fint * (a[index] - a[index - 1])));
                tauice[b] = cldice * extco;
                ssaice[b] = tauice[b] * ssaco;
                asyice[b] = ssaice[b] * asyco;
            }
        }
        for (int b = 0; b < 14; ++b) {
            c->taucw[b][k] = tauliq[b] + tauice[b] + tauran + 0.0;
            c->ssacw[b][k] = ssaliq[b] + ssaice[b] + ssaran[b] + 0.0;
            c->asycw[b][k] = asyliq[b] + asyice[b] + asyran[b] + 0.0;
        }
    }
    for (int k = 0; k < nl; ++k) {
        const double cf = c->cfrac[k];
        if (mcica)
            c->cldf[k] = (cf < O_FTINY) ? 0.0 : cf;
        else
            c->cldfrc[k] = cf / c->zcf1;
    }
}

/* ---- SW gas optics taumol16..29 ---------------------------------------------------------------- */
#define S2(tab, ind, g, ng) (tab)[((size_t)(ind)-1) * (ng) + (g)]
static double sum8(const double *A, int ng, int g, const int *id, const double *fc)
{
    return fc[0] * S2(A, id[0], g, ng) + fc[1] * S2(A, id[1], g, ng) + fc[2] * S2(A, id[2], g, ng) +
           fc[3] * S2(A, id[3], g, ng) + fc[4] * S2(A, id[4], g, ng) + fc[5] * S2(A, id[5], g, ng) +
           fc[6] * S2(A, id[6], g, ng) + fc[7] * S2(A, id[7], g, ng);
}
/* fac000 fac100 fac010 fac110 fac001 fac101 fac011 fac111 with the table offsets of nspa/nspb */
static void fac8(double fs, double f00, double f01, double f10, double f11, double *fc)
{
    const double fs1 = 1.0 - fs;
    fc[0] = fs1 * f00, fc[1] = fs * f00, fc[2] = fs1 * f10, fc[3] = fs * f10;
    fc[4] = fs1 * f01, fc[5] = fs * f01, fc[6] = fs1 * f11, fc[7] = fs * f11;
}
static void ind8(int id0, int id1, int js, int nrow, int d1, int d2, int *id)
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
static double abs4(const double *A, int ng, int g, int id0, int id1, int nrow, double f00, double f01, double f10,
                   double f11)
{
    const int i01 = imax(1, imin(nrow, id0 + 1)), i02 = imin(nrow, i01 + 1);
    const int i11 = imax(1, imin(nrow, id1 + 1)), i12 = imin(nrow, i11 + 1);
    return f00 * S2(A, i01, g, ng) + f10 * S2(A, i02, g, ng) + f01 * S2(A, i11, g, ng) + f11 * S2(A, i12, g, ng);
}

static void sw_band(const OT *t, const SwCol *c, int bi, int k, double *tg, double *tr)
{
    const SwBand *B = &t->sw[bi];
    const int ng = B->ng, jb = bi + 16;
    const int kk = k + 1, L = c->laytrop;
    const int jp = c->jp[k], jt = c->jt[k], jt1 = c->jt1[k];
    const double f00 = c->fac00[k], f01 = c->fac01[k], f10 = c->fac10[k], f11 = c->fac11[k];
    const int inds = c->indself[k], indsp = inds + 1, indf = c->indfor[k], indfp = indf + 1;
    const double sf = c->selffac[k], sfr = c->selffrac[k], ff = c->forfac[k], ffr = c->forfrac[k];
    const double *ca = c->colamt[k];
    const double cm = c->colmol[k];
    /* rrtm_sw_taumol_index: lower formula where kk <= laytrop, else upper */
    int id0, id1;
    if (kk <= L) {
        id0 = ((jp - 1) * 5 + (jt - 1)) * SW_NSPA[bi];
        id1 = (jp * 5 + (jt1 - 1)) * SW_NSPA[bi];
    } else {
        id0 = ((jp - 13) * 5 + (jt - 1)) * SW_NSPB[bi];
        id1 = ((jp - 12) * 5 + (jt1 - 1)) * SW_NSPB[bi];
    }
#define SELFI(g) (S2(B->selfref, inds, g, ng) + sfr * (S2(B->selfref, indsp, g, ng) - S2(B->selfref, inds, g, ng)))
#define FORI(g) (S2(B->forref, indf, g, ng) + ffr * (S2(B->forref, indfp, g, ng) - S2(B->forref, indf, g, ng)))
    double fc[8];
    int id[8];
    switch (jb) {
    case 16: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * B->rayl;
        if (kk <= L) {
            const double speccomb = ca[0] + B->strrat * ca[4];
            const double specmult = 8.0 * fmin(O_ONEMINUS, ca[0] / speccomb);
            const int js = 1 + (int)specmult;
            const double fs = specmult - (double)((int)specmult);
            fac8(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g)
                tg[g] = speccomb * sum8(B->absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        }
        if (kk >= L)
            for (int g = 0; g < ng; ++g) tg[g] = ca[4] * abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
        break;
    }
    case 17: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * 6.86e-10;
        const double speccomb = ca[0] + B->strrat * ca[1];
        if (kk <= L) {
            const double specmult = 8.0 * fmin(O_ONEMINUS, ca[0] / speccomb);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(((jp - 1) * 5 + (jt - 1)) * 9, (jp * 5 + (jt1 - 1)) * 9, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) {
                const double forpart = FORI(g);
                const double selfpart = (inds > 0) ? sf * SELFI(g) : 0.0;
                tg[g] = speccomb * sum8(B->absa, ng, g, id, fc) + ca[0] * (selfpart + ff * forpart);
            }
        } else {
            const double specmult = 4.0 * fmin(O_ONEMINUS, ca[0] / speccomb);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(((jp - 13) * 5 + (jt - 1)) * 5, ((jp - 12) * 5 + (jt1 - 1)) * 5, js, 1175, 5, 6, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8(B->absb, ng, g, id, fc) + (ca[0] * ff) * FORI(g);
        }
        break;
    }
    case 18: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * B->rayl;
        if (kk <= L) {
            const double speccomb = ca[0] + B->strrat * ca[4];
            const double ratio = (speccomb > 0.0) ? (ca[0] / speccomb) : 0.0;
            const double specmult = 8.0 * fmin(O_ONEMINUS, ratio);
            const int jsi = (int)specmult;
            fac8(specmult - (double)jsi, f00, f01, f10, f11, fc);
            ind8(id0, id1, 1 + jsi, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g)
                tg[g] = speccomb * sum8(B->absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = ca[4] * abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
        }
        break;
    }
    case 19: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * 2.29e-09;
        if (kk <= L) {
            const double speccomb = ca[0] + B->strrat * ca[1];
            const double specmult = 8.0 * fmin(0.999999, ca[0] / speccomb);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g)
                tg[g] = speccomb * sum8(B->absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = ca[1] * abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
        }
        break;
    }
    case 20: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * B->rayl;
        if (kk <= L) {
            for (int g = 0; g < ng; ++g) {
                const double absterm = abs4(B->absa, ng, g, id0, id1, 65, f00, f01, f10, f11);
                tg[g] = ca[0] * ((absterm) + sf * (SELFI(g)) + ff * (FORI(g))) + ca[4] * B->absch4[g];
            }
        } else {
            for (int g = 0; g < ng; ++g) {
                const double absterm = abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
                tg[g] = ca[0] * ((absterm) + ff * (FORI(g))) + ca[4] * B->absch4[g];
            }
        }
        break;
    }
    case 21: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * 9.41e-09;
        const double speccomb = ca[0] + B->strrat * ca[1];
        if (kk < L) {
            const double specmult = 8.0 * fmin(O_ONEMINUS, ca[0] / speccomb);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g)
                tg[g] = speccomb * sum8(B->absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g));
        }
        if (kk >= L) { /* at kk == L the lower-formula id0/id1 are used, as in the reference */
            const double specmult = 4.0 * fmin(O_ONEMINUS, ca[0] / speccomb);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 1175, 5, 6, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8(B->absb, ng, g, id, fc) + ca[0] * ff * FORI(g);
        }
        break;
    }
    case 22: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * 1.54e-08;
        const double o2cont = (4.35e-4 / (350.0 * 2.0)) * ca[5];
        if (kk <= L) {
            const double speccomb = ca[0] + B->strrat * ca[5];
            const double ratio = (speccomb > 0.0) ? (ca[0] / speccomb) : 0.0;
            const double specmult = 8.0 * fmin(O_ONEMINUS, ratio);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g)
                tg[g] = speccomb * sum8(B->absa, ng, g, id, fc) + ca[0] * (sf * SELFI(g) + ff * FORI(g)) + o2cont;
        } else {
            for (int g = 0; g < ng; ++g)
                tg[g] = ca[5] * 1.6 * abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11) + o2cont;
        }
        break;
    }
    case 23: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * B->rayl_v[g];
        if (kk <= L) {
            for (int g = 0; g < ng; ++g)
                tg[g] = ca[0] * (1.029 * abs4(B->absa, ng, g, id0, id1, 65, f00, f01, f10, f11) + sf * SELFI(g) + ff * FORI(g));
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = 0.0;
        }
        break;
    }
    case 24: {
        if (kk <= L) {
            const double speccomb = ca[0] + B->strrat * ca[5];
            const double specmult = 8.0 * fmin(O_ONEMINUS, ca[0] / speccomb);
            const int jsi = (int)specmult;
            const int js = 1 + jsi;
            const double fs = specmult - (double)jsi;
            fac8(fs, f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) {
                tg[g] = speccomb * sum8(B->absa, ng, g, id, fc) + ca[2] * B->abso3a[g] + ca[0] * (sf * SELFI(g) + ff * FORI(g));
                const double ray0 = B->rayla[(js - 1) * 8 + g], ray1 = B->rayla[js * 8 + g];
                tr[g] = cm * (ray0 + fs * (ray1 - ray0));
            }
        } else {
            for (int g = 0; g < ng; ++g) {
                tg[g] = ca[5] * abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11) + ca[2] * B->abso3b[g];
                tr[g] = cm * B->raylb[g];
            }
        }
        break;
    }
    case 25: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * B->rayl_v[g];
        if (kk <= L) {
            for (int g = 0; g < ng; ++g)
                tg[g] = ca[0] * abs4(B->absa, ng, g, id0, id1, 65, f00, f01, f10, f11) + ca[2] * B->abso3a[g];
        } else {
            for (int g = 0; g < ng; ++g) tg[g] = ca[2] * B->abso3b[g];
        }
        break;
    }
    case 26: {
        for (int g = 0; g < ng; ++g) tg[g] = 0.0, tr[g] = cm * B->rayl_v[g];
        break;
    }
    case 27: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * B->rayl_v[g];
        if (kk < L)
            for (int g = 0; g < ng; ++g) tg[g] = ca[2] * abs4(B->absa, ng, g, id0, id1, 65, f00, f01, f10, f11);
        if (kk >= L)
            for (int g = 0; g < ng; ++g) tg[g] = ca[2] * abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11);
        break;
    }
    case 28: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * B->rayl;
        const double speccomb = ca[2] + B->strrat * ca[5];
        if (kk <= L) {
            const double specmult = 8.0 * fmin(O_ONEMINUS, ca[2] / speccomb);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 585, 9, 10, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8(B->absa, ng, g, id, fc);
        } else {
            const double specmult = 4.0 * fmin(O_ONEMINUS, ca[2] / speccomb);
            const int js = 1 + (int)specmult;
            fac8(specmult - (double)((int)specmult), f00, f01, f10, f11, fc);
            ind8(id0, id1, js, 1175, 5, 6, id);
            for (int g = 0; g < ng; ++g) tg[g] = speccomb * sum8(B->absb, ng, g, id, fc);
        }
        break;
    }
    case 29: {
        for (int g = 0; g < ng; ++g) tr[g] = cm * 9.30e-11;
        if (kk <= L) {
            const int isp = imin(inds + 1, 10), ifp = imin(indf + 1, 4);
            for (int g = 0; g < ng; ++g) {
                const double selfv = S2(B->selfref, inds, g, ng) + sfr * (S2(B->selfref, isp, g, ng) - S2(B->selfref, inds, g, ng));
                const double forv = S2(B->forref, indf, g, ng) + ffr * (S2(B->forref, ifp, g, ng) - S2(B->forref, indf, g, ng));
                tg[g] = ca[0] * ((abs4(B->absa, ng, g, id0, id1, 65, f00, f01, f10, f11)) + sf * selfv + ff * forv) +
                        ca[1] * B->absco2[g];
            }
        } else {
            for (int g = 0; g < ng; ++g)
                tg[g] = ca[1] * abs4(B->absb, ng, g, id0, id1, 235, f00, f01, f10, f11) + ca[0] * B->absh2o[g];
        }
        break;
    }
    }
#undef SELFI
#undef FORI
}

void osw_taumol(const OT *t, const SwCol *c, double (*taug)[O_NL], double (*taur)[O_NL])
{
    double tg[16], tr[16];
    for (int b = 0; b < 14; ++b)
        for (int k = 0; k < c->nlay; ++k) {
            for (int g = 0; g < 16; ++g) tg[g] = 0.0, tr[g] = 0.0;
            sw_band(t, c, b, k, tg, tr);
            for (int g = 0; g < SW_NGA[b]; ++g) {
                taug[SW_NSA[b] + g][k] = tg[g];
                taur[SW_NSA[b] + g][k] = tr[g];
            }
        }
}

/* ---- spcvrtc (phys_rrtm_sw_spcvrtc.cu:217-547), iswmode = 2 ------------------------------------ */
static double expm(double zb, const double *tbl, int clamped)
{
    if (zb <= 0.06) return 1.0 - zb + 0.5 * zb * zb;
    const double ftind = zb / ((1.0 / 0.278) + zb);
    int itind = clamped ? (int)(0.5 + 10000.0 * ftind) : (int)(ftind * 10000.0 + 0.5);
    if (clamped) itind = imax(0, imin(10000, itind));
    return tbl[itind];
}

/* swc_layer_rt for zssaw < zcrit (the conservative branch needs zssaw >= 0.9999995, which the
 * 1 - 1e-6 clamp excludes; it is kept for completeness) */
static void layer_rt(double ztau1, double zssaw, double zssa1, double zgam1, double zgam2, double zgam3, double zgam4,
                     double cz, double snt, const double *tbl, double *rb, double *rd, double *tb, double *td)
{
    const double eps1 = 1.0e-8;
    if (zssaw >= 0.9999995) {
        const double za1 = zgam1 * cz - zgam3;
        const double za2 = zgam1 * ztau1;
        const double zb1 = fmin(ztau1 * snt, 500.0);
        const double zb2 = expm(zb1, tbl, 0);
        *rb = fmax(0.0, fmin(1.0, (za2 - za1 * (1.0 - zb2)) / (1.0 + za2)));
        *tb = fmax(0.0, fmin(1.0, 1.0 - *rb));
        *rd = fmax(0.0, fmin(1.0, za2 / (1.0 + za2)));
        *td = fmax(0.0, fmin(1.0, 1.0 - *rd));
        return;
    }
    const double za1 = zgam1 * zgam4 + zgam2 * zgam3;
    const double za2 = zgam1 * zgam3 + zgam2 * zgam4;
    const double zrk = sqrt((zgam1 - zgam2) * (zgam1 + zgam2));
    const double zrk2 = 2.0 * zrk;
    const double zrp = zrk * cz;
    const double zrp1 = 1.0 + zrp;
    const double zrm1 = 1.0 - zrp;
    double zrpp = 1.0 - zrp * zrp;
    if (fabs(zrpp) < eps1) zrpp = (zrpp < 0.0) ? -eps1 : eps1;
    const double zrkg1 = zrk + zgam1;
    const double zrkg3 = zrk * zgam3;
    const double zrkg4 = zrk * zgam4;
    const double zr1 = zrm1 * (za2 + zrkg3);
    const double zr2 = zrp1 * (za2 - zrkg3);
    const double zr3 = zrk2 * (zgam3 - za2 * cz);
    const double zr4 = zrpp * zrkg1;
    const double zr5 = zrpp * (zrk - zgam1);
    const double zt1 = zrp1 * (za1 + zrkg4);
    const double zt2 = zrm1 * (za1 - zrkg4);
    const double zt3 = zrk2 * (zgam4 + za1 * cz);
    const double zb1 = fmin(zrk * ztau1, 500.0);
    const double zexm1 = expm(zb1, tbl, 0);
    const double zexp1 = 1.0 / zexm1;
    const double zb2 = fmin(ztau1 * snt, 500.0);
    const double zexm2 = expm(zb2, tbl, 0);
    const double zexp2 = 1.0 / zexm2;
    const double ze1r45 = zr4 * zexp1 + zr5 * zexm1;
    if (ze1r45 >= -eps1 && ze1r45 <= eps1) {
        *rb = eps1;
        *tb = zexm2;
    } else {
        const double zden1 = zssa1 / ze1r45;
        *rb = fmax(0.0, fmin(1.0, (zr1 * zexp1 - zr2 * zexm1 - zr3 * zexm2) * zden1));
        *tb = fmax(0.0, fmin(1.0, zexm2 * (1.0 - (zt1 * zexp1 - zt2 * zexm1 - zt3 * zexp2) * zden1)));
    }
    const double zden1d = zr4 / (ze1r45 * zrkg1);
    *rd = fmax(0.0, fmin(1.0, zgam2 * (zexp1 - zexm1) * zden1d));
    *td = fmax(0.0, fmin(1.0, zrk2 * zden1d));
}

/* delta scaling + PIFM (iswmode 2) for one layer */
static void layer_opt(double ztau0, double zssa0, double zasy0, double cz, double snt, const double *tbl,
                      double *rb, double *rd, double *tb, double *td, double *ztau1_out)
{
    const double zssaw = fmin(O_ONEMINUS, zssa0 / ztau0);
    const double zasyw = zasy0 / fmax(O_FTINY, zssa0);
    const double za1s = zasyw * zasyw;
    const double za2s = zssaw * za1s;
    const double ztau1 = (1.0 - za2s) * ztau0;
    const double zssa1 = (zssaw - za2s) / (1.0 - za2s);
    const double zasy1 = zasyw / (1.0 + zasyw);
    const double zasy3 = 0.75 * zasy1;
    const double zgam1 = 2.0 - zssa1 * (1.25 + zasy3);
    const double zgam2 = 0.75 * zssa1 * (1.0 - zasy1);
    const double zgam3 = 0.5 - zasy3 * cz;
    const double zgam4 = 1.0 - zgam3;
    layer_rt(ztau1, zssaw, zssa1, zgam1, zgam2, zgam3, zgam4, cz, snt, tbl, rb, rd, tb, td);
    *ztau1_out = ztau1;
}

static void swflux(const double *zrefb, const double *zrefd, const double *ztrab, const double *ztrad,
                   const double *zldbt, const double *ztdbt, int nlay, double *zfu, double *zfd)
{
    const int nlp1 = nlay + 1;
    double zrupb[O_NL + 3], zrupd[O_NL + 3], zrdnd[O_NL + 3], ztdn[O_NL + 3];
    zrupb[1] = zrefb[1];
    zrupd[1] = zrefd[1];
    for (int k = 1; k <= nlay; ++k) {
        const int kp = k + 1;
        const double zden1 = 1.0 / (1.0 - zrupd[k] * zrefd[kp]);
        zrupb[kp] = zrefb[kp] + (ztrad[kp] * ((ztrab[kp] - zldbt[kp]) * zrupd[k] + zldbt[kp] * zrupb[k])) * zden1;
        zrupd[kp] = zrefd[kp] + ztrad[kp] * ztrad[kp] * zrupd[k] * zden1;
    }
    ztdn[nlp1] = 1.0;
    zrdnd[nlp1] = 0.0;
    ztdn[nlay] = ztrab[nlp1];
    zrdnd[nlay] = zrefd[nlp1];
    for (int k = nlay; k >= 2; --k) {
        const double zden1 = 1.0 / (1.0 - zrefd[k] * zrdnd[k]);
        ztdn[k - 1] = ztdbt[k] * ztrab[k] + (ztrad[k] * ((ztdn[k] - ztdbt[k]) + ztdbt[k] * zrefb[k] * zrdnd[k])) * zden1;
        zrdnd[k - 1] = zrefd[k] + ztrad[k] * ztrad[k] * zrdnd[k] * zden1;
    }
    for (int k = 1; k <= nlp1; ++k) {
        const double zden1 = 1.0 / (1.0 - zrdnd[k] * zrupd[k]);
        zfu[k] = (ztdbt[k] * zrupb[k] + (ztdn[k] - ztdbt[k]) * zrupd[k]) * zden1;
        zfd[k] = ztdbt[k] + (ztdn[k] - ztdbt[k] + ztdbt[k] * zrupb[k] * zrdnd[k]) * zden1;
    }
}

static const int SW_IDXSFC[14] = {1, 1, 1, 1, 1, 1, 1, 1, 0, 2, 2, 2, 2, 1}; /* bands 16..29 */

/* Total-sky and clear-sky band fluxes summed to level fluxes fup/fdn [nlay+1] (surface-first).
 * mcica = 0: the reference's spcvrtc (cldfrc mixing, cf1 weighting); mcica = 1: the same two-stream
 * with the per-g 0/1 McICA mask in place of cldfrc and no cf1 weighting. */
/* One g-point of spcvrtc for one column, both passes (the body of the reference's per-g launch,
 * phys_rrtm_sw_spcvrtc.cu:217-484).  Per-layer inputs are surface-first (index k-1 for Fortran layer
 * k): tr/tg the Rayleigh and gas optical depths; zc the cloud weight of the total-sky pass (cldfrc
 * on the NAM route, the McICA 0/1 mask bit otherwise) and tc/sc/ac the band's cloud optics.  Writes
 * the clear-sky level fluxes f0u/f0d and, when cloudy != 0, the total-sky ones fcu/fcd, per unit
 * incident flux, levels 1..nlay+1 (index 0 unused). */
void osw_gpoint(const double *tbl, int nlay, double cz, double snt, double albb, double albd, const double *tr,
                const double *tg, const double *ta, const double *sa, const double *ya, int cloudy, const double *zc,
                const double *tc, const double *sc, const double *ac, double *f0u, double *f0d, double *fcu, double *fcd)
{
    const int nlp1 = nlay + 1;
    double zrefb[O_NL + 3], zrefd[O_NL + 3], ztrab[O_NL + 3], ztrad[O_NL + 3], ztdbt[O_NL + 3], zldbt[O_NL + 3];
    double ztaus[O_NL + 3], zssas[O_NL + 3], zasys[O_NL + 3], zldbt0[O_NL + 3];
    double ztdbt0 = 1.0;
    ztdbt[nlp1] = 1.0;
    zldbt[1] = 0.0;
    zrefb[1] = albb;
    zrefd[1] = albd;
    ztrab[1] = 0.0;
    ztrad[1] = 0.0;
    for (int k = nlay; k >= 1; --k) {
        const int kp = k + 1;
        /* clear-sky optical parameters with the aerosol (phys_rrtm_sw_spcvrtc.cu:323-329) */
        const double a_t = ta ? ta[k - 1] : 0.0, a_s = sa ? sa[k - 1] : 0.0, a_y = ya ? ya[k - 1] : 0.0;
        const double ztau0 = fmax(O_FTINY, tr[k - 1] + tg[k - 1] + a_t);
        const double zssa0 = tr[k - 1] + a_t * a_s;
        const double zasy0 = a_y * a_s * a_t;
        ztaus[k] = ztau0;
        zssas[k] = zssa0;
        zasys[k] = zasy0;
        double ztau1;
        layer_opt(ztau0, zssa0, zasy0, cz, snt, tbl, &zrefb[kp], &zrefd[kp], &ztrab[kp], &ztrad[kp], &ztau1);
        const double zexp3 = expm(ztau1 * snt, tbl, 1);
        ztdbt[k] = zexp3 * ztdbt[kp];
        zldbt[kp] = zexp3;
        const double zexp4 = expm(ztau0 * snt, tbl, 1);
        zldbt0[k] = zexp4;
        ztdbt0 = zexp4 * ztdbt0;
    }
    swflux(zrefb, zrefd, ztrab, ztrad, zldbt, ztdbt, nlay, f0u, f0d);
    if (!cloudy) return;
    ztdbt0 = 1.0;
    zldbt[1] = 0.0;
    for (int k = nlay; k >= 1; --k) {
        const int kp = k + 1;
        const double zc1 = zc[k - 1];
        const double zc0 = 1.0 - zc1;
        if (zc1 > O_FTINY) {
            const double ztau0 = ztaus[k] + tc[k - 1];
            const double zssa0 = zssas[k] + sc[k - 1];
            const double zasy0 = zasys[k] + ac[k - 1];
            const double zrefb1 = zrefb[kp], zrefd1 = zrefd[kp], ztrab1 = ztrab[kp], ztrad1 = ztrad[kp];
            double ztau1;
            layer_opt(ztau0, zssa0, zasy0, cz, snt, tbl, &zrefb[kp], &zrefd[kp], &ztrab[kp], &ztrad[kp], &ztau1);
            zrefb[kp] = zc0 * zrefb1 + zc1 * zrefb[kp];
            zrefd[kp] = zc0 * zrefd1 + zc1 * zrefd[kp];
            ztrab[kp] = zc0 * ztrab1 + zc1 * ztrab[kp];
            ztrad[kp] = zc0 * ztrad1 + zc1 * ztrad[kp];
            const double zexp3 = expm(ztau1 * snt, tbl, 1);
            zldbt[kp] = zc0 * zldbt[kp] + zc1 * zexp3;
            ztdbt[k] = zldbt[kp] * ztdbt[kp];
            const double zexp4 = expm(ztau0 * snt, tbl, 1);
            ztdbt0 = (zc0 * zldbt0[k] + zc1 * zexp4) * ztdbt0;
        } else {
            ztdbt[k] = zldbt[kp] * ztdbt[kp];
            ztdbt0 = zldbt0[k] * ztdbt0;
        }
    }
    swflux(zrefb, zrefd, ztrab, ztrad, zldbt, ztdbt, nlay, fcu, fcd);
}

/* Total-sky and clear-sky band fluxes summed to level fluxes fup/fdn [nlay+1] (surface-first).
 * mcica = 0: the reference's spcvrtc (cldfrc mixing, cf1 weighting); mcica = 1: the same two-stream
 * with the per-g 0/1 McICA mask in place of cldfrc and no cf1 weighting. */
void osw_spcvrtc(const OT *t, const SwCol *c, double (*taug)[O_NL], double (*taur)[O_NL], const uint32_t *mask,
                 int mcica, double *fup, double *fdn, double *fdn0_sfc)
{
    const int nlay = c->nlay, nlp1 = nlay + 1;
    const double *tbl = t->sw_exp_tbl;
    const double cz = c->cosz, snt = c->sntz, cf1v = c->zcf1, cf0v = c->zcf0;
    double fxupc[14][O_NL + 1], fxdnc[14][O_NL + 1], fxup0[14][O_NL + 1], fxdn0[14][O_NL + 1];
    for (int b = 0; b < 14; ++b)
        for (int k = 0; k < nlp1; ++k) fxupc[b][k] = fxdnc[b][k] = fxup0[b][k] = fxdn0[b][k] = 0.0;
    double f0u[O_NL + 3], f0d[O_NL + 3], fcu[O_NL + 3], fcd[O_NL + 3], zc[O_NL];
    const int cloudy = mcica || cf1v > O_EPS;
    for (int bi = 0; bi < 14; ++bi) {
        const int ibd = SW_IDXSFC[bi];
        double albb, albd;
        if (ibd != 0) {
            albb = c->albbm[ibd - 1];
            albd = c->albdf[ibd - 1];
        } else {
            albb = 0.5 * (c->albbm[0] + c->albbm[1]);
            albd = 0.5 * (c->albdf[0] + c->albdf[1]);
        }
        for (int jg = SW_NSA[bi]; jg < SW_NSA[bi] + SW_NGA[bi]; ++jg) {
            const double zsolar = c->ssolar * c->sfluxzen[jg];
            for (int k = 0; k < nlay; ++k)
                zc[k] = mcica ? (((mask[k * 4 + (jg >> 5)] >> (jg & 31)) & 1u) ? 1.0 : 0.0) : c->cldfrc[k];
            osw_gpoint(tbl, nlay, cz, snt, albb, albd, taur[jg], taug[jg], c->tauae[bi], c->ssaae[bi], c->asyae[bi], cloudy,
                       zc, c->taucw[bi], c->ssacw[bi],
                       c->asycw[bi], f0u, f0d, fcu, fcd);
            for (int k = 1; k <= nlp1; ++k) {
                fxup0[bi][k - 1] = fxup0[bi][k - 1] + zsolar * f0u[k];
                fxdn0[bi][k - 1] = fxdn0[bi][k - 1] + zsolar * f0d[k];
            }
            if (cloudy)
                for (int k = 1; k <= nlp1; ++k) {
                    fxupc[bi][k - 1] = fxupc[bi][k - 1] + zsolar * fcu[k];
                    fxdnc[bi][k - 1] = fxdnc[bi][k - 1] + zsolar * fcd[k];
                }
        }
    }
    if (!mcica) {
        if (cf1v <= O_EPS) {
            for (int b = 0; b < 14; ++b)
                for (int k = 0; k < nlp1; ++k) fxupc[b][k] = fxup0[b][k], fxdnc[b][k] = fxdn0[b][k];
        } else {
            for (int b = 0; b < 14; ++b)
                for (int k = 0; k < nlp1; ++k) {
                    fxupc[b][k] = cf1v * fxupc[b][k] + cf0v * fxup0[b][k];
                    fxdnc[b][k] = cf1v * fxdnc[b][k] + cf0v * fxdn0[b][k];
                }
        }
    }
    /* swrad_post (phys_rrtm_sw_swrad.cu:319): band sums */
    for (int k = 0; k < nlp1; ++k) {
        double su = 0.0, sd = 0.0;
        for (int b = 0; b < 14; ++b) {
            su = su + fxupc[b][k];
            sd = sd + fxdnc[b][k];
        }
        fup[k] = su;
        fdn[k] = sd;
    }
    /* clear-sky surface downward flux (NAM route; the McICA route reports 0) */
    if (fdn0_sfc) {
        double s0 = 0.0;
        if (!mcica)
            for (int b = 0; b < 14; ++b) s0 = s0 + fxdn0[b][0];
        *fdn0_sfc = s0;
    }
}

/* ---- pieces for the oracle's own pins (tests) ------------------------------------------------- */
void osw_band(const OT *t, const SwCol *c, int bi, int k, double *tg, double *tr) { sw_band(t, c, bi, k, tg, tr); }
double osw_expm(const OT *t, double z, int clamped) { return expm(z, t->sw_exp_tbl, clamped); }
void osw_layer_opt(const OT *t, double tau0, double ssa0, double asy0, double cz, double *rb, double *rd, double *tb,
                   double *td, double *tau1)
{
    layer_opt(tau0, ssa0, asy0, cz, 1.0 / cz, t->sw_exp_tbl, rb, rd, tb, td, tau1);
}
