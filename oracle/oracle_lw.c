/* oracle_lw.c -- CPU FP64 restatement of the reference's RRTMG longwave path.
 *
 * TEST INFRASTRUCTURE ONLY: this is the checker the HIP library is compared against; nothing in
 * the product links it.  Every routine restates the float32 CUDA kernel it cites, term by term and
 * in the same order of operations, in double precision (compiled with -ffp-contract=off).
 * Layer index k is 0-based in the reference's internal SURFACE-FIRST order (k = 0 lowest layer);
 * "lower atmosphere" is k < laytrop (Fortran k <= laytrop).
 */
#include "oracle.h"

/* ---- rlwinit tables (phys_rrtm_lw_rtrnmr.py:97-132 lw_tables; phys_rrtm_lw_rlwinit.cu:296) ---- */
void olw_tables(double *exp_tbl, double *tfn_tbl, double *tau_tbl)
{
    const double bpade = 1.0 / 0.278;
    const double expeps = 1.0e-20;
    const double explimit = 87.0; /* aint(-log(tiny(real*4))) as the reference builds it */
    exp_tbl[0] = 1.0;
    tau_tbl[0] = 0.0;
    tfn_tbl[0] = 0.0;
    for (int i = 1; i < O_NTBL; ++i) {
        const double tfn = (double)i / (double)(O_NTBL - i);
        const double tau = bpade * tfn;
        const double e = (tau >= explimit) ? expeps : om_exp(-tau);
        tau_tbl[i] = tau;
        exp_tbl[i] = e;
        tfn_tbl[i] = (tau < 0.06) ? tau / 6.0 : 1.0 - 2.0 * ((1.0 / tau) - (e / (1.0 - e)));
    }
    exp_tbl[O_NTBL] = expeps;
    tau_tbl[O_NTBL] = 1.0e10;
    tfn_tbl[O_NTBL] = 1.0;
}

/* ---- lwrad_prep (phys_rrtm_lw_lwrad.cu:98-306) for one column --------------------------------
 * Inputs are the top-first radiation arrays of the prep (Rad); outputs surface-first LwCol. */
static const double LW_A0[16] = {1.66, 1.55, 1.58, 1.66, 1.54, 1.454, 1.89, 1.33,
                                 1.668, 1.66, 1.66, 1.66, 1.66, 1.66, 1.66, 1.66};
static const double LW_A1[16] = {0.00, 0.25, 0.22, 0.00, 0.13, 0.446, -0.10, 0.40,
                                 -0.006, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00};
static const double LW_A2[16] = {0.00, -12.0, -11.7, 0.00, -0.72, -0.243, 0.19, -0.062,
                                 0.414, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00, 0.00};

void olw_prep(const Rad *r, double sfemis, LwCol *c)
{
    const int nl = r->nrl;
    c->nlay = nl;
    const double tem1g = 100.0 * O_CON_G;
    const double tem2c = 1.0e-20 * 1.0e3 * O_CON_AVGD;
    c->tz[0] = r->tlvl[nl]; /* ivflip 0: tz(0) = tlvl(nlp1) */
    for (int k = 0; k < nl; ++k) {
        const int kin = nl - 1 - k;
        const double dlp = r->plvl[nl - k] - r->plvl[nl - k - 1];
        const double qv = r->qlyr[kin];
        const double ov = r->olyr[kin];
        c->pavel[k] = r->plyr[kin];
        c->tavel[k] = r->tlyr[kin];
        c->delp[k] = dlp;
        /* aerosol: tauaer = tau (1 - ssa), band by band (phys_rrtm_lw_lwrad.cu:214-219) */
        for (int b = 0; b < 16; ++b) c->tauaer[b][k] = r->aerlw[0][b][kin] * (1.0 - r->aerlw[1][b][kin]);
        c->tz[k + 1] = r->tlvl[nl - k - 1];
        const double h2ovmr = fmax(0.0, qv * O_AMDW / (1.0 - qv));
        const double o3vmr = fmax(0.0, ov * O_AMDO3);
        const double tem0 = (1.0 - h2ovmr) * O_CON_AMD + h2ovmr * O_CON_AMW;
        const double cdry = tem2c * dlp / (tem1g * tem0 * (1.0 + h2ovmr));
        const double temcol = 1.0e-12 * cdry;
        c->coldry[k] = cdry;
        c->colamt[k][0] = fmax(0.0, cdry * h2ovmr);
        c->colamt[k][1] = fmax(temcol, cdry * r->vmr[0]);
        c->colamt[k][2] = fmax(temcol, cdry * o3vmr);
        c->colamt[k][3] = fmax(temcol, cdry * r->vmr[1]);
        c->colamt[k][4] = fmax(temcol, cdry * r->vmr[2]);
        c->colamt[k][5] = fmax(0.0, cdry * r->vmr[3]);
        c->colamt[k][6] = fmax(0.0, cdry * r->vmr[4]);
        c->wx[k][0] = fmax(0.0, cdry * r->vmr[8]); /* ccl4  */
        c->wx[k][1] = fmax(0.0, cdry * r->vmr[5]); /* cfc11 */
        c->wx[k][2] = fmax(0.0, cdry * r->vmr[6]); /* cfc12 */
        c->wx[k][3] = fmax(0.0, cdry * r->vmr[7]); /* cfc22 */
        c->cldfrc[k + 1] = r->cld[kin][0];
        c->clwp[k] = r->cld[kin][1];
        c->relw[k] = r->cld[kin][2];
        c->ciwp[k] = r->cld[kin][3];
        c->reiw[k] = r->cld[kin][4];
        c->crp[k] = r->cld[kin][5];
    }
    c->cldfrc[0] = 1.0;
    c->cldfrc[nl + 1] = 0.0;
    for (int k = 0; k < nl; ++k) {
        double summol = 0.0;
        for (int m = 1; m < 7; ++m) summol = summol + c->colamt[k][m];
        c->colbrd[k] = c->coldry[k] - summol;
        /* setcoef's h2ovmr is recovered from the column amounts (lw_core.py:238) */
        c->h2ovmr[k] = c->colamt[k][0] / fmax(c->coldry[k], 1e-30);
    }
    double tem1 = 0.0, tem2 = 0.0;
    for (int k = 0; k < nl; ++k) {
        tem1 = tem1 + c->coldry[k] + c->colamt[k][0];
        tem2 = tem2 + c->colamt[k][0];
    }
    const double tem0p = 10.0 * tem2 / (O_AMDW * tem1 * O_CON_G);
    c->pwvcm = tem0p * r->plvl[nl];
    for (int b = 0; b < 16; ++b) {
        const int j1 = b + 1;
        if (j1 == 1 || j1 == 4 || j1 == 10)
            c->secdiff[b] = 1.66;
        else
            c->secdiff[b] = fmin(1.80, fmax(1.50, LW_A0[b] + LW_A1[b] * om_exp(LW_A2[b] * c->pwvcm)));
    }
    int lcf = 0;
    for (int k = 1; k <= nl; ++k)
        if (c->cldfrc[k] > O_EPS) {
            lcf = 1;
            break;
        }
    c->lcf1 = lcf;
    for (int b = 0; b < 16; ++b) c->semiss[b] = (sfemis > O_EPS && sfemis <= 1.0) ? sfemis : 1.0;
    c->stemp = r->tskn;
}

/* ---- LW cldprop (phys_rrtm_lw_cldprop.cu:79), ilwcliq = 1, ilwcice = 1 (NAM) ------------------ */
void olw_cldprop(const OT *t, LwCol *c)
{
    const int nl = c->nlay;
    for (int b = 0; b < 16; ++b)
        for (int k = 0; k < nl; ++k) c->taucld[b][k] = 0.0;
    for (int k = 0; k < nl; ++k) c->cldf[k] = 0.0;
    if (!c->lcf1) return;
    const double absrain = 0.33e-3;
    for (int k = 0; k < nl; ++k) {
        const double cf = c->cldfrc[k + 1];
        c->cldf[k] = (cf < 0.0) ? 0.0 : cf;
        if (!(cf > 0.0)) continue;
        const double tauran = absrain * c->crp[k];
        const double tausnw = 0.0; /* snow path input is zero on this path */
        const double cldliq = c->clwp[k], cldice = c->ciwp[k];
        const double refliq = c->relw[k];
        double refice = c->reiw[k];
        double tauliq[16], tauice[16];
        for (int b = 0; b < 16; ++b) tauliq[b] = 0.0, tauice[b] = 0.0;
        if (cldliq > 0.0) {
            const double factor = refliq - 1.5;
            const int index = imax(1, imin(57, (int)factor));
            const double fint = factor - (double)index;
            for (int b = 0; b < 16; ++b) {
                const double a1 = t->absliq1[b * 58 + (index - 1)];
                const double a2 = t->absliq1[b * 58 + index];
                tauliq[b] = fmax(0.0, cldliq * (a1 + fint * (a2 - a1)));
            }
        }
        if (cldice > 0.0) {
            refice = fmin(130.0, fmax(13.0, refice));
            for (int b = 0; b < 16; ++b) {
                const int ia0 = t->ipat[b] - 1;
                tauice[b] = fmax(0.0, cldice * (t->absice1[ia0 * 2 + 0] + t->absice1[ia0 * 2 + 1] / refice));
            }
        }
        for (int b = 0; b < 16; ++b) c->taucld[b][k] = tauice[b] + tauliq[b] + tauran + tausnw;
    }
}

/* ---- LW setcoef (phys_rrtm_lw_setcoef.cu:82-244) ----------------------------------------------- */
static double planck(const OT *t, int b, double temp) /* delwave * interpolated totplnk */
{
    const int ind = imin(180, imax(1, (int)(temp - 159.0)));
    const double fr = temp - (double)((int)temp);
    const double *tp = t->totplnk + (size_t)b * 181;
    return t->delwave[b] * (tp[ind - 1] + fr * (tp[ind] - tp[ind - 1]));
}

void olw_setcoef(const OT *t, LwCol *c)
{
    const int nl = c->nlay;
    for (int b = 0; b < 16; ++b) {
        c->pklay[0][b] = planck(t, b, c->stemp);
        c->pklev[0][b] = planck(t, b, c->tz[0]);
    }
    int lt = 0;
    for (int k = 0; k < nl; ++k) {
        const double pav = c->pavel[k], tav = c->tavel[k], tzk = c->tz[k + 1];
        for (int b = 0; b < 16; ++b) {
            c->pklay[k + 1][b] = planck(t, b, tav);
            c->pklev[k + 1][b] = planck(t, b, tzk);
        }
        const double plog = om_log(pav);
        const int jp = imax(1, imin(58, (int)(36.0 - 5.0 * (plog + 0.04))));
        const int jp1 = jp + 1;
        const double fp = fmax(0.0, fmin(1.0, 5.0 * (t->preflog[jp - 1] - plog)));
        const double tem1t = (tav - t->tref[jp - 1]) / 15.0;
        const double tem2t = (tav - t->tref[jp1 - 1]) / 15.0;
        const int jt = imax(1, imin(4, (int)(3.0 + tem1t)));
        const int jt1 = imax(1, imin(4, (int)(3.0 + tem2t)));
        const double ft = fmax(-0.5, fmin(1.5, tem1t - (double)(jt - 3)));
        const double ft1 = fmax(-0.5, fmin(1.5, tem2t - (double)(jt1 - 3)));
        c->jp[k] = jp;
        c->jt[k] = jt;
        c->jt1[k] = jt1;
        const double tem1 = 1.0 - fp;
        c->fac10[k] = tem1 * ft;
        c->fac00[k] = tem1 * (1.0 - ft);
        c->fac11[k] = fp * ft1;
        c->fac01[k] = fp * (1.0 - ft1);
        const double h2o = c->h2ovmr[k];
        const double forf = pav * O_STPFAC / (tav * (1.0 + h2o));
        const double selff = h2o * forf;
        c->scaleminor[k] = pav / tav;
        c->scaleminorn2[k] = (pav / tav) * (c->colbrd[k] / (c->coldry[k] + c->colamt[k][0]));
        const double tem1m = (tav - 180.8) / 7.2;
        const int indmin = imin(18, imax(1, (int)tem1m));
        c->indminor[k] = indmin;
        c->minorfrac[k] = tem1m - (double)indmin;
#define CHI(g, lev) t->chi_mls[((lev)-1) * 7 + ((g)-1)]
        for (int m = 0; m < 6; ++m) c->rfrate[k][m][0] = c->rfrate[k][m][1] = 0.0;
        c->rfrate[k][0][0] = CHI(1, jp) / CHI(2, jp);
        c->rfrate[k][0][1] = CHI(1, jp1) / CHI(2, jp1);
        if (plog > 4.56) {
            lt += 1;
            const double tem1f = (332.0 - tav) / 36.0;
            const int indf = imin(2, imax(1, (int)tem1f));
            c->indfor[k] = indf;
            c->forfrac[k] = tem1f - (double)indf;
            const double tem1s = (tav - 188.0) / 7.2;
            const int inds = imin(9, imax(1, (int)tem1s - 7));
            c->indself[k] = inds;
            c->selffrac[k] = tem1s - (double)(inds + 7);
            c->rfrate[k][1][0] = CHI(1, jp) / CHI(3, jp);
            c->rfrate[k][1][1] = CHI(1, jp1) / CHI(3, jp1);
            c->rfrate[k][2][0] = CHI(1, jp) / CHI(4, jp);
            c->rfrate[k][2][1] = CHI(1, jp1) / CHI(4, jp1);
            c->rfrate[k][3][0] = CHI(1, jp) / CHI(6, jp);
            c->rfrate[k][3][1] = CHI(1, jp1) / CHI(6, jp1);
            c->rfrate[k][4][0] = CHI(4, jp) / CHI(2, jp);
            c->rfrate[k][4][1] = CHI(4, jp1) / CHI(2, jp1);
        } else {
            const double tem1f = (tav - 188.0) / 36.0;
            c->indfor[k] = 3;
            c->forfrac[k] = tem1f - 1.0;
            c->indself[k] = 0;
            c->selffrac[k] = 0.0;
            c->rfrate[k][5][0] = CHI(3, jp) / CHI(2, jp);
            c->rfrate[k][5][1] = CHI(3, jp1) / CHI(2, jp1);
        }
#undef CHI
        c->selffac[k] = c->colamt[k][0] * selff;
        c->forfac[k] = c->colamt[k][0] * forf;
    }
    c->laytrop = lt;
}

/* ---- LW gas optics: taugb01..16 ----------------------------------------------------------------
 * Each band writes tau[NG], fr[NG] for layer k.  Tables are [index][g] with g fastest (1-based
 * Fortran indices are kept and subtracted at access, as in the reference kernels). */
#define T2(tab, ind, g, ng) (tab)[((size_t)(ind)-1) * (ng) + (g)]
#define T3(tab, m, j, nj, g, ng) (tab)[(((size_t)(m)-1) * (nj) + ((j)-1)) * (ng) + (g)]

typedef struct {
    double fk0, fk1, fk2;
    int id[6]; /* id000 id010 id100 id110 id200 id210 */
} Sel;

/* lw4_lower_select (phys_rrtm_lw_taugb04.cu) / taugb03 inline selection: one eta side */
static Sel sel_one(double specparm, double fs, int ind)
{
    Sel s;
    if (specparm < 0.125) {
        const double p = fs - 1.0;
        const double p4 = (p * p) * (p * p);
        s.fk0 = p4;
        s.fk1 = 1.0 - p - 2.0 * p4;
        s.fk2 = p + p4;
        s.id[0] = ind, s.id[1] = ind + 9, s.id[2] = ind + 1, s.id[3] = ind + 10, s.id[4] = ind + 2, s.id[5] = ind + 11;
    } else if (specparm > 0.875) {
        const double p = -fs;
        const double p4 = (p * p) * (p * p);
        s.fk0 = p4;
        s.fk1 = 1.0 - p - 2.0 * p4;
        s.fk2 = p + p4;
        s.id[0] = ind + 1, s.id[1] = ind + 10, s.id[2] = ind, s.id[3] = ind + 9, s.id[4] = ind - 1, s.id[5] = ind + 8;
    } else {
        s.fk0 = 1.0 - fs;
        s.fk1 = fs;
        s.fk2 = 0.0;
        s.id[0] = ind, s.id[1] = ind + 9, s.id[2] = ind + 1, s.id[3] = ind + 10, s.id[4] = ind, s.id[5] = ind;
    }
    return s;
}

/* joint selection of bands 5, 7, 9, 12, 13, 15, 16: both sides cubic only when both are */
static void sel_joint(double sp0, double fs0, int ind0, double sp1, double fs1, int ind1, Sel *s0, Sel *s1)
{
    if (sp0 < 0.125 && sp1 < 0.125) {
        *s0 = sel_one(0.0, fs0, ind0);
        *s1 = sel_one(0.0, fs1, ind1);
    } else if (sp0 > 0.875 && sp1 > 0.875) {
        *s0 = sel_one(1.0, fs0, ind0);
        *s1 = sel_one(1.0, fs1, ind1);
    } else {
        *s0 = sel_one(0.5, fs0, ind0);
        *s1 = sel_one(0.5, fs1, ind1);
    }
}

typedef struct { /* binary-species eta parameter */
    double speccomb, specparm, fs;
    int js;
} Eta;

static Eta eta(double a, double ratio, double b, double mult)
{
    Eta e;
    e.speccomb = a + ratio * b;
    e.specparm = a / e.speccomb;
    const double specmult = mult * fmin(e.specparm, O_ONEMINUS);
    e.js = 1 + (int)specmult;
    e.fs = specmult - (double)((int)specmult);
    return e;
}

/* six-point major-species sum in the reference's order 000 010 100 110 200 210 */
static double maj6(const double *tab, int ng, int g, const Sel *s, double fa, double fb)
{
    return fs_mul(s->fk0, fa) * T2(tab, s->id[0], g, ng) + fs_mul(s->fk0, fb) * T2(tab, s->id[1], g, ng) +
           fs_mul(s->fk1, fa) * T2(tab, s->id[2], g, ng) + fs_mul(s->fk1, fb) * T2(tab, s->id[3], g, ng) +
           fs_mul(s->fk2, fa) * T2(tab, s->id[4], g, ng) + fs_mul(s->fk2, fb) * T2(tab, s->id[5], g, ng);
}

static void lw_band(const OT *t, const LwCol *c, int b, int k, double *tau, double *fr)
{
    const LwBand *B = &t->lw[b];
    const int ng = B->ng;
    const int lower = k < c->laytrop;
    const int jp = c->jp[k], jt = c->jt[k], jt1 = c->jt1[k];
    const double f00 = c->fac00[k], f01 = c->fac01[k], f10 = c->fac10[k], f11 = c->fac11[k];
    const int inds = c->indself[k], indsp = inds + 1;
    const int indf = c->indfor[k], indfp = indf + 1;
    const int indm = c->indminor[k], indmp = indm + 1;
    const double sf = c->selffac[k], sfr = c->selffrac[k];
    const double ff = c->forfac[k], ffr = c->forfrac[k];
    const double mf = c->minorfrac[k];
    const double *ca = c->colamt[k];
    const double(*rf)[2] = c->rfrate[k];
#define SELF(g) (sf * (T2(B->selfref, inds, g, ng) + sfr * (T2(B->selfref, indsp, g, ng) - T2(B->selfref, inds, g, ng))))
#define FOR(g) (ff * (T2(B->forref, indf, g, ng) + ffr * (T2(B->forref, indfp, g, ng) - T2(B->forref, indf, g, ng))))
#define MIN1(tab, g) (T2(tab, indm, g, ng) + mf * (T2(tab, indmp, g, ng) - T2(tab, indm, g, ng)))
#define ABS4(tab, i0, i1, g) \
    (f00 * T2(tab, i0, g, ng) + f10 * T2(tab, (i0) + 1, g, ng) + f01 * T2(tab, i1, g, ng) + f11 * T2(tab, (i1) + 1, g, ng))
    const int l0 = ((jp - 1) * 5 + (jt - 1)); /* lower rows: ind0 = l0*nspa + js */
    const int l1 = (jp * 5 + (jt1 - 1));
    const int u0 = ((jp - 13) * 5 + (jt - 1)); /* upper rows */
    const int u1 = ((jp - 12) * 5 + (jt1 - 1));
    switch (b + 1) {
    case 1: { /* taugb01.cu:99-171 */
        const double pp = c->pavel[k];
        const double scalen2 = c->colbrd[k] * c->scaleminorn2[k];
        if (lower) {
            const int i0 = l0 + 1, i1 = l1 + 1;
            const double corradj = (pp < 250.0) ? (1.0 - 0.15 * (250.0 - pp) / 154.4) : 1.0;
            for (int g = 0; g < ng; ++g) {
                const double taun2 = scalen2 * MIN1(B->ka_m[0], g);
                tau[g] = corradj * (ca[0] * ABS4(B->absa, i0, i1, g) + SELF(g) + FOR(g) + taun2);
                fr[g] = B->fracrefa[g];
            }
        } else {
            const int i0 = u0 + 1, i1 = u1 + 1;
            const double corradj = 1.0 - 0.15 * (pp / 95.6);
            for (int g = 0; g < ng; ++g) {
                const double taun2 = scalen2 * MIN1(B->kb_m[0], g);
                tau[g] = corradj * (ca[0] * ABS4(B->absb, i0, i1, g) + FOR(g) + taun2);
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 2: { /* taugb02.cu */
        if (lower) {
            const int i0 = l0 + 1, i1 = l1 + 1;
            const double corradj = 1.0 - 0.05 * (c->pavel[k] - 100.0) / 900.0;
            for (int g = 0; g < ng; ++g) {
                tau[g] = corradj * (ca[0] * ABS4(B->absa, i0, i1, g) + SELF(g) + FOR(g));
                fr[g] = B->fracrefa[g];
            }
        } else {
            const int i0 = u0 + 1, i1 = u1 + 1;
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[0] * ABS4(B->absb, i0, i1, g) + FOR(g);
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 3: { /* taugb03.cu: independent eta selection per side */
        const double p = c->coldry[k] * O_CHI(t, 4, jp + 1);
        const double ratn2o = ca[3] / p;
        const double adjcoln2o = (ratn2o > 1.5) ? (0.5 + om_pow(ratn2o - 0.5, 0.65)) * p : ca[3];
        if (lower) {
            const Eta e0 = eta(ca[0], rf[0][0], ca[1], 8.0), e1 = eta(ca[0], rf[0][1], ca[1], 8.0);
            const Eta em = eta(ca[0], B->refrat_m_a, ca[1], 8.0), ep = eta(ca[0], B->refrat_planck_a, ca[1], 8.0);
            const Sel s0 = sel_one(e0.specparm, e0.fs, l0 * 9 + e0.js);
            const Sel s1 = sel_one(e1.specparm, e1.fs, l1 * 9 + e1.js);
            for (int g = 0; g < ng; ++g) {
                const double n2om1 = T3(B->ka_m[0], indm, em.js, 9, g, ng) +
                                     em.fs * (T3(B->ka_m[0], indm, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indm, em.js, 9, g, ng));
                const double n2om2 = T3(B->ka_m[0], indmp, em.js, 9, g, ng) +
                                     em.fs * (T3(B->ka_m[0], indmp, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indmp, em.js, 9, g, ng));
                const double absn2o = n2om1 + mf * (n2om2 - n2om1);
                const double tmaj = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10);
                const double tmaj1 = e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11);
                tau[g] = tmaj + tmaj1 + SELF(g) + FOR(g) + adjcoln2o * absn2o;
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            const Eta e0 = eta(ca[0], rf[0][0], ca[1], 4.0), e1 = eta(ca[0], rf[0][1], ca[1], 4.0);
            const Eta em = eta(ca[0], B->refrat_m_b, ca[1], 4.0), ep = eta(ca[0], B->refrat_planck_b, ca[1], 4.0);
            const int i0 = u0 * 5 + e0.js, i1 = u1 * 5 + e1.js;
            for (int g = 0; g < ng; ++g) {
                const double n2om1 = T3(B->kb_m[0], indm, em.js, 5, g, ng) +
                                     em.fs * (T3(B->kb_m[0], indm, em.js + 1, 5, g, ng) - T3(B->kb_m[0], indm, em.js, 5, g, ng));
                const double n2om2 = T3(B->kb_m[0], indmp, em.js, 5, g, ng) +
                                     em.fs * (T3(B->kb_m[0], indmp, em.js + 1, 5, g, ng) - T3(B->kb_m[0], indmp, em.js, 5, g, ng));
                const double absn2o = n2om1 + mf * (n2om2 - n2om1);
                const double tmaj = e0.speccomb * (((1.0 - e0.fs) * f00) * T2(B->absb, i0, g, ng) + ((1.0 - e0.fs) * f10) * T2(B->absb, i0 + 5, g, ng) +
                                                   (e0.fs * f00) * T2(B->absb, i0 + 1, g, ng) + (e0.fs * f10) * T2(B->absb, i0 + 6, g, ng));
                const double tmaj1 = e1.speccomb * (((1.0 - e1.fs) * f01) * T2(B->absb, i1, g, ng) + ((1.0 - e1.fs) * f11) * T2(B->absb, i1 + 5, g, ng) +
                                                    (e1.fs * f01) * T2(B->absb, i1 + 1, g, ng) + (e1.fs * f11) * T2(B->absb, i1 + 6, g, ng));
                tau[g] = tmaj + tmaj1 + FOR(g) + adjcoln2o * absn2o;
                fr[g] = T2(B->fracrefb, ep.js, g, ng) + ep.fs * (T2(B->fracrefb, ep.js + 1, g, ng) - T2(B->fracrefb, ep.js, g, ng));
            }
        }
        break;
    }
    case 4: { /* taugb04.cu */
        if (lower) {
            const Eta e0 = eta(ca[0], rf[0][0], ca[1], 8.0), e1 = eta(ca[0], rf[0][1], ca[1], 8.0);
            const Eta ep = eta(ca[0], B->refrat_planck_a, ca[1], 8.0);
            const Sel s0 = sel_one(e0.specparm, e0.fs, l0 * 9 + e0.js);
            const Sel s1 = sel_one(e1.specparm, e1.fs, l1 * 9 + e1.js);
            for (int g = 0; g < ng; ++g) {
                const double tmaj = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10);
                const double tmaj1 = e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11);
                tau[g] = tmaj + tmaj1 + SELF(g) + FOR(g);
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            const Eta e0 = eta(ca[2], rf[5][0], ca[1], 4.0), e1 = eta(ca[2], rf[5][1], ca[1], 4.0);
            const Eta ep = eta(ca[2], B->refrat_planck_b, ca[1], 4.0);
            const int i0 = u0 * 5 + e0.js, i1 = u1 * 5 + e1.js;
            static const double scl[14] = {1, 1, 1, 1, 1, 1, 1, 0.92, 0.88, 1.07, 1.1, 0.99, 0.88, 0.943};
            for (int g = 0; g < ng; ++g) {
                const double tmaj = e0.speccomb * (((1.0 - e0.fs) * f00) * T2(B->absb, i0, g, ng) + ((1.0 - e0.fs) * f10) * T2(B->absb, i0 + 5, g, ng) +
                                                   (e0.fs * f00) * T2(B->absb, i0 + 1, g, ng) + (e0.fs * f10) * T2(B->absb, i0 + 6, g, ng));
                const double tmaj1 = e1.speccomb * (((1.0 - e1.fs) * f01) * T2(B->absb, i1, g, ng) + ((1.0 - e1.fs) * f11) * T2(B->absb, i1 + 5, g, ng) +
                                                    (e1.fs * f01) * T2(B->absb, i1 + 1, g, ng) + (e1.fs * f11) * T2(B->absb, i1 + 6, g, ng));
                double tg = tmaj + tmaj1;
                if (g >= 7) tg = tg * scl[g];
                tau[g] = tg;
                fr[g] = T2(B->fracrefb, ep.js, g, ng) + ep.fs * (T2(B->fracrefb, ep.js + 1, g, ng) - T2(B->fracrefb, ep.js, g, ng));
            }
        }
        break;
    }
    case 5: { /* taugb05.cu: joint selection */
        const double wxk = c->wx[k][0];
        if (lower) {
            const Eta e0 = eta(ca[0], rf[0][0], ca[1], 8.0), e1 = eta(ca[0], rf[0][1], ca[1], 8.0);
            const Eta em = eta(ca[0], B->refrat_m_a, ca[1], 8.0), ep = eta(ca[0], B->refrat_planck_a, ca[1], 8.0);
            Sel s0, s1;
            sel_joint(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, &s0, &s1);
            for (int g = 0; g < ng; ++g) {
                const double o3m1 = T3(B->ka_m[0], indm, em.js, 9, g, ng) +
                                    em.fs * (T3(B->ka_m[0], indm, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indm, em.js, 9, g, ng));
                const double o3m2 = T3(B->ka_m[0], indmp, em.js, 9, g, ng) +
                                    em.fs * (T3(B->ka_m[0], indmp, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indmp, em.js, 9, g, ng));
                const double abso3 = o3m1 + mf * (o3m2 - o3m1);
                tau[g] = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10) + e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11) +
                         SELF(g) + FOR(g) + abso3 * ca[2] + wxk * B->x1[g];
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            const Eta e0 = eta(ca[2], rf[5][0], ca[1], 4.0), e1 = eta(ca[2], rf[5][1], ca[1], 4.0);
            const Eta ep = eta(ca[2], B->refrat_planck_b, ca[1], 4.0);
            const int i0 = u0 * 5 + e0.js, i1 = u1 * 5 + e1.js;
            for (int g = 0; g < ng; ++g) {
                tau[g] = e0.speccomb * (((1.0 - e0.fs) * f00) * T2(B->absb, i0, g, ng) + ((1.0 - e0.fs) * f10) * T2(B->absb, i0 + 5, g, ng) +
                                        (e0.fs * f00) * T2(B->absb, i0 + 1, g, ng) + (e0.fs * f10) * T2(B->absb, i0 + 6, g, ng)) +
                         e1.speccomb * (((1.0 - e1.fs) * f01) * T2(B->absb, i1, g, ng) + ((1.0 - e1.fs) * f11) * T2(B->absb, i1 + 5, g, ng) +
                                        (e1.fs * f01) * T2(B->absb, i1 + 1, g, ng) + (e1.fs * f11) * T2(B->absb, i1 + 6, g, ng)) +
                         wxk * B->x1[g];
                fr[g] = T2(B->fracrefb, ep.js, g, ng) + ep.fs * (T2(B->fracrefb, ep.js + 1, g, ng) - T2(B->fracrefb, ep.js, g, ng));
            }
        }
        break;
    }
    case 6: { /* taugb06.cu */
        const double wx2 = c->wx[k][1], wx3 = c->wx[k][2];
        if (lower) {
            const int i0 = l0 + 1, i1 = l1 + 1;
            const double temp = c->coldry[k] * O_CHI(t, 2, jp + 1);
            const double ratco2 = ca[1] / temp;
            const double adjcolco2 = (ratco2 > 3.0) ? (2.0 + om_pow(ratco2 - 2.0, 0.77)) * temp : ca[1];
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[0] * ABS4(B->absa, i0, i1, g) + SELF(g) + FOR(g) + adjcolco2 * MIN1(B->ka_m[0], g) +
                         wx2 * B->x1[g] + wx3 * B->x2[g];
                fr[g] = B->fracrefa[g];
            }
        } else {
            for (int g = 0; g < ng; ++g) {
                tau[g] = wx2 * B->x1[g] + wx3 * B->x2[g];
                fr[g] = B->fracrefa[g];
            }
        }
        break;
    }
    case 7: { /* taugb07.cu: joint selection, rfrate rate 2 */
        const double temp = c->coldry[k] * O_CHI(t, 2, jp + 1);
        const double ratco2 = ca[1] / temp;
        if (lower) {
            const double adjcolco2 = (ratco2 > 3.0) ? (3.0 + om_pow(ratco2 - 3.0, 0.79)) * temp : ca[1];
            const Eta e0 = eta(ca[0], rf[1][0], ca[2], 8.0), e1 = eta(ca[0], rf[1][1], ca[2], 8.0);
            const Eta em = eta(ca[0], B->refrat_m_a, ca[2], 8.0), ep = eta(ca[0], B->refrat_planck_a, ca[2], 8.0);
            Sel s0, s1;
            sel_joint(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, &s0, &s1);
            for (int g = 0; g < ng; ++g) {
                const double co2m1 = T3(B->ka_m[0], indm, em.js, 9, g, ng) +
                                     em.fs * (T3(B->ka_m[0], indm, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indm, em.js, 9, g, ng));
                const double co2m2 = T3(B->ka_m[0], indmp, em.js, 9, g, ng) +
                                     em.fs * (T3(B->ka_m[0], indmp, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indmp, em.js, 9, g, ng));
                const double absco2 = co2m1 + mf * (co2m2 - co2m1);
                tau[g] = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10) + e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11) +
                         SELF(g) + FOR(g) + adjcolco2 * absco2;
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            const double adjcolco2 = (ratco2 > 3.0) ? (2.0 + om_pow(ratco2 - 2.0, 0.79)) * temp : ca[1];
            const int i0 = u0 + 1, i1 = u1 + 1;
            static const double scl[12] = {1, 1, 1, 1, 1, 0.92, 0.88, 1.07, 1.1, 0.99, 0.855, 1};
            for (int g = 0; g < ng; ++g) {
                double tg = ca[2] * ABS4(B->absb, i0, i1, g) + adjcolco2 * MIN1(B->kb_m[0], g);
                if (g >= 5 && g <= 10) tg = tg * scl[g];
                tau[g] = tg;
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 8: { /* taugb08.cu */
        const double temp = c->coldry[k] * O_CHI(t, 2, jp + 1);
        const double ratco2 = ca[1] / temp;
        const double adjcolco2 = (ratco2 > 3.0) ? (2.0 + om_pow(ratco2 - 2.0, 0.65)) * temp : ca[1];
        const double w3 = c->wx[k][2], w4 = c->wx[k][3];
        if (lower) {
            const int r0 = l0, r1 = l1;
            for (int g = 0; g < ng; ++g) {
                const double glow = ca[0] * (f00 * B->absa[r0 * ng + g] + f10 * B->absa[(r0 + 1) * ng + g] + f01 * B->absa[r1 * ng + g] +
                                             f11 * B->absa[(r1 + 1) * ng + g]) +
                                    SELF(g) + FOR(g) + adjcolco2 * MIN1(B->ka_m[0], g) + ca[2] * MIN1(B->ka_m[1], g) +
                                    ca[3] * MIN1(B->ka_m[2], g) + w3 * B->x1[g] + w4 * B->x2[g];
                tau[g] = glow;
                fr[g] = B->fracrefa[g];
            }
        } else {
            const int r0 = u0, r1 = u1;
            for (int g = 0; g < ng; ++g) {
                const double gup = ca[2] * (f00 * B->absb[r0 * ng + g] + f10 * B->absb[(r0 + 1) * ng + g] + f01 * B->absb[r1 * ng + g] +
                                            f11 * B->absb[(r1 + 1) * ng + g]) +
                                   adjcolco2 * MIN1(B->kb_m[0], g) + ca[3] * MIN1(B->kb_m[2], g) + w3 * B->x1[g] + w4 * B->x2[g];
                tau[g] = gup;
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 9: { /* taugb09.cu: joint selection, rfrate rate 4 */
        const double temp = c->coldry[k] * O_CHI(t, 4, jp + 1);
        const double ratn2o = ca[3] / temp;
        const double adjcoln2o = (ratn2o > 1.5) ? (0.5 + om_pow(ratn2o - 0.5, 0.65)) * temp : ca[3];
        if (lower) {
            const Eta e0 = eta(ca[0], rf[3][0], ca[4], 8.0), e1 = eta(ca[0], rf[3][1], ca[4], 8.0);
            const Eta em = eta(ca[0], B->refrat_m_a, ca[4], 8.0), ep = eta(ca[0], B->refrat_planck_a, ca[4], 8.0);
            Sel s0, s1;
            sel_joint(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, &s0, &s1);
            for (int g = 0; g < ng; ++g) {
                const double n2om1 = T3(B->ka_m[0], indm, em.js, 9, g, ng) +
                                     em.fs * (T3(B->ka_m[0], indm, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indm, em.js, 9, g, ng));
                const double n2om2 = T3(B->ka_m[0], indmp, em.js, 9, g, ng) +
                                     em.fs * (T3(B->ka_m[0], indmp, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indmp, em.js, 9, g, ng));
                const double absn2o = n2om1 + mf * (n2om2 - n2om1);
                tau[g] = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10) + e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11) +
                         SELF(g) + FOR(g) + adjcoln2o * absn2o;
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            const int i0 = u0 + 1, i1 = u1 + 1;
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[4] * ABS4(B->absb, i0, i1, g) + adjcoln2o * MIN1(B->kb_m[0], g);
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 10: { /* taugb10.cu */
        if (lower) {
            const int i0 = l0 + 1, i1 = l1 + 1;
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[0] * ABS4(B->absa, i0, i1, g) + SELF(g) + FOR(g);
                fr[g] = B->fracrefa[g];
            }
        } else {
            const int i0 = u0 + 1, i1 = u1 + 1;
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[0] * ABS4(B->absb, i0, i1, g) + FOR(g);
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 11: { /* taugb11.cu */
        const double scaleo2 = ca[5] * c->scaleminor[k];
        if (lower) {
            const int i0 = l0 + 1, i1 = l1 + 1;
            for (int g = 0; g < ng; ++g) {
                const double tauo2 = scaleo2 * MIN1(B->ka_m[0], g);
                tau[g] = ca[0] * ABS4(B->absa, i0, i1, g) + SELF(g) + FOR(g) + tauo2;
                fr[g] = B->fracrefa[g];
            }
        } else {
            const int i0 = u0 + 1, i1 = u1 + 1;
            for (int g = 0; g < ng; ++g) {
                const double tauo2 = scaleo2 * MIN1(B->kb_m[0], g);
                tau[g] = ca[0] * ABS4(B->absb, i0, i1, g) + FOR(g) + tauo2;
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 12: { /* taugb12.cu: upper atmosphere has no absorption */
        if (lower) {
            const Eta e0 = eta(ca[0], rf[0][0], ca[1], 8.0), e1 = eta(ca[0], rf[0][1], ca[1], 8.0);
            const Eta ep = eta(ca[0], B->refrat_planck_a, ca[1], 8.0);
            Sel s0, s1;
            sel_joint(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, &s0, &s1);
            for (int g = 0; g < ng; ++g) {
                tau[g] = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10) + e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11) +
                         SELF(g) + FOR(g);
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            for (int g = 0; g < ng; ++g) tau[g] = 0.0, fr[g] = 0.0;
        }
        break;
    }
    case 13: { /* taugb13.cu: joint selection, rfrate rate 3 */
        if (lower) {
            const double temp = c->coldry[k] * 3.55e-4;
            const double ratco2 = ca[1] / temp;
            const double adjcolco2 = (ratco2 > 3.0) ? (2.0 + om_pow(ratco2 - 2.0, 0.68)) * temp : ca[1];
            const Eta e0 = eta(ca[0], rf[2][0], ca[3], 8.0), e1 = eta(ca[0], rf[2][1], ca[3], 8.0);
            const Eta emc2 = eta(ca[0], B->refrat_m_a, ca[3], 8.0), emc = eta(ca[0], B->refrat_m_a3, ca[3], 8.0);
            const Eta ep = eta(ca[0], B->refrat_planck_a, ca[3], 8.0);
            Sel s0, s1;
            sel_joint(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, &s0, &s1);
            for (int g = 0; g < ng; ++g) {
                const double co2m1 = T3(B->ka_m[0], indm, emc2.js, 9, g, ng) +
                                     emc2.fs * (T3(B->ka_m[0], indm, emc2.js + 1, 9, g, ng) - T3(B->ka_m[0], indm, emc2.js, 9, g, ng));
                const double co2m2 = T3(B->ka_m[0], indmp, emc2.js, 9, g, ng) +
                                     emc2.fs * (T3(B->ka_m[0], indmp, emc2.js + 1, 9, g, ng) - T3(B->ka_m[0], indmp, emc2.js, 9, g, ng));
                const double absco2 = co2m1 + mf * (co2m2 - co2m1);
                const double com1 = T3(B->ka_m[1], indm, emc.js, 9, g, ng) +
                                    emc.fs * (T3(B->ka_m[1], indm, emc.js + 1, 9, g, ng) - T3(B->ka_m[1], indm, emc.js, 9, g, ng));
                const double com2 = T3(B->ka_m[1], indmp, emc.js, 9, g, ng) +
                                    emc.fs * (T3(B->ka_m[1], indmp, emc.js + 1, 9, g, ng) - T3(B->ka_m[1], indmp, emc.js, 9, g, ng));
                const double absco = com1 + mf * (com2 - com1);
                tau[g] = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10) + e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11) +
                         SELF(g) + FOR(g) + adjcolco2 * absco2 + ca[6] * absco;
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[2] * MIN1(B->kb_m[0], g);
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 14: { /* taugb14.cu */
        if (lower) {
            const int i0 = l0 + 1, i1 = l1 + 1;
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[1] * ABS4(B->absa, i0, i1, g) + SELF(g) + FOR(g);
                fr[g] = B->fracrefa[g];
            }
        } else {
            const int i0 = u0 + 1, i1 = u1 + 1;
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[1] * ABS4(B->absb, i0, i1, g);
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    case 15: { /* taugb15.cu: joint selection, rfrate rate 5; upper atmosphere zero */
        if (lower) {
            const double scalen2 = c->colbrd[k] * c->scaleminor[k];
            const Eta e0 = eta(ca[3], rf[4][0], ca[1], 8.0), e1 = eta(ca[3], rf[4][1], ca[1], 8.0);
            const Eta em = eta(ca[3], B->refrat_m_a, ca[1], 8.0), ep = eta(ca[3], B->refrat_planck_a, ca[1], 8.0);
            Sel s0, s1;
            sel_joint(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, &s0, &s1);
            for (int g = 0; g < ng; ++g) {
                const double n2m1 = T3(B->ka_m[0], indm, em.js, 9, g, ng) +
                                    em.fs * (T3(B->ka_m[0], indm, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indm, em.js, 9, g, ng));
                const double n2m2 = T3(B->ka_m[0], indmp, em.js, 9, g, ng) +
                                    em.fs * (T3(B->ka_m[0], indmp, em.js + 1, 9, g, ng) - T3(B->ka_m[0], indmp, em.js, 9, g, ng));
                const double taun2 = scalen2 * (n2m1 + mf * (n2m2 - n2m1));
                tau[g] = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10) + e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11) +
                         SELF(g) + FOR(g) + taun2;
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            for (int g = 0; g < ng; ++g) tau[g] = 0.0, fr[g] = 0.0;
        }
        break;
    }
    case 16: { /* taugb16.cu: joint selection, rfrate rate 4; nspb(16) = 0 upper */
        if (lower) {
            const Eta e0 = eta(ca[0], rf[3][0], ca[4], 8.0), e1 = eta(ca[0], rf[3][1], ca[4], 8.0);
            const Eta ep = eta(ca[0], B->refrat_planck_a, ca[4], 8.0);
            Sel s0, s1;
            sel_joint(e0.specparm, e0.fs, l0 * 9 + e0.js, e1.specparm, e1.fs, l1 * 9 + e1.js, &s0, &s1);
            for (int g = 0; g < ng; ++g) {
                tau[g] = e0.speccomb * maj6(B->absa, ng, g, &s0, f00, f10) + e1.speccomb * maj6(B->absa, ng, g, &s1, f01, f11) +
                         SELF(g) + FOR(g);
                fr[g] = T2(B->fracrefa, ep.js, g, ng) + ep.fs * (T2(B->fracrefa, ep.js + 1, g, ng) - T2(B->fracrefa, ep.js, g, ng));
            }
        } else {
            for (int g = 0; g < ng; ++g) {
                tau[g] = ca[4] * (f00 * T2(B->absb, 1, g, ng) + f10 * T2(B->absb, 2, g, ng) + f01 * T2(B->absb, 1, g, ng) +
                                  f11 * T2(B->absb, 2, g, ng));
                fr[g] = B->fracrefb[g];
            }
        }
        break;
    }
    }
#undef SELF
#undef FOR
#undef MIN1
#undef ABS4
}

/* all bands of one column: taug/fracs [140][nlay], tautot = taug + tauaer (radlw taumol) */
void olw_taumol(const OT *t, const LwCol *c, double (*taug)[O_NL], double (*fracs)[O_NL])
{
    double tau[16], fr[16];
    for (int b = 0; b < 16; ++b)
        for (int k = 0; k < c->nlay; ++k) {
            lw_band(t, c, b, k, tau, fr);
            for (int g = 0; g < t->lw[b].ng; ++g) {
                taug[t->lw[b].ns + g][k] = tau[g] + c->tauaer[b][k];
                fracs[t->lw[b].ns + g][k] = fr[g];
            }
        }
}

/* ---- LW McICA masks: generator of phys_rrtm_lw_mcica_subcol.cu:105-168 with the MT19937
 * stream replaced by Philox4x32-10 keyed (seed, column), counter (column, g | kind << 16, k/4). */
void o_mcica(uint64_t col, uint64_t seed, int kind, int ngpt, int nl, int iovr, const double *cf,
             uint32_t *bits /* [nl][words] */, int words)
{
    for (int k = 0; k < nl; ++k)
        for (int w = 0; w < words; ++w) bits[k * words + w] = 0u;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (int g = 0; g < ngpt; ++g) {
        double prev = 0.0;
        uint32_t out[4] = {0, 0, 0, 0};
        for (int k = 0; k < nl; ++k) {
            if ((k & 3) == 0 && (iovr != 2 || k == 0)) {
                const uint32_t ctr[4] = {(uint32_t)col, (uint32_t)(col >> 32), (uint32_t)g | ((uint32_t)kind << 16),
                                         (uint32_t)(k >> 2)};
                om_philox4x32_10(ctr, key, out);
            }
            const double r = (double)out[iovr == 2 ? 0 : (k & 3)] * 0x1p-32;
            double cd;
            if (iovr == 1) {
                if (k == 0)
                    cd = r;
                else {
                    const double tem1 = 1.0 - cf[k - 1];
                    cd = (prev > tem1) ? prev : r * tem1;
                }
                prev = cd;
            } else {
                cd = r;
            }
            if (cd >= 1.0 - cf[k]) bits[k * words + (g >> 5)] |= 1u << (g & 31);
        }
    }
}

/* ---- rtrnmc (phys_rrtm_lw_rtrnmc.cu:85-310), McICA transfer --------------------------------- */
void olw_rtrnmc(const OT *t, const LwCol *c, double (*taug)[O_NL], double (*fracs)[O_NL], const uint32_t *mask,
                double *totuflux, double *totdflux, double *htr)
{
    const int nl = c->nlay;
    const double bpade = 1.0 / 0.278, tblint = (double)O_NTBL, rec6 = 0.166667;
    const double flxfac = 0.5 * O_FLUXFAC;
    double gassrcu[O_NL], totsrcu[O_NL], trngas[O_NL], efclrfr[O_NL];
    double tur[O_NL + 1], tdr[O_NL + 1], cur[O_NL + 1], cdr[O_NL + 1];
    double uc[O_NL + 1], dc[O_NL + 1];
    for (int k = 0; k <= nl; ++k) totuflux[k] = totdflux[k] = uc[k] = dc[k] = 0.0;
    for (int b = 0; b < 16; ++b) {
        const double semis = c->semiss[b], sdif = c->secdiff[b];
        for (int k = 0; k <= nl; ++k) tur[k] = tdr[k] = cur[k] = cdr[k] = 0.0;
        for (int g = t->lw[b].ns; g < t->lw[b].ns + t->lw[b].ng; ++g) {
            double radtotd = 0.0, radclrd = 0.0;
            for (int k = nl; k >= 1; --k) {
                double odepth = fmax(0.0, sdif * taug[g][k - 1]);
                double atrgas, trng, gasfac;
                if (odepth <= 0.06) {
                    atrgas = odepth - 0.5 * odepth * odepth;
                    trng = 1.0 - atrgas;
                    gasfac = rec6 * odepth;
                } else {
                    const double tblind = odepth / (bpade + odepth);
                    const int itgas = (int)(tblint * tblind + 0.5);
                    trng = t->exp_tbl[itgas];
                    atrgas = 1.0 - trng;
                    gasfac = t->tfn_tbl[itgas];
                    odepth = t->tau_tbl[itgas];
                }
                const double plfrac = fracs[g][k - 1];
                const double blay = c->pklay[k][b];
                const double dplnku = c->pklev[k][b] - blay;
                const double dplnkd = c->pklev[k - 1][b] - blay;
                const double bbdgas = plfrac * (blay + dplnkd * gasfac);
                const double bbugas = plfrac * (blay + dplnku * gasfac);
                const double gassrcd = bbdgas * atrgas;
                gassrcu[k - 1] = bbugas * atrgas;
                trngas[k - 1] = trng;
                const double clfm = ((mask[(k - 1) * 5 + (g >> 5)] >> (g & 31)) & 1u) ? 1.0 : 0.0;
                if (clfm >= O_EPS) {
                    const double odcld = sdif * c->taucld[b][k - 1];
                    efclrfr[k - 1] = 1.0 - (1.0 - om_exp(-odcld)) * clfm;
                    const double odtot = odepth + odcld;
                    double totfac, atrtot;
                    if (odtot < 0.06) {
                        totfac = rec6 * odtot;
                        atrtot = odtot - 0.5 * odtot * odtot;
                    } else {
                        const double tblind = odtot / (bpade + odtot);
                        const int ittot = (int)(tblint * tblind + 0.5);
                        totfac = t->tfn_tbl[ittot];
                        atrtot = 1.0 - t->exp_tbl[ittot];
                    }
                    const double bbdtot = plfrac * (blay + dplnkd * totfac);
                    const double bbutot = plfrac * (blay + dplnku * totfac);
                    const double totsrcd = bbdtot * atrtot;
                    totsrcu[k - 1] = bbutot * atrtot;
                    radtotd = radtotd * trng * efclrfr[k - 1] + gassrcd + clfm * (totsrcd - gassrcd);
                    tdr[k - 1] = tdr[k - 1] + radtotd;
                    radclrd = radclrd * trng + gassrcd;
                    cdr[k - 1] = cdr[k - 1] + radclrd;
                } else {
                    radtotd = radtotd * trng + gassrcd;
                    tdr[k - 1] = tdr[k - 1] + radtotd;
                    radclrd = radclrd * trng + gassrcd;
                    cdr[k - 1] = cdr[k - 1] + radclrd;
                }
            }
            const double reflct = 1.0 - semis;
            const double rad0 = semis * fracs[g][0] * c->pklay[0][b];
            double radtotu = rad0 + reflct * radtotd;
            tur[0] = tur[0] + radtotu;
            double radclru = rad0 + reflct * radclrd;
            cur[0] = cur[0] + radclru;
            for (int k = 1; k <= nl; ++k) {
                const double clfm = ((mask[(k - 1) * 5 + (g >> 5)] >> (g & 31)) & 1u) ? 1.0 : 0.0;
                const double trng = trngas[k - 1];
                const double gasu = gassrcu[k - 1];
                if (clfm > O_EPS) {
                    radtotu = radtotu * trng * efclrfr[k - 1] + gasu + clfm * (totsrcu[k - 1] - gasu);
                } else {
                    radtotu = radtotu * trng + gasu;
                }
                tur[k] = tur[k] + radtotu;
                radclru = radclru * trng + gasu;
                cur[k] = cur[k] + radclru;
            }
        }
        for (int k = 0; k <= nl; ++k) {
            totuflux[k] = totuflux[k] + tur[k];
            totdflux[k] = totdflux[k] + tdr[k];
            uc[k] = uc[k] + cur[k];
            dc[k] = dc[k] + cdr[k];
        }
    }
    for (int k = 0; k <= nl; ++k) {
        totuflux[k] = totuflux[k] * flxfac;
        totdflux[k] = totdflux[k] * flxfac;
    }
    double fnet_prev = totuflux[0] - totdflux[0];
    for (int k = 1; k <= nl; ++k) {
        const double rfdelp = O_HEATFAC / c->delp[k - 1];
        const double fnet = totuflux[k] - totdflux[k];
        htr[k - 1] = (fnet_prev - fnet) * rfdelp;
        fnet_prev = fnet;
    }
}

/* ---- rtrnmr (phys_rrtm_lw_rtrnmr_body.cuh:9-365), max-random overlap (NAM route) ------------- */
void olw_rtrnmr(const OT *t, const LwCol *c, double (*taug)[O_NL], double (*fracs)[O_NL], double *totuflux,
                double *totdflux, double *htr)
{
    const int nl = c->nlay;
    const double eps = 1.0e-6, bpade = 1.0 / 0.278, tblint = (double)O_NTBL, rec6 = 0.166667, c06 = 0.06;
    const double *cf = c->cldfrc; /* 0..nl+1 */
    double faccld1u[O_NL + 3] = {0}, faccld2u[O_NL + 3] = {0}, facclr1u[O_NL + 3] = {0}, facclr2u[O_NL + 3] = {0},
           faccmb1u[O_NL + 3] = {0}, faccmb2u[O_NL + 3] = {0};
    int lstu[O_NL + 3] = {0}, lstd[O_NL + 3] = {0};
    lstu[1] = (cf[1] > eps) ? 1 : 0;
    double rat1 = 0.0, rat2 = 0.0;
    for (int k = 1; k <= nl - 1; ++k) {
        lstu[k + 1] = (cf[k + 1] > eps && cf[k] <= eps) ? 1 : 0;
        if (cf[k] > eps) {
            if (cf[k + 1] >= cf[k]) {
                if (lstu[k]) {
                    if (cf[k] < 1.0) facclr2u[k + 1] = (cf[k + 1] - cf[k]) / (1.0 - cf[k]);
                    facclr2u[k] = 0.0;
                    faccld2u[k] = 0.0;
                } else {
                    const double fmx = fmax(cf[k], cf[k - 1]);
                    if (cf[k + 1] > fmx) {
                        facclr1u[k + 1] = rat2;
                        facclr2u[k + 1] = (cf[k + 1] - fmx) / (1.0 - fmx);
                    } else if (cf[k + 1] < fmx) {
                        facclr1u[k + 1] = (cf[k + 1] - cf[k]) / (cf[k - 1] - cf[k]);
                    } else {
                        facclr1u[k + 1] = rat2;
                    }
                }
                if (facclr1u[k + 1] > 0.0 || facclr2u[k + 1] > 0.0) rat1 = 1.0, rat2 = 0.0;
                else rat1 = 0.0, rat2 = 0.0;
            } else {
                if (lstu[k]) {
                    faccld2u[k + 1] = (cf[k] - cf[k + 1]) / cf[k];
                    facclr2u[k] = 0.0;
                    faccld2u[k] = 0.0;
                } else {
                    const double fmn = fmin(cf[k], cf[k - 1]);
                    if (cf[k + 1] <= fmn) {
                        faccld1u[k + 1] = rat1;
                        faccld2u[k + 1] = (fmn - cf[k + 1]) / fmn;
                    } else {
                        faccld1u[k + 1] = (cf[k] - cf[k + 1]) / (cf[k] - fmn);
                    }
                }
                if (faccld1u[k + 1] > 0.0 || faccld2u[k + 1] > 0.0) rat1 = 0.0, rat2 = 1.0;
                else rat1 = 0.0, rat2 = 0.0;
            }
            faccmb1u[k + 1] = facclr1u[k + 1] * faccld2u[k] * cf[k - 1];
            faccmb2u[k + 1] = faccld1u[k + 1] * facclr2u[k] * (1.0 - cf[k - 1]);
        }
    }
    double faccld1d[O_NL + 3] = {0}, faccld2d[O_NL + 3] = {0}, facclr1d[O_NL + 3] = {0}, facclr2d[O_NL + 3] = {0},
           faccmb1d[O_NL + 3] = {0}, faccmb2d[O_NL + 3] = {0};
    lstd[nl] = (cf[nl] > eps) ? 1 : 0;
    rat1 = 0.0, rat2 = 0.0;
    for (int k = nl; k >= 2; --k) {
        lstd[k - 1] = (cf[k - 1] > eps && cf[k] <= eps) ? 1 : 0;
        if (cf[k] > eps) {
            if (cf[k - 1] >= cf[k]) {
                if (lstd[k]) {
                    if (cf[k] < 1.0) facclr2d[k - 1] = (cf[k - 1] - cf[k]) / (1.0 - cf[k]);
                    facclr2d[k] = 0.0;
                    faccld2d[k] = 0.0;
                } else {
                    const double fmx = fmax(cf[k], cf[k + 1]);
                    if (cf[k - 1] > fmx) {
                        facclr1d[k - 1] = rat2;
                        facclr2d[k - 1] = (cf[k - 1] - fmx) / (1.0 - fmx);
                    } else if (cf[k - 1] < fmx) {
                        facclr1d[k - 1] = (cf[k - 1] - cf[k]) / (cf[k + 1] - cf[k]);
                    } else {
                        facclr1d[k - 1] = rat2;
                    }
                }
                if (facclr1d[k - 1] > 0.0 || facclr2d[k - 1] > 0.0) rat1 = 1.0, rat2 = 0.0;
                else rat1 = 0.0, rat2 = 0.0;
            } else {
                if (lstd[k]) {
                    faccld2d[k - 1] = (cf[k] - cf[k - 1]) / cf[k];
                    facclr2d[k] = 0.0;
                    faccld2d[k] = 0.0;
                } else {
                    const double fmn = fmin(cf[k], cf[k + 1]);
                    if (cf[k - 1] <= fmn) {
                        faccld1d[k - 1] = rat1;
                        faccld2d[k - 1] = (fmn - cf[k - 1]) / fmn;
                    } else {
                        faccld1d[k - 1] = (cf[k] - cf[k - 1]) / (cf[k] - fmn);
                    }
                }
                if (faccld1d[k - 1] > 0.0 || faccld2d[k - 1] > 0.0) rat1 = 0.0, rat2 = 1.0;
                else rat1 = 0.0, rat2 = 0.0;
            }
            faccmb1d[k - 1] = facclr1d[k - 1] * faccld2d[k] * cf[k + 1];
            faccmb2d[k - 1] = faccld1d[k - 1] * facclr2d[k] * (1.0 - cf[k + 1]);
        }
    }
    const double flxfac = 0.5 * O_FLUXFAC;
    double tuf[O_NL + 2] = {0}, tdf[O_NL + 2] = {0};
    double toturad[O_NL + 2], totdrad[O_NL + 2];
    double totsrcu[O_NL + 2], trntot[O_NL + 2], gassrcu[O_NL + 2], trngas[O_NL + 2];
    double totradd = 0.0, clrradd = 0.0, rad = 0.0, totradu = 0.0, clrradu = 0.0;
    for (int b = 0; b < 16; ++b) {
        for (int k = 0; k <= nl; ++k) toturad[k] = totdrad[k] = 0.0;
        const double semi = c->semiss[b], sdif = c->secdiff[b];
        for (int g = t->lw[b].ns; g < t->lw[b].ns + t->lw[b].ng; ++g) {
            double radtotd = 0.0, radclrd = 0.0;
            for (int k = nl; k >= 1; --k) {
                double odepth = fmax(0.0, sdif * taug[g][k - 1]);
                double atrgas, trng, gasfac;
                if (odepth <= c06) {
                    atrgas = odepth - 0.5 * odepth * odepth;
                    trng = 1.0 - atrgas;
                    gasfac = rec6 * odepth;
                } else {
                    const double tblind = odepth / (bpade + odepth);
                    const int itgas = (int)(tblint * tblind + 0.5);
                    trng = t->exp_tbl[itgas];
                    atrgas = 1.0 - trng;
                    gasfac = t->tfn_tbl[itgas];
                    odepth = t->tau_tbl[itgas];
                }
                const double plfrac = fracs[g][k - 1];
                const double blay = c->pklay[k][b];
                const double dplnku = c->pklev[k][b] - blay;
                const double dplnkd = c->pklev[k - 1][b] - blay;
                const double bbdgas = plfrac * (blay + dplnkd * gasfac);
                const double bbugas = plfrac * (blay + dplnku * gasfac);
                const double gassrcd = bbdgas * atrgas;
                gassrcu[k] = bbugas * atrgas;
                trngas[k] = trng;
                const double clfr = cf[k];
                if (lstd[k]) {
                    totradd = clfr * radtotd;
                    clrradd = radtotd - totradd;
                    rad = 0.0;
                }
                if (clfr >= eps) {
                    const double odcld = sdif * c->taucld[b][k - 1];
                    const double odtot = odepth + odcld;
                    double totfac, atrtot, trnt;
                    if (odtot < c06) {
                        totfac = rec6 * odtot;
                        atrtot = odtot - 0.5 * odtot * odtot;
                        trnt = 1.0 - atrtot;
                    } else {
                        const double tblind = odtot / (bpade + odtot);
                        const int ittot = (int)(tblint * tblind + 0.5);
                        totfac = t->tfn_tbl[ittot];
                        trnt = t->exp_tbl[ittot];
                        atrtot = 1.0 - trnt;
                    }
                    const double bbdtot = plfrac * (blay + dplnkd * totfac);
                    const double bbutot = plfrac * (blay + dplnku * totfac);
                    const double totsrcd = bbdtot * atrtot;
                    totsrcu[k] = bbutot * atrtot;
                    trntot[k] = trnt;
                    totradd = totradd * trnt + clfr * totsrcd;
                    clrradd = clrradd * trng + (1.0 - clfr) * gassrcd;
                    radtotd = totradd + clrradd;
                    totdrad[k - 1] = totdrad[k - 1] + radtotd;
                    radclrd = radclrd * trng + gassrcd;
                    const double radmod = rad * (facclr1d[k - 1] * trng + faccld1d[k - 1] * trnt) - faccmb1d[k - 1] * gassrcd +
                                          faccmb2d[k - 1] * totsrcd;
                    rad = -radmod + facclr2d[k - 1] * (clrradd + radmod) - faccld2d[k - 1] * (totradd - radmod);
                    totradd = totradd + rad;
                    clrradd = clrradd - rad;
                } else {
                    radtotd = radtotd * trng + gassrcd;
                    totdrad[k - 1] = totdrad[k - 1] + radtotd;
                    radclrd = radclrd * trng + gassrcd;
                }
            }
            const double reflct = 1.0 - semi;
            const double rad0 = semi * fracs[g][0] * c->pklay[0][b];
            double radtotu = rad0 + reflct * radtotd;
            toturad[0] = toturad[0] + radtotu;
            double radclru = rad0 + reflct * radclrd;
            for (int k = 1; k <= nl; ++k) {
                const double clfr = cf[k];
                const double trng = trngas[k];
                const double gasu = gassrcu[k];
                if (lstu[k]) {
                    totradu = clfr * radtotu;
                    clrradu = radtotu - totradu;
                    rad = 0.0;
                }
                if (clfr >= eps) {
                    const double trnt = trntot[k];
                    const double totu = totsrcu[k];
                    totradu = totradu * trnt + clfr * totu;
                    clrradu = clrradu * trng + (1.0 - clfr) * gasu;
                    radtotu = totradu + clrradu;
                    toturad[k] = toturad[k] + radtotu;
                    radclru = radclru * trng + gasu;
                    const double radmod = rad * (facclr1u[k + 1] * trng + faccld1u[k + 1] * trnt) - faccmb1u[k + 1] * gasu +
                                          faccmb2u[k + 1] * totu;
                    rad = -radmod + facclr2u[k + 1] * (clrradu + radmod) - faccld2u[k + 1] * (totradu - radmod);
                    totradu = totradu + rad;
                    clrradu = clrradu - rad;
                } else {
                    radtotu = radtotu * trng + gasu;
                    toturad[k] = toturad[k] + radtotu;
                    radclru = radclru * trng + gasu;
                }
            }
            (void)radclru;
        }
        for (int k = 0; k <= nl; ++k) {
            tuf[k] = tuf[k] + toturad[k];
            tdf[k] = tdf[k] + totdrad[k];
        }
    }
    for (int k = 0; k <= nl; ++k) {
        totuflux[k] = tuf[k] * flxfac;
        totdflux[k] = tdf[k] * flxfac;
    }
    double fprev = totuflux[0] - totdflux[0];
    for (int k = 1; k <= nl; ++k) {
        const double fn = totuflux[k] - totdflux[k];
        htr[k - 1] = (fprev - fn) * (O_HEATFAC / c->delp[k - 1]);
        fprev = fn;
    }
}

/* one band of one layer from a hand-built setcoef record (the oracle's own pins, tests) */
void olw_band(const OT *t, const LwCol *c, int b, int k, double *tau, double *fr) { lw_band(t, c, b, k, tau, fr); }
