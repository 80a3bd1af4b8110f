/* oracle_aer.c -- CPU restatement of the climatological aerosol optics of the radiation call (TEST
 * INFRASTRUCTURE ONLY): setaer's prep, grrad setup's layer relative humidity and aer_property with
 * radclimaer, in the NAM's setting (layers from the top, ivflip = 0; SW and LW bands; no volcanic
 * aerosol).  FP64 on the float32 tables with the library's own exp/log/pow, one column at a time in
 * the reference's loop order (whole-column arrays, the smoothing done in place), so that the
 * library's wr_aer_optics agrees with it bit for bit. */
#include "oracle.h"

#define A_NDM 5
#define A_IMXAE 72
#define A_JMXAE 37
#define A_NXC 5
#define A_NRHLEV 8
#define A_NCM1 6
#define A_NCM 10
#define A_NSW 14
#define A_NLW 16
#define A_NB 30
#define A_NL 64

/* physcons of setaer (phys_rrtm_setaer.cu:38-43) and grrad setup (phys_rrtm_grrad_setup.cu:3-24) */
#define A_PI 3.1415926535897931
#define A_RD 287.05
#define A_RV 461.5
#define A_G 9.80665
#define A_FVIRT (A_RV / A_RD - 1.0)
#define A_EPS (A_RD / A_RV)
#define A_EPSM1 (A_RD / A_RV - 1.0)
#define A_QMIN 1.0e-10
#define A_QME5 1.0e-7
#define A_PRSMIN 1.0e-6

/* GFS fpvsx (phys_rrtm_grrad_setup.cu:39-60): saturation vapour pressure, Pa, over liquid above the
 * triple point, over ice below 20 K under it, blended linearly between */
static double a_fpvsx(double t)
{
    const double ttp = 2.7316e+2, psat = 6.1078e+2, cvap = 1.8460e+3, cliq = 4.1855e+3, csol = 2.1060e+3,
                 hvap = 2.5000e+6, hfus = 3.3358e+5;
    const double tliq = ttp, tice = ttp - 20.0;
    const double dldtl = cvap - cliq, heatl = hvap;
    const double xponal = -dldtl / A_RV, xponbl = -dldtl / A_RV + heatl / (A_RV * ttp);
    const double dldti = cvap - csol, heati = hvap + hfus;
    const double xponai = -dldti / A_RV, xponbi = -dldti / A_RV + heati / (A_RV * ttp);
    const double tr = ttp / t;
    if (t >= tliq) return psat * om_pow(tr, xponal) * om_exp(xponbl * (1.0 - tr));
    if (t < tice) return psat * om_pow(tr, xponai) * om_exp(xponbi * (1.0 - tr));
    const double w = (t - tice) / (tliq - tice);
    const double pvl = psat * om_pow(tr, xponal) * om_exp(xponbl * (1.0 - tr));
    const double pvi = psat * om_pow(tr, xponai) * om_exp(xponbi * (1.0 - tr));
    return w * pvl + (1.0 - w) * pvi;
}

/* GFS fpvs (phys_rrtm_grrad_setup.cu:62-71): the 7501-node fpvsx table over 180..330 K, interpolated
 * linearly (its two nodes evaluated here), the index clamped to the table */
static double a_fpvs(double t)
{
    const double xmin = 180.0, nx = 7501.0, xinc = (330.0 - 180.0) / (nx - 1.0);
    const double c2 = 1.0 / xinc, c1 = 1.0 - xmin * c2;
    const double xj = fmin(fmax(c1 + c2 * t, 1.0), nx);
    const int jx = (int)fmin(xj, nx - 1.0);
    const double f0 = a_fpvsx(xmin + (double)(jx - 1) * xinc), f1 = a_fpvsx(xmin + (double)jx * xinc);
    return f0 + (xj - (double)jx) * (f1 - f0);
}

typedef struct {
    int nrl;                                   /* radiation layers (model layers + 1)        */
    double pl[A_NL + 1];                       /* levels, mb (setaer's prsi, top first)       */
    double dz[A_NL], hz[A_NL + 1], rh[A_NL];   /* thickness, height (km), relative humidity   */
    int idom[A_NL];                            /* aerosol domain of each layer                */
    double delz[A_NL];                         /* effective thickness                         */
    int ih1[A_NL], ih2[A_NL];
    double rdrh[A_NL];
    int kp;
    double denn[3], cmix[A_NCM + 1];
} AerCol;

/* table accessors, 1-based as the reference's Fortran indices */
#define HAER(idm, kp) ((double)tab->haer[((kp) - 1) * A_NDM + (idm) - 1])
#define PRSREF(idm, kp) ((double)tab->prsref[((kp) - 1) * A_NDM + (idm) - 1])
#define SIGREF(idm, kp) ((double)tab->sigref[((kp) - 1) * A_NDM + (idm) - 1])
#define RHLEV(ir) ((double)tab->rhlev[(ir) - 1])
#define RHI(a, ic, ib) ((double)tab->a[((ib) - 1) * A_NCM1 + (ic) - 1])
#define RHD(a, ir, ic1, ib) ((double)tab->a[(((ib) - 1) * 4 + (ic1) - 1) * A_NRHLEV + (ir) - 1])
#define CMIXG(m, i, j) ((double)tab->cmixg[(((j) - 1) * A_IMXAE + (i) - 1) * A_NXC + (m) - 1])
#define IDXCG(m, i, j) (tab->idxcg[(((j) - 1) * A_IMXAE + (i) - 1) * A_NXC + (m) - 1])
#define DENNG(m, i, j) ((double)tab->denng[(((j) - 1) * A_IMXAE + (i) - 1) * 2 + (m) - 1])
/* profile type of a cell; a valid table holds 1..7, anything else is clamped (as the library does) */
static int a_kprf(const wr_aer_tables *tab, int i, int j)
{
    const int k = tab->kprfg[(j - 1) * A_IMXAE + i - 1];
    return k < 1 ? 1 : (k > 7 ? 7 : k);
}
#define KPRFG(i, j) a_kprf(tab, i, j)

static void aer_column(const wr_aer_tables *tab, int ncol, int nlay, const double *p_lev, const double *t_lay,
                       const double *q_lay, double xlon, double xlat, double slm, int c, AerCol *a)
{
    const int nrl = nlay + 1;
    a->nrl = nrl;
    /* grrad setup's levels (phys_rrtm_grrad_setup.cu:147-190): mb, the added top level at PRSMIN */
    a->pl[0] = A_PRSMIN;
    for (int l = 1; l <= nrl; ++l) a->pl[l] = 10.0 * (p_lev[(size_t)(l - 1) * ncol + c] * 1.0e-3);
    if (a->pl[1] <= A_PRSMIN) a->pl[1] = 2.0 * A_PRSMIN;
    /* layer temperature, moisture (QME5 floor) and relative humidity; the added layer takes layer 1's */
    double tl[A_NL] = {0.0}, ql[A_NL] = {0.0};
    for (int l = 1; l < nrl; ++l) {
        const size_t s = (size_t)(l - 1) * ncol + c;
        const double t = t_lay[s], q = q_lay[s];
        tl[l] = t;
        ql[l] = fmax(A_QME5, q);
        const double pl_cb = 0.5 * (p_lev[s] + p_lev[s + ncol]) * 1.0e-3;
        const double es = fmin(pl_cb, 0.001 * a_fpvs(t));
        const double qs = fmax(A_QMIN, A_EPS * es / (pl_cb + A_EPSM1 * es));
        a->rh[l] = fmax(0.0, fmin(1.0, fmax(A_QMIN, q) / qs));
    }
    tl[0] = tl[1], ql[0] = ql[1], a->rh[0] = a->rh[1];
    /* setaer prep, ivflip = 0 (phys_rrtm_setaer.cu:140-153): hypsometric thickness, the top layer from
     * its mid pressure 0.5 pl(1) (grrad setup's plyr), doubled; heights summed from the surface */
    const double rovg = 0.001 * A_RD / A_G;
    for (int l = 0; l < nrl; ++l) {
        const double lnk = (l == 0) ? om_log(0.5 * a->pl[1]) : om_log(a->pl[l]);
        const double lnk1 = om_log(a->pl[l + 1]);
        a->dz[l] = rovg * (lnk1 - lnk) * tl[l] * (1.0 + A_FVIRT * ql[l]);
    }
    a->dz[0] = 2.0 * a->dz[0];
    a->hz[nrl] = 0.0;
    for (int l = nrl - 1; l >= 0; --l) a->hz[l] = a->hz[l + 1] + a->dz[l];

    /* aer_property (phys_rrtm_aer_property.cu:126-223): the grid cell around the column */
    const double rdg = 180.0 / A_PI;
    double alon = xlon * rdg;
    if (alon < 0.0) alon = alon + 360.0;
    double alat = xlat * rdg;
    if (alat < -90.0) alat = -90.0;
    if (alat > 90.0) alat = 90.0;
    const double dltg = 360.0 / (double)A_IMXAE, hdlt = 0.5 * dltg, rdlt = 1.0 / dltg;
    int i3 = 1;
    double dtmp = alon;
    if (dtmp < 0.0) dtmp = 0.0;
    while (dtmp > dltg && i3 < A_IMXAE) {
        ++i3;
        dtmp = alon - dltg * (double)(i3 - 1);
    }
    const int i1 = i3, i2 = i3 % A_IMXAE + 1;
    const double wi = dtmp * rdlt;
    const int kpi = (dtmp <= hdlt) ? i3 : i2;
    int j3 = 1;
    double dtl = 90.0 - alat;
    while (dtl > dltg && j3 < A_JMXAE) {
        ++j3;
        dtl = (90.0 - dltg * (double)(j3 - 1)) - alat;
    }
    const int j1 = j3, j2 = (j3 + 1 > A_JMXAE) ? A_JMXAE : j3 + 1;
    const double wj = dtl * rdlt;
    const int kpj = (dtl <= hdlt) ? j3 : j2;

    int kp = KPRFG(kpi, kpj);
    const int kpa = imax(imax(KPRFG(i1, j1), KPRFG(i1, j2)), imax(KPRFG(i2, j1), KPRFG(i2, j2)));
    double h1 = HAER(1, kp);
    int ii = 1;
    if (kp != kpa) {
        if (kpa == 6) {
            ii = 2;
            if (slm > 0.0) {
                kp = 7;
                h1 = 0.5 * (HAER(1, 6) + HAER(1, 7));
            } else {
                kp = kpa;
                h1 = HAER(1, 6);
            }
        } else if (kpa == 7) {
            ii = 2;
            if (slm <= 0.0) {
                kp = 6;
                h1 = 0.5 * (HAER(1, 6) + HAER(1, 7));
            } else {
                kp = kpa;
                h1 = HAER(1, 7);
            }
        } else {
            h1 = HAER(1, kpa);
            kp = kpa;
        }
    }
    a->kp = kp;
    const double w11 = (1.0 - wi) * (1.0 - wj), w12 = (1.0 - wi) * wj, w21 = wi * (1.0 - wj), w22 = wi * wj;
    a->denn[0] = a->denn[1] = a->denn[2] = 0.0;
    for (int m = 1; m <= ii; ++m)
        a->denn[m] = w11 * DENNG(m, i1, j1) + w12 * DENNG(m, i1, j2) + w21 * DENNG(m, i2, j1) + w22 * DENNG(m, i2, j2);
    for (int n = 0; n <= A_NCM; ++n) a->cmix[n] = 0.0;
    for (int m = 1; m <= A_NXC; ++m) {
        int ic = IDXCG(m, i1, j1);
        if (ic > 0 && ic <= A_NCM) a->cmix[ic] = a->cmix[ic] + w11 * CMIXG(m, i1, j1);
        ic = IDXCG(m, i1, j2);
        if (ic > 0 && ic <= A_NCM) a->cmix[ic] = a->cmix[ic] + w12 * CMIXG(m, i1, j2);
        ic = IDXCG(m, i2, j1);
        if (ic > 0 && ic <= A_NCM) a->cmix[ic] = a->cmix[ic] + w21 * CMIXG(m, i2, j1);
        ic = IDXCG(m, i2, j2);
        if (ic > 0 && ic <= A_NCM) a->cmix[ic] = a->cmix[ic] + w22 * CMIXG(m, i2, j2);
    }

    /* domains from the surface up (phys_rrtm_aer_property.cu:251-267, ivflip = 0) */
    const double rps = 1.0 / a->pl[nrl];
    ii = 1;
    for (int l = nrl - 1; l >= 0; --l) {
        if (ii <= A_NDM && a->pl[l] * rps < SIGREF(ii, kp)) {
            ii = ii + 1;
            if (ii == 2 && PRSREF(2, kp) == PRSREF(3, kp)) ii = 3;
        }
        a->idom[l] = ii;
        const double tmp1 = (ii > 1) ? ((ii <= A_NDM) ? HAER(ii, kp) : 0.0) : h1;
        if (tmp1 > 0.0) {
            const double tmp2 = 1.0 / tmp1;
            a->delz[l] = tmp1 * (om_exp(-a->hz[l + 1] * tmp2) - om_exp(-a->hz[l] * tmp2));
        } else {
            a->delz[l] = a->dz[l];
        }
    }
    /* radclimaer's humidity bracket (phys_rrtm_aer_property.cu:271-282) */
    for (int l = 0; l < nrl; ++l) {
        const double rh1 = a->rh[l];
        int ih2 = 1;
        while (ih2 <= A_NRHLEV && rh1 > RHLEV(ih2)) ++ih2;
        const int ih1 = imax(1, ih2 - 1);
        ih2 = imin(A_NRHLEV, ih2);
        const double drh0 = RHLEV(ih2) - RHLEV(ih1), drh1 = rh1 - RHLEV(ih1);
        a->rdrh[l] = (ih1 == ih2) ? 0.0 : drh1 / drh0;
        a->ih1[l] = ih1, a->ih2[l] = ih2;
    }
}

/* radclimaer, one band over the column (phys_rrtm_aer_property.cu:293-400): tau, ssa, asy per layer,
 * then the smoothing from the top down (ivflip = 0) */
static void aer_band(const wr_aer_tables *tab, const AerCol *a, int ib, double *tau, double *ssa, double *asy)
{
    const int nrl = a->nrl;
    for (int l = 0; l < nrl; ++l) {
        const int idom = a->idom[l], ih1 = a->ih1[l], ih2 = a->ih2[l];
        const double rdrh = a->rdrh[l], dzk = a->delz[l];
        if (idom == 5 || idom == 4) {
            tau[l] = (idom == 5) ? 0.0 : (double)tab->extstra[ib - 1] * dzk;
            if (ib <= A_NSW) ssa[l] = 0.99, asy[l] = 0.696;
            else ssa[l] = 0.5, asy[l] = 0.3;
        } else if (idom == 3) {
            const double ex01 = RHI(extrhi, 1, ib), sc01 = RHI(scarhi, 1, ib), ss01 = RHI(ssarhi, 1, ib),
                         as01 = RHI(asyrhi, 1, ib);
            const double ex02 = RHI(extrhi, 2, ib), sc02 = RHI(scarhi, 2, ib), ss02 = RHI(ssarhi, 2, ib),
                         as02 = RHI(asyrhi, 2, ib);
            const double ex03 = RHD(extrhd, ih1, 1, ib) + rdrh * (RHD(extrhd, ih2, 1, ib) - RHD(extrhd, ih1, 1, ib));
            const double sc03 = RHD(scarhd, ih1, 1, ib) + rdrh * (RHD(scarhd, ih2, 1, ib) - RHD(scarhd, ih1, 1, ib));
            const double ss03 = RHD(ssarhd, ih1, 1, ib) + rdrh * (RHD(ssarhd, ih2, 1, ib) - RHD(ssarhd, ih1, 1, ib));
            const double as03 = RHD(asyrhd, ih1, 1, ib) + rdrh * (RHD(asyrhd, ih2, 1, ib) - RHD(asyrhd, ih1, 1, ib));
            const double ext1 = 0.17e-3 * ex01 + 0.4 * ex02 + 0.59983 * ex03;
            const double sca1 = 0.17e-3 * sc01 + 0.4 * sc02 + 0.59983 * sc03;
            const double ssa1 = 0.17e-3 * ss01 * ex01 + 0.4 * ss02 * ex02 + 0.59983 * ss03 * ex03;
            const double asy1 = 0.17e-3 * as01 * sc01 + 0.4 * as02 * sc02 + 0.59983 * as03 * sc03;
            tau[l] = ext1 * 730.0 * dzk;
            ssa[l] = (ext1 > 0.0) ? fmin(1.0, ssa1 / ext1) : 1.0;
            asy[l] = (sca1 > 0.0) ? fmin(1.0, asy1 / sca1) : 0.0;
        } else if (idom == 1) {
            double ext1 = 0.0, sca1 = 0.0, ssa1 = 0.0, asy1 = 0.0;
            for (int icmp = 1; icmp <= A_NCM; ++icmp) {
                const double cm = a->cmix[icmp];
                if (!(cm > 0.0)) continue;
                double ex, sc, ss, as;
                if (icmp <= A_NCM1) {
                    ex = RHI(extrhi, icmp, ib), sc = RHI(scarhi, icmp, ib), ss = RHI(ssarhi, icmp, ib);
                    as = RHI(asyrhi, icmp, ib);
                } else {
                    const int ic1 = icmp - A_NCM1;
                    ex = RHD(extrhd, ih1, ic1, ib) + rdrh * (RHD(extrhd, ih2, ic1, ib) - RHD(extrhd, ih1, ic1, ib));
                    sc = RHD(scarhd, ih1, ic1, ib) + rdrh * (RHD(scarhd, ih2, ic1, ib) - RHD(scarhd, ih1, ic1, ib));
                    ss = RHD(ssarhd, ih1, ic1, ib) + rdrh * (RHD(ssarhd, ih2, ic1, ib) - RHD(ssarhd, ih1, ic1, ib));
                    as = RHD(asyrhd, ih1, ic1, ib) + rdrh * (RHD(asyrhd, ih2, ic1, ib) - RHD(asyrhd, ih1, ic1, ib));
                }
                ext1 = ext1 + cm * ex;
                sca1 = sca1 + cm * sc;
                ssa1 = ssa1 + cm * ss * ex;
                asy1 = asy1 + cm * as * sc;
            }
            tau[l] = ext1 * a->denn[1] * dzk;
            ssa[l] = (ext1 > 0.0) ? fmin(1.0, ssa1 / ext1) : 1.0;
            asy[l] = (sca1 > 0.0) ? fmin(1.0, asy1 / sca1) : 0.0;
        } else if (idom == 2) {
            tau[l] = RHI(extrhi, 6, ib) * a->denn[2] * dzk;
            ssa[l] = RHI(ssarhi, 6, ib);
            asy[l] = RHI(asyrhi, 6, ib);
        } else {
            tau[l] = 0.0, ssa[l] = 1.0, asy[l] = 0.0;
        }
    }
    const double crt1 = 30.0, crt2 = 0.03333;
    for (int l = 1; l < nrl; ++l) {
        const double ratio = (tau[l] > 0.0) ? tau[l - 1] / tau[l] : 1.0;
        const double tt0 = tau[l] + tau[l - 1], tt1 = 0.2 * tt0, tt2 = tt0 - tt1;
        if (ratio > crt1) tau[l] = tt1, tau[l - 1] = tt2;
        if (ratio < crt2) tau[l] = tt2, tau[l - 1] = tt1;
    }
}

int oracle_aer_optics(const wr_aer_tables *tab, int ncol, int nlay, const double *p_lev, const double *t_lay,
                      const double *q_lay, const float *xlon, const float *xlat, const float *slmsk, float *aer_sw,
                      float *aer_lw, int64_t ld)
{
    if (!tab || ncol < 0 || nlay < 1 || nlay + 1 > A_NL || ld < ncol) return 1;
    const int nrl = nlay + 1;
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
    for (int c = 0; c < ncol; ++c) {
        AerCol a;
        aer_column(tab, ncol, nlay, p_lev, t_lay, q_lay, (double)xlon[c], (double)xlat[c], (double)slmsk[c], c, &a);
        double tau[A_NL], ssa[A_NL], asy[A_NL];
        for (int ib = 1; ib <= A_NB; ++ib) {
            aer_band(tab, &a, ib, tau, ssa, asy);
            float *o = (ib <= A_NSW) ? aer_sw : aer_lw;
            const int nb = (ib <= A_NSW) ? A_NSW : A_NLW, b = (ib <= A_NSW) ? ib - 1 : ib - 1 - A_NSW;
            for (int l = 0; l < nrl; ++l) {
                o[(((size_t)0 * nb + b) * nrl + l) * ld + c] = (float)tau[l];
                o[(((size_t)1 * nb + b) * nrl + l) * ld + c] = (float)ssa[l];
                o[(((size_t)2 * nb + b) * nrl + l) * ld + c] = (float)asy[l];
            }
        }
    }
    return bad;
}

/* One column's intermediate fields (for the oracle's own checks): dz, hz (nrl+1), rh, delz, rdrh as
 * doubles; idom, ih1 as ints; cmix[11], denn[3]; returns the profile type kp. */
int oracle_aer_column(const wr_aer_tables *tab, int ncol, int nlay, const double *p_lev, const double *t_lay,
                      const double *q_lay, double xlon, double xlat, double slmsk, int c, double *dz, double *hz,
                      double *rh, double *delz, double *rdrh, int *idom, int *ih1, double *cmix, double *denn)
{
    if (!tab || c < 0 || c >= ncol || nlay < 1 || nlay + 1 > A_NL) return -1;
    AerCol a;
    aer_column(tab, ncol, nlay, p_lev, t_lay, q_lay, xlon, xlat, slmsk, c, &a);
    for (int l = 0; l < a.nrl; ++l) {
        dz[l] = a.dz[l], rh[l] = a.rh[l], delz[l] = a.delz[l], rdrh[l] = a.rdrh[l];
        idom[l] = a.idom[l], ih1[l] = a.ih1[l];
    }
    for (int l = 0; l <= a.nrl; ++l) hz[l] = a.hz[l];
    for (int n = 0; n <= A_NCM; ++n) cmix[n] = a.cmix[n];
    for (int n = 0; n < 3; ++n) denn[n] = a.denn[n];
    return a.kp;
}

double oracle_aer_fpvs(double t) { return a_fpvs(t); }
