/* oracle_pins.c -- entry points that expose single pieces of the CPU oracle to its own pins.
 *
 * TEST INFRASTRUCTURE ONLY (tests/test_oracle_pins*.py).  Each entry point runs exactly the oracle
 * code that oracle_run uses, on one hand-built layer or column, so that the tests can put the
 * pieces on table nodes and mathematical limits:
 *   oracle_lw_setcoef_layer / oracle_sw_setcoef_layer -- setcoef of one layer -> a flat record;
 *   oracle_lw_band_layer / oracle_sw_band_layer       -- taugbNN / taumolNN of one layer from a record;
 *   oracle_lw_cldprop_layer / oracle_sw_cldprop_layer -- cloud optics of one cloudy layer;
 *   oracle_sw_expm, oracle_sw_layer_opt, oracle_sw_gpoint -- the spcvrtc pieces.
 * Record layouts are the enums below (doubles; integers stored as doubles).
 */
#include "oracle.h"

enum { LR_LOWER, LR_JP, LR_JT, LR_JT1, LR_INDS, LR_INDF, LR_INDM, LR_F00, LR_F01, LR_F10, LR_F11, LR_SF, LR_SFR, LR_FF,
       LR_FFR, LR_MF, LR_SCM, LR_SCMN2, LR_CA = 18, LR_WX = 25, LR_COLDRY = 29, LR_COLBRD = 30, LR_PAVEL = 31,
       LR_RF = 32, LR_N = 44 };
enum { SR_LAYTROP, SR_JP, SR_JT, SR_JT1, SR_INDS, SR_INDF, SR_F00, SR_F01, SR_F10, SR_F11, SR_SF, SR_SFR, SR_FF, SR_FFR,
       SR_CA = 14, SR_COLMOL = 20, SR_N = 21 };

int oracle_rec_sizes(int which) { return which == 0 ? LR_N : SR_N; }

static void lw_rec_out(const LwCol *c, double *r)
{
    r[LR_LOWER] = c->laytrop > 0;
    r[LR_JP] = c->jp[0], r[LR_JT] = c->jt[0], r[LR_JT1] = c->jt1[0];
    r[LR_INDS] = c->indself[0], r[LR_INDF] = c->indfor[0], r[LR_INDM] = c->indminor[0];
    r[LR_F00] = c->fac00[0], r[LR_F01] = c->fac01[0], r[LR_F10] = c->fac10[0], r[LR_F11] = c->fac11[0];
    r[LR_SF] = c->selffac[0], r[LR_SFR] = c->selffrac[0], r[LR_FF] = c->forfac[0], r[LR_FFR] = c->forfrac[0];
    r[LR_MF] = c->minorfrac[0], r[LR_SCM] = c->scaleminor[0], r[LR_SCMN2] = c->scaleminorn2[0];
    for (int m = 0; m < 7; ++m) r[LR_CA + m] = c->colamt[0][m];
    for (int m = 0; m < 4; ++m) r[LR_WX + m] = c->wx[0][m];
    r[LR_COLDRY] = c->coldry[0], r[LR_COLBRD] = c->colbrd[0], r[LR_PAVEL] = c->pavel[0];
    for (int m = 0; m < 6; ++m) r[LR_RF + 2 * m] = c->rfrate[0][m][0], r[LR_RF + 2 * m + 1] = c->rfrate[0][m][1];
}

static void lw_rec_in(const double *r, LwCol *c)
{
    memset(c, 0, sizeof(*c));
    c->nlay = 1;
    c->laytrop = r[LR_LOWER] != 0.0 ? 1 : 0;
    c->jp[0] = (int)r[LR_JP], c->jt[0] = (int)r[LR_JT], c->jt1[0] = (int)r[LR_JT1];
    c->indself[0] = (int)r[LR_INDS], c->indfor[0] = (int)r[LR_INDF], c->indminor[0] = (int)r[LR_INDM];
    c->fac00[0] = r[LR_F00], c->fac01[0] = r[LR_F01], c->fac10[0] = r[LR_F10], c->fac11[0] = r[LR_F11];
    c->selffac[0] = r[LR_SF], c->selffrac[0] = r[LR_SFR], c->forfac[0] = r[LR_FF], c->forfrac[0] = r[LR_FFR];
    c->minorfrac[0] = r[LR_MF], c->scaleminor[0] = r[LR_SCM], c->scaleminorn2[0] = r[LR_SCMN2];
    for (int m = 0; m < 7; ++m) c->colamt[0][m] = r[LR_CA + m];
    for (int m = 0; m < 4; ++m) c->wx[0][m] = r[LR_WX + m];
    c->coldry[0] = r[LR_COLDRY], c->colbrd[0] = r[LR_COLBRD], c->pavel[0] = r[LR_PAVEL];
    for (int m = 0; m < 6; ++m) c->rfrate[0][m][0] = r[LR_RF + 2 * m], c->rfrate[0][m][1] = r[LR_RF + 2 * m + 1];
}

/* LW setcoef of one layer: pressure pav (hPa), temperature tav (K), the column amounts colamt[7]
 * (molecules/cm2), coldry and colbrd; h2ovmr = colamt[0] / coldry as lwrad_prep forms it. */
int oracle_lw_setcoef_layer(double pav, double tav, const double *colamt, double coldry, double colbrd, double *rec)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    LwCol *c = (LwCol *)calloc(1, sizeof(LwCol));
    c->nlay = 1;
    c->pavel[0] = pav, c->tavel[0] = tav, c->tz[0] = tav, c->tz[1] = tav, c->stemp = tav;
    for (int m = 0; m < 7; ++m) c->colamt[0][m] = colamt[m];
    c->coldry[0] = coldry, c->colbrd[0] = colbrd;
    c->h2ovmr[0] = colamt[0] / fmax(coldry, 1e-30);
    olw_setcoef(t, c);
    lw_rec_out(c, rec);
    free(c);
    return WR_OK;
}

/* taugbNN of LW band b (0..15) for one layer given as a record: tau[ng], fr[ng] */
int oracle_lw_band_layer(const double *rec, int b, double *tau, double *fr)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    if (b < 0 || b > 15) return WR_ERR_ARG;
    LwCol *c = (LwCol *)malloc(sizeof(LwCol));
    lw_rec_in(rec, c);
    olw_band(t, c, b, 0, tau, fr);
    free(c);
    return WR_OK;
}

/* SW setcoef of one layer (day column): colamt[6], the layer at index 0 of a one-layer column */
int oracle_sw_setcoef_layer(double pav, double tav, double h2ovmr, const double *colamt, double *rec)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    SwCol *c = (SwCol *)calloc(1, sizeof(SwCol));
    c->nlay = 1;
    c->pavel[0] = pav, c->tavel[0] = tav, c->h2ovmr[0] = h2ovmr;
    for (int m = 0; m < 6; ++m) c->colamt[0][m] = colamt[m];
    osw_setcoef(t, c);
    rec[SR_LAYTROP] = c->laytrop;
    rec[SR_JP] = c->jp[0], rec[SR_JT] = c->jt[0], rec[SR_JT1] = c->jt1[0];
    rec[SR_INDS] = c->indself[0], rec[SR_INDF] = c->indfor[0];
    rec[SR_F00] = c->fac00[0], rec[SR_F01] = c->fac01[0], rec[SR_F10] = c->fac10[0], rec[SR_F11] = c->fac11[0];
    rec[SR_SF] = c->selffac[0], rec[SR_SFR] = c->selffrac[0], rec[SR_FF] = c->forfac[0], rec[SR_FFR] = c->forfrac[0];
    for (int m = 0; m < 6; ++m) rec[SR_CA + m] = c->colamt[0][m];
    rec[SR_COLMOL] = c->colmol[0];
    free(c);
    return WR_OK;
}

/* taumolNN of SW band bi (0..13 = bands 16..29) for layer 1 of a one-layer column whose laytrop is
 * rec[SR_LAYTROP] (2: lower atmosphere in every band, 0: upper): tg[ng], tr[ng] */
int oracle_sw_band_layer(const double *rec, int bi, double *tg, double *tr)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    if (bi < 0 || bi > 13) return WR_ERR_ARG;
    SwCol *c = (SwCol *)calloc(1, sizeof(SwCol));
    c->nlay = 1;
    c->laytrop = (int)rec[SR_LAYTROP];
    c->jp[0] = (int)rec[SR_JP], c->jt[0] = (int)rec[SR_JT], c->jt1[0] = (int)rec[SR_JT1];
    c->indself[0] = (int)rec[SR_INDS], c->indfor[0] = (int)rec[SR_INDF];
    c->fac00[0] = rec[SR_F00], c->fac01[0] = rec[SR_F01], c->fac10[0] = rec[SR_F10], c->fac11[0] = rec[SR_F11];
    c->selffac[0] = rec[SR_SF], c->selffrac[0] = rec[SR_SFR], c->forfac[0] = rec[SR_FF], c->forfrac[0] = rec[SR_FFR];
    for (int m = 0; m < 6; ++m) c->colamt[0][m] = rec[SR_CA + m];
    c->colmol[0] = rec[SR_COLMOL];
    for (int g = 0; g < 16; ++g) tg[g] = 0.0, tr[g] = 0.0;
    osw_band(t, c, bi, 0, tg, tr);
    free(c);
    return WR_OK;
}

/* LW cldprop of one cloudy layer (cloud fraction cf > 0): taucld[16] */
int oracle_lw_cldprop_layer(double cf, double cldliq, double cldice, double relw, double reiw, double crp, double *taucld)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    LwCol *c = (LwCol *)calloc(1, sizeof(LwCol));
    c->nlay = 1;
    c->lcf1 = 1;
    c->cldfrc[1] = cf;
    c->clwp[0] = cldliq, c->ciwp[0] = cldice, c->relw[0] = relw, c->reiw[0] = reiw, c->crp[0] = crp;
    olw_cldprop(t, c);
    for (int b = 0; b < 16; ++b) taucld[b] = c->taucld[b][0];
    free(c);
    return WR_OK;
}

/* SW cldprop of one cloudy layer: per band tau, tau * ssa and tau * ssa * asy (taucw, ssacw, asycw) */
int oracle_sw_cldprop_layer(double cf, double cliqp, double reliq, double cicep, double reice, double rainp, double *tau,
                            double *ssa, double *asy)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    SwCol *c = (SwCol *)calloc(1, sizeof(SwCol));
    c->nlay = 1;
    c->zcf1 = 1.0;
    c->cfrac[0] = cf;
    c->cliqp[0] = cliqp, c->reliq[0] = reliq, c->cicep[0] = cicep, c->reice[0] = reice, c->cdat1[0] = rainp;
    osw_cldprop(t, c, 1);
    for (int b = 0; b < 14; ++b) tau[b] = c->taucw[b][0], ssa[b] = c->ssacw[b][0], asy[b] = c->asycw[b][0];
    free(c);
    return WR_OK;
}

/* spcvrtc's exponential: 1 - z + z^2/2 for z <= 0.06, else the rswinit table at the rounded index */
double oracle_sw_expm(double z, int clamped)
{
    const OT *t = oracle_tables();
    return t ? osw_expm(t, z, clamped) : -1.0;
}

/* delta scaling + PIFM two-stream of one layer: out = rb, rd, tb, td, tau1 */
int oracle_sw_layer_opt(double tau0, double ssa0, double asy0, double cz, double *out)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    osw_layer_opt(t, tau0, ssa0, asy0, cz, &out[0], &out[1], &out[2], &out[3], &out[4]);
    return WR_OK;
}

/* One g-point of spcvrtc on a column of nlay layers (surface-first per-layer inputs; ta / sa / ya the
 * aerosol optical depth, single-scattering albedo and asymmetry, NULL: none), per unit incident flux:
 * f[0..3][nlay+1] = clear up, clear down, total up, total down, levels surface first. */
int oracle_sw_gpoint(int nlay, double cz, double albb, double albd, const double *tr, const double *tg,
                     const double *ta, const double *sa, const double *ya, int cloudy, const double *zc,
                     const double *tc, const double *sc, const double *ac, double *f)
{
    const OT *t = oracle_tables();
    if (!t) return WR_ERR_TABLES;
    if (nlay < 1 || nlay > O_NL) return WR_ERR_ARG;
    double f0u[O_NL + 3], f0d[O_NL + 3], fcu[O_NL + 3], fcd[O_NL + 3];
    memset(fcu, 0, sizeof fcu);
    memset(fcd, 0, sizeof fcd);
    osw_gpoint(t->sw_exp_tbl, nlay, cz, 1.0 / cz, albb, albd, tr, tg, ta, sa, ya, cloudy, zc, tc, sc, ac, f0u, f0d, fcu,
               fcd);
    const int n = nlay + 1;
    for (int k = 0; k < n; ++k) {
        f[0 * n + k] = f0u[k + 1];
        f[1 * n + k] = f0d[k + 1];
        f[2 * n + k] = fcu[k + 1];
        f[3 * n + k] = fcd[k + 1];
    }
    return WR_OK;
}
