/* oracle.h -- internal types of the CPU FP64 RRTMG restatement (TEST INFRASTRUCTURE ONLY). */
#ifndef WOOFRAD_ORACLE_H
#define WOOFRAD_ORACLE_H

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/woofrad.h"
#include "omath.h"
#include "otables.h"

#define O_NL 66      /* max radiation layers (model layers + 1) */
#define O_NTBL 10000 /* LW/SW exponential look-up tables (NTBL / NTBMX) */

/* physcons (phys_rrtm_lw_lwrad.cu:55-62, phys_rrtm_sw_swrad.cu:22-34) in double */
#define O_CON_G 9.80665
#define O_CON_AVGD 6.0221415e23
#define O_CON_AMD 28.9644
#define O_CON_AMW 18.0154
#define O_CON_AMO3 47.9982
#define O_CON_CP 1.0046e3
#define O_AMDW (O_CON_AMD / O_CON_AMW)
#define O_AMDO3 (O_CON_AMD / O_CON_AMO3)
#define O_EPS 1.0e-6
#define O_ONEMINUS (1.0 - 1.0e-6)
#define O_FTINY 1.0e-12
#define O_STPFAC (296.0 / 1013.0)
#define O_HEATFAC (O_CON_G * 1.0e-2 / O_CON_CP) /* iswrate = ilwrate = 2: K/s */
#define O_FLUXFAC (3.141592653589793 * 2.0e4)  /* pival * 2e4, pival = 2 asin(1) */
#define O_S0 1368.22

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline double fs_mul(double a, double b) { return a * b; }

typedef struct {
    int ng, ns;
    const double *absa, *absb, *selfref, *forref, *fracrefa, *fracrefb;
    const double *ka_m[3], *kb_m[3]; /* minor-gas tables */
    const double *x1, *x2;           /* cross-section tables (cfc / ccl4) */
    double refrat_planck_a, refrat_planck_b, refrat_m_a, refrat_m_b, refrat_m_a3;
} LwBand;

typedef struct {
    int ng, ns;
    const double *absa, *absb, *selfref, *forref;
    const double *rayl_v, *rayla, *raylb, *abso3a, *abso3b, *absch4, *absh2o, *absco2;
    double rayl, strrat, givfac;
    int nabsa, nabsb, nself, nfor; /* rows */
} SwBand;

typedef struct {
    ot_file f;
    /* LW */
    const double *preflog, *tref, *chi_mls, *totplnk, *delwave;
    const double *absliq1, *absice1, *absice2, *absice3;
    const int32_t *ipat;
    LwBand lw[16];
    double exp_tbl[O_NTBL + 1], tfn_tbl[O_NTBL + 1], tau_tbl[O_NTBL + 1];
    /* SW */
    const double *sw_preflog, *sw_tref, *sf01, *sf02, *sf03, *sw_strrat, *sw_specwt;
    double scalekur;
    const int32_t *layreffr, *ix1, *ix2, *ibx;
    const double *extliq1, *ssaliq1, *asyliq1, *extice3, *ssaice3, *asyice3;
    SwBand sw[14];
    double sw_exp_tbl[O_NTBL + 1];
} OT;

#define O_CHI(t, g, lev) ((t)->chi_mls[((lev)-1) * 7 + ((g)-1)])

/* Radiation-grid column after the driver prep (grrad setup / gases / clouds), TOP-FIRST:
 * layer 0 is the added layer, levels 0..nrl (level 0 = 1e-6 hPa). */
typedef struct {
    int nrl;
    double plvl[O_NL + 1], plyr[O_NL], tlyr[O_NL], tlvl[O_NL + 1], qlyr[O_NL], olyr[O_NL];
    double cld[O_NL][9]; /* clouds(:, n): frac, cwp, rew, cip, rei, crp, rer, snow path, snow radius */
    double aerlw[2][16][O_NL]; /* aerosol optical depth, single-scattering albedo per LW band (0: none) */
    double aersw[3][14][O_NL]; /* aerosol optical depth, ssa, asymmetry per SW band (0: none)           */
    double tskn;
    double vmr[10];
} Rad;

/* LW column, SURFACE-FIRST (the reference's lwrad internal order) */
typedef struct {
    int nlay, laytrop, lcf1;
    double pavel[O_NL], tavel[O_NL], delp[O_NL], tz[O_NL + 1], coldry[O_NL], colbrd[O_NL], h2ovmr[O_NL];
    double colamt[O_NL][7], wx[O_NL][4];
    double cldfrc[O_NL + 2], clwp[O_NL], relw[O_NL], ciwp[O_NL], reiw[O_NL], crp[O_NL], cldf[O_NL];
    double secdiff[16], semiss[16], stemp, pwvcm;
    int jp[O_NL], jt[O_NL], jt1[O_NL], indself[O_NL], indfor[O_NL], indminor[O_NL];
    double fac00[O_NL], fac01[O_NL], fac10[O_NL], fac11[O_NL], selffac[O_NL], selffrac[O_NL];
    double forfac[O_NL], forfrac[O_NL], minorfrac[O_NL], scaleminor[O_NL], scaleminorn2[O_NL];
    double rfrate[O_NL][6][2];
    double pklay[O_NL + 1][16], pklev[O_NL + 1][16];
    double taucld[16][O_NL];
    double tauaer[16][O_NL]; /* aerosol absorption optical depth tau (1 - ssa) */
} LwCol;

/* SW day column, SURFACE-FIRST */
typedef struct {
    int nlay, laytrop;
    double pavel[O_NL], tavel[O_NL], delp[O_NL], coldry[O_NL], colmol[O_NL], h2ovmr[O_NL];
    double colamt[O_NL][6];
    double cfrac[O_NL], cliqp[O_NL], reliq[O_NL], cicep[O_NL], reice[O_NL], cdat1[O_NL], cdat3[O_NL], cdat4[O_NL];
    double zcf0, zcf1, cosz, sntz, ssolar, albbm[2], albdf[2];
    int jp[O_NL], jt[O_NL], jt1[O_NL], indself[O_NL], indfor[O_NL];
    double fac00[O_NL], fac01[O_NL], fac10[O_NL], fac11[O_NL], selffac[O_NL], selffrac[O_NL];
    double forfac[O_NL], forfrac[O_NL];
    double sfluxzen[112];
    double taucw[14][O_NL], ssacw[14][O_NL], asycw[14][O_NL];
    double tauae[14][O_NL], ssaae[14][O_NL], asyae[14][O_NL]; /* aerosol optics */
    double cldfrc[O_NL], cldf[O_NL];
} SwCol;

int ot_bind(OT *t, const char *path);

void olw_tables(double *exp_tbl, double *tfn_tbl, double *tau_tbl);
void olw_prep(const Rad *r, double sfemis, LwCol *c);
void olw_cldprop(const OT *t, LwCol *c);
void olw_setcoef(const OT *t, LwCol *c);
void olw_taumol(const OT *t, const LwCol *c, double (*taug)[O_NL], double (*fracs)[O_NL]);
void o_mcica(uint64_t col, uint64_t seed, int kind, int ngpt, int nl, int iovr, const double *cf, uint32_t *bits,
             int words);
void olw_rtrnmc(const OT *t, const LwCol *c, double (*taug)[O_NL], double (*fracs)[O_NL], const uint32_t *mask,
                double *totuflux, double *totdflux, double *htr);
void olw_rtrnmr(const OT *t, const LwCol *c, double (*taug)[O_NL], double (*fracs)[O_NL], double *totuflux,
                double *totdflux, double *htr);

void osw_prep(const Rad *r, double alb, double cosz, double solcon, SwCol *c);
void osw_setcoef(const OT *t, SwCol *c);
void osw_sfluxzen(const OT *t, SwCol *c);
void osw_cldprop(const OT *t, SwCol *c, int mcica);
void osw_taumol(const OT *t, const SwCol *c, double (*taug)[O_NL], double (*taur)[O_NL]);
void osw_spcvrtc(const OT *t, const SwCol *c, double (*taug)[O_NL], double (*taur)[O_NL], const uint32_t *mask,
                 int mcica, double *fup, double *fdn, double *fdn0_sfc);
/* pieces exported for the oracle's own pins (oracle_pins.c) */
void olw_band(const OT *t, const LwCol *c, int b, int k, double *tau, double *fr);
void osw_band(const OT *t, const SwCol *c, int bi, int k, double *tg, double *tr);
double osw_expm(const OT *t, double z, int clamped);
void osw_layer_opt(const OT *t, double tau0, double ssa0, double asy0, double cz, double *rb, double *rd, double *tb,
                   double *td, double *tau1);
void osw_gpoint(const double *tbl, int nlay, double cz, double snt, double albb, double albd, const double *tr,
                const double *tg, const double *ta, const double *sa, const double *ya, int cloudy, const double *zc,
                const double *tc, const double *sc, const double *ac, double *f0u, double *f0d, double *fcu, double *fcd);
const OT *oracle_tables(void);
#endif
