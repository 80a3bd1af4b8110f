/* oracle_nmmb.c -- CPU restatement of the NMMB-state side of the radiation call (TEST
 * INFRASTRUCTURE ONLY): the prep that turns the model state into column inputs and the post that
 * writes the radiation outputs back into the state.  Same operation sequence as the library's
 * wr_nmmb_prep / wr_nmmb_post (FP64 on the single-precision state, own exp/log/pow), so the two
 * agree bit for bit.  Each step cites the reference kernel it restates. */
#include "oalbd.h"
#include "oracle.h"

/* physcons of progcld2 (phys_rrtm_progcld2.cu:49-55) and grrad setup (phys_rrtm_grrad_setup.cu:16) */
#define N_G 9.80665
#define N_RD 287.05
#define N_RV 461.5
#define N_FVIRT (N_RV / N_RD - 1.0)
#define N_T0C 273.15
#define N_TTP 273.16
/* module_CONSTANTS of the driver call (nmmb_common.cuh:36-40) */
#define N_RD_MC 287.04
#define N_RV_MC 461.6
#define N_EP2 (N_RD_MC / N_RV_MC)
#define N_EPSQ 1.0e-12
#define N_TRAD_ICE (273.15 - 30.0)
#define N_STBOLT 5.67051e-8

/* FPVSX of module_MP_FER_HIRES.F90:2915-2971 (fer_fpvs_common.cuh:28-45): saturation vapour
 * pressure over water (T >= TTP) or ice, kPa */
static double n_fpvsx(double t)
{
    const double ttp = 2.7316e+2, hvap = 2.5000e+6, psat = 6.1078e+2, cliq = 4.1855e+3, cvap = 1.8460e+3,
                 cice = 2.1060e+3, hsub = 2.8340e+6;
    const double psatk = psat * 1.0e-3;
    const double dldt = cvap - cliq, xa = -dldt / N_RV_MC, xb = xa + hvap / (N_RV_MC * ttp);
    const double dldti = cvap - cice, xai = -dldti / N_RV_MC, xbi = xai + hsub / (N_RV_MC * ttp);
    const double tr = ttp / t;
    if (t >= ttp) return psatk * om_pow(tr, xa) * om_exp(xb * (1.0 - tr));
    return psatk * om_pow(tr, xai) * om_exp(xbi * (1.0 - tr));
}

/* FPVS (phys_fer_fpvs.cu:31-46): linear interpolation in the 7501-entry FPVSX table over 180..330 K
 * (0.02 K steps; the table nodes are evaluated where they are needed), Magnus-Tetens outside */
static double n_fpvs(double t)
{
    const double xmin = 180.0, xmax = 330.0, nx = 7501.0;
    const double xinc = (xmax - xmin) / (nx - 1.0);
    if (t >= xmin && t <= xmax) {
        const double c2 = 1.0 / xinc, c1 = 1.0 - xmin / xinc;
        const double xj = fmin(fmax(c1 + c2 * t, 1.0), nx);
        const int jx = (int)fmin(xj, nx - 1.0);
        const double f0 = n_fpvsx(xmin + (double)(jx - 1) * xinc), f1 = n_fpvsx(xmin + (double)jx * xinc);
        return f0 + (xj - (double)jx) * (f1 - f0);
    }
    if (t > xmax) return 0.61078 * om_exp(17.2694 * (t - 273.16) / (t - 35.86));
    return 0.61078 * om_exp(21.8746 * (t - 273.16) / (t - 7.66));
}

#define S3(a, k, j, i) (a)[((size_t)(k) * st->jm + (j)) * st->im + (i)]
#define S2(a, j, i) (a)[(size_t)(j) * st->im + (i)]

/* the Exner function (p / 1e5 Pa)^(con_rd / con_cp) as funcphys' fpkap table gives it: linear between
 * nodes every 10 Pa (phys_rrtm_gas_init.cu:52-72), the nodes evaluated here */
static double n_fpkap(double p)
{
    const double rocp = (double)(287.05f / 1004.6f);
    const double xinc = 110000.0 / 11000.0;
    const double xj = fmin(fmax(1.0 + p / xinc, 1.0), 11001.0);
    const int jx = (int)fmin(xj, 11000.0);
    const double t0 = om_pow((double)(jx - 1) * xinc / 1.0e5, rocp);
    const double t1 = om_pow((double)jx * xinc / 1.0e5, rocp);
    return t0 + (xj - (double)jx) * (t1 - t0);
}

/* getozn (phys_rrtm_getozn.cu:57-110) at one layer of Exner function wk: linear in the Exner function
 * between the climatology's levels pk, the end values outside */
static double n_ozone_at(const double *o3i, const double *pk, int loz, double wk)
{
    if (wk > pk[loz - 1]) return o3i[loz - 1];
    if (wk <= pk[0]) return o3i[0];
    double v = o3i[0];
    for (int k = 0; k < loz - 1; ++k)
        if (wk > pk[k] && wk <= pk[k + 1]) {
            const double tem = (pk[k + 1] - wk) * (1.0 / (pk[k + 1] - pk[k]));
            v = tem * o3i[k] + (1.0 - tem) * o3i[k + 1];
        }
    return v;
}

/* One column c of the prep (see include/woofrad.h, wr_nmmb_prep). */
static void n_prep_col(const wr_nmmb_state *st, const wr_nmmb_opts *o, const wr_nmmb_cols *w, int c)
{
    const int im = st->im, lm = st->lm;
    const int i = c % im, j = st->j0 + c / im;
    const size_t nc = (size_t)(st->j1 - st->j0) * im;
    const int nrl = lm + 1;
    /* driver prep (phys_rrtm_driver_prep.cu:82-98), single precision as there */
    const float pd = S2(st->pd, j, i);
    float p8 = st->pt;
    w->p_lev[c] = (double)p8;
    for (int k = 0; k < lm; ++k) {
        p8 = p8 + st->pdsg1[k] + st->dsg2[k] * pd;
        w->p_lev[(size_t)(k + 1) * nc + c] = (double)p8;
        if (w->p_phy) S3(w->p_phy, k, j, i) = st->sgml2[k] * pd + st->psgml1[k];
    }
    /* climatological ozone (getozn): this latitude's climatology column, between the two months */
    const wr_ozone *oz = o->ozone;
    int oz_on = oz != NULL;
    double o3i[64], pk[64];
    if (oz_on) {
        const double deglat = (double)S2(st->glat, j, i) * 57.29578;
        const double blte = oz->blte, dlte = oz->dlte, elte = blte + (double)(oz->jmr - 1) * dlte;
        int j1, j2;
        double tem1;
        if (deglat > blte && deglat < elte) {
            tem1 = (deglat - blte) / dlte + 1.0;
            j1 = (int)tem1;
            j2 = j1 + 1;
            tem1 = tem1 - (double)j1;
        } else if (deglat <= blte) {
            j1 = j2 = 1, tem1 = 1.0;
        } else if (deglat >= elte) {
            j1 = j2 = oz->jmr, tem1 = 1.0;
        } else {
            j1 = j2 = 1, tem1 = 1.0, oz_on = 0;
        }
        const double tem2 = 1.0 - tem1, facoz = (double)oz->facoz;
        const double rocp = (double)(287.05f / 1004.6f);
        for (int k = 0; k < oz->loz; ++k) {
            const size_t m1 = ((size_t)(oz->k1oz - 1) * oz->loz + k) * oz->jmr;
            const size_t m2 = ((size_t)(oz->k2oz - 1) * oz->loz + k) * oz->jmr;
            const double tem3 = tem2 * (double)oz->o3r[m1 + j1 - 1] + tem1 * (double)oz->o3r[m1 + j2 - 1];
            const double tem4 = tem2 * (double)oz->o3r[m2 + j1 - 1] + tem1 * (double)oz->o3r[m2 + j2 - 1];
            o3i[k] = tem4 * facoz + tem3 * (1.0 - facoz);
            const double ps = (double)oz->pstr[k];
            pk[k] = ps >= 0.1 ? n_fpkap(ps * 100.0) : om_pow(ps / 1000.0, rocp);
        }
    }
    /* radiation layers, top-first, layer 0 = the added layer above the model top: plvl, plyr in hPa
     * as the column call forms them from p_lev */
    double plvl[O_NL + 1], plyr[O_NL], ct[O_NL];
    plvl[0] = 1.0e-6;
    for (int l = 1; l <= nrl; ++l) plvl[l] = 10.0 * (w->p_lev[(size_t)(l - 1) * nc + c] * 1.0e-3);
    if (plvl[1] <= 1.0e-6) plvl[1] = 2.0e-6;
    plyr[0] = 0.5 * plvl[1];
    if (w->o3_top)
        w->o3_top[c] = oz_on ? n_ozone_at(o3i, pk, oz->loz, om_pow(plyr[0] * 100.0 / 1.0e5, (double)(287.05f / 1004.6f)))
                             : (double)S3(st->o3, 0, j, i);
    for (int l = 1; l < nrl; ++l)
        plyr[l] = 10.0 * (0.5 * (w->p_lev[(size_t)(l - 1) * nc + c] + w->p_lev[(size_t)l * nc + c]) * 1.0e-3);
    ct[0] = 0.0;
    const double recw1 = 620.35 / om_pow(200.0, 1.0 / 3.0);   /* rsipath2: 620.35 / TNW^(1/3) */
    for (int k = 0; k < lm; ++k) {
        const size_t kc = (size_t)k * nc + c;
        const double t = (double)S3(st->t, k, j, i), q = (double)S3(st->q, k, j, i);
        w->t_lay[kc] = t;
        w->q_lay[kc] = q;
        if (oz_on) { /* the model layer's Exner function as the driver call forms it (PRSLK) */
            const float prsl = (st->sgml2[k] * pd + st->psgml1[k]) / 1000.0f;
            w->o3_lay[kc] = n_ozone_at(o3i, pk, oz->loz, om_pow((double)prsl * 0.01, (double)(287.04f / 1004.6f)));
        } else {
            w->o3_lay[kc] = (double)S3(st->o3, k, j, i);
        }
        double cw = (double)S3(st->cw, k, j, i);
        double fice = (double)S3(st->f_ice, k, j, i), frain = (double)S3(st->f_rain, k, j, i);
        /* driver call SUBGRID (phys_rrtm_driver_call.cu:248-275): RH > 95 % without condensate */
        if (o->subgrid) {
            const double wv = q / (1.0 - q);
            const double p1 = 0.5 * (w->p_lev[kc] + w->p_lev[kc + nc]) * 1.0e-3;  /* kPa */
            const double esat = fmin(n_fpvs(t), 0.99 * p1);
            const double qsat = N_EP2 * esat / (p1 - esat);
            const double rhum = wv / qsat;
            if (cw < N_EPSQ && rhum > 0.95) {
                cw = fmin(0.01e-3, fmin(0.01, rhum - 0.95) * qsat);
                fice = (t > N_TRAD_ICE) ? 0.0 : 1.0;
                frain = 0.0;
            }
        }
        /* grrad clouds: clw below EPSQ is zero (phys_rrtm_grrad_clouds.cu:397-412) */
        const double clw = cw < N_EPSQ ? 0.0 : cw;
        /* progcld2 (phys_rrtm_progcld2.cu:145-305), NAM flags: no lcrick smoothing */
        const int l = k + 1;
        const double pl = plyr[l], plo = plvl[l], plp = plvl[l + 1];
        const double qlyr = fmax(1.0e-7, q);                  /* grrad setup floor QME5/QME6 */
        const double tvly = t * (1.0 + N_FVIRT * qlyr);
        const double fc = fmax(0.0, fmin(1.0, fice)), fr = fmax(0.0, fmin(1.0, frain));
        double qcice, qrain, qcwat, clw2;
        if (t - N_T0C > -40.0) {
            qcice = clw * fc;
            const double tem1 = clw - qcice;
            qrain = tem1 * fr;
            qcwat = tem1 - qrain;
            clw2 = qcwat + qcice;
        } else {
            qcice = clw;
            qrain = 0.0;
            qcwat = 0.0;
            clw2 = clw;
        }
        double cwp = 0.0, cip = 0.0, crp = 0.0, rew = 5.0, rei = 10.0;
        if (qcwat + qcice + qrain > 1.0e-12) {
            const double rho = 100.0 * pl / (N_RD * t * (1.0 + N_FVIRT * qlyr));
            const double cpath = fabs(plp - plo) * (100000.0 / N_G);
            if (qcwat > 0.0) {
                rew = fmax(5.0, recw1 * om_pow(rho * qcwat, 1.0 / 3.0));
                cwp = cpath * qcwat;
            }
            if (qrain > 0.0) crp = cpath * qrain;
            cip = cpath * qcice;
        }
        const double tem2d = (N_G * pl) / (N_RD * (plp - plo));
        double cf = 0.0;
        if (clw2 > 1.0e-8 * (pl * 0.001)) cf = fmin(1.0, 1.0e5 * clw2 * pl * 0.001);
        if (cf < 0.001) {
            cf = 0.0;
            cwp = 0.0;
            cip = 0.0;
            crp = 0.0;
        }
        if (cip > 0.0) {
            const double tem3 = tem2d * cip / tvly;
            const double tem1 = t - N_TTP;
            double tem2;
            if (tem1 < -50.0)
                tem2 = (1250.0 / 9.917) * om_pow(tem3, 0.109);
            else if (tem1 < -40.0)
                tem2 = (1250.0 / 9.337) * om_pow(tem3, 0.08);
            else if (tem1 < -30.0)
                tem2 = (1250.0 / 9.208) * om_pow(tem3, 0.055);
            else
                tem2 = (1250.0 / 9.387) * om_pow(tem3, 0.031);
            rei = fmax(10.0, tem2 * 0.6495);
        }
        ct[l] = cf;
        w->cldf[kc] = cf;
        w->cwp[kc] = cwp;
        w->cip[kc] = cip;
        w->crp[kc] = crp;
        w->rew[kc] = rew;
        w->rei[kc] = rei;
    }
    /* gethml (phys_rrtm_progcld2.cu:307-450, ivflip = 0), Fortran 1-based top-first layers k = 1..nrl.
     * OVCST is 1 - 1e-8 in single precision, which is 1. */
    double ptop1[4];
    {
        static const double p1c[4] = {1050.0, 650.0, 400.0, 0.0}, p2c[4] = {1050.0, 750.0, 500.0, 0.0};
        const double tem2 = fmax(0.0, 4.0 * fabs((double)S2(st->glat, j, i)) / 3.1415926535897931 - 1.0);
        for (int id = 0; id < 4; ++id) ptop1[id] = p1c[id] + (p2c[id] - p1c[id]) * tem2;
    }
    double cl[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int mt[4] = {0, 0, 0, 0}, mb[4] = {0, 0, 0, 0};
#define CT1(k) fmin(1.0, fmax(ct[(k)-1], 0.0))
    {
        double cl1 = 1.0, cl2 = 1.0;
        for (int k = nrl; k >= 1; --k) {
            const double ccur = CT1(k);
            if (o->iovr == 0) {
                if (ccur >= 0.001) cl1 = cl1 * (1.0 - ccur);
                if (k == o->llyr) cl[4] = 1.0 - cl1;
            } else {
                if (ccur >= 0.001) {
                    cl2 = fmin(cl2, 1.0 - ccur);
                } else {
                    cl1 = cl1 * cl2;
                    cl2 = 1.0;
                }
                if (k == o->llyr) cl[4] = 1.0 - cl1 * cl2;
            }
        }
        cl[3] = (o->iovr == 0) ? 1.0 - cl1 : 1.0 - cl1 * cl2;
    }
    {
        double cl1 = 0.0, cl2 = 0.0;
        int kbt1 = nrl, kbt2 = nrl, kth1 = 0, kth2 = 0, idom = 1;
        mb[0] = nrl;
        mt[0] = nrl;
        for (int k = nrl; k >= 1; --k) {
            int id = idom, id1 = id + 1;
            const double pcur = plyr[k - 1];
            const double ccur = CT1(k);
            double pnxt, cnxt;
            if (k > 1) {
                pnxt = plyr[k - 2];
                cnxt = CT1(k - 1);
            } else {
                pnxt = -1.0;
                cnxt = 0.0;
            }
            if (pcur < ptop1[id1 - 1]) {
                id += 1;
                id1 += 1;
                idom = id;
            }
            if (ccur >= 0.001) {
                if (kth2 == 0) kbt2 = k;
                kth2 += 1;
                cl2 = (o->iovr == 0) ? cl2 + ccur - cl2 * ccur : fmax(cl2, ccur);
                if (cnxt < 0.001 || pnxt < ptop1[id1 - 1]) {
                    kbt1 = (int)round((cl1 * (double)kbt1 + cl2 * (double)kbt2) / (cl1 + cl2));
                    kth1 = (int)round((cl1 * (double)kth1 + cl2 * (double)kth2) / (cl1 + cl2));
                    cl1 = cl1 + cl2 - cl1 * cl2;
                    kbt2 = k - 1;
                    kth2 = 0;
                    cl2 = 0.0;
                }
            }
            if (pnxt < ptop1[id1 - 1]) {
                cl[id - 1] = cl1;
                mt[id - 1] = imin(kbt1, kbt1 - kth1 + 1);
                mb[id - 1] = kbt1;
                cl1 = 0.0;
                kbt1 = k - 1;
                kth1 = 0;
                if (id1 <= 3) {
                    mb[id1 - 1] = kbt1;
                    mt[id1 - 1] = kbt1;
                }
            }
        }
    }
#undef CT1
    for (int n = 0; n < 5; ++n) w->clds[(size_t)n * nc + c] = (float)cl[n];
    for (int n = 0; n < 3; ++n) {
        w->mtop[(size_t)n * nc + c] = mt[n];
        w->mbot[(size_t)n * nc + c] = mb[n];
    }
}

int oracle_nmmb_prep(const wr_nmmb_state *st, const wr_nmmb_opts *o, const wr_nmmb_cols *w)
{
    if (!st || !o || !w || st->lm + 1 > O_NL - 1 || st->j0 < 0 || st->j1 > st->jm || st->j0 >= st->j1) return 1;
    if (o->ozone && (o->ozone->loz < 2 || o->ozone->loz > 32 || o->ozone->k1oz < 1 || o->ozone->k1oz > 12 ||
                     o->ozone->k2oz < 1 || o->ozone->k2oz > 12 || !w->o3_top))
        return 1;
    const int n = (st->j1 - st->j0) * st->im;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) n_prep_col(st, o, w, c);
    return 0;
}

/* driver post (phys_rrtm_driver_post.cu:71-150), icwp /= -1, lssav, CU_CLOUD = 0 (CUCLD off) */
int oracle_nmmb_post(const wr_nmmb_state *st, const wr_nmmb_cols *w, const wr_outputs *rad, const double *czmean,
                     int lw_step, int sw_step, const wr_nmmb_out *o)
{
    if (!st || !w || !rad || !o || !czmean) return 1;
    const int im = st->im, lm = st->lm, nrl = lm + 1;
    const size_t nc = (size_t)(st->j1 - st->j0) * im;
    for (size_t c = 0; c < nc; ++c) {
        const int i = (int)(c % im), j = st->j0 + (int)(c / im);
        if (lw_step) {
            for (int k = 0; k < lm; ++k) S3(o->rlwtt, k, j, i) = (float)rad->hr_lw[(size_t)k * nc + c];
            S2(o->rlwin, j, i) = (float)rad->flw_dn[(size_t)nrl * nc + c];
            S2(o->rlwtoa, j, i) = (float)rad->flw_up[c];
            const float tl = S3(st->t, lm - 1, j, i);
            S2(o->sigt4, j, i) = (float)(N_STBOLT * (double)tl * (double)tl * (double)tl * (double)tl);
        }
        if (sw_step) {
            for (int k = 0; k < lm; ++k) S3(o->rswtt, k, j, i) = (float)rad->hr_sw[(size_t)k * nc + c];
            S2(o->rswin, j, i) = (float)rad->fsw_dn[(size_t)nrl * nc + c];
            S2(o->rswout, j, i) = (float)rad->fsw_up[(size_t)nrl * nc + c];
            S2(o->rswinc, j, i) = (float)rad->fsw_dn_sfc_clr[c];
            S2(o->rswtoa, j, i) = (float)rad->fsw_up[c];
        }
        S2(o->czmean, j, i) = (float)czmean[c];
        /* cloud cover: zero when the column holds no condensate (max CW < EPSQ2 = 1e-8) */
        float cwmax = -3.402823466e38f;
        for (int k = 0; k < lm; ++k) cwmax = fmaxf(cwmax, S3(st->cw, k, j, i));
        const int clear = cwmax < 1.0e-8f;
        for (int k = 0; k < lm; ++k) S3(o->cldfra, k, j, i) = clear ? 0.0f : (float)w->cldf[(size_t)k * nc + c];
        const float cs[4] = {clear ? 0.0f : w->clds[c], clear ? 0.0f : w->clds[nc + c], clear ? 0.0f : w->clds[2 * nc + c],
                             clear ? 0.0f : w->clds[3 * nc + c]};
        S2(o->cfracl, j, i) = cs[0];
        S2(o->cfracm, j, i) = cs[1];
        S2(o->cfrach, j, i) = cs[2];
        S2(o->acfrst, j, i) = S2(o->acfrst, j, i) + cs[3];
        S2(o->ncfrst, j, i) = S2(o->ncfrst, j, i) + 1;
        S2(o->ncfrcv, j, i) = S2(o->ncfrcv, j, i) + 1;
    }
    return 0;
}
#undef S3
#undef S2

/* NAM surface albedo (the NMMB radiation driver's IALB = 2 block, phys_rrtm_grrad_clouds.cu:345-396;
 * setalb ialbflg 2 copies it): the state's albedo; over open water (sm > 0.5) with the sun up, the
 * Payne (1972) table at transmission 0.537, linear in zenith angle between its nodes and in
 * transmission between its 0.50 and 0.55 columns. */
int oracle_nmmb_albedo(int im, int jm, int j0, int j1, const float *albedo, const float *sm, const double *coszen,
                       double *alb)
{
    static const double za[20] = WR_ALBD_ZA, dza[19] = WR_ALBD_DZA, lo[20] = WR_ALBD_LO, hi[20] = WR_ALBD_HI;
    if (im <= 0 || jm <= 0 || j0 < 0 || j1 > jm || j0 >= j1 || !albedo || !sm || !coszen || !alb) return WR_ERR_ARG;
    const int nc = (j1 - j0) * im;
    for (int c = 0; c < nc; ++c) {
        const size_t p = (size_t)(j0 + c / im) * im + c % im;
        double a = (double)albedo[p];
        const double cz = coszen[c];
        if (cz > 0.0 && (double)sm[p] > 0.5) {
            const double zen = (180.0 / 3.14159265) * acos(fmax(cz, 0.0));
            int jx;
            if (zen >= 74.0)
                jx = (int)(0.5 * (90.0 - zen) + 1.0);
            else if (zen >= 50.0)
                jx = (int)(0.25 * (74.0 - zen) + 9.0);
            else
                jx = (int)(0.1 * (50.0 - zen) + 15.0);
            jx = imin(19, imax(1, jx));
            const double dzen = -(zen - za[jx - 1]) / dza[jx - 1];
            const double a1 = lo[jx - 1] + dzen * (lo[jx] - lo[jx - 1]);
            const double a2 = hi[jx - 1] + dzen * (hi[jx] - hi[jx - 1]);
            a = a1 + 20.0 * (a2 - a1) * (0.537 - WR_ALBD_TRN_LO);
        }
        alb[c] = a;
    }
    return WR_OK;
}
