/* oracle_column.c -- driver prep, table binding and the C entry points of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  Built into oracle/liboracle.so by oracle/Makefile; loaded by tests/,
 * __graft_entry__.smoke() and bench.py's CPU baseline, never by the product library.
 *
 *   oracle_run      -- the whole column path (prep, LW, SW, heating rates) for wr_columns input,
 *                      same outputs and diagnostics as wr_run_fp64 of include/woofrad.h;
 *   oracle_lw_tables, oracle_sw_table, oracle_exp/log, oracle_philox -- pieces pinned by tests.
 */
#include <stdio.h>

#include "oracle.h"

void osw_table(double *exp_tbl);

/* ---- table binding --------------------------------------------------------------------------- */
static const int LW_NGA[16] = {10, 12, 16, 14, 16, 8, 12, 8, 12, 6, 8, 8, 4, 2, 2, 2};
static const int SW_NG_[14] = {6, 12, 8, 8, 10, 10, 2, 10, 8, 6, 6, 8, 6, 12};

static const double *need_d(const OT *t, const char *nm, int *bad)
{
    const double *p = ot_d(&t->f, nm, 0);
    if (!p) {
        fprintf(stderr, "oracle: table %s missing\n", nm);
        *bad = 1;
    }
    return p;
}
static const double *opt_d(const OT *t, const char *nm) { return ot_d(&t->f, nm, 0); }

int ot_bind(OT *t, const char *path)
{
    memset(t, 0, sizeof(*t));
    if (ot_load(path, &t->f) != 0) return WR_ERR_TABLES;
    int bad = 0;
    char nm[64];
    t->preflog = need_d(t, "lw/preflog", &bad);
    t->tref = need_d(t, "lw/tref", &bad);
    t->chi_mls = need_d(t, "lw/chi_mls", &bad);
    t->totplnk = need_d(t, "lw/totplnk", &bad);
    t->delwave = need_d(t, "lw/delwave", &bad);
    t->absliq1 = need_d(t, "lw/absliq1", &bad);
    t->absice1 = need_d(t, "lw/absice1", &bad);
    t->absice2 = need_d(t, "lw/absice2", &bad);
    t->absice3 = need_d(t, "lw/absice3", &bad);
    t->ipat = ot_i(&t->f, "lw/ipat", 16);
    if (!t->ipat) bad = 1;
    int ns = 0;
    for (int b = 0; b < 16; ++b) {
        LwBand *B = &t->lw[b];
        B->ng = LW_NGA[b];
        B->ns = ns;
        ns += B->ng;
#define LWT(field, key, req)                                   \
    snprintf(nm, sizeof nm, "lw/b%02d/%s", b + 1, key);        \
    B->field = (req) ? need_d(t, nm, &bad) : opt_d(t, nm);
        LWT(absa, "absa", 1);
        LWT(absb, "absb", b != 5 && b != 11 && b != 12 && b != 14);
        LWT(selfref, "selfref", 1);
        LWT(forref, "forref", 1);
        LWT(fracrefa, "fracrefa", 1);
        LWT(fracrefb, "fracrefb", b != 5 && b != 11 && b != 14);
#undef LWT
    }
    /* minor-gas and cross-section tables per band (slot order as the band code reads them) */
#define MT(b, slot, key) snprintf(nm, sizeof nm, "lw/b%02d/%s", b, key), t->lw[(b)-1].slot = need_d(t, nm, &bad)
    MT(1, ka_m[0], "ka_mn2");
    MT(1, kb_m[0], "kb_mn2");
    MT(3, ka_m[0], "ka_mn2o");
    MT(3, kb_m[0], "kb_mn2o");
    MT(5, ka_m[0], "ka_mo3");
    MT(5, x1, "ccl4");
    MT(6, ka_m[0], "ka_mco2");
    MT(6, x1, "cfc11adj");
    MT(6, x2, "cfc12");
    MT(7, ka_m[0], "ka_mco2");
    MT(7, kb_m[0], "kb_mco2");
    MT(8, ka_m[0], "ka_mco2");
    MT(8, kb_m[0], "kb_mco2");
    MT(8, ka_m[1], "ka_mo3");
    MT(8, ka_m[2], "ka_mn2o");
    MT(8, kb_m[2], "kb_mn2o");
    MT(8, x1, "cfc12");
    MT(8, x2, "cfc22adj");
    MT(9, ka_m[0], "ka_mn2o");
    MT(9, kb_m[0], "kb_mn2o");
    MT(11, ka_m[0], "ka_mo2");
    MT(11, kb_m[0], "kb_mo2");
    MT(13, ka_m[0], "ka_mco2");
    MT(13, ka_m[1], "ka_mco");
    MT(13, kb_m[0], "kb_mo3");
    MT(15, ka_m[0], "ka_mn2");
#undef MT
    if (bad) return WR_ERR_TABLES;
    /* reference ratios of the binary-species bands, from chi_mls (taugbNN.cu constants) */
#define RR(g1, l1, g2, l2) (O_CHI(t, g1, l1) / O_CHI(t, g2, l2))
    t->lw[2].refrat_planck_a = RR(1, 9, 2, 9);
    t->lw[2].refrat_planck_b = RR(1, 13, 2, 13);
    t->lw[2].refrat_m_a = RR(1, 3, 2, 3);
    t->lw[2].refrat_m_b = RR(1, 13, 2, 13);
    t->lw[3].refrat_planck_a = RR(1, 11, 2, 11);
    t->lw[3].refrat_planck_b = RR(3, 13, 2, 13);
    t->lw[4].refrat_planck_a = RR(1, 5, 2, 5);
    t->lw[4].refrat_planck_b = RR(3, 43, 2, 43);
    t->lw[4].refrat_m_a = RR(1, 7, 2, 7);
    t->lw[6].refrat_planck_a = RR(1, 3, 3, 3);
    t->lw[6].refrat_m_a = RR(1, 3, 3, 3);
    t->lw[8].refrat_planck_a = RR(1, 9, 6, 9);
    t->lw[8].refrat_m_a = RR(1, 3, 6, 3);
    t->lw[11].refrat_planck_a = RR(1, 10, 2, 10);
    t->lw[12].refrat_planck_a = RR(1, 5, 4, 5);
    t->lw[12].refrat_m_a = RR(1, 1, 4, 1);
    t->lw[12].refrat_m_a3 = RR(1, 3, 4, 3);
    t->lw[14].refrat_planck_a = RR(4, 1, 2, 1);
    t->lw[14].refrat_m_a = RR(4, 1, 2, 1);
    t->lw[15].refrat_planck_a = RR(1, 6, 6, 6);
#undef RR
    olw_tables(t->exp_tbl, t->tfn_tbl, t->tau_tbl);

    /* ---- SW ---- */
    t->sw_preflog = need_d(t, "sw/preflog", &bad);
    t->sw_tref = need_d(t, "sw/tref", &bad);
    t->sf01 = need_d(t, "sw/sfluxref01", &bad);
    t->sf02 = need_d(t, "sw/sfluxref02", &bad);
    t->sf03 = need_d(t, "sw/sfluxref03", &bad);
    t->sw_strrat = need_d(t, "sw/strrat", &bad);
    t->sw_specwt = need_d(t, "sw/specwt", &bad);
    const double *sk = need_d(t, "sw/scalekur", &bad);
    t->layreffr = ot_i(&t->f, "sw/layreffr", 14);
    t->ix1 = ot_i(&t->f, "sw/ix1", 14);
    t->ix2 = ot_i(&t->f, "sw/ix2", 14);
    t->ibx = ot_i(&t->f, "sw/ibx", 14);
    t->extliq1 = need_d(t, "sw/extliq1", &bad);
    t->ssaliq1 = need_d(t, "sw/ssaliq1", &bad);
    t->asyliq1 = need_d(t, "sw/asyliq1", &bad);
    t->extice3 = need_d(t, "sw/extice3", &bad);
    t->ssaice3 = need_d(t, "sw/ssaice3", &bad);
    t->asyice3 = need_d(t, "sw/asyice3", &bad);
    if (bad || !sk || !t->layreffr || !t->ix1 || !t->ix2 || !t->ibx) return WR_ERR_TABLES;
    t->scalekur = sk[0];
    ns = 0;
    for (int b = 0; b < 14; ++b) {
        SwBand *B = &t->sw[b];
        const int jb = b + 16;
        B->ng = SW_NG_[b];
        B->ns = ns;
        ns += B->ng;
#define SWT(field, key) snprintf(nm, sizeof nm, "sw/b%d/%s", jb, key), B->field = opt_d(t, nm)
        SWT(absa, "absa");
        SWT(absb, "absb");
        SWT(selfref, "selfref");
        SWT(forref, "forref");
        SWT(rayla, "rayla");
        SWT(raylb, "raylb");
        SWT(abso3a, "abso3a");
        SWT(abso3b, "abso3b");
        SWT(absch4, "absch4");
        SWT(absh2o, "absh2o");
        SWT(absco2, "absco2");
#undef SWT
        snprintf(nm, sizeof nm, "sw/b%d/rayl", jb);
        const ot_entry *e = NULL;
        for (int q = 0; q < t->f.n; ++q)
            if (strcmp(t->f.e[q].name, nm) == 0) e = &t->f.e[q];
        if (e && e->n == 1)
            B->rayl = e->d[0];
        else if (e)
            B->rayl_v = e->d;
        B->strrat = t->sw_strrat[b];
    }
    /* strrat(16)/(28) come with their band files in the reference (kgb16/kgb28 .npz) */
    const double *s16 = opt_d(t, "sw/b16/strrat16"), *s28 = opt_d(t, "sw/b28/strrat28");
    if (s16) t->sw[0].strrat = s16[0];
    if (s28) t->sw[12].strrat = s28[0];
    osw_table(t->sw_exp_tbl);
    return WR_OK;
}

/* ---- driver prep: grrad_setup (phys_rrtm_grrad_setup.cu:102), grrad_gases (.cu:56), getgases,
 *      progcld2 with the cloud fraction given (phys_rrtm_progcld2.cu:92, cld_fraction != 0 branch) */
static void o_prep(const wr_columns *in, const wr_params *p, int i, Rad *r)
{
    const int nm = in->nlay, nrl = nm + 1, nc = in->ncol;
    r->nrl = nrl;
#define F(a, k) (in->a[(size_t)(k)*nc + i])
    for (int k0 = 0; k0 < nm; ++k0) {
        const int l = k0 + 1;
        const double prsi = F(p_lev, k0);
        const double prsl = 0.5 * (F(p_lev, k0) + F(p_lev, k0 + 1));
        r->plvl[l] = 10.0 * (prsi * 1.0e-3);
        r->plyr[l] = 10.0 * (prsl * 1.0e-3);
        r->tlyr[l] = F(t_lay, k0);
        r->olyr[l] = fmax(1.0e-10, F(o3_lay, k0));
    }
    r->plvl[nm + 1] = 10.0 * (F(p_lev, nm) * 1.0e-3);
    double tem1d = 1.0e-7;
    for (int s = 0; s < nm; ++s) {
        const double ql = fmax(tem1d, F(q_lay, s));
        tem1d = fmin(1.0e-7, ql);
        r->qlyr[s + 1] = ql;
    }
    r->plvl[0] = 1.0e-6;
    if (r->plvl[1] <= 1.0e-6) r->plvl[1] = 2.0e-6;
    r->plyr[0] = 0.5 * r->plvl[1];
    r->tlyr[0] = r->tlyr[1];
    r->qlyr[0] = r->qlyr[1];
    r->olyr[0] = in->o3_top ? fmax(1.0e-10, in->o3_top[i]) : r->olyr[1];
    r->tskn = in->t_sfc[i];
    /* tlvl by log-pressure interpolation; tlvl(top) = tlyr(top), tlvl(surface) = skin T */
    double la[O_NL], lb[O_NL];
    for (int l = 0; l < nrl; ++l) {
        la[l] = om_log(r->plyr[l]);
        lb[l] = om_log(r->plvl[l]);
    }
    r->tlvl[0] = r->tlyr[0];
    r->tlvl[nrl] = r->tskn;
    for (int l = 1; l < nrl; ++l)
        r->tlvl[l] = r->tlyr[l] + (r->tlyr[l - 1] - r->tlyr[l]) * (lb[l] - la[l]) / (la[l - 1] - la[l]);
    for (int g = 0; g < 10; ++g) r->vmr[g] = p->vmr[g];
    /* clouds [9]: frac, cwp, rew, cip, rei, crp, rer, snow path, snow radius; the added top layer
     * carries no cloud */
    for (int n = 0; n < 9; ++n) r->cld[0][n] = 0.0;
    r->cld[0][2] = p->re_liq;
    r->cld[0][4] = p->re_ice;
    r->cld[0][6] = 100.0;
    for (int l = 1; l < nrl; ++l) {
        const int k0 = l - 1;
        double ct = F(cldf, k0);
        if (in->cwp) { /* path form: progcld2's layer values as given */
            const int cl = ct >= 0.001;
            r->cld[l][0] = cl ? ct : 0.0;
            r->cld[l][1] = cl ? F(cwp, k0) : 0.0;
            r->cld[l][2] = F(rew, k0);
            r->cld[l][3] = cl ? F(cip, k0) : 0.0;
            r->cld[l][4] = F(rei, k0);
            r->cld[l][5] = (cl && in->crp) ? F(crp, k0) : 0.0;
            r->cld[l][6] = 100.0;
            r->cld[l][7] = 0.0;
            r->cld[l][8] = 0.0;
            continue;
        }
        const double qc = F(qc, k0), qi = F(qi, k0);
        double qcwat, qcice;
        if (r->tlyr[l] - 273.15 > -40.0) {
            qcwat = qc;
            qcice = qi;
        } else {
            qcwat = 0.0;
            qcice = qc + qi;
        }
        double cwp = 0.0, cip = 0.0;
        if (qcwat + qcice > 1.0e-12) {
            const double cpath = fabs(r->plvl[l + 1] - r->plvl[l]) * (100000.0 / O_CON_G);
            if (qcwat > 0.0) cwp = cpath * qcwat;
            cip = cpath * qcice;
        }
        if (ct < 0.001) {
            ct = 0.0;
            cwp = 0.0;
            cip = 0.0;
        }
        r->cld[l][0] = ct;
        r->cld[l][1] = cwp;
        r->cld[l][2] = p->re_liq;
        r->cld[l][3] = cip;
        r->cld[l][4] = p->re_ice;
        r->cld[l][5] = 0.0;
        r->cld[l][6] = 100.0;
        r->cld[l][7] = 0.0;
        r->cld[l][8] = 0.0;
    }
    /* aerosol optics, setaer's layout [field][band][radiation layer, top first][column] (float32) */
    const size_t ld = in->aer_ld > 0 ? (size_t)in->aer_ld : (size_t)nc;
    for (int l = 0; l < nrl; ++l) {
        for (int b = 0; b < 16; ++b)
            for (int f = 0; f < 2; ++f)
                r->aerlw[f][b][l] = in->aer_lw ? (double)in->aer_lw[((size_t)(f * 16 + b) * nrl + l) * ld + i] : 0.0;
        for (int b = 0; b < 14; ++b)
            for (int f = 0; f < 3; ++f)
                r->aersw[f][b][l] = in->aer_sw ? (double)in->aer_sw[((size_t)(f * 14 + b) * nrl + l) * ld + i] : 0.0;
    }
#undef F
}

/* ---- one column ------------------------------------------------------------------------------- */
typedef struct {
    LwCol lw;
    SwCol sw;
    Rad r;
    double taug[140][O_NL], fracs[140][O_NL];
    double staug[112][O_NL], staur[112][O_NL];
    uint32_t lwmask[O_NL * 5], swmask[O_NL * 4];
    double fu[O_NL + 1], fd[O_NL + 1], htr[O_NL];
} Work;

static void o_column(const OT *t, const wr_columns *in, const wr_params *p, const wr_outputs *out, const wr_diag *dg,
                     int i, Work *w)
{
    const int nm = in->nlay, nrl = nm + 1, nc = in->ncol;
    const uint64_t col = in->col_id ? in->col_id[i] : (uint64_t)i;
    o_prep(in, p, i, &w->r);
    /* LW: cloud paths scaled by CPATHFAC4LW = 1.5 (np3d = 3, phys_rrtm_grrad_lw.cu:58) */
    Rad rl = w->r;
    for (int l = 0; l < nrl; ++l) {
        rl.cld[l][1] = rl.cld[l][1] * 1.5;
        rl.cld[l][3] = rl.cld[l][3] * 1.5;
    }
    LwCol *c = &w->lw;
    olw_prep(&rl, in->emis[i], c);
    olw_cldprop(t, c);
    olw_setcoef(t, c);
    olw_taumol(t, c, w->taug, w->fracs);
    if (p->mcica) {
        o_mcica(col, p->seed, 0, 140, nrl, p->iovr, c->cldf, w->lwmask, 5);
        olw_rtrnmc(t, c, w->taug, w->fracs, w->lwmask, w->fu, w->fd, w->htr);
    } else {
        memset(w->lwmask, 0, sizeof w->lwmask);
        olw_rtrnmr(t, c, w->taug, w->fracs, w->fu, w->fd, w->htr);
    }
    /* lwrad_post / driver_post: top-first; model layer k0 is radiation layer k0 + 1 */
    for (int L = 0; L <= nrl; ++L) {
        out->flw_up[(size_t)L * nc + i] = w->fu[nrl - L];
        out->flw_dn[(size_t)L * nc + i] = w->fd[nrl - L];
    }
    for (int k0 = 0; k0 < nm; ++k0) out->hr_lw[(size_t)k0 * nc + i] = w->htr[nm - 1 - k0];
    if (dg) {
        for (int k = 0; k < nrl; ++k) {
            if (dg->lw_jp) dg->lw_jp[(size_t)k * nc + i] = c->jp[k];
            if (dg->lw_jt) dg->lw_jt[(size_t)k * nc + i] = c->jt[k];
            if (dg->lw_jt1) dg->lw_jt1[(size_t)k * nc + i] = c->jt1[k];
            if (dg->lw_mask)
                for (int wd = 0; wd < 5; ++wd) dg->lw_mask[((size_t)k * 5 + wd) * nc + i] = w->lwmask[k * 5 + wd];
        }
        if (dg->lw_laytrop) dg->lw_laytrop[i] = c->laytrop;
    }

    /* SW on day columns (coszen >= 1e-4, phys_rrtm_grrad_sw.cu:69) */
    const double cz = in->cosz[i];
    if (cz >= 1.0e-4) {
        SwCol *s = &w->sw;
        osw_prep(&w->r, in->alb[i], cz, p->solcon, s);
        osw_setcoef(t, s);
        osw_sfluxzen(t, s);
        osw_cldprop(t, s, p->mcica);
        osw_taumol(t, s, w->staug, w->staur);
        if (p->mcica)
            o_mcica(col, p->seed, 1, 112, nrl, p->iovr, s->cldf, w->swmask, 4);
        else
            memset(w->swmask, 0, sizeof w->swmask);
        double fdn0 = 0.0;
        osw_spcvrtc(t, s, w->staug, w->staur, w->swmask, p->mcica, w->fu, w->fd, &fdn0);
        if (out->fsw_dn_sfc_clr) out->fsw_dn_sfc_clr[i] = fdn0;
        for (int L = 0; L <= nrl; ++L) {
            out->fsw_up[(size_t)L * nc + i] = w->fu[nrl - L];
            out->fsw_dn[(size_t)L * nc + i] = w->fd[nrl - L];
        }
        /* swrad_post heating: hswc(k) = (fnet(k+1) - fnet(k)) * heatfac / delp(k), fnet = dn - up */
        for (int k = 0; k < nrl; ++k) {
            const double rf = O_HEATFAC / s->delp[k];
            w->htr[k] = ((w->fd[k + 1] - w->fu[k + 1]) - (w->fd[k] - w->fu[k])) * rf;
        }
        for (int k0 = 0; k0 < nm; ++k0) out->hr_sw[(size_t)k0 * nc + i] = w->htr[nm - 1 - k0];
        if (dg) {
            for (int k = 0; k < nrl; ++k) {
                if (dg->sw_jp) dg->sw_jp[(size_t)k * nc + i] = s->jp[k];
                if (dg->sw_jt) dg->sw_jt[(size_t)k * nc + i] = s->jt[k];
                if (dg->sw_jt1) dg->sw_jt1[(size_t)k * nc + i] = s->jt1[k];
                if (dg->sw_mask)
                    for (int wd = 0; wd < 4; ++wd) dg->sw_mask[((size_t)k * 4 + wd) * nc + i] = w->swmask[k * 4 + wd];
            }
            if (dg->sw_laytrop) dg->sw_laytrop[i] = s->laytrop;
        }
    } else {
        for (int L = 0; L <= nrl; ++L) out->fsw_up[(size_t)L * nc + i] = out->fsw_dn[(size_t)L * nc + i] = 0.0;
        if (out->fsw_dn_sfc_clr) out->fsw_dn_sfc_clr[i] = 0.0;
        for (int k0 = 0; k0 < nm; ++k0) out->hr_sw[(size_t)k0 * nc + i] = 0.0;
        if (dg) {
            for (int k = 0; k < nrl; ++k) {
                if (dg->sw_jp) dg->sw_jp[(size_t)k * nc + i] = -1;
                if (dg->sw_jt) dg->sw_jt[(size_t)k * nc + i] = -1;
                if (dg->sw_jt1) dg->sw_jt1[(size_t)k * nc + i] = -1;
                if (dg->sw_mask)
                    for (int wd = 0; wd < 4; ++wd) dg->sw_mask[((size_t)k * 4 + wd) * nc + i] = 0u;
            }
            if (dg->sw_laytrop) dg->sw_laytrop[i] = -1;
        }
    }
}

/* ---- entry points ----------------------------------------------------------------------------- */
static OT *g_tab = NULL;
static char g_path[1024];

int oracle_init(const char *table_path)
{
    if (g_tab && strcmp(g_path, table_path) == 0) return WR_OK;
    OT *t = (OT *)calloc(1, sizeof(OT));
    const int rc = ot_bind(t, table_path);
    if (rc != WR_OK) {
        ot_free(&t->f);
        free(t);
        return rc;
    }
    if (g_tab) {
        ot_free(&g_tab->f);
        free(g_tab);
    }
    g_tab = t;
    snprintf(g_path, sizeof g_path, "%s", table_path);
    return WR_OK;
}

int oracle_run(const wr_columns *in, const wr_params *p, const wr_outputs *out, const wr_diag *dg, int nthreads)
{
    if (!g_tab) return WR_ERR_TABLES;
    if (!in || !p || !out || in->ncol < 0 || in->nlay < 2 || in->nlay + 1 > O_NL) return WR_ERR_ARG;
    const int nc = in->ncol;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        Work *w = (Work *)malloc(sizeof(Work));
#pragma omp for schedule(dynamic, 16)
        for (int i = 0; i < nc; ++i) o_column(g_tab, in, p, out, dg, i, w);
        free(w);
    }
    return WR_OK;
}

/* pieces for tests */
const OT *oracle_tables(void) { return g_tab; }
void oracle_lw_tables(double *e, double *f, double *t) { olw_tables(e, f, t); }
void oracle_sw_table(double *e) { osw_table(e); }
double oracle_exp(double x) { return om_exp(x); }
double oracle_log(double x) { return om_log(x); }
double oracle_pow(double x, double y) { return om_pow(x, y); }
void oracle_philox(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { om_philox4x32_10(ctr, key, out); }
void oracle_mcica(uint64_t col, uint64_t seed, int kind, int ngpt, int nl, int iovr, const double *cf, uint32_t *bits,
                  int words)
{
    o_mcica(col, seed, kind, ngpt, nl, iovr, cf, bits, words);
}

/* CAL_CLDFRA (phys_rad_cal_cldfra.cu:29-55) */
void oracle_cal_cldfra(int im, int jm, int lm, int i_lo, int i_hi, int j_lo, int j_hi, int f_qc, int f_qi, float *cldfra,
                       const float *qc, const float *qi)
{
    for (int l = 0; l < lm; ++l)
        for (int j = 0; j < jm; ++j)
            for (int i = 0; i < im; ++i) {
                if (i < i_lo || i > i_hi || j < j_lo || j > j_hi) continue;
                const size_t p = ((size_t)l * jm + j) * im + i;
                if (f_qi && f_qc)
                    cldfra[p] = (qc[p] + qi[p] > 1.0e-6f) ? 1.0f : 0.0f;
                else if (f_qc)
                    cldfra[p] = (qc[p] > 1.0e-6f) ? 1.0f : 0.0f;
                else
                    cldfra[p] = 0.0f;
            }
}

void wr_default_params(wr_params *p)
{
    memset(p, 0, sizeof(*p));
    p->mcica = 1;
    p->iovr = 1;
    p->seed = 0;
    p->solcon = 1368.22;
    p->re_liq = 10.0;
    p->re_ice = 50.0;
    const double v[10] = {410.0e-6, 0.33e-6, 1.9e-6, 0.209, 1.50e-8, 3.520e-10, 6.358e-10, 1.500e-10, 1.397e-10, 8.2000e-11};
    for (int g = 0; g < 10; ++g) p->vmr[g] = v[g];
}
